"""Host-side loading set-up for the 2D meshes: Neumann traction assembly and damage preseeding.

These run once per problem on the host (numpy) and hand device tensors to the solvers; they
are not on the per-step path.  Reference files (paths relative to the reference's
``src/phast``): ``physics/boundary_conditions.py`` (``get_neumann_forces``),
``physics/initial_conditions.py`` and ``io/config/run_config.py`` (preseeding).
"""
from __future__ import annotations

from typing import Iterable

import numpy as np
import torch

F64 = torch.float64


def _edges(elements: np.ndarray, element_type: str, reference_q4_edges: bool):
    """Element edges as two endpoint arrays, edge k of every element before edge k+1.
    T3: (0,1), (1,2), (2,0).  Q4: (0,1), (1,2), (2,3), (3,0); with ``reference_q4_edges`` the
    three triangle edges (0,1), (1,2), (2,0) only, as the reference enumerates them for any
    element (boundary_conditions.py:536-541)."""
    nv = elements.shape[1]
    if element_type == "Q4" and not reference_q4_edges:
        pairs = [(0, 1), (1, 2), (2, 3), (3, 0)]
    else:
        pairs = [(0, 1), (1, 2), (2, 0)]
    if max(max(p) for p in pairs) >= nv:
        raise ValueError("element connectivity has too few corners for its type")
    a = np.concatenate([elements[:, i] for i, _ in pairs])
    b = np.concatenate([elements[:, j] for _, j in pairs])
    return a, b


def neumann_forces(mesh, node_indices, traction, load_factor: float = 1.0, ramp_factor: float = 1.0,
                   reference_q4_edges: bool = False, device=None) -> torch.Tensor:
    """Nodal forces (N,2) of a uniform traction on the boundary edges whose two endpoints are
    in ``node_indices`` (boundary_conditions.py:499-580): each edge owned by exactly one
    element gives ``traction * length / 2 * load_factor * ramp_factor`` to both endpoints.

    Edges are deduplicated by ``min * N + max`` and visited in increasing order of that key;
    the forces are added first at the lower endpoints, then at the upper ones, as the
    reference's two ``scatter_add_`` calls do.  On Q4 meshes all four cell edges are used;
    ``reference_q4_edges=True`` reproduces the reference, which visits only edges (0,1),
    (1,2), (2,0) of a cell and so never loads the top edge of a structured Q4 plate."""
    nodes = np.asarray(mesh.nodes, dtype=np.float64)
    els = np.asarray(mesh.elements, dtype=np.int64)
    n = nodes.shape[0]
    in_set = np.zeros(n, dtype=bool)
    in_set[np.asarray(node_indices, dtype=np.int64)] = True
    ea, eb = _edges(els, getattr(mesh, "element_type", "T3"), reference_q4_edges)
    both = in_set[ea] & in_set[eb]
    ea, eb = ea[both], eb[both]
    lo, hi = np.minimum(ea, eb), np.maximum(ea, eb)
    ids, first, counts = np.unique(lo * n + hi, return_index=True, return_counts=True)
    keep = first[counts == 1]  # boundary edges belong to one element only
    lo, hi = lo[keep], hi[keep]
    f = np.zeros((n, 2))
    if len(lo):
        t = np.asarray(traction, dtype=np.float64)
        length = np.linalg.norm(nodes[lo] - nodes[hi], axis=1)
        per_node = (length / 2.0 * load_factor * ramp_factor)[:, None] * t[None, :]
        np.add.at(f, lo, per_node)
        np.add.at(f, hi, per_node)
    dev = getattr(mesh, "device", None) if device is None else device
    return torch.as_tensor(f, dtype=F64, device=dev)


# ----------------------------------------------------------------------------------------
# damage preseeding (physics/initial_conditions.py, io/config/run_config.py:427-478)
# ----------------------------------------------------------------------------------------

def _mean_edge_length(mesh) -> float:
    """Mean over the triangle edges (0,1), (1,2), (2,0) of every element (initial_conditions.py:105-113)."""
    nodes, els = np.asarray(mesh.nodes), np.asarray(mesh.elements)
    e = [np.linalg.norm(nodes[els[:, i]] - nodes[els[:, j]], axis=1) for i, j in ((0, 1), (1, 2), (2, 0))]
    return float(np.concatenate(e).mean())


def _region_mask(nodes: np.ndarray, region: dict, mesh) -> np.ndarray:
    rtype = region["type"]
    if rtype == "line_segment":  # initial_conditions.py:38-54
        thickness = region.get("thickness")
        thickness = _mean_edge_length(mesh) if thickness is None else float(thickness)
        p0, p1 = np.asarray(region["from"], float), np.asarray(region["to"], float)
        seg = p1 - p0
        L2 = float(seg @ seg)
        if L2 == 0.0:
            return np.linalg.norm(nodes - p0, axis=1) <= thickness
        t = np.clip(((nodes - p0) @ seg) / L2, 0.0, 1.0)
        return np.linalg.norm(nodes - (p0 + t[:, None] * seg), axis=1) <= thickness
    if rtype == "rectangle":  # :57-63
        o, s = np.asarray(region["origin"], float), np.asarray(region["size"], float)
        up = o + s
        return (nodes[:, 0] >= o[0]) & (nodes[:, 0] <= up[0]) & (nodes[:, 1] >= o[1]) & (nodes[:, 1] <= up[1])
    if rtype == "circle":  # :66-68
        return np.linalg.norm(nodes - np.asarray(region["center"], float), axis=1) <= float(region["radius"])
    if rtype == "polygon":  # :71-102, even-odd ray cast
        V = np.asarray(region["vertices"], float)
        if V.shape[0] < 3:
            raise ValueError("polygon requires at least 3 vertices")
        x, y = nodes[:, 0], nodes[:, 1]
        inside = np.zeros(len(nodes), dtype=bool)
        j = len(V) - 1
        for i in range(len(V)):
            (xi, yi), (xj, yj) = V[i], V[j]
            cond = (yi > y) != (yj > y)
            den = (yj - yi) if yj != yi else 1.0
            inside ^= cond & (x < (xj - xi) * (y - yi) / den + xi)
            j = i
        return inside
    raise ValueError(f"Unknown preseed_damage region type '{rtype}'. "
                     "Supported: line_segment, rectangle, circle, polygon")


def resolve_preseed_specs(mesh, specs: Iterable[dict]):
    """``(node_mask, value_per_node)`` of preseed specs (initial_conditions.py:116-196): each
    spec is ``{nodes: <node set>}`` or ``{region: {type: ...}}`` with a ``value`` in [0, 1]
    (default 1); overlapping specs keep the maximum value."""
    nodes = np.asarray(mesh.nodes, dtype=np.float64)
    n = nodes.shape[0]
    mask_all = np.zeros(n, dtype=bool)
    vals = np.zeros(n)
    for spec in specs:
        if not isinstance(spec, dict):
            raise TypeError(f"preseed_damage entry must be a dict, got {type(spec)}")
        value = float(spec.get("value", 1.0))
        if not 0.0 <= value <= 1.0:
            raise ValueError(f"preseed_damage value must be in [0, 1], got {value}")
        if "nodes" in spec:
            sets = getattr(mesh, "node_sets", {})
            if spec["nodes"] not in sets:
                raise RuntimeError(f"preseed_damage: mesh has no node set '{spec['nodes']}'. "
                                   f"Available: {list(sets)}")
            mask = np.zeros(n, dtype=bool)
            mask[np.asarray(sets[spec["nodes"]], dtype=np.int64)] = True
        elif "region" in spec:
            region = spec["region"]
            if not isinstance(region, dict) or "type" not in region:
                raise ValueError(f"preseed_damage region must be a dict with 'type': {region}")
            mask = _region_mask(nodes, region, mesh)
        else:
            raise ValueError(f"preseed_damage entry must specify 'region' or 'nodes': {spec}")
        mask_all |= mask
        vals = np.where(mask, np.maximum(vals, value), vals)
    return mask_all, vals


def normalise_legacy_preseed(notch_nodesets):
    """``preseed_notch_nodesets: [a, b]`` -> ``[{nodes: a, value: 1}, ...]`` (initial_conditions.py:199-203)."""
    return [{"nodes": name, "value": 1.0} for name in notch_nodesets]


def value_to_H_seed(value: float, Gc: float, l0: float, saturation: float = 0.999) -> float:
    """History giving damage ``value`` at the AT2 zero-gradient equilibrium
    d = 2H / (Gc/l0 + 2H), i.e. H = d (Gc/l0) / (2 (1 - d)); ``1e4 Gc/l0`` at or above
    ``saturation`` (initial_conditions.py:206-226)."""
    if value >= saturation:
        return 1.0e4 * Gc / l0
    if value <= 0.0:
        return 0.0
    return value * (Gc / l0) / (2.0 * (1.0 - value))


def preseed_history(mesh, H_elem: torch.Tensor, specs: Iterable[dict], Gc: float, l0: float) -> int:
    """Raise the element history of the one-ring of seeded nodes to ``value_to_H_seed`` of the
    largest value among each element's seeded nodes (run_config.py:440-470).  ``H_elem`` is
    (E,) (T3) or (E, 4) (Q4 Gauss points, all raised together); it is updated in place.
    Returns the number of seeded elements.  The caller then solves the damage once."""
    mask, vals = resolve_preseed_specs(mesh, specs)
    if not mask.any():
        return 0
    els = np.asarray(mesh.elements, dtype=np.int64)
    elem_in = mask[els]
    elem_value = (vals[els] * elem_in).max(axis=1)
    elem_mask = elem_in.any(axis=1)
    seed = np.zeros(len(els))
    for v in np.unique(elem_value[elem_mask]):
        if v > 0.0:
            seed[elem_mask & (elem_value == v)] = value_to_H_seed(float(v), Gc, l0)
    s = torch.as_tensor(seed, dtype=H_elem.dtype, device=H_elem.device)
    sel = torch.as_tensor(seed > 0.0, device=H_elem.device)
    if H_elem.dim() == 2:
        s, sel = s.unsqueeze(1).expand_as(H_elem), sel.unsqueeze(1).expand_as(H_elem)
    H_elem.copy_(torch.where(sel, torch.maximum(H_elem, s), H_elem))
    return int(elem_mask.sum())
