"""Linear-triangle (T3) phase-field path on MI355X: host mirror of the reference interface.

The classes keep the names, arguments and return conventions of the reference's T3 branch
of ``FEMOperators``, of ``ExplicitDynamics`` and of ``PhaseFieldDamageSolver`` (element history,
Jacobi-preconditioned projected / post-clamp CG).  All arithmetic runs in ``libphast_mi.so``; this module uploads
the mesh once (the plan), allocates outputs with ``torch.empty`` on the fields' device and
passes raw pointers.  There is no CPU fallback.

Reference files (paths relative to the reference's ``src/phast``):
``core/mesh.py`` (``_precompute_t3``), ``core/fem_operators.py``, ``physics/material.py``,
``solvers/mechanics_solver.py``, ``solvers/damage_solver.py``.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr, stream
from .q4 import F64, Material, StaggeredSolver, _dev_check, history_update  # noqa: F401  (re-exported)

_SPLIT_ID = {"amor": 0, "isotropic": 1, "spectral": 2}


class T3Plan:
    """Owner of a ``phast_t3_plan``: the mesh on the device plus its per-node incidence lists."""

    def __init__(self, nodes: np.ndarray, elements: np.ndarray, material: Material):
        _lib.require_device()
        nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        els = np.ascontiguousarray(elements, dtype=np.int32)
        h = C.c_void_p()
        check(_lib.lib().phast_t3_plan_create(
            len(nodes), len(els), nodes.ctypes.data, els.ctypes.data, material.E, material.nu, material.Gc,
            material.l0, material.eta_residual, int(bool(material.plane_stress)),
            _SPLIT_ID[material.energy_split], _SPLIT_ID[material.stress_split], 1 if material.pf_model == "AT1" else 0,
            material.at1_source, C.byref(h)), "phast_t3_plan_create")
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _lib.lib().phast_t3_plan_destroy(h)
            self.handle = None


class FEMMesh:
    """Unstructured T3 mesh (``core/mesh.py:FEMMesh`` with ``element_type='T3'``).

    ``nodes`` (N,2) and ``elements`` (E,3) stay on the CPU for host-side set-up; the
    geometry the kernels use (areas, shape-function gradients) is recomputed on the device
    from the node coordinates with the formulas of ``_precompute_t3`` (mesh.py:226-300).
    """

    element_type = "T3"

    def __init__(self, nodes, elements, device=None):
        nodes = np.asarray(nodes, dtype=np.float64)
        elements = np.asarray(elements, dtype=np.int64)
        if nodes.ndim != 2 or nodes.shape[1] != 2 or elements.ndim != 2 or elements.shape[1] != 3:
            raise ValueError("T3 mesh needs nodes (N,2) and elements (E,3)")
        if elements.size and (elements.min() < 0 or elements.max() >= len(nodes)):
            raise ValueError("element connectivity refers to nodes outside the mesh")
        if 3 * len(elements) >= 2 ** 31:
            raise ValueError("T3 meshes are limited to 2^31 / 3 elements (32-bit incidence indices)")
        self.nodes, self.elements = nodes, elements
        self.n_nodes, self.n_elems = len(nodes), len(elements)
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self.dtype = F64
        p = nodes[elements]
        v01, v02, v12 = p[:, 1] - p[:, 0], p[:, 2] - p[:, 0], p[:, 2] - p[:, 1]
        self.areas_np = 0.5 * np.abs(v01[:, 0] * v02[:, 1] - v01[:, 1] * v02[:, 0])
        if np.any(self.areas_np <= 0.0):
            raise ValueError("degenerate (zero-area) triangle in the mesh")
        per = np.linalg.norm(v01, axis=1) + np.linalg.norm(v12, axis=1) + np.linalg.norm(v02, axis=1)
        self.h_min = float((4.0 * self.areas_np / (per + 1e-30)).min())  # mesh.py:276-283
        self.node_sets: dict[str, np.ndarray] = {}

    @classmethod
    def sent_split_quad(cls, W: float, H: float, nx: int, ny: int, a: float | None = None,
                        device=None) -> "FEMMesh":
        """Structured split-quad plate ``[0,W] x [0,H]`` of ``nx x ny`` cells, each split into
        two triangles with the checkerboard diagonal mirrored about mid-height
        (core/mesh_generator.py:1552-1633 ``rectangular_sent_liu_structured``: nodes
        ``j*(nx+1)+i``, triangles cell by cell in row order).  Node sets ``bottom``, ``top``,
        ``left``, ``right`` and, when ``a`` is given, the notch line ``crack`` (nodes with
        ``x <= a`` on ``y = H/2``)."""
        if nx < 2 or ny < 2:
            raise ValueError("structured SENT mesh needs at least 2 cells in x/y")
        x, y = np.linspace(0.0, W, nx + 1), np.linspace(0.0, H, ny + 1)
        xx, yy = np.meshgrid(x, y, indexing="xy")
        nodes = np.column_stack([xx.ravel(), yy.ravel()])
        j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
        n00 = j * (nx + 1) + i
        n10, n01, n11 = n00 + 1, n00 + nx + 1, n00 + nx + 2
        even = ((i + np.minimum(j, ny - 1 - j)) % 2 == 0)[..., None]
        t0 = np.where(even, np.stack([n00, n10, n11], -1), np.stack([n00, n10, n01], -1))
        t1 = np.where(even, np.stack([n00, n11, n01], -1), np.stack([n10, n11, n01], -1))
        els = np.stack([t0, t1], axis=2).reshape(-1, 3)
        mesh = cls(nodes, els, device=device)
        ids = np.arange(len(nodes)).reshape(ny + 1, nx + 1)
        mesh.node_sets = {"bottom": ids[0], "top": ids[-1], "left": ids[:, 0], "right": ids[:, -1]}
        if a is not None:
            jm, ic = int(round((H / 2.0) / (H / ny))), int(round(a / (W / nx)))
            if abs(y[jm] - H / 2.0) > 1e-10 or abs(x[ic] - a) > 1e-10:  # mesh_generator.py:1647-1652
                raise ValueError("H/2 and a must lie on the structured grid")
            mesh.node_sets["crack"] = ids[jm, : ic + 1]
        return mesh


class FEMOperators:
    """T3 branch of ``core/fem_operators.py:FEMOperators`` on the HIP library."""

    def __init__(self, mesh: FEMMesh, material: Material, ctx=None):
        if material.energy_split not in _SPLIT_ID:
            raise NotImplementedError(f"energy_split={material.energy_split!r}")
        self.mesh, self.material = mesh, material
        self.device, self.dtype = mesh.device, F64
        with torch.cuda.device(self.device):
            self.plan = T3Plan(mesh.nodes, mesh.elements, material)
        M = np.empty(mesh.n_nodes)
        hmin, area = C.c_double(), C.c_double()
        check(_lib.lib().phast_t3_mesh_info(self.plan.handle, M.ctypes.data, C.byref(hmin), C.byref(area)),
              "mesh_info")
        self.M_scalar = torch.from_numpy(M).to(self.device)  # mesh.py:287-289
        # fem_operators.py:83-86: interleaved vector mass and its guarded inverse
        self.M_vec = material.rho * self.M_scalar.unsqueeze(1).expand(-1, 2).flatten()
        self.M_vec_inv = 1.0 / (self.M_vec + 1e-30)
        self.M_inv_node = 1.0 / (material.rho * self.M_scalar + 1e-30)
        self.dt_cfl = hmin.value / material.c_p  # fem_operators.py:88-89
        self.path = self.set_path(0)

    def set_path(self, path: int) -> int:
        """0: element patches in LDS (default when the mesh allows), 1: every incidence
        re-evaluates its triangle.  Both give the same bits; returns the organisation in use."""
        act, mx = C.c_int32(), C.c_int32()
        check(_lib.lib().phast_t3_set_path(self.plan.handle, int(path), C.byref(act), C.byref(mx)), "set_path")
        self.path, self.max_patch_list = int(act.value), int(mx.value)
        return self.path

    def _u(self, u, name="u"):
        return _dev_check(u, (self.mesh.n_nodes, 2), name, self.device)

    def _d(self, d, name="d"):
        return _dev_check(d, (self.mesh.n_nodes,), name, self.device)

    def _e(self, h, name="H"):
        return _dev_check(h, (self.mesh.n_elems,), name, self.device)

    def internal_force(self, u: torch.Tensor, d: torch.Tensor) -> torch.Tensor:
        """``f_int (N,2)`` (fem_operators.py:598-612, T3: 223-242, 614-653, 661-679)."""
        u, d = self._u(u), self._d(d)
        f = torch.empty_like(u)
        check(_lib.lib().phast_t3_internal_force(self.plan.handle, ptr(u), ptr(d), ptr(f), stream()),
              "internal_force")
        return f

    def internal_force_and_history(self, u: torch.Tensor, d: torch.Tensor, H_old: torch.Tensor):
        """``(f_int, psi+, max(H_old, psi+))`` of one pass over the mesh."""
        u, d, H_old = self._u(u), self._d(d), self._e(H_old, "H_old")
        f, psi, H_new = torch.empty_like(u), torch.empty_like(H_old), torch.empty_like(H_old)
        check(_lib.lib().phast_t3_force_history(self.plan.handle, ptr(u), ptr(d), ptr(H_old), ptr(f), ptr(H_new),
                                                ptr(psi), stream()), "force_history")
        return f, psi, H_new

    def _fields(self, u, d=None, strain=False, stress=False, psi=False):
        u = self._u(u)
        d = None if d is None else self._d(d)
        E = self.mesh.n_elems
        eps = torch.empty((3, E), dtype=F64, device=self.device) if strain else None
        sig = torch.empty((3, E), dtype=F64, device=self.device) if stress else None
        ps = torch.empty(E, dtype=F64, device=self.device) if psi else None
        check(_lib.lib().phast_t3_element_fields(self.plan.handle, ptr(u), ptr(d), ptr(eps), ptr(sig), ptr(ps),
                                                 stream()), "element_fields")
        return eps, sig, ps

    def compute_strain(self, u: torch.Tensor):
        """Centroid strains ``(eps_xx, eps_yy, gam_xy)``, each (E,) (fem_operators.py:223-242)."""
        eps, _, _ = self._fields(u, strain=True)
        return eps[0], eps[1], eps[2]

    def compute_stress(self, u: torch.Tensor, d: torch.Tensor, strain=None):
        """Degraded element stresses ``(sxx, syy, sxy)``, each (E,), for the material's split
        (fem_operators.py:614-653); ``strain`` is recomputed from ``u``."""
        _, sig, _ = self._fields(u, d, stress=True)
        return sig[0], sig[1], sig[2]

    def compute_psi_plus(self, u: torch.Tensor, strain=None, d=None) -> torch.Tensor:
        """psi+ (E,) at the centroids (fem_operators.py:951-1030); ``strain`` is recomputed."""
        return self._fields(u, psi=True)[2]

    def compute_driving_force(self, u: torch.Tensor, strain=None, d=None) -> torch.Tensor:
        """'strain_energy' driving force == psi+ (fem_operators.py:1173-1204)."""
        return self.compute_psi_plus(u, strain=strain, d=d)

    def elem_to_node(self, field_e: torch.Tensor) -> torch.Tensor:
        """Area-weighted element -> node projection (core/mesh.py:407-433)."""
        field_e = self._e(field_e.contiguous(), "field")
        out = torch.empty(self.mesh.n_nodes, dtype=F64, device=self.device)
        check(_lib.lib().phast_t3_elem_to_node(self.plan.handle, ptr(field_e), ptr(self.M_scalar), ptr(out),
                                               stream()), "elem_to_node")
        return out

    def compute_reaction_force(self, u, d, node_indices, component: int = 1, f_ext=None) -> float:
        """fem_operators.py:1502-1530."""
        f = self.internal_force(u, d)
        idx = torch.as_tensor(node_indices, dtype=torch.long, device=self.device)
        r = f[idx, component]
        if f_ext is not None:
            r = r - f_ext[idx, component]
        return float(r.sum().item())


class ExplicitDynamics:
    """Central difference with lumped mass (solvers/mechanics_solver.py:86-239) on T3 meshes.

    ``step`` returns fresh ``(u, v, a)``; ``step_with_history`` also returns
    ``H = max(H_old, psi+(u1))`` and ``psi+`` of the staggered explicit branch.
    """

    def __init__(self, fem: FEMOperators, dt: float | None = None, dt_safety: float = 1.0,
                 differentiable: bool = False, damping_ratio_max: float = 0.0):
        if differentiable:
            raise NotImplementedError("differentiable explicit dynamics are not on the HIP path")
        if damping_ratio_max < 0.0:
            raise ValueError("damping_ratio_max must be non-negative")
        self.fem = fem
        # Kelvin-Voigt damping f_damp = beta f_int(v_n, d), beta = zeta dt_cfl; the CFL step
        # shrinks by sqrt(1 + zeta^2) - zeta (mechanics_solver.py:110-148)
        self.damping_ratio_max = float(damping_ratio_max)
        self._damping_beta = self.damping_ratio_max * fem.dt_cfl
        z = self.damping_ratio_max
        cfl_factor = math.sqrt(1.0 + z * z) - z if z > 0.0 else 1.0
        self.dt_cfl = fem.dt_cfl * dt_safety * cfl_factor
        self.dt = self.dt_cfl if dt is None else float(dt)
        self.last_iter = 1

    def _bc(self, bc_mask, bc_vals):
        if bc_mask is None:
            return None, None
        n = self.fem.mesh.n_nodes
        if not isinstance(bc_mask, torch.Tensor) or bc_mask.dtype not in (torch.bool, torch.uint8):
            raise TypeError("bc_mask must be a bool or uint8 tensor")
        if tuple(bc_mask.shape) != (n, 2) or bc_mask.device != self.fem.device:
            raise ValueError(f"bc_mask must have shape ({n}, 2) on {self.fem.device}")
        return bc_mask.to(torch.uint8).contiguous(), self.fem._u(bc_vals, "bc_vals")

    def _launch(self, u, v, a, d, f_ext, bc_mask, bc_vals, H_old=None, want_psi=False):
        fem = self.fem
        u, v, a, d = fem._u(u), fem._u(v, "v"), fem._u(a, "a"), fem._d(d)
        fe = None if f_ext is None else fem._u(f_ext, "f_ext")
        m, bv = self._bc(bc_mask, bc_vals)
        u1, v1, a1 = torch.empty_like(u), torch.empty_like(v), torch.empty_like(a)
        keep = H_old is not None
        if keep:
            H_old = fem._e(H_old, "H_old")
        else:
            H_old = torch.zeros(fem.mesh.n_elems, dtype=F64, device=fem.device)
        H_new = torch.empty_like(H_old)
        psi = torch.empty_like(H_old) if want_psi else None
        fd = None
        if self._damping_beta > 0.0:  # mechanics_solver.py:192-195
            fd = torch.empty_like(v)
            check(_lib.lib().phast_t3_damping_force(fem.plan.handle, ptr(v), ptr(d), self._damping_beta, ptr(fd),
                                                    stream()), "damping_force")
        check(_lib.lib().phast_t3_explicit_step(
            fem.plan.handle, self.dt, ptr(u), ptr(v), ptr(a), ptr(d), ptr(fe), ptr(fd), ptr(fem.M_inv_node), ptr(m),
            ptr(bv), ptr(H_old), ptr(u1), ptr(v1), ptr(a1), ptr(H_new), ptr(psi), stream()), "explicit_step")
        return u1, v1, a1, (H_new if keep else None), psi

    def step(self, u, v, a, d, f_ext, bc_mask=None, bc_vals=None, rigid_connectors=None):
        """mechanics_solver.py:204-239."""
        if rigid_connectors:
            raise NotImplementedError("rigid connectors are not on the HIP path")
        self.last_iter = 1
        u1, v1, a1, _, _ = self._launch(u, v, a, d, f_ext, bc_mask, bc_vals)
        return u1, v1, a1

    def step_with_history(self, u, v, a, d, f_ext, bc_mask, bc_vals, H_old, want_psi=True):
        """One step plus ``psi = psi+(u1)`` and ``H = max(H_old, psi)``: returns (u, v, a, H, psi)."""
        self.last_iter = 1
        return self._launch(u, v, a, d, f_ext, bc_mask, bc_vals, H_old=H_old, want_psi=want_psi)


class PhaseFieldDamageSolver:
    """AT2 damage solve of ``solvers/damage_solver.py:PhaseFieldDamageSolver`` on T3 meshes
    with element history H (E,) and no viscosity (damage_solver.py:1157-1487, 2160-2268,
    2493-2684, 3377-3775).

    ``bounds_method`` 'projected_cg' runs the projected Jacobi-PCG on ``[d_prev, 1]``;
    'post_clamp' (the reference default) runs unconstrained PCG and clamps.  The iteration
    lives on the device; the host reads a convergence flag every few iterations.
    ``Gc_field`` (E,) selects the per-element Gc branch.
    """

    def __init__(self, fem: FEMOperators, tol: float = 1e-5, max_iter: int = 5000, ctx=None,
                 use_multigrid: bool = True, bounds_method: str = "post_clamp", preconditioner: str | None = None,
                 nodal_H: bool = False, Gc_field: torch.Tensor | None = None):
        if bounds_method not in ("post_clamp", "projected_cg", "direct"):
            raise ValueError(f"bounds_method must be 'post_clamp', 'projected_cg', or 'direct', got '{bounds_method}'")
        if bounds_method == "direct":
            raise NotImplementedError("the T3 HIP damage path runs the matrix-free CG bounds methods "
                                      "('post_clamp' and 'projected_cg')")
        if preconditioner not in (None, "jacobi"):
            raise NotImplementedError("the T3 HIP damage path uses the Jacobi preconditioner")
        if nodal_H:
            raise NotImplementedError("the T3 HIP damage path takes element history H (E,)")
        self.fem, self.mesh, self.material = fem, fem.mesh, fem.material
        self.tol, self.max_iter = float(tol), int(max_iter)
        self._bounds_method = bounds_method
        self.Gc_field = None if Gc_field is None else fem._e(Gc_field, "Gc_field")
        self.last_iter = 0
        self.last_converged = False
        self.last_status = ""
        self.last_active = None
        self.n_solves = 0

    def damage_rhs_diag(self, H_input):
        """``(b, diag)`` of the AT2 system at element history H (damage_solver.py:2112-2119, 2249-2258)."""
        H = self.fem._e(H_input, "H_input")
        b = torch.empty(self.mesh.n_nodes, dtype=F64, device=self.fem.device)
        dg = torch.empty_like(b)
        check(_lib.lib().phast_t3_damage_rhs_diag(self.fem.plan.handle, ptr(H), ptr(self.Gc_field), ptr(b), ptr(dg),
                                                  stream()), "damage_rhs_diag")
        return b, dg

    def damage_apply(self, H_input, x):
        """``A x`` of the AT2 operator at history H (``_Ax`` T3, damage_solver.py:1715-1755)."""
        H = self.fem._e(H_input, "H_input")
        x = self.fem._d(x, "x")
        y = torch.empty_like(x)
        check(_lib.lib().phast_t3_damage_apply(self.fem.plan.handle, ptr(H), ptr(self.Gc_field), ptr(x), ptr(y),
                                               stream()), "damage_apply")
        return y

    def solve(self, H_input, d_prev, Gc=None, l0=None, Gc_field=None, pf_dirichlet_mask=None,
              pf_dirichlet_values=None, initial_guess=None):
        """damage_solver.py:2493-2640 (forward, non-differentiable)."""
        self.n_solves += 1
        if Gc is not None or l0 is not None:
            raise NotImplementedError("scalar Gc/l0 overrides: rebuild the operators with the new material")
        if any(isinstance(t, torch.Tensor) and t.requires_grad for t in (H_input, d_prev, Gc_field)):
            raise NotImplementedError("differentiable T3 damage solves are not on the HIP path")
        fem = self.fem
        d_prev = fem._d(d_prev, "d_prev")
        if not bool(torch.isfinite(d_prev).all()):  # damage_solver.py:2647-2650
            self.last_iter, self.last_converged = 0, False
            return d_prev.clone()
        H = fem._e(H_input, "H_input")
        gcf = self.Gc_field if Gc_field is None else fem._e(Gc_field, "Gc_field")
        fixed = vals = None
        if pf_dirichlet_mask is not None and bool(pf_dirichlet_mask.any()):
            if pf_dirichlet_values is None:
                raise ValueError("pf_dirichlet_values must be supplied with mask")
            fixed = fem._d(pf_dirichlet_mask.to(F64), "pf_dirichlet_mask").to(torch.uint8)
            vals = fem._d(pf_dirichlet_values.to(F64), "pf_dirichlet_values")
        init = None if initial_guess is None else fem._d(initial_guess.to(F64), "initial_guess")
        d_out = torch.empty_like(d_prev)
        method = 0 if self._bounds_method == "projected_cg" else 1
        active = torch.empty(self.mesh.n_nodes, dtype=torch.uint8, device=fem.device) if method == 0 else None
        iters, status = C.c_int32(), C.c_int32()
        check(_lib.lib().phast_t3_damage_solve(
            fem.plan.handle, method, ptr(H), ptr(gcf), ptr(d_prev), ptr(init), ptr(fixed), ptr(vals), self.tol,
            self.max_iter, ptr(d_out), ptr(active), C.byref(iters), C.byref(status), stream()), "damage_solve")
        self.last_iter = int(iters.value)
        self.last_status = _lib.CG_STATUS_NAMES.get(int(status.value), str(status.value))
        self.last_converged = int(status.value) == _lib.CG_CONVERGED
        self.last_active = None if active is None else active.bool()
        return d_out
