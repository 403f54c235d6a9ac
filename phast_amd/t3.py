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
import warnings

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr, stream
from .q4 import F64, Material, StaggeredSolver, _dev_check, _gamma_Gc_field, gamma_factor, history_update  # noqa: F401

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
        self.elem_h_np = 4.0 * self.areas_np / (per + 1e-30)  # mesh.py:275-276
        self.h_min = float(self.elem_h_np.min())
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
    ``Gc_field`` (E,) selects the per-element Gc branch.  The preconditioner is Jacobi whatever
    ``use_multigrid`` says (the reference's ``preconditioner='jacobi'`` setting; its GPU default
    'auto' would build a two-level aggregation multigrid, which is not on the HIP path).
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
        self._gamma_correction = bool(getattr(self.material, "gamma_correction", False))
        if self._gamma_correction:
            self.Gc_field = _gamma_Gc_field(self, Gc_field)
        # the preconditioner actually used (damage_solver.py:1446-1480 'jacobi'); no multigrid.
        # The reference resolves use_multigrid=True with no preconditioner to 'auto' for AT2
        # (aggregation multigrid, check_every 5), so its CG iterates differ from these.
        if use_multigrid and preconditioner is None and getattr(self.material, "pf_model", "AT2") == "AT2":
            warnings.warn("T3 damage solve: the reference's default preconditioner for AT2 ('auto', aggregation "
                          "multigrid) is not on the HIP path; using Jacobi (pass preconditioner='jacobi' or "
                          "use_multigrid=False to select it explicitly).  The converged damage agrees to the "
                          "solver tolerance; CG iteration counts and iterates differ from the reference's.",
                          stacklevel=2)
        self._preconditioner = "jacobi"
        self._use_multigrid = False
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

    # ------------------------------------------------------------------------------------
    # solve: dispatch to a forward solve or one of the three adjoint rules
    # ------------------------------------------------------------------------------------

    def solve(self, H_input, d_prev, Gc=None, l0=None, Gc_field=None, pf_dirichlet_mask=None,
              pf_dirichlet_values=None, initial_guess=None):
        """damage_solver.py:2493-2640.  Without ``differentiable`` or a gradient-tracked Gc, l0 or
        Gc_field the solve is forward only; otherwise it runs under the matching adjoint rule
        (``_AdjointDamageSolve``, ``_AdjointDamageSolveScalar``, ``_AdjointDamageSolveField``),
        whose backward costs two more device-resident CG solves."""
        self.n_solves += 1
        self._initial_guess = initial_guess
        try:
            tg = _needs_grad
            Gc_g, l0_g, field_g = tg(Gc), tg(l0), tg(Gc_field)
            diff = bool(getattr(self, "differentiable", False))
            if pf_dirichlet_mask is not None and not (diff or Gc_g or l0_g or field_g):
                return self._forward(H_input, d_prev, Gc, l0, Gc_field, pf_dirichlet_mask, pf_dirichlet_values)
            if Gc_field is not None:
                if Gc is not None:
                    raise ValueError("Pass either Gc (scalar) or Gc_field (per-element), not both.")
                if l0 is None:
                    l0 = torch.tensor(self.material.l0, dtype=F64)
                if not (diff or field_g or l0_g or tg(H_input) or tg(d_prev)):
                    return self._forward(H_input, d_prev, None, l0, Gc_field, pf_dirichlet_mask,
                                         pf_dirichlet_values)
                return _AdjointDamageSolveField.apply(self, H_input, d_prev, Gc_field, l0, pf_dirichlet_mask,
                                                      pf_dirichlet_values)
            if Gc is not None or l0 is not None:
                if Gc is None:
                    Gc = torch.tensor(self.material.Gc, dtype=F64)
                if l0 is None:
                    l0 = torch.tensor(self.material.l0, dtype=F64)
                if not (diff or Gc_g or l0_g or tg(H_input) or tg(d_prev)):
                    return self._forward(H_input, d_prev, Gc, l0, None, pf_dirichlet_mask, pf_dirichlet_values)
                return _AdjointDamageSolveScalar.apply(self, H_input, d_prev, Gc, l0, pf_dirichlet_mask,
                                                       pf_dirichlet_values)
            if not diff:
                return self._forward(H_input, d_prev, None, None, None, pf_dirichlet_mask, pf_dirichlet_values)
            return _AdjointDamageSolve.apply(self, H_input, d_prev, pf_dirichlet_mask, pf_dirichlet_values)
        finally:
            self._initial_guess = None

    def _pins(self, mask, values, need_values=True):
        """(uint8 mask, values) on the device, or (None, None) (damage_solver.py:2350-2360).  The
        adjoint solve takes a mask without values (pins held at 0)."""
        if mask is None or not bool(mask.any()):
            return None, None
        if values is None and need_values:
            raise ValueError("pf_dirichlet_values must be supplied with mask")
        fem = self.fem
        fixed = fem._d(mask.detach().to(F64), "pf_dirichlet_mask").to(torch.uint8)
        return fixed, None if values is None else fem._d(values.detach().to(F64), "pf_dirichlet_values")

    def _install(self, Gc=None, l0=None):
        """Temporary fracture constants (damage_solver.py:574-585): AT1 then uses
        S_H = 3 Gc/(8 l0).  Returns True when constants were replaced."""
        if Gc is None and l0 is None:
            return False
        m = self.material
        Gcv = m.Gc if Gc is None else float(Gc)
        l0v = m.l0 if l0 is None else float(l0)
        check(_lib.lib().phast_t3_set_fracture(self.fem.plan.handle, Gcv, l0v, 3.0 * Gcv / (8.0 * l0v)),
              "set_fracture")
        return True

    def _restore(self):
        m = self.material
        check(_lib.lib().phast_t3_set_fracture(self.fem.plan.handle, m.Gc, m.l0, m.at1_source), "set_fracture")

    def _forward(self, H_input, d_prev, Gc, l0, Gc_field, mask, values):
        """One device solve with optional temporary Gc / l0 (damage_solver.py:2376-2492, 2642-2684).
        Under gamma correction the per-element Gc built at construction governs a solve without a
        Gc field, as the reference's per-element arrays do; scalar overrides then change nothing."""
        if self._gamma_correction and Gc_field is None:
            Gc = l0 = None
        installed = self._install(None if Gc is None else _scalar(Gc), None if l0 is None else _scalar(l0))
        try:
            return self._solve_device(H_input, d_prev, Gc_field, mask, values)
        finally:
            if installed:
                self._restore()

    def _solve_device(self, H_input, d_prev, Gc_field, mask, values):
        fem = self.fem
        d_prev = fem._d(d_prev.detach(), "d_prev")
        if not bool(torch.isfinite(d_prev).all()):  # damage_solver.py:2647-2650
            self.last_iter, self.last_converged = 0, False
            return d_prev.clone()
        H = fem._e(H_input.detach(), "H_input")
        gcf = self.Gc_field if Gc_field is None else fem._e(Gc_field.detach(), "Gc_field")
        fixed, vals = self._pins(mask, values)
        ig = getattr(self, "_initial_guess", None)
        init = None if ig is None else fem._d(ig.detach().to(F64), "initial_guess")
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

    def adjoint_pcg(self, H_input, rhs=None, pf_dirichlet_mask=None, pf_dirichlet_values=None, rel_tol=1e-22,
                    Gc_field=None):
        """``_pcg_solve_adjoint_system`` (damage_solver.py:423-478) on the device: A x = rhs with the
        forward operator, pins eliminated (held at their values, 0 without values); ``rhs=None``
        solves for the unconstrained damage (rhs = b, pins at their values as its start).
        Returns x; ``last_adjoint_iter`` holds the iteration count."""
        fem = self.fem
        H = fem._e(H_input.detach(), "H_input")
        gcf = self.Gc_field if Gc_field is None else fem._e(Gc_field.detach(), "Gc_field")
        fixed, vals = self._pins(pf_dirichlet_mask, pf_dirichlet_values, need_values=False)
        if rhs is not None:
            rhs = fem._d(rhs.detach().to(F64), "rhs")
        x = torch.empty(self.mesh.n_nodes, dtype=F64, device=fem.device)
        iters, status = C.c_int32(), C.c_int32()
        check(_lib.lib().phast_t3_adjoint_solve(
            fem.plan.handle, ptr(H), ptr(gcf), ptr(rhs), ptr(fixed), ptr(vals), None, float(rel_tol),
            min(self.max_iter, 1000), ptr(x), C.byref(iters), C.byref(status), stream()), "adjoint_solve")
        self.last_adjoint_iter = int(iters.value)
        return x

    def _backward(self, H_input, d_prev, d_sol, grad_output, mode, mask, values, Gc_field=None):
        """The shared backward of the three rules (damage_solver.py:700-810, 931-1028, 1074-1152):
        the active-set split of the upstream gradient, the unconstrained damage and the adjoint
        by device PCG, then the per-triangle history gradient and quadratic forms."""
        fem = self.fem
        n, ne = self.mesh.n_nodes, self.mesh.n_elems
        d_sol, d_prev = d_sol.detach(), d_prev.detach()
        g0 = fem._d(grad_output.detach().to(F64).contiguous(), "grad_output")
        fixed, _ = self._pins(mask, values, need_values=False)
        # classify_damage_active_set (damage_solver.py:70-105) and the gradient split (:400-421)
        upper = d_sol >= 1.0 - 1e-12
        fx = torch.zeros_like(upper) if fixed is None else fixed.bool()
        lower = (d_sol <= d_prev + 1e-12) & ~upper & ~fx
        interior = ~(lower | (upper & ~fx) | fx)
        g = g0 * interior.to(F64)
        grad_prev = g0 * lower.to(F64)
        d_unc = self.adjoint_pcg(H_input, None, mask, values, 1e-24, Gc_field)
        lam = self.adjoint_pcg(H_input, g, mask, None, 1e-10 if mode == "H" else 1e-22, Gc_field)
        d_pull = d_unc if self._bounds_method in ("post_clamp", "direct") else d_sol  # :269-281
        gH = torch.empty(ne, dtype=F64, device=fem.device)
        forms = [torch.empty(ne, dtype=F64, device=fem.device) for _ in range(3)] if mode != "H" else [None] * 3
        check(_lib.lib().phast_t3_adjoint_terms(fem.plan.handle, ptr(lam), ptr(d_pull), ptr(d_unc), ptr(gH),
                                                *[ptr(f) for f in forms], stream()), "adjoint_terms")
        self.last_adjoint = {"lam": lam, "d_unc": d_unc}
        return gH, grad_prev, forms


def _needs_grad(t) -> bool:
    return bool(torch.is_grad_enabled() and isinstance(t, torch.Tensor) and t.requires_grad)


def _scalar(t) -> float:
    return float(t.detach().item()) if isinstance(t, torch.Tensor) else float(t)


class _AdjointDamageSolve(torch.autograd.Function):
    """Implicit differentiation of the damage solve w.r.t. H and d_prev
    (damage_solver.py:1038-1152): forward = the device solve, backward = two device PCG solves
    (unconstrained damage, adjoint at relative tolerance 1e-10) and the per-triangle pullback."""

    @staticmethod
    def forward(ctx, solver, H_input, d_prev, pf_dirichlet_mask=None, pf_dirichlet_values=None):
        d_new = solver._solve_device(H_input, d_prev, None, pf_dirichlet_mask, pf_dirichlet_values)
        ctx.solver = solver
        ctx.pins = (pf_dirichlet_mask, pf_dirichlet_values)
        ctx.save_for_backward(H_input, d_prev, d_new)
        return d_new

    @staticmethod
    def backward(ctx, grad_output):
        H_input, d_prev, d_sol = ctx.saved_tensors
        gH, gprev, _ = ctx.solver._backward(H_input, d_prev, d_sol, grad_output, "H", *ctx.pins)
        return None, gH, gprev, None, None


class _AdjointDamageSolveScalar(torch.autograd.Function):
    """Implicit differentiation w.r.t. H, d_prev and scalar Gc, l0 (damage_solver.py:482-856,
    without gamma correction).  Gc and l0 are installed on the plan for the forward solve and
    again for the backward's solves.  AT2: dL/dGc = -(l0 lKd + lMd/l0),
    dL/dl0 = -(Gc lKd - Gc/l0^2 lMd); AT1: dL/dGc = -3/(8 l0) lM1 - 0.75 l0 lKd,
    dL/dl0 = 3 Gc/(8 l0^2) lM1 - 0.75 Gc lKd."""

    @staticmethod
    def forward(ctx, solver, H_input, d_prev, Gc_t, l0_t, pf_dirichlet_mask=None, pf_dirichlet_values=None):
        Gcv, l0v = _scalar(Gc_t), _scalar(l0_t)
        ctx.gamma = None
        if solver._gamma_correction:  # damage_solver.py:545-585: per-element Gc_e = Gc gamma_e(l0)
            if l0_t.requires_grad:
                warnings.warn("_AdjointDamageSolveScalar with gamma_correction=True returns the l0 gradient "
                              "through the direct model dependence only; gamma_e(l0) is held fixed",
                              RuntimeWarning, stacklevel=3)
            gam = torch.as_tensor(gamma_factor(solver.mesh.elem_h_np, l0v, solver.material.pf_model), dtype=F64,
                                  device=solver.fem.device)
            ctx.gamma = gam
            d_new = solver._forward(H_input, d_prev, None, l0v, Gcv * gam, pf_dirichlet_mask, pf_dirichlet_values)
        else:
            d_new = solver._forward(H_input, d_prev, Gcv, l0v, None, pf_dirichlet_mask, pf_dirichlet_values)
        ctx.solver, ctx.Gc, ctx.l0 = solver, Gcv, l0v
        ctx.pins = (pf_dirichlet_mask, pf_dirichlet_values)
        ctx.like = (Gc_t.dtype, Gc_t.device, l0_t.dtype, l0_t.device)
        ctx.save_for_backward(H_input, d_prev, d_new)
        return d_new

    @staticmethod
    def backward(ctx, grad_output):
        H_input, d_prev, d_sol = ctx.saved_tensors
        s, Gcv, l0v, gam = ctx.solver, ctx.Gc, ctx.l0, ctx.gamma
        gd, gdev, ld, ldev = ctx.like
        if gam is not None:  # damage_solver.py:766-781: per-element forms reduced through gamma_e
            Ge = Gcv * gam
            s._install(None, l0v)
            try:
                gH, gprev, (lKd, lMd, lM1) = s._backward(H_input, d_prev, d_sol, grad_output, "scalar", *ctx.pins,
                                                         Gc_field=Ge)
            finally:
                s._restore()
            if s.material.pf_model == "AT1":
                gGc_e = -3.0 / (8.0 * l0v) * lM1 - 0.75 * l0v * lKd
                gl0 = 3.0 / (8.0 * l0v ** 2) * (Ge * lM1).sum() - 0.75 * (Ge * lKd).sum()
            else:
                gGc_e = -(l0v * lKd + (1.0 / l0v) * lMd)
                gl0 = -((Ge * lKd).sum() - (Ge / l0v ** 2 * lMd).sum())
            gGc = (gam * gGc_e).sum()
            return None, gH, gprev, gGc.to(gd).to(gdev), gl0.to(ld).to(ldev), None, None
        s._install(Gcv, l0v)
        try:
            gH, gprev, (lKd, lMd, lM1) = s._backward(H_input, d_prev, d_sol, grad_output, "scalar", *ctx.pins)
        finally:
            s._restore()
        sKd, sMd, sM1 = lKd.sum(), lMd.sum(), lM1.sum()
        if s.material.pf_model == "AT1":
            gGc = -3.0 / (8.0 * l0v) * sM1 - 0.75 * l0v * sKd
            gl0 = 3.0 * Gcv / (8.0 * l0v ** 2) * sM1 - 0.75 * Gcv * sKd
        else:
            gGc = -(l0v * sKd + (1.0 / l0v) * sMd)
            gl0 = -(Gcv * sKd - (Gcv / l0v ** 2) * sMd)
        return None, gH, gprev, gGc.to(gd).to(gdev), gl0.to(ld).to(ldev), None, None


class _AdjointDamageSolveField(torch.autograd.Function):
    """Implicit differentiation w.r.t. H, d_prev, a per-element Gc field and l0
    (damage_solver.py:859-1035).  AT2: dL/dGc_e = -(l0 lKd_e + lMd_e/l0),
    dL/dl0 = -(sum Gc_e lKd_e - sum Gc_e/l0^2 lMd_e); AT1: dL/dGc_e = -3/(8 l0) lM1_e - 0.75 l0 lKd_e,
    dL/dl0 = 3/(8 l0^2) sum Gc_e lM1_e - 0.75 sum Gc_e lKd_e."""

    @staticmethod
    def forward(ctx, solver, H_input, d_prev, Gc_e_t, l0_t, pf_dirichlet_mask=None, pf_dirichlet_values=None):
        l0v = _scalar(l0_t)
        d_new = solver._forward(H_input, d_prev, None, l0v, Gc_e_t, pf_dirichlet_mask, pf_dirichlet_values)
        ctx.solver, ctx.l0 = solver, l0v
        ctx.pins = (pf_dirichlet_mask, pf_dirichlet_values)
        ctx.like = (l0_t.dtype, l0_t.device)
        ctx.save_for_backward(H_input, d_prev, d_new, Gc_e_t)
        return d_new

    @staticmethod
    def backward(ctx, grad_output):
        H_input, d_prev, d_sol, Gc_e = ctx.saved_tensors
        s, l0v = ctx.solver, ctx.l0
        s._install(None, l0v)
        try:
            gH, gprev, (lKd, lMd, lM1) = s._backward(H_input, d_prev, d_sol, grad_output, "field", *ctx.pins,
                                                     Gc_field=Gc_e)
        finally:
            s._restore()
        Ge = Gc_e.detach()
        if s.material.pf_model == "AT1":
            gGc = -3.0 / (8.0 * l0v) * lM1 - 0.75 * l0v * lKd
            gl0 = 3.0 / (8.0 * l0v ** 2) * (Ge * lM1).sum() - 0.75 * (Ge * lKd).sum()
        else:
            gGc = -(l0v * lKd + (1.0 / l0v) * lMd)
            gl0 = -((Ge * lKd).sum() - (Ge / l0v ** 2 * lMd).sum())
        ld, ldev = ctx.like
        return None, gH, gprev, gGc, gl0.to(ld).to(ldev), None, None


def make_gamma_corrected_Gc_field(solver: PhaseFieldDamageSolver, Gc_scalar: torch.Tensor) -> torch.Tensor:
    """``Gc_scalar * gamma_e`` for ``solve(..., Gc_field=...)`` (damage_solver.py:219-266): a
    differentiable way to a scalar-Gc gradient under gamma correction."""
    if not solver._gamma_correction:
        raise ValueError("Solver was not built with gamma_correction=True. Either enable it via "
                         "material.gamma_correction=True, or use the scalar Gc path directly.")
    gam = gamma_factor(solver.mesh.elem_h_np, solver.material.l0, solver.material.pf_model)
    return Gc_scalar * torch.as_tensor(gam, dtype=Gc_scalar.dtype, device=Gc_scalar.device)

