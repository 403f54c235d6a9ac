"""Native-Q4 explicit phase-field path on MI355X: host mirror of the reference interface.

Classes keep the reference's names, arguments and return conventions so that callers of
``FEMOperators`` / ``ExplicitDynamics`` / ``PhaseFieldDamageSolver`` /
``StaggeredSolver`` (explicit branch) can switch over.  All arithmetic on the hot path
runs in ``libphast_mi.so``; this module only prepares geometry constants, allocates
outputs with ``torch.empty`` on the tensors' device and passes raw pointers.

Reference files (paths relative to the reference's ``src/phast``):
``core/fem_operators.py``, ``core/mesh.py``, ``core/quad_elements.py``,
``physics/material.py``, ``solvers/mechanics_solver.py``, ``solvers/damage_solver.py``,
``solvers/staggered_solver.py``.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import Q4Desc, check, ptr, stream

F64 = torch.float64


# ----------------------------------------------------------------------------------------
# material and mesh
# ----------------------------------------------------------------------------------------

@dataclass
class Material:
    """Subset of ``physics/material.py:Material`` used by the native-Q4 AT2 path."""
    E: float
    nu: float
    rho: float
    Gc: float
    l0: float
    eta_residual: float = 1e-7
    plane_stress: bool = False
    energy_split: str = "amor"
    pf_model: str = "AT2"
    degradation_type: str = "standard"
    stress_degradation: str = "split"
    at1_threshold: object = "auto"
    gamma_correction: bool = False

    def __post_init__(self):
        if self.stress_degradation not in ("split", "full"):  # material.py:185-188
            raise ValueError(f"stress_degradation must be 'split' or 'full', got {self.stress_degradation!r}")
        if self.energy_split not in ("amor", "isotropic", "spectral"):
            raise NotImplementedError(
                f"energy_split={self.energy_split!r}: the HIP path implements 'amor', 'isotropic' and "
                "'spectral' (T3 only)")
        if self.pf_model not in ("AT2", "AT1") or self.degradation_type != "standard":
            raise NotImplementedError("the HIP path implements AT2 (and AT1 on T3 meshes) with the standard "
                                      "degradation")

    @property
    def at1_source(self) -> float:
        """AT1 source S_H of the damage right-hand side 2H - S_H: the override or 3 Gc/(8 l0);
        0 for AT2 (material.py:260-271)."""
        if self.pf_model != "AT1":
            return 0.0
        if isinstance(self.at1_threshold, str) and self.at1_threshold == "auto":
            return 3.0 * self.Gc / (8.0 * self.l0)
        return float(self.at1_threshold)

    @property
    def stress_split(self) -> str:
        """Split of the stress (material.py:367-376): 'isotropic' under
        ``stress_degradation='full'``; psi+ and the history keep ``energy_split``."""
        return "isotropic" if self.stress_degradation == "full" else self.energy_split

    @property
    def lam(self) -> float:  # material.py:379-388
        if self.plane_stress:
            return self.E * self.nu / (1.0 - self.nu ** 2)
        return self.E * self.nu / ((1 + self.nu) * (1 - 2 * self.nu))

    @property
    def mu(self) -> float:  # material.py:390-393
        return self.E / (2 * (1 + self.nu))

    @property
    def kappa(self) -> float:  # material.py:395-402
        return self.E / (3.0 * (1.0 - 2.0 * self.nu))

    @property
    def c_p(self) -> float:  # material.py:413-430
        return math.sqrt((self.lam + 2 * self.mu) / self.rho)

    @property
    def Gc_over_l0(self) -> float:
        return self.Gc / self.l0

    @property
    def Gc_times_l0(self) -> float:
        return self.Gc * self.l0


def _q4_reference_geometry(coords: np.ndarray):
    """Quadrature data of one cell with corner coordinates ``coords`` (4,2), computed as in
    ``core/quad_elements.py:315-349`` (Jacobian, inverse, weighted determinant)."""
    a = 1.0 / np.sqrt(np.float64(3.0))
    pts = [(-a, -a), (a, -a), (-a, a), (a, a)]
    N = np.zeros((4, 4))
    grad = np.zeros((4, 4, 2))
    wdet = np.zeros(4)
    for q, (xi, eta) in enumerate(pts):
        N[q] = 0.25 * np.array([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta), (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)])
        dN = np.stack([0.25 * np.array([-(1 - eta), (1 - eta), (1 + eta), -(1 + eta)]),
                       0.25 * np.array([-(1 - xi), -(1 + xi), (1 + xi), (1 - xi)])], axis=1)
        J = dN.T @ coords
        det = np.linalg.det(J)
        if det <= 0:
            raise ValueError("Q4 element has non-positive Jacobian")
        grad[q] = dN @ np.linalg.inv(J).T
        wdet[q] = det
    return N, grad, wdet


def _structured_layout(nodes: np.ndarray, elements: np.ndarray):
    """(nx, ny, x, y) when the mesh is the structured plate of ``rectangular_sent_q4_structured``
    (congruent axis-aligned cells on a uniform tensor grid, nodes j*(nx+1)+i, cells j*nx+i with
    corners (i,j),(i+1,j),(i+1,j+1),(i,j+1)); None otherwise."""
    x = np.unique(nodes[:, 0])
    y = np.unique(nodes[:, 1])
    nx, ny = len(x) - 1, len(y) - 1
    if nx < 1 or ny < 1 or (nx + 1) * (ny + 1) != nodes.shape[0] or nx * ny != elements.shape[0]:
        return None
    jj, ii = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    if not (np.array_equal(nodes[:, 0], x[ii.ravel()]) and np.array_equal(nodes[:, 1], y[jj.ravel()])):
        return None
    je, ie = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    n0 = (je * (nx + 1) + ie).ravel()
    expect = np.stack([n0, n0 + 1, n0 + nx + 2, n0 + nx + 1], axis=1)
    if not np.array_equal(elements, expect):
        return None
    hx, hy = np.diff(x), np.diff(y)
    if np.ptp(hx) > 1e-12 * hx.max() or np.ptp(hy) > 1e-12 * hy.max():
        return None
    return nx, ny, x, y


class FEMMesh:
    """Q4 mesh (core/mesh.py:305-360 precompute).

    ``nodes`` (N,2) and ``elements`` (E,4) are kept on the CPU for host-side set-up
    (boundary sets, Neumann loads).  A structured plate (the mesh of
    ``core/mesh_generator.py:1695-1818``: congruent axis-aligned cells numbered as the
    generator numbers them) runs on the strip kernels, which exploit the one congruent cell;
    any other quadrilateral mesh (``structured`` False) runs on the general-quadrilateral plan,
    whose passes recompute each cell's Gauss-point geometry from its corners.  Every
    Gauss-point Jacobian must be positive (counter-clockwise corners).  ``device`` is the ROCm
    device the fields live on.
    """

    element_type = "Q4"

    def __init__(self, nodes, elements, device=None):
        nodes = np.asarray(nodes, dtype=np.float64)
        elements = np.asarray(elements, dtype=np.int64)
        if nodes.ndim != 2 or nodes.shape[1] != 2 or elements.ndim != 2 or elements.shape[1] != 4:
            raise ValueError("Q4 mesh needs nodes (N,2) and elements (E,4)")
        if elements.size and (elements.min() < 0 or elements.max() >= nodes.shape[0]):
            raise ValueError("Q4 elements refer to nodes outside the mesh")
        layout = _structured_layout(nodes, elements)
        self.structured = layout is not None
        self.nodes, self.elements = nodes, elements
        self.n_nodes, self.n_elems = nodes.shape[0], elements.shape[0]
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self.dtype = F64
        c = nodes[elements]
        self._elem_h = np.min(np.stack([np.linalg.norm(c[:, (k + 1) % 4] - c[:, k], axis=1) for k in range(4)],
                                       axis=1), axis=1)
        self.h_min = float(self._elem_h.min())  # mesh.py:331-339
        if self.structured:
            nx, ny, x, y = layout
            self.nx, self.ny = nx, ny
            self.quad_N, self.quad_grad_phi_cell, self.quad_wdetJ_cell = _q4_reference_geometry(nodes[elements[0]])
            self.hx, self.hy = float(x[1] - x[0]), float(y[1] - y[0])
            self.node_sets = {
                "bottom": np.nonzero(nodes[:, 1] == y[0])[0], "top": np.nonzero(nodes[:, 1] == y[-1])[0],
                "left": np.nonzero(nodes[:, 0] == x[0])[0], "right": np.nonzero(nodes[:, 0] == x[-1])[0],
            }
        else:
            self.nx = self.ny = None
            # boundary sets by the bounding box (mesh.py:380-405 identify_boundaries)
            lo, hi = nodes.min(axis=0), nodes.max(axis=0)
            tol = 1e-6 * max(float((hi - lo).max()), 1.0)
            self.node_sets = {
                "bottom": np.nonzero(np.abs(nodes[:, 1] - lo[1]) < tol)[0],
                "top": np.nonzero(np.abs(nodes[:, 1] - hi[1]) < tol)[0],
                "left": np.nonzero(np.abs(nodes[:, 0] - lo[0]) < tol)[0],
                "right": np.nonzero(np.abs(nodes[:, 0] - hi[0]) < tol)[0],
            }

    @property
    def elem_h_np(self) -> np.ndarray:
        """Per-cell minimum edge length (mesh.py:331-338)."""
        return self._elem_h

    @classmethod
    def rectangle(cls, W: float, H: float, nx: int, ny: int, device=None) -> "FEMMesh":
        """Plate ``[0,W] x [0,H]`` with ``nx x ny`` cells (``rectangular_sent_q4_structured``)."""
        x = np.linspace(0.0, W, nx + 1)
        y = np.linspace(0.0, H, ny + 1)
        xx, yy = np.meshgrid(x, y, indexing="xy")
        nodes = np.column_stack([xx.ravel(), yy.ravel()])
        je, ie = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
        n0 = (je * (nx + 1) + ie).ravel()
        elements = np.stack([n0, n0 + 1, n0 + nx + 2, n0 + nx + 1], axis=1)
        return cls(nodes, elements, device=device)


def q4_desc(mesh: FEMMesh, mat: Material, dt: float = 0.0) -> Q4Desc:
    """C descriptor of one plate."""
    if not mesh.structured:
        raise NotImplementedError("the strip kernels (and the batched plans) need structured Q4 plates; a general "
                                  "quadrilateral mesh runs through FEMOperators on the general-quadrilateral plan")
    d = Q4Desc()
    d.nx, d.ny = mesh.nx, mesh.ny
    for i, v in enumerate(mesh.quad_grad_phi_cell.ravel()):
        d.gradN[i] = float(v)
    for i, v in enumerate(mesh.quad_N.ravel()):
        d.N[i] = float(v)
    for i, v in enumerate(mesh.quad_wdetJ_cell):
        d.wdetJ[i] = float(v)
    d.lam, d.mu, d.kappa, d.nu = mat.lam, mat.mu, mat.kappa, mat.nu
    d.plane_stress = int(bool(mat.plane_stress))
    _check_q4_material(mat)
    d.split = 0 if mat.energy_split == "amor" else 1
    d.stress_split = 0 if mat.stress_split == "amor" else 1
    d.eta_residual, d.Gc, d.l0, d.dt = mat.eta_residual, mat.Gc, mat.l0, float(dt)
    d.hx, d.hy = mesh.hx, mesh.hy
    return d


def _check_q4_material(mat: Material) -> None:
    if mat.energy_split not in ("amor", "isotropic"):
        raise NotImplementedError(f"energy_split={mat.energy_split!r}: the Q4 HIP path implements 'amor' and "
                                  "'isotropic'")
    if mat.pf_model != "AT2":  # damage_solver.py:1256-1263: the native Q4 damage is AT2 only
        raise NotImplementedError("native Q4 damage supports only pf_model='AT2'")


class GeneralQ4Plan:
    """Owner of a general-quadrilateral plan (``phast_q4u_plan_create``): any Q4 mesh with
    positive Gauss-point Jacobians; element passes recompute each cell's geometry."""

    def __init__(self, mesh: FEMMesh, mat: Material):
        _lib.require_device()
        _check_q4_material(mat)
        nodes = np.ascontiguousarray(mesh.nodes, dtype=np.float64)
        els = np.ascontiguousarray(mesh.elements, dtype=np.int32)
        h = C.c_void_p()
        check(_lib.lib().phast_q4u_plan_create(
            mesh.n_nodes, mesh.n_elems, nodes.ctypes.data, els.ctypes.data, mat.E, mat.nu, mat.Gc, mat.l0,
            mat.eta_residual, int(bool(mat.plane_stress)), 0 if mat.energy_split == "amor" else 1,
            0 if mat.stress_split == "amor" else 1, C.byref(h)), "phast_q4u_plan_create")
        self.handle = h
        M = np.empty(mesh.n_nodes)
        hmin, area = C.c_double(), C.c_double()
        check(_lib.lib().phast_t3_mesh_info(h, M.ctypes.data, C.byref(hmin), C.byref(area)), "mesh_info")
        self.M_scalar_host, self.h_min, self.total_area = M, hmin.value, area.value

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _lib.lib().phast_t3_plan_destroy(h)
            self.handle = None


class Q4Plan:
    """Owner of a ``phast_q4_plan`` handle over one or more packed plates."""

    def __init__(self, descs: list[Q4Desc], tile=(128, 8)):
        _lib.require_device()
        arr = (Q4Desc * len(descs))(*descs)
        h = C.c_void_p()
        check(_lib.lib().phast_q4_plan_create(arr, len(descs), int(tile[0]), int(tile[1]), C.byref(h)),
              "phast_q4_plan_create")
        self.handle = h
        self.n_inst = len(descs)
        self.tile = tuple(tile)
        nn, nc = C.c_int64(), C.c_int64()
        check(_lib.lib().phast_q4_plan_sizes(h, C.byref(nn), C.byref(nc)))
        self.n_nodes, self.n_cells = nn.value, nc.value
        no = (C.c_int64 * self.n_inst)()
        co = (C.c_int64 * self.n_inst)()
        check(_lib.lib().phast_q4_plan_offsets(h, no, co))
        self.node_off = list(no)
        self.cell_off = list(co)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _lib.lib().phast_q4_plan_destroy(h)
            self.handle = None


def pf_c_alpha(pf_model: str) -> float:
    """c_w of the crack-surface density (damage_solver.py:1491-1498): 2 for AT2, 8/3 for AT1."""
    return 8.0 / 3.0 if pf_model == "AT1" else 2.0


def gamma_factor(elem_h: np.ndarray, l0: float, pf_model: str) -> np.ndarray:
    """Per-element gamma-convergence factor 1 / (1 + h_e / (c_w l0)) (Bourdin et al. 2000;
    damage_solver.py:1414-1417)."""
    return 1.0 / (1.0 + elem_h / (pf_c_alpha(pf_model) * l0))


def _gamma_Gc_field(solver, Gc_field) -> torch.Tensor:
    """Gc_eff_e = Gc * gamma_e of a solver built with ``material.gamma_correction``
    (damage_solver.py:1409-1433); the per-element Gc branch of the kernels then applies it."""
    if Gc_field is not None:
        raise ValueError("gamma_correction builds its own per-element Gc; pass Gc_field to solve() instead")
    m = solver.material
    g = m.Gc * gamma_factor(solver.mesh.elem_h_np, m.l0, m.pf_model)
    return torch.as_tensor(g, dtype=F64, device=solver.fem.device)


def _dev_check(t: torch.Tensor, shape, name: str, device) -> torch.Tensor:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")
    if t.dtype != F64:
        raise TypeError(f"{name} must be float64, got {t.dtype}")
    if t.device != device:
        raise ValueError(f"{name} must be on {device}, got {t.device}")
    return t.contiguous()


# ----------------------------------------------------------------------------------------
# FEMOperators (Q4 branch)
# ----------------------------------------------------------------------------------------

class FEMOperators:
    """Q4 branch of ``core/fem_operators.py:FEMOperators`` on the HIP library."""

    def __init__(self, mesh: FEMMesh, material: Material, ctx=None, tile=(128, 8)):
        self.mesh, self.material = mesh, material
        self.device, self.dtype = mesh.device, F64
        self._tile = tile
        self._plans: dict[float, Q4Plan] = {}
        self.general = None if mesh.structured else GeneralQ4Plan(mesh, material)
        if self.general is not None:
            M = torch.as_tensor(self.general.M_scalar_host, dtype=F64, device=self.device)
        else:
            M = torch.empty(mesh.n_nodes, dtype=F64, device=self.device)
            check(_lib.lib().phast_q4_lumped_mass(self.plan(0.0).handle, ptr(M), stream()), "lumped mass")
        self.M_scalar = M  # mesh.py:346-352
        # fem_operators.py:83-85 (interleaved vector mass and its guarded inverse)
        self.M_vec = material.rho * M.unsqueeze(1).expand(-1, 2).flatten()
        self.M_vec_inv = 1.0 / (self.M_vec + 1e-30)
        self.M_inv_node = 1.0 / (material.rho * M + 1e-30)
        self.dt_cfl = mesh.h_min / material.c_p  # fem_operators.py:88

    def plan(self, dt: float) -> Q4Plan:
        p = self._plans.get(float(dt))
        if p is None:
            p = Q4Plan([q4_desc(self.mesh, self.material, dt)], tile=self._tile)
            self._plans[float(dt)] = p
        return p

    def _u(self, u, name="u"):
        return _dev_check(u, (self.mesh.n_nodes, 2), name, self.device)

    def _d(self, d, name="d"):
        return _dev_check(d, (self.mesh.n_nodes,), name, self.device)

    def internal_force(self, u: torch.Tensor, d: torch.Tensor) -> torch.Tensor:
        """``f_int (N,2)`` (fem_operators.py:598-612)."""
        u, d = self._u(u), self._d(d)
        f = torch.empty_like(u)
        if self.general is not None:
            check(_lib.lib().phast_t3_internal_force(self.general.handle, ptr(u), ptr(d), ptr(f), stream()),
                  "internal_force")
            return f
        check(_lib.lib().phast_q4_internal_force(self.plan(0.0).handle, ptr(u), ptr(d), ptr(f), stream()),
              "internal_force")
        return f

    def compute_strain(self, u: torch.Tensor):
        """Gauss-point strains ``(eps_xx, eps_yy, gam_xy)``, each (E,4) (fem_operators.py:244-254)."""
        u = self._u(u)
        eps = torch.empty((3, self.mesh.n_elems, 4), dtype=F64, device=self.device)
        if self.general is not None:
            check(_lib.lib().phast_t3_element_fields(self.general.handle, ptr(u), None, ptr(eps), None, None,
                                                     stream()), "compute_strain")
            return eps[0], eps[1], eps[2]
        check(_lib.lib().phast_q4_strain(self.plan(0.0).handle, ptr(u), ptr(eps), stream()), "compute_strain")
        return eps[0], eps[1], eps[2]

    def compute_psi_plus(self, u: torch.Tensor, strain=None, d=None) -> torch.Tensor:
        """psi+ (E,4) (fem_operators.py:951-1006); ``strain`` is recomputed from ``u``."""
        u = self._u(u)
        psi = torch.empty((self.mesh.n_elems, 4), dtype=F64, device=self.device)
        if self.general is not None:
            check(_lib.lib().phast_t3_element_fields(self.general.handle, ptr(u), None, None, None, ptr(psi),
                                                     stream()), "compute_psi_plus")
            return psi
        check(_lib.lib().phast_q4_driving_force(self.plan(0.0).handle, ptr(u), ptr(psi), None, None, stream()),
              "compute_psi_plus")
        return psi

    def compute_driving_force(self, u: torch.Tensor, strain=None, d=None) -> torch.Tensor:
        """'strain_energy' driving force == psi+ (fem_operators.py:1173-1204)."""
        return self.compute_psi_plus(u, strain=strain, d=d)

    def driving_force_and_history(self, u: torch.Tensor, H_old: torch.Tensor):
        """Fused psi+ and ``H = max(H_old, psi+)``; returns ``(psi, H_new)``."""
        u = self._u(u)
        H_old = _dev_check(H_old, (self.mesh.n_elems, 4), "H_old", self.device)
        psi = torch.empty_like(H_old)
        H_new = torch.empty_like(H_old)
        if self.general is not None:
            psi = self.compute_psi_plus(u)
            return psi, history_update(H_old, psi)
        check(_lib.lib().phast_q4_driving_force(self.plan(0.0).handle, ptr(u), ptr(psi), ptr(H_old), ptr(H_new),
                                                stream()), "driving_force")
        return psi, H_new

    def elem_to_node(self, field_q: torch.Tensor) -> torch.Tensor:
        """Gauss-point field -> nodes (core/mesh.py:407-428)."""
        if field_q.shape == (self.mesh.n_elems,):
            field_q = field_q.unsqueeze(1).expand(-1, 4)
        field_q = _dev_check(field_q.contiguous(), (self.mesh.n_elems, 4), "field", self.device)
        out = torch.empty(self.mesh.n_nodes, dtype=F64, device=self.device)
        if self.general is not None:
            check(_lib.lib().phast_t3_elem_to_node(self.general.handle, ptr(field_q), ptr(self.M_scalar), ptr(out),
                                                   stream()), "elem_to_node")
            return out
        check(_lib.lib().phast_q4_elem_to_node(self.plan(0.0).handle, ptr(field_q), ptr(self.M_scalar), ptr(out),
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


def history_update(H_old: torch.Tensor, psi: torch.Tensor) -> torch.Tensor:
    """``H = max(H_old, psi)`` (physics/h_update.py:47-49), each entry one operand bit-exactly."""
    if H_old.shape != psi.shape or H_old.dtype != F64 or psi.dtype != F64:
        raise ValueError("H_old and psi must be float64 tensors of one shape")
    H_old, psi = H_old.contiguous(), psi.contiguous()
    out = torch.empty_like(H_old)
    check(_lib.lib().phast_history_max(H_old.numel(), ptr(H_old), ptr(psi), ptr(out), stream()), "history_max")
    return out


# ----------------------------------------------------------------------------------------
# ExplicitDynamics
# ----------------------------------------------------------------------------------------

class ExplicitDynamics:
    """Central difference with lumped mass (solvers/mechanics_solver.py:86-239).

    ``step`` returns fresh ``(u, v, a)``; ``step_with_history`` additionally fuses the
    driving force and the history maximum of the staggered explicit branch.
    Kelvin-Voigt damping (``damping_ratio_max``) is supported; rigid connectors are not.
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
        m = bc_mask.to(torch.uint8).contiguous()
        v = self.fem._u(bc_vals, "bc_vals")
        return m, v

    def _launch(self, u, v, a, d, f_ext, bc_mask, bc_vals, H_old=None, want_psi=False):
        fem = self.fem
        u, v, a, d = fem._u(u), fem._u(v, "v"), fem._u(a, "a"), fem._d(d)
        fe = None if f_ext is None else fem._u(f_ext, "f_ext")
        m, bv = self._bc(bc_mask, bc_vals)
        u1, v1, a1 = torch.empty_like(u), torch.empty_like(v), torch.empty_like(a)
        H_new = psi = None
        if H_old is not None:
            H_old = _dev_check(H_old, (fem.mesh.n_elems, 4), "H_old", fem.device)
            H_new = torch.empty_like(H_old)
            psi = torch.empty_like(H_old) if want_psi else None
        fd = None
        if fem.general is not None:  # general quadrilaterals: history array always written
            g = fem.general.handle
            if H_old is None:
                H_old = torch.zeros((fem.mesh.n_elems, 4), dtype=F64, device=fem.device)
                H_scr = torch.empty_like(H_old)
            else:
                H_scr = H_new
            if self._damping_beta > 0.0:
                fd = torch.empty_like(v)
                check(_lib.lib().phast_t3_damping_force(g, ptr(v), ptr(d), self._damping_beta, ptr(fd), stream()),
                      "damping_force")
            check(_lib.lib().phast_t3_explicit_step(
                g, float(self.dt), ptr(u), ptr(v), ptr(a), ptr(d), ptr(fe), ptr(fd), ptr(fem.M_inv_node), ptr(m),
                ptr(bv), ptr(H_old), ptr(u1), ptr(v1), ptr(a1), ptr(H_scr), ptr(psi), stream()), "explicit_step")
            return u1, v1, a1, H_new, psi
        plan = fem.plan(self.dt)
        if self._damping_beta > 0.0:  # mechanics_solver.py:192-195
            fd = torch.empty_like(v)
            check(_lib.lib().phast_q4_damping_force(plan.handle, ptr(v), ptr(d), self._damping_beta, ptr(fd),
                                                    stream()), "damping_force")
        check(_lib.lib().phast_q4_explicit_step(
            plan.handle, ptr(u), ptr(v), ptr(a), ptr(d), ptr(fe), ptr(fd), ptr(fem.M_inv_node), ptr(m), ptr(bv),
            ptr(H_old), ptr(u1), ptr(v1), ptr(a1), ptr(H_new), ptr(psi), stream()), "explicit_step")
        return u1, v1, a1, H_new, psi

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


# ----------------------------------------------------------------------------------------
# PhaseFieldDamageSolver (native Q4 AT2)
# ----------------------------------------------------------------------------------------

class PhaseFieldDamageSolver:
    """Native-Q4 AT2 damage solve (solvers/damage_solver.py:1157-1487, 2493-2684, 3377-3775).

    ``bounds_method`` 'projected_cg' runs the projected Jacobi-PCG on ``[d_prev, 1]``;
    'post_clamp' (the reference default) runs unconstrained PCG and clamps.  The whole
    iteration lives on the device; the host reads a convergence flag every few
    iterations.  ``Gc_field`` (E,) selects the per-cell Gc branch.
    """

    def __init__(self, fem: FEMOperators, tol: float = 1e-5, max_iter: int = 5000, ctx=None,
                 use_multigrid: bool = True, bounds_method: str = "post_clamp", preconditioner: str | None = None,
                 nodal_H: bool = False, Gc_field: torch.Tensor | None = None):
        if bounds_method not in ("post_clamp", "projected_cg", "direct"):
            raise ValueError(f"bounds_method must be 'post_clamp', 'projected_cg', or 'direct', got '{bounds_method}'")
        if bounds_method == "direct":
            raise NotImplementedError("Native Q4 AT2 damage supports matrix-free CG bounds methods "
                                      "('post_clamp' and 'projected_cg')")
        if nodal_H:
            raise NotImplementedError("Native Q4 damage dispatch supports only element/Gauss-point H")
        self.fem, self.mesh, self.material = fem, fem.mesh, fem.material
        self.tol, self.max_iter = float(tol), int(max_iter)
        self._bounds_method = bounds_method
        self.Gc_field = None if Gc_field is None else _dev_check(Gc_field, (fem.mesh.n_elems,), "Gc_field", fem.device)
        self._gamma_correction = bool(getattr(self.material, "gamma_correction", False))
        if self._gamma_correction:
            self.Gc_field = _gamma_Gc_field(self, Gc_field)
        self.last_iter = 0
        self.last_converged = False
        self.last_status = ""
        self.last_active = None
        self.n_solves = 0

    def _normalize_H(self, H):
        E = self.mesh.n_elems
        if H.shape == (E,):
            H = H.unsqueeze(1).expand(-1, 4)
        return _dev_check(H.contiguous(), (E, 4), "H_input", self.fem.device)

    def damage_rhs_diag(self, H_input):
        H = self._normalize_H(H_input)
        n = self.mesh.n_nodes
        b = torch.empty(n, dtype=F64, device=self.fem.device)
        diag = torch.empty_like(b)
        if self.fem.general is not None:
            check(_lib.lib().phast_t3_damage_rhs_diag(self.fem.general.handle, ptr(H), ptr(self.Gc_field), ptr(b),
                                                      ptr(diag), stream()), "damage_rhs_diag")
            return b, diag
        check(_lib.lib().phast_q4_damage_rhs_diag(self.fem.plan(0.0).handle, ptr(H), ptr(self.Gc_field), ptr(b),
                                                  ptr(diag), stream()), "damage_rhs_diag")
        return b, diag

    def damage_apply(self, H_input, x):
        """``A x`` of the AT2 operator at history H (damage_solver.py:1797-1838)."""
        H = self._normalize_H(H_input)
        x = self.fem._d(x, "x")
        y = torch.empty_like(x)
        if self.fem.general is not None:
            check(_lib.lib().phast_t3_damage_apply(self.fem.general.handle, ptr(H), ptr(self.Gc_field), ptr(x),
                                                   ptr(y), stream()), "damage_apply")
            return y
        check(_lib.lib().phast_q4_damage_apply(self.fem.plan(0.0).handle, ptr(H), ptr(self.Gc_field), ptr(x),
                                               ptr(y), stream()), "damage_apply")
        return y

    def solve(self, H_input, d_prev, Gc=None, l0=None, Gc_field=None, pf_dirichlet_mask=None,
              pf_dirichlet_values=None, initial_guess=None):
        """damage_solver.py:2493-2640 (forward, non-differentiable)."""
        self.n_solves += 1
        if Gc is not None or l0 is not None:
            raise NotImplementedError("scalar Gc/l0 overrides: rebuild the operators with the new material")
        if any(isinstance(t, torch.Tensor) and t.requires_grad for t in (H_input, d_prev, Gc_field)):
            raise NotImplementedError("Differentiable native Q4 damage solves are not implemented "
                                      "(use phast_amd.adjoint for gradients)")
        d_prev = self.fem._d(d_prev, "d_prev")
        if not bool(torch.isfinite(d_prev).all()):  # damage_solver.py:2647-2650
            self.last_iter, self.last_converged = 0, False
            return d_prev.clone()
        H = self._normalize_H(H_input)
        gcf = self.Gc_field if Gc_field is None else _dev_check(Gc_field, (self.mesh.n_elems,), "Gc_field",
                                                                 self.fem.device)
        fixed = vals = None
        if pf_dirichlet_mask is not None and bool(pf_dirichlet_mask.any()):
            if pf_dirichlet_values is None:
                raise ValueError("pf_dirichlet_values must be supplied with mask")
            fixed = self.fem._d(pf_dirichlet_mask.to(F64), "pf_dirichlet_mask").to(torch.uint8)
            vals = self.fem._d(pf_dirichlet_values.to(F64), "pf_dirichlet_values")
        init = None if initial_guess is None else self.fem._d(initial_guess.to(F64), "initial_guess")
        d_out = torch.empty_like(d_prev)
        method = 0 if self._bounds_method == "projected_cg" else 1
        active = torch.empty(self.mesh.n_nodes, dtype=torch.uint8, device=self.fem.device) if method == 0 else None
        iters = (C.c_int32 * 1)()
        status = (C.c_int32 * 1)()
        if self.fem.general is not None:
            check(_lib.lib().phast_t3_damage_solve(
                self.fem.general.handle, method, ptr(H), ptr(gcf), ptr(d_prev), ptr(init), ptr(fixed), ptr(vals),
                self.tol, self.max_iter, ptr(d_out), ptr(active), iters, status, stream()), "damage_solve")
        else:
            check(_lib.lib().phast_q4_damage_solve(
                self.fem.plan(0.0).handle, method, ptr(H), ptr(gcf), ptr(d_prev), ptr(init), ptr(fixed), ptr(vals),
                self.tol, self.max_iter, ptr(d_out), ptr(active), iters, status, stream()), "damage_solve")
        self.last_iter = int(iters[0])
        self.last_status = _lib.CG_STATUS_NAMES.get(int(status[0]), str(status[0]))
        self.last_converged = int(status[0]) == _lib.CG_CONVERGED
        self.last_active = None if active is None else active.bool()
        return d_out


# ----------------------------------------------------------------------------------------
# explicit staggered step (StaggeredSolver.step_full, explicit branch)
# ----------------------------------------------------------------------------------------

class StaggeredSolver:
    """Explicit branch of ``solvers/staggered_solver.py:StaggeredSolver``.

    State: ``u, v, a (N,2)``, ``d (N,)``, ``H_elem`` ((E,4) Gauss-point history on Q4
    plates, (E,) element history on T3 meshes), ``H_nodal (N,)`` (lazy).
    ``step_full`` = mechanics step fused with psi+ and the history update, then the
    damage solve gated by ``damage_every`` (staggered_solver.py:1115-1165).  The same class
    drives the Q4 and the T3 operators (``phast_amd.t3``).
    """

    def __init__(self, fem: FEMOperators, dyn: ExplicitDynamics, damage: PhaseFieldDamageSolver,
                 f_ext=None, bc_mask=None, bc_vals=None, damage_every: int = 3, enable_damage: bool = True):
        self.fem, self.dyn, self.damage_solver = fem, dyn, damage
        n, E, dev = fem.mesh.n_nodes, fem.mesh.n_elems, fem.device
        self.u = torch.zeros((n, 2), dtype=F64, device=dev)
        self.v = torch.zeros_like(self.u)
        self.a = torch.zeros_like(self.u)
        self.d = torch.zeros(n, dtype=F64, device=dev)
        hist = (E, 4) if getattr(fem.mesh, "element_type", "Q4") == "Q4" else (E,)
        self.H_elem = torch.zeros(hist, dtype=F64, device=dev)
        self.f_ext = torch.zeros_like(self.u) if f_ext is None else f_ext
        self.bc_mask, self.bc_vals = bc_mask, bc_vals
        self.damage_every = int(damage_every)
        self.enable_damage = enable_damage
        self._explicit_step_count = 0
        self._step_count = 0
        self.dt = dyn.dt

    def _H_update(self, H_old, psi):
        """staggered_solver.py:804-845 ('hard_max')."""
        return history_update(H_old, psi)

    @property
    def H_nodal(self):
        return self.fem.elem_to_node(self.H_elem)

    def step_full(self) -> torch.Tensor:
        u, v, a, H, psi = self.dyn.step_with_history(self.u, self.v, self.a, self.d, self.f_ext, self.bc_mask,
                                                     self.bc_vals, self.H_elem, want_psi=True)
        self.u, self.v, self.a = u, v, a
        if not self.enable_damage:
            self._step_count += 1
            return psi
        self.H_elem = H
        self._explicit_step_count += 1
        k = self._explicit_step_count
        if self.damage_every <= 1 or k <= 5 or k % self.damage_every == 0:
            self.d = self.damage_solver.solve(self.H_elem, self.d)
        self._step_count += 1
        return psi

    def get_state(self) -> dict:
        return {"u": self.u.clone(), "v": self.v.clone(), "a": self.a.clone(), "d": self.d.clone(),
                "H_elem": self.H_elem.clone(), "step": self._step_count,
                "explicit_step": self._explicit_step_count}

    def set_state(self, st: dict) -> None:
        self.u, self.v, self.a, self.d = (st[k].clone() for k in ("u", "v", "a", "d"))
        self.H_elem = st["H_elem"].clone()
        self._step_count = int(st["step"])
        self._explicit_step_count = int(st["explicit_step"])
