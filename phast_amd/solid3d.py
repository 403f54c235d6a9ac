"""Periodic hex8 voxel path on MI355X: host mirror of ``phast.solid3d`` operator calls.

``PeriodicVoxelPhaseFieldOperator`` keeps the reference's method names and return
conventions (solid3d/operator.py:133-1544); ``pcg`` / ``solve_constrained`` mirror
solid3d/linalg.py:99-287 and ``solve_h8_scalar_box_kkt`` mirrors
solvers/box_kkt_h8.py:603-778.  Operator-bound actions dispatch to the device-resident
solvers of ``libphast_mi.so``; every apply is a HIP kernel.  Set-up (phase maps, material
bank, pinned node) runs once on the host.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Callable, Mapping, Sequence

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr, stream

F64 = torch.float64
VOIGT = ("xx", "yy", "zz", "yz", "xz", "xy")
PHASE_NAMES = ("fibre", "bn_band", "cvi_sic", "mi_matrix", "pore")  # solid3d/model.py:40-49
PHASE_ID = {n: i for i, n in enumerate(PHASE_NAMES)}
DRIVER_NONE, DRIVER_FULL, DRIVER_SPECTRAL = 0, 1, 2  # solid3d/model.py:52
DEFAULT_LINEAR_TOLERANCE = 1.0e-10  # box_kkt_h8.py:13
DEFAULT_KKT_TOLERANCE = 1.0e-9      # box_kkt_h8.py:16
H8_SCALAR_BOX_KKT_IDENTIFIER = "h8_scalar_box_kkt_active_set_v1"

_lib.lib()  # the module is unusable without the HIP library
_L = _lib.lib()
for _name, _args in {
    "phast_hex8_mech_solve": [C.c_void_p] * 8 + [C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_int64,
                                                  C.c_void_p, C.POINTER(C.c_double), C.c_void_p],
    "phast_hex8_damage_dirichlet_pcg": [C.c_void_p, C.c_void_p, C.c_double] + [C.c_void_p] * 5 +
                                       [C.c_double, C.c_int32, C.c_int64, C.c_void_p, C.POINTER(C.c_double),
                                        C.c_void_p],
    "phast_hex8_box_update": [C.c_void_p, C.c_void_p, C.c_double] + [C.c_void_p] * 7 +
                             [C.c_double, C.c_int32, C.c_int64, C.c_void_p, C.POINTER(C.c_int64), C.c_void_p],
    "phast_hex8_kkt_stats": [C.c_void_p] + [C.c_void_p] * 7 + [C.c_double, C.c_int64, C.POINTER(C.c_double),
                                                              C.c_void_p],
    "phast_hex8_solver_reserve": [C.c_void_p],
    "phast_hex8_solver_release": [C.c_void_p],
    "phast_hex8_pfczm_setup": [C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_double],
    "phast_hex8_pfczm_degradation": [C.c_void_p] * 6,
    "phast_hex8_pfczm_energy": [C.c_void_p] * 5,
    "phast_hex8_pfczm_gradient": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p],
    "phast_hex8_pfczm_tangent_weights": [C.c_void_p] * 5,
    "phast_hex8_pfczm_tangent_action": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p],
    "phast_hex8_pfczm_tangent_diagonal": [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p],
    "phast_hex8_pfczm_dirichlet_pcg": [C.c_void_p, C.c_void_p, C.c_double] + [C.c_void_p] * 5 +
                                      [C.c_double, C.c_int32, C.c_int64, C.c_void_p, C.POINTER(C.c_double),
                                       C.c_void_p],
    "phast_hex8_pfczm_box_update": [C.c_void_p, C.c_void_p, C.c_double] + [C.c_void_p] * 7 +
                                   [C.c_double, C.c_int32, C.c_int64, C.c_void_p, C.POINTER(C.c_int64), C.c_void_p],
    "phast_hex8_closure_setup": [C.c_void_p, C.c_void_p],
    "phast_hex8_closure_setup_rows": [C.c_void_p] * 3,
    "phast_hex8_closure_pattern": [C.c_void_p] * 7,
    "phast_hex8_mech_action_cl": [C.c_void_p] * 4 + [C.c_int32, C.c_void_p, C.c_void_p],
    "phast_hex8_stress_sums_cl": [C.c_void_p] * 4 + [C.c_int32, C.c_void_p, C.c_void_p],
    "phast_hex8_mech_diagonal_cl": [C.c_void_p] * 3 + [C.c_int32, C.c_void_p, C.c_void_p],
    "phast_hex8_mech_solve_cl": [C.c_void_p] * 8 + [C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_int64,
                                                     C.c_void_p, C.POINTER(C.c_double), C.c_void_p],
}.items():
    _fn = getattr(_L, _name)
    _fn.argtypes = _args
    _fn.restype = C.c_int


@dataclass
class FieldSnapshot:
    """CPU numpy fields at one step (solid3d/model.py:1340-1375): nodal ``displacement``,
    ``fluctuation``, ``damage`` over the full nodes; element Gauss-point means ``strain``,
    ``stress``, ``history_mean``, ``psi_elastic`` (mean of g psi0), ``psi_driving`` (mean of Y)
    and ``history_max``.  The crack-closure fields are not produced (no closure here)."""
    step: int
    time_value: float
    macro_strain: np.ndarray
    macro_stress: np.ndarray
    displacement: np.ndarray
    fluctuation: np.ndarray
    damage: np.ndarray
    strain: np.ndarray
    stress: np.ndarray
    history_mean: np.ndarray
    history_max: np.ndarray
    psi_elastic: np.ndarray
    psi_driving: np.ndarray
    psi_closure: np.ndarray | None = None
    closure_active: np.ndarray | None = None
    det_F_min: np.ndarray | None = None


def _full_to_reduced(shape) -> np.ndarray:
    """Reduced node of every full node i + (nx+1)(j + (ny+1) k) of a triply periodic grid
    (solid3d/periodic.py:61-90)."""
    nx, ny, nz = shape
    k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    return ((i % nx) + nx * ((j % ny) + ny * (k % nz))).reshape(-1)


def _full_node_offsets(grid) -> np.ndarray:
    """Full node coordinates relative to the origin, (i hx, j hy, k hz) in full-node order."""
    nx, ny, nz = grid.shape
    k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    return np.stack([i.reshape(-1) * grid.h[0], j.reshape(-1) * grid.h[1], k.reshape(-1) * grid.h[2]], axis=1)


class PCGBreakdownError(RuntimeError):
    """Non-positive curvature, a non-positive diagonal or non-finite values (linalg.py:41)."""


@dataclass(frozen=True)
class PCGInfo:
    """Outcome of one PCG solve (solid3d/linalg.py:54-68)."""
    iterations: int
    converged: bool
    rhs_norm: float
    residual_norm: float
    relative_residual: float


# ----------------------------------------------------------------------------------------
# model: grid, phases, element material arrays
# ----------------------------------------------------------------------------------------

@dataclass(frozen=True)
class VoxelGrid:
    """Triply periodic voxel grid (solid3d/model.py:183-262)."""
    shape: tuple[int, int, int]
    h: tuple[float, float, float]

    @property
    def n_elements(self) -> int:
        return int(np.prod(self.shape))

    @property
    def lengths(self):
        return tuple(n * h for n, h in zip(self.shape, self.h))

    @property
    def volume(self) -> float:
        lx, ly, lz = self.lengths
        return lx * ly * lz

    @property
    def voxel_volume(self) -> float:
        return self.h[0] * self.h[1] * self.h[2]

    @property
    def array_shape(self):
        nx, ny, nz = self.shape
        return (nz, ny, nx)


@dataclass(frozen=True)
class PhaseMaterial:
    """One phase (solid3d/model.py:340-420 PhaseMaterial): ``kind`` 'isotropic', or
    'smeared_laminate' -- isotropic ``layers`` ``(fraction, E, nu)`` stacked along the voxel's
    tangent, with ``E``/``nu`` the isotropic equivalent that only the damage driver and the
    PF-CZM ``E_ref`` use.  ``tensile_strength`` f_t is needed by the PF-CZM law only."""
    name: str
    E: float
    nu: float
    Gc: float
    driver: str = "spectral"
    tensile_strength: float | None = None
    kind: str = "isotropic"
    layers: tuple = ()

    def __post_init__(self):
        if self.kind not in ("isotropic", "smeared_laminate"):
            raise ValueError(f"{self.name}: kind must be 'isotropic' or 'smeared_laminate', got {self.kind!r}")
        if self.kind == "smeared_laminate" and not self.layers:
            raise ValueError(f"{self.name}: smeared_laminate phases need layers")
        if self.kind == "isotropic" and self.layers:
            raise ValueError(f"{self.name}: layers apply to smeared_laminate phases only")

    @property
    def is_laminate(self) -> bool:
        return self.kind == "smeared_laminate"


def isotropic_stiffness_3d(E: float, nu: float) -> np.ndarray:
    """Engineering-Voigt isotropic stiffness (core/hex8.py:296-320)."""
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    C6 = np.zeros((6, 6))
    C6[:3, :3] = lam
    C6[0, 0] += 2.0 * mu
    C6[1, 1] += 2.0 * mu
    C6[2, 2] += 2.0 * mu
    C6[3, 3] = C6[4, 4] = C6[5, 5] = mu
    return C6


# Voigt order (xx, yy, zz, yz, xz, xy); engineering strains carry doubled shears.
_VOIGT_PAIRS = ((0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1))
LAMINATE_NORMAL = (0, 4, 5)     # stress components continuous across layers stacked along local axis 1
LAMINATE_IN_PLANE = (1, 2, 3)   # strain components continuous along the layers


def _eng_strain(T: np.ndarray) -> np.ndarray:
    """Engineering strain (..., 6) of symmetric tensors (..., 3, 3): shears doubled."""
    return np.stack([T[..., i, j] * (1.0 if i == j else 2.0) for i, j in _VOIGT_PAIRS], axis=-1)


def _eng_stress(T: np.ndarray) -> np.ndarray:
    """Voigt stress (..., 6) of symmetric tensors (..., 3, 3)."""
    return np.stack([T[..., i, j] for i, j in _VOIGT_PAIRS], axis=-1)


def _tensor_of(v: np.ndarray, shear: float) -> np.ndarray:
    """Symmetric tensor (..., 3, 3) of a Voigt vector (..., 6) whose shears are scaled by ``shear``."""
    T = np.zeros(v.shape[:-1] + (3, 3))
    for k, (i, j) in enumerate(_VOIGT_PAIRS):
        T[..., i, j] = v[..., k] * (1.0 if i == j else shear)
        T[..., j, i] = T[..., i, j]
    return T


def mandel(C: np.ndarray) -> np.ndarray:
    """Mandel form of engineering-Voigt stiffnesses (..., 6, 6): the shear rows and columns
    scaled by sqrt 2, so that the eigenvalues are those of the stiffness tensor."""
    w = np.array([1.0, 1.0, 1.0, math.sqrt(2.0), math.sqrt(2.0), math.sqrt(2.0)])
    return C * w[:, None] * w[None, :]


def laminate_stiffness(layers) -> np.ndarray:
    """Exact stiffness of isotropic layers ``(fraction, E, nu)`` stacked along local axis 1
    (solid3d/materials.py:99-144).  Normal components (stress xx, xz, xy) are uniform across
    the layers and in-plane components (strain yy, zz, yz) uniform along them:
    C_nn = <C_nn^-1>^-1, C_np = C_nn <C_nn^-1 C_np>, C_pn = <C_pn C_nn^-1> C_nn and
    C_pp = <C_pp - C_pn C_nn^-1 C_np> + <C_pn C_nn^-1> C_nn <C_nn^-1 C_np>."""
    rows = [tuple(float(v) for v in layer) for layer in layers]
    if not rows or any(len(r) != 3 for r in rows):
        raise ValueError("layers must be a non-empty sequence of (fraction, E, nu)")
    f = np.array([r[0] for r in rows])
    if np.any(f <= 0.0) or abs(float(f.sum()) - 1.0) > 1.0e-12:
        raise ValueError("layer fractions must be positive and sum to 1")
    Cl = np.stack([isotropic_stiffness_3d(E, nu) for _, E, nu in rows])
    n, pl = np.array(LAMINATE_NORMAL), np.array(LAMINATE_IN_PLANE)
    Cnn, Cnp = Cl[:, n][:, :, n], Cl[:, n][:, :, pl]
    Cpn, Cpp = Cl[:, pl][:, :, n], Cl[:, pl][:, :, pl]
    Cnn_inv = np.linalg.inv(Cnn)
    avg = lambda v: np.einsum("l,lij->ij", f, v)  # noqa: E731
    A = np.linalg.inv(avg(Cnn_inv))
    Bm = avg(Cnn_inv @ Cnp)
    Pm = avg(Cpn @ Cnn_inv)
    out = np.zeros((6, 6))
    out[np.ix_(n, n)] = A
    out[np.ix_(n, pl)] = A @ Bm
    out[np.ix_(pl, n)] = Pm @ A
    out[np.ix_(pl, pl)] = avg(Cpp - Cpn @ Cnn_inv @ Cnp) + Pm @ A @ Bm
    out = 0.5 * (out + out.T)
    if float(np.linalg.eigvalsh(mandel(out)).min()) <= 0.0:
        raise ValueError("laminate stiffness is not symmetric positive definite")
    return out


def frame_from_tangent(t) -> np.ndarray:
    """Right-handed frames (..., 3, 3) with columns e1, e2, e3 (solid3d/materials.py:175-203):
    e1 = t/|t| (a zero tangent counts as (1, 0, 0)); e2 the normalised cross product of the
    global axis least aligned with e1 (lowest index on ties) with e1; e3 = e1 x e2."""
    t = np.asarray(t, dtype=np.float64)
    if t.shape[-1] != 3:
        raise ValueError(f"tangent must end in 3 entries, got {t.shape}")
    norm = np.linalg.norm(t, axis=-1, keepdims=True)
    valid = norm > 0.0
    e1 = np.where(valid, t / np.where(valid, norm, 1.0), np.array([1.0, 0.0, 0.0]))
    axis = np.eye(3)[np.argmin(np.abs(e1), axis=-1)]
    e2 = np.cross(axis, e1)
    e2 = e2 / np.linalg.norm(e2, axis=-1, keepdims=True)
    e3 = np.cross(e1, e2)
    return np.stack((e1, e2, e3), axis=-1)


def rotate_stiffness(C_local: np.ndarray, R: np.ndarray) -> np.ndarray:
    """Engineering stiffnesses (n, 6, 6) in global axes of ``C_local`` (6, 6) or (n, 6, 6)
    given in frames ``R`` (n, 3, 3) whose columns are the material axes
    (solid3d/materials.py:211-245): C = T_s C_local T_e with T_e the map of global to local
    engineering strain (eps_l = R^T eps R) and T_s that of local to global stress
    (sigma = R sigma_l R^T)."""
    R = np.asarray(R, dtype=np.float64)
    if R.ndim != 3 or R.shape[1:] != (3, 3):
        raise ValueError(f"R must have shape (n, 3, 3), got {R.shape}")
    if R.shape[0] and float(np.abs(np.einsum("nji,njk->nik", R, R) - np.eye(3)).max()) > 1.0e-10:
        raise ValueError("every R must be orthogonal")
    I6 = np.eye(6)
    Te = _eng_strain(np.einsum("nji,cjk,nkl->ncil", R, _tensor_of(I6, 0.5), R)).transpose(0, 2, 1)
    Ts = _eng_stress(np.einsum("nij,cjk,nlk->ncil", R, _tensor_of(I6, 1.0), R)).transpose(0, 2, 1)
    return Ts @ C_local @ Te


@dataclass
class ElementMaterialArrays:
    """Per-voxel material data (solid3d/model.py:1137-1196; built as materials.py:474-548)."""
    stiffness_bank: np.ndarray   # (K,6,6)
    stiffness_index: np.ndarray  # (E,) int
    Gc: np.ndarray
    lam: np.ndarray
    mu: np.ndarray
    driver: np.ndarray           # (E,) int8
    phase: np.ndarray            # (E,) int
    E: np.ndarray | None = None     # (E,) Young's modulus of the voxel's phase (PF-CZM E_ref)
    f_t: np.ndarray | None = None   # (E,) tensile strength (NaN where the phase has none)
    tangent: np.ndarray | None = None  # (E,3) interface normal of each voxel (laminate frames)
    laminate: np.ndarray | None = None  # (E,) bool: voxel of a smeared-laminate phase

    @property
    def all_isotropic(self) -> bool:
        return self.laminate is None or not bool(np.any(self.laminate))


def build_element_arrays(phase_id: np.ndarray, phases: Mapping[str, PhaseMaterial],
                         energy_split_isotropic: str = "spectral",
                         Gc_field: np.ndarray | None = None, tangent: np.ndarray | None = None) -> ElementMaterialArrays:
    """Per-voxel material arrays (materials.py:474-548): a stiffness bank of the present
    isotropic phases in phase-id order, then one entry per smeared-laminate voxel in voxel
    order, its laminate stiffness rotated into the frame of its ``tangent`` (nz,ny,nx,3) or
    (E,3); lam/mu of every phase (the isotropic equivalent of a laminate).

    ``Gc_field`` (E,) optionally replaces the per-phase Gc (per-voxel fracture energy)."""
    phase_flat = np.asarray(phase_id).reshape(-1).astype(np.int64)
    n_el = phase_flat.shape[0]
    present = [n for n in PHASE_NAMES if PHASE_ID[n] in set(np.unique(phase_flat).tolist())]
    for n in present:
        if n not in phases:
            raise ValueError(f"no material for phase {n!r}")
    index = np.empty(n_el, dtype=np.int64)
    bank = []
    n_ids = len(PHASE_NAMES)
    Gc_id, lam_id, mu_id = np.zeros(n_ids), np.zeros(n_ids), np.zeros(n_ids)
    E_id, ft_id = np.zeros(n_ids), np.full(n_ids, np.nan)
    drv_id = np.zeros(n_ids, dtype=np.int8)
    iso_names = [n for n in present if not phases[n].is_laminate]
    lam_names = [n for n in present if phases[n].is_laminate]
    for slot, name in enumerate(iso_names):
        m = phases[name]
        bank.append(isotropic_stiffness_3d(m.E, m.nu)[None])
        index[phase_flat == PHASE_ID[name]] = slot
    for name in present:
        m = phases[name]
        pid = PHASE_ID[name]
        Gc_id[pid] = m.Gc
        lam_id[pid] = m.E * m.nu / ((1.0 + m.nu) * (1.0 - 2.0 * m.nu))
        mu_id[pid] = m.E / (2.0 * (1.0 + m.nu))
        E_id[pid] = m.E
        if m.tensile_strength is not None:
            ft_id[pid] = float(m.tensile_strength)
        if name == "pore" or m.driver == "none":  # materials.py:450-471
            drv_id[pid] = DRIVER_NONE
        else:
            drv_id[pid] = DRIVER_SPECTRAL if energy_split_isotropic == "spectral" else DRIVER_FULL
    tan = None
    laminate = np.zeros(n_el, dtype=bool)
    if tangent is not None:
        tan = np.asarray(tangent, dtype=np.float64).reshape(-1, 3)
        if tan.shape[0] != n_el:
            raise ValueError(f"tangent must hold one vector per voxel ({n_el}), got {tan.shape[0]}")
    if lam_names:
        if tan is None:
            raise ValueError("smeared_laminate phases need the voxel tangent field")
        ids = np.array([PHASE_ID[n] for n in lam_names])
        framed = np.nonzero(np.isin(phase_flat, ids))[0]
        local = np.stack([laminate_stiffness(phases[n].layers) for n in lam_names])
        slot_of = np.full(n_ids, -1, dtype=np.int64)
        slot_of[ids] = np.arange(len(lam_names))
        bank.append(rotate_stiffness(local[slot_of[phase_flat[framed]]], frame_from_tangent(tan[framed])))
        index[framed] = len(iso_names) + np.arange(framed.shape[0])
        laminate[framed] = True
    Gc = Gc_id[phase_flat] if Gc_field is None else np.asarray(Gc_field, dtype=np.float64).reshape(-1)
    return ElementMaterialArrays(np.concatenate(bank, axis=0), index, Gc, lam_id[phase_flat], mu_id[phase_flat],
                                 drv_id[phase_flat], phase_flat, E_id[phase_flat], ft_id[phase_flat], tan,
                                 laminate)


# -- crack closure rows of the two groups (solid3d/splits.py:424-742) --------------------
CLOSURE_MODULI_TOLERANCE = 1.0e-10      # splits.py:411
CLOSURE_STIFFENING_TOLERANCE = 1.0e-12  # splits.py:414
INTERFACE_DIAGONAL_TOLERANCE = 1.0e-10  # splits.py:183
CLOSURE_ISO_ROWS = np.array([[1.0, 1.0, 1.0, 0.0, 0.0, 0.0],
                             [1.0, 0.0, 0.0, 0.0, 0.0, 0.0],
                             [0.0, 1.0, 0.0, 0.0, 0.0, 0.0],
                             [0.0, 0.0, 1.0, 0.0, 0.0, 0.0]])  # splits.py:489-495: trace, e_x, e_y, e_z


def _strain_of_dyad(a: np.ndarray) -> np.ndarray:
    """Engineering strain of a (x) a: [a1^2, a2^2, a3^2, 2 a2 a3, 2 a1 a3, 2 a1 a2]."""
    a1, a2, a3 = a[..., 0], a[..., 1], a[..., 2]
    return np.stack((a1 * a1, a2 * a2, a3 * a3, 2.0 * a2 * a3, 2.0 * a1 * a3, 2.0 * a1 * a2), axis=-1)


def _normal_strain_vector(a: np.ndarray) -> np.ndarray:
    """q with q . eps = a.eps.a for engineering eps: [a1^2, a2^2, a3^2, a2 a3, a1 a3, a1 a2]
    (splits.py:547-559)."""
    a1, a2, a3 = a[..., 0], a[..., 1], a[..., 2]
    return np.stack((a1 * a1, a2 * a2, a3 * a3, a2 * a3, a1 * a3, a1 * a2), axis=-1)


def _sym_outer(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return 0.5 * (a[..., :, None] * b[..., None, :] + b[..., :, None] * a[..., None, :])


def interface_traction_rows(tangent: np.ndarray) -> np.ndarray:
    """Rows m_k (..., 3, 6) with t_k = m_k . sigma for the local traction components
    (11, 13, 12) of the frame of ``tangent`` (splits.py:190-209): the engineering strain of
    sym(a n^T) for a = n, e3, e2."""
    F = frame_from_tangent(tangent)
    n, e2, e3 = F[..., :, 0], F[..., :, 1], F[..., :, 2]
    return np.stack([_eng_strain(_sym_outer(a, n)) for a in (n, e3, e2)], axis=-2)


def band_closure_rows(C: np.ndarray, tangent: np.ndarray):
    """Closure rows (nb, 4, 6) and moduli (nb, 4) of smeared-laminate voxels
    (operator.py:669-687; splits.py:212-258, 603-670): rows (C m1, q2 + q3, q2, q3) with m1
    the normal-traction row and q_k the normal-strain vectors of the in-plane axes e2, e3;
    moduli (1/a11, lam_S, 2 mu_S, 2 mu_S) with a11 the normal entry of the (diagonal)
    normal block M C M^T and (lam_S, mu_S) = (S23, (S22 - S23)/2) of the in-plane Schur
    block S = C_pp - C_pn C_nn^-1 C_np, which must be two-dimensionally isotropic."""
    F = frame_from_tangent(tangent)
    M = interface_traction_rows(tangent)
    Cnn = M @ C @ M.transpose(0, 2, 1)
    diag = np.diagonal(Cnn, axis1=-2, axis2=-1)
    if diag.size:
        if not np.all(diag > 0.0):
            raise ValueError("the interface normal block C_nn must have a positive diagonal")
        off = np.abs(Cnn - diag[..., :, None] * np.eye(3)).max(axis=(-2, -1)) / np.abs(diag).max(axis=-1)
        if not float(off.max()) <= INTERFACE_DIAGONAL_TOLERANCE:
            raise ValueError(f"crack_closure 'axes' needs a band frame with a diagonal C_nn (relative "
                             f"off-diagonal {float(off.max()):.3e})")
    e2, e3 = F[..., :, 1], F[..., :, 2]
    P = np.stack((_strain_of_dyad(e2), _strain_of_dyad(e3), _eng_strain(_sym_outer(e2, e3))), axis=-2)
    Cnp = M @ C @ P.transpose(0, 2, 1)
    Cpp = P @ C @ P.transpose(0, 2, 1)
    S = Cpp - Cnp.transpose(0, 2, 1) @ np.linalg.solve(Cnn, Cnp)
    s22, s33, s23 = S[:, 0, 0], S[:, 1, 1], S[:, 0, 1]
    if s22.size:
        scale = np.maximum(np.abs(s22), np.abs(s33))
        defect = float((np.abs(s22 - s33) / scale).max())
        coupling = float((np.maximum(np.abs(S[:, 0, 2]), np.abs(S[:, 1, 2])) / scale).max())
        if not (defect <= CLOSURE_MODULI_TOLERANCE and coupling <= CLOSURE_MODULI_TOLERANCE):
            raise ValueError("crack_closure 'axes' needs a band whose in-plane Schur block is two-dimensionally "
                             f"isotropic (relative S22 - S33 {defect:.3e}, normal-shear coupling {coupling:.3e})")
        if np.any(s23 < 0.0):
            raise ValueError("crack_closure 'axes' needs a band Schur block with S23 >= 0 (lam_S >= 0)")
    lam_S, mu_S = s23, 0.5 * (s22 - s23)
    q2, q3 = _normal_strain_vector(e2), _normal_strain_vector(e3)
    Cm = np.einsum("nij,nj->ni", C, M[:, 0, :])
    rows = np.stack((Cm, q2 + q3, q2, q3), axis=-2)
    moduli = np.stack((1.0 / diag[:, 0], lam_S, 2.0 * mu_S, 2.0 * mu_S), axis=-1)
    return rows, moduli


def closure_stiffening_margin(C: np.ndarray, rows: np.ndarray, moduli: np.ndarray) -> np.ndarray:
    """lambda_min(C - D) / max C_ii in Mandel form, D = sum_k k_k v_k v_k^T with every bit
    active (splits.py:456-486)."""
    D = np.einsum("nki,nk,nkj->nij", rows, moduli, rows)
    diff = mandel(C - D)
    diff = 0.5 * (diff + diff.transpose(0, 2, 1))
    return np.linalg.eigvalsh(diff)[:, 0] / np.diagonal(C, axis1=-2, axis2=-1).max(axis=-1)


@dataclass(frozen=True)
class LoadCase:
    """Mixed macro-strain control (solid3d/model.py:589-671)."""
    name: str
    controlled: int
    targets: tuple
    free: tuple
    fixed_zero: tuple

    @classmethod
    def from_schedule(cls, name, *, controlled, schedule, free, fixed_zero):
        targets, start = [], 0.0
        for upto, inc in schedule:
            n_steps = int(round(abs(upto - start) / abs(inc)))
            span = upto - start
            for step in range(1, n_steps):
                targets.append(start + span * (step / n_steps))
            targets.append(float(upto))
            start = float(upto)
        return cls(name, int(controlled), tuple(targets), tuple(free), tuple(fixed_zero))


def choose_pinned_node(shape, phase_flat: np.ndarray, pore_id: int | None = PHASE_ID["pore"]) -> int:
    """First reduced node none of whose 8 voxels is a pore (solid3d/periodic.py:224-249)."""
    if pore_id is None:
        return 0
    nx, ny, nz = shape
    pore = (np.asarray(phase_flat).reshape(nz, ny, nx) == pore_id).astype(np.int64)
    count = np.zeros_like(pore)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):  # node (i,j,k) touches voxels (i-1+dx, j-1+dy, k-1+dz)
                count += np.roll(pore, shift=(1 - dz, 1 - dy, 1 - dx), axis=(0, 1, 2))
    return int(np.argmin(count.reshape(-1)))


# ----------------------------------------------------------------------------------------
# the operator
# ----------------------------------------------------------------------------------------

_PATTERN_SHIFTS = torch.arange(32, dtype=torch.int64).view(8, 4)   # bit 4 q + k


@dataclass(eq=False)
class ClosurePattern:
    """Sign pattern of the crack closure (solid3d/model.py:1200-1223): ``iso`` (E, 8, 4) bool
    holds the bits ``[v_k . eps < 0]`` of each voxel's four closure rows per Gauss point
    (trace, eps_xx, eps_yy, eps_zz in the isotropic group; t11, p2 + p3, p2, p3 of the band
    frame for a laminate band voxel, whose bits the reference keeps in a separate ``band``
    array).  Kept packed on the device, one 32-bit word per voxel (bit 4 q + k), which is the
    form the kernels read; ``iso`` unpacks it."""
    bits: torch.Tensor            # (E,) int32
    band: None = None

    @classmethod
    def from_iso(cls, iso: torch.Tensor) -> "ClosurePattern":
        if not isinstance(iso, torch.Tensor) or iso.dtype != torch.bool or iso.dim() != 3 or iso.shape[1:] != (8, 4):
            raise TypeError("iso must be a torch.bool tensor of shape (E, 8, 4)")
        words = (iso.to(torch.int64) << _PATTERN_SHIFTS.to(iso.device)).sum(dim=(1, 2))
        words = torch.where(words >= 2 ** 31, words - 2 ** 32, words)
        return cls(words.to(torch.int32).contiguous())

    @property
    def iso(self) -> torch.Tensor:
        words = self.bits.to(torch.int64) & 0xFFFFFFFF
        return ((words.view(-1, 1, 1) >> _PATTERN_SHIFTS.to(words.device)) & 1).to(torch.bool)

    def groups(self):
        return (self.iso, None)

    def equals(self, other: "ClosurePattern") -> bool:
        return torch.equal(self.bits, other.bits)


@dataclass(frozen=True)
class H8ScalarSPDSystem:
    """Matrix-free SPD data of one scalar subproblem (box_kkt_h8.py:63-122).  ``_form`` 0: the
    AT2 damage system with history ``_H``; 1: the PF-CZM Newton tangent with the curvatures
    ``_H``."""
    action: Callable[[torch.Tensor], torch.Tensor]
    rhs: torch.Tensor
    diagonal: torch.Tensor
    _op: object = None
    _H: torch.Tensor | None = None
    _scale: float = 1.0
    _form: int = 0

    def __post_init__(self):
        if not callable(self.action):
            raise TypeError("action must be callable")
        if self.rhs.ndim != 1 or self.rhs.numel() == 0:
            raise ValueError("rhs must be a non-empty one-dimensional tensor")
        if self.diagonal.shape != self.rhs.shape:
            raise ValueError("diagonal must be strictly positive")
        finite, positive = (bool(v) for v in torch.stack((torch.isfinite(self.rhs).all(),
                                                          (self.diagonal > 0.0).all())).cpu())
        if not finite:
            raise ValueError("rhs contains a non-finite value")
        if not positive:
            raise ValueError("diagonal must be strictly positive")

    @property
    def n_dofs(self) -> int:
        return self.rhs.numel()

    def apply(self, field: torch.Tensor) -> torch.Tensor:
        return self.action(field)


PFCZM_C_ALPHA = math.pi          # physics/pfczm.py:6
PFCZM_TANGENT_FLOOR = 0.25       # solid3d/operator.py:124


def pfczm_softening_parameters(softening: str):
    """``(a2, a3)`` of Wu's softening laws (physics/pfczm.py:55-63)."""
    if softening == "linear":
        return -0.5, 0.0
    if softening == "exponential":
        return 2.0 ** (5.0 / 3.0) - 3.0, 0.0
    raise ValueError(f"PF-CZM softening must be one of ('linear', 'exponential'), got {softening!r}")


def default_tiles(shape) -> tuple[int, int, int]:
    """Launch shape of the voxel kernels when the caller names none: 16x14 node tiles in z
    slabs of 8 layers when that already gives 1024 blocks (four per CU of a 256-CU device),
    else slabs of 3 layers, the measured best of the mechanics and damage applies at 48^3 and
    64^3 (thinner slabs add launch-wide parallelism but evaluate more halo and seam layers).
    Any kz gives the same nodal rows, bit for bit."""
    nx, ny, nz = (int(v) for v in shape)
    tiles_xy = -(-nx // 16) * -(-ny // 14)
    return 16, 14, max(1, min(nz, 8 if tiles_xy * nz >= 1024 * 8 else 3))


class PeriodicVoxelPhaseFieldOperator:
    """Matrix-free mechanics and damage operators of one triply periodic voxel cell.

    Mirrors solid3d/operator.py:133-1544: the AT2 law or (``damage_law='pfczm'``) Wu's PF-CZM
    law, with or without the crack closure (``crack_closure='axes'``: the isotropic phases on the
    cell axes, smeared-laminate voxels on their band frame); no interface
    drivers.  x = [u~ (3 n), E (6)] with node-major u~ (``3 node + c``).
    """

    def __init__(self, grid: VoxelGrid, materials: ElementMaterialArrays, load: LoadCase, *, kappa: float = 1e-6,
                 ell: float, pinned_node: int | None = None, device=None, tiles=None,
                 damage_law: str = "at2", pfczm_softening: str = "linear", pfczm_p: float = 2,
                 crack_closure: str = "none"):
        self.device = _lib.require_device() if device is None else torch.device(device)
        self.grid, self.mat, self.load = grid, materials, load
        self.dtype = F64
        nx, ny, nz = grid.shape
        self.n_elements = grid.n_elements
        self.n_red = self.n_elements  # triply periodic: one reduced node per voxel
        self.n_mech = 3 * self.n_red + 6
        self.volume = float(grid.volume)
        self.ell, self.kappa = float(ell), float(kappa)
        Gc = np.asarray(materials.Gc, dtype=np.float64)
        if not np.all(np.isfinite(Gc)) or not np.all(Gc > 0):
            raise ValueError("Gc must be finite and positive at every element")
        self.damage_scale = self.ell / (float(Gc.min()) * float(grid.voxel_volume))  # operator.py:238
        self._c_history = 2.0 * (1.0 - self.kappa)
        if pinned_node is None:
            pinned_node = choose_pinned_node(grid.shape, materials.phase)
        if not 0 <= int(pinned_node) < self.n_red:
            raise ValueError(f"pinned_node must be a reduced node in [0, {self.n_red}), got {pinned_node!r}")
        self.pinned_node = int(pinned_node)
        bank = np.ascontiguousarray(materials.stiffness_bank, dtype=np.float64)
        idx = np.ascontiguousarray(materials.stiffness_index, dtype=np.int32)
        lam = np.ascontiguousarray(materials.lam, dtype=np.float64)
        mu = np.ascontiguousarray(materials.mu, dtype=np.float64)
        drv = np.ascontiguousarray(materials.driver, dtype=np.int8)
        Gc = np.ascontiguousarray(Gc)
        h = C.c_void_p()
        check(_L.phast_hex8_plan_create(nx, ny, nz, grid.h[0], grid.h[1], grid.h[2], bank.ctypes.data,
                                        bank.shape[0], idx.ctypes.data, Gc.ctypes.data, lam.ctypes.data,
                                        mu.ctypes.data, drv.ctypes.data, self.kappa, self.ell, C.byref(h)),
              "phast_hex8_plan_create")
        self._plan = h
        check(_L.phast_hex8_solver_reserve(h), "phast_hex8_solver_reserve")  # solves then allocate nothing
        self.set_tiles(*(default_tiles(grid.shape) if tiles is None else tiles))
        self.Gc = torch.as_tensor(Gc, device=self.device)
        if damage_law not in ("at2", "pfczm"):
            raise ValueError(f"damage_law must be 'at2' or 'pfczm', got {damage_law!r}")
        if crack_closure not in ("none", "axes"):
            raise ValueError(f"crack_closure must be 'none' or 'axes', got {crack_closure!r}")
        self.damage_law = damage_law
        self._pfczm = damage_law == "pfczm"
        self._closure = crack_closure == "axes"
        self.n_closure_band = 0
        self.band_closure = None   # (voxel ids, rows (nb,4,6), moduli (nb,4)) of the laminate band
        self.a1 = None
        self.pfczm_bound = None
        if self._pfczm:
            self._build_pfczm(materials, pfczm_softening, pfczm_p, drv)
        if self._closure:
            self._build_closure(materials, drv, lam)

    # -- PF-CZM law and crack closure set-up (operator.py:300-383, 631-690) --------------
    def _build_pfczm(self, materials, softening, p, drv):
        """Per-voxel a1 = 4 E Gc/(pi ell f_t^2 (1 - kappa)) (physics/pfczm.py:17-32) and the
        onset check a1 >= p + a2 on voxels with a damage driver (operator.py:300-383)."""
        a2, a3 = pfczm_softening_parameters(softening)
        if float(p) < 2.0:
            raise ValueError(f"pfczm_p must be >= 2, got {p!r}")
        self._pfczm_a2, self._pfczm_a3, self._pfczm_p = a2, a3, p
        if materials.E is None or materials.f_t is None:
            raise ValueError("damage_law 'pfczm' needs the element E and tensile strength arrays")
        f_t = np.asarray(materials.f_t, dtype=np.float64)
        E_ref = np.asarray(materials.E, dtype=np.float64)
        missing = sorted({PHASE_NAMES[int(i)] for i in np.unique(np.asarray(materials.phase)[~np.isfinite(f_t)])})
        if missing:
            raise ValueError("damage_law 'pfczm' needs a tensile strength for every phase of the cell "
                             f"(material.tensile_strength); missing: {', '.join(missing)}")
        if not np.all(f_t > 0.0):
            raise ValueError("PFCZM a1 requires f_t > 0")
        Gc = np.asarray(materials.Gc, dtype=np.float64)
        a1 = 4.0 * E_ref * Gc / (math.pi * self.ell * f_t ** 2 * max(1.0 - self.kappa, 1.0e-30))
        bound = float(p) + a2
        self.pfczm_bound = bound
        active = drv != DRIVER_NONE
        if active.any() and float(a1[active].min()) < bound:
            raise ValueError(f"damage_law 'pfczm': the onset condition a1 >= p + a2 = {bound:g} (g''(0) >= 0) fails "
                             f"at ell = {self.ell:g} mm: smallest a1 {float(a1[active].min()):.4g}")
        a1 = np.ascontiguousarray(a1)
        check(_L.phast_hex8_pfczm_setup(self._plan, a1.ctypes.data, a2, a3, float(p)), "pfczm_setup")
        self.a1 = torch.as_tensor(a1, device=self.device)

    def _build_closure(self, materials, drv, lam):
        """The closure groups (operator.py:631-690): the isotropic phases' voxels with a damage
        driver (cell axes, lam >= 0 there) and the smeared-laminate voxels with a driver (the
        band frame: rows and moduli of ``band_closure_rows``); neither group may stiffen a
        voxel beyond its intact stiffness."""
        on = drv != DRIVER_NONE
        laminate = np.zeros_like(on) if materials.laminate is None else np.asarray(materials.laminate, dtype=bool)
        iso = on & ~laminate
        band = on & laminate
        if iso.any() and float(lam[iso].min()) < 0.0:
            raise ValueError("crack_closure 'axes' needs lam >= 0 in every isotropic phase")
        self.cl_on = torch.as_tensor(on, device=self.device)
        self.n_closure_band = int(band.sum())
        if materials.all_isotropic:
            group = np.ascontiguousarray(iso.astype(np.uint8))
            check(_L.phast_hex8_closure_setup(self._plan, group.ctypes.data), "closure_setup")
            return
        bank = np.asarray(materials.stiffness_bank, dtype=np.float64)
        index = np.asarray(materials.stiffness_index, dtype=np.int64)
        mu = np.asarray(materials.mu, dtype=np.float64)
        tab = np.zeros((bank.shape[0], 28))
        entries = np.unique(index[iso])
        for slot in entries:  # isotropic phases: one bank entry each, shared lam and mu
            members = iso & (index == slot)
            lam_s, mu_s = lam[members], mu[members]
            if np.ptp(lam_s) > 0.0 or np.ptp(mu_s) > 0.0:
                raise ValueError("isotropic elements sharing a stiffness must share lam and mu")
            tab[slot, :24] = CLOSURE_ISO_ROWS.reshape(-1)
            tab[slot, 24:] = (lam_s[0], 2.0 * mu_s[0], 2.0 * mu_s[0], 2.0 * mu_s[0])
        if entries.size:
            margin = closure_stiffening_margin(bank[entries], tab[entries, :24].reshape(-1, 4, 6), tab[entries, 24:])
            if float(margin.min()) < -CLOSURE_STIFFENING_TOLERANCE:
                raise ValueError(f"the isotropic closure stiffens a phase (margin {float(margin.min()):.3e})")
        if band.any():
            if materials.tangent is None:
                raise ValueError("the band closure needs the voxel tangent field")
            ids = np.nonzero(band)[0]
            slots = index[ids]
            if np.unique(slots).size != slots.size:
                raise ValueError("every band voxel needs its own bank entry")
            rows, moduli = band_closure_rows(bank[slots], np.asarray(materials.tangent)[ids])
            margin = closure_stiffening_margin(bank[slots], rows, moduli)
            if float(margin.min()) < -CLOSURE_STIFFENING_TOLERANCE:
                raise ValueError("the band closure stiffens an element beyond its intact stiffness "
                                 f"(lambda_min(C - D) / max C_ii = {float(margin.min()):.3e})")
            tab[slots, :24] = rows.reshape(-1, 24)
            tab[slots, 24:] = moduli
            self.band_closure = (ids, rows, moduli)
        group = np.ascontiguousarray((iso.astype(np.uint8) + 2 * band.astype(np.uint8)))
        tab = np.ascontiguousarray(tab)
        check(_L.phast_hex8_closure_setup_rows(self._plan, group.ctypes.data, tab.ctypes.data), "closure_setup_rows")

    @property
    def is_pfczm(self) -> bool:
        return self._pfczm

    @property
    def has_closure(self) -> bool:
        return self._closure

    @property
    def has_interface_stress(self) -> bool:
        return False

    def __del__(self):
        h = getattr(self, "_plan", None)
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _L.phast_hex8_plan_destroy(h)
            self._plan = None

    def set_tiles(self, tx: int, ty: int, kz: int) -> None:
        """Launch shape of the apply: (16,16) / (32,8) voxel-parallel, (16,8) / (8,8) marching."""
        check(_L.phast_hex8_set_tiles(self._plan, int(tx), int(ty), int(kz)), "set_tiles")
        self.tiles = (int(tx), int(ty), int(kz))

    # -- checks --------------------------------------------------------------------------
    def _check(self, name, value, shape):
        if not isinstance(value, torch.Tensor):
            raise TypeError(f"{name} must be a torch.Tensor")
        if tuple(value.shape) != tuple(shape):
            raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(value.shape)}")
        if value.dtype != F64 or value.device != self.device:
            raise ValueError(f"{name} must be {F64} on {self.device}")
        return value.contiguous()

    def _check_qp(self, name, value):
        return None if value is None else self._check(name, value, (self.n_elements, 8))

    # -- vectors -------------------------------------------------------------------------
    def split_mech(self, x):
        x = self._check("x", x, (self.n_mech,))
        n = 3 * self.n_red
        return x[:n].reshape(self.n_red, 3), x[n:]

    def pack_mech(self, u, E):
        u = torch.as_tensor(u, dtype=F64, device=self.device).reshape(-1)
        E = torch.as_tensor(E, dtype=F64, device=self.device).reshape(-1)
        if u.shape != (3 * self.n_red,) or E.shape != (6,):
            raise ValueError(f"u must hold {3 * self.n_red} entries and E 6, got {u.numel()} and {E.numel()}")
        return torch.cat((u, E))

    def mech_fixed_mask(self) -> torch.Tensor:
        """Pinned-node dofs, the controlled and fixed_zero E components (operator.py:914-925)."""
        mask = torch.zeros(self.n_mech, dtype=torch.bool, device=self.device)
        b = 3 * self.pinned_node
        mask[b:b + 3] = True
        off = 3 * self.n_red
        mask[off + self.load.controlled] = True
        for c in self.load.fixed_zero:
            mask[off + c] = True
        return mask

    def mech_fixed_values(self, e_c) -> torch.Tensor:
        values = torch.zeros(self.n_mech, dtype=F64, device=self.device)
        values[3 * self.n_red + self.load.controlled] = float(e_c)
        return values

    # -- mechanics -----------------------------------------------------------------------
    def degradation(self, d: torch.Tensor) -> torch.Tensor:
        """g (E,8) at d_q = d_e N^T (operator.py:939-946): (1-kappa)(1-d_q)^2 + kappa, or Wu's
        rational PF-CZM degradation with the voxel's a1."""
        d = self._check("d", d, (self.n_red,))
        g = torch.empty((self.n_elements, 8), dtype=F64, device=self.device)
        if self._pfczm:
            check(_L.phast_hex8_pfczm_degradation(self._plan, ptr(d), ptr(g), None, None, stream()), "degradation")
        else:
            check(_L.phast_hex8_degradation(self._plan, ptr(d), ptr(g), stream()), "degradation")
        return g

    def _check_closure(self, closure):
        """operator.py:735-753."""
        if closure is None:
            return None
        if not self._closure:
            raise ValueError("closure applies with crack_closure 'axes' only; this operator has none")
        if not isinstance(closure, ClosurePattern):
            raise TypeError("closure must be a ClosurePattern")
        b = closure.bits
        if b.dtype != torch.int32 or b.shape != (self.n_elements,) or b.device != self.device:
            raise ValueError(f"closure must hold ({self.n_elements},) int32 words on {self.device}")
        return closure

    def _closure_mode(self, g, closure):
        """0: no add-back (no closure, or undamaged g); 1: own pattern; 2: `closure`."""
        closure = self._check_closure(closure)
        if not self._closure or g is None:
            return 0, None
        return (1, None) if closure is None else (2, closure.bits.contiguous())

    def mech_action(self, x, g=None, *, opening=None, closure=None) -> torch.Tensor:
        """A(g) x (operator.py:1171-1194); g=None means undamaged.  With the crack closure the
        action is that of the pattern ``closure`` (linear), or the internal force at each
        point's own pattern when it is None."""
        if opening is not None:
            raise ValueError("opening applies to the variational band stress only; this cell has none")
        x = self._check("x", x, (self.n_mech,))
        g = self._check_qp("g", g)
        mode, bits = self._closure_mode(g, closure)
        y = torch.empty_like(x)
        if mode:
            check(_L.phast_hex8_mech_action_cl(self._plan, ptr(x), ptr(g), ptr(bits), mode, ptr(y), stream()),
                  "mech_action")
        else:
            check(_L.phast_hex8_mech_action(self._plan, ptr(x), ptr(g), ptr(y), stream()), "mech_action")
        return y

    def mech_action_d(self, x, d) -> torch.Tensor:
        """A(g(d)) x; for the AT2 law without closure g is formed in-kernel from nodal damage
        (fused degradation)."""
        x = self._check("x", x, (self.n_mech,))
        d = self._check("d", d, (self.n_red,))
        if self._pfczm or self._closure:
            return self.mech_action(x, self.degradation(d))
        y = torch.empty_like(x)
        check(_L.phast_hex8_mech_action_d(self._plan, ptr(x), ptr(d), ptr(y), stream()), "mech_action_d")
        return y

    def mech_diagonal(self, g=None, *, opening=None, closure=None) -> torch.Tensor:
        """Jacobi diagonal (operator.py:1196-1300); with the crack closure every closure bit
        (all of them when ``closure`` is None) adds its (1 - g) term."""
        if opening is not None:
            raise ValueError("opening applies to the variational band stress only; this cell has none")
        g = self._check_qp("g", g)
        mode, bits = self._closure_mode(g, closure)
        out = torch.empty(self.n_mech, dtype=F64, device=self.device)
        if mode:
            check(_L.phast_hex8_mech_diagonal_cl(self._plan, ptr(g), ptr(bits), mode, ptr(out), stream()),
                  "mech_diagonal")
        else:
            check(_L.phast_hex8_mech_diagonal(self._plan, ptr(g), ptr(out), stream()), "mech_diagonal")
        return out

    def _stress_sums(self, x, g, closure=None):
        x = self._check("x", x, (self.n_mech,))
        g = self._check_qp("g", g)
        mode, bits = self._closure_mode(g, closure)
        out = torch.empty(7, dtype=F64, device=self.device)
        if mode:
            check(_L.phast_hex8_stress_sums_cl(self._plan, ptr(x), ptr(g), ptr(bits), mode, ptr(out), stream()),
                  "stress_sums")
        else:
            check(_L.phast_hex8_stress_sums(self._plan, ptr(x), ptr(g), ptr(out), stream()), "stress_sums")
        return out

    def macro_stress(self, x, g=None, *, opening=None, closure=None) -> torch.Tensor:
        """Sigma = r_E / V (operator.py:1303-1316)."""
        return self._stress_sums(x, g, closure)[:6] / self.volume

    def elastic_energy(self, x, g=None, *, opening=None, closure=None) -> torch.Tensor:
        """0.5 x.A(g) x (operator.py:1318-1340); with the closure the degraded energy at the
        pattern of x (or ``closure``)."""
        return self._stress_sums(x, g, closure)[6]

    def hill_mandel(self, x, g=None, *, closure=None) -> dict:
        """<sigma.eps> against Sigma.E (operator.py:1390-1413)."""
        sums = self._stress_sums(x, g, closure)
        E = self._check("x", x, (self.n_mech,))[3 * self.n_red:]
        mean_work = float(2.0 * sums[6] / self.volume)
        macro_work = float(torch.dot(sums[:6] / self.volume, E))
        scale = max(abs(mean_work), abs(macro_work))
        return {"mean_work": mean_work, "macro_work": macro_work,
                "relative": abs(mean_work - macro_work) / scale if scale > 0.0 else 0.0}

    # -- crack closure patterns (operator.py:786-841) --------------------------------------
    def closure_pattern(self, x) -> ClosurePattern:
        """Sign pattern v . eps < 0 of every closure term at state x (operator.py:792-806)."""
        if not self._closure:
            raise ValueError("closure_pattern needs crack_closure 'axes'")
        x = self._check("x", x, (self.n_mech,))
        bits = torch.empty(self.n_elements, dtype=torch.int32, device=self.device)
        check(_L.phast_hex8_closure_pattern(self._plan, ptr(x), None, None, ptr(bits), None, stream()),
              "closure_pattern")
        return ClosurePattern(bits)

    def closure_pattern_change(self, x, pattern: ClosurePattern, g, *, tolerance: float = 1.0e-8):
        """New pattern of x and whether a flipped bit moves the stress by more than
        ``tolerance`` times the group's largest k |v . eps| (operator.py:808-841)."""
        pattern = self._check_closure(pattern)
        if pattern is None:
            raise ValueError("closure_pattern_change needs a pattern")
        x = self._check("x", x, (self.n_mech,))
        g = self._check_qp("g", g)
        if g is None:
            raise ValueError("closure_pattern_change needs g")
        bits = torch.empty(self.n_elements, dtype=torch.int32, device=self.device)
        out = torch.empty(4 if self.n_closure_band else 2, dtype=F64, device=self.device)
        check(_L.phast_hex8_closure_pattern(self._plan, ptr(x), ptr(g), ptr(pattern.bits.contiguous()), ptr(bits),
                                            ptr(out), stream()), "closure_pattern_change")
        vals = [float(v) for v in out.cpu()]
        changed = any(jump >= 0.0 and jump > tolerance * strength for strength, jump in zip(vals[0::2], vals[1::2]))
        return ClosurePattern(bits), changed

    def driving_force(self, x) -> torch.Tensor:
        """Undamaged driving energy Y (E,8) (operator.py:1415-1444)."""
        x = self._check("x", x, (self.n_mech,))
        Y = torch.empty((self.n_elements, 8), dtype=F64, device=self.device)
        check(_L.phast_hex8_driving_force(self._plan, ptr(x), ptr(Y), stream()), "driving_force")
        return Y

    # -- damage --------------------------------------------------------------------------
    def damage_action(self, d, H=None) -> torch.Tensor:
        """Unscaled AT2 damage action (operator.py:1450-1466)."""
        return self._damage_action(d, H, 1.0)

    def _damage_action(self, d, H, scale):
        d = self._check("d", d, (self.n_red,))
        H = self._check_qp("H", H)
        y = torch.empty_like(d)
        check(_L.phast_hex8_damage_action(self._plan, ptr(d), ptr(H), float(scale), ptr(y), stream()),
              "damage_action")
        return y

    def damage_rhs(self, H, scale: float = 1.0) -> torch.Tensor:
        """2(1-kappa) sum_q w H_q N_q (operator.py:1468-1480)."""
        H = self._check_qp("H", H)
        out = torch.empty(self.n_red, dtype=F64, device=self.device)
        check(_L.phast_hex8_damage_rhs(self._plan, ptr(H), float(scale), ptr(out), stream()), "damage_rhs")
        return out

    def damage_diagonal(self, H=None, scale: float = 1.0) -> torch.Tensor:
        """Jacobi diagonal of the damage operator (operator.py:1482-1494)."""
        H = self._check_qp("H", H)
        out = torch.empty(self.n_red, dtype=F64, device=self.device)
        check(_L.phast_hex8_damage_diagonal(self._plan, ptr(H), float(scale), ptr(out), stream()),
              "damage_diagonal")
        return out

    def damage_system(self, H, *, scaled: bool = True) -> H8ScalarSPDSystem:
        """Scaled SPD system (operator.py:1496-1509)."""
        H = self._check_qp("H", H)
        scale = self.damage_scale if scaled else 1.0
        return H8ScalarSPDSystem(action=lambda f: self._damage_action(f, H, scale),
                                 rhs=self.damage_rhs(H, scale), diagonal=self.damage_diagonal(H, scale),
                                 _op=self, _H=H, _scale=scale)

    # -- PF-CZM damage (operator.py:461-631) ----------------------------------------------
    def _require_pfczm(self):
        if not self._pfczm:
            raise ValueError("the PF-CZM damage operators need damage_law 'pfczm'")

    def _pfczm_parts(self, d, H):
        self._require_pfczm()
        d = self._check("d", d, (self.n_red,))
        H = self._check_qp("H", H)
        out = torch.empty(3, dtype=F64, device=self.device)
        check(_L.phast_hex8_pfczm_energy(self._plan, ptr(d), ptr(H), ptr(out), stream()), "pfczm_energy")
        return out

    def pfczm_damage_energy(self, d, H) -> torch.Tensor:
        """sum_e [sum_q w H_q g(d_q) + Gc/(pi ell)(2 m.d_e - d_e.M d_e) + Gc ell/pi d_e.L d_e]
        in mJ, a 0-dim tensor (operator.py:461-488); H None drops the history term."""
        return self._pfczm_parts(d, H).sum()

    def pfczm_damage_gradient(self, d, H, scale: float = 1.0) -> torch.Tensor:
        """Gradient (n_red,) of :meth:`pfczm_damage_energy` times ``scale`` (operator.py:490-512)."""
        self._require_pfczm()
        d = self._check("d", d, (self.n_red,))
        H = self._check_qp("H", H)
        out = torch.empty_like(d)
        check(_L.phast_hex8_pfczm_gradient(self._plan, ptr(d), ptr(H), float(scale), ptr(out), stream()),
              "pfczm_gradient")
        return out

    def pfczm_tangent_weights(self, d, H) -> torch.Tensor:
        """Gauss-point curvatures max(H g''(d_q) - 2 Gc/(pi ell), 0.25 * 2 Gc/(pi ell)) (E,8)
        (operator.py:514-538)."""
        self._require_pfczm()
        d = self._check("d", d, (self.n_red,))
        H = self._check_qp("H", H)
        W = torch.empty((self.n_elements, 8), dtype=F64, device=self.device)
        check(_L.phast_hex8_pfczm_tangent_weights(self._plan, ptr(d), ptr(H), ptr(W), stream()), "pfczm_weights")
        return W

    def pfczm_tangent_action(self, field, weights, scale: float = 1.0) -> torch.Tensor:
        """scale (Gc (2 ell/pi) L f_e + sum_q w c_q N_q N_q^T f_e) (operator.py:540-552)."""
        self._require_pfczm()
        field = self._check("field", field, (self.n_red,))
        weights = self._check("weights", weights, (self.n_elements, 8))
        y = torch.empty_like(field)
        check(_L.phast_hex8_pfczm_tangent_action(self._plan, ptr(field), ptr(weights), float(scale), ptr(y),
                                                 stream()), "pfczm_tangent_action")
        return y

    def pfczm_tangent_diagonal(self, weights, scale: float = 1.0) -> torch.Tensor:
        """Jacobi diagonal of :meth:`pfczm_tangent_action` (operator.py:554-564)."""
        self._require_pfczm()
        weights = self._check("weights", weights, (self.n_elements, 8))
        out = torch.empty(self.n_red, dtype=F64, device=self.device)
        check(_L.phast_hex8_pfczm_tangent_diagonal(self._plan, ptr(weights), float(scale), ptr(out), stream()),
              "pfczm_tangent_diagonal")
        return out

    def pfczm_newton_system(self, d, H, *, scaled: bool = True, correction: bool = False):
        """SPD Newton subproblem at d and the gradient there, both times ``damage_scale`` when
        scaled (operator.py:566-609).  rhs = K d - gradient (``correction``: -gradient, the
        unknown being the step)."""
        scale = self.damage_scale if scaled else 1.0
        W = self.pfczm_tangent_weights(d, H)
        gradient = self.pfczm_damage_gradient(d, H, scale)
        action = lambda f: self.pfczm_tangent_action(f, W, scale)   # noqa: E731
        diagonal = self.pfczm_tangent_diagonal(W, scale)
        rhs = -gradient if correction else action(d) - gradient
        return H8ScalarSPDSystem(action=action, rhs=rhs, diagonal=diagonal, _op=self, _H=W, _scale=scale,
                                 _form=1), gradient

    def fracture_energy(self, d) -> dict:
        """Phase-field surface functional in mJ (operator.py:1511-1544; PF-CZM: 611-631):
        ``total``, ``local``, ``gradient``."""
        d = self._check("d", d, (self.n_red,))
        if self._pfczm:
            local, grad, _ = (float(v) for v in self._pfczm_parts(d, None).cpu())
            return {"total": local + grad, "local": local, "gradient": grad}
        return {"total": 0.5 * float(torch.dot(d, self.damage_action(d)))}

    # -- adjoint pieces (transposed derivatives of the residuals) --------------------------
    def set_gc(self, Gc: torch.Tensor) -> None:
        """Replace the per-voxel Gc of the damage operator (and the damage scale built on it)."""
        Gc = self._check("Gc", Gc, (self.n_elements,))
        gmin = float(Gc.min())
        if not (gmin > 0.0 and math.isfinite(gmin) and bool(torch.isfinite(Gc).all())):
            raise ValueError("Gc must be finite and positive at every element")
        check(_L.phast_hex8_set_gc(self._plan, ptr(Gc), stream()), "set_gc")
        self.Gc = Gc.clone()
        self.damage_scale = self.ell / (gmin * float(self.grid.voxel_volume))

    def adj_coupling(self, x, l, d) -> torch.Tensor:
        """(E,8): w g'(d_q) eps(l)_q . C eps(x)_q (the Gauss-point form of (dF/dd)^T l)."""
        x = self._check("x", x, (self.n_mech,))
        l = self._check("l", l, (self.n_mech,))
        d = self._check("d", d, (self.n_red,))
        out = torch.empty((self.n_elements, 8), dtype=F64, device=self.device)
        check(_L.phast_hex8_adj_coupling(self._plan, ptr(x), ptr(l), ptr(d), ptr(out), stream()), "adj_coupling")
        return out

    def adj_driving(self, x, t) -> torch.Tensor:
        """(E,8,6): t_q dY/deps_q."""
        x = self._check("x", x, (self.n_mech,))
        t = self._check_qp("t", t)
        out = torch.empty((self.n_elements, 8, 6), dtype=F64, device=self.device)
        check(_L.phast_hex8_adj_driving(self._plan, ptr(x), ptr(t), ptr(out), stream()), "adj_driving")
        return out

    def field_snapshot(self, step: int, time_value: float, x, d, H=None, g=None) -> "FieldSnapshot":
        """CPU fields of one state (solid3d/operator.py:1561-1652; no interface drivers or
        closure).  Element Gauss-point means of the strain, the degraded stress g C eps, H, the
        degraded energy and the driving energy Y (from ``phast_hex8_field_snapshot`` and
        ``driving_force``), the maximum of H; nodal fluctuation, damage and total displacement
        ``E_tensor (X - origin) + P u~`` over the (nx+1)(ny+1)(nz+1) full nodes."""
        x = self._check("x", x, (self.n_mech,))
        d = self._check("d", d, (self.n_red,))
        H = self._check_qp("H", H)
        g = self.degradation(d) if g is None else self._check_qp("g", g)
        ne = self.n_elements
        strain = torch.empty((ne, 6), dtype=F64, device=self.device)
        stress = torch.empty_like(strain)
        psi = torch.empty(ne, dtype=F64, device=self.device)
        check(_L.phast_hex8_field_snapshot(self._plan, ptr(x), ptr(g), ptr(strain), ptr(stress), ptr(psi),
                                           stream()), "field_snapshot")
        Y = self.driving_force(x)
        if H is None:
            H = torch.zeros((ne, 8), dtype=F64, device=self.device)
        u, E = x[:3 * self.n_red].reshape(-1, 3), x[3 * self.n_red:]
        red = _full_to_reduced(self.grid.shape)
        fluct = u.cpu().numpy()[red]
        e = E.cpu().numpy()
        Et = np.array([[e[0], e[5] / 2, e[4] / 2], [e[5] / 2, e[1], e[3] / 2], [e[4] / 2, e[3] / 2, e[2]]])
        disp = _full_node_offsets(self.grid) @ Et.T + fluct
        return FieldSnapshot(
            step=step, time_value=time_value, macro_strain=e.copy(),
            macro_stress=self.macro_stress(x, g).cpu().numpy(), displacement=disp, fluctuation=fluct,
            damage=d.cpu().numpy()[red], strain=strain.cpu().numpy(), stress=stress.cpu().numpy(),
            history_mean=H.mean(dim=1).cpu().numpy(), history_max=H.amax(dim=1).cpu().numpy(),
            psi_elastic=psi.cpu().numpy(), psi_driving=Y.mean(dim=1).cpu().numpy())

    def qp_to_node(self, f, scale: float = 1.0) -> torch.Tensor:
        """(n,): scale sum_e sum_q N_a(q) f_q."""
        f = self._check_qp("f", f)
        out = torch.empty(self.n_red, dtype=F64, device=self.device)
        check(_L.phast_hex8_qp_to_node(self._plan, ptr(f), float(scale), ptr(out), stream()), "qp_to_node")
        return out

    def qp_stress_apply(self, S) -> torch.Tensor:
        """(n_mech,): [sum_q B_q^T S_q, sum_q S_q] for S (E,8,6)."""
        S = self._check("S", S, (self.n_elements, 8, 6))
        out = torch.empty(self.n_mech, dtype=F64, device=self.device)
        check(_L.phast_hex8_qp_stress_apply(self._plan, ptr(S), ptr(out), stream()), "qp_stress_apply")
        return out

    def adj_damage_history(self, d, l, scale: float = 1.0) -> torch.Tensor:
        """(E,8): scale 2(1-kappa) w (N_q.d - 1)(N_q.l) ((dG/dH)^T l)."""
        d = self._check("d", d, (self.n_red,))
        l = self._check("l", l, (self.n_red,))
        out = torch.empty((self.n_elements, 8), dtype=F64, device=self.device)
        check(_L.phast_hex8_adj_damage_history(self._plan, ptr(d), ptr(l), float(scale), ptr(out), stream()),
              "adj_damage_history")
        return out

    def adj_damage_gc(self, d, l, scale: float = 1.0) -> torch.Tensor:
        """(E,): scale l_e.(M/ell + ell L) d_e ((dG/dGc)^T l)."""
        d = self._check("d", d, (self.n_red,))
        l = self._check("l", l, (self.n_red,))
        out = torch.empty(self.n_elements, dtype=F64, device=self.device)
        check(_L.phast_hex8_adj_damage_gc(self._plan, ptr(d), ptr(l), float(scale), ptr(out), stream()),
              "adj_damage_gc")
        return out

    # -- device-resident solves ----------------------------------------------------------
    def solve_mechanics(self, *, g=None, d=None, fixed_mask, fixed_values, diag, rhs=None, x0=None,
                        rtol=1e-8, atol=0.0, maxiter=5000, check_every=10, closure=None):
        """solve_constrained(mech_action(., g), ...) on the device (linalg.py:223-287); with
        ``closure`` (a ClosurePattern, g required) the action of that fixed pattern."""
        if g is not None and d is not None:
            raise ValueError("pass g or d, not both")
        if d is not None and (self._pfczm or self._closure):
            g, d = self.degradation(d), None
        mode, bits = self._closure_mode(g, closure)
        if self._closure and g is not None and mode != 2:
            raise ValueError("a mechanics solve with the crack closure needs a fixed pattern (closure=...)")
        g = self._check_qp("g", g)
        d = None if d is None else self._check("d", d, (self.n_red,))
        fm = self._check_mask(fixed_mask, self.n_mech)
        fv = self._check("fixed_values", fixed_values, (self.n_mech,))
        dg = self._check("diag", diag, (self.n_mech,))
        rhs = None if rhs is None else self._check("rhs", rhs, (self.n_mech,))
        x0 = None if x0 is None else self._check("x0", x0, (self.n_mech,))
        if rtol < 0 or atol < 0 or (rtol == 0 and atol == 0):
            raise ValueError("rtol and atol must be non-negative and not both zero")
        x = torch.empty(self.n_mech, dtype=F64, device=self.device)
        info = (C.c_double * 5)()
        if mode == 2:
            rc = _L.phast_hex8_mech_solve_cl(self._plan, ptr(g), ptr(bits), ptr(fm), ptr(fv), ptr(rhs), ptr(x0),
                                             ptr(dg), float(rtol), float(atol), int(maxiter), int(check_every),
                                             self.n_mech, ptr(x), info, stream())
        else:
            rc = _L.phast_hex8_mech_solve(self._plan, ptr(g), ptr(d), ptr(fm), ptr(fv), ptr(rhs), ptr(x0), ptr(dg),
                                          float(rtol), float(atol), int(maxiter), int(check_every), self.n_mech,
                                          ptr(x), info, stream())
        _raise_pcg(rc)
        return x, PCGInfo(int(info[0]), bool(info[1]), info[2], info[3], info[4])

    def _check_mask(self, m, n):
        if not isinstance(m, torch.Tensor) or m.dtype != torch.bool or m.shape != (n,) or m.device != self.device:
            raise TypeError(f"mask must be a torch.bool tensor of shape ({n},) on {self.device}")
        return m.contiguous().view(torch.uint8)


def _raise_pcg(rc: int) -> None:
    if rc == 0:
        return
    msg = _L.phast_last_error().decode(errors="replace")
    if rc in (_lib.PHAST_ERR_NONFINITE, _lib.PHAST_ERR_NONPOSITIVE):
        raise PCGBreakdownError(msg)
    check(rc)


def _raise_scalar(rc: int) -> None:
    """box_kkt_h8._pcg breakdowns are RuntimeError / ValueError with fixed messages."""
    if rc == 0:
        return
    msg = _L.phast_last_error().decode(errors="replace")
    if rc == _lib.PHAST_ERR_NONFINITE:
        raise ValueError(msg)
    if rc == _lib.PHAST_ERR_NONPOSITIVE:
        raise RuntimeError(msg)
    check(rc)


# ----------------------------------------------------------------------------------------
# linalg.pcg / solve_constrained
# ----------------------------------------------------------------------------------------

_APPLY_CB = C.CFUNCTYPE(C.c_int, C.c_void_p)
_L.phast_pcg_callback.argtypes = [_APPLY_CB, C.c_void_p, C.c_int64] + [C.c_void_p] * 7 + [
    C.c_double, C.c_double, C.c_int32, C.c_int32, C.POINTER(C.c_double), C.c_void_p]
_L.phast_pcg_callback.restype = C.c_int


def pcg(action, rhs, *, diag, x0=None, rtol=1e-8, atol=0.0, maxiter=5000, check_every=10, fixed=None):
    """Jacobi PCG of ``linalg.pcg`` (solid3d/linalg.py:99-220) for any device action.

    The Krylov loop is the library's (``phast_pcg_callback``): vector updates, the
    flag-masked steps, the fixed-order dot products and the flag reads every
    ``check_every`` iterations all run there.  Python is entered once per iteration, to
    evaluate ``action`` on the direction the library places in a device buffer.
    ``fixed`` (bool, optional) zeroes those entries of every image.  Operator-bound solves
    are faster through :meth:`PeriodicVoxelPhaseFieldOperator.solve_mechanics`."""
    rtol, atol = float(rtol), float(atol)
    if rtol < 0 or atol < 0 or (rtol == 0 and atol == 0):
        raise ValueError("rtol and atol must be non-negative and not both zero")
    if not isinstance(rhs, torch.Tensor) or rhs.dim() != 1 or rhs.dtype != F64 or rhs.device.type != "cuda":
        raise TypeError("rhs must be a float64 vector on a ROCm device")
    n, dev = rhs.shape[0], rhs.device
    for name, t in (("diag", diag), ("x0", x0)):
        if t is not None and (not isinstance(t, torch.Tensor) or t.shape != rhs.shape or t.dtype != F64
                              or t.device != dev):
            raise TypeError(f"{name} must be a float64 tensor of shape ({n},) on {dev}")
    rhs, diag = rhs.contiguous(), diag.contiguous()
    x0 = None if x0 is None else x0.contiguous()
    fx = None if fixed is None else fixed.to(device=dev, dtype=torch.uint8).contiguous()
    inbuf, outbuf = torch.empty_like(rhs), torch.empty_like(rhs)
    failure = []

    def evaluate(_ctx):
        try:
            y = action(inbuf)
            if not isinstance(y, torch.Tensor) or y.shape != inbuf.shape:
                raise ValueError(f"PCG action must return a tensor of shape ({n},)")
            outbuf.copy_(y)
            return 0
        except BaseException as exc:  # re-raised below, after the library has unwound
            failure.append(exc)
            return 1

    cb = _APPLY_CB(evaluate)
    x = torch.empty_like(rhs)
    info = (C.c_double * 5)()
    rc = _L.phast_pcg_callback(cb, None, n, ptr(inbuf), ptr(outbuf), ptr(fx), ptr(rhs), ptr(diag), ptr(x0), ptr(x),
                               rtol, atol, int(maxiter), int(check_every), info, stream())
    if failure:
        raise failure[0]
    _raise_pcg(rc)
    return x, PCGInfo(int(info[0]), bool(info[1]), info[2], info[3], info[4])


def solve_constrained(action, *, diag, fixed_mask, fixed_values, rhs=None, x0=None, rtol=1e-8, atol=0.0,
                      maxiter=5000, check_every=10):
    """Dirichlet-eliminated PCG of ``linalg.solve_constrained`` (solid3d/linalg.py:223-287):
    the prescribed entries are moved to the right-hand side and the PCG runs on the rest.

    An action made by :func:`mech_action_of` runs the whole solve in the library; any
    other callable goes through :func:`pcg` (library loop, Python apply)."""
    bound = getattr(action, "_phast_mech", None)
    if bound is not None:
        op, g, d, closure = bound
        return op.solve_mechanics(g=g, d=d, fixed_mask=fixed_mask, fixed_values=fixed_values, diag=diag, rhs=rhs,
                                  x0=x0, rtol=rtol, atol=atol, maxiter=maxiter, check_every=check_every,
                                  closure=closure)
    zero = torch.zeros((), dtype=fixed_values.dtype, device=fixed_values.device)
    prescribed = torch.where(fixed_mask, fixed_values, zero)   # known part of the solution
    lifted = -action(prescribed) if rhs is None else rhs - action(prescribed)
    y, info = pcg(action, torch.where(fixed_mask, zero, lifted),
                  diag=torch.where(fixed_mask, torch.ones_like(diag), diag),
                  x0=None if x0 is None else torch.where(fixed_mask, zero, x0),
                  rtol=rtol, atol=atol, maxiter=maxiter, check_every=check_every, fixed=fixed_mask)
    return torch.where(fixed_mask, prescribed, y), info


def mech_action_of(op: PeriodicVoxelPhaseFieldOperator, g=None, d=None, closure=None):
    """``lambda v: op.mech_action(v, g, closure=closure)`` tagged so that solve_constrained
    runs on the device."""
    if d is None:
        fn = lambda v: op.mech_action(v, g, closure=closure)   # noqa: E731
    else:
        fn = lambda v: op.mech_action_d(v, d)   # noqa: E731
    fn._phast_mech = (op, g, d, closure)
    return fn


# ----------------------------------------------------------------------------------------
# box-KKT damage solve (solvers/box_kkt_h8.py)
# ----------------------------------------------------------------------------------------

@dataclass(frozen=True)
class ScalarPCGResult:
    solution: torch.Tensor
    iterations: int
    converged: bool
    initial_residual_norm: float
    residual_norm: float
    relative_residual: float


@dataclass(frozen=True)
class ScalarDirichletResult:
    solution: torch.Tensor
    krylov: ScalarPCGResult
    n_free_dofs: int

    @property
    def converged(self) -> bool:
        return self.krylov.converged


@dataclass(frozen=True)
class BoxKKTIteration:
    iteration: int
    active_lower_dofs: int
    active_upper_dofs: int
    prescribed_equal_bound_dofs: int
    free_dofs: int
    pcg_iterations: int
    pcg_converged: bool
    added_lower_dofs: int
    added_upper_dofs: int
    released_lower_dofs: int
    released_upper_dofs: int


@dataclass(frozen=True)
class BoxKKTDiagnostics:
    kkt_tolerance: float
    free_stationarity_maximum: float
    projected_gradient_maximum: float
    complementarity_maximum: float
    lower_sign_violation_dofs: int
    upper_sign_violation_dofs: int
    satisfies_kkt: bool


@dataclass(frozen=True)
class H8ScalarBoxKKTResult:
    formulation_identifier: str
    solution: torch.Tensor
    lower_bound: torch.Tensor
    upper_bound: torch.Tensor
    lower_active: torch.Tensor
    upper_active: torch.Tensor
    prescribed_equal_bound: torch.Tensor
    last_linear_solve: ScalarDirichletResult
    active_set_iterations: int
    total_pcg_iterations: int
    active_set_trace: tuple
    kkt: BoxKKTDiagnostics


class H8ScalarBoxKKTConvergenceError(RuntimeError):
    """Raised instead of returning a non-converged or non-KKT field (box_kkt_h8.py:239-253)."""

    def __init__(self, reason, *, last_solution, last_linear_solve, evidence):
        super().__init__(reason)
        self.last_solution = last_solution
        self.last_linear_solve = last_linear_solve
        self.evidence = evidence


def solve_h8_scalar_dirichlet_pcg(system: H8ScalarSPDSystem, *, fixed, values, initial=None,
                                  tolerance=DEFAULT_LINEAR_TOLERANCE, max_iterations=4000,
                                  _n_free: int | None = None) -> ScalarDirichletResult:
    """Exact-Dirichlet PCG of the damage system on the device (box_kkt_h8.py:350-432).
    ``_n_free`` (internal): the number of free entries when the caller already counted them."""
    op = system._op
    if op is None:
        raise NotImplementedError("the device solve needs a system built by PeriodicVoxelPhaseFieldOperator")
    n = system.n_dofs
    fx = op._check_mask(fixed, n)
    values = op._check("values", values, (n,))
    init = None if initial is None else op._check("initial", initial, (n,))
    n_free = int(n - fixed.sum().item()) if _n_free is None else int(_n_free)
    if n_free == 0:
        known = torch.where(fixed, values, torch.zeros_like(values))
        z = torch.empty(0, dtype=F64, device=op.device)
        return ScalarDirichletResult(known, ScalarPCGResult(z, 0, True, 0.0, 0.0, 0.0), 0)
    if not (tolerance > 0 and math.isfinite(tolerance)):
        raise ValueError(f"tolerance must be finite and positive, got {tolerance!r}")
    out = torch.empty(n, dtype=F64, device=op.device)
    info = (C.c_double * 5)()
    fn = _L.phast_hex8_pfczm_dirichlet_pcg if system._form == 1 else _L.phast_hex8_damage_dirichlet_pcg
    rc = fn(op._plan, ptr(system._H), float(system._scale), ptr(system.rhs), ptr(system.diagonal), ptr(fx),
            ptr(values), ptr(init), float(tolerance), int(max_iterations), n, ptr(out), info, stream())
    _raise_scalar(rc)
    krylov = ScalarPCGResult(out, int(info[0]), bool(info[1]), info[2], info[3], info[4])
    return ScalarDirichletResult(out, krylov, n_free)


def _kkt_diagnostics(system, solution, lower, upper, lact, uact, presc, kkt_tol, grad=None) -> BoxKKTDiagnostics:
    """Audit of box_kkt_h8.py:459-558 (the quantities its ``satisfies_kkt`` depends on), in
    one device pass (``phast_hex8_kkt_stats``) for a system built by the voxel operator.
    ``grad`` (optional) is A solution - b when the caller already has it."""
    if grad is None:
        grad = system.apply(solution) - system.rhs
    op = system._op
    if op is not None:
        n = system.n_dofs
        out = (C.c_double * 8)()
        u8 = lambda m: m.contiguous().view(torch.uint8)  # noqa: E731
        check(_L.phast_hex8_kkt_stats(op._plan, ptr(grad.contiguous()), ptr(solution.contiguous()),
                                      ptr(lower.contiguous()), ptr(upper.contiguous()), ptr(u8(lact)), ptr(u8(uact)),
                                      ptr(u8(presc)), float(kkt_tol), n, out, stream()), "kkt_stats")
        stats = list(out)
        ok = (stats[5] == 0 and stats[6] == 0 and stats[7] == 0 and stats[3] == 0 and stats[4] == 0
              and stats[0] <= kkt_tol and stats[1] <= kkt_tol and stats[2] <= kkt_tol)
        return BoxKKTDiagnostics(kkt_tol, stats[0], stats[1], stats[2], int(stats[3]), int(stats[4]), bool(ok))
    fixed = lact | uact | presc
    free = ~fixed
    zero = torch.zeros_like(grad)
    free_g = torch.where(free, grad, zero).abs().max()
    proj = torch.where(free, grad, zero)
    proj = torch.where(lact, torch.minimum(grad, zero), proj)
    proj = torch.where(uact, torch.maximum(grad, zero), proj)
    lmult = torch.where(lact, grad.clamp_min(0.0), zero)
    umult = torch.where(uact, (-grad).clamp_min(0.0), zero)
    comp = torch.maximum((lmult * (solution - lower)).abs().max(), (umult * (upper - solution)).abs().max())
    lsv = (lact & (grad < -kkt_tol)).sum()
    usv = (uact & (-grad < -kkt_tol)).sum()
    stats = torch.stack([free_g, proj.abs().max(), comp, lsv.to(F64), usv.to(F64),
                         (solution < lower).sum().to(F64), (solution > upper).sum().to(F64),
                         (lower > upper).sum().to(F64)]).cpu().tolist()
    ok = (stats[5] == 0 and stats[6] == 0 and stats[7] == 0 and stats[3] == 0 and stats[4] == 0
          and stats[0] <= kkt_tol and stats[1] <= kkt_tol and stats[2] <= kkt_tol)
    return BoxKKTDiagnostics(kkt_tol, stats[0], stats[1], stats[2], int(stats[3]), int(stats[4]), bool(ok))


def solve_h8_scalar_box_kkt(system: H8ScalarSPDSystem, *, lower_bound=0.0, upper_bound=1.0, initial=None,
                            linear_tolerance=DEFAULT_LINEAR_TOLERANCE, kkt_tolerance=DEFAULT_KKT_TOLERANCE,
                            max_linear_iterations=4000, max_active_set_iterations=64) -> H8ScalarBoxKKTResult:
    """Primal active-set box-KKT solve (box_kkt_h8.py:603-778) on the device."""
    op = system._op
    if op is None:
        raise NotImplementedError("the device solve needs a system built by PeriodicVoxelPhaseFieldOperator")
    n = system.n_dofs
    dev = op.device

    def bound(v, name):
        if isinstance(v, torch.Tensor):
            return op._check(name, v, (n,))
        return torch.full((n,), float(v), dtype=F64, device=dev)

    lower, upper = bound(lower_bound, "lower_bound"), bound(upper_bound, "upper_bound")
    if bool((lower > upper).any()):
        raise ValueError("lower_bound must not exceed upper_bound at any degree of freedom")
    if initial is not None:
        initial = op._check("initial", initial, (n,))
        if bool((initial < lower).any()) or bool((initial > upper).any()):
            raise ValueError("initial must already satisfy the requested box bounds; it is never clipped")
    presc = lower == upper
    lact = torch.zeros(n, dtype=torch.bool, device=dev)
    uact = torch.zeros_like(lact)
    cand_init = initial
    trace, total = [], 0
    grad = torch.empty(n, dtype=F64, device=dev)
    counts = (C.c_int64 * 8)()
    last = None
    cand = None
    n_pr = int(presc.sum())
    free_next = n - n_pr  # free entries of the next solve (the update pass counts them)
    for it in range(1, max_active_set_iterations + 1):
        fixed = lact | uact | presc
        values = torch.where(uact, upper, lower)
        last = solve_h8_scalar_dirichlet_pcg(system, fixed=fixed, values=values, initial=cand_init,
                                             tolerance=linear_tolerance, max_iterations=max_linear_iterations,
                                             _n_free=free_next)
        cand = last.solution
        total += last.krylov.iterations
        apply = 1 if it < max_active_set_iterations else 0
        la8, ua8 = lact.view(torch.uint8).clone(), uact.view(torch.uint8).clone()
        upd = _L.phast_hex8_pfczm_box_update if system._form == 1 else _L.phast_hex8_box_update
        rc = upd(op._plan, ptr(system._H), float(system._scale), ptr(system.rhs), ptr(cand), ptr(lower), ptr(upper),
                 ptr(la8), ptr(ua8), ptr(presc.view(torch.uint8)), float(kkt_tolerance), apply, n, ptr(grad), counts,
                 stream())
        check(rc, "box_update")
        add_l, add_u, rel_l, rel_u, n_free, n_la, n_ua, free_next = (int(c) for c in counts)
        trace.append(BoxKKTIteration(it, n_la, n_ua, n_pr, n_free, last.krylov.iterations, last.krylov.converged,
                                     add_l, add_u, rel_l, rel_u))
        if not last.krylov.converged:
            raise H8ScalarBoxKKTConvergenceError("H8 scalar box-KKT PCG did not converge", last_solution=cand,
                                                 last_linear_solve=last, evidence=(it, total, tuple(trace)))
        if add_l + add_u + rel_l + rel_u == 0:  # the update pass left grad = A cand - b for these masks
            kkt = _kkt_diagnostics(system, cand, lower, upper, lact, uact, presc, kkt_tolerance, grad=grad)
            if not kkt.satisfies_kkt:
                raise H8ScalarBoxKKTConvergenceError("H8 scalar box-KKT active set reached a non-KKT state",
                                                     last_solution=cand, last_linear_solve=last,
                                                     evidence=(it, total, tuple(trace), kkt))
            return H8ScalarBoxKKTResult(H8_SCALAR_BOX_KKT_IDENTIFIER, cand, lower, upper, lact, uact, presc, last, it,
                                        total, tuple(trace), kkt)
        if it == max_active_set_iterations:
            break
        lact, uact = la8.view(torch.bool), ua8.view(torch.bool)
        cand_init = cand
    raise H8ScalarBoxKKTConvergenceError("H8 scalar box-KKT active set did not stabilise", last_solution=cand,
                                         last_linear_solve=last, evidence=(max_active_set_iterations, total,
                                                                           tuple(trace)))


# ----------------------------------------------------------------------------------------
# geometric multigrid on the periodic voxel hierarchy (new; no reference counterpart)
# ----------------------------------------------------------------------------------------

for _name, _args in {
    "phast_hex8_mg_create": [C.c_void_p, C.c_int64, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                             C.c_double, C.POINTER(C.c_void_p)],
    "phast_hex8_mg_destroy": [C.c_void_p],
    "phast_hex8_mg_update": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "phast_hex8_mg_apply": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "phast_hex8_mg_levels": [C.c_void_p, C.POINTER(C.c_int32), C.c_void_p],
    "phast_hex8_mg_weights": [C.c_void_p, C.c_void_p, C.c_void_p],
    "phast_hex8_mech_solve_mg": [C.c_void_p, C.c_void_p] + [C.c_void_p] * 7 +
                                [C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_int64, C.c_void_p,
                                 C.POINTER(C.c_double), C.c_void_p],
}.items():
    _fn = getattr(_L, _name)
    _fn.argtypes = _args
    _fn.restype = C.c_int


class VoxelMultigrid:
    """V-cycle preconditioner of the voxel mechanics operator on the 2:1 periodic hierarchy.

    Coarse levels are rediscretisations with voxel-averaged moduli ``mean(g) lam``,
    ``mean(g) mu`` (the Galerkin product for constant coefficients); damped Jacobi
    smoothing (``nu`` sweeps before and after, weight ``omega / lambda_max(D^-1 A)`` per
    level from a power iteration), ``coarse_sweeps`` sweeps
    on the coarsest level; trilinear prolongation, restriction its transpose.  The
    preconditioned PCG keeps the flag semantics of solid3d/linalg.py:99-220, so the
    converged solution agrees with the Jacobi-PCG one to the solver tolerance."""

    def __init__(self, op: PeriodicVoxelPhaseFieldOperator, *, max_levels: int = 6, min_coarse: int = 4, nu: int = 2,
                 coarse_sweeps: int = 40, omega: float = 1.2):
        self.op = op
        macro = np.zeros(6, dtype=np.uint8)
        macro[op.load.controlled] = 1
        for c in op.load.fixed_zero:
            macro[c] = 1
        h = C.c_void_p()
        check(_L.phast_hex8_mg_create(op._plan, op.pinned_node, macro.ctypes.data, int(max_levels), int(min_coarse),
                                      int(nu), int(coarse_sweeps), float(omega), C.byref(h)), "mg_create")
        self._mg = h
        lv = C.c_int32()
        dims = (C.c_int32 * 24)()
        check(_L.phast_hex8_mg_levels(h, C.byref(lv), dims))
        self.levels = [tuple(dims[3 * l:3 * l + 3]) for l in range(lv.value)]
        self._src = None

    def __del__(self):
        h = getattr(self, "_mg", None)
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _L.phast_hex8_mg_destroy(h)
            self._mg = None

    def update(self, *, g=None, d=None):
        """Rebuild coarse moduli and diagonals for the degradation g (E,8) or nodal d."""
        op = self.op
        g = op._check_qp("g", g)
        d = None if d is None else op._check("d", d, (op.n_red,))
        check(_L.phast_hex8_mg_update(self._mg, ptr(g), ptr(d), stream()), "mg_update")
        self._src = (g, d)
        lam = (C.c_double * 8)()
        om = (C.c_double * 8)()
        check(_L.phast_hex8_mg_weights(self._mg, lam, om))
        self.lambda_max = list(lam)[:len(self.levels)]
        self.omega = list(om)[:len(self.levels)]

    def apply(self, r: torch.Tensor) -> torch.Tensor:
        r = self.op._check("r", r, (self.op.n_mech,))
        z = torch.empty_like(r)
        check(_L.phast_hex8_mg_apply(self._mg, ptr(r), ptr(z), stream()), "mg_apply")
        return z

    def solve(self, *, g=None, d=None, fixed_mask, fixed_values, diag, rhs=None, x0=None, rtol=1e-8, atol=0.0,
              maxiter=5000, check_every=10, refresh: bool = True):
        """MG-preconditioned solve_constrained of mech_action(., g(d)); ``refresh=False`` keeps
        the hierarchy of the previous call (same degradation)."""
        op = self.op
        if refresh or self._src is None:
            self.update(g=g if d is None else None, d=d)
        g = op._check_qp("g", g) if d is None else None
        d = None if d is None else op._check("d", d, (op.n_red,))
        fm = op._check_mask(fixed_mask, op.n_mech)
        fv = op._check("fixed_values", fixed_values, (op.n_mech,))
        dg = op._check("diag", diag, (op.n_mech,))
        rhs = None if rhs is None else op._check("rhs", rhs, (op.n_mech,))
        x0 = None if x0 is None else op._check("x0", x0, (op.n_mech,))
        x = torch.empty(op.n_mech, dtype=F64, device=op.device)
        info = (C.c_double * 5)()
        rc = _L.phast_hex8_mech_solve_mg(op._plan, self._mg, ptr(g), ptr(d), ptr(fm), ptr(fv), ptr(rhs), ptr(x0),
                                         ptr(dg), float(rtol), float(atol), int(maxiter), int(check_every), op.n_mech,
                                         ptr(x), info, stream())
        _raise_pcg(rc)
        return x, PCGInfo(int(info[0]), bool(info[1]), info[2], info[3], info[4])
