"""Oracle (numpy, float64) for the reference's native-Q4 explicit phase-field path.

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.  Imported by ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg; never by the product.

The functions restate, op for op, the reference files cited in each docstring
(paths relative to the reference's ``src/phast``).  Scatter-adds use
``np.bincount`` (sequential accumulation in index order, like ``scatter_add_`` on CPU).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

GAUSS_A = 1.0 / math.sqrt(3.0)


# ----------------------------------------------------------------------------------------
# a1: quadrature geometry and mesh precompute
# ----------------------------------------------------------------------------------------

def q4_gauss_points():
    """2x2 Gauss-Legendre points, eta outer / xi inner; weights 1 (core/quad_elements.py:223-250)."""
    a = 1.0 / np.sqrt(np.float64(3.0))
    pts = np.array([[-a, -a], [a, -a], [-a, a], [a, a]], dtype=np.float64)
    return pts, np.ones(4, dtype=np.float64)


def q4_shape_functions(xi: float, eta: float) -> np.ndarray:
    """Bilinear shape functions, counter-clockwise corners (core/quad_elements.py:90-99)."""
    return 0.25 * np.array([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta),
                            (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)], dtype=np.float64)


def q4_shape_function_derivs(xi: float, eta: float) -> np.ndarray:
    """(4, 2) derivatives w.r.t. (xi, eta) (core/quad_elements.py:102-118)."""
    dxi = 0.25 * np.array([-(1 - eta), (1 - eta), (1 + eta), -(1 + eta)])
    deta = 0.25 * np.array([-(1 - xi), -(1 + xi), (1 + xi), (1 - xi)])
    return np.stack([dxi, deta], axis=1)


def q4_quadrature_geometry(nodes: np.ndarray, quads: np.ndarray):
    """Return ``N_q (4,4)``, ``gradN (E,4,4,2)``, ``wdetJ (E,4)`` (core/quad_elements.py:315-349)."""
    coords = nodes[quads]
    pts, wts = q4_gauss_points()
    E = coords.shape[0]
    grad = np.zeros((E, 4, 4, 2))
    wdet = np.zeros((E, 4))
    N_all = []
    for q in range(4):
        xi, eta = pts[q]
        N = q4_shape_functions(xi, eta)
        dN = q4_shape_function_derivs(xi, eta)
        J = np.einsum("ai,eaj->eij", dN, coords)
        det = np.linalg.det(J)
        if np.any(det <= 0.0):
            raise ValueError("Q4 element has non-positive Jacobian")
        invJ = np.linalg.inv(J)
        grad[:, q] = np.einsum("ai,eji->eaj", dN, invJ)
        wdet[:, q] = wts[q] * det
        N_all.append(N)
    return np.stack(N_all), grad, wdet


def structured_q4_mesh(W: float, H: float, nx: int, ny: int):
    """Structured Q4 plate: ``node(i, j) = j (nx+1) + i``, cells row-major with corners
    ``(i,j),(i+1,j),(i+1,j+1),(i,j+1)`` (core/mesh_generator.py:1735-1752)."""
    x = np.linspace(0.0, W, nx + 1)
    y = np.linspace(0.0, H, ny + 1)
    xx, yy = np.meshgrid(x, y, indexing="xy")
    nodes = np.column_stack([xx.ravel(), yy.ravel()])
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    n0 = (jj * (nx + 1) + ii).ravel()
    quads = np.stack([n0, n0 + 1, n0 + nx + 2, n0 + nx + 1], axis=1).astype(np.int64)
    return nodes, quads


@dataclass
class Q4Mesh:
    """Q4 precompute (core/mesh.py:305-360): quadrature data, element h, lumped scalar mass."""
    nodes: np.ndarray
    elements: np.ndarray
    N_q: np.ndarray = field(init=False)
    grad: np.ndarray = field(init=False)
    wdet: np.ndarray = field(init=False)

    def __post_init__(self):
        self.nodes = np.asarray(self.nodes, dtype=np.float64)
        self.elements = np.asarray(self.elements, dtype=np.int64)
        self.N_q, self.grad, self.wdet = q4_quadrature_geometry(self.nodes, self.elements)
        self.areas = self.wdet.sum(axis=1)
        c = self.nodes[self.elements]
        edges = np.stack([np.linalg.norm(c[:, 1] - c[:, 0], axis=1),
                          np.linalg.norm(c[:, 2] - c[:, 1], axis=1),
                          np.linalg.norm(c[:, 3] - c[:, 2], axis=1),
                          np.linalg.norm(c[:, 0] - c[:, 3], axis=1)], axis=1)
        self.elem_h = edges.min(axis=1)
        self.h_min = float(self.elem_h.min())
        self.elem_flat = self.elements.ravel()
        local_mass = np.einsum("qi,eq->ei", self.N_q, self.wdet)
        self.M_scalar = scatter(self.elem_flat, local_mass.ravel(), self.n_nodes)

    @property
    def n_nodes(self) -> int:
        return self.nodes.shape[0]

    @property
    def n_elems(self) -> int:
        return self.elements.shape[0]

    def elem_to_node(self, field_q: np.ndarray) -> np.ndarray:
        """Quadrature-weighted projection of an (E,) or (E,4) field (core/mesh.py:407-428)."""
        if field_q.shape == (self.n_elems,):
            field_q = np.repeat(field_q[:, None], 4, axis=1)
        local = np.einsum("qa,eq,eq->ea", self.N_q, field_q, self.wdet)
        ns = scatter(self.elem_flat, local.ravel(), self.n_nodes)
        return ns / (self.M_scalar + 1e-30)


def scatter(index: np.ndarray, values: np.ndarray, n: int) -> np.ndarray:
    """``zeros(n).scatter_add_(0, index, values)``."""
    return np.bincount(index, weights=values, minlength=n).astype(np.float64)


# ----------------------------------------------------------------------------------------
# a2-a8: material, element apply, driving force, history
# ----------------------------------------------------------------------------------------

@dataclass
class Material:
    """Constants and degradation (physics/material.py:379-430, 497-529; 'standard' AT2)."""
    E: float
    nu: float
    rho: float
    Gc: float
    l0: float
    eta_residual: float = 1e-7
    plane_stress: bool = False
    energy_split: str = "amor"
    stress_degradation: str = "split"
    pf_model: str = "AT2"
    at1_threshold: object = "auto"

    @property
    def at1_source(self) -> float:
        """AT1 source term S_H of the damage right-hand side 2H - S_H: the override or
        3 Gc / (8 l0); 0 for AT2 (material.py:260-271)."""
        if self.pf_model != "AT1":
            return 0.0
        if isinstance(self.at1_threshold, str) and self.at1_threshold == "auto":
            return 3.0 * self.Gc / (8.0 * self.l0)
        return float(self.at1_threshold)

    @property
    def stress_split(self) -> str:
        """Split of the stress: 'isotropic' under stress_degradation='full', which degrades the
        whole stress by one factor; the history keeps energy_split (material.py:367-376)."""
        return "isotropic" if self.stress_degradation == "full" else self.energy_split

    @property
    def lam(self) -> float:
        if self.plane_stress:
            return self.E * self.nu / (1.0 - self.nu ** 2)
        return self.E * self.nu / ((1 + self.nu) * (1 - 2 * self.nu))

    @property
    def mu(self) -> float:
        return self.E / (2 * (1 + self.nu))

    @property
    def kappa(self) -> float:
        return self.E / (3.0 * (1.0 - 2.0 * self.nu))

    @property
    def c_p(self) -> float:
        return math.sqrt((self.lam + 2 * self.mu) / self.rho)

    def degradation(self, d: np.ndarray) -> np.ndarray:
        return (1.0 - d) ** 2 * (1.0 - self.eta_residual) + self.eta_residual


class FEMOperators:
    """Q4 branch of ``FEMOperators`` (core/fem_operators.py)."""

    def __init__(self, mesh: Q4Mesh, material: Material):
        self.mesh, self.material = mesh, material
        # fem_operators.py:83-85
        self.M_vec = material.rho * np.repeat(mesh.M_scalar, 2)
        self.M_vec_inv = 1.0 / (self.M_vec + 1e-30)
        self.dt_cfl = mesh.h_min / material.c_p  # fem_operators.py:88

    def compute_strain(self, u: np.ndarray):
        """fem_operators.py:244-254."""
        gp = self.mesh.grad
        u_e = u[self.mesh.elements]
        exx = np.einsum("eqa,ea->eq", gp[..., 0], u_e[..., 0])
        eyy = np.einsum("eqa,ea->eq", gp[..., 1], u_e[..., 1])
        gxy = (np.einsum("eqa,ea->eq", gp[..., 1], u_e[..., 0])
               + np.einsum("eqa,ea->eq", gp[..., 0], u_e[..., 1]))
        return exx, eyy, gxy

    def degradation_q(self, d: np.ndarray) -> np.ndarray:
        """d_q = N d_e then g(d_q) (fem_operators.py:626-630)."""
        d_q = np.einsum("qa,ea->eq", self.mesh.N_q, d[self.mesh.elements])
        return self.material.degradation(d_q)

    def compute_stress(self, u, d, strain=None):
        """fem_operators.py:614-653 -> amor 301-333 / isotropic 277-290."""
        g = self.degradation_q(d)
        if strain is None:
            strain = self.compute_strain(u)
        exx, eyy, gxy = strain
        m = self.material
        if m.stress_split == "isotropic":
            lam, mu = m.lam, m.mu
            c11 = lam + 2.0 * mu
            return g * (c11 * exx + lam * eyy), g * (lam * exx + c11 * eyy), g * (mu * gxy)
        if m.stress_split != "amor":
            raise ValueError(m.stress_split)
        mu, kappa = m.mu, m.kappa
        tr2 = exx + eyy
        ezz = (-m.nu / (1.0 - m.nu) * tr2) if m.plane_stress else 0.0
        tr = tr2 + ezz
        trp = np.maximum(tr, 0.0)
        trm = tr - trp
        dxx = exx - tr / 3.0
        dyy = eyy - tr / 3.0
        sxx = g * (kappa * trp + 2 * mu * dxx) + kappa * trm
        syy = g * (kappa * trp + 2 * mu * dyy) + kappa * trm
        sxy = g * mu * gxy
        return sxx, syy, sxy

    def assemble_force(self, sxx, syy, sxy) -> np.ndarray:
        """fem_operators.py:681-699."""
        gp, w = self.mesh.grad, self.mesh.wdet
        fx = ((gp[..., 0] * sxx[..., None] + gp[..., 1] * sxy[..., None]) * w[..., None]).sum(axis=1)
        fy = ((gp[..., 1] * syy[..., None] + gp[..., 0] * sxy[..., None]) * w[..., None]).sum(axis=1)
        n = self.mesh.n_nodes
        ef = self.mesh.elem_flat
        return np.stack([scatter(ef, fx.ravel(), n), scatter(ef, fy.ravel(), n)], axis=1)

    def internal_force(self, u: np.ndarray, d: np.ndarray) -> np.ndarray:
        """fem_operators.py:598-612."""
        strain = self.compute_strain(u)
        return self.assemble_force(*self.compute_stress(u, d, strain=strain))

    def compute_psi_plus(self, u, strain=None) -> np.ndarray:
        """Amor psi+ (fem_operators.py:983-1006) / isotropic psi (974-981)."""
        if strain is None:
            strain = self.compute_strain(u)
        exx, eyy, gxy = strain
        m = self.material
        if m.energy_split == "isotropic":
            exy = gxy / 2.0
            tr = exx + eyy
            return 0.5 * m.lam * tr ** 2 + m.mu * (exx ** 2 + eyy ** 2 + 2 * exy ** 2)
        tr2 = exx + eyy
        ezz = (-m.nu / (1.0 - m.nu) * tr2) if m.plane_stress else np.zeros_like(exx)
        tr = tr2 + ezz
        trp = np.maximum(tr, 0.0)
        dxx, dyy, dzz = exx - tr / 3.0, eyy - tr / 3.0, ezz - tr / 3.0
        exy = gxy / 2.0
        return 0.5 * m.kappa * trp ** 2 + m.mu * (dxx ** 2 + dyy ** 2 + dzz ** 2 + 2.0 * exy ** 2)

    compute_driving_force = compute_psi_plus  # fem_operators.py:1173-1204 ('strain_energy')

    def compute_reaction_force(self, u, d, node_indices, component=1, f_ext=None) -> float:
        """fem_operators.py:1502-1530."""
        f = self.internal_force(u, d)
        r = f[node_indices, component]
        if f_ext is not None:
            r = r - f_ext[node_indices, component]
        return float(r.sum())


def history_update(H_old: np.ndarray, psi: np.ndarray) -> np.ndarray:
    """``hard_max`` (physics/h_update.py:47-49) via StaggeredSolver._H_update (staggered_solver.py:804-845)."""
    return np.maximum(H_old, psi)


# ----------------------------------------------------------------------------------------
# a9: central-difference explicit step
# ----------------------------------------------------------------------------------------

class ExplicitDynamics:
    """Predictor / corrector with lumped mass (solvers/mechanics_solver.py:153-239), with the
    optional Kelvin-Voigt damping force beta f_int(v_n, d), beta = zeta dt_cfl."""

    def __init__(self, fem: FEMOperators, dt: float | None = None, damping_ratio_max: float = 0.0):
        self.fem = fem
        self.dt = fem.dt_cfl if dt is None else dt
        self.beta = damping_ratio_max * fem.dt_cfl  # mechanics_solver.py:139-142

    def predictor(self, u, v, a, bc_mask, bc_vals):
        dt = self.dt
        u_new = (u + dt * v) + (0.5 * dt * dt) * a
        if bc_mask is not None:
            u_new = np.where(bc_mask, bc_vals, u_new)
        return u_new

    def corrector(self, u_new, v, a, d, f_ext, bc_mask):
        dt = self.dt
        f_int = self.fem.internal_force(u_new, d)
        f_net = f_ext - f_int
        if self.beta > 0.0:  # Kelvin-Voigt damping (mechanics_solver.py:192-195)
            f_net = f_net - self.beta * self.fem.internal_force(v, d)
        a_new = self.fem.M_vec_inv.reshape(-1, 2) * f_net
        v_new = v + 0.5 * dt * (a + a_new)
        if bc_mask is not None:
            a_new = np.where(bc_mask, 0.0, a_new)
            v_new = np.where(bc_mask, 0.0, v_new)
        return v_new, a_new

    def step(self, u, v, a, d, f_ext, bc_mask=None, bc_vals=None):
        u_new = self.predictor(u, v, a, bc_mask, bc_vals)
        v_new, a_new = self.corrector(u_new, v, a, d, f_ext, bc_mask)
        return u_new, v_new, a_new


# ----------------------------------------------------------------------------------------
# a10-a13: Q4 AT2 damage operator and the projected / post-clamp CG solves
# ----------------------------------------------------------------------------------------

def projected_damage_active_mask(d, lb, r, upper=1.0, atol=1e-14, fixed=None):
    """solvers/damage_solver.py:109-124."""
    active = ((d <= lb + atol) & (r < 0)) | ((d >= upper - atol) & (r > 0))
    if fixed is not None:
        active = active | fixed
    return active


class PhaseFieldDamageSolver:
    """Native-Q4 AT2 CG damage solver (solvers/damage_solver.py:1157-1487, 1797-1838,
    2121-2160, 2270-2301, 3377-3775).  ``Gc_e`` is the optional per-element Gc field
    (the ``_Gc_l0_e`` / ``_Gc_over_l0_e`` branch, damage_solver.py:1412-1431)."""

    def __init__(self, fem: FEMOperators, tol=1e-5, max_iter=5000,
                 bounds_method="projected_cg", Gc_e: np.ndarray | None = None):
        self.fem, self.mesh = fem, fem.mesh
        self.tol, self.max_iter, self.bounds_method = tol, max_iter, bounds_method
        m = fem.material
        self.Gc_l0 = m.Gc * m.l0
        self.Gc_over_l0 = m.Gc / m.l0
        self.Gc_l0_e = None if Gc_e is None else Gc_e * m.l0
        self.Gc_over_l0_e = None if Gc_e is None else Gc_e / m.l0
        gp, w = self.mesh.grad, self.mesh.wdet
        self.diag_lap = (w[..., None] * (gp[..., 0] ** 2 + gp[..., 1] ** 2)).sum(axis=1)  # 1397-1402
        self.last_iter = 0
        self.last_status = ""

    def Ax(self, d, rho_q):
        """_Ax_q4 (damage_solver.py:1797-1838); rho_q = 2H_q + Gc/l0."""
        mesh = self.mesh
        N, gp, w = mesh.N_q, mesh.grad, mesh.wdet
        d_e = d[mesh.elements]
        gdx = np.einsum("eqa,ea->eq", gp[..., 0], d_e)
        gdy = np.einsum("eqa,ea->eq", gp[..., 1], d_e)
        lap = (w[..., None] * (gp[..., 0] * gdx[..., None] + gp[..., 1] * gdy[..., None])).sum(axis=1)
        if self.Gc_l0_e is not None:
            lap = self.Gc_l0_e[:, None] * lap
        out = scatter(mesh.elem_flat, lap.ravel(), mesh.n_nodes)
        d_q = np.einsum("qa,ea->eq", N, d_e)
        mass = np.einsum("qa,eq->ea", N, rho_q * d_q * w)
        react = scatter(mesh.elem_flat, mass.ravel(), mesh.n_nodes)
        if self.Gc_l0_e is not None:
            return out + react
        return out * self.Gc_l0 + react

    def prepare(self, H_q, d_prev):
        """_prepare_cg_q4 (damage_solver.py:2270-2301): rho_q, b, M_inv."""
        mesh = self.mesh
        ratio = self.Gc_over_l0_e[:, None] if self.Gc_over_l0_e is not None else self.Gc_over_l0
        rho = 2.0 * H_q + ratio
        N, w = mesh.N_q, mesh.wdet
        b = scatter(mesh.elem_flat, np.einsum("qa,eq->ea", N, 2.0 * H_q * w).ravel(), mesh.n_nodes)
        mass_diag = np.einsum("qa,eq->ea", N * N, rho * w)
        lap_diag = (self.Gc_l0_e[:, None] if self.Gc_l0_e is not None else self.Gc_l0) * self.diag_lap
        A_diag = scatter(mesh.elem_flat, (lap_diag + mass_diag).ravel(), mesh.n_nodes)
        return rho, b, 1.0 / (A_diag + 1e-30)

    def solve(self, H_q, d_prev, initial_guess=None, fixed=None, vals=None):
        """``solve`` -> ``_solve_dispatch`` (damage_solver.py:2493-2684)."""
        if not np.all(np.isfinite(d_prev)):
            self.last_iter = 0
            return d_prev.copy()
        if self.bounds_method == "projected_cg":
            return self._projected_cg(H_q, d_prev, initial_guess, fixed, vals)
        return self._post_clamp(H_q, d_prev, initial_guess, fixed, vals)

    def _start(self, H_q, d_prev, initial_guess, fixed, vals):
        rho, b, M_inv = self.prepare(H_q, d_prev)
        d = d_prev.copy() if initial_guess is None else initial_guess.copy()
        lb = d_prev.copy()
        if fixed is not None:
            d = np.where(fixed, vals, d)
            lb = np.where(fixed, vals, lb)
        return rho, b, M_inv, d, lb

    def _projected_cg(self, H_q, d_prev, initial_guess, fixed, vals):
        """_solve_projected_cg (damage_solver.py:3527-3775), non-deterministic-mode branch."""
        rho, b, M_inv, d, lb = self._start(H_q, d_prev, initial_guess, fixed, vals)
        ub = 1.0
        check_every = 10
        self.last_iter = self.max_iter
        self.last_status = "max_iter"
        r = b - self.Ax(d, rho)
        if fixed is not None:
            r[fixed] = 0.0
            M_inv = np.where(fixed, 0.0, M_inv)
        rr0 = float(r @ r)
        ref = max(math.sqrt(float(b @ b)), math.sqrt(rr0), 1e-30)
        tol_sq = (self.tol * ref) ** 2
        active = projected_damage_active_mask(d, lb, r, ub, fixed=fixed)
        r_free = np.where(active, 0.0, r)
        z = r_free * M_inv
        z[active] = 0.0
        p = z.copy()
        rz = float(r_free @ z)
        for i in range(self.max_iter):
            it = i + 1
            Ap = self.Ax(p, rho)
            pAp = float(p @ Ap)
            if pAp <= 0:
                self.last_iter, self.last_status = it, "nonpositive"
                break
            alpha_cg = rz / (pAp + 1e-30)
            alpha_max = alpha_cg
            pos = p > 0
            if pos.any():
                up = float(((ub - d[pos]) / p[pos]).min())
                if up < alpha_max:
                    alpha_max = up
            neg = p < 0
            if neg.any():
                dn = float(((lb[neg] - d[neg]) / p[neg]).min())
                if dn < alpha_max:
                    alpha_max = dn
            if alpha_max < alpha_cg and alpha_max >= 0:
                d = d + alpha_max * p
                d = np.minimum(np.maximum(d, lb), ub)
                r = b - self.Ax(d, rho)
                if fixed is not None:
                    r[fixed] = 0.0
                active = projected_damage_active_mask(d, lb, r, ub, fixed=fixed)
                r_free = np.where(active, 0.0, r)
                rfn = float(r_free @ r_free)
                if rfn < tol_sq:
                    self.last_iter, self.last_status = it, "converged"
                    break
                if rfn > 1e12 * max(rr0, 1e-30):
                    self.last_iter, self.last_status = it, "diverged"
                    break
                z = r_free * M_inv
                z[active] = 0.0
                p = z.copy()
                rz = float(r_free @ z)
                continue
            d = d + alpha_cg * p
            r = r - alpha_cg * Ap
            if it % check_every == 0:
                r = b - self.Ax(d, rho)
                if fixed is not None:
                    r[fixed] = 0.0
                old = active
                active = projected_damage_active_mask(d, lb, r, ub, fixed=fixed)
                changed = bool((active != old).any())
                r_free = np.where(active, 0.0, r)
                rfn = float(r_free @ r_free)
                if rfn < tol_sq:
                    self.last_iter, self.last_status = it, "converged"
                    break
                if rfn > 1e12 * max(rr0, 1e-30):
                    self.last_iter, self.last_status = it, "diverged"
                    break
                if changed:
                    z = r_free * M_inv
                    z[active] = 0.0
                    p = z.copy()
                    rz = float(r_free @ z)
                    continue
            else:
                r_free = np.where(active, 0.0, r)
            z_new = r_free * M_inv
            z_new[active] = 0.0
            rz_new = float(r_free @ z_new)
            beta = rz_new / (rz + 1e-30)
            p = p * beta + z_new
            p[active] = 0.0
            rz = rz_new
        self.last_active = active
        d = np.minimum(np.maximum(d, lb), ub)
        if fixed is not None:
            d = np.where(fixed, vals, d)
        return d

    def _post_clamp(self, H_q, d_prev, initial_guess, fixed, vals):
        """_solve_post_clamp (damage_solver.py:3377-3521), Jacobi branch."""
        rho, b, M_inv, d, lb = self._start(H_q, d_prev, initial_guess, fixed, vals)
        check_every = 50
        self.last_iter = self.max_iter
        self.last_status = "max_iter"
        r = b - self.Ax(d, rho)
        if fixed is not None:
            r[fixed] = 0.0
            M_inv = np.where(fixed, 0.0, M_inv)
        rr0 = float(r @ r)
        ref = max(math.sqrt(float(b @ b)), math.sqrt(rr0), 1e-30)
        tol_sq = (self.tol * ref) ** 2
        z = r * M_inv
        rz = float(r @ z)
        p = z.copy()
        for i in range(self.max_iter):
            it = i + 1
            Ap = self.Ax(p, rho)
            pAp = float(p @ Ap)
            if pAp <= 0:
                self.last_iter, self.last_status = it, "nonpositive"
                break
            alpha = rz / (pAp + 1e-30)
            d = d + alpha * p
            r = r - alpha * Ap
            if it % check_every == 0:
                r = b - self.Ax(d, rho)
                if fixed is not None:
                    r[fixed] = 0.0
            rr = float(r @ r)
            if rr < tol_sq:
                self.last_iter, self.last_status = it, "converged"
                break
            if rr > 1e12 * max(rr0, 1e-30):
                self.last_iter, self.last_status = it, "diverged"
                break
            z = r * M_inv
            rz_new = float(r @ z)
            beta = rz_new / (rz + 1e-30)
            p = p * beta + z
            rz = rz_new
        d = np.clip(np.maximum(d, lb), 0.0, 1.0)
        if fixed is not None:
            d = np.where(fixed, vals, d)
        return d


# ----------------------------------------------------------------------------------------
# a14: one explicit staggered step
# ----------------------------------------------------------------------------------------

@dataclass
class ExplicitState:
    u: np.ndarray
    v: np.ndarray
    a: np.ndarray
    d: np.ndarray
    H: np.ndarray
    count: int = 0


def step_full_explicit(dyn: ExplicitDynamics, dmg: PhaseFieldDamageSolver, st: ExplicitState,
                       f_ext, bc_mask=None, bc_vals=None, damage_every: int = 1) -> np.ndarray:
    """Explicit branch of StaggeredSolver.step_full (solvers/staggered_solver.py:1115-1165):
    mechanics step, strain, psi+ and H = max(H, psi+), then the gated damage solve."""
    st.u, st.v, st.a = dyn.step(st.u, st.v, st.a, st.d, f_ext, bc_mask, bc_vals)
    fem = dyn.fem
    strain = fem.compute_strain(st.u)
    psi = fem.compute_driving_force(st.u, strain=strain)
    st.H = history_update(st.H, psi)
    st.count += 1
    if damage_every <= 1 or st.count <= 5 or st.count % damage_every == 0:
        st.d = dmg.solve(st.H, st.d)
    return psi
