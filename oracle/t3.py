"""Oracle (numpy, float64) for the reference's linear-triangle (T3) path.

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.  Imported by ``tests/`` and
``bench.py``'s ``cpu_baseline`` leg; never by the product.

The functions restate, op for op, the reference files cited in each docstring (paths
relative to the reference's ``src/phast``).  Scatter-adds use ``np.bincount`` (sequential
accumulation in index order, like ``scatter_add_`` on CPU).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .q4 import Material, PhaseFieldDamageSolver

SPECTRAL_EPS = 1e-12  # core/fem_operators.py:126 (float64)


def scatter(idx: np.ndarray, vals: np.ndarray, n: int) -> np.ndarray:
    return np.bincount(idx, weights=vals, minlength=n).astype(np.float64)


class T3Mesh:
    """T3 precompute of ``FEMMesh._precompute_t3`` (core/mesh.py:226-300)."""

    def __init__(self, nodes: np.ndarray, elements: np.ndarray):
        self.nodes = np.asarray(nodes, dtype=np.float64)
        self.elements = np.asarray(elements, dtype=np.int64)
        self.n_nodes, self.n_elems = len(self.nodes), len(self.elements)
        p0, p1, p2 = (self.nodes[self.elements[:, k]] for k in range(3))
        v01, v02, v12 = p1 - p0, p2 - p0, p2 - p1
        s2a = v01[:, 0] * v02[:, 1] - v01[:, 1] * v02[:, 0]
        self.areas = 0.5 * np.abs(s2a)
        sgn = np.sign(s2a)
        sgn[sgn == 0] = 1.0
        inv = 1.0 / np.where(np.abs(s2a) > 1e-10, s2a, sgn * 1e-10)
        gp = np.zeros((self.n_elems, 3, 2))
        gp[:, 0, 0] = (p1[:, 1] - p2[:, 1]) * inv
        gp[:, 0, 1] = (p2[:, 0] - p1[:, 0]) * inv
        gp[:, 1, 0] = (p2[:, 1] - p0[:, 1]) * inv
        gp[:, 1, 1] = (p0[:, 0] - p2[:, 0]) * inv
        gp[:, 2, 0] = (p0[:, 1] - p1[:, 1]) * inv
        gp[:, 2, 1] = (p1[:, 0] - p0[:, 0]) * inv
        self.grad_phi = gp
        per = np.linalg.norm(v01, axis=1) + np.linalg.norm(v12, axis=1) + np.linalg.norm(v02, axis=1)
        self.elem_h = 4.0 * self.areas / (per + 1e-30)
        self.h_min = float(self.elem_h.min())
        self.elem_flat = self.elements.reshape(-1)
        self.M_scalar = scatter(self.elem_flat, np.repeat(self.areas / 3.0, 3), self.n_nodes)

    def elem_to_node(self, field_e: np.ndarray) -> np.ndarray:
        """Area-weighted projection of an element field (core/mesh.py:430-433)."""
        ns = scatter(self.elem_flat, np.repeat(field_e * self.areas, 3), self.n_nodes)
        return ns / (3.0 * self.M_scalar + 1e-30)


def structured_t3(nx: int, ny: int, W: float, H: float, jitter: float = 0.0, seed: int = 0):
    """A rectangle of nx x ny cells, each split into two triangles along alternating
    diagonals, nodes numbered row by row; interior nodes optionally jittered."""
    rng = np.random.default_rng(seed)
    xs, ys = np.linspace(0.0, W, nx + 1), np.linspace(0.0, H, ny + 1)
    X, Y = np.meshgrid(xs, ys)
    nodes = np.stack([X.ravel(), Y.ravel()], axis=1)
    if jitter:
        inner = np.ones((ny + 1, nx + 1), dtype=bool)
        inner[[0, -1], :] = False
        inner[:, [0, -1]] = False
        h = min(W / nx, H / ny)
        nodes[inner.ravel()] += rng.uniform(-jitter, jitter, (int(inner.sum()), 2)) * h
    els = []
    for j in range(ny):
        for i in range(nx):
            a = j * (nx + 1) + i
            b, c, d = a + 1, a + nx + 2, a + nx + 1
            if (i + j) % 2 == 0:
                els += [(a, b, c), (a, c, d)]
            else:
                els += [(a, b, d), (b, c, d)]
    return nodes, np.asarray(els, dtype=np.int64)


class T3Operators:
    """T3 branch of ``FEMOperators`` (core/fem_operators.py)."""

    def __init__(self, mesh: T3Mesh, material: Material):
        self.mesh, self.material = mesh, material
        self.M_vec = material.rho * np.repeat(mesh.M_scalar, 2)       # fem_operators.py:85
        self.M_vec_inv = 1.0 / (self.M_vec + 1e-30)                   # :86
        self.dt_cfl = mesh.h_min / material.c_p                        # :89

    def compute_strain(self, u: np.ndarray):
        """Centroid strain (fem_operators.py:223-242)."""
        gp = self.mesh.grad_phi
        u_e = u[self.mesh.elements]
        exx = (gp[:, :, 0] * u_e[:, :, 0]).sum(1)
        eyy = (gp[:, :, 1] * u_e[:, :, 1]).sum(1)
        gxy = (gp[:, :, 1] * u_e[:, :, 0]).sum(1) + (gp[:, :, 0] * u_e[:, :, 1]).sum(1)
        return exx, eyy, gxy

    def degradation_e(self, d: np.ndarray) -> np.ndarray:
        """g of the element-mean damage (fem_operators.py:628-630)."""
        return self.material.degradation(d[self.mesh.elements].mean(1))

    def compute_stress(self, u, d, strain=None):
        """Split dispatch (fem_operators.py:614-653)."""
        g = self.degradation_e(d)
        exx, eyy, gxy = self.compute_strain(u) if strain is None else strain
        m = self.material
        lam, mu, kappa = m.lam, m.mu, m.kappa
        split = m.stress_split  # fem_operators.py:24-33, 641-653
        if split == "isotropic":  # fem_operators.py:277-290
            c11 = lam + 2.0 * mu
            return g * (c11 * exx + lam * eyy), g * (lam * exx + c11 * eyy), g * (mu * gxy)
        if split == "amor":  # fem_operators.py:301-333
            tr2 = exx + eyy
            ezz = -m.nu / (1.0 - m.nu) * tr2 if m.plane_stress else 0.0
            tr = tr2 + ezz
            trp = np.maximum(tr, 0.0)
            trm = tr - trp
            dxx, dyy = exx - tr / 3.0, eyy - tr / 3.0
            sxx = g * (kappa * trp + 2 * mu * dxx) + kappa * trm
            syy = g * (kappa * trp + 2 * mu * dyy) + kappa * trm
            return sxx, syy, g * mu * gxy
        if split == "spectral":  # fem_operators.py:335-390
            exy = gxy / 2.0
            trace = exx + eyy
            hd = (exx - eyy) / 2.0
            delta = np.sqrt(hd ** 2 + exy ** 2 + SPECTRAL_EPS)
            e1, e2 = trace / 2.0 + delta, trace / 2.0 - delta
            e1p, e2p = np.maximum(e1, 0.0), np.maximum(e2, 0.0)
            trp = np.maximum(trace, 0.0)
            trm = trace - trp
            inv = 1.0 / (2.0 * delta + 1e-30)
            p_xx, p_yy, p_xy = (hd + delta) * inv, (-hd + delta) * inv, exy * inv
            exx_p = e1p * p_xx + e2p * (1.0 - p_xx)
            eyy_p = e1p * p_yy + e2p * (1.0 - p_yy)
            exy_p = (e1p - e2p) * p_xy
            exx_m, eyy_m, exy_m = exx - exx_p, eyy - eyy_p, exy - exy_p
            sxx = g * (lam * trp + 2 * mu * exx_p) + (lam * trm + 2 * mu * exx_m)
            syy = g * (lam * trp + 2 * mu * eyy_p) + (lam * trm + 2 * mu * eyy_m)
            sxy = g * (2 * mu * exy_p) + (2 * mu * exy_m)
            return sxx, syy, sxy
        raise ValueError(split)

    def assemble_force(self, sxx, syy, sxy):
        """Nodal force (fem_operators.py:661-679)."""
        gp, A = self.mesh.grad_phi, self.mesh.areas[:, None]
        fx = A * (sxx[:, None] * gp[:, :, 0] + sxy[:, None] * gp[:, :, 1])
        fy = A * (sxy[:, None] * gp[:, :, 0] + syy[:, None] * gp[:, :, 1])
        n, idx = self.mesh.n_nodes, self.mesh.elem_flat
        return np.stack([scatter(idx, fx.ravel(), n), scatter(idx, fy.ravel(), n)], axis=1)

    def internal_force(self, u, d):
        """fem_operators.py:598-612."""
        strain = self.compute_strain(u)
        return self.assemble_force(*self.compute_stress(u, d, strain))

    def compute_psi_plus(self, u, strain=None):
        """Tensile energy density at the centroids (fem_operators.py:951-1030)."""
        exx, eyy, gxy = self.compute_strain(u) if strain is None else strain
        m = self.material
        lam, mu, kappa = m.lam, m.mu, m.kappa
        exy = gxy / 2.0
        if m.energy_split == "isotropic":  # :983-989
            tr = exx + eyy
            return 0.5 * lam * tr ** 2 + mu * (exx ** 2 + eyy ** 2 + 2 * exy ** 2)
        if m.energy_split == "amor":  # :991-1013
            tr2 = exx + eyy
            ezz = -m.nu / (1.0 - m.nu) * tr2 if m.plane_stress else np.zeros_like(exx)
            tr = tr2 + ezz
            trp = np.maximum(tr, 0.0)
            dxx, dyy, dzz = exx - tr / 3.0, eyy - tr / 3.0, ezz - tr / 3.0
            return 0.5 * kappa * trp ** 2 + mu * (dxx ** 2 + dyy ** 2 + dzz ** 2 + 2.0 * exy ** 2)
        if m.energy_split == "spectral":  # :1015-1030
            trace = exx + eyy
            delta = np.sqrt(((exx - eyy) / 2.0) ** 2 + exy ** 2 + SPECTRAL_EPS)
            e1p, e2p = np.maximum(trace / 2.0 + delta, 0.0), np.maximum(trace / 2.0 - delta, 0.0)
            trp = np.maximum(trace, 0.0)
            return 0.5 * lam * trp ** 2 + mu * (e1p ** 2 + e2p ** 2)
        raise ValueError(m.energy_split)

    def compute_driving_force(self, u, strain=None, d=None):
        """'strain_energy' driving force == psi+ (fem_operators.py:1173-1204)."""
        return self.compute_psi_plus(u, strain)


class T3Damage:
    """AT2 / AT1 damage operator of the element-H path (solvers/damage_solver.py).  ``Gc_e`` is
    the optional per-element Gc field (the ``_Gc_l0_e`` / ``_Gc_over_l0_e`` branch,
    damage_solver.py:1412-1433, 591-611).  AT1: 3/4 Gc l0 on the Laplacian, no Gc/l0 in the
    reaction, and the source S_H in the right-hand side (damage_solver.py:1321-1336)."""

    def __init__(self, mesh: T3Mesh, material: Material, Gc_e: np.ndarray | None = None):
        self.mesh, self.material = mesh, material
        gp = mesh.grad_phi
        self.diag_lap = mesh.areas[:, None] * (gp[:, :, 0] ** 2 + gp[:, :, 1] ** 2)  # :1404-1406
        at1 = getattr(material, "pf_model", "AT2") == "AT1"
        if at1:
            self.Gc_l0 = 0.75 * material.Gc * material.l0
            self.Gc_over_l0 = 0.0
            self.at1_source = material.at1_source
        else:
            self.Gc_l0 = material.Gc * material.l0
            self.Gc_over_l0 = material.Gc / material.l0
            self.at1_source = 0.0
        self.at1_source_e = None
        if Gc_e is None:
            self.Gc_l0_e = self.Gc_over_l0_e = None
        elif at1:
            self.Gc_l0_e = 0.75 * Gc_e * material.l0
            self.Gc_over_l0_e = np.zeros_like(Gc_e)
            self.at1_source_e = 3.0 * Gc_e / (8.0 * material.l0)
        else:
            self.Gc_l0_e = Gc_e * material.l0
            self.Gc_over_l0_e = Gc_e / material.l0

    def reaction_coeff(self, H):
        """(2 H + Gc/l0) area / 12 (damage_solver.py:2168-2177, no viscosity)."""
        ratio = self.Gc_over_l0 if self.Gc_over_l0_e is None else self.Gc_over_l0_e
        return (2.0 * H + ratio) * self.mesh.areas / 12.0

    def rhs(self, H):
        """b = scatter((2 H - S_H) area/3) (damage_solver.py:2249-2258; S_H = 0 for AT2)."""
        m = self.mesh
        src = self.at1_source if self.at1_source_e is None else self.at1_source_e
        return scatter(m.elem_flat, np.repeat((2.0 * H - src) * (m.areas / 3.0), 3), m.n_nodes)

    def diagonal(self, H):
        """Jacobi diagonal Gc l0 diag_lap + 2 rc (damage_solver.py:2112-2119)."""
        m = self.mesh
        lap = self.Gc_l0 * self.diag_lap if self.Gc_l0_e is None else self.Gc_l0_e[:, None] * self.diag_lap
        contrib = lap + 2.0 * self.reaction_coeff(H)[:, None]
        return scatter(m.elem_flat, contrib.ravel(), m.n_nodes)

    def apply(self, d, H):
        """(Gc l0 K_lap + M_react) d (damage_solver.py:1719-1755)."""
        m = self.mesh
        d_e = d[m.elements]
        gp = m.grad_phi
        gdx, gdy = (gp[:, :, 0] * d_e).sum(1), (gp[:, :, 1] * d_e).sum(1)
        lap = m.areas[:, None] * (gp[:, :, 0] * gdx[:, None] + gp[:, :, 1] * gdy[:, None])
        if self.Gc_l0_e is not None:
            lap = self.Gc_l0_e[:, None] * lap
        out = scatter(m.elem_flat, lap.ravel(), m.n_nodes)
        rc = self.reaction_coeff(H)
        react = scatter(m.elem_flat, (rc[:, None] * (d_e + d_e.sum(1)[:, None])).ravel(), m.n_nodes)
        if self.Gc_l0_e is not None:
            return out + react
        return self.Gc_l0 * out + react

    def dense(self, H):
        """Dense matrix of ``apply`` (columns of the identity) for checks on small meshes."""
        n = self.mesh.n_nodes
        return np.stack([self.apply(np.eye(n)[k], H) for k in range(n)], axis=1)


class T3DamageSolver(PhaseFieldDamageSolver):
    """Projected / post-clamp Jacobi CG of ``PhaseFieldDamageSolver`` on the T3 operator:
    the solve loops (damage_solver.py:3377-3775) are the element-type independent ones of
    the Q4 oracle; ``Ax`` is ``_Ax`` T3 and ``prepare`` is ``_prepare_cg`` for element H
    (damage_solver.py:2160-2268: reaction coefficient, b, M_inv = 1/(diag + 1e-30))."""

    def __init__(self, mesh: T3Mesh, material: Material, tol=1e-5, max_iter=5000,
                 bounds_method="projected_cg", Gc_e: np.ndarray | None = None):
        self.mesh = mesh
        self.op = T3Damage(mesh, material, Gc_e)
        self.tol, self.max_iter, self.bounds_method = tol, max_iter, bounds_method
        self.last_iter = 0
        self.last_status = ""

    def Ax(self, d, H):
        return self.op.apply(d, H)

    def prepare(self, H, d_prev):
        return H, self.op.rhs(H), 1.0 / (self.op.diagonal(H) + 1e-30)
