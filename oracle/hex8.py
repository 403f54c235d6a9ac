"""Oracle (numpy, float64) for the reference's periodic hex8 voxel path.

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

Restates, with the reference's own data layout (local H8 node order, engineering-strain
``B (8,6,24)`` matrices, element dof gather and scatter-add), the files
``solid3d/periodic.py``, ``core/hex8.py``, ``solid3d/operator.py``, ``solid3d/splits.py``,
``solid3d/linalg.py`` and ``solvers/box_kkt_h8.py`` (paths relative to ``src/phast``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

H8_SIGNS = np.array([(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1),
                     (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)], dtype=np.float64)  # core/hex8.py:16-36


def h8_gauss_points():
    """Tensor Gauss points, xi fastest, weights 1 (core/hex8.py:157-179)."""
    a = 1.0 / math.sqrt(3.0)
    p = np.array([-a, a])
    zeta, eta, xi = np.meshgrid(p, p, p, indexing="ij")
    return np.stack((xi, eta, zeta), axis=-1).reshape(-1, 3), np.ones(8)


def reference_voxel(h):
    """N (8q,8a), gradN (8,8,3), w (8), B (8,6,24), M, L (solid3d/periodic.py:193-221; hex8.py:182-293)."""
    hx, hy, hz = h
    pts, wts = h8_gauss_points()
    coords = (H8_SIGNS + 1.0) / 2.0 * np.array([hx, hy, hz])
    N = np.zeros((8, 8))
    grad = np.zeros((8, 8, 3))
    w = np.zeros(8)
    for q in range(8):
        xi, eta, zeta = pts[q]
        s = H8_SIGNS
        N[q] = 0.125 * (1 + s[:, 0] * xi) * (1 + s[:, 1] * eta) * (1 + s[:, 2] * zeta)
        dN = np.stack([0.125 * s[:, 0] * (1 + s[:, 1] * eta) * (1 + s[:, 2] * zeta),
                       0.125 * (1 + s[:, 0] * xi) * s[:, 1] * (1 + s[:, 2] * zeta),
                       0.125 * (1 + s[:, 0] * xi) * (1 + s[:, 1] * eta) * s[:, 2]], axis=-1)
        J = np.einsum("ar,aj->rj", dN, coords)
        grad[q] = np.einsum("ar,jr->aj", dN, np.linalg.inv(J))
        w[q] = wts[q] * np.linalg.det(J)
    B = np.zeros((8, 6, 24))
    for a in range(8):
        gx, gy, gz = grad[:, a, 0], grad[:, a, 1], grad[:, a, 2]
        d = 3 * a
        B[:, 0, d], B[:, 1, d + 1], B[:, 2, d + 2] = gx, gy, gz
        B[:, 3, d + 1], B[:, 3, d + 2] = gz, gy
        B[:, 4, d], B[:, 4, d + 2] = gz, gx
        B[:, 5, d], B[:, 5, d + 1] = gy, gx
    M = np.einsum("q,qa,qb->ab", w, N, N)
    L = np.einsum("q,qaj,qbj->ab", w, grad, grad)
    return N, grad, w, B, 0.5 * (M + M.T), 0.5 * (L + L.T)


def periodic_map(shape):
    """element_nodes (E,8) in local H8 order and element_dofs (E,24) (solid3d/periodic.py:61-108)."""
    nx, ny, nz = shape
    ke, je, ie = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    off = ((H8_SIGNS + 1) / 2).astype(np.int64)
    i = (ie.reshape(-1, 1) + off[:, 0]) % nx
    j = (je.reshape(-1, 1) + off[:, 1]) % ny
    k = (ke.reshape(-1, 1) + off[:, 2]) % nz
    nodes = i + nx * (j + ny * k)
    dofs = (3 * nodes[..., None] + np.arange(3)).reshape(-1, 24)
    return nodes, dofs


def scatter(index, values, n):
    return np.bincount(index.ravel(), weights=values.ravel(), minlength=n).astype(np.float64)


def isotropic_stiffness_3d(E, nu):
    """core/hex8.py:296-320."""
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[0, 0] += 2 * mu
    C[1, 1] += 2 * mu
    C[2, 2] += 2 * mu
    C[3, 3] = C[4, 4] = C[5, 5] = mu
    return C


def symmetric_eigvalsh_3x3(A):
    """Closed-form ascending eigenvalues (solid3d/splits.py:93-125)."""
    a11, a22, a33 = A[..., 0, 0], A[..., 1, 1], A[..., 2, 2]
    a12, a13, a23 = A[..., 0, 1], A[..., 0, 2], A[..., 1, 2]
    q = (a11 + a22 + a33) / 3.0
    b11, b22, b33 = a11 - q, a22 - q, a33 - q
    p1 = a12 ** 2 + a13 ** 2 + a23 ** 2
    p = np.sqrt((b11 ** 2 + b22 ** 2 + b33 ** 2 + 2.0 * p1) / 6.0)
    det_b = b11 * (b22 * b33 - a23 ** 2) - a12 * (a12 * b33 - a23 * a13) + a13 * (a12 * a23 - b22 * a13)
    iso = p <= np.finfo(np.float64).tiny
    ps = np.where(iso, 1.0, p)
    r = np.clip(det_b / (2.0 * ps ** 3), -1.0, 1.0)
    phi = np.arccos(r) / 3.0
    largest = q + 2.0 * p * np.cos(phi)
    smallest = q + 2.0 * p * np.cos(phi + 2.0 * math.pi / 3.0)
    middle = 3.0 * q - largest - smallest
    out = np.stack((smallest, middle, largest), axis=-1)
    return np.where(iso[..., None], q[..., None], out)


def engineering_strain_to_tensor(e):
    """core/voigt.py:12-30 (xx,yy,zz,yz,xz,xy; halved shears)."""
    T = np.zeros(e.shape[:-1] + (3, 3))
    T[..., 0, 0], T[..., 1, 1], T[..., 2, 2] = e[..., 0], e[..., 1], e[..., 2]
    T[..., 1, 2] = T[..., 2, 1] = 0.5 * e[..., 3]
    T[..., 0, 2] = T[..., 2, 0] = 0.5 * e[..., 4]
    T[..., 0, 1] = T[..., 1, 0] = 0.5 * e[..., 5]
    return T


def spectral_tensile_energy(eps, lam, mu, method="analytic"):
    """Miehe tensile energy (solid3d/splits.py:146-161)."""
    T = engineering_strain_to_tensor(eps)
    ev = symmetric_eigvalsh_3x3(T) if method == "analytic" else np.linalg.eigvalsh(T)
    trp = np.maximum(eps[..., :3].sum(axis=-1), 0.0)
    return 0.5 * lam * trp ** 2 + mu * (np.maximum(ev, 0.0) ** 2).sum(axis=-1)


@dataclass
class VoxelOperator:
    """Restatement of PeriodicVoxelPhaseFieldOperator, AT2, no closure (solid3d/operator.py)."""
    shape: tuple
    h: tuple
    bank: np.ndarray       # (K,6,6)
    index: np.ndarray      # (E,)
    Gc: np.ndarray         # (E,)
    lam: np.ndarray
    mu: np.ndarray
    driver: np.ndarray     # (E,) int
    kappa: float
    ell: float
    pinned_node: int = 0
    controlled: int = 0
    fixed_zero: tuple = ()

    def __post_init__(self):
        self.nodes, self.dofs = periodic_map(self.shape)
        self.N, self.grad, self.w, self.B, M, L = reference_voxel(self.h)
        self.B_flat = self.B.transpose(2, 0, 1).reshape(24, 48)   # operator.py:242
        self.K_damage = M / self.ell + self.ell * L                # operator.py:246
        self.n_red = int(np.prod(self.shape))
        self.n_mech = 3 * self.n_red + 6
        self.E = self.n_red
        hx, hy, hz = self.h
        self.volume = float(np.prod(self.shape) * hx * hy * hz)
        self.damage_scale = self.ell / (float(self.Gc.min()) * hx * hy * hz)
        self.c_history = 2.0 * (1.0 - self.kappa)

    def degradation(self, d):
        """operator.py:939-946; splits.py:56-59."""
        d_q = d[self.nodes] @ self.N.T
        return (1.0 - self.kappa) * (1.0 - d_q) ** 2 + self.kappa

    def qp_strain(self, x):
        u, E = x[:3 * self.n_red], x[3 * self.n_red:]
        return (u[self.dofs] @ self.B_flat).reshape(-1, 8, 6) + E  # operator.py:948-952

    def stress_pass(self, x, g=None, assemble=True):
        """operator.py:1108-1169 (one batch): r_u, r_E, energy."""
        eps = self.qp_strain(x)
        C = self.bank[self.index]
        sigma = np.einsum("eqm,ekm->eqk", eps, C)
        if g is not None:
            sigma = g[..., None] * sigma
        weighted = sigma * self.w[None, :, None]
        r_E = weighted.sum(axis=(0, 1))
        energy = 0.5 * (weighted * eps).sum()
        r_u = None
        if assemble:
            local = weighted.reshape(-1, 48) @ self.B_flat.T
            r_u = scatter(self.dofs, local, 3 * self.n_red)
        return r_u, r_E, energy

    def mech_action(self, x, g=None):
        r_u, r_E, _ = self.stress_pass(x, g)
        return np.concatenate((r_u, r_E))

    def mech_diagonal(self, g=None):
        """operator.py:1196-1254."""
        BB = np.einsum("qia,qja->qaij", self.B, self.B).reshape(8 * 24, 36)
        C = self.bank[self.index]
        gw = np.broadcast_to(self.w, (self.E, 8)) if g is None else g * self.w
        per_qp = (C.reshape(self.E, 36) @ BB.T).reshape(self.E, 8, 24)
        local = (gw[..., None] * per_qp).sum(axis=1)
        diag_u = scatter(self.dofs, local, 3 * self.n_red)
        diag_E = (gw.sum(axis=1)[:, None] * np.diagonal(C, axis1=1, axis2=2)).sum(axis=0)
        return np.concatenate((diag_u, diag_E))

    def macro_stress(self, x, g=None):
        return self.stress_pass(x, g, assemble=False)[1] / self.volume

    def elastic_energy(self, x, g=None):
        return self.stress_pass(x, g, assemble=False)[2]

    def driving_force(self, x, method="analytic"):
        """operator.py:1415-1444 (drivers none / full / spectral)."""
        eps = self.qp_strain(x)
        Y = np.zeros((self.E, 8))
        full = self.driver == 1
        if full.any():
            C = self.bank[self.index[full]][:, None]
            Y[full] = 0.5 * (eps[full] * np.einsum("eqkm,eqm->eqk", np.broadcast_to(C, (C.shape[0], 8, 6, 6)),
                                                   eps[full])).sum(-1)
        spec = self.driver == 2
        if spec.any():
            Y[spec] = spectral_tensile_energy(eps[spec], self.lam[spec, None], self.mu[spec, None], method)
        return Y

    def damage_action(self, d, H=None):
        """operator.py:1450-1466."""
        d_e = d[self.nodes]
        y = self.Gc[:, None] * (d_e @ self.K_damage)
        if H is not None:
            d_q = d_e @ self.N.T
            y = y + self.c_history * ((H * self.w) * d_q) @ self.N
        return scatter(self.nodes, y, self.n_red)

    def damage_rhs(self, H):
        return scatter(self.nodes, self.c_history * (H * self.w) @ self.N, self.n_red)

    def damage_diagonal(self, H=None):
        y = self.Gc[:, None] * np.diagonal(self.K_damage)[None, :]
        if H is not None:
            y = y + self.c_history * (H * self.w) @ (self.N ** 2)
        return scatter(self.nodes, y, self.n_red)

    def mech_fixed_mask(self):
        m = np.zeros(self.n_mech, dtype=bool)
        m[3 * self.pinned_node:3 * self.pinned_node + 3] = True
        m[3 * self.n_red + self.controlled] = True
        for c in self.fixed_zero:
            m[3 * self.n_red + c] = True
        return m


# ---------------------------------------------------------------------------------------
# Adjoint pieces of the converged staggered state.  The reference has no adjoint for the
# voxel path; these are the transposed derivatives of its residuals (operator.py:939-946,
# 1108-1194, 1415-1509), written from the mathematics and pinned by finite differences of
# the functions above.
# ---------------------------------------------------------------------------------------

def field_snapshot_elements(op: VoxelOperator, x, g=None):
    """Element Gauss-point means of a field snapshot (solid3d/operator.py:1589-1600): strain,
    degraded stress g C eps and 0.5 g eps . C eps; g None = 1."""
    eps = op.qp_strain(x)
    sigma0 = np.einsum("eqm,ekm->eqk", eps, op.bank[op.index])
    gg = np.ones(eps.shape[:2]) if g is None else g
    return eps.mean(axis=1), (gg[..., None] * sigma0).mean(axis=1), (0.5 * gg * (eps * sigma0).sum(-1)).mean(axis=1)


def driving_strain_gradient(op: VoxelOperator, eps):
    """dY/d(engineering strain) per Gauss point (E,8,6): drivers none / full (C eps) /
    spectral (lam <tr>+ I + 2 mu eps+, shear entries the tensor off-diagonals)."""
    T = np.zeros(eps.shape)
    full = op.driver == 1
    if full.any():
        C = op.bank[op.index[full]]
        T[full] = np.einsum("ekm,eqm->eqk", C, eps[full])
    spec = op.driver == 2
    if spec.any():
        e = eps[spec]
        Et = engineering_strain_to_tensor(e)
        val, vec = np.linalg.eigh(Et)
        pos = np.einsum("...ia,...a,...ja->...ij", vec, np.maximum(val, 0.0), vec)
        trp = np.maximum(e[..., :3].sum(-1), 0.0)
        lam, mu = op.lam[spec, None], op.mu[spec, None]
        Tt = 2.0 * mu[..., None, None] * pos
        Tt[..., 0, 0] += lam * trp
        Tt[..., 1, 1] += lam * trp
        Tt[..., 2, 2] += lam * trp
        T[spec] = np.stack((Tt[..., 0, 0], Tt[..., 1, 1], Tt[..., 2, 2], Tt[..., 1, 2], Tt[..., 0, 2],
                            Tt[..., 0, 1]), axis=-1)
    return T


def adjoint_coupling(op: VoxelOperator, x, lam_x, d):
    """(dF/dd)^T lam for F = A(g(d)) x: sum_q N_a w g'(d_q) eps(lam)_q . C eps(x)_q (n,)."""
    d_q = d[op.nodes] @ op.N.T
    gp = -2.0 * (1.0 - op.kappa) * (1.0 - d_q)
    ex, el = op.qp_strain(x), op.qp_strain(lam_x)
    C = op.bank[op.index]
    s = np.einsum("eqk,ekm,eqm->eq", el, C, ex) * gp * op.w
    return scatter(op.nodes, s @ op.N, op.n_red)


def qp_stress_apply(op: VoxelOperator, S):
    """sum_q B_q^T S_q scattered, plus the E rows sum_q S_q (no quadrature weight)."""
    local = S.reshape(-1, 48) @ op.B_flat.T
    return np.concatenate((scatter(op.dofs, local, 3 * op.n_red), S.sum(axis=(0, 1))))


def adjoint_driving(op: VoxelOperator, x, t):
    """(dY/dx)^T t for t (E,8): sum_q B_q^T t_q dY/deps_q (mech space incl. E rows)."""
    T = driving_strain_gradient(op, op.qp_strain(x))
    return qp_stress_apply(op, t[..., None] * T)


def adjoint_damage_history(op: VoxelOperator, d, lam_d):
    """(dG/dH)^T lam for G = A_H d - b_H (unscaled): c_h w (N_q.d_e - 1)(N_q.lam_e) (E,8)."""
    d_q = d[op.nodes] @ op.N.T
    l_q = lam_d[op.nodes] @ op.N.T
    return op.c_history * op.w * (d_q - 1.0) * l_q


def adjoint_damage_gc(op: VoxelOperator, d, lam_d):
    """(dG/dGc)^T lam (unscaled): lam_e . (M/ell + ell L) d_e per voxel (E,)."""
    return np.einsum("ea,ab,eb->e", lam_d[op.nodes], op.K_damage, d[op.nodes])


def qp_to_node(op: VoxelOperator, f):
    """sum over the voxels of a node of sum_q N_a(q) f_q (n,)."""
    return scatter(op.nodes, f @ op.N, op.n_red)


class PCGBreakdownError(RuntimeError):
    pass


def pcg(action, rhs, diag, x0=None, rtol=1e-8, atol=0.0, maxiter=5000, check_every=10):
    """solid3d/linalg.py:99-220 (flag semantics)."""
    b_norm = float(np.linalg.norm(rhs))
    if not np.all(np.isfinite(rhs)):
        raise PCGBreakdownError("PCG right-hand side contains a non-finite value")
    if not np.all(diag > 0):
        raise PCGBreakdownError("PCG diagonal has a non-positive entry (non-positive curvature)")
    if b_norm == 0.0:
        return np.zeros_like(rhs), (0, True, 0.0, 0.0, 0.0)
    thr = max(rtol * b_norm, atol)
    if x0 is None:
        x, r, r_norm = np.zeros_like(rhs), rhs.copy(), b_norm
    else:
        x = x0.copy()
        r = rhs - action(x)
        r_norm = float(np.linalg.norm(r))
    if r_norm <= thr:
        return x, (0, True, b_norm, r_norm, r_norm / b_norm)
    inv = 1.0 / diag
    z = r * inv
    p = z.copy()
    rz = float(r @ z)
    active, breakdown, done, rr = True, False, 0, float(r @ r)
    for it in range(1, maxiter + 1):
        Ap = action(p)
        pAp = float(p @ Ap)
        positive = pAp > 0.0
        breakdown = breakdown or (active and not positive)
        step = active and positive
        alpha = rz / pAp if step else 0.0
        x = x + alpha * p
        r = r - alpha * Ap
        done += int(step)
        z = r * inv
        rz_new = float(r @ z)
        rr = float(r @ r)
        active = active and rr > thr * thr
        beta = rz_new / rz if (active and rz > 0.0) else 0.0
        p = z + beta * p
        rz = rz_new
        if it % check_every == 0 or it == maxiter:
            if breakdown:
                raise PCGBreakdownError(f"PCG met non-positive curvature by iteration {it}")
            r_norm = math.sqrt(rr)
            if not active:
                return x, (done, True, b_norm, r_norm, r_norm / b_norm)
    return x, (maxiter, False, b_norm, r_norm, r_norm / b_norm)


def solve_constrained(action, diag, fixed_mask, fixed_values, rhs=None, x0=None, **kw):
    """solid3d/linalg.py:223-287."""
    free = ~fixed_mask
    x_c = np.where(fixed_mask, fixed_values, 0.0)
    image = action(x_c)
    b = -image if rhs is None else rhs - image
    b_f = np.where(free, b, 0.0)
    diag_f = np.where(free, diag, 1.0)
    x0_f = None if x0 is None else np.where(free, x0, 0.0)
    y, info = pcg(lambda v: np.where(free, action(v), 0.0), b_f, diag_f, x0=x0_f, **kw)
    return np.where(free, y, x_c), info


def scalar_pcg(action, rhs, diag, initial, tol, max_iter):
    """solvers/box_kkt_h8.py:256-347 on the reduced (free) system."""
    x = np.zeros_like(rhs) if initial is None else initial.copy()
    res = rhs - action(x)
    n0 = float(np.linalg.norm(res))
    ref = max(n0, 1.0)
    if n0 <= tol * ref:
        return x, 0, True
    pre = res / diag
    dr = pre.copy()
    rz = float(res @ pre)
    if not rz > 0:
        raise RuntimeError("PCG encountered a non-positive initial residual product")
    for it in range(1, max_iter + 1):
        im = action(dr)
        den = float(dr @ im)
        if not den > 0:
            raise RuntimeError("PCG operator is not positive definite on the free space")
        a = rz / den
        x = x + a * dr
        res = res - a * im
        if float(np.linalg.norm(res)) <= tol * ref:
            return x, it, True
        pre = res / diag
        nd = float(res @ pre)
        if not nd > 0:
            raise RuntimeError("PCG encountered a non-positive residual product")
        dr = pre + (nd / rz) * dr
        rz = nd
    return x, max_iter, False


def dirichlet_pcg(action, rhs, diag, fixed, values, initial=None, tol=1e-10, max_iter=4000):
    """solvers/box_kkt_h8.py:350-432."""
    known = np.where(fixed, values, 0.0)
    free = ~fixed
    if not free.any():
        return known, 0, True
    red_rhs = (rhs - action(known))[free]

    def red_action(v):
        full = np.zeros_like(rhs)
        full[free] = v
        return action(full)[free]

    sol, its, conv = scalar_pcg(red_action, red_rhs, diag[free], None if initial is None else initial[free], tol,
                                max_iter)
    out = known.copy()
    out[free] = sol
    return out, its, conv


def box_kkt(action, rhs, diag, lower, upper, initial=None, lin_tol=1e-10, kkt_tol=1e-9, max_lin=4000, max_as=64):
    """solvers/box_kkt_h8.py:603-778 (primal active set); returns (d, lower_active, upper_active, its, trace)."""
    presc = lower == upper
    la = np.zeros(rhs.shape, dtype=bool)
    ua = np.zeros(rhs.shape, dtype=bool)
    cand_init, total, trace = initial, 0, []
    for it in range(1, max_as + 1):
        fixed = la | ua | presc
        values = np.where(ua, upper, lower)
        cand, its, conv = dirichlet_pcg(action, rhs, diag, fixed, values, cand_init, lin_tol, max_lin)
        total += its
        g = action(cand) - rhs
        free = ~fixed
        add_l, add_u = free & (cand < lower), free & (cand > upper)
        rel_l, rel_u = la & (g < -kkt_tol), ua & (g > kkt_tol)
        trace.append((int(add_l.sum()), int(add_u.sum()), int(rel_l.sum()), int(rel_u.sum()), its))
        if not conv:
            raise RuntimeError("H8 scalar box-KKT PCG did not converge")
        if not (add_l | add_u | rel_l | rel_u).any():
            return cand, la, ua, it, total, trace
        if it == max_as:
            break
        la = (la | add_l) & ~rel_l
        ua = (ua | add_u) & ~rel_u
        cand_init = cand
    raise RuntimeError("H8 scalar box-KKT active set did not stabilise")


def fracture_energy_at2(op: VoxelOperator, d):
    """AT2 crack-surface energy (solid3d/operator.py:1511-1544): 1/2 sum_e Gc_e d_e.K d_e with
    K = M/ell + ell L, i.e. half of d . (damage operator without the history term) d."""
    return 0.5 * float(d @ op.damage_action(d))


def anderson_mix(pairs, lower, upper=1.0, regularisation=1e-10):
    """Projected type II Anderson mix of the stored pairs [(d_i, G(d_i)), ...]
    (solid3d/staggered.py:336-374): residuals f_i = G_i - d_i; least squares
    min ||f_k - dF gamma|| over unit-scaled columns of the residual differences with the
    normal equations regularised by ``regularisation``; G_k - dG gamma clamped to the box.
    None with fewer than two pairs or when no column differs."""
    if len(pairs) < 2:
        return None
    G = np.stack([o for _, o in pairs], axis=1)
    F = np.stack([o - i for i, o in pairs], axis=1)
    dF, dG = np.diff(F, axis=1), np.diff(G, axis=1)
    nrm = np.linalg.norm(dF, axis=0)
    keep = nrm > 0.0
    if not keep.any():
        return None
    dF, dG, nrm = dF[:, keep], dG[:, keep], nrm[keep]
    A = dF / nrm
    gamma = np.linalg.solve(A.T @ A + regularisation * np.eye(A.shape[1]), A.T @ F[:, -1]) / nrm
    mixed = G[:, -1] - dG @ gamma
    if not np.isfinite(mixed).all():
        return None
    return np.minimum(np.maximum(mixed, lower), upper)


def staggered_increment(op: VoxelOperator, state: dict, target: float, *, stagger_tol=1e-4, min_stagger=2,
                        max_stagger=2000, rtol=1e-8, damage_linear_tol=1e-10, damage_kkt_tol=1e-8,
                        check_every=10, history="max", irreversibility="box_kkt", anderson_depth=0,
                        anderson_start=2):
    """One accepted increment of solid3d/staggered.py:_attempt (1259-1517), AT2 route, final
    mechanics re-solve.  ``history`` 'max' (H = max(H_prev, Y)) or 'current' (H = Y);
    ``irreversibility`` 'box_kkt' (d_prev <= d <= 1) or 'history_only' (one unconstrained
    solve, staggered.py:804-839); ``anderson_depth`` > 0 mixes the damage fixed point from
    iteration ``anderson_start`` on, each mixed input judged by the end energy and fixed-point
    residual of its own iteration (staggered.py:1283-1401).  ``state`` holds x, d, H, E.
    Returns the new state and (stagger_iters, macro_stress, anderson (accepted, rejected))."""
    c = op.controlled
    n = 3 * op.n_red
    previous = float(state["x"][n + c])
    x_old = state["x"] * (target / previous) if previous != 0.0 else np.zeros_like(state["x"])
    x_old[n + c] = target
    d_old, lower = state["d"].copy(), state["d"]
    fixed = op.mech_fixed_mask()
    vals = np.zeros(op.n_mech)
    vals[n + c] = target
    step = abs(target - previous)
    u_floor = min(op.h) * math.sqrt(n)
    H = state["H"]

    def mech(d, x0):
        g = op.degradation(d)
        x, info = solve_constrained(lambda v: op.mech_action(v, g), op.mech_diagonal(g), fixed, vals, x0=x0,
                                    rtol=rtol, maxiter=20000, check_every=check_every)
        assert info[1]
        return x

    pairs, fallback, best, accepted, rejected_n = [], None, math.inf, 0, 0
    for k in range(1, max_stagger + 1):
        x_new = mech(d_old, x_old)
        Y = op.driving_force(x_new)
        H = np.maximum(state["H"], Y) if history == "max" else Y
        s = op.damage_scale
        act = lambda v: s * op.damage_action(v, H)
        rhs, dg = s * op.damage_rhs(H), s * op.damage_diagonal(H)
        r0 = float(np.linalg.norm(rhs - act(d_old)))
        tol = max(min(damage_linear_tol, 0.1 * damage_kkt_tol / max(r0, 1.0)), 4.0 * np.finfo(np.float64).eps)
        if irreversibility == "box_kkt":
            d_new = box_kkt(act, rhs, dg, lower, np.ones(op.n_red), initial=d_old, lin_tol=tol,
                            kkt_tol=damage_kkt_tol)[0]
        else:
            d_new, _, conv = dirichlet_pcg(act, rhs, dg, np.zeros(op.n_red, dtype=bool), np.zeros(op.n_red),
                                           d_old, tol, 4000)
            assert conv
        du = np.linalg.norm(x_new[:n] - x_old[:n])
        u = np.linalg.norm(x_new[:n])
        dE = np.abs(x_new[n:] - x_old[n:]).max()
        E_inf = np.abs(x_new[n:]).max()
        dd = np.abs(d_new - d_old).max()
        scale_E = max(E_inf, step)
        u_ref = max(u, scale_E * u_floor)
        change = max(du / u_ref if u_ref > 0 else 0.0, dE / scale_E if scale_E > 0 else 0.0, dd)
        if anderson_depth > 0:
            residual = float(np.linalg.norm(d_new - d_old))
            energy = op.elastic_energy(x_new, op.degradation(d_new)) + fracture_energy_at2(op, d_new)
            bad = False
            if fallback is not None:
                bad = bool(energy > best + 1e-12 * abs(best) or residual > fallback[2])
                accepted, rejected_n = accepted + int(not bad), rejected_n + int(bad)
            if not bad:
                best = min(best, energy)
            if bad:  # continue from the plain iterate the mix had replaced
                d_old, x_old, fallback, pairs = fallback[0], fallback[1], None, []
                continue
            fallback = None
        if change <= stagger_tol and k >= min_stagger:
            x_old, d_old = x_new, d_new
            break
        if anderson_depth > 0:
            pairs = (pairs + [(d_old.copy(), d_new.copy())])[-(anderson_depth + 1):]
            mixed = anderson_mix(pairs, lower) if k >= anderson_start else None
            if mixed is not None and not np.array_equal(mixed, d_new):
                fallback = (d_new, x_new, residual)
                x_old, d_old = x_new, mixed
                continue
        x_old, d_old = x_new, d_new
    x = mech(d_old, x_old)
    g = op.degradation(d_old)
    sigma = op.macro_stress(x, g)
    return {"x": x, "d": d_old, "H": H}, (k, sigma, (accepted, rejected_n))
