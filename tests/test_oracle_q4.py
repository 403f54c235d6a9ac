"""Pins the Q4 oracle against closed forms, invariants and independent brute-force assemblies.

Each check is chosen so that a plausible mistake in the oracle (a dropped term, a wrong
sign or index, a transposed operand) fails it.  Runs on the CPU.
"""
import math

import numpy as np
import pytest

from oracle import q4 as O


def _ops(nx=3, ny=2, W=3.0, H=2.0, split="amor", plane_stress=False):
    nodes, quads = O.structured_q4_mesh(W, H, nx, ny)
    mesh = O.Q4Mesh(nodes, quads)
    mat = O.Material(E=32000.0, nu=0.2, rho=2.45e-9, Gc=3e-3, l0=0.5, energy_split=split, plane_stress=plane_stress)
    return mesh, mat, O.FEMOperators(mesh, mat)


def test_rectangle_geometry_closed_form():
    mesh, _, _ = _ops(nx=4, ny=3, W=2.0, H=0.9)
    hx, hy = 0.5, 0.3
    a = 1.0 / math.sqrt(3.0)
    sx = np.array([-1, 1, 1, -1.0])
    sy = np.array([-1, -1, 1, 1.0])
    pts = [(-a, -a), (a, -a), (-a, a), (a, a)]
    for q, (xi, eta) in enumerate(pts):
        gx = sx * (1 + sy * eta) / (2 * hx)
        gy = sy * (1 + sx * xi) / (2 * hy)
        np.testing.assert_allclose(mesh.grad[:, q, :, 0], np.broadcast_to(gx, (mesh.n_elems, 4)), rtol=1e-14,
                                   atol=1e-13)
        np.testing.assert_allclose(mesh.grad[:, q, :, 1], np.broadcast_to(gy, (mesh.n_elems, 4)), rtol=1e-14,
                                   atol=1e-13)
    np.testing.assert_allclose(mesh.wdet, hx * hy / 4.0, rtol=1e-14)
    np.testing.assert_allclose(mesh.grad.sum(axis=2), 0.0, atol=1e-13)   # sum_a gradN_a = 0
    np.testing.assert_allclose(mesh.N_q.sum(axis=1), 1.0, rtol=1e-15)     # partition of unity
    assert math.isclose(mesh.M_scalar.sum(), 2.0 * 0.9, rel_tol=1e-14)    # lumped mass = area
    # corner node carries a quarter cell, an interior node a full cell
    assert math.isclose(mesh.M_scalar[0], hx * hy / 4.0, rel_tol=1e-14)
    assert math.isclose(mesh.M_scalar[1 * 5 + 1], hx * hy, rel_tol=1e-14)


def _independent_K(mesh, mat, g_q):
    """Textbook sum_q w g B^T D B with plane-strain D, assembled by loops (isotropic split)."""
    lam, mu = mat.lam, mat.mu
    D = np.array([[lam + 2 * mu, lam, 0.0], [lam, lam + 2 * mu, 0.0], [0.0, 0.0, mu]])
    n = mesh.n_nodes
    K = np.zeros((2 * n, 2 * n))
    for e, conn in enumerate(mesh.elements):
        for q in range(4):
            B = np.zeros((3, 8))
            for a in range(4):
                dx, dy = mesh.grad[e, q, a]
                B[0, 2 * a] = dx
                B[1, 2 * a + 1] = dy
                B[2, 2 * a] = dy
                B[2, 2 * a + 1] = dx
            Ke = mesh.wdet[e, q] * g_q[e, q] * B.T @ D @ B
            dofs = np.ravel([[2 * c, 2 * c + 1] for c in conn])
            K[np.ix_(dofs, dofs)] += Ke
    return K


def test_internal_force_matches_dense_stiffness_isotropic():
    mesh, mat, fem = _ops(split="isotropic")
    rng = np.random.default_rng(1)
    u = rng.uniform(-1e-3, 1e-3, (mesh.n_nodes, 2))
    d = rng.uniform(0, 1, mesh.n_nodes)
    g_q = fem.degradation_q(d)
    K = _independent_K(mesh, mat, g_q)
    f = fem.internal_force(u, d)
    np.testing.assert_allclose(f.ravel(), K @ u.ravel(), rtol=1e-12, atol=1e-12 * np.abs(K @ u.ravel()).max())


def _energy(fem, u, d):
    """sum_q w (g psi+ + psi-) with psi- = kappa/2 <tr>_-^2 (Amor, plane strain)."""
    exx, eyy, gxy = fem.compute_strain(u)
    g = fem.degradation_q(d)
    psi_p = fem.compute_psi_plus(u)
    tr = exx + eyy
    psi_m = 0.5 * fem.material.kappa * np.minimum(tr, 0.0) ** 2
    return float((fem.mesh.wdet * (g * psi_p + psi_m)).sum())


@pytest.mark.parametrize("plane_stress", [False, True])
def test_amor_force_is_energy_gradient(plane_stress):
    """Amor is variationally consistent: f_int = dPsi/du, Psi = sum w (g psi+ + psi-)."""
    if plane_stress:
        pytest.skip("plane-stress Amor is not an exact energy gradient (eps_zz is not a free variable)")
    mesh, mat, fem = _ops(split="amor", plane_stress=plane_stress)
    rng = np.random.default_rng(2)
    u = rng.uniform(-1e-3, 1e-3, (mesh.n_nodes, 2))
    d = rng.uniform(0, 1, mesh.n_nodes)
    f = fem.internal_force(u, d)
    h = 1e-7
    fd = np.zeros_like(u)
    for i in range(mesh.n_nodes):
        for c in range(2):
            up, um = u.copy(), u.copy()
            up[i, c] += h
            um[i, c] -= h
            fd[i, c] = (_energy(fem, up, d) - _energy(fem, um, d)) / (2 * h)
    np.testing.assert_allclose(f, fd, rtol=1e-6, atol=1e-6 * np.abs(f).max())


def test_amor_branches_by_hand():
    mesh, mat, fem = _ops(nx=1, ny=1, W=1.0, H=1.0)
    kap, mu = mat.kappa, mat.mu
    g = 0.25
    e = np.array([[-2e-3, -1e-3, 0.0]])  # compression: tr < 0 keeps kappa tr undegraded
    exx, eyy, gxy = (np.full((1, 4), v) for v in e[0])
    d = np.full(4, 0.5)  # g(0.5) = 0.25 (1-eta) + eta
    sxx, syy, sxy = fem.compute_stress(None, d, strain=(exx, eyy, gxy))
    gg = 0.25 * (1 - mat.eta_residual) + mat.eta_residual
    tr = -3e-3
    assert math.isclose(sxx[0, 0], gg * 2 * mu * (-2e-3 - tr / 3) + kap * tr, rel_tol=1e-13)
    assert math.isclose(syy[0, 0], gg * 2 * mu * (-1e-3 - tr / 3) + kap * tr, rel_tol=1e-13)
    # pure shear: psi+ = mu * 2 exy^2 and sigma_xy degraded
    z = np.zeros((1, 4))
    psi = fem.compute_psi_plus(None, strain=(z, z, np.full((1, 4), 2e-3)))
    assert math.isclose(psi[0, 0], mu * 2 * (1e-3) ** 2, rel_tol=1e-13)
    _, _, sxy = fem.compute_stress(None, d, strain=(z, z, np.full((1, 4), 2e-3)))
    assert math.isclose(sxy[0, 0], gg * mu * 2e-3, rel_tol=1e-13)
    del g


def test_psi_split_sums_to_total_energy():
    mesh, mat, fem = _ops()
    rng = np.random.default_rng(3)
    u = rng.uniform(-1e-3, 1e-3, (mesh.n_nodes, 2))
    exx, eyy, gxy = fem.compute_strain(u)
    psi_p = fem.compute_psi_plus(u)
    tr = exx + eyy
    psi_m = 0.5 * mat.kappa * np.minimum(tr, 0) ** 2
    psi0 = 0.5 * mat.lam * tr ** 2 + mat.mu * (exx ** 2 + eyy ** 2 + 0.5 * gxy ** 2)
    np.testing.assert_allclose(psi_p + psi_m, psi0, rtol=1e-12)


def test_patch_test_linear_field_zero_interior_force():
    mesh, mat, fem = _ops(nx=4, ny=4, W=1.0, H=1.0)
    x, y = mesh.nodes[:, 0], mesh.nodes[:, 1]
    u = np.stack([1e-3 * x + 2e-4 * y + 0.1, -3e-4 * x + 5e-4 * y - 0.2], axis=1)
    d = np.full(mesh.n_nodes, 0.3)
    f = fem.internal_force(u, d)
    interior = (x > 0) & (x < 1) & (y > 0) & (y < 1)
    assert np.abs(f[interior]).max() < 1e-12 * np.abs(f).max()
    # rigid motion: zero everywhere
    u_r = np.stack([np.full_like(x, 0.3) - 1e-3 * y, np.full_like(x, -0.1) + 1e-3 * x], axis=1)
    assert np.abs(fem.internal_force(u_r, d)).max() < 1e-9


def test_history_update_bitwise_operand():
    rng = np.random.default_rng(4)
    H = rng.uniform(0, 1, (50, 4))
    psi = rng.uniform(0, 1, (50, 4))
    Hn = O.history_update(H, psi)
    assert np.all((Hn == H) | (Hn == psi)) and np.all(Hn >= H) and np.all(Hn >= psi)


def test_central_difference_recurrence():
    """Velocity-Verlet == u_{n+1} - 2u_n + u_{n-1} = dt^2 M^-1 (f_ext - f_int(u_n))."""
    mesh, mat, fem = _ops(nx=3, ny=3, W=3.0, H=3.0, split="isotropic")
    dyn = O.ExplicitDynamics(fem, dt=0.5 * fem.dt_cfl)
    rng = np.random.default_rng(5)
    u = rng.uniform(-1e-4, 1e-4, (mesh.n_nodes, 2))
    v = rng.uniform(-1e-2, 1e-2, (mesh.n_nodes, 2))
    d = np.zeros(mesh.n_nodes)
    f_ext = rng.uniform(-1, 1, (mesh.n_nodes, 2)) * 1e-6
    a = fem.M_vec_inv.reshape(-1, 2) * (f_ext - fem.internal_force(u, d))
    us = [u]
    for _ in range(3):
        u, v, a = dyn.step(u, v, a, d, f_ext)
        us.append(u)
    for n in (1, 2):
        lhs = us[n + 1] - 2 * us[n] + us[n - 1]
        rhs = dyn.dt ** 2 * fem.M_vec_inv.reshape(-1, 2) * (f_ext - fem.internal_force(us[n], d))
        np.testing.assert_allclose(lhs, rhs, rtol=1e-7, atol=1e-9 * np.abs(rhs).max())


def test_rigid_translation_is_preserved():
    mesh, mat, fem = _ops()
    dyn = O.ExplicitDynamics(fem)
    n = mesh.n_nodes
    u = np.tile([0.1, -0.2], (n, 1))
    v = np.tile([1.0, 2.0], (n, 1))
    a = np.zeros((n, 2))
    u1, v1, a1 = dyn.step(u, v, a, np.zeros(n), np.zeros((n, 2)))
    # f_int of a translation is rounding noise only (relative to a unit-strain force scale)
    np.testing.assert_allclose(u1, u + dyn.dt * v, rtol=1e-15, atol=1e-12)
    np.testing.assert_allclose(v1, v, rtol=0, atol=1e-8)
    assert np.abs(a1).max() * mesh.M_scalar.max() * mat.rho < 1e-12 * mat.E


def _independent_damage_matrix(mesh, Gc, l0, H):
    n = mesh.n_nodes
    A = np.zeros((n, n))
    b = np.zeros(n)
    for e, conn in enumerate(mesh.elements):
        for q in range(4):
            w = mesh.wdet[e, q]
            Nq = mesh.N_q[q]
            G = mesh.grad[e, q]
            rho = 2 * H[e, q] + Gc / l0
            A[np.ix_(conn, conn)] += w * (Gc * l0 * G @ G.T + rho * np.outer(Nq, Nq))
            b[conn] += w * 2 * H[e, q] * Nq
    return A, b


def test_damage_operator_matches_dense_assembly():
    mesh, mat, fem = _ops(nx=3, ny=2)
    dmg = O.PhaseFieldDamageSolver(fem, tol=1e-12)
    rng = np.random.default_rng(6)
    H = rng.uniform(0, 1e-3, (mesh.n_elems, 4))
    A, b_ref = _independent_damage_matrix(mesh, mat.Gc, mat.l0, H)
    rho, b, M_inv = dmg.prepare(H, np.zeros(mesh.n_nodes))
    x = rng.uniform(0, 1, mesh.n_nodes)
    np.testing.assert_allclose(dmg.Ax(x, rho), A @ x, rtol=1e-13)
    np.testing.assert_allclose(b, b_ref, rtol=1e-13)
    np.testing.assert_allclose(1.0 / M_inv - 1e-30, np.diag(A), rtol=1e-13)
    assert np.allclose(A, A.T, rtol=1e-14) and np.linalg.eigvalsh(A).min() > 0


def test_uniform_history_closed_form_damage():
    """Uniform H and d_prev = 0: d = 2H / (Gc/l0 + 2H) (physics/initial_conditions.py:206-224)."""
    mesh, mat, fem = _ops(nx=6, ny=4, W=6.0, H=4.0)
    for method in ("projected_cg", "post_clamp"):
        dmg = O.PhaseFieldDamageSolver(fem, tol=1e-13, max_iter=500, bounds_method=method)
        Hv = 0.7 * mat.Gc / mat.l0
        d = dmg.solve(np.full((mesh.n_elems, 4), Hv), np.zeros(mesh.n_nodes))
        np.testing.assert_allclose(d, 2 * Hv / (mat.Gc / mat.l0 + 2 * Hv), rtol=1e-10)


def test_projected_cg_matches_bruteforce_box_qp():
    """Enumerate active sets of min 1/2 d.A d - b.d on [lb,1] for 8 dofs and compare."""
    mesh, mat, fem = _ops(nx=3, ny=1, W=1.5, H=0.5)
    rng = np.random.default_rng(7)
    H = rng.uniform(0, 40, (mesh.n_elems, 4)) * mat.Gc / mat.l0
    lb = rng.uniform(0, 0.9, mesh.n_nodes)
    lb[::3] = 0.0
    A, b = _independent_damage_matrix(mesh, mat.Gc, mat.l0, H)
    n = mesh.n_nodes
    best, best_val = None, np.inf
    import itertools
    for code in itertools.product((0, 1, 2), repeat=n):  # 0 free, 1 at lb, 2 at 1
        code = np.array(code)
        fixed = code > 0
        x = np.where(code == 1, lb, np.where(code == 2, 1.0, 0.0))
        free = ~fixed
        if free.any():
            x[free] = np.linalg.solve(A[np.ix_(free, free)], b[free] - A[np.ix_(free, fixed)] @ x[fixed])
        if np.any(x < lb - 1e-12) or np.any(x > 1 + 1e-12):
            continue
        val = 0.5 * x @ A @ x - b @ x
        if val < best_val - 1e-15:
            best, best_val = x, val
    dmg = O.PhaseFieldDamageSolver(fem, tol=1e-14, max_iter=2000, bounds_method="projected_cg")
    d = dmg.solve(H, lb)
    np.testing.assert_allclose(d, best, atol=1e-9)


# -- general (distorted, renumbered) quadrilateral meshes ---------------------------------

def distorted_q4_mesh(nx=4, ny=3, W=4.0, H=3.0, amp=0.22, seed=0, shuffle=True):
    """A structured grid whose interior nodes are moved by up to ``amp`` cells and whose
    nodes and cells are renumbered at random (corners stay counter-clockwise)."""
    nodes, quads = O.structured_q4_mesh(W, H, nx, ny)
    rng = np.random.default_rng(seed)
    hx, hy = W / nx, H / ny
    x, y = nodes[:, 0], nodes[:, 1]
    inner = (x > 0) & (x < W) & (y > 0) & (y < H)
    nodes = nodes.copy()
    nodes[inner, 0] += amp * hx * rng.uniform(-1, 1, inner.sum())
    nodes[inner, 1] += amp * hy * rng.uniform(-1, 1, inner.sum())
    if shuffle:
        perm = rng.permutation(nodes.shape[0])       # new id of old node i: inv[i]
        inv = np.empty_like(perm)
        inv[perm] = np.arange(perm.size)
        nodes = nodes[perm]
        quads = inv[quads]
        quads = quads[rng.permutation(quads.shape[0])]
        quads = np.stack([np.roll(q, int(r)) for q, r in zip(quads, rng.integers(0, 4, quads.shape[0]))])
    return nodes, quads


def _shoelace(p):
    x, y = p[:, 0], p[:, 1]
    return 0.5 * float(np.dot(x, np.roll(y, -1)) - np.dot(np.roll(x, -1), y))


def test_distorted_mesh_geometry_and_patch_test():
    nodes, quads = distorted_q4_mesh()
    mesh = O.Q4Mesh(nodes, quads)
    for e in range(mesh.n_elems):   # sum_q wdetJ is the cell's polygon area (bilinear map, exact)
        assert math.isclose(mesh.wdet[e].sum(), _shoelace(nodes[quads[e]]), rel_tol=1e-13)
    assert (mesh.wdet > 0).all()
    np.testing.assert_allclose(mesh.grad.sum(axis=2), 0.0, atol=1e-12)
    assert math.isclose(mesh.M_scalar.sum(), 12.0, rel_tol=1e-13)
    # a linear field has the exact constant gradient at every Gauss point of every cell
    G = np.array([[1.3e-3, -0.4e-3], [0.7e-3, 2.1e-3]])   # du_i/dx_j
    u = nodes @ G.T + np.array([0.1, -0.2])
    mat = O.Material(E=32000.0, nu=0.2, rho=2.45e-9, Gc=3e-3, l0=0.5, energy_split="isotropic")
    fem = O.FEMOperators(mesh, mat)
    exx, eyy, gxy = fem.compute_strain(u)
    np.testing.assert_allclose(exx, G[0, 0], rtol=1e-11)
    np.testing.assert_allclose(eyy, G[1, 1], rtol=1e-11)
    np.testing.assert_allclose(gxy, G[0, 1] + G[1, 0], rtol=1e-11)
    x, y = nodes[:, 0], nodes[:, 1]
    interior = (x > 1e-9) & (x < 4 - 1e-9) & (y > 1e-9) & (y < 3 - 1e-9)
    f = fem.internal_force(u, np.full(mesh.n_nodes, 0.2))
    assert np.abs(f[interior]).max() < 1e-11 * np.abs(f).max()


def test_distorted_amor_force_is_energy_gradient_and_damage_operator_is_the_assembly():
    nodes, quads = distorted_q4_mesh(seed=3)
    mesh = O.Q4Mesh(nodes, quads)
    mat = O.Material(E=32000.0, nu=0.2, rho=2.45e-9, Gc=3e-3, l0=0.5, energy_split="amor")
    fem = O.FEMOperators(mesh, mat)
    rng = np.random.default_rng(4)
    u = rng.uniform(-1e-3, 1e-3, (mesh.n_nodes, 2))
    d = rng.uniform(0, 1, mesh.n_nodes)
    f = fem.internal_force(u, d)
    h = 1e-7
    fd = np.zeros_like(u)
    for i in range(mesh.n_nodes):
        for c in range(2):
            up, um = u.copy(), u.copy()
            up[i, c] += h
            um[i, c] -= h
            fd[i, c] = (_energy(fem, up, d) - _energy(fem, um, d)) / (2 * h)
    np.testing.assert_allclose(f, fd, rtol=1e-6, atol=1e-6 * np.abs(f).max())
    dmg = O.PhaseFieldDamageSolver(fem, tol=1e-12)
    H = rng.uniform(0, 1e-3, (mesh.n_elems, 4))
    A, b_ref = _independent_damage_matrix(mesh, mat.Gc, mat.l0, H)
    rho, b, M_inv = dmg.prepare(H, np.zeros(mesh.n_nodes))
    x = rng.uniform(0, 1, mesh.n_nodes)
    np.testing.assert_allclose(dmg.Ax(x, rho), A @ x, rtol=1e-13)
    np.testing.assert_allclose(b, b_ref, rtol=1e-13)
    # the Laplacian part annihilates constants and, away from the boundary, linear fields
    lin = 0.3 * nodes[:, 0] - 0.2 * nodes[:, 1] + 1.0
    lap = dmg.Ax(lin, np.zeros_like(rho))
    xs, ys = nodes[:, 0], nodes[:, 1]
    interior = (xs > 1e-9) & (xs < 4 - 1e-9) & (ys > 1e-9) & (ys < 3 - 1e-9)
    assert np.abs(lap[interior]).max() < 1e-12 * max(np.abs(lap).max(), 1e-30) + 1e-15
    assert np.abs(dmg.Ax(np.ones(mesh.n_nodes), np.zeros_like(rho))).max() < 1e-13
