"""Pins of the T3 oracle (CPU): closed forms, patch tests, energy gradients, dense assembly."""
import numpy as np
import pytest

from oracle import t3 as O
from oracle.q4 import Material

SPLITS = ["isotropic", "amor", "spectral"]


def mesh(nx=5, ny=4, jitter=0.2, seed=0):
    nodes, els = O.structured_t3(nx, ny, 5.0, 4.0, jitter=jitter, seed=seed)
    return O.T3Mesh(nodes, els)


def mat(split="amor", plane_stress=False):
    return Material(E=32000.0, nu=0.2, rho=2.45e-9, Gc=3e-3, l0=0.5, plane_stress=plane_stress,
                    energy_split=split)


def test_geometry_closed_forms():
    m = O.T3Mesh(np.array([[0.0, 0.0], [2.0, 0.0], [0.0, 3.0]]), np.array([[0, 1, 2]]))
    assert m.areas[0] == pytest.approx(3.0, rel=1e-15)
    # phi0 = 1 - x/2 - y/3, phi1 = x/2, phi2 = y/3
    np.testing.assert_allclose(m.grad_phi[0], [[-0.5, -1 / 3], [0.5, 0.0], [0.0, 1 / 3]], rtol=1e-15, atol=1e-16)
    np.testing.assert_allclose(m.M_scalar, [1.0, 1.0, 1.0], rtol=1e-15)
    mm = mesh()
    np.testing.assert_allclose(mm.grad_phi.sum(1), 0.0, atol=1e-13)
    assert mm.M_scalar.sum() == pytest.approx(20.0, rel=1e-13)
    assert mm.areas.sum() == pytest.approx(20.0, rel=1e-13)


@pytest.mark.parametrize("split", SPLITS)
def test_linear_field_strain_and_patch(split):
    mm = mesh()
    op = O.T3Operators(mm, mat(split))
    G = np.array([[1.1e-4, -0.4e-4], [0.7e-4, -0.5e-4]])
    u = mm.nodes @ G.T + np.array([1e-3, -2e-3])
    exx, eyy, gxy = op.compute_strain(u)
    np.testing.assert_allclose(exx, G[0, 0], rtol=1e-9)
    np.testing.assert_allclose(eyy, G[1, 1], rtol=1e-9)
    np.testing.assert_allclose(gxy, G[0, 1] + G[1, 0], rtol=1e-9)
    f = op.internal_force(u, np.zeros(mm.n_nodes))  # constant stress: interior nodes balanced
    xs, ys = mm.nodes[:, 0], mm.nodes[:, 1]
    inner = (xs > 1e-9) & (xs < 5 - 1e-9) & (ys > 1e-9) & (ys < 4 - 1e-9)
    assert np.abs(f[inner]).max() < 1e-10 * np.abs(f).max()


@pytest.mark.parametrize("split", ["isotropic", "amor"])
def test_rigid_motion_has_no_force(split):
    mm = mesh()
    op = O.T3Operators(mm, mat(split))
    th = 1e-9
    u = np.stack([-th * mm.nodes[:, 1], th * mm.nodes[:, 0]], 1) + np.array([3e-4, -1e-4])
    d = np.random.default_rng(1).uniform(0, 1, mm.n_nodes)
    assert np.abs(op.internal_force(u, d)).max() < 1e-12


def energy(op, u, d):
    """sum_e A_e (g_e psi+_e + psi-_e) with psi- = psi_total - psi+ (plane strain)."""
    s = op.compute_strain(u)
    m = op.material
    exx, eyy, gxy = s
    psi_tot = 0.5 * m.lam * (exx + eyy) ** 2 + m.mu * (exx ** 2 + eyy ** 2 + 0.5 * gxy ** 2)
    if m.energy_split == "isotropic":
        return float((op.mesh.areas * op.degradation_e(d) * psi_tot).sum())
    pp = op.compute_psi_plus(u, s)
    if m.energy_split == "amor":  # psi- = 0.5 kappa <tr>-^2 (plane strain)
        tr = exx + eyy
        pm = 0.5 * m.kappa * np.minimum(tr, 0.0) ** 2
    else:
        pm = psi_tot - pp
    return float((op.mesh.areas * (op.degradation_e(d) * pp + pm)).sum())


@pytest.mark.parametrize("split", SPLITS)
def test_force_is_energy_gradient(split):
    mm = mesh(4, 3)
    op = O.T3Operators(mm, mat(split))
    rng = np.random.default_rng(2)
    u = rng.uniform(-1e-4, 1e-4, (mm.n_nodes, 2))
    d = rng.uniform(0, 0.9, mm.n_nodes)
    f = op.internal_force(u, d)
    h = 1e-9
    fd = np.zeros_like(f)
    for n in range(mm.n_nodes):
        for c in range(2):
            up, um = u.copy(), u.copy()
            up[n, c] += h
            um[n, c] -= h
            fd[n, c] = (energy(op, up, d) - energy(op, um, d)) / (2 * h)
    assert np.abs(f - fd).max() < 1e-5 * np.abs(f).max()


def test_amor_hand_values_and_psi_split():
    mm = O.T3Mesh(np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]]), np.array([[0, 1, 2]]))
    op = O.T3Operators(mm, mat("amor"))
    m = op.material
    u = np.array([[0.0, 0.0], [-1e-4, 0.0], [0.0, -2e-4]])  # compression: exx=-1e-4, eyy=-2e-4
    sxx, syy, sxy = op.compute_stress(u, np.array([0.5, 0.5, 0.5]))
    g = m.degradation(np.array([0.5]))[0]
    tr = -3e-4
    assert sxx[0] == pytest.approx(g * 2 * m.mu * (-1e-4 - tr / 3) + m.kappa * tr, rel=1e-13)
    assert syy[0] == pytest.approx(g * 2 * m.mu * (-2e-4 - tr / 3) + m.kappa * tr, rel=1e-13)
    assert sxy[0] == 0.0
    iso = O.T3Operators(mm, mat("isotropic"))
    u2 = np.array([[0.0, 0.0], [1e-4, 3e-5], [2e-5, -5e-5]])
    s = iso.compute_strain(u2)
    sx, sy, sxy2 = iso.compute_stress(u2, np.zeros(3))
    psi = iso.compute_psi_plus(u2, s)
    assert psi[0] == pytest.approx(0.5 * (sx[0] * s[0][0] + sy[0] * s[1][0] + sxy2[0] * s[2][0]), rel=1e-12)


def test_spectral_principal_values():
    mm = O.T3Mesh(np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]]), np.array([[0, 1, 2]]))
    op = O.T3Operators(mm, mat("spectral"))
    m = op.material
    u = np.array([[0.0, 0.0], [2e-4, 0.0], [0.0, -1e-4]])  # exx=2e-4, eyy=-1e-4, gxy=0
    psi = op.compute_psi_plus(u)
    e1 = 0.5e-4 + np.sqrt(1.5e-4 ** 2 + O.SPECTRAL_EPS)
    assert psi[0] == pytest.approx(0.5 * m.lam * 1e-4 ** 2 + m.mu * e1 ** 2, rel=1e-12)


def dense_damage(mm, matl, H):
    """Textbook assembly: Gc l0 int grad phi_a . grad phi_b + (2H + Gc/l0) int phi_a phi_b."""
    n = mm.n_nodes
    K = np.zeros((n, n))
    Mc = np.array([[2.0, 1.0, 1.0], [1.0, 2.0, 1.0], [1.0, 1.0, 2.0]]) / 12.0
    for e, nodes in enumerate(mm.elements):
        gp = mm.grad_phi[e]
        ke = matl.Gc * matl.l0 * mm.areas[e] * gp @ gp.T + (2.0 * H[e] + matl.Gc / matl.l0) * mm.areas[e] * Mc
        K[np.ix_(nodes, nodes)] += ke
    return K


def test_damage_operator_dense_assembly():
    mm = mesh(4, 3)
    matl = mat()
    dm = O.T3Damage(mm, matl)
    H = np.random.default_rng(3).uniform(0, 1e-3, mm.n_elems)
    K = dense_damage(mm, matl, H)
    D = dm.dense(H)
    assert np.abs(D - K).max() < 1e-13 * np.abs(K).max()
    np.testing.assert_allclose(dm.diagonal(H), np.diag(K), rtol=1e-13)
    assert np.abs(D - D.T).max() < 1e-14 * np.abs(D).max()
    assert dm.rhs(H).sum() == pytest.approx(float((2.0 * H * mm.areas).sum()), rel=1e-13)


def test_damage_uniform_history_closed_form():
    mm = mesh()
    matl = mat()
    dm = O.T3Damage(mm, matl)
    Hc = 2e-4
    H = np.full(mm.n_elems, Hc)
    d = np.full(mm.n_nodes, 2 * Hc / (matl.Gc / matl.l0 + 2 * Hc))
    np.testing.assert_allclose(dm.apply(d, H), dm.rhs(H), rtol=1e-12)


def dense_damage_gc(mm, matl, H, Gc_e):
    n = mm.n_nodes
    K = np.zeros((n, n))
    Mc = np.array([[2.0, 1.0, 1.0], [1.0, 2.0, 1.0], [1.0, 1.0, 2.0]]) / 12.0
    for e, nodes in enumerate(mm.elements):
        gp = mm.grad_phi[e]
        ke = Gc_e[e] * matl.l0 * mm.areas[e] * gp @ gp.T + (2.0 * H[e] + Gc_e[e] / matl.l0) * mm.areas[e] * Mc
        K[np.ix_(nodes, nodes)] += ke
    return K


def test_damage_per_element_gc_dense_and_uniform():
    mm = mesh(4, 3)
    matl = mat()
    rng = np.random.default_rng(4)
    H = rng.uniform(0, 1e-3, mm.n_elems)
    Gc_e = rng.uniform(1e-3, 5e-3, mm.n_elems)
    dm = O.T3Damage(mm, matl, Gc_e)
    K = dense_damage_gc(mm, matl, H, Gc_e)
    assert np.abs(dm.dense(H) - K).max() < 1e-13 * np.abs(K).max()
    np.testing.assert_allclose(dm.diagonal(H), np.diag(K), rtol=1e-13)
    flat = O.T3Damage(mm, matl, np.full(mm.n_elems, matl.Gc))
    d = rng.uniform(0, 1, mm.n_nodes)
    np.testing.assert_allclose(flat.apply(d, H), O.T3Damage(mm, matl).apply(d, H), rtol=1e-14)


def test_damage_solve_unconstrained_matches_dense_solve():
    """With d_prev = 0 and moderate H the bounds stay inactive: both CG variants reach the
    solution of the assembled linear system."""
    mm = mesh(6, 5)
    matl = mat()
    H = np.random.default_rng(5).uniform(0, 2e-3, mm.n_elems)
    x = np.linalg.solve(dense_damage(mm, matl, H), O.T3Damage(mm, matl).rhs(H))
    assert 0 < x.min() and x.max() < 1
    for method in ("projected_cg", "post_clamp"):
        s = O.T3DamageSolver(O.T3Mesh(mm.nodes, mm.elements), matl, tol=1e-12, bounds_method=method)
        d = s.solve(H, np.zeros(mm.n_nodes))
        assert s.last_status == "converged"
        np.testing.assert_allclose(d, x, rtol=1e-9, atol=1e-12)


def test_projected_damage_solve_is_the_bound_constrained_minimiser():
    """Brute force on a small mesh: L-BFGS-B on 1/2 d.K d - b.d over lb <= d <= 1."""
    from scipy.optimize import minimize
    mm = mesh(5, 4)
    matl = mat()
    rng = np.random.default_rng(6)
    H = rng.uniform(0, 2e-3, mm.n_elems)
    lb = rng.uniform(0, 0.6, mm.n_nodes)
    lb[rng.permutation(mm.n_nodes)[:3]] = 1.0  # fully damaged nodes sit on both bounds
    K = dense_damage(mm, matl, H)
    b = O.T3Damage(mm, matl).rhs(H)
    res = minimize(lambda z: 0.5 * z @ K @ z - b @ z, lb.copy(), jac=lambda z: K @ z - b,
                   bounds=list(zip(lb, np.ones_like(lb))), method="L-BFGS-B",
                   options=dict(ftol=1e-15, gtol=1e-13, maxiter=20000))
    s = O.T3DamageSolver(O.T3Mesh(mm.nodes, mm.elements), matl, tol=1e-12)
    d = s.solve(H, lb)
    assert s.last_status == "converged"
    assert (d >= lb).all() and (d <= 1.0).all()
    assert np.isclose(d, lb, rtol=0, atol=1e-12)[lb < 1].any() and (d > lb + 1e-3).any()
    np.testing.assert_allclose(d, res.x, atol=1e-6)


@pytest.mark.parametrize("split", ["amor", "spectral"])
def test_full_stress_degradation_uses_the_isotropic_stress(split):
    """stress_degradation='full': the stress is the isotropic one, psi+ keeps the split."""
    mm = mesh()
    rng = np.random.default_rng(7)
    u = rng.uniform(-1e-4, 1e-4, (mm.n_nodes, 2))
    d = rng.uniform(0, 1, mm.n_nodes)
    kw = dict(E=32000.0, nu=0.2, rho=2.45e-9, Gc=3e-3, l0=0.5)
    full = O.T3Operators(mm, Material(energy_split=split, stress_degradation="full", **kw))
    iso = O.T3Operators(mm, Material(energy_split="isotropic", **kw))
    plain = O.T3Operators(mm, Material(energy_split=split, **kw))
    for a, b in zip(full.compute_stress(u, d), iso.compute_stress(u, d)):
        assert np.array_equal(a, b)
    assert np.array_equal(full.compute_psi_plus(u), plain.compute_psi_plus(u))
    assert not np.allclose(full.internal_force(u, d), plain.internal_force(u, d))


def test_kelvin_voigt_damping_dissipates():
    """Free vibration on a T3 plate with both edges x = 0, x = W clamped: the undamped
    central-difference energy stays bounded, the damped one decays (Kelvin-Voigt force
    beta f_int(v, d), mechanics_solver.py:192-195)."""
    from oracle import q4 as OQ
    mm = mesh(8, 4, jitter=0.0)
    ops = O.T3Operators(mm, mat("isotropic"))
    n = mm.n_nodes
    bc = np.zeros((n, 2), dtype=bool)
    xs = mm.nodes[:, 0]
    bc[(xs < 1e-9) | (xs > 5 - 1e-9)] = True
    rng = np.random.default_rng(9)
    u0 = np.where(bc, 0.0, rng.uniform(-1e-5, 1e-5, (n, 2)))
    z = np.zeros((n, 2))
    d = np.zeros(n)
    M = ops.M_vec.reshape(-1, 2)

    def run(zeta):
        dyn = OQ.ExplicitDynamics(ops, dt=0.5 * ops.dt_cfl, damping_ratio_max=zeta)
        u, v, a, E = u0.copy(), z.copy(), z.copy(), []
        a = ops.M_vec_inv.reshape(-1, 2) * (-ops.internal_force(u, d))
        a[bc] = 0.0
        for _ in range(400):
            u, v, a = dyn.step(u, v, a, d, z, bc, z)
            E.append(0.5 * (M * v * v).sum() + 0.5 * (u * ops.internal_force(u, d)).sum())
        return np.array(E)

    e0, e1 = run(0.0), run(0.2)
    assert e0[-50:].mean() > 0.5 * e0[:50].mean()
    assert e1[-50:].mean() < 0.1 * e1[:50].mean()


def test_at1_operator_dense_and_uniform_history():
    """AT1: 3/4 Gc l0 Laplacian, reaction 2H only, source S_H = 3 Gc/(8 l0) (damage_solver.py:1321-1336)."""
    mm = mesh(5, 4)
    kw = dict(E=32000.0, nu=0.2, rho=2.45e-9, Gc=3e-3, l0=0.5)
    m1 = Material(pf_model="AT1", **kw)
    dm = O.T3Damage(mm, m1)
    rng = np.random.default_rng(11)
    H = rng.uniform(0, 1e-2, mm.n_elems)
    n = mm.n_nodes
    K = np.zeros((n, n))
    Mc = np.array([[2.0, 1.0, 1.0], [1.0, 2.0, 1.0], [1.0, 1.0, 2.0]]) / 12.0
    for e, nodes in enumerate(mm.elements):
        gp = mm.grad_phi[e]
        K[np.ix_(nodes, nodes)] += 0.75 * 3e-3 * 0.5 * mm.areas[e] * gp @ gp.T + 2.0 * H[e] * mm.areas[e] * Mc
    assert np.abs(dm.dense(H) - K).max() < 1e-13 * np.abs(K).max()
    src = 3.0 * 3e-3 / (8.0 * 0.5)
    assert dm.at1_source == pytest.approx(src, rel=1e-15)
    assert dm.rhs(H).sum() == pytest.approx(float(((2.0 * H - src) * mm.areas).sum()), rel=1e-12)
    # uniform H above the threshold: d = 1 - S_H/(2H); below it the bound d >= 0 holds d = 0
    for Hc, expect in ((4.0 * src, 1.0 - 1.0 / 8.0), (0.2 * src, 0.0)):
        s = O.T3DamageSolver(O.T3Mesh(mm.nodes, mm.elements), m1, tol=1e-13)
        d = s.solve(np.full(mm.n_elems, Hc), np.zeros(n))
        np.testing.assert_allclose(d, expect, atol=1e-10)
    # a uniform per-element Gc reproduces the scalar-Gc AT1 system
    dme = O.T3Damage(mm, m1, np.full(mm.n_elems, 3e-3))
    x = np.linspace(0.0, 1.0, n)
    np.testing.assert_allclose(dme.apply(x, H), dm.apply(x, H), rtol=1e-13)
    np.testing.assert_allclose(dme.rhs(H), dm.rhs(H), rtol=1e-13)
    np.testing.assert_allclose(dme.diagonal(H), dm.diagonal(H), rtol=1e-13)
