"""Pins of the oracle's step-level functions (CPU): the explicit Q4 step
(``oracle.q4.step_full_explicit``), the reaction force, the plane-stress Amor branch and one
voxel staggered increment (``oracle.hex8.staggered_increment``).

The checks come from what the mathematics fixes: closed-form fixed points of homogeneous and
laminated cells, uniform-strain patches, plane-stress Hooke's law, and the step order and
damage gating of ``solvers/staggered_solver.py:1115-1165``.
"""
import math

import numpy as np
import pytest

from oracle import hex8 as OH
from oracle import q4 as O

MAT = dict(E=32000.0, nu=0.2, rho=2.45e-9, Gc=3e-3, l0=0.5)


def _q4(nx=4, ny=3, W=2.0, H=1.5, **kw):
    nodes, quads = O.structured_q4_mesh(W, H, nx, ny)
    mesh = O.Q4Mesh(nodes, quads)
    mat = O.Material(**{**MAT, **kw})
    return mesh, mat, O.FEMOperators(mesh, mat)


def _ps_hooke(E, nu, exx, eyy, gxy):
    c = E / (1.0 - nu * nu)
    return c * (exx + nu * eyy), c * (nu * exx + eyy), c * 0.5 * (1.0 - nu) * gxy


# ----------------------------------------------------------------------------------------
# plane-stress Amor (fem_operators.py:301-333, 983-1006)
# ----------------------------------------------------------------------------------------

@pytest.mark.parametrize("d", [0.0, 0.3, 0.9])
def test_plane_stress_amor_tension_is_degraded_plane_stress_hooke(d):
    """With a positive 3D trace nothing is left undegraded, and eps_zz = -nu/(1-nu) tr makes
    sigma_zz vanish: sigma = g(d) D_ps eps and psi+ = 1/2 eps . D_ps eps."""
    mesh, mat, fem = _q4(plane_stress=True)
    E, nu = mat.E, mat.nu
    strain = (np.full((1, 4), 2e-3), np.full((1, 4), -1e-4), np.full((1, 4), 5e-4))
    sxx, syy, sxy = fem.compute_stress(None, np.full(mesh.n_nodes, d), strain=strain)
    g = (1 - d) ** 2 * (1 - mat.eta_residual) + mat.eta_residual
    hx, hy, hxy = _ps_hooke(E, nu, 2e-3, -1e-4, 5e-4)
    assert math.isclose(sxx[0, 0], g * hx, rel_tol=1e-12)
    assert math.isclose(syy[0, 0], g * hy, rel_tol=1e-12)
    assert math.isclose(sxy[0, 0], g * hxy, rel_tol=1e-12)
    psi = fem.compute_psi_plus(None, strain=strain)[0, 0]
    assert math.isclose(psi, 0.5 * (hx * 2e-3 + hy * -1e-4 + hxy * 5e-4), rel_tol=1e-12)


def test_plane_stress_amor_compression_hand_values():
    """exx = eyy = -1e-3, nu = 0.2: eps_zz = +5e-4, tr = -1.5e-3, dev_xx = -5e-4; only the
    deviatoric part is degraded.  mu = 13333.3.., kappa = 17777.7.. (E = 32 GPa)."""
    mesh, mat, fem = _q4(plane_stress=True)
    strain = (np.full((1, 4), -1e-3), np.full((1, 4), -1e-3), np.zeros((1, 4)))
    mu, kap = 32000.0 / 2.4, 32000.0 / 1.8
    for d in (0.0, 0.5, 1.0):
        g = (1 - d) ** 2 * (1 - mat.eta_residual) + mat.eta_residual
        sxx, syy, _ = fem.compute_stress(None, np.full(mesh.n_nodes, d), strain=strain)
        expect = g * 2.0 * mu * -5e-4 + kap * -1.5e-3
        assert math.isclose(sxx[0, 0], expect, rel_tol=1e-12) and math.isclose(syy[0, 0], expect, rel_tol=1e-12)
    # at d = 0 the split is whole: plane-stress Hooke for any sign of the trace
    hx, _, _ = _ps_hooke(32000.0, 0.2, -1e-3, -1e-3, 0.0)
    sxx, _, _ = fem.compute_stress(None, np.zeros(mesh.n_nodes), strain=strain)
    assert math.isclose(sxx[0, 0], hx, rel_tol=1e-12)
    psi = fem.compute_psi_plus(None, strain=strain)[0, 0]
    # psi+ = mu dev:dev = mu (dxx^2 + dyy^2 + dzz^2), dzz = 5e-4 + 5e-4
    assert math.isclose(psi, mu * (2 * (5e-4) ** 2 + (1e-3) ** 2), rel_tol=1e-12)


# ----------------------------------------------------------------------------------------
# reaction force (fem_operators.py:1502-1530)
# ----------------------------------------------------------------------------------------

@pytest.mark.parametrize("plane_stress", [False, True])
def test_reaction_force_of_a_uniform_strain_is_stress_times_edge(plane_stress):
    """u = (exx x, eyy y) on a W x H plate with uniform d: f_int is the nodal image of the
    boundary traction sigma n, so the top-edge y reaction is sigma_yy W and the right-edge x
    reaction sigma_xx H; interior nodes carry none."""
    W, H = 2.0, 1.5
    mesh, mat, fem = _q4(W=W, H=H, plane_stress=plane_stress)
    exx, eyy, d = 1e-3, 2e-3, 0.3
    u = np.stack([exx * mesh.nodes[:, 0], eyy * mesh.nodes[:, 1]], axis=1)
    dn = np.full(mesh.n_nodes, d)
    g = (1 - d) ** 2 * (1 - mat.eta_residual) + mat.eta_residual
    if plane_stress:
        sxx, syy, _ = _ps_hooke(mat.E, mat.nu, exx, eyy, 0.0)
    else:
        lam, mu = mat.lam, mat.mu
        sxx, syy = lam * (exx + eyy) + 2 * mu * exx, lam * (exx + eyy) + 2 * mu * eyy
    top = np.nonzero(mesh.nodes[:, 1] == H)[0]
    right = np.nonzero(mesh.nodes[:, 0] == W)[0]
    assert math.isclose(fem.compute_reaction_force(u, dn, top, 1), g * syy * W, rel_tol=1e-12)
    assert math.isclose(fem.compute_reaction_force(u, dn, right, 0), g * sxx * H, rel_tol=1e-12)
    f_ext = np.zeros((mesh.n_nodes, 2))
    f_ext[top, 1] = 0.25
    R = fem.compute_reaction_force(u, dn, top, 1, f_ext=f_ext)
    assert math.isclose(R, g * syy * W - 0.25 * len(top), rel_tol=1e-12)
    inner = np.nonzero((mesh.nodes[:, 0] > 0) & (mesh.nodes[:, 0] < W) & (mesh.nodes[:, 1] > 0)
                       & (mesh.nodes[:, 1] < H))[0]
    assert abs(fem.compute_reaction_force(u, dn, inner, 1)) < 1e-12 * g * syy * W


# ----------------------------------------------------------------------------------------
# explicit staggered step (staggered_solver.py:1115-1165)
# ----------------------------------------------------------------------------------------

def _explicit(nx=6, ny=4):
    mesh, mat, fem = _q4(nx=nx, ny=ny, W=3.0, H=2.0)
    dyn = O.ExplicitDynamics(fem, dt=0.5 * fem.dt_cfl)
    dmg = O.PhaseFieldDamageSolver(fem, tol=1e-12, max_iter=500, bounds_method="projected_cg")
    return mesh, mat, fem, dyn, dmg


def test_explicit_step_damage_uses_the_updated_history():
    """At rest (u = v = a = f = 0) psi+ = 0, H keeps its uniform value Hv, and the gated solve
    gives the uniform AT2 equilibrium d = 2 Hv / (Gc/l0 + 2 Hv)."""
    mesh, mat, fem, dyn, dmg = _explicit()
    n, E = mesh.n_nodes, mesh.n_elems
    Hv = 0.7 * mat.Gc / mat.l0
    z = np.zeros((n, 2))
    st = O.ExplicitState(z.copy(), z.copy(), z.copy(), np.zeros(n), np.full((E, 4), Hv))
    psi = O.step_full_explicit(dyn, dmg, st, z)
    assert np.all(psi == 0.0) and np.all(st.H == Hv)
    np.testing.assert_allclose(st.d, 2 * Hv / (mat.Gc / mat.l0 + 2 * Hv), rtol=1e-11)


def test_explicit_step_order_mechanics_sees_the_previous_damage():
    """Step 1 from rest gives a1 = M^-1 f, u1 = 0, v1 = dt/2 a1; step 2 moves
    u2 = dt v1 + dt^2/2 a1 = dt^2 a1, and its acceleration is M^-1 (f - f_int(u2, d1)) with d1
    the damage solved in step 1; psi+ and H are those of u2 (mechanics -> strain -> psi+/H ->
    damage)."""
    mesh, mat, fem, dyn, dmg = _explicit()
    n, E = mesh.n_nodes, mesh.n_elems
    rng = np.random.default_rng(5)
    f = rng.uniform(-1.0, 1.0, (n, 2))
    H0 = rng.uniform(0.0, 3.0, (E, 4)) * mat.Gc / mat.l0
    z = np.zeros((n, 2))
    st = O.ExplicitState(z.copy(), z.copy(), z.copy(), np.zeros(n), H0.copy())
    O.step_full_explicit(dyn, dmg, st, f)
    Minv = fem.M_vec_inv.reshape(-1, 2)
    assert np.all(st.u == 0.0)
    np.testing.assert_allclose(st.a, Minv * f, rtol=1e-15)
    d1 = st.d.copy()
    assert d1.max() > 0.5
    a1 = st.a.copy()
    O.step_full_explicit(dyn, dmg, st, f)
    u2 = dyn.dt ** 2 * a1
    np.testing.assert_allclose(st.u, u2, rtol=1e-14)
    a_d1 = Minv * (f - fem.internal_force(u2, d1))
    a_d0 = Minv * (f - fem.internal_force(u2, np.zeros(n)))
    np.testing.assert_allclose(st.a, a_d1, rtol=1e-13, atol=1e-13 * np.abs(a_d1).max())
    assert np.abs(a_d1 - a_d0).max() > 1e-6 * np.abs(a_d1).max()
    np.testing.assert_array_equal(st.H, np.maximum(H0, fem.compute_psi_plus(u2)))
    assert np.all(st.d >= d1)


def test_explicit_step_damage_gating():
    """damage_every = 3 solves at steps 1-5 and then every third step; psi+ and H are updated
    every step."""
    mesh, mat, fem, dyn, dmg = _explicit()
    n, E = mesh.n_nodes, mesh.n_elems
    rng = np.random.default_rng(6)
    f = rng.uniform(-50.0, 50.0, (n, 2))
    z = np.zeros((n, 2))
    st = O.ExplicitState(z.copy(), z.copy(), z.copy(), np.zeros(n), np.zeros((E, 4)))
    solved = []
    for k in range(1, 16):
        calls = []
        orig = dmg.solve

        def counting(*a, **kw):
            calls.append(1)
            return orig(*a, **kw)
        dmg.solve = counting
        H_before = st.H.copy()
        O.step_full_explicit(dyn, dmg, st, f, damage_every=3)
        dmg.solve = orig
        if calls:
            solved.append(k)
        assert np.all(st.H >= H_before)
    assert solved == [1, 2, 3, 4, 5, 6, 9, 12, 15]


# ----------------------------------------------------------------------------------------
# one voxel staggered increment (solid3d/staggered.py:1259-1517)
# ----------------------------------------------------------------------------------------

LAM = lambda E, nu: E * nu / ((1 + nu) * (1 - 2 * nu))
MU = lambda E, nu: E / (2 * (1 + nu))


def _cell(shape, index, phases, h=(1.0, 1.0, 1.0), ell=2.0):
    """phases: list of (E, nu, Gc) per bank slot; index (E,) into it."""
    bank = np.stack([OH.isotropic_stiffness_3d(E, nu) for E, nu, _ in phases])
    lam = np.array([LAM(E, nu) for E, nu, _ in phases])[index]
    mu = np.array([MU(E, nu) for E, nu, _ in phases])[index]
    Gc = np.array([g for _, _, g in phases])[index]
    return OH.VoxelOperator(shape, h, bank, index, Gc, lam, mu, np.full(index.shape[0], 2), kappa=1e-6, ell=ell)


def _state(op):
    return {"x": np.zeros(op.n_mech), "d": np.zeros(op.n_red), "H": np.zeros((op.n_red, 8))}


def test_staggered_increment_homogeneous_cell_fixed_point():
    """One isotropic phase under E_xx control with the other five components free: the
    fluctuation vanishes, E_yy = E_zz = -nu E_xx (uniaxial stress), H = Y = lam/2 <tr>^2 +
    mu E_xx^2 (one positive principal strain), the damage is uniform,
    d = c H / (Gc/l + c H) with c = 2 (1 - kappa), and Sigma_xx = g(d) E E_xx.  The fixed
    point is reached in the minimum two stagger iterations.  Unloading keeps H and d."""
    E, nu, Gc, ell, kap = 3000.0, 0.35, 5e-4, 2.0, 1e-6
    op = _cell((4, 3, 5), np.zeros(60, dtype=np.int64), [(E, nu, Gc)], h=(1.0, 1.1, 0.9), ell=ell)
    n = 3 * op.n_red
    st = _state(op)
    H_max = 0.0
    for t in (2e-4, 4e-4, 6e-4, 3e-4):
        st, (k, sigma, _) = OH.staggered_increment(op, st, t, rtol=1e-13)
        Y = 0.5 * LAM(E, nu) * ((1 - 2 * nu) * t) ** 2 + MU(E, nu) * t ** 2
        H_max = max(H_max, Y)
        c = 2 * (1 - kap)
        d = c * H_max / (Gc / ell + c * H_max)
        g = (1 - kap) * (1 - d) ** 2 + kap
        assert k == 2
        assert np.abs(st["x"][:n]).max() < 1e-12 * t * 1.1
        np.testing.assert_allclose(st["x"][n:], [t, -nu * t, -nu * t, 0, 0, 0], rtol=1e-11, atol=1e-13 * t)
        np.testing.assert_allclose(st["H"], H_max, rtol=1e-10)
        np.testing.assert_allclose(st["d"], d, rtol=1e-10)
        assert math.isclose(sigma[0], g * E * t, rel_tol=1e-10)
        assert np.abs(sigma[1:]).max() < 1e-10 * abs(sigma[0])
    assert st["d"].max() > 0.3


def test_staggered_increment_elastic_laminate():
    """Layers normal to z (stiff/soft), E_xx controlled, the rest free, Gc so large that
    d ~ 1e-13: the in-plane strains are uniform, each layer carries sigma_zz = 0, so
    eps_zz_i = -lam_i (E_xx + E_yy)/(lam_i + 2 mu_i) and Sigma_yy = sum f_i sigma_yy_i = 0
    fixes E_yy.  The piecewise-linear u_z(z) is in the trilinear space, so the Galerkin
    solution is exact: E_yy, E_zz = <eps_zz> and Sigma_xx equal the closed form."""
    nx, ny, nz = 3, 2, 7
    layer = np.array([0, 0, 1, 1, 1, 0, 1])
    index = np.repeat(layer, nx * ny)
    phases = [(3000.0, 0.35, 1e12), (70000.0, 0.3, 1e12)]
    op = _cell((nx, ny, nz), index, phases, h=(1.0, 1.2, 0.8))
    f = np.array([(layer == 0).mean(), (layer == 1).mean()])
    lam = np.array([LAM(E, nu) for E, nu, _ in phases])
    mu = np.array([MU(E, nu) for E, nu, _ in phases])
    a = 2 * lam * mu / (lam + 2 * mu)
    t = 1e-3
    Eyy = -(f @ a) * t / (f @ a + 2 * (f @ mu))
    s = t + Eyy
    ezz = -lam * s / (lam + 2 * mu)
    sxx = f @ (a * s + 2 * mu * t)
    st, (k, sigma, _) = OH.staggered_increment(op, _state(op), t, rtol=1e-13)
    n = 3 * op.n_red
    assert k == 2
    assert st["d"].max() < 1e-12
    np.testing.assert_allclose(st["x"][n:n + 3], [t, Eyy, f @ ezz], rtol=1e-9)
    assert np.abs(st["x"][n + 3:]).max() < 1e-12 * t
    assert math.isclose(sigma[0], sxx, rel_tol=1e-9)
    assert np.abs(sigma[1:]).max() < 1e-9 * abs(sigma[0])
    # u_x, u_y fluctuations vanish; u_z is constant on every node plane
    u = st["x"][:n].reshape(nz, ny, nx, 3)
    assert np.abs(u[..., :2]).max() < 1e-12 * t
    assert np.ptp(u[..., 2], axis=(1, 2)).max() < 1e-9 * np.abs(u[..., 2]).max()


# ----------------------------------------------------------------------------------------
# stagger-loop options: history 'current', irreversibility 'history_only', Anderson mixing
# ----------------------------------------------------------------------------------------

def _uniform_d(H, Gc, ell, kap=1e-6):
    c = 2 * (1 - kap)
    return c * H / (Gc / ell + c * H)


@pytest.mark.parametrize("history,irreversibility", [("max", "box_kkt"), ("max", "history_only"),
                                                     ("current", "box_kkt"), ("current", "history_only")])
def test_staggered_increment_history_and_irreversibility_options(history, irreversibility):
    """Homogeneous cell, loaded to 6e-4 and unloaded to 3e-4.  The damage of the uniform
    problem is d(H) = c H / (Gc/l + c H).  On unloading H is the loading maximum
    (history 'max') or the current Y ('current'); the box solve then keeps d >= d_prev,
    'history_only' does not: only history 'current' with 'history_only' lets the damage
    heal, to d(Y(3e-4))."""
    E, nu, Gc, ell = 3000.0, 0.35, 5e-4, 2.0
    op = _cell((4, 3, 5), np.zeros(60, dtype=np.int64), [(E, nu, Gc)], h=(1.0, 1.1, 0.9), ell=ell)
    Y = lambda t: 0.5 * LAM(E, nu) * ((1 - 2 * nu) * t) ** 2 + MU(E, nu) * t ** 2
    st = _state(op)
    for t in (2e-4, 6e-4):
        st, _ = OH.staggered_increment(op, st, t, rtol=1e-13, history=history, irreversibility=irreversibility)
        np.testing.assert_allclose(st["d"], _uniform_d(Y(t), Gc, ell), rtol=1e-10)
    d_load = _uniform_d(Y(6e-4), Gc, ell)
    st, (k, sigma, _) = OH.staggered_increment(op, st, 3e-4, rtol=1e-13, history=history,
                                               irreversibility=irreversibility)
    H_expect = Y(3e-4) if history == "current" else Y(6e-4)
    np.testing.assert_allclose(st["H"], H_expect, rtol=1e-10)
    d_expect = _uniform_d(Y(3e-4), Gc, ell) if (history, irreversibility) == ("current", "history_only") \
        else d_load
    np.testing.assert_allclose(st["d"], d_expect, rtol=1e-10)
    g = (1 - 1e-6) * (1 - d_expect) ** 2 + 1e-6
    assert math.isclose(sigma[0], g * E * 3e-4, rel_tol=1e-10)


@pytest.mark.parametrize("reg", [0.0, 1e-10])
def test_anderson_mix_solves_an_affine_fixed_point(reg):
    """For d -> M d + c (a contraction), n + 1 plain iterates give n residual differences
    that span R^n, so the type II mix is the fixed point (I - M)^-1 c itself (exactly
    without regularisation); far inside the box the projection does nothing."""
    rng = np.random.default_rng(3)
    n = 5
    M = rng.uniform(-1, 1, (n, n))
    M *= 0.8 / np.abs(np.linalg.eigvals(M)).max()
    c = rng.uniform(-1, 1, n)
    fixed = np.linalg.solve(np.eye(n) - M, c)
    d, pairs = rng.uniform(-1, 1, n), []
    for _ in range(n + 1):
        out = M @ d + c
        pairs.append((d.copy(), out))
        d = out
    mixed = OH.anderson_mix(pairs, -1e9, 1e9, regularisation=reg)
    np.testing.assert_allclose(mixed, fixed, rtol=1e-9 if reg else 1e-11, atol=1e-9 if reg else 1e-11)
    # fewer pairs than the dimension: a least-squares mix, not the fixed point, but projected
    lo = fixed - 1e-3
    part = OH.anderson_mix(pairs[:3], lo, 1e9, regularisation=reg)
    assert part is not None and np.all(part >= lo)
    assert OH.anderson_mix(pairs[:1], -1e9) is None
    same = [(pairs[0][0], pairs[0][1]), (pairs[0][0], pairs[0][1])]
    assert OH.anderson_mix(same, -1e9) is None  # every residual difference vanishes


def test_anderson_mix_matches_the_product_host_class():
    """The host class ProjectedAnderson (torch) against the oracle's mix on random pairs,
    with box clamping active on both sides; depth 3 keeps the last four pairs."""
    import torch
    from phast_amd.staggered3d import ProjectedAnderson
    rng = np.random.default_rng(7)
    n = 40
    lower = rng.uniform(0.0, 0.3, n)
    acc = ProjectedAnderson(3)
    pairs = []
    for i in range(6):
        a, b = rng.uniform(0, 1, n), rng.uniform(0, 1, n)
        pairs.append((a, b))
        acc.push(torch.tensor(a), torch.tensor(b))
        got = acc.mix(torch.tensor(lower))
        want = OH.anderson_mix(pairs[-4:], lower)
        if want is None:
            assert got is None
        else:
            np.testing.assert_allclose(got.numpy(), want, rtol=1e-9, atol=1e-12)
            assert (got.numpy() >= lower).all() and (got.numpy() <= 1.0).all()
    acc.reset()
    assert len(acc) == 0 and acc.mix(torch.tensor(lower)) is None


def test_staggered_increment_anderson_reaches_the_plain_fixed_point():
    """A two-phase cell past the damage peak, where the plain staggered iteration needs many
    iterations: with Anderson mixing (depth 3) the increment converges to the same fixed
    point within the stagger tolerance, and every mixed input is either kept or discarded."""
    rng = np.random.default_rng(5)
    shape = (5, 4, 4)
    index = (rng.uniform(0, 1, 80) < 0.3).astype(np.int64)
    op = _cell(shape, index, [(3000.0, 0.35, 5e-4), (70000.0, 0.3, 5e-3)], ell=2.0)
    plain, mixed = _state(op), _state(op)
    totals = [0, 0]
    for t in (4e-3, 8e-3, 1.2e-2):
        plain, (kp, sp, ap) = OH.staggered_increment(op, plain, t, rtol=1e-13, stagger_tol=1e-8)
        mixed, (km, sm, am) = OH.staggered_increment(op, mixed, t, rtol=1e-13, stagger_tol=1e-8, anderson_depth=3)
        assert ap == (0, 0)
        totals[0] += kp
        totals[1] += km
        np.testing.assert_allclose(mixed["d"], plain["d"], atol=1e-6)
        np.testing.assert_allclose(sm, sp, rtol=1e-5, atol=1e-6 * abs(sp[0]))
    assert plain["d"].max() > 0.3
    assert totals[0] > 12, totals  # the plain iteration is slow here


def test_fracture_energy_at2_uniform_and_quadratic():
    """Gamma(c 1) = Gc c^2 V / (2 ell) (the stiffness part vanishes on constants, the mass
    sums to the volume); Gamma is a quadratic form: Gamma(a u) = a^2 Gamma(u)."""
    Gc, ell = 5e-4, 2.0
    op = _cell((4, 3, 5), np.zeros(60, dtype=np.int64), [(3000.0, 0.35, Gc)], h=(1.0, 1.1, 0.9), ell=ell)
    V = 4 * 3 * 5 * 1.0 * 1.1 * 0.9
    for c in (0.3, 1.0):
        assert math.isclose(OH.fracture_energy_at2(op, np.full(op.n_red, c)), Gc * c * c * V / (2 * ell),
                            rel_tol=1e-12)
    u = np.random.default_rng(1).uniform(0, 1, op.n_red)
    assert math.isclose(OH.fracture_energy_at2(op, 0.5 * u), 0.25 * OH.fracture_energy_at2(op, u), rel_tol=1e-12)
