"""Pins the PF-CZM / crack-closure oracle against closed forms, finite differences, dense
assemblies and an independently solved homogeneous cell (CPU)."""
import math

import numpy as np
import pytest
from scipy.optimize import brentq

from oracle import hex8 as O
from oracle import pfczm as P

E_M, NU, F_T, GC, ELL, KAPPA = 345e3, 0.2, 600.0, 25e-3, 1.5e-3, 1e-6
LAM = E_M * NU / ((1 + NU) * (1 - 2 * NU))
MU = E_M / (2 * (1 + NU))


def _op(shape=(3, 3, 3), h=(0.5e-3,) * 3, pore_voxels=(), closure=True, seed=None):
    """A matrix cell (the pore-in-cube material) with optional pore voxels."""
    n = int(np.prod(shape))
    pore = np.zeros(n, dtype=bool)
    pore[list(pore_voxels)] = True
    Ep = np.where(pore, E_M * 1e-6, E_M)
    bank = np.stack([O.isotropic_stiffness_3d(E_M * 1e-6, NU), O.isotropic_stiffness_3d(E_M, NU)])
    lam = Ep * NU / ((1 + NU) * (1 - 2 * NU))
    mu = Ep / (2 * (1 + NU))
    Gc = np.full(n, GC) if seed is None else GC * (1 + 0.2 * np.random.default_rng(seed).random(n))
    base = O.VoxelOperator(shape, h, bank, np.where(pore, 0, 1), Gc, lam, mu, np.where(pore, 0, 2),
                           kappa=KAPPA, ell=ELL, pinned_node=0)
    return P.PfczmOperator(base, Ep, np.full(n, F_T), closure=closure)


# -- the law -------------------------------------------------------------------------------

@pytest.mark.parametrize("softening,p", [("linear", 2), ("exponential", 2.5)])
def test_degradation_closed_forms_and_finite_differences(softening, p):
    a2, a3 = P.softening_parameters(softening)
    a1 = 7.3
    g0, gp0, gpp0 = (float(v) for v in P.degradation_derivatives(np.array(0.0), a1, a2, a3, p, KAPPA))
    assert g0 == 1.0
    assert math.isclose(gp0, -(1 - KAPPA) * a1, rel_tol=1e-14)
    assert math.isclose(gpp0, 2 * (1 - KAPPA) * a1 * (a1 - p - a2), rel_tol=1e-12)   # physics/pfczm.py:66-76
    g1 = float(P.degradation_derivatives(np.array(1.0), a1, a2, a3, p, KAPPA)[0])
    assert math.isclose(g1, KAPPA, rel_tol=1e-12)
    d = np.linspace(0.02, 0.97, 40)
    hstep = 1e-6
    gp = P.degradation_derivatives(d, a1, a2, a3, p, KAPPA)
    up = P.degradation_derivatives(d + hstep, a1, a2, a3, p, KAPPA)
    dn = P.degradation_derivatives(d - hstep, a1, a2, a3, p, KAPPA)
    np.testing.assert_allclose(gp[1], (up[0] - dn[0]) / (2 * hstep), rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(gp[2], (up[1] - dn[1]) / (2 * hstep), rtol=1e-6, atol=1e-7)


def test_linear_softening_a1_and_onset_bound():
    a1 = P.pfczm_a1(E_M, GC, ELL, F_T, KAPPA)
    assert math.isclose(a1, 4 * E_M * GC / (math.pi * ELL * F_T ** 2 * (1 - KAPPA)), rel_tol=1e-15)
    with pytest.raises(ValueError):   # ell beyond 8 l_ch / (3 pi): a1 < 3/2
        base = _op().b
        P.PfczmOperator(base, np.full(base.E, E_M), np.full(base.E, 5000.0))


def test_onset_history_gives_zero_gradient_at_zero_damage():
    """With g'(0) = -(1-eta) a1 and a1 = 4 E Gc/(pi ell f_t^2 (1-eta)) the gradient at d = 0
    vanishes for the uniform history H = f_t^2/(2E): the damage onset."""
    op = _op()
    H = np.full((op.b.E, 8), F_T ** 2 / (2 * E_M))
    g = op.damage_gradient(np.zeros(op.b.n_red), H)
    scale = op.damage_gradient(np.zeros(op.b.n_red), None)
    assert np.abs(g).max() < 1e-12 * np.abs(scale).max()
    assert (op.damage_gradient(np.zeros(op.b.n_red), 1.01 * H) < 0).all()


def test_damage_energy_uniform_closed_form():
    op = _op(pore_voxels=(4,))
    d0, Hv = 0.37, 0.8
    H = np.full((op.b.E, 8), Hv)
    g = float(P.degradation_derivatives(np.array(d0), op.a1[0], op.a2, op.a3, 2, KAPPA)[0])
    Ve = np.prod(op.b.h)
    # voxel 4 is the pore: its own a1
    gp = float(P.degradation_derivatives(np.array(d0), op.a1[4], op.a2, op.a3, 2, KAPPA)[0])
    expect = (op.b.Gc.sum() / (math.pi * ELL) * (2 * d0 - d0 ** 2) * Ve
              + Hv * Ve * (g * (op.b.E - 1) + gp))
    assert math.isclose(op.damage_energy(np.full(op.b.n_red, d0), H), expect, rel_tol=1e-13)


def test_gradient_is_the_derivative_of_the_energy_and_the_tangent_its_hessian():
    op = _op(seed=3)   # pores (a1 ~ 1e-5) have g''(0) < 0: the floor would be active there
    rng = np.random.default_rng(0)
    d = 0.05 * rng.random(op.b.n_red)
    H = 40.0 + 10.0 * rng.random((op.b.E, 8))   # well above the floor: the tangent is the Hessian
    W = op.tangent_weights(d, H)
    floor = P.TANGENT_FLOOR * 2 / (math.pi * ELL) * op.b.Gc[:, None]
    assert (W > floor).all()
    f = rng.standard_normal(op.b.n_red)
    hstep = 1e-7
    fd = (op.damage_energy(d + hstep * f, H) - op.damage_energy(d - hstep * f, H)) / (2 * hstep)
    assert math.isclose(fd, float(op.damage_gradient(d, H) @ f), rel_tol=1e-7)
    fd2 = (op.damage_gradient(d + hstep * f, H) - op.damage_gradient(d - hstep * f, H)) / (2 * hstep)
    np.testing.assert_allclose(op.tangent_action(f, W), fd2, rtol=1e-6, atol=1e-8 * np.abs(fd2).max())
    # the diagonal is the diagonal of the assembled tangent
    n = op.b.n_red
    K = np.stack([op.tangent_action(np.eye(n)[i], W) for i in range(n)], axis=1)
    np.testing.assert_allclose(op.tangent_diagonal(W), np.diag(K), rtol=1e-13)
    np.testing.assert_allclose(K, K.T, rtol=1e-12, atol=1e-12 * np.abs(K).max())
    # below the floor the weights are the floor
    Wl = op.tangent_weights(d, 0.0 * H)
    np.testing.assert_allclose(Wl, np.broadcast_to(floor, Wl.shape), rtol=1e-15)


# -- the closure ---------------------------------------------------------------------------

def _homogeneous_x(op, E6):
    return np.concatenate([np.zeros(3 * op.b.n_red), E6])


@pytest.mark.parametrize("E6", [[-1e-3, 2e-4, 5e-4, 0, 0, 0], [2e-3, -1e-3, -3e-3, 0, 0, 0],
                                [-1e-3, -2e-3, -5e-4, 0, 0, 0], [1e-3, 2e-3, 4e-4, 0, 0, 0]])
def test_closure_energy_of_diagonal_strain_is_the_miehe_compressive_energy(E6):
    """g = 0: the degraded energy is the closure energy, which for strains diagonal in the
    cell axes equals psi0 - (Miehe tensile energy) (splits.py:506-525)."""
    op = _op(shape=(2, 2, 2))
    E6 = np.array(E6)
    x = _homogeneous_x(op, E6)
    C = O.isotropic_stiffness_3d(E_M, NU)
    psi0 = 0.5 * E6 @ C @ E6
    psi_plus = float(O.spectral_tensile_energy(E6[None], LAM, MU)[0])
    g0 = np.zeros((op.b.E, 8))
    assert math.isclose(op.elastic_energy(x, g0) / op.b.volume, psi0 - psi_plus, rel_tol=1e-12, abs_tol=1e-12)


def test_hydrostatic_and_confined_compression_keep_the_intact_stress():
    op = _op(shape=(2, 2, 2))
    C = O.isotropic_stiffness_3d(E_M, NU)
    g = np.full((op.b.E, 8), 0.3)
    for E6 in ([-1e-3, -1e-3, -1e-3, 0, 0, 0], [0, 0, -2e-3, 0, 0, 0]):
        E6 = np.array(E6, dtype=float)
        np.testing.assert_allclose(op.macro_stress(_homogeneous_x(op, E6), g), C @ E6, rtol=1e-13, atol=1e-9)
    # uniaxial tension: the normal stress is degraded, the compressed lateral axes are not
    E6 = np.array([2e-3, -4e-4, -4e-4, 0, 0, 0])
    sigma = op.macro_stress(_homogeneous_x(op, E6), g)
    tr = E6[:3].sum()
    assert math.isclose(sigma[0], 0.3 * (LAM * tr + 2 * MU * E6[0]), rel_tol=1e-13)
    assert math.isclose(sigma[1], 0.3 * (LAM * tr + 2 * MU * E6[1]) + 0.7 * 2 * MU * E6[1], rel_tol=1e-13)


def test_internal_force_is_the_derivative_of_the_degraded_energy():
    op = _op(shape=(3, 2, 2), pore_voxels=(3,))
    rng = np.random.default_rng(5)
    x = np.concatenate([1e-6 * rng.standard_normal(3 * op.b.n_red), [1e-3, -2e-3, 5e-4, 3e-4, -1e-4, 2e-4]])
    g = 0.2 + 0.7 * rng.random((op.b.E, 8))
    f = op.mech_action(x, g)
    v = rng.standard_normal(op.b.n_mech)
    v[-6:] *= 1e-3
    v[:-6] *= 1e-6
    hstep = 1e-4
    fd = (op.elastic_energy(x + hstep * v, g) - op.elastic_energy(x - hstep * v, g)) / (2 * hstep)
    assert math.isclose(fd, float(f @ v), rel_tol=1e-8)
    # at its own pattern the fixed-pattern action is the internal force
    np.testing.assert_allclose(op.mech_action(x, g, op.closure_pattern(x)), f, rtol=1e-14, atol=1e-14 * np.abs(f).max())


def test_pattern_action_diagonal_and_symmetry_from_dense_assembly():
    op = _op(shape=(2, 2, 2), pore_voxels=(6,))
    rng = np.random.default_rng(7)
    g = 0.1 + 0.8 * rng.random((op.b.E, 8))
    pattern = rng.random((op.b.E, 8, 4)) < 0.5
    pattern &= op.cl_on[:, None, None]
    n = op.b.n_mech
    K = np.stack([op.mech_action(np.eye(n)[i], g, pattern) for i in range(n)], axis=1)
    np.testing.assert_allclose(K, K.T, rtol=1e-12, atol=1e-12 * np.abs(K).max())
    # on 2 voxels per axis each reduced node meets its own periodic images, so the exact
    # assembled diagonal is not the Jacobi estimate; compare the element-level sum instead
    expect = np.zeros(n)
    N, grad, w, B, _, _ = O.reference_voxel(op.b.h)
    for e in range(op.b.E):
        C = op.b.bank[op.b.index[e]]
        dofs = list(op.b.dofs[e]) + list(range(n - 6, n))
        for q in range(8):
            Bq = np.hstack([B[q], np.eye(6)])
            D = (P.CLOSURE_ROWS.T * (op.cl_k[e] * pattern[e, q])) @ P.CLOSURE_ROWS
            Kq = w[q] * Bq.T @ (g[e, q] * C + (1 - g[e, q]) * D) @ Bq
            expect[dofs] += np.diag(Kq)
    np.testing.assert_allclose(op.mech_diagonal(g, pattern), expect, rtol=1e-12)
    # all bits active without a pattern
    full = op.cl_on[:, None, None] & np.ones((op.b.E, 8, 4), dtype=bool)
    np.testing.assert_allclose(op.mech_diagonal(g), op.mech_diagonal(g, full), rtol=1e-15)


def test_pattern_change_tolerance_rule():
    op = _op(shape=(2, 2, 2))
    E6 = np.array([1e-3, -1e-3, 5e-4, 0, 0, 0])
    x = _homogeneous_x(op, E6)
    g = np.full((op.b.E, 8), 0.5)
    own = op.closure_pattern(x)
    assert own[:, :, 2].all() and not own[:, :, [0, 1, 3]].any()
    assert op.closure_pattern_change(x, own, g)[1] is False
    old = own.copy()
    old[3, 5, 2] = False       # a bit of strength 2 mu |eps_yy| flipped: jump (1-g) of it
    assert op.closure_pattern_change(x, old, g)[1] is True
    assert op.closure_pattern_change(x, old, np.full_like(g, 1.0 - 1e-9))[1] is False   # jump 1e-9 x strength


# -- the solves ----------------------------------------------------------------------------

def test_projected_newton_uniform_history_matches_the_scalar_root():
    """Uniform H on a matrix cell: d is uniform and solves H g'(d) + 2 Gc/(pi ell)(1 - d) = 0."""
    op = _op(shape=(2, 2, 2))
    for Hv, lower in ((1.5, 0.0), (4.0, 0.0), (4.0, 0.6)):
        H = np.full((op.b.E, 8), Hv)
        d, nit, _, _ = P.damage_pfczm(op, H, np.full(op.b.n_red, lower), np.full(op.b.n_red, lower))
        f = lambda s: Hv * float(P.degradation_derivatives(np.array(s), op.a1[0], op.a2, op.a3, 2, KAPPA)[1]) + \
            2 * GC / (math.pi * ELL) * (1 - s)
        root = brentq(f, 1e-12, 1 - 1e-12, xtol=1e-15)
        expect = max(root, lower)
        # Newton stops at a scaled projected gradient <= 1e-8, about 3e-8 of its scale here
        np.testing.assert_allclose(d, expect, rtol=1e-7, atol=1e-12)
        assert (nit >= 1) == (lower < root)   # a start at a KKT point takes no step


def _homogeneous_solution(eps, Hprev):
    """Uniaxial macro strain eps with free lateral stresses on a homogeneous cell, closure on:
    the lateral strain t solves g lam (eps + 2t) + 2 mu t = 0 (yy, zz compressed, their closure
    bits set, the trace positive) and d solves H(eps, t) g'(d) + 2 Gc/(pi ell)(1 - d) = 0 with
    H = max(Hprev, Miehe tensile energy)."""
    a1 = P.pfczm_a1(E_M, GC, ELL, F_T, KAPPA)
    a2, a3 = P.softening_parameters("linear")

    def gfun(s, k=0):
        return float(P.degradation_derivatives(np.array(s), a1, a2, a3, 2, KAPPA)[k])

    def lateral(g):
        return -g * LAM * eps / (2 * g * LAM + 2 * MU)

    def history(s):
        t = lateral(gfun(s))
        return max(Hprev, 0.5 * LAM * (eps + 2 * t) ** 2 + MU * eps ** 2)

    f = lambda s: history(s) * gfun(s, 1) + 2 * GC / (math.pi * ELL) * (1 - s)
    d = 0.0 if f(0.0) >= 0 else brentq(f, 0.0, 1 - 1e-9, xtol=1e-15)
    g = gfun(d)
    t = lateral(g)
    sigma = g * (LAM * (eps + 2 * t) + 2 * MU * eps)
    return d, t, sigma, history(d)


def test_homogeneous_cell_stagger_reaches_the_independent_fixed_point():
    op = _op(shape=(2, 2, 2))
    st = P.initial_state(op)
    Hprev = 0.0
    seen_damage = False
    for k, eps in enumerate((1.0e-3, 1.9e-3, 2.1e-3, 2.4e-3), start=1):
        # tolerances far below the defaults so that the fixed point is reached to ~1e-12
        st, rec = P.staggered_increment(op, st, eps, stagger_tol=1e-12, damage_kkt_tol=1e-13,
                                       rtol=1e-14)
        d, t, sigma, Hprev = _homogeneous_solution(eps, Hprev)
        seen_damage |= d > 0
        np.testing.assert_allclose(st["d"], d, rtol=1e-9, atol=1e-12)
        n3 = 3 * op.b.n_red
        np.testing.assert_allclose(st["x"][n3 + 1], t, rtol=1e-9)
        np.testing.assert_allclose(st["x"][n3 + 2], t, rtol=1e-9)
        assert math.isclose(rec["sigma"][0], sigma, rel_tol=1e-9)
        assert abs(rec["sigma"][1]) < 1e-7 * sigma
    assert seen_damage


# -- the laminate band group -----------------------------------------------------------------

LAYERS = ((0.10397175686880507, 20000.0, 0.2), (0.44801412156559745, 270000.0, 0.2),
          (0.44801412156559745, 345000.0, 0.2))


def _band_op(shape=(3, 2, 2), band_voxels=(1, 4, 7, 10), seed=0):
    """The pore-in-cube matrix with some voxels replaced by a laminate band of random normals,
    with the band rows and moduli of the host set-up."""
    from phast_amd import solid3d as S
    n = int(np.prod(shape))
    rng = np.random.default_rng(seed)
    ids = np.array(band_voxels)
    t = rng.standard_normal((ids.size, 3))
    Cb = S.rotate_stiffness(S.laminate_stiffness(LAYERS), S.frame_from_tangent(t))
    bank = np.concatenate([O.isotropic_stiffness_3d(E_M, NU)[None], Cb])
    index = np.zeros(n, dtype=np.int64)
    index[ids] = 1 + np.arange(ids.size)
    Eb = 122601.74
    Ep = np.full(n, E_M)
    Ep[ids] = Eb
    lam = Ep * NU / ((1 + NU) * (1 - 2 * NU))
    mu = Ep / (2 * (1 + NU))
    f_t = np.full(n, F_T)
    f_t[ids] = 63.0
    Gc = np.full(n, GC)
    Gc[ids] = 6.46e-4
    base = O.VoxelOperator(shape, (0.5e-3,) * 3, bank, index, Gc, lam, mu, np.full(n, 2), kappa=KAPPA, ell=ELL,
                           pinned_node=0)
    rows, k = S.band_closure_rows(Cb, t)
    return P.PfczmOperator(base, Ep, f_t, closure=True, band=(ids, rows, k))


def test_band_internal_force_is_the_derivative_of_the_degraded_energy():
    op = _band_op()
    assert (op.cl_group == 2).sum() == 4 and (op.cl_group == 1).sum() == op.b.E - 4
    rng = np.random.default_rng(5)
    x = np.concatenate([1e-6 * rng.standard_normal(3 * op.b.n_red), [1e-3, -2e-3, 5e-4, 3e-4, -1e-4, 2e-4]])
    g = 0.2 + 0.7 * rng.random((op.b.E, 8))
    f = op.mech_action(x, g)
    assert op.closure_pattern(x)[op.cl_group == 2].any()
    v = rng.standard_normal(op.b.n_mech)
    v[-6:] *= 1e-3
    v[:-6] *= 1e-6
    hstep = 1e-4
    fd = (op.elastic_energy(x + hstep * v, g) - op.elastic_energy(x - hstep * v, g)) / (2 * hstep)
    assert math.isclose(fd, float(f @ v), rel_tol=1e-8)
    np.testing.assert_allclose(op.mech_action(x, g, op.closure_pattern(x)), f, rtol=1e-14, atol=1e-14 * np.abs(f).max())


def test_band_pattern_action_and_diagonal_from_dense_assembly():
    op = _band_op(shape=(2, 2, 2), band_voxels=(0, 5, 6))
    rng = np.random.default_rng(7)
    g = 0.1 + 0.8 * rng.random((op.b.E, 8))
    pattern = (rng.random((op.b.E, 8, 4)) < 0.5) & op.cl_on[:, None, None]
    n = op.b.n_mech
    K = np.stack([op.mech_action(np.eye(n)[i], g, pattern) for i in range(n)], axis=1)
    np.testing.assert_allclose(K, K.T, rtol=1e-12, atol=1e-12 * np.abs(K).max())
    expect = np.zeros(n)
    Kd = np.zeros((n, n))
    N, grad, w, B, _, _ = O.reference_voxel(op.b.h)
    for e in range(op.b.E):
        C = op.b.bank[op.b.index[e]]
        dofs = list(op.b.dofs[e]) + list(range(n - 6, n))
        R = op.cl_rows[e]
        for q in range(8):
            Bq = np.hstack([B[q], np.eye(6)])
            D = (R.T * (op.cl_k[e] * pattern[e, q])) @ R
            Kq = w[q] * Bq.T @ (g[e, q] * C + (1 - g[e, q]) * D) @ Bq
            expect[dofs] += np.diag(Kq)
            Kd[np.ix_(dofs, dofs)] += Kq
    np.testing.assert_allclose(op.mech_diagonal(g, pattern), expect, rtol=1e-12)
    # the action is the dense assembly (duplicate periodic dofs add up)
    np.testing.assert_allclose(K, Kd, rtol=1e-11, atol=1e-12 * np.abs(Kd).max())


def test_pattern_change_is_judged_per_group():
    """A flipped bit counts against its own group's largest strength (operator.py:826-840):
    flip the strongest bit of the group with the smaller maximum, with a tolerance between
    the two groups' scales -- a change for the per-group rule, none for a global one."""
    op = _band_op(shape=(2, 2, 2), band_voxels=(3,))
    x = _homogeneous_x(op, np.array([1e-3, -1e-3, 5e-4, 0, 0, 0]))
    g = np.full((op.b.E, 8), 0.5)
    own = op.closure_pattern(x)
    assert op.closure_pattern_change(x, own, g)[1] is False
    strength = op.cl_k[:, None, :] * np.abs(op.closure_values(x))
    smax = {grp: strength[op.cl_group == grp].max() for grp in (1, 2)}
    small, big = sorted((1, 2), key=lambda grp: smax[grp])
    assert 0.0 < smax[small] < smax[big]
    cand = np.where(own & (op.cl_group == small)[:, None, None], strength, -1.0)
    e, q, k = np.unravel_index(np.argmax(cand), cand.shape)
    s_bit = strength[e, q, k]
    old = own.copy()
    old[e, q, k] = False
    tol = 0.5 * s_bit / math.sqrt(smax[small] * smax[big])
    assert op.closure_pattern_change(x, old, g, tolerance=tol)[1] is True
    assert op.closure_pattern_change(x, old, g, tolerance=0.5 * s_bit / smax[small] * 1.01)[1] is False
