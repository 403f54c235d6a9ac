"""Pins the hex8 voxel oracle against closed forms, invariants and independent assemblies (CPU)."""
import itertools
import math

import numpy as np
import pytest

from oracle import hex8 as O
from phast_amd import geometry as G


def _op(shape=(3, 3, 3), h=(1.0, 1.2, 0.8), two_phase=True, seed=0, driver=2):
    rng = np.random.default_rng(seed)
    E = int(np.prod(shape))
    bank = np.stack([O.isotropic_stiffness_3d(3000.0, 0.35), O.isotropic_stiffness_3d(70000.0, 0.3)])
    index = rng.integers(0, 2, E) if two_phase else np.zeros(E, dtype=np.int64)
    lam = np.where(index == 0, 3000 * 0.35 / (1.35 * 0.3), 70000 * 0.3 / (1.3 * 0.4))
    mu = np.where(index == 0, 3000 / 2.7, 70000 / 2.6)
    return O.VoxelOperator(shape, h, bank, index, np.where(index == 0, 5e-4, 5e-3), lam, mu,
                           np.full(E, driver), kappa=1e-6, ell=2.0)


def test_reference_voxel_invariants():
    N, grad, w, B, M, L = O.reference_voxel((1.0, 2.0, 0.5))
    np.testing.assert_allclose(N.sum(axis=1), 1.0, rtol=1e-15)
    np.testing.assert_allclose(grad.sum(axis=1), 0.0, atol=1e-14)
    np.testing.assert_allclose(w, 1.0 * 2.0 * 0.5 / 8, rtol=1e-14)
    assert math.isclose(M.sum(), 1.0, rel_tol=1e-13)          # sum of consistent mass = volume
    np.testing.assert_allclose(L @ np.ones(8), 0.0, atol=1e-13)  # constants in the kernel of L
    # B reproduces a linear displacement's constant strain
    x = (O.H8_SIGNS + 1) / 2 * np.array([1.0, 2.0, 0.5])
    Gm = np.array([[1e-3, 2e-3, 0], [-1e-3, 3e-3, 5e-4], [0, 1e-3, -2e-3]])
    u = (x @ Gm.T).ravel()
    eps = B @ u
    expect = [Gm[0, 0], Gm[1, 1], Gm[2, 2], Gm[1, 2] + Gm[2, 1], Gm[0, 2] + Gm[2, 0], Gm[0, 1] + Gm[1, 0]]
    np.testing.assert_allclose(eps, np.broadcast_to(expect, (8, 6)), rtol=1e-12, atol=1e-16)


def test_homogeneous_cell_closed_form():
    """u~ = 0, macro strain E: r_u = 0, r_E = V C E, Sigma = C E."""
    op = _op(two_phase=False)
    E6 = np.array([1e-3, -2e-4, 3e-4, 1e-4, -5e-4, 2e-4])
    x = np.concatenate([np.zeros(3 * op.n_red), E6])
    y = op.mech_action(x)
    C = op.bank[0]
    assert np.abs(y[:-6]).max() < 1e-12 * np.abs(y[-6:]).max()
    np.testing.assert_allclose(y[-6:], op.volume * C @ E6, rtol=1e-13)
    np.testing.assert_allclose(op.macro_stress(x), C @ E6, rtol=1e-13)


def _independent_K(op, g):
    """Dense periodic stiffness incl. the E block: sum_e sum_q w g [B 1]^T C [B 1], by loops."""
    n = op.n_mech
    K = np.zeros((n, n))
    N, grad, w, B, _, _ = O.reference_voxel(op.h)
    for e in range(op.E):
        C = op.bank[op.index[e]]
        dofs = list(op.dofs[e]) + list(range(3 * op.n_red, 3 * op.n_red + 6))
        for q in range(8):
            Bq = np.hstack([B[q], np.eye(6)])
            K[np.ix_(dofs, dofs)] += w[q] * g[e, q] * Bq.T @ C @ Bq
    return K


def test_mech_action_matches_dense_assembly_and_diagonal():
    op = _op(shape=(2, 3, 2))
    rng = np.random.default_rng(1)
    d = rng.uniform(0, 1, op.n_red)
    g = op.degradation(d)
    K = _independent_K(op, g)
    x = rng.uniform(-1e-3, 1e-3, op.n_mech)
    np.testing.assert_allclose(op.mech_action(x, g), K @ x, rtol=1e-11, atol=1e-11 * np.abs(K @ x).max())
    # single-voxel periodic axes couple two local nodes; the diagonal is exact for n >= 2
    np.testing.assert_allclose(op.mech_diagonal(g), np.diag(K), rtol=1e-12)
    assert np.allclose(K, K.T, rtol=1e-12, atol=1e-12 * np.abs(K).max())
    # Hill-Mandel style energy identity: x.A x = 2 * energy
    assert math.isclose(x @ op.mech_action(x, g), 2 * op.elastic_energy(x, g), rel_tol=1e-11)


def test_degradation_by_hand():
    op = _op()
    d = np.full(op.n_red, 0.25)
    np.testing.assert_allclose(op.degradation(d), (1 - 1e-6) * 0.75 ** 2 + 1e-6, rtol=1e-14)


def test_spectral_energy_diagonal_strain_and_eigvalsh():
    lam, mu = 1.7, 0.9
    eps = np.array([[2e-3, -1e-3, 5e-4, 0, 0, 0]])
    Y = O.spectral_tensile_energy(eps, lam, mu)
    assert math.isclose(Y[0], 0.5 * lam * (1.5e-3) ** 2 + mu * ((2e-3) ** 2 + (5e-4) ** 2), rel_tol=1e-12)
    rng = np.random.default_rng(2)
    e = rng.normal(size=(500, 6)) * 1e-3
    np.testing.assert_allclose(O.symmetric_eigvalsh_3x3(O.engineering_strain_to_tensor(e)),
                               np.linalg.eigvalsh(O.engineering_strain_to_tensor(e)), rtol=1e-7, atol=1e-12)
    # all principal strains positive: sum <e_a>^2 = |eps|^2 (Frobenius)
    T = O.engineering_strain_to_tensor(e)
    pos = np.linalg.eigvalsh(T).min(axis=1) > 0
    if pos.any():
        Yp = O.spectral_tensile_energy(e[pos], 0.0, 1.0)
        np.testing.assert_allclose(Yp, (T[pos] ** 2).sum(axis=(1, 2)), rtol=1e-10)


def test_full_driver_is_half_eps_C_eps():
    op = _op(driver=1, two_phase=False)
    E6 = np.array([1e-3, 0, 0, 0, 0, 0])
    x = np.concatenate([np.zeros(3 * op.n_red), E6])
    Y = op.driving_force(x)
    np.testing.assert_allclose(Y, 0.5 * E6 @ op.bank[0] @ E6, rtol=1e-13)


def _independent_damage(op, H):
    n = op.n_red
    A = np.zeros((n, n))
    b = np.zeros(n)
    N, grad, w, _, _, _ = O.reference_voxel(op.h)
    for e in range(op.E):
        nd = op.nodes[e]
        for q in range(8):
            A[np.ix_(nd, nd)] += w[q] * (op.Gc[e] * (np.outer(N[q], N[q]) / op.ell + op.ell * grad[q] @ grad[q].T)
                                         + op.c_history * H[e, q] * np.outer(N[q], N[q]))
            b[nd] += op.c_history * w[q] * H[e, q] * N[q]
    return A, b


def test_damage_operator_matches_dense_assembly():
    op = _op(shape=(2, 2, 3))
    rng = np.random.default_rng(3)
    H = rng.uniform(0, 1e-3, (op.E, 8))
    A, b = _independent_damage(op, H)
    x = rng.uniform(0, 1, op.n_red)
    np.testing.assert_allclose(op.damage_action(x, H), A @ x, rtol=1e-12)
    np.testing.assert_allclose(op.damage_rhs(H), b, rtol=1e-13)
    np.testing.assert_allclose(op.damage_diagonal(H), np.diag(A), rtol=1e-13)


def test_uniform_history_closed_form():
    """Uniform Gc, H, no bounds active: d = cH / (Gc/ell + cH), c = 2(1-kappa)."""
    op = _op(shape=(4, 4, 4), two_phase=False)
    Hv = 3e-4
    H = np.full((op.E, 8), Hv)
    s = op.damage_scale
    act = lambda v: s * op.damage_action(v, H)
    d, its, conv = O.dirichlet_pcg(act, s * op.damage_rhs(H), s * op.damage_diagonal(H), np.zeros(op.n_red, bool),
                                   np.zeros(op.n_red), tol=1e-13)
    c = op.c_history
    np.testing.assert_allclose(d, c * Hv / (5e-4 / op.ell + c * Hv), rtol=1e-10)


def test_box_kkt_matches_bruteforce():
    op = _op(shape=(2, 2, 2))
    rng = np.random.default_rng(4)
    H = rng.uniform(0, 2e-3, (op.E, 8))
    A, b = _independent_damage(op, H)
    lower = rng.uniform(0, 0.8, op.n_red)
    lower[::3] = 0.0
    best, best_val = None, np.inf
    for code in itertools.product((0, 1, 2), repeat=op.n_red):
        code = np.array(code)
        fixed = code > 0
        x = np.where(code == 1, lower, np.where(code == 2, 1.0, 0.0))
        free = ~fixed
        if free.any():
            x[free] = np.linalg.solve(A[np.ix_(free, free)], b[free] - A[np.ix_(free, fixed)] @ x[fixed])
        if np.any(x < lower - 1e-12) or np.any(x > 1 + 1e-12):
            continue
        val = 0.5 * x @ A @ x - b @ x
        if val < best_val - 1e-18:
            best, best_val = x, val
    act = lambda v: op.damage_action(v, H)
    d, la, ua, its, total, trace = O.box_kkt(act, op.damage_rhs(H), op.damage_diagonal(H), lower,
                                             np.ones(op.n_red), initial=lower, lin_tol=1e-14, kkt_tol=1e-12)
    np.testing.assert_allclose(d, best, atol=1e-9)


def test_pcg_solves_dense_system():
    op = _op(shape=(2, 2, 2))
    rng = np.random.default_rng(5)
    g = op.degradation(rng.uniform(0, 0.5, op.n_red))
    K = _independent_K(op, g)
    fixed = op.mech_fixed_mask()
    vals = np.zeros(op.n_mech)
    vals[3 * op.n_red] = 1e-3
    x, info = O.solve_constrained(lambda v: op.mech_action(v, g), op.mech_diagonal(g), fixed, vals, rtol=1e-13,
                                  maxiter=2000)
    free = ~fixed
    ref = vals.copy()
    ref[free] = np.linalg.solve(K[np.ix_(free, free)], -K[np.ix_(free, fixed)] @ vals[fixed])
    np.testing.assert_allclose(x, ref, rtol=1e-9, atol=1e-12)
    assert info[1]


def test_sphere_pore_geometry():
    phase = G.sphere_pore((16, 16, 16), (1.0, 1.0, 1.0), 0.5)
    r = 4.0
    c = (np.arange(16) + 0.5) - 8
    dd = np.sqrt(c[None, None, :] ** 2 + c[None, :, None] ** 2 + c[:, None, None] ** 2)
    assert (phase == G.PORE).sum() == (dd < r).sum()


def test_field_snapshot_closed_form_and_invariants():
    """Snapshot element means: a homogeneous cell gives strain E, stress g C E and energy
    0.5 g E.C E in every voxel; for any state the voxel-mean stress averages to the macro
    stress and the energies sum to the elastic energy 0.5 x.A(g) x."""
    op = _op(two_phase=False)
    E6 = np.array([1e-3, -2e-4, 3e-4, 1e-4, -5e-4, 2e-4])
    x = np.concatenate([np.zeros(3 * op.n_red), E6])
    C = op.bank[0]
    strain, stress, psi = O.field_snapshot_elements(op, x)
    np.testing.assert_allclose(strain, np.broadcast_to(E6, strain.shape), rtol=1e-13)
    np.testing.assert_allclose(stress, np.broadcast_to(C @ E6, stress.shape), rtol=1e-13)
    np.testing.assert_allclose(psi, 0.5 * E6 @ C @ E6, rtol=1e-13)
    op = _op(shape=(4, 3, 5))
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.uniform(-1e-3, 1e-3, 3 * op.n_red), rng.uniform(-1e-3, 1e-3, 6)])
    g = op.degradation(rng.uniform(0, 1, op.n_red))
    strain, stress, psi = O.field_snapshot_elements(op, x, g)
    np.testing.assert_allclose(stress.mean(axis=0), op.macro_stress(x, g), rtol=1e-12)
    np.testing.assert_allclose(psi.sum() * op.volume / op.E, op.elastic_energy(x, g), rtol=1e-12)
    np.testing.assert_allclose(strain.mean(axis=0), x[-6:], rtol=1e-12, atol=1e-18)  # periodic fluctuation averages out
