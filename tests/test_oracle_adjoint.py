"""Pins the oracle's adjoint pieces of the voxel path by central finite differences of the
oracle's own forward operators (CPU).  A dropped term, a wrong sign or a transposed operand
in any of them changes the derivative by O(1) and fails here."""
import numpy as np

from oracle import hex8 as O


def _op(shape=(3, 3, 2), h=(1.0, 1.2, 0.8), seed=0):
    rng = np.random.default_rng(seed)
    E = int(np.prod(shape))
    bank = np.stack([O.isotropic_stiffness_3d(3000.0, 0.35), O.isotropic_stiffness_3d(70000.0, 0.3)])
    index = rng.integers(0, 2, E)
    lam = np.where(index == 0, 3000 * 0.35 / (1.35 * 0.3), 70000 * 0.3 / (1.3 * 0.4))
    mu = np.where(index == 0, 3000 / 2.7, 70000 / 2.6)
    driver = rng.integers(0, 3, E)  # none / full / spectral mixed
    Gc = rng.uniform(4e-4, 6e-4, E)
    return O.VoxelOperator(shape, h, bank, index, Gc, lam, mu, driver, kappa=1e-6, ell=2.0), rng


def _fd(f, v, h):
    g = np.zeros_like(v)
    for i in range(v.size):
        e = np.zeros_like(v)
        e.flat[i] = h
        g.flat[i] = (f(v + e) - f(v - e)) / (2 * h)
    return g


def test_coupling_is_the_damage_derivative_of_the_mechanics_residual():
    op, rng = _op()
    x = rng.uniform(-1e-3, 1e-3, op.n_mech)
    lam = rng.uniform(-1, 1, op.n_mech)
    d = rng.uniform(0.0, 0.9, op.n_red)
    got = O.adjoint_coupling(op, x, lam, d)
    ref = _fd(lambda dd: lam @ op.mech_action(x, op.degradation(dd)), d, 1e-3)  # quadratic in d: exact
    np.testing.assert_allclose(got, ref, rtol=1e-7, atol=1e-9 * np.abs(ref).max())


def test_driving_vjp_is_the_strain_derivative_of_the_driving_energy():
    op, rng = _op(seed=1)
    x = rng.uniform(-1e-3, 1e-3, op.n_mech)
    t = rng.uniform(0.5, 1.5, (op.E, 8))
    got = O.adjoint_driving(op, x, t)
    ref = _fd(lambda xx: (t * op.driving_force(xx, method="eigh")).sum(), x, 1e-8)
    np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-7 * np.abs(ref).max())


def test_history_vjp_is_the_history_derivative_of_the_damage_residual():
    op, rng = _op(seed=2)
    d = rng.uniform(0, 1, op.n_red)
    lam = rng.uniform(-1, 1, op.n_red)
    H = rng.uniform(0, 1, (op.E, 8))
    got = O.adjoint_damage_history(op, d, lam)
    ref = _fd(lambda HH: lam @ (op.damage_action(d, HH) - op.damage_rhs(HH)), H, 0.1)  # linear in H: exact
    np.testing.assert_allclose(got, ref, rtol=1e-8, atol=1e-12)


def test_gc_vjp_is_the_gc_derivative_of_the_damage_residual():
    op, rng = _op(seed=3)
    d = rng.uniform(0, 1, op.n_red)
    lam = rng.uniform(-1, 1, op.n_red)
    H = rng.uniform(0, 1, (op.E, 8))
    got = O.adjoint_damage_gc(op, d, lam)
    Gc0 = op.Gc.copy()

    def f(G):
        op.Gc = G
        v = lam @ op.damage_action(d, H)
        op.Gc = Gc0
        return v

    ref = _fd(f, Gc0, 1e-5)  # linear in Gc
    np.testing.assert_allclose(got, ref, rtol=1e-7, atol=1e-12)


def test_qp_to_node_is_the_transpose_of_interpolation():
    op, rng = _op(seed=4)
    f = rng.uniform(-1, 1, (op.E, 8))
    v = rng.uniform(-1, 1, op.n_red)
    lhs = (f * (v[op.nodes] @ op.N.T)).sum()
    assert abs(lhs - v @ O.qp_to_node(op, f)) < 1e-12 * abs(lhs) + 1e-14
