"""Pins of the T3 damage-adjoint oracle (CPU): central differences of dense solve-and-project
damage, the reference's two-triangle regression cases, the per-element / scalar identity, the
adjoint PCG against a dense solve, and the active-set precedence rules."""
import dataclasses

import numpy as np
import pytest

from oracle import t3 as O
from oracle import t3_adjoint as A
from oracle.q4 import Material


def dense_unconstrained(op, H, fixed=None, vals=None):
    """A d = b by a dense solve; pinned nodes eliminated and held at vals."""
    K, b = op.dense(H), op.rhs(H)
    if fixed is None:
        return np.linalg.solve(K, b)
    free = ~fixed
    x = np.where(fixed, vals, 0.0)
    x[free] = np.linalg.solve(K[np.ix_(free, free)], b[free] - K[np.ix_(free, fixed)] @ vals[fixed])
    return x


def projected(op, H, d_prev, fixed=None, vals=None):
    """Post-clamp damage: clamp(max(d_unc, d_prev), 0, 1), pins at their values."""
    d = np.clip(np.maximum(dense_unconstrained(op, H, fixed, vals), d_prev), 0.0, 1.0)
    return d if fixed is None else np.where(fixed, vals, d)


def two_triangles():
    """The unit square of two triangles of the reference's adjoint regression tests, with
    Gc l0 = 0.7 and Gc/l0 = 0.6 and node 0 pinned at 0.8."""
    m = O.T3Mesh(np.array([[0., 0.], [1., 0.], [1., 1.], [0., 1.]]), np.array([[0, 1, 2], [0, 2, 3]]))
    mat = Material(E=1.0, nu=0.3, rho=1.0, Gc=np.sqrt(0.42), l0=np.sqrt(0.7 / 0.6))
    fixed = np.array([True, False, False, False])
    vals = np.array([0.8, 0.0, 0.0, 0.0])
    return m, mat, fixed, vals


def irreversibility_state(op, fixed, vals):
    """H, d_prev with node 1 held at d_prev and nodes 2, 3 strictly free."""
    H = np.array([0.3, 0.8])
    x = dense_unconstrained(op, H, fixed, vals)
    d_prev = 0.5 * x
    d_prev[1] = 0.5 * (1.0 + x[1])
    assert d_prev[1] > x[1] + 1e-3 and np.all(x[2:] > d_prev[2:] + 1e-3)
    return H, d_prev


@pytest.mark.parametrize("weights", [[.5, -.7, 1.2, -.4], [-.3, .8, -.6, 1.1], [1., 1., 1., 1.]])
def test_two_triangle_history_and_previous_damage_gradients(weights):
    m, mat, fixed, vals = two_triangles()
    op = O.T3Damage(m, mat)
    H, d_prev = irreversibility_state(op, fixed, vals)
    w = np.asarray(weights)
    d_sol = projected(op, H, d_prev, fixed, vals)
    out = A.T3DamageAdjoint(m, mat, max_iter=50).backward(H, d_prev, d_sol, w, mode="H", fixed=fixed, vals=vals)
    h = 1e-5  # keeps every node on its projection branch
    for i in range(2):
        e = np.eye(2)[i]
        fd = w @ (projected(op, H + h * e, d_prev, fixed, vals) - projected(op, H - h * e, d_prev, fixed, vals)) / (2 * h)
        np.testing.assert_allclose(out["H"][i], fd, rtol=1e-8, atol=1e-10)
    fd_prev = np.array([w @ (projected(op, H, d_prev + h * e, fixed, vals)
                             - projected(op, H, d_prev - h * e, fixed, vals)) / (2 * h) for e in np.eye(4)])
    np.testing.assert_allclose(out["d_prev"], fd_prev, rtol=0, atol=1e-9)
    assert out["d_prev"][1] == w[1] and np.all(out["d_prev"][[0, 2, 3]] == 0.0)


def plate(nx=4, ny=3, seed=0):
    nodes, els = O.structured_t3(nx, ny, 5.0, 4.0, jitter=0.2, seed=seed)
    return O.T3Mesh(nodes, els)


def material(pf_model):
    return Material(E=32000.0, nu=0.2, rho=2.45e-9, Gc=2.7, l0=0.5, pf_model=pf_model)


PINS = [None, "pins"]


def pin_set(m):
    fixed = np.zeros(m.n_nodes, dtype=bool)
    fixed[[0, m.n_nodes - 1]] = True
    return fixed, np.where(fixed, 0.35, 0.0)


def interior_state(m, mat, fixed, vals, seed=1):
    rng = np.random.default_rng(seed)
    H = rng.uniform(2.0, 6.0, m.n_elems)
    d_prev = np.zeros(m.n_nodes)
    d = projected(O.T3Damage(m, mat), H, d_prev, fixed, vals)
    lower, upper, _, interior = A.classify_active_set(d, d_prev, fixed)
    assert not lower.any() and not upper.any()
    return H, d_prev, d, rng.uniform(-1.0, 1.0, m.n_nodes)


@pytest.mark.parametrize("pf_model", ["AT2", "AT1"])
@pytest.mark.parametrize("pins", PINS)
def test_scalar_gradients_match_central_differences(pf_model, pins):
    m, mat = plate(), material(pf_model)
    fixed, vals = pin_set(m) if pins else (None, None)
    H, d_prev, d_sol, w = interior_state(m, mat, fixed, vals)
    adj = A.T3DamageAdjoint(m, mat)
    out = adj.backward(H, d_prev, d_sol, w, mode="scalar", Gc=mat.Gc, l0=mat.l0, fixed=fixed, vals=vals)

    def L(Gc=mat.Gc, l0=mat.l0, Hx=H):
        return w @ projected(O.T3Damage(m, A.override_material(mat, Gc, l0)), Hx, d_prev, fixed, vals)

    for name, base in (("Gc", mat.Gc), ("l0", mat.l0)):
        h = 1e-6 * base
        fd = (L(**{name: base + h}) - L(**{name: base - h})) / (2 * h)
        np.testing.assert_allclose(out[name], fd, rtol=1e-7)
    for e in (0, m.n_elems // 2, m.n_elems - 1):
        h = 1e-6 * H[e]
        Hp, Hm = H.copy(), H.copy()
        Hp[e] += h
        Hm[e] -= h
        np.testing.assert_allclose(out["H"][e], (L(Hx=Hp) - L(Hx=Hm)) / (2 * h), rtol=1e-7, atol=1e-12)
    np.testing.assert_array_equal(out["d_prev"], 0.0)


@pytest.mark.parametrize("pf_model", ["AT2", "AT1"])
def test_field_gradients_match_central_differences(pf_model):
    m, mat = plate(), material(pf_model)
    rng = np.random.default_rng(3)
    Gc_e = mat.Gc * (1.0 + 0.3 * rng.uniform(-1.0, 1.0, m.n_elems))
    H = rng.uniform(2.5, 6.0, m.n_elems)
    d_prev = np.zeros(m.n_nodes)

    def dvec(Ge=Gc_e, l0=mat.l0):
        return projected(O.T3Damage(m, A.override_material(mat, None, l0), Ge), H, d_prev)

    d_sol = dvec()
    assert A.classify_active_set(d_sol, d_prev)[3].all()
    w = rng.uniform(-1.0, 1.0, m.n_nodes)
    out = A.T3DamageAdjoint(m, mat).backward(H, d_prev, d_sol, w, mode="field", l0=mat.l0, Gc_e=Gc_e)
    for e in (1, m.n_elems // 3, m.n_elems - 2):
        h = 1e-6 * Gc_e[e]
        Gp, Gm = Gc_e.copy(), Gc_e.copy()
        Gp[e] += h
        Gm[e] -= h
        np.testing.assert_allclose(out["Gc"][e], w @ (dvec(Ge=Gp) - dvec(Ge=Gm)) / (2 * h), rtol=1e-6, atol=1e-13)
    h = 1e-6 * mat.l0
    np.testing.assert_allclose(out["l0"], w @ (dvec(l0=mat.l0 + h) - dvec(l0=mat.l0 - h)) / (2 * h), rtol=1e-7)


@pytest.mark.parametrize("pf_model", ["AT2", "AT1"])
def test_field_gradient_sums_to_the_scalar_one(pf_model):
    """damage_solver.py:880-885: the per-element gradient at a uniform field sums to the
    scalar gradient."""
    m, mat = plate(), material(pf_model)
    H, d_prev, d_sol, w = interior_state(m, mat, None, None)
    adj = A.T3DamageAdjoint(m, mat)
    sc = adj.backward(H, d_prev, d_sol, w, mode="scalar", Gc=mat.Gc, l0=mat.l0)
    fi = adj.backward(H, d_prev, d_sol, w, mode="field", l0=mat.l0, Gc_e=np.full(m.n_elems, mat.Gc))
    np.testing.assert_allclose(fi["Gc"].sum(), sc["Gc"], rtol=1e-9)
    np.testing.assert_allclose(fi["l0"], sc["l0"], rtol=1e-9)
    np.testing.assert_allclose(fi["H"], sc["H"], rtol=1e-9, atol=1e-14)


def test_adjoint_pcg_matches_a_dense_solve_with_pins():
    m, mat = plate(6, 5), material("AT2")
    op = O.T3Damage(m, mat)
    H = np.random.default_rng(4).uniform(0.0, 5.0, m.n_elems)
    fixed, vals = pin_set(m)
    rhs = np.random.default_rng(5).uniform(-1.0, 1.0, m.n_nodes)
    x, it = A.pcg_adjoint(lambda v: op.apply(v, H), rhs, 1.0 / (op.diagonal(H) + 1e-30), fixed, vals,
                          None, 1e-26, 1000)
    K = op.dense(H)
    free = ~fixed
    ref = vals.copy()
    ref[free] = np.linalg.solve(K[np.ix_(free, free)], rhs[free] - K[np.ix_(free, fixed)] @ vals[fixed])
    np.testing.assert_allclose(x, ref, rtol=1e-11, atol=1e-13)
    assert np.all(x[fixed] == vals[fixed]) and it <= m.n_nodes
    x0, _ = A.pcg_adjoint(lambda v: op.apply(v, H), rhs, 1.0 / (op.diagonal(H) + 1e-30), fixed, None)
    assert np.all(x0[fixed] == 0.0)


def test_active_set_precedence():
    """damage_solver.py:70-105: saturated nodes are upper-active, pins take precedence."""
    d_new = np.array([1.0, 0.2, 0.2, 0.5, 0.3])
    d_prev = np.array([1.0, 0.2, 0.2 - 1e-13, 0.1, 0.3])
    fixed = np.array([False, False, False, False, True])
    lower, upper, fx, interior = A.classify_active_set(d_new, d_prev, fixed)
    np.testing.assert_array_equal(upper, [True, False, False, False, False])
    np.testing.assert_array_equal(lower, [False, True, True, False, False])
    np.testing.assert_array_equal(interior, [False, False, False, True, False])
    g, gp = A.masked_adjoint_inputs(np.arange(1.0, 6.0), d_new, d_prev, fixed)
    np.testing.assert_array_equal(g, [0, 0, 0, 4, 0])
    np.testing.assert_array_equal(gp, [0, 2, 3, 0, 0])


def test_scalar_override_resets_the_at1_threshold():
    """The scalar rule installs S_H = 3 Gc/(8 l0) (damage_solver.py:580-583)."""
    mat = dataclasses.replace(material("AT1"), at1_threshold=0.4)
    assert mat.at1_source == 0.4
    o = A.override_material(mat, 3.0, 0.25)
    assert o.at1_source == pytest.approx(3.0 * 3.0 / (8.0 * 0.25), rel=1e-15)


def test_adjoint_pcg_stopping_edges():
    """max_iter = 0 returns the start (pins at their values, 0 elsewhere); a zero right-hand
    side converges at once; a zero operator stops on p.Ap <= 0 before any update."""
    m, mat = plate(), material("AT2")
    op = O.T3Damage(m, mat)
    H = np.full(m.n_elems, 1.0)
    Minv = 1.0 / (op.diagonal(H) + 1e-30)
    fixed, vals = pin_set(m)
    rhs = np.random.default_rng(9).uniform(-1.0, 1.0, m.n_nodes)
    x, it = A.pcg_adjoint(lambda v: op.apply(v, H), rhs, Minv, fixed, vals, None, 1e-22, 0)
    assert it == 0 and np.array_equal(x, np.where(fixed, vals, 0.0))
    x, it = A.pcg_adjoint(lambda v: op.apply(v, H), np.zeros(m.n_nodes), Minv, None, None, None, 1e-22, 50)
    assert it == 1 and np.all(x == 0.0)
    x, it = A.pcg_adjoint(lambda v: np.zeros_like(v), rhs, Minv, None, None, None, 1e-22, 50)
    assert it == 1 and np.all(x == 0.0)
