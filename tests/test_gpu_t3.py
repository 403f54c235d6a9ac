"""Parity of the T3 (linear triangle) HIP path with the oracle (GPU).

The T3 kernels round every product and sum on its own (no FMA contraction) and sum each
node's corners in the reference's scatter order, so the operators -- internal force, psi+,
the history max, the explicit step, the damage apply / rhs / diagonal and the element-to-node
projection -- equal the oracle bit for bit.  Solves, whose dot products sum in another
order, are held to 1e-10 relative (max-norm relative to the field's max-norm) and to the
oracle's iteration counts.
"""
import numpy as np
import pytest
import torch

from oracle import t3 as O
from oracle.q4 import Material as OMaterial

pytestmark = pytest.mark.gpu
RTOL = 1e-10
SPLITS = [("amor", False), ("amor", True), ("isotropic", False), ("isotropic", True), ("spectral", False)]


def exact(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return np.array_equal(a, np.asarray(b))


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    b = b.detach().cpu().numpy() if isinstance(b, torch.Tensor) else np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


T = lambda x, dev: torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device=dev)


def scrambled_mesh(nx, ny, seed=0, W=5.0, H=4.0):
    """Jittered structured triangles with node numbering, element order and corner order
    randomly permuted (so half the triangles are clockwise and incidence lists are ragged)."""
    rng = np.random.default_rng(seed)
    nodes, els = O.structured_t3(nx, ny, W, H, jitter=0.2, seed=seed)
    perm = rng.permutation(len(nodes))
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(nodes))
    nodes = nodes[perm]
    els = inv[els][rng.permutation(len(els))]
    rot = rng.integers(0, 3, len(els))
    els = np.stack([np.roll(e, r) for e, r in zip(els, rot)])
    flip = rng.random(len(els)) < 0.5
    els[flip] = els[flip][:, ::-1]
    return nodes, els


def setup(nodes, els, split="amor", plane_stress=False):
    from phast_amd import t3 as P
    kw = dict(E=32000.0, nu=0.2, rho=2.45e-9, Gc=3e-3, l0=0.5, energy_split=split, plane_stress=plane_stress)
    fem = P.FEMOperators(P.FEMMesh(nodes, els), P.Material(**kw))
    ofem = O.T3Operators(O.T3Mesh(nodes, els), OMaterial(**kw))
    return P, fem, ofem


def state(n, E, seed=0):
    rng = np.random.default_rng(seed)
    u = rng.uniform(-1e-4, 1e-4, (n, 2))
    v = rng.uniform(-1e-1, 1e-1, (n, 2))
    a = rng.uniform(-1e3, 1e3, (n, 2))
    d = rng.uniform(0, 1, n)
    H = rng.uniform(0, 1e-4, E)
    return u, v, a, d, H


@pytest.mark.parametrize("split,ps", SPLITS)
def test_internal_force_psi_history(dev, split, ps):
    nodes, els = scrambled_mesh(23, 17, seed=1)
    P, fem, ofem = setup(nodes, els, split, ps)
    u, v, a, d, H = state(len(nodes), len(els))
    f = fem.internal_force(T(u, dev), T(d, dev))
    assert exact(f, ofem.internal_force(u, d))
    f2, psi, Hn = fem.internal_force_and_history(T(u, dev), T(d, dev), T(H, dev))
    assert torch.equal(f2, f)
    assert exact(psi, ofem.compute_psi_plus(u))
    assert np.array_equal(Hn.cpu().numpy(), np.maximum(H, psi.cpu().numpy()))
    assert (Hn.cpu().numpy() > H).any() and (Hn.cpu().numpy() == H).any()


@pytest.mark.parametrize("split,ps", [("amor", True), ("spectral", False), ("isotropic", False)])
def test_element_fields(dev, split, ps):
    nodes, els = scrambled_mesh(21, 19, seed=22)
    P, fem, ofem = setup(nodes, els, split, ps)
    u, v, a, d, H = state(len(nodes), len(els), seed=23)
    s = fem.compute_strain(T(u, dev))
    for x, y in zip(s, ofem.compute_strain(u)):
        assert exact(x, y)
    for x, y in zip(fem.compute_stress(T(u, dev), T(d, dev)), ofem.compute_stress(u, d)):
        assert exact(x, y)
    assert exact(fem.compute_psi_plus(T(u, dev)), ofem.compute_psi_plus(u))


@pytest.mark.parametrize("split", ["amor", "spectral"])
def test_full_stress_degradation(dev, split):
    """stress_degradation='full': isotropic stress, the split's psi+ (material.py:367-376)."""
    from phast_amd import t3 as P
    nodes, els = scrambled_mesh(17, 15, seed=25)
    kw = dict(E=32000.0, nu=0.2, rho=2.45e-9, Gc=3e-3, l0=0.5, energy_split=split, stress_degradation="full")
    fem = P.FEMOperators(P.FEMMesh(nodes, els), P.Material(**kw))
    ofem = O.T3Operators(O.T3Mesh(nodes, els), OMaterial(**kw))
    u, v, a, d, H = state(len(nodes), len(els), seed=26)
    assert exact(fem.internal_force(T(u, dev), T(d, dev)), ofem.internal_force(u, d))
    for x, y in zip(fem.compute_stress(T(u, dev), T(d, dev)), ofem.compute_stress(u, d)):
        assert exact(x, y)
    assert exact(fem.compute_psi_plus(T(u, dev)), ofem.compute_psi_plus(u))


def test_mesh_precompute(dev):
    nodes, els = scrambled_mesh(31, 12, seed=2)
    P, fem, ofem = setup(nodes, els)
    assert rel(fem.M_scalar, ofem.mesh.M_scalar) < 1e-14
    assert fem.dt_cfl == pytest.approx(ofem.dt_cfl, rel=1e-14)


@pytest.mark.parametrize("split,ps", [("amor", False), ("spectral", False), ("isotropic", True)])
def test_explicit_step(dev, split, ps):
    nodes, els = scrambled_mesh(19, 26, seed=3)
    P, fem, ofem = setup(nodes, els, split, ps)
    n, E = len(nodes), len(els)
    u, v, a, d, H = state(n, E, seed=4)
    rng = np.random.default_rng(5)
    f_ext = rng.uniform(-1e-2, 1e-2, (n, 2))
    bc = rng.random((n, 2)) < 0.1
    bcv = rng.uniform(-1e-4, 1e-4, (n, 2))
    dt = 0.4 * fem.dt_cfl
    dyn = P.ExplicitDynamics(fem, dt=dt)
    u1, v1, a1, Hn, psi = dyn.step_with_history(T(u, dev), T(v, dev), T(a, dev), T(d, dev), T(f_ext, dev),
                                                torch.as_tensor(bc, device=dev), T(bcv, dev), T(H, dev))
    # oracle step with the T3 operators (mechanics_solver.py:153-239)
    ou = (u + dt * v) + (0.5 * dt * dt) * a
    ou = np.where(bc, bcv, ou)
    oa = ofem.M_vec_inv.reshape(-1, 2) * (f_ext - ofem.internal_force(ou, d))
    ov = v + 0.5 * dt * (a + oa)
    oa, ov = np.where(bc, 0.0, oa), np.where(bc, 0.0, ov)
    assert exact(u1, ou) and exact(v1, ov) and exact(a1, oa)
    assert np.array_equal(u1.cpu().numpy()[bc], bcv[bc])
    assert (a1.cpu().numpy()[bc] == 0).all() and (v1.cpu().numpy()[bc] == 0).all()
    assert exact(psi, ofem.compute_psi_plus(ou))
    assert np.array_equal(Hn.cpu().numpy(), np.maximum(H, psi.cpu().numpy()))
    # the plain step (no f_ext, no BC) agrees with step_with_history's kinematics
    s1, s2, s3 = dyn.step(T(u, dev), T(v, dev), T(a, dev), T(d, dev), None)
    ou = (u + dt * v) + (0.5 * dt * dt) * a
    oa = ofem.M_vec_inv.reshape(-1, 2) * (-ofem.internal_force(ou, d))
    assert exact(s1, ou) and exact(s3, oa) and exact(s2, v + 0.5 * dt * (a + oa))


def test_kelvin_voigt_damped_step(dev):
    """damping_ratio_max > 0 (mechanics_solver.py:192-195), bit-exact against the oracle."""
    from oracle import q4 as OQ
    nodes, els = scrambled_mesh(21, 16, seed=27)
    P, fem, ofem = setup(nodes, els, "amor")
    n, E = len(nodes), len(els)
    u, v, a, d, H = state(n, E, seed=28)
    f_ext = np.random.default_rng(29).uniform(-1e-2, 1e-2, (n, 2))
    dyn = P.ExplicitDynamics(fem, damping_ratio_max=0.1)
    u1, v1, a1 = dyn.step(T(u, dev), T(v, dev), T(a, dev), T(d, dev), T(f_ext, dev))
    ou, ov, oa = OQ.ExplicitDynamics(ofem, dt=dyn.dt, damping_ratio_max=0.1).step(u, v, a, d, f_ext)
    assert exact(u1, ou) and exact(v1, ov) and exact(a1, oa)


def test_damage_operator(dev):
    nodes, els = scrambled_mesh(27, 21, seed=6)
    P, fem, ofem = setup(nodes, els)
    _, _, _, d, H = state(len(nodes), len(els), seed=7)
    H = H * 30.0
    Gc_e = np.random.default_rng(8).uniform(1e-3, 5e-3, len(els))
    for gc in (None, Gc_e):
        dmg = P.PhaseFieldDamageSolver(fem, Gc_field=None if gc is None else T(gc, dev))
        odmg = O.T3Damage(ofem.mesh, ofem.material, gc)
        assert exact(dmg.damage_apply(T(H, dev), T(d, dev)), odmg.apply(d, H))
        b, dg = dmg.damage_rhs_diag(T(H, dev))
        assert exact(b, odmg.rhs(H)) and exact(dg, odmg.diagonal(H))


def test_at1_damage(dev):
    """AT1 (damage_solver.py:1321-1336): operator pieces bit-exact, solve to the oracle's iterations."""
    from phast_amd import t3 as P
    nodes, els = scrambled_mesh(23, 19, seed=30)
    kw = dict(E=32000.0, nu=0.2, rho=2.45e-9, Gc=3e-3, l0=0.5, pf_model="AT1")
    fem = P.FEMOperators(P.FEMMesh(nodes, els), P.Material(**kw))
    omesh, omat = O.T3Mesh(nodes, els), OMaterial(**kw)
    rng = np.random.default_rng(31)
    n, E = len(nodes), len(els)
    H = rng.uniform(0, 0.02, E)
    d = rng.uniform(0, 1, n)
    Gc_e = rng.uniform(1e-3, 5e-3, E)
    for gc in (None, Gc_e):
        dmg = P.PhaseFieldDamageSolver(fem, Gc_field=None if gc is None else T(gc, dev))
        od = O.T3Damage(omesh, omat, gc)
        assert exact(dmg.damage_apply(T(H, dev), T(d, dev)), od.apply(d, H))
        b, dg = dmg.damage_rhs_diag(T(H, dev))
        assert exact(b, od.rhs(H)) and exact(dg, od.diagonal(H))
    d_prev = np.clip(rng.uniform(-0.3, 0.5, n), 0, 1)
    dmg = P.PhaseFieldDamageSolver(fem, tol=1e-9, max_iter=3000, bounds_method="projected_cg")
    dd = dmg.solve(T(H, dev), T(d_prev, dev))
    osol = O.T3DamageSolver(omesh, omat, tol=1e-9, max_iter=3000)
    ods = osol.solve(H, d_prev)
    assert dmg.last_iter == osol.last_iter and rel(dd, ods) < RTOL
    assert np.all(dd.cpu().numpy() >= d_prev)


def test_damage_operator_dense_small(dev):
    """Columns of the device operator on a tiny mesh form the symmetric textbook matrix."""
    nodes, els = scrambled_mesh(3, 2, seed=8)
    P, fem, ofem = setup(nodes, els)
    dmg = P.PhaseFieldDamageSolver(fem)
    H = np.random.default_rng(9).uniform(0, 1e-3, len(els))
    n = len(nodes)
    cols = [dmg.damage_apply(T(H, dev), T(np.eye(n)[k], dev)).cpu().numpy() for k in range(n)]
    D = np.stack(cols, axis=1)
    K = O.T3Damage(ofem.mesh, ofem.material).dense(H)
    assert np.abs(D - K).max() < 1e-13 * np.abs(K).max()
    assert np.abs(D - D.T).max() < 1e-14 * np.abs(D).max()
    assert rel(dmg.damage_rhs_diag(T(H, dev))[1], np.diag(K)) < 1e-13


def test_single_triangle_and_unused_node(dev):
    """Ragged edge cases: one element, and a node no element touches (zero force/mass)."""
    nodes = np.array([[0.0, 0.0], [2.0, 0.0], [0.0, 3.0], [5.0, 5.0]])
    els = np.array([[0, 2, 1]])  # clockwise
    P, fem, ofem = setup(nodes, els, "spectral")
    u = np.array([[0.0, 0.0], [2e-4, 1e-5], [-3e-5, -1e-4], [1.0, 1.0]])
    d = np.array([0.1, 0.2, 0.3, 0.9])
    f = fem.internal_force(T(u, dev), T(d, dev)).cpu().numpy()
    assert exact(f, ofem.internal_force(u, d))
    assert (f[3] == 0.0).all()
    assert fem.M_scalar.cpu().numpy()[3] == 0.0


def test_large_mesh(dev):
    """A 1024 x 512-cell mesh (1M triangles) against the oracle, plus the patch and
    uniform-history properties that hold at any size."""
    nodes, els = O.structured_t3(1024, 512, 20.0, 10.0, jitter=0.25, seed=10)
    P, fem, ofem = setup(nodes, els, "amor")
    n, E = len(nodes), len(els)
    u, v, a, d, H = state(n, E, seed=11)
    assert exact(fem.internal_force(T(u, dev), T(d, dev)), ofem.internal_force(u, d))
    G = np.array([[1.1e-4, -0.4e-4], [0.7e-4, -0.5e-4]])
    ul = nodes @ G.T
    f = fem.internal_force(T(ul, dev), T(d * 0, dev)).cpu().numpy()
    xs, ys = nodes[:, 0], nodes[:, 1]
    inner = (xs > 1e-9) & (xs < 20 - 1e-9) & (ys > 1e-9) & (ys < 10 - 1e-9)
    assert np.abs(f[inner]).max() < 1e-9 * np.abs(f).max()
    dmg = P.PhaseFieldDamageSolver(fem)
    Hc = 2e-4
    Hu = T(np.full(E, Hc), dev)
    dc = 2 * Hc / (3e-3 / 0.5 + 2 * Hc)
    assert rel(dmg.damage_apply(Hu, T(np.full(n, dc), dev)), dmg.damage_rhs_diag(Hu)[0]) < 1e-12


def test_uniform_history_solve_closed_form(dev):
    """Uniform H: the solve returns d = 2H / (Gc/l0 + 2H) at every node."""
    nodes, els = scrambled_mesh(40, 30, seed=16)
    P, fem, ofem = setup(nodes, els)
    n, E = len(nodes), len(els)
    Hc = 1.3 * 3e-3 / 0.5
    dmg = P.PhaseFieldDamageSolver(fem, tol=1e-13, max_iter=1000, bounds_method="projected_cg")
    d = dmg.solve(torch.full((E,), Hc, dtype=torch.float64, device=dev),
                  torch.zeros(n, dtype=torch.float64, device=dev))
    assert dmg.last_status == "converged"
    assert rel(d, np.full(n, 2 * Hc / (3e-3 / 0.5 + 2 * Hc))) < 1e-10


@pytest.mark.parametrize("method", ["projected_cg", "post_clamp"])
def test_damage_solve(dev, method):
    nodes, els = scrambled_mesh(33, 21, seed=12)
    P, fem, ofem = setup(nodes, els)
    n, E = len(nodes), len(els)
    rng = np.random.default_rng(13)
    H = rng.uniform(0, 0.05, E) * (rng.uniform(0, 1, E) > 0.7)
    d_prev = np.clip(rng.uniform(-0.5, 0.8, n), 0, 1)
    dmg = P.PhaseFieldDamageSolver(fem, tol=1e-8, max_iter=3000, bounds_method=method)
    odmg = O.T3DamageSolver(ofem.mesh, ofem.material, tol=1e-8, max_iter=3000, bounds_method=method)
    d = dmg.solve(T(H, dev), T(d_prev, dev))
    od = odmg.solve(H, d_prev)
    assert dmg.last_iter == odmg.last_iter, (dmg.last_iter, odmg.last_iter, dmg.last_status, odmg.last_status)
    assert dmg.last_status == odmg.last_status
    assert rel(d, od) < RTOL
    dc = d.cpu().numpy()
    assert np.all(dc >= d_prev) and np.all(dc <= 1.0)  # bound clamps exact
    if method == "projected_cg":
        assert np.array_equal(dmg.last_active.cpu().numpy(), odmg.last_active)


def test_damage_solve_gc_field_pins_and_initial_guess(dev):
    nodes, els = scrambled_mesh(21, 14, seed=14)
    P, fem, ofem = setup(nodes, els)
    n, E = len(nodes), len(els)
    rng = np.random.default_rng(15)
    H = rng.uniform(0, 0.02, E)
    d_prev = rng.uniform(0, 0.3, n)
    Gc_e = rng.uniform(1e-3, 6e-3, E)
    fixed = nodes[:, 0] < 1e-9
    vals = np.where(fixed, 1.0, 0.0)
    init = np.clip(d_prev + 0.1, 0, 1)
    dmg = P.PhaseFieldDamageSolver(fem, tol=1e-12, max_iter=2000, bounds_method="projected_cg",
                                   Gc_field=T(Gc_e, dev))
    d = dmg.solve(T(H, dev), T(d_prev, dev), pf_dirichlet_mask=torch.as_tensor(fixed, device=dev),
                  pf_dirichlet_values=T(vals, dev), initial_guess=T(init, dev))
    od = O.T3DamageSolver(ofem.mesh, ofem.material, tol=1e-12, max_iter=2000, Gc_e=Gc_e).solve(
        H, d_prev, initial_guess=init, fixed=fixed, vals=vals)
    assert rel(d, od) < RTOL
    assert np.all(d.cpu().numpy()[fixed] == 1.0)


def test_elem_to_node(dev):
    nodes, els = scrambled_mesh(17, 13, seed=17)
    P, fem, ofem = setup(nodes, els)
    f = np.random.default_rng(18).uniform(0, 1, len(els))
    assert exact(fem.elem_to_node(T(f, dev)), ofem.mesh.elem_to_node(f))
    np.testing.assert_allclose(fem.elem_to_node(T(np.full(len(els), 2.5), dev)).cpu().numpy(), 2.5, rtol=1e-13)


def test_explicit_staggered_steps_track_oracle(dev):
    """30 explicit steps of the staggered solver (mechanics + history + projected damage solve)
    on a notched split-quad plate pulled at top and bottom."""
    from oracle import q4 as OQ
    P3 = __import__("phast_amd.t3", fromlist=["t3"])
    mesh = P3.FEMMesh.sent_split_quad(10.0, 4.0, 40, 16, a=5.0)
    kw = dict(E=32000.0, nu=0.2, rho=2.45e-9, Gc=3e-3, l0=0.5)
    fem = P3.FEMOperators(mesh, P3.Material(**kw))
    ofem = O.T3Operators(O.T3Mesh(mesh.nodes, mesh.elements), OMaterial(**kw))
    n, E = mesh.n_nodes, mesh.n_elems
    f_ext = np.zeros((n, 2))
    f_ext[mesh.node_sets["top"], 1] = 40.0
    f_ext[mesh.node_sets["bottom"], 1] = -40.0
    dt = 0.5 * fem.dt_cfl
    dyn = P3.ExplicitDynamics(fem, dt=dt)
    dmg = P3.PhaseFieldDamageSolver(fem, tol=1e-12, max_iter=2000, bounds_method="projected_cg")
    st = P3.StaggeredSolver(fem, dyn, dmg, f_ext=T(f_ext, dev), damage_every=1)
    d0 = np.zeros(n)
    d0[mesh.node_sets["crack"]] = 1.0  # preseeded notch
    st.d = T(d0, dev)
    ost = OQ.ExplicitState(u=np.zeros((n, 2)), v=np.zeros((n, 2)), a=np.zeros((n, 2)), d=d0.copy(), H=np.zeros(E))
    odyn = OQ.ExplicitDynamics(ofem, dt=dt)
    odmg = O.T3DamageSolver(ofem.mesh, ofem.material, tol=1e-12, max_iter=2000)
    for k in range(30):
        psi = st.step_full()
        opsi = OQ.step_full_explicit(odyn, odmg, ost, f_ext)
    print(f"30 steps: psi {rel(psi, opsi):.1e} u {rel(st.u, ost.u):.1e} v {rel(st.v, ost.v):.1e} "
          f"d {rel(st.d, ost.d):.1e} H {rel(st.H_elem, ost.H):.1e}")
    assert rel(psi, opsi) < 1e-10
    assert rel(st.u, ost.u) < 1e-10 and rel(st.v, ost.v) < 1e-10
    assert rel(st.d, ost.d) < 1e-10 and rel(st.H_elem, ost.H) < 1e-10
    assert float(st.d[torch.as_tensor(d0 == 0, device=dev)].max()) > 0.0
    assert rel(st.H_nodal, ofem.mesh.elem_to_node(ost.H)) < 1e-10


@pytest.mark.parametrize("which", ["scrambled", "split_quad", "one_triangle"])
def test_patch_and_recompute_organisations_agree_bitwise(dev, which):
    """The element-patch passes (default) and the per-incidence recompute passes give the same
    bits for every operator; the damage solve agrees to its tolerance."""
    from phast_amd import t3 as P3
    if which == "scrambled":
        nodes, els = scrambled_mesh(45, 37, seed=19)
    elif which == "split_quad":
        m = P3.FEMMesh.sent_split_quad(12.0, 6.0, 96, 48, a=6.0)
        nodes, els = m.nodes, m.elements
    else:
        nodes, els = np.array([[0.0, 0.0], [2.0, 0.0], [0.0, 3.0]]), np.array([[0, 2, 1]])
    n, E = len(nodes), len(els)
    rng = np.random.default_rng(20)
    u, v, a, d, H = state(n, E, seed=21)
    Gc_e = rng.uniform(1e-3, 5e-3, E)
    bc = rng.random((n, 2)) < 0.1
    outs = {}
    for split in ("amor", "spectral"):
        P, fem, _ = setup(nodes, els, split, split == "amor")
        assert fem.path == 0 and fem.max_patch_list > 0
        for path in (0, 1):
            assert fem.set_path(path) == path
            r = [fem.internal_force(T(u, dev), T(d, dev))]
            r += list(fem.internal_force_and_history(T(u, dev), T(d, dev), T(H, dev)))
            dyn = P.ExplicitDynamics(fem, dt=0.3 * fem.dt_cfl)
            r += list(dyn.step_with_history(T(u, dev), T(v, dev), T(a, dev), T(d, dev), T(u * 1e-2, dev),
                                            torch.as_tensor(bc, device=dev), T(u, dev), T(H, dev)))
            for gc in (None, T(Gc_e, dev)):
                dmg = P.PhaseFieldDamageSolver(fem, Gc_field=gc)
                r.append(dmg.damage_apply(T(H * 30, dev), T(d, dev)))
                r += list(dmg.damage_rhs_diag(T(H * 30, dev)))
            outs[(split, path)] = [t.cpu().numpy() for t in r]
            dmg = P.PhaseFieldDamageSolver(fem, tol=1e-10, max_iter=3000, bounds_method="projected_cg")
            dsol = dmg.solve(T(H * 300, dev), T(np.clip(d - 0.5, 0, 1), dev))
            outs[(split, path, "solve")] = (dsol.cpu().numpy(), dmg.last_iter, dmg.last_status)
        for k, (x0, x1) in enumerate(zip(outs[(split, 0)], outs[(split, 1)])):
            assert np.array_equal(x0, x1), (split, k)
        s0, s1 = outs[(split, 0, "solve")], outs[(split, 1, "solve")]
        assert s0[2] == s1[2] == "converged" and abs(s0[1] - s1[1]) <= 1
        assert rel(s0[0], s1[0]) < 1e-8


def test_fan_mesh_beyond_the_patch_budget_uses_the_recompute_passes(dev):
    """A fan of 3000 triangles around one node: that node's patch lists more triangles than
    the LDS budget holds, so the plan has no patches and every pass recomputes per incidence;
    results still equal the oracle bit for bit."""
    k = 3000
    ang = np.linspace(0.0, 2 * np.pi, k, endpoint=False)
    nodes = np.vstack([[0.0, 0.0], np.stack([np.cos(ang), np.sin(ang)], 1)])
    els = np.stack([np.zeros(k, int), 1 + np.arange(k), 1 + (np.arange(k) + 1) % k], 1)
    P, fem, ofem = setup(nodes, els, "amor")
    assert fem.path == 1 and fem.set_path(0) == 1 and fem.max_patch_list > 1280
    u, v, a, d, H = state(len(nodes), len(els), seed=24)
    u *= 1e-3
    assert exact(fem.internal_force(T(u, dev), T(d, dev)), ofem.internal_force(u, d))
    dmg = P.PhaseFieldDamageSolver(fem)
    assert exact(dmg.damage_apply(T(H, dev), T(d, dev)), O.T3Damage(ofem.mesh, ofem.material).apply(d, H))


def test_plan_rejects_bad_connectivity(dev):
    from phast_amd import t3 as P
    from phast_amd._lib import PhastLibraryError
    nodes = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    with pytest.raises(ValueError):
        P.FEMMesh(nodes, np.array([[0, 1, 3]]))
    mat = P.Material(E=1.0, nu=0.2, rho=1.0, Gc=1.0, l0=1.0)
    with pytest.raises((ValueError, PhastLibraryError)):
        P.T3Plan(nodes, np.array([[0, 1, 7]]), mat)
