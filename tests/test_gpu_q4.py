"""Parity of the native-Q4 HIP path with the oracle (GPU).

Tolerances: fields within 1e-10 relative (max-norm relative to the
field's max-norm); history maxima, masks and clamps bit-exact.
"""
import numpy as np
import pytest
import torch

from oracle import q4 as O

pytestmark = pytest.mark.gpu
RTOL = 1e-10


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    b = b.detach().cpu().numpy() if isinstance(b, torch.Tensor) else np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def setup(nx=37, ny=23, W=3.7, H=2.3, split="amor", plane_stress=False, tile=(128, 8)):
    from phast_amd import q4 as P
    mat_kw = dict(E=32000.0, nu=0.2, rho=2.45e-9, Gc=3e-3, l0=0.5, energy_split=split, plane_stress=plane_stress)
    mesh = P.FEMMesh.rectangle(W, H, nx, ny)
    fem = P.FEMOperators(mesh, P.Material(**mat_kw), tile=tile)
    nodes, quads = O.structured_q4_mesh(W, H, nx, ny)
    omesh = O.Q4Mesh(nodes, quads)
    ofem = O.FEMOperators(omesh, O.Material(**mat_kw))
    return P, fem, ofem


def state(n, E, seed=0):
    rng = np.random.default_rng(seed)
    u = rng.uniform(-1e-4, 1e-4, (n, 2))
    v = rng.uniform(-1e-1, 1e-1, (n, 2))
    a = rng.uniform(-1e3, 1e3, (n, 2))
    d = rng.uniform(0, 1, n)
    H = rng.uniform(0, 1, (E, 4))
    return u, v, a, d, H


T = lambda x, dev: torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device=dev)


@pytest.mark.parametrize("tile", [(64, 4), (128, 8), (256, 3), (64, 100)])
@pytest.mark.parametrize("split,ps", [("amor", False), ("amor", True), ("isotropic", False)])
def test_internal_force(dev, tile, split, ps):
    P, fem, ofem = setup(split=split, plane_stress=ps, tile=tile)
    u, v, a, d, H = state(fem.mesh.n_nodes, fem.mesh.n_elems)
    f = fem.internal_force(T(u, dev), T(d, dev))
    assert rel(f, ofem.internal_force(u, d)) < RTOL


@pytest.mark.parametrize("tile", [(64, 7), (128, 16), (256, 1)])
def test_multi_strip_paths(dev, tile):
    """Plates wider and taller than one strip: strip seams in x and y, every path."""
    P, fem, ofem = setup(nx=200, ny=45, W=20.0, H=4.5, tile=tile)
    n, E = fem.mesh.n_nodes, fem.mesh.n_elems
    u, v, a, d, H = state(n, E, seed=11)
    assert rel(fem.internal_force(T(u, dev), T(d, dev)), ofem.internal_force(u, d)) < RTOL
    # the oracle's node coordinates (multiples of 0.1) carry rounding the plan's constants do not
    assert rel(fem.M_scalar, ofem.mesh.M_scalar) < RTOL
    dt = 0.3 * fem.dt_cfl
    dyn = P.ExplicitDynamics(fem, dt=dt)
    z = np.zeros((n, 2))
    u1, v1, a1, Hn, psi = dyn.step_with_history(T(u, dev), T(v, dev), T(a, dev), T(d, dev), T(z, dev), None, None,
                                                T(H * 1e-4, dev))
    ou, ov, oa = O.ExplicitDynamics(ofem, dt=dt).step(u, v, a, d, z, None, None)
    assert rel(u1, ou) < RTOL and rel(v1, ov) < RTOL and rel(a1, oa) < RTOL
    assert rel(psi, ofem.compute_psi_plus(ou)) < RTOL
    assert np.array_equal(Hn.cpu().numpy(), np.maximum(H * 1e-4, psi.cpu().numpy()))
    dmg = P.PhaseFieldDamageSolver(fem)
    odmg = O.PhaseFieldDamageSolver(ofem)
    Hd = H * 5e-2
    rho, ob, ominv = odmg.prepare(Hd, np.zeros(n))
    b, diag = dmg.damage_rhs_diag(T(Hd, dev))
    assert rel(b, ob) < RTOL and rel(diag, 1.0 / ominv - 1e-30) < RTOL
    assert rel(dmg.damage_apply(T(Hd, dev), T(d, dev)), odmg.Ax(d, rho)) < RTOL
    assert rel(fem.elem_to_node(T(H, dev)), ofem.mesh.elem_to_node(H)) < RTOL


def test_mass_strain_psi_history(dev):
    P, fem, ofem = setup()
    u, v, a, d, H = state(fem.mesh.n_nodes, fem.mesh.n_elems, seed=1)
    assert rel(fem.M_scalar, ofem.mesh.M_scalar) < 1e-14
    assert rel(fem.M_vec_inv, ofem.M_vec_inv) < 1e-14
    for g, o in zip(fem.compute_strain(T(u, dev)), ofem.compute_strain(u)):
        assert rel(g, o) < RTOL
    psi, Hn = fem.driving_force_and_history(T(u, dev), T(H * 1e-3, dev))
    assert rel(psi, ofem.compute_psi_plus(u)) < RTOL
    # H = max(H_old, psi) entrywise one operand bit-exactly, against the GPU's own psi
    Hc, pc, Ho = Hn.cpu().numpy(), psi.cpu().numpy(), H * 1e-3
    assert np.array_equal(Hc, np.maximum(Ho, pc))
    Hm = P.history_update(T(Ho, dev), psi)
    assert torch.equal(Hm, Hn)
    assert rel(fem.elem_to_node(T(H, dev)), ofem.mesh.elem_to_node(H)) < RTOL


def test_explicit_step_with_bc_and_history(dev):
    P, fem, ofem = setup()
    n, E = fem.mesh.n_nodes, fem.mesh.n_elems
    u, v, a, d, H = state(n, E, seed=2)
    rng = np.random.default_rng(3)
    f_ext = rng.uniform(-1, 1, (n, 2))
    bcm = np.zeros((n, 2), dtype=bool)
    bcm[fem.mesh.node_sets["left"], 0] = True
    bcm[fem.mesh.node_sets["bottom"][:5], 1] = True
    bcv = rng.uniform(-1e-5, 1e-5, (n, 2))
    dt = 0.3 * fem.dt_cfl
    dyn = P.ExplicitDynamics(fem, dt=dt)
    u1, v1, a1, Hn, psi = dyn.step_with_history(T(u, dev), T(v, dev), T(a, dev), T(d, dev), T(f_ext, dev),
                                                torch.as_tensor(bcm, device=dev), T(bcv, dev), T(H * 1e-4, dev))
    odyn = O.ExplicitDynamics(ofem, dt=dt)
    ou, ov, oa = odyn.step(u, v, a, d, f_ext, bcm, bcv)
    assert rel(u1, ou) < RTOL and rel(v1, ov) < RTOL and rel(a1, oa) < RTOL
    assert np.array_equal(u1.cpu().numpy()[bcm], bcv[bcm])
    assert np.all(a1.cpu().numpy()[bcm] == 0) and np.all(v1.cpu().numpy()[bcm] == 0)
    opsi = ofem.compute_psi_plus(ou)
    assert rel(psi, opsi) < RTOL
    assert np.array_equal(Hn.cpu().numpy(), np.maximum(H * 1e-4, psi.cpu().numpy()))
    u2, v2, a2 = dyn.step(T(u, dev), T(v, dev), T(a, dev), T(d, dev), T(f_ext, dev), torch.as_tensor(bcm, device=dev),
                          T(bcv, dev))
    # the kernel without the history epilogue may contract FMAs differently: agree to rounding
    assert rel(u2, u1) < 1e-14 and rel(v2, v1) < 1e-14 and rel(a2, a1) < 1e-14


@pytest.mark.parametrize("gc_field", [False, True])
def test_damage_operator(dev, gc_field):
    P, fem, ofem = setup(nx=29, ny=17)
    n, E = fem.mesh.n_nodes, fem.mesh.n_elems
    rng = np.random.default_rng(4)
    H = rng.uniform(0, 5e-2, (E, 4))
    Gc_e = rng.lognormal(np.log(3e-3), 0.2, E) if gc_field else None
    dmg = P.PhaseFieldDamageSolver(fem, Gc_field=None if Gc_e is None else T(Gc_e, dev))
    odmg = O.PhaseFieldDamageSolver(ofem, Gc_e=Gc_e)
    rho, ob, ominv = odmg.prepare(H, np.zeros(n))
    b, diag = dmg.damage_rhs_diag(T(H, dev))
    assert rel(b, ob) < RTOL and rel(diag, 1.0 / ominv - 1e-30) < RTOL
    x = rng.uniform(0, 1, n)
    assert rel(dmg.damage_apply(T(H, dev), T(x, dev)), odmg.Ax(x, rho)) < RTOL


@pytest.mark.parametrize("method", ["projected_cg", "post_clamp"])
@pytest.mark.parametrize("tol", [1e-5, 1e-12])
def test_damage_solve(dev, method, tol):
    P, fem, ofem = setup(nx=41, ny=19)
    n, E = fem.mesh.n_nodes, fem.mesh.n_elems
    rng = np.random.default_rng(5)
    H = rng.uniform(0, 0.5, (E, 4)) * (rng.uniform(0, 1, (E, 1)) > 0.7)
    d_prev = np.clip(rng.uniform(-0.5, 0.8, n), 0, 1)
    dmg = P.PhaseFieldDamageSolver(fem, tol=tol, max_iter=3000, bounds_method=method)
    odmg = O.PhaseFieldDamageSolver(ofem, tol=tol, max_iter=3000, bounds_method=method)
    d = dmg.solve(T(H, dev), T(d_prev, dev))
    od = odmg.solve(H, d_prev)
    assert dmg.last_iter == odmg.last_iter, (dmg.last_iter, odmg.last_iter, dmg.last_status, odmg.last_status)
    assert rel(d, od) < RTOL
    dc = d.cpu().numpy()
    assert np.all(dc >= d_prev) and np.all(dc <= 1.0)  # bound clamps exact
    if method == "projected_cg":
        assert np.array_equal(dmg.last_active.cpu().numpy(), odmg.last_active)


def test_damage_solve_pf_dirichlet_and_initial_guess(dev):
    P, fem, ofem = setup(nx=21, ny=11)
    n, E = fem.mesh.n_nodes, fem.mesh.n_elems
    rng = np.random.default_rng(6)
    H = rng.uniform(0, 0.2, (E, 4))
    d_prev = rng.uniform(0, 0.3, n)
    fixed = np.zeros(n, dtype=bool)
    fixed[fem.mesh.node_sets["left"]] = True
    vals = np.where(fixed, 1.0, 0.0)
    init = np.clip(d_prev + 0.1, 0, 1)
    dmg = P.PhaseFieldDamageSolver(fem, tol=1e-12, max_iter=2000, bounds_method="projected_cg")
    d = dmg.solve(T(H, dev), T(d_prev, dev), pf_dirichlet_mask=torch.as_tensor(fixed, device=dev),
                  pf_dirichlet_values=T(vals, dev), initial_guess=T(init, dev))
    od = O.PhaseFieldDamageSolver(ofem, tol=1e-12, max_iter=2000).solve(H, d_prev, initial_guess=init, fixed=fixed,
                                                                        vals=vals)
    assert rel(d, od) < RTOL
    assert np.all(d.cpu().numpy()[fixed] == 1.0)


def test_uniform_history_closed_form_on_gpu(dev):
    P, fem, ofem = setup(nx=64, ny=32, W=6.4, H=3.2)
    E, n = fem.mesh.n_elems, fem.mesh.n_nodes
    Hv = 1.3 * 3e-3 / 0.5
    dmg = P.PhaseFieldDamageSolver(fem, tol=1e-13, max_iter=500, bounds_method="projected_cg")
    d = dmg.solve(torch.full((E, 4), Hv, dtype=torch.float64, device=dev), torch.zeros(n, dtype=torch.float64,
                                                                                       device=dev))
    assert rel(d, np.full(n, 2 * Hv / (3e-3 / 0.5 + 2 * Hv))) < 1e-10


def test_explicit_staggered_steps_track_oracle(dev):
    """30 explicit steps with a projected damage solve each step (tight tolerance)."""
    P, fem, ofem = setup(nx=40, ny=16, W=10.0, H=4.0)
    n, E = fem.mesh.n_nodes, fem.mesh.n_elems
    f_ext = np.zeros((n, 2))
    top, bot = fem.mesh.node_sets["top"], fem.mesh.node_sets["bottom"]
    f_ext[top, 1] = 40.0
    f_ext[bot, 1] = -40.0
    dt = 0.5 * fem.dt_cfl
    dyn = P.ExplicitDynamics(fem, dt=dt)
    dmg = P.PhaseFieldDamageSolver(fem, tol=1e-12, max_iter=2000, bounds_method="projected_cg")
    st = P.StaggeredSolver(fem, dyn, dmg, f_ext=T(f_ext, dev), damage_every=1)
    d0 = np.zeros(n)
    d0[(fem.mesh.nodes[:, 1] == 2.0) & (fem.mesh.nodes[:, 0] <= 5.0)] = 1.0
    st.d = T(d0, dev)
    ost = O.ExplicitState(np.zeros((n, 2)), np.zeros((n, 2)), np.zeros((n, 2)), d0.copy(), np.zeros((E, 4)))
    odyn = O.ExplicitDynamics(ofem, dt=dt)
    odmg = O.PhaseFieldDamageSolver(ofem, tol=1e-12, max_iter=2000, bounds_method="projected_cg")
    for _ in range(30):
        st.step_full()
        O.step_full_explicit(odyn, odmg, ost, f_ext)
    print(f"30 steps: u {rel(st.u, ost.u):.1e} v {rel(st.v, ost.v):.1e} d {rel(st.d, ost.d):.1e} "
          f"H {rel(st.H_elem, ost.H):.1e}")
    assert rel(st.u, ost.u) < RTOL and rel(st.v, ost.v) < RTOL and rel(st.d, ost.d) < RTOL
    assert rel(st.H_elem, ost.H) < RTOL


def test_batched_plan_is_bitwise_equal_to_single_plans(dev):
    from phast_amd import q4 as P
    from phast_amd import batch as B
    mat = P.Material(E=32000.0, nu=0.2, rho=2.45e-9, Gc=3e-3, l0=0.5)
    meshes = [P.FEMMesh.rectangle(0.5 * nx, 0.25 * nx, nx, nx // 2) for nx in (13, 40, 7)]
    rng = np.random.default_rng(8)
    bat = B.BatchedExplicit([(m, mat) for m in meshes], dt_factor=0.5)
    states = []
    for m in meshes:
        u, v, a, d, H = state(m.n_nodes, m.n_elems, seed=int(rng.integers(1000)))
        states.append((u * 1e-2, v * 1e-3, a * 0, d * 0.2, H * 1e-4))
    bat.load_states([tuple(T(x, dev) for x in s) for s in states])
    bat.step(n_steps=3)
    for k, m in enumerate(meshes):
        fem = P.FEMOperators(m, mat)
        dyn = P.ExplicitDynamics(fem, dt=bat.dt[k])
        dmg = P.PhaseFieldDamageSolver(fem, tol=bat.tol, max_iter=bat.max_iter, bounds_method="projected_cg")
        st = P.StaggeredSolver(fem, dyn, dmg, damage_every=1)
        u, v, a, d, H = (T(x, dev) for x in states[k])
        st.u, st.v, st.a, st.d, st.H_elem = u, v, a, d, H
        for _ in range(3):
            st.step_full()
        bu, bv, ba, bd, bH = bat.instance_state(k)
        assert torch.equal(bu, st.u) and torch.equal(bd, st.d) and torch.equal(bH, st.H_elem)


@pytest.mark.parametrize("ps", [False, True])
def test_full_stress_degradation(dev, ps):
    """stress_degradation='full': isotropic stress, Amor psi+ and history (material.py:367-376)."""
    from phast_amd import q4 as P
    kw = dict(E=32000.0, nu=0.2, rho=2.45e-9, Gc=3e-3, l0=0.5, energy_split="amor", plane_stress=ps,
              stress_degradation="full")
    mesh = P.FEMMesh.rectangle(3.7, 2.3, 37, 23)
    fem = P.FEMOperators(mesh, P.Material(**kw))
    ofem = O.FEMOperators(O.Q4Mesh(*O.structured_q4_mesh(3.7, 2.3, 37, 23)), O.Material(**kw))
    u, v, a, d, H = state(mesh.n_nodes, mesh.n_elems, seed=3)
    assert rel(fem.internal_force(T(u, dev), T(d, dev)), ofem.internal_force(u, d)) < RTOL
    assert rel(fem.compute_psi_plus(T(u, dev)), ofem.compute_psi_plus(u)) < RTOL
    dt = 0.3 * fem.dt_cfl
    z = np.zeros((mesh.n_nodes, 2))
    u1, v1, a1, Hn, psi = P.ExplicitDynamics(fem, dt=dt).step_with_history(
        T(u, dev), T(v, dev), T(a, dev), T(d, dev), T(z, dev), None, None, T(H * 1e-4, dev))
    ou, ov, oa = O.ExplicitDynamics(ofem, dt=dt).step(u, v, a, d, z)
    assert rel(a1, oa) < RTOL and rel(psi, ofem.compute_psi_plus(ou)) < RTOL


def test_kelvin_voigt_damped_step(dev):
    """damping_ratio_max > 0: f_net = f_ext - f_int(u1) - beta f_int(v) (mechanics_solver.py:192-195)."""
    P, fem, ofem = setup(nx=31, ny=17, W=3.1, H=1.7)
    n, E = fem.mesh.n_nodes, fem.mesh.n_elems
    u, v, a, d, H = state(n, E, seed=9)
    f_ext = np.random.default_rng(10).uniform(-1e-2, 1e-2, (n, 2))
    dyn = P.ExplicitDynamics(fem, damping_ratio_max=0.15)
    odyn = O.ExplicitDynamics(ofem, dt=dyn.dt, damping_ratio_max=0.15)
    assert dyn.dt == pytest.approx(fem.dt_cfl * (np.sqrt(1 + 0.15 ** 2) - 0.15), rel=1e-14)
    u1, v1, a1 = dyn.step(T(u, dev), T(v, dev), T(a, dev), T(d, dev), T(f_ext, dev))
    ou, ov, oa = odyn.step(u, v, a, d, f_ext)
    assert rel(u1, ou) < RTOL and rel(v1, ov) < RTOL and rel(a1, oa) < RTOL
    und = O.ExplicitDynamics(ofem, dt=dyn.dt).step(u, v, a, d, f_ext)[2]
    assert rel(und, oa) > 1e-6  # the damping term is not negligible here



def test_damage_solve_gamma_correction(dev):
    """material.gamma_correction: Gc_eff = Gc / (1 + h / (2 l0)) per cell with h the shortest
    edge (damage_solver.py:1409-1433), run through the per-cell Gc branch."""
    from phast_amd import q4 as P
    P_, fem0, ofem = setup(nx=29, ny=13)
    mat_kw = dict(E=32000.0, nu=0.2, rho=2.45e-9, Gc=3e-3, l0=0.5)
    fem = P.FEMOperators(P.FEMMesh.rectangle(3.7, 2.3, 29, 13), P.Material(gamma_correction=True, **mat_kw),
                         tile=(128, 8))
    n, E = fem.mesh.n_nodes, fem.mesh.n_elems
    rng = np.random.default_rng(41)
    H = rng.uniform(0, 0.2, (E, 4))
    d_prev = np.clip(rng.uniform(-0.5, 0.5, n), 0, 1)
    h = min(3.7 / 29, 2.3 / 13)
    Gc_e = np.full(E, 3e-3 / (1.0 + h / (2.0 * 0.5)))
    d = P.PhaseFieldDamageSolver(fem, tol=1e-10, max_iter=3000).solve(T(H, dev), T(d_prev, dev))
    od = O.PhaseFieldDamageSolver(ofem, tol=1e-10, max_iter=3000, bounds_method="post_clamp", Gc_e=Gc_e).solve(
        H, d_prev)
    assert rel(d, od) < RTOL
