"""The five benchmark workloads at their own sizes on the GPU.

Each is checked against the oracle where the oracle finishes in seconds, and otherwise
through size-independent properties (KKT conditions, the Hill-Mandel identity, symmetry,
irreversibility).  Fields are compared in the max norm relative to the field's max norm.

The projected damage CG of the reference (damage_solver.py:3527-3775) branches on
rounding-level quantities: step limits of exactly zero at nodes on their bound, residual
signs at weakly active nodes, the stop test relative to ||b||, which the preseeded notch
dominates.  Two correct implementations therefore follow different iterate paths and agree
on the damage field only to the solver tolerance.  The mechanics, the history, the reaction
and the clamps do not iterate and agree to rounding.
"""
import time

import numpy as np
import pytest
import torch

from oracle import hex8 as OH
from oracle import q4 as OQ

pytestmark = pytest.mark.gpu
RTOL = 1e-10


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    b = b.detach().cpu().numpy() if isinstance(b, torch.Tensor) else np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def T(x, dev):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device=dev)


def _say(msg):
    print(msg, flush=True)


# ----------------------------------------------------------------------------------------
# single-edge-notched plate, 200 x 80 Q4, 1000 explicit steps
# ----------------------------------------------------------------------------------------

def _sent(dev, tol):
    from phast_amd import q4 as P
    from phast_amd import workloads as W
    p = W.sent_plate()
    mesh = P.FEMMesh(p.nodes, p.elements)
    fem = P.FEMOperators(mesh, P.Material(**p.material))
    ofem = OQ.FEMOperators(OQ.Q4Mesh(p.nodes, p.elements), OQ.Material(**p.material))
    dmg = P.PhaseFieldDamageSolver(fem, tol=tol, max_iter=5000, bounds_method="projected_cg")
    odmg = OQ.PhaseFieldDamageSolver(ofem, tol=tol, max_iter=5000, bounds_method="projected_cg")
    st = P.StaggeredSolver(fem, P.ExplicitDynamics(fem, dt=p.dt), dmg, f_ext=T(p.f_ext, dev), damage_every=1)
    return p, fem, ofem, dmg, odmg, st


def test_sent_plate_every_step_from_the_oracle_state(dev):
    """300 steps of the SENT run (the stress wave reaches the notch near step 105): at every
    step the GPU advances the oracle's state by one step.  u, v, a and H equal the oracle's to
    1e-10.  The damage equals it to the solve's tolerance (tol 1e-12, whose stop test is
    relative to ||b||, which the notch's preseeded history dominates by ~1e4 over the crack
    tip), and so does the top-edge reaction, which sees the small AT2 damage of the loaded
    edge through g(d)."""
    p, fem, ofem, dmg, odmg, st = _sent(dev, 1e-12)
    n = p.nodes.shape[0]
    d0 = odmg.solve(p.H0, np.zeros(n))
    assert rel(dmg.solve(T(p.H0, dev), T(np.zeros(n), dev)), d0) < 1e-12  # the preseeding solve
    ost = OQ.ExplicitState(np.zeros((n, 2)), np.zeros((n, 2)), np.zeros((n, 2)), d0, p.H0.copy())
    odyn = OQ.ExplicitDynamics(ofem, dt=p.dt)
    worst = dict.fromkeys(("u", "v", "a", "H", "R", "d"), 0.0)
    same_iters = mask_steps = 0
    t0 = time.time()
    steps = 300
    for k in range(steps):
        st.u, st.v, st.a, st.d, st.H_elem = (T(x, dev) for x in (ost.u, ost.v, ost.a, ost.d, ost.H))
        d_prev, H_prev = ost.d.copy(), ost.H.copy()
        st.step_full()
        OQ.step_full_explicit(odyn, odmg, ost, p.f_ext)
        R = fem.compute_reaction_force(st.u, st.d, p.top, 1)
        oR = ofem.compute_reaction_force(ost.u, ost.d, p.top, 1)
        e = {"u": rel(st.u, ost.u), "v": rel(st.v, ost.v), "a": rel(st.a, ost.a), "H": rel(st.H_elem, ost.H),
             "R": abs(R - oR) / abs(oR) if oR else abs(R), "d": rel(st.d, ost.d)}
        for key, val in e.items():
            worst[key] = max(worst[key], val)
        assert dmg.last_converged, (k, dmg.last_status)
        d = st.d.cpu().numpy()
        assert np.all(d >= d_prev) and np.all(d <= 1.0), k             # bound clamps
        assert np.all(st.H_elem.cpu().numpy() >= H_prev), k             # history max
        same_iters += dmg.last_iter == odmg.last_iter
        # final active sets, on the nodes whose residual is not at rounding level
        rho, b, _ = odmg.prepare(ost.H, d_prev)
        r = b - odmg.Ax(ost.d, rho)
        strict = np.abs(r) > 1e-9 * np.abs(r).max()
        mask_steps += bool(np.array_equal(dmg.last_active.cpu().numpy()[strict], odmg.last_active[strict]))
        if k % 50 == 0:
            _say(f"step {k}: " + " ".join(f"{key} {val:.1e}" for key, val in e.items())
                 + f" R {oR:.3f} its {dmg.last_iter}/{odmg.last_iter} ({time.time() - t0:.0f}s)")
    _say(f"worst over {steps} steps: " + " ".join(f"{k} {v:.2e}" for k, v in worst.items())
         + f"; equal CG iteration counts {same_iters}/{steps}; equal strict active sets {mask_steps}/{steps}")
    for key in ("u", "v", "a", "H"):
        assert worst[key] < RTOL, (key, worst[key])
    assert worst["d"] < 1e-7 and worst["R"] < 1e-7
    assert mask_steps >= 0.95 * steps


def test_sent_plate_1000_steps(dev):
    """The whole run at the reference's tolerance (1e-5) on the GPU: every damage solve
    converges, H never decreases, d never falls below its previous value; the first 100
    steps (before the wave reaches the notch) equal the oracle's free-running trajectory."""
    p, fem, ofem, dmg, odmg, st = _sent(dev, 1e-5)
    n = p.nodes.shape[0]
    st.d = dmg.solve(T(p.H0, dev), T(np.zeros(n), dev))
    st.H_elem = T(p.H0, dev)
    ost = OQ.ExplicitState(np.zeros((n, 2)), np.zeros((n, 2)), np.zeros((n, 2)), odmg.solve(p.H0, np.zeros(n)),
                           p.H0.copy())
    odyn = OQ.ExplicitDynamics(ofem, dt=p.dt)
    reactions, its = [], []
    for k in range(p.steps):
        d_prev, H_prev = st.d, st.H_elem
        st.step_full()
        assert dmg.last_converged, (k, dmg.last_status)
        assert bool((st.d >= d_prev).all()) and bool((st.H_elem >= H_prev).all()), k
        its.append(dmg.last_iter)
        reactions.append(fem.compute_reaction_force(st.u, st.d, p.top, 1))
        if k < 100:
            OQ.step_full_explicit(odyn, odmg, ost, p.f_ext)
            for a, b in ((st.u, ost.u), (st.v, ost.v), (st.d, ost.d), (st.H_elem, ost.H)):
                assert rel(a, b) < RTOL, k
            oR = ofem.compute_reaction_force(ost.u, ost.d, p.top, 1)
            assert abs(reactions[-1] - oR) <= RTOL * abs(oR), k
    R = np.array(reactions)
    _say(f"1000 steps: reaction {R.min():.2f}..{R.max():.2f} N, mean damage CG iterations {np.mean(its):.2f}, "
         f"d max {float(st.d.max()):.3f}")
    assert np.all(np.isfinite(R)) and float(st.d.max()) == 1.0


def test_sent_plate_persistent_graph_equals_the_per_kernel_loop(dev):
    """The SENT run through the persistent kernel with one instance, replayed from a captured
    graph of 20 steps, gives the per-kernel loop's bits for 200 steps."""
    from phast_amd import batch as B
    p, fem, ofem, dmg, odmg, st = _sent(dev, 1e-5)
    n = p.nodes.shape[0]
    d0 = dmg.solve(T(p.H0, dev), T(np.zeros(n), dev))
    st.d, st.H_elem = d0.clone(), T(p.H0, dev)
    for _ in range(200):
        st.step_full()
    from phast_amd import q4 as P
    bat = B.BatchedExplicit([(fem.mesh, fem.material)], dts=[p.dt], f_ext=[p.f_ext], tol=1e-5, max_iter=5000,
                            bounds_method="projected_cg")
    z = torch.zeros((n, 2), dtype=torch.float64, device=dev)
    bat.load_states([(z, z, z, d0, T(p.H0, dev))])
    g = bat.capture(20)
    g.replay(10)
    bat.check_persistent(g.status[-1])
    for a, b in zip(bat.instance_state(0), (st.u, st.v, st.a, st.d, st.H_elem)):
        assert torch.equal(a, b)
    assert isinstance(fem.material, P.Material)


# ----------------------------------------------------------------------------------------
# 64^3 pore cube
# ----------------------------------------------------------------------------------------

def _voxel(dev, phase, phases, h, load, Gc_field=None):
    from phast_amd import solid3d as S
    mats = S.build_element_arrays(phase, phases, Gc_field=Gc_field)
    n = phase.shape[::-1]
    op = S.PeriodicVoxelPhaseFieldOperator(S.VoxelGrid(n, h), mats, load, kappa=1e-6, ell=2e-3, device=dev)
    oop = OH.VoxelOperator(n, h, mats.stiffness_bank, mats.stiffness_index, mats.Gc, mats.lam, mats.mu,
                           mats.driver.astype(np.int64), kappa=1e-6, ell=2e-3, pinned_node=op.pinned_node)
    return S, op, oop


def _operators_match_oracle(dev, op, oop, strain, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1e-6, 1e-6, op.n_mech)
    x[-6:] = np.array([1.0, -0.3, -0.3, 0.1, -0.05, 0.02]) * strain
    d = rng.uniform(0, 1, op.n_red)
    H = rng.uniform(0, 1e-3, (op.n_elements, 8))
    g = oop.degradation(d)
    xt, dt, Ht = T(x, dev), T(d, dev), T(H, dev)
    assert rel(op.degradation(dt), g) < 1e-14
    assert rel(op.mech_action_d(xt, dt), oop.mech_action(x, g)) < RTOL
    assert rel(op.mech_diagonal(T(g, dev)), oop.mech_diagonal(g)) < RTOL
    assert rel(op.driving_force(xt), oop.driving_force(x, method="eigh")) < RTOL
    assert rel(op.damage_action(dt, Ht), oop.damage_action(d, H)) < RTOL
    assert rel(op.damage_rhs(Ht), oop.damage_rhs(H)) < RTOL
    assert rel(op.damage_diagonal(Ht), oop.damage_diagonal(H)) < RTOL


def _increments(S, op, targets, rtol=1e-10):
    from phast_amd import staggered3d as ST
    mg = S.VoxelMultigrid(op)
    solver = ST.StaggeredSolver(op, cfg=ST.SolverConfig(pcg_rtol_mech=rtol, damage_kkt_tol=1e-9), preconditioner=mg)
    st = solver.initial_state()
    out = []
    for t in targets:
        prev = st
        st, rec = solver.advance(st, t)
        lower, upper = st.active
        sysm = op.damage_system(st.H)
        kkt = S._kkt_diagnostics(sysm, st.d, prev.d, torch.ones_like(st.d), lower, upper,
                                 torch.zeros_like(lower), 1e-9)
        _say(f"target {t:.2e}: stagger {rec.stagger_iters}, mech PCG {rec.pcg_iters_mech}, damage PCG "
             f"{rec.pcg_iters_pf}, max d {rec.max_damage:.3e}, Hill-Mandel {rec.hill_mandel_rel:.1e}, "
             f"KKT {kkt.satisfies_kkt} ({rec.wall_s:.1f}s)")
        assert kkt.satisfies_kkt
        assert rec.hill_mandel_rel < 1e-8  # an identity at equilibrium: holds to the mechanics tolerance
        assert bool((st.d >= prev.d).all()) and bool((st.H >= prev.H).all())
        out.append((st, rec))
    return out


def test_pore_cube_64_operators_and_increments(dev):
    from phast_amd import workloads as W
    phase, phases, h = W.pore_cube(64, 12.0)
    from phast_amd import solid3d as S
    load = S.LoadCase("xx", 0, tuple(5e-4 * k for k in range(1, 41)), (1, 2, 3, 4, 5), ())
    S, op, oop = _voxel(dev, phase, phases, h, load)
    _operators_match_oracle(dev, op, oop, 5e-4, 1)
    res = _increments(S, op, (5e-4, 1e-3, 1.5e-3))
    st, rec = res[-1]
    # the centred pore and the load are symmetric under y <-> z: so are E and the damage
    E = st.x[-6:].cpu().numpy()
    assert abs(E[1] - E[2]) <= 1e-8 * abs(E[1])
    d = st.d.cpu().numpy().reshape(64, 64, 64)
    assert np.abs(d - d.transpose(1, 0, 2)).max() <= 1e-6 * d.max()
    assert d.max() > 1e-3


# ----------------------------------------------------------------------------------------
# 128^3 particle composite
# ----------------------------------------------------------------------------------------

def test_particle_composite_128_operators_and_first_increment(dev):
    from phast_amd import geometry as G
    from phast_amd import workloads as W
    phase, phases, h, centres, radii = W.particle_composite(128)
    assert len(radii) > 100
    frac = (phase != G.MATRIX).mean()
    assert 0.25 < frac <= 0.36, frac
    from phast_amd import solid3d as S
    load = S.LoadCase("xx", 0, tuple(2.5e-4 * k for k in range(1, 61)), (1, 2, 3, 4, 5), ())
    S, op, oop = _voxel(dev, phase, phases, h, load)
    t0 = time.time()
    _operators_match_oracle(dev, op, oop, 2.5e-4, 2)
    _say(f"128^3 operators against the oracle: {time.time() - t0:.1f}s")
    del oop
    (st, rec), = _increments(S, op, (2.5e-4,))
    # the weak interface shell (Gc x 0.1) damages first
    shell = (phase == G.INTERFACE).astype(np.int8)
    touch = np.zeros_like(shell)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):  # node (i,j,k) touches voxels (i-1+dx, j-1+dy, k-1+dz)
                touch |= np.roll(shell, shift=(1 - dz, 1 - dy, 1 - dx), axis=(0, 1, 2))
    d = st.d.cpu().numpy()
    assert touch.reshape(-1)[int(np.argmax(d))] == 1


# ----------------------------------------------------------------------------------------
# 64 plates of uneven sizes, batched
# ----------------------------------------------------------------------------------------

def test_specimen_batch_every_specimen_against_the_oracle(dev):
    """The 64-specimen batch (3.03 M nodes) advanced one step by the persistent kernel;
    every specimen equals the oracle's step (Gc field, notch pinned at d = 1)."""
    from phast_amd import batch as B
    from phast_amd import q4 as P
    from phast_amd import workloads as W
    specs = W.specimen_batch()
    mat = P.Material(**W.Q4_MATERIAL)
    items = [(P.FEMMesh.rectangle(s.h * s.nx, s.h * s.ny, s.nx, s.ny), mat) for s in specs]
    bat = B.BatchedExplicit(items, dt_factor=0.4, Gc_fields=[s.Gc_field for s in specs],
                            f_ext=[s.f_ext for s in specs], fixed=[(s.notch, 1.0) for s in specs],
                            bounds_method="projected_cg", tol=1e-12, max_iter=5000, tile=(256, 16))
    bat.step_persistent(1)
    bat.check_persistent()
    t0 = time.time()
    worst = {"u": 0.0, "v": 0.0, "a": 0.0, "H": 0.0, "d": 0.0}
    for k, (s, (mesh, _)) in enumerate(zip(specs, items)):
        nodes, quads = OQ.structured_q4_mesh(s.h * s.nx, s.h * s.ny, s.nx, s.ny)
        ofem = OQ.FEMOperators(OQ.Q4Mesh(nodes, quads), OQ.Material(**W.Q4_MATERIAL))
        odyn = OQ.ExplicitDynamics(ofem, dt=bat.dt[k])
        odmg = OQ.PhaseFieldDamageSolver(ofem, tol=1e-12, max_iter=5000, bounds_method="projected_cg",
                                         Gc_e=s.Gc_field)
        n = nodes.shape[0]
        fixed = np.zeros(n, dtype=bool)
        fixed[s.notch] = True
        vals = np.where(fixed, 1.0, 0.0)
        z = np.zeros((n, 2))
        u, v, a = odyn.step(z, z, z, vals, s.f_ext)
        H = np.maximum(np.zeros((quads.shape[0], 4)), ofem.compute_psi_plus(u))
        d = odmg.solve(H, vals, fixed=fixed, vals=vals)
        got = bat.instance_state(k)
        for key, g, o in zip(("u", "v", "a", "d", "H"), got, (u, v, a, d, H)):
            e = rel(g, o)
            worst[key] = max(worst[key], e)
        assert np.array_equal(got[3].cpu().numpy()[fixed], vals[fixed])
    _say(f"64 specimens, oracle {time.time() - t0:.0f}s, worst " + " ".join(f"{k} {v:.1e}" for k, v in worst.items()))
    for key in ("u", "v", "a", "H"):
        assert worst[key] < RTOL, key
    assert worst["d"] < 1e-8


# ----------------------------------------------------------------------------------------
# reaction-curve gradient on a porous cube with a per-voxel Gc field
# ----------------------------------------------------------------------------------------

def test_porous_cube_gradient_at_32_random_voxels(dev):
    """48^3 cube, 20 seeded pores (radii U[2, 4.5] voxels, the 96^3 workload at half scale),
    Gc = 5e-4 (1 + 0.1 N(0,1)); the adjoint gradient of the squared reaction misfit over 6
    load steps, with load-step checkpointing and recomputation, against central differences
    with the workload's step h = 1e-6 N/mm (2e-3 of Gc) on 32 seeded random voxels:
    |adjoint - FD| <= 1e-6 max(|FD|, 1e-2 max|grad|).  Most random voxels carry gradients far
    below the largest; there a per-entry relative error would measure the difference
    quotient's rounding, not the adjoint.  (A step of 1e-6 Gc_e is rounding-dominated at this
    loss, ~7e-9: it leaves 7e-5 relative error on an entry whose differences at 1e-3 and 1e-4
    of Gc_e agree with the adjoint to 3e-8.)"""
    from phast_amd import adjoint3d as A
    from phast_amd import workloads as W
    from phast_amd import solid3d as S
    phase, phases, h, Gc, Gc_t = W.porous_gc_cube(48, 20, (2.0, 4.5), seed=48, gc_seed=5)
    load = S.LoadCase("xx", 0, tuple(5e-4 * k for k in range(1, 7)), (1, 2, 3, 4, 5), ())
    S, op, _ = _voxel(dev, phase, phases, h, load)
    curve = A.ReactionCurve(op, preconditioner=S.VoxelMultigrid(op), checkpoint_every=3)
    R_hat, _ = curve.forward(T(Gc_t, dev))
    R_hat = R_hat.detach().clone()
    Gt = T(Gc, dev)
    R, tape = curve.forward(Gt)
    grad = curve.backward(tape, 2.0 * (R - R_hat))
    d_max = float(tape["checkpoints"][6].d.max())
    assert d_max > 0.05, d_max

    def loss(g):
        return float(((curve.forward(g)[0] - R_hat) ** 2).sum())

    rng = np.random.default_rng(32)
    gmax = float(grad.abs().max())
    picks = rng.choice(op.n_elements, 32, replace=False)
    t0 = time.time()
    errs = []
    for e in picks:
        step = 1e-6
        gp, gm = Gt.clone(), Gt.clone()
        gp[e] += step
        gm[e] -= step
        fd = (loss(gp) - loss(gm)) / (2 * step)
        errs.append(abs(float(grad[e]) - fd))
        assert abs(float(grad[e]) - fd) <= 1e-6 * max(abs(fd), 1e-2 * gmax), (int(e), float(grad[e]), fd, gmax)
    _say(f"32 voxels: max |adjoint - FD| / max|grad| = {max(errs) / gmax:.1e}, max|grad| {gmax:.3e} "
         f"({time.time() - t0:.0f}s)")
