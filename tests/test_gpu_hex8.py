"""Parity of the periodic hex8 HIP path with the oracle, and the reference's recorded
pore-in-cube elastic rows (GPU).  Fields within 1e-10 relative (max-norm); masks exact."""
import json
import math
from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import hex8 as O

pytestmark = pytest.mark.gpu
RTOL = 1e-10
GOLDEN = Path(__file__).with_name("golden") / "pore_in_cube_elastic.json"


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    b = b.detach().cpu().numpy() if isinstance(b, torch.Tensor) else np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


T = lambda x, dev: torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device=dev)


def build(shape, h=(1e-3, 1.1e-3, 0.9e-3), seed=0, pores=True, tiles=(32, 8, 8), energy_split="spectral"):
    from phast_amd import solid3d as S
    from phast_amd import geometry as G
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    phase = np.full((nz, ny, nx), G.MATRIX, dtype=np.uint8)
    phase[rng.uniform(0, 1, phase.shape) < 0.3] = G.INCLUSION
    if pores:
        phase[rng.uniform(0, 1, phase.shape) < 0.1] = G.PORE
    phases = {"mi_matrix": S.PhaseMaterial("mi_matrix", 3000.0, 0.35, 5e-4),
              "fibre": S.PhaseMaterial("fibre", 70000.0, 0.3, 5e-3),
              "pore": S.PhaseMaterial("pore", 3e-3, 0.35, 5e-4)}
    mats = S.build_element_arrays(phase, phases, energy_split_isotropic=energy_split)
    load = S.LoadCase("uni", 0, (1e-3,), (1, 2, 3, 4, 5), ())
    op = S.PeriodicVoxelPhaseFieldOperator(S.VoxelGrid(shape, h), mats, load, kappa=1e-6, ell=2e-3, tiles=tiles)
    oop = O.VoxelOperator(shape, h, mats.stiffness_bank, mats.stiffness_index, mats.Gc, mats.lam, mats.mu,
                          mats.driver.astype(np.int64), kappa=1e-6, ell=2e-3, pinned_node=op.pinned_node,
                          controlled=0)
    return S, op, oop


@pytest.mark.parametrize("shape", [(6, 5, 7), (16, 16, 16), (17, 9, 10)])
@pytest.mark.parametrize("tiles", [(16, 8, 8), (8, 8, 4), (16, 16, 3), (32, 8, 16), (16, 16, 1), (32, 8, 2),
                                   (16, 16, 100), (16, 14, 5), (32, 14, 4), (32, 6, 3), (64, 6, 7)])
def test_mech_action(dev, shape, tiles):
    S, op, oop = build(shape, tiles=tiles)
    rng = np.random.default_rng(1)
    x = rng.uniform(-1e-6, 1e-6, op.n_mech)
    x[-6:] = rng.uniform(-1e-3, 1e-3, 6)
    d = rng.uniform(0, 1, op.n_red)
    g = oop.degradation(d)
    assert rel(op.degradation(T(d, dev)), g) < 1e-14
    ref = oop.mech_action(x, g)
    assert rel(op.mech_action(T(x, dev), T(g, dev)), ref) < RTOL
    assert rel(op.mech_action_d(T(x, dev), T(d, dev)), ref) < RTOL
    assert rel(op.mech_action(T(x, dev)), oop.mech_action(x)) < RTOL


@pytest.mark.parametrize("tiles", [(16, 16, 2), (32, 8, 3), (16, 16, 7), (32, 8, 1), (16, 14, 6), (32, 14, 5),
                                   (32, 6, 4), (64, 6, 3)])
def test_voxel_parallel_apply_matches_marching_apply(dev, tiles):
    """Same per-voxel arithmetic and node summation order as the marching kernel: the
    nodal rows agree bitwise; the macro rows and energy (block partials in another order)
    to rounding."""
    S, op, oop = build((20, 19, 14), seed=5)
    rng = np.random.default_rng(6)
    x = T(rng.uniform(-1e-6, 1e-6, op.n_mech), dev)
    x[-6:] = T(rng.uniform(-1e-3, 1e-3, 6), dev)
    d = T(rng.uniform(0, 1, op.n_red), dev)
    g = op.degradation(d)
    op.set_tiles(16, 8, 8)
    ref, ref_d = op.mech_action(x, g), op.mech_action_d(x, d)
    ref_s = op._stress_sums(x, g)
    op.set_tiles(*tiles)
    for _ in range(2):  # repeated launches reuse the partial and seam buffers
        y, y_d = op.mech_action(x, g), op.mech_action_d(x, d)
        n = 3 * op.n_red
        assert torch.equal(y[:n], ref[:n]) and torch.equal(y_d[:n], ref_d[:n])
        assert rel(y[n:], ref[n:]) < 1e-13 and rel(y_d[n:], ref_d[n:]) < 1e-13
        assert rel(op._stress_sums(x, g), ref_s) < 1e-13
    xo, go = x.cpu().numpy(), g.cpu().numpy()
    assert rel(op.mech_action(x, g), oop.mech_action(xo, go)) < RTOL


@pytest.mark.parametrize("n_moduli", [256, 257, 0])
def test_isotropic_moduli_per_voxel(dev, n_moduli):
    """Isotropic cells with n_moduli distinct (E, nu) pairs drawn per voxel (0: every voxel
    its own): up to 256 pairs the voxel-parallel apply reads a 1-byte phase and a table,
    beyond that the per-voxel lam/mu arrays.  Either way the nodal rows equal the marching
    kernel's bitwise (same doubles, same arithmetic) and the oracle's to RTOL."""
    import dataclasses
    from phast_amd import solid3d as S
    from phast_amd import geometry as G
    shape, h = (20, 19, 14), (1e-3, 1.1e-3, 0.9e-3)
    n = shape[0] * shape[1] * shape[2]
    rng = np.random.default_rng(8)
    k = n if n_moduli == 0 else n_moduli
    E, nu = rng.uniform(2e3, 8e4, k), rng.uniform(0.2, 0.4, k)
    idx = np.arange(n) if n_moduli == 0 else np.concatenate([np.arange(k), rng.integers(0, k, n - k)])
    bank = np.stack([S.isotropic_stiffness_3d(a, b) for a, b in zip(E, nu)])
    base = S.build_element_arrays(np.full((shape[2], shape[1], shape[0]), G.MATRIX, np.uint8),
                                  {"mi_matrix": S.PhaseMaterial("mi_matrix", 3000.0, 0.35, 5e-4)})
    mats = dataclasses.replace(base, stiffness_bank=bank, stiffness_index=idx.astype(np.int64),
                               lam=bank[idx, 0, 1].copy(), mu=bank[idx, 3, 3].copy())
    assert len({(a, b) for a, b in zip(mats.lam, mats.mu)}) == k
    load = S.LoadCase("uni", 0, (1e-3,), (1, 2, 3, 4, 5), ())
    op = S.PeriodicVoxelPhaseFieldOperator(S.VoxelGrid(shape, h), mats, load, kappa=1e-6, ell=2e-3, tiles=(16, 8, 8))
    oop = O.VoxelOperator(shape, h, mats.stiffness_bank, mats.stiffness_index, mats.Gc, mats.lam, mats.mu,
                          mats.driver.astype(np.int64), kappa=1e-6, ell=2e-3, pinned_node=op.pinned_node,
                          controlled=0)
    x = T(rng.uniform(-1e-6, 1e-6, op.n_mech), dev)
    x[-6:] = T(rng.uniform(-1e-3, 1e-3, 6), dev)
    d = T(rng.uniform(0, 1, op.n_red), dev)
    g = op.degradation(d)
    ref, ref_d, ref_1 = op.mech_action(x, g), op.mech_action_d(x, d), op.mech_action(x)
    m = 3 * n
    for tiles in [(16, 14, 8), (16, 14, 3), (32, 8, 5)]:
        op.set_tiles(*tiles)
        y, y_d, y_1 = op.mech_action(x, g), op.mech_action_d(x, d), op.mech_action(x)
        assert torch.equal(y[:m], ref[:m]) and torch.equal(y_d[:m], ref_d[:m]) and torch.equal(y_1[:m], ref_1[:m])
        assert rel(y[m:], ref[m:]) < 1e-13
    assert rel(op.mech_action(x, g), oop.mech_action(x.cpu().numpy(), g.cpu().numpy())) < RTOL


@pytest.mark.parametrize("tiles", [(16, 8, 8), (32, 14, 3), (16, 14, 6), (64, 6, 5)])
def test_general_stiffness_bank(dev, tiles):
    """Bank entries that are not isotropic take the kernels' general 6x6 stiffness path
    (per-voxel bank index instead of lam/mu): apply, fused apply, stress sums, diagonal."""
    import dataclasses
    from phast_amd import solid3d as S
    from phast_amd import geometry as G
    shape, h = (13, 11, 9), (1e-3, 1.1e-3, 0.9e-3)
    rng = np.random.default_rng(11)
    phase = np.full(shape[::-1], G.MATRIX, dtype=np.uint8)
    phase[rng.uniform(0, 1, phase.shape) < 0.4] = G.INCLUSION
    phases = {"mi_matrix": S.PhaseMaterial("mi_matrix", 3000.0, 0.35, 5e-4),
              "fibre": S.PhaseMaterial("fibre", 70000.0, 0.3, 5e-3),
              "pore": S.PhaseMaterial("pore", 3e-3, 0.35, 5e-4)}
    mats = S.build_element_arrays(phase, phases, energy_split_isotropic="spectral")
    bank = np.array(mats.stiffness_bank, dtype=np.float64)
    for k in range(bank.shape[0]):  # symmetric positive perturbation couples every component
        a = rng.uniform(-1, 1, (6, 6))
        bank[k] += 0.05 * np.abs(bank[k]).max() * (a @ a.T) / 6.0
    mats = dataclasses.replace(mats, stiffness_bank=bank)
    load = S.LoadCase("uni", 0, (1e-3,), (1, 2, 3, 4, 5), ())
    op = S.PeriodicVoxelPhaseFieldOperator(S.VoxelGrid(shape, h), mats, load, kappa=1e-6, ell=2e-3, tiles=tiles)
    oop = O.VoxelOperator(shape, h, bank, mats.stiffness_index, mats.Gc, mats.lam, mats.mu,
                          mats.driver.astype(np.int64), kappa=1e-6, ell=2e-3, pinned_node=op.pinned_node,
                          controlled=0)
    x = rng.uniform(-1e-6, 1e-6, op.n_mech)
    x[-6:] = rng.uniform(-1e-3, 1e-3, 6)
    d = rng.uniform(0, 1, op.n_red)
    g = oop.degradation(d)
    ref = oop.mech_action(x, g)
    assert rel(op.mech_action(T(x, dev), T(g, dev)), ref) < RTOL
    assert rel(op.mech_action_d(T(x, dev), T(d, dev)), ref) < RTOL
    assert rel(op.macro_stress(T(x, dev), T(g, dev)), oop.macro_stress(x, g)) < RTOL
    assert rel(op.elastic_energy(T(x, dev), T(g, dev)), oop.elastic_energy(x, g)) < RTOL
    assert rel(op.mech_diagonal(T(g, dev)), oop.mech_diagonal(g)) < RTOL


def test_diag_stress_energy_driving(dev):
    S, op, oop = build((9, 7, 8))
    rng = np.random.default_rng(2)
    x = rng.uniform(-1e-6, 1e-6, op.n_mech)
    x[-6:] = rng.uniform(-1e-3, 1e-3, 6)
    g = oop.degradation(rng.uniform(0, 1, op.n_red))
    assert rel(op.mech_diagonal(T(g, dev)), oop.mech_diagonal(g)) < RTOL
    assert rel(op.mech_diagonal(), oop.mech_diagonal()) < RTOL
    assert rel(op.macro_stress(T(x, dev), T(g, dev)), oop.macro_stress(x, g)) < RTOL
    assert rel(op.elastic_energy(T(x, dev), T(g, dev)), oop.elastic_energy(x, g)) < RTOL
    assert rel(op.driving_force(T(x, dev)), oop.driving_force(x)) < RTOL
    assert rel(op.driving_force(T(x, dev)), oop.driving_force(x, method="eigh")) < RTOL


def test_full_driver(dev):
    S, op, oop = build((6, 6, 6), energy_split="full")
    rng = np.random.default_rng(3)
    x = rng.uniform(-1e-6, 1e-6, op.n_mech)
    assert rel(op.driving_force(T(x, dev)), oop.driving_force(x)) < RTOL


@pytest.mark.parametrize("tiles", [(32, 8, 8), (32, 6, 3), (64, 6, 2)])
def test_damage_operator(dev, tiles):
    S, op, oop = build((10, 6, 8), tiles=tiles)
    rng = np.random.default_rng(4)
    H = rng.uniform(0, 1e-3, (op.n_elements, 8))
    x = rng.uniform(0, 1, op.n_red)
    assert rel(op.damage_action(T(x, dev), T(H, dev)), oop.damage_action(x, H)) < RTOL
    assert rel(op.damage_action(T(x, dev)), oop.damage_action(x)) < RTOL
    assert rel(op.damage_rhs(T(H, dev)), oop.damage_rhs(H)) < RTOL
    assert rel(op.damage_diagonal(T(H, dev)), oop.damage_diagonal(H)) < RTOL
    sysm = op.damage_system(T(H, dev))
    assert rel(sysm.apply(T(x, dev)), oop.damage_scale * oop.damage_action(x, H)) < RTOL


def test_mechanics_solve_constrained(dev):
    S, op, oop = build((8, 8, 8))
    rng = np.random.default_rng(5)
    g = oop.degradation(rng.uniform(0, 0.6, op.n_red))
    fixed = op.mech_fixed_mask()
    vals = op.mech_fixed_values(1e-3)
    diag = op.mech_diagonal(T(g, dev))
    x, info = S.solve_constrained(S.mech_action_of(op, g=T(g, dev)), diag=diag, fixed_mask=fixed, fixed_values=vals,
                                  rtol=1e-12, maxiter=5000)
    ox, oinfo = O.solve_constrained(lambda v: oop.mech_action(v, g), oop.mech_diagonal(g), fixed.cpu().numpy(),
                                    vals.cpu().numpy(), rtol=1e-12, maxiter=5000)
    assert info.converged and oinfo[1]
    assert abs(info.iterations - oinfo[0]) <= 2
    assert rel(x, ox) < RTOL
    # the generic (python-callable) form agrees with the device-resident loop
    xg, ig = S.solve_constrained(lambda v: op.mech_action(v, T(g, dev)), diag=diag, fixed_mask=fixed,
                                 fixed_values=vals, rtol=1e-12, maxiter=5000)
    assert rel(xg, ox) < RTOL


def test_box_kkt(dev):
    S, op, oop = build((8, 6, 7))
    rng = np.random.default_rng(6)
    H = rng.uniform(0, 4e-3, (op.n_elements, 8)) * (rng.uniform(0, 1, (op.n_elements, 1)) > 0.5)
    lower = np.clip(rng.uniform(-0.5, 0.6, op.n_red), 0, 1)
    sysm = op.damage_system(T(H, dev))
    res = S.solve_h8_scalar_box_kkt(sysm, lower_bound=T(lower, dev), upper_bound=1.0, initial=T(lower, dev),
                                    linear_tolerance=1e-13, kkt_tolerance=1e-9)
    s = oop.damage_scale
    od, ola, oua, its, total, trace = O.box_kkt(lambda v: s * oop.damage_action(v, H), s * oop.damage_rhs(H),
                                                s * oop.damage_diagonal(H), lower, np.ones(op.n_red), initial=lower,
                                                lin_tol=1e-13, kkt_tol=1e-9)
    assert res.kkt.satisfies_kkt
    assert rel(res.solution, od) < 1e-10
    assert np.array_equal(res.lower_active.cpu().numpy(), ola)
    assert np.array_equal(res.upper_active.cpu().numpy(), oua)
    assert res.active_set_iterations == its


def _kkt_stats_numpy(grad, sol, lo, up, la, ua, pr, tol):
    """box_kkt_h8.py:459-558's audit quantities, restated with numpy (NaN-propagating max)."""
    free = ~(la | ua | pr)
    proj = np.where(free, grad, 0.0)
    proj = np.where(la, np.minimum(grad, 0.0), proj)
    proj = np.where(ua, np.maximum(grad, 0.0), proj)
    lm = np.where(la, np.maximum(grad, 0.0), 0.0)
    um = np.where(ua, np.maximum(-grad, 0.0), 0.0)
    comp = float(np.maximum(np.max(np.abs(lm * (sol - lo))), np.max(np.abs(um * (up - sol)))))
    return (np.max(np.abs(np.where(free, grad, 0.0))), np.max(np.abs(proj)), comp,
            int((la & (grad < -tol)).sum()), int((ua & (-grad < -tol)).sum()))


def test_kkt_audit_one_pass(dev):
    """The one-pass device audit (phast_hex8_kkt_stats) against the numpy restatement on
    random states: every mask combination, sign violations, bound violations, and a NaN
    gradient entry, which must fail the audit as torch's max would."""
    S, op, oop = build((8, 6, 7))
    rng = np.random.default_rng(21)
    n = op.n_red
    H = T(rng.uniform(0, 4e-3, (op.n_elements, 8)), dev)
    sysm = op.damage_system(H)
    for trial in range(4):
        sol = rng.uniform(-0.1, 1.1, n)
        lo, up = np.clip(rng.uniform(-0.2, 0.5, n), 0, 1), np.ones(n)
        la, ua = rng.uniform(0, 1, n) < 0.3, rng.uniform(0, 1, n) < 0.2
        ua &= ~la
        pr = rng.uniform(0, 1, n) < 0.05
        grad = rng.normal(0, 1e-8 if trial % 2 else 1e-3, n)
        if trial == 3:
            grad[17] = np.nan
        tt = lambda a: T(a, dev)  # noqa: E731
        tb = lambda a: torch.as_tensor(a, device=dev)  # noqa: E731
        d = S._kkt_diagnostics(sysm, tt(sol), tt(lo), tt(up), tb(la), tb(ua), tb(pr), 1e-9, grad=tt(grad))
        fg, pm, cm, lsv, usv = _kkt_stats_numpy(grad, sol, lo, up, la, ua, pr, 1e-9)
        same = lambda a, b: a == b or (math.isnan(a) and math.isnan(b))  # noqa: E731
        assert same(d.free_stationarity_maximum, fg) and same(d.projected_gradient_maximum, pm)
        assert same(d.complementarity_maximum, cm)
        if trial == 3:  # the NaN reaches at least one maximum, and the audit fails
            assert math.isnan(fg) or math.isnan(pm) or math.isnan(cm)
        assert (d.lower_sign_violation_dofs, d.upper_sign_violation_dofs) == (lsv, usv)
        assert d.satisfies_kkt is False  # random states violate the bounds
    x = T(np.full(n, 0.5), dev)  # an interior point with zero gradient satisfies the audit
    z = torch.zeros(n, dtype=torch.bool, device=dev)
    d = S._kkt_diagnostics(sysm, x, T(np.zeros(n), dev), T(np.ones(n), dev), z, z, z, 1e-9,
                           grad=torch.zeros(n, dtype=torch.float64, device=dev))
    assert d.satisfies_kkt and d.free_stationarity_maximum == 0.0


def test_pore_in_cube_recorded_elastic_rows(dev):
    """The reference's stored history (48^3 pore cell) at E_xx = 2e-5, 4e-5, 6e-5."""
    from phast_amd import solid3d as S
    from phast_amd import geometry as G
    gold = json.loads(GOLDEN.read_text())
    c = gold["cell"]
    shape = tuple(c["shape"])
    h = (c["voxel_mm"],) * 3
    phase = G.sphere_pore(shape, h, c["pore_diameter_fraction"])
    phases = {"mi_matrix": S.PhaseMaterial("mi_matrix", c["E"], c["nu"], 25e-3),
              "pore": S.PhaseMaterial("pore", c["E"] * c["pore_stiffness_ratio"], c["nu"], 25e-3)}
    mats = S.build_element_arrays(phase, phases)
    load = S.LoadCase("xx", 0, (2e-5,), (1, 2, 3, 4, 5), ())
    op = S.PeriodicVoxelPhaseFieldOperator(S.VoxelGrid(shape, h), mats, load, kappa=1e-6, ell=1.5e-3)
    diag = op.mech_diagonal()
    fixed = op.mech_fixed_mask()
    x = None
    for row in gold["rows"]:
        x, info = op.solve_mechanics(fixed_mask=fixed, fixed_values=op.mech_fixed_values(row["macro_strain_xx"]),
                                     diag=diag, x0=x, rtol=c["pcg_rtol_mech"], maxiter=20000)
        assert info.converged
        sig = float(op.macro_stress(x)[0])
        assert abs(sig - row["macro_stress_xx"]) / row["macro_stress_xx"] < 2e-7, (sig, row)


def test_staggered_increments_track_oracle(dev):
    """Three AT2 increments of the voxel staggered solver on a small damaged cell."""
    from phast_amd import staggered3d as ST
    S, op, oop = build((8, 8, 8), h=(1e-3,) * 3, seed=11, pores=False)
    cfg = ST.SolverConfig(pcg_rtol_mech=1e-12, damage_kkt_tol=1e-9)
    solver = ST.StaggeredSolver(op, cfg=cfg)
    st = solver.initial_state()
    ost = {"x": np.zeros(op.n_mech), "d": np.zeros(op.n_red), "H": np.zeros((op.n_elements, 8))}
    for target in (4e-3, 8e-3, 1.2e-2):
        st, rec = solver.advance(st, target)
        ost, (k, sigma, _) = O.staggered_increment(oop, ost, target, rtol=1e-12, damage_kkt_tol=1e-9)
        assert rec.stagger_iters == k
        assert rel(st.d, ost["d"]) < RTOL and rel(st.x, ost["x"]) < RTOL
        assert rel(st.macro_stress, sigma) < RTOL and rel(st.H, ost["H"]) < RTOL
    assert float(st.d.max()) > 0.05  # damage actually evolved
    snap = solver.snapshot(st)
    assert snap.time_value == 1.2e-2 and snap.step == st.step
    np.testing.assert_allclose(snap.macro_stress, st.macro_stress.cpu().numpy(), rtol=1e-12)
    assert rel(snap.history_max, st.H.amax(dim=1)) == 0.0


@pytest.mark.parametrize("opts", [dict(history="current", irreversibility="history_only"),
                                  dict(history="current"), dict(anderson_depth=3)])
def test_staggered_options_track_oracle(dev, opts):
    """The stagger-loop options of solid3d/model.py:SolverConfig on the small damaged cell:
    history 'current', irreversibility 'history_only' and Anderson mixing (depth 3), loaded
    and then unloaded, against the oracle's increment with the same options; the fracture
    energy the Anderson safeguard uses against the oracle's."""
    from phast_amd import staggered3d as ST
    S, op, oop = build((8, 8, 8), h=(1e-3,) * 3, seed=11, pores=False)
    cfg = ST.SolverConfig(pcg_rtol_mech=1e-12, damage_kkt_tol=1e-9, **opts)
    solver = ST.StaggeredSolver(op, cfg=cfg)
    seen = []
    solver.iterate_observer, solver.observe_steps = seen.append, frozenset({2})
    st = solver.initial_state()
    ost = {"x": np.zeros(op.n_mech), "d": np.zeros(op.n_red), "H": np.zeros((op.n_elements, 8))}
    for target in (4e-3, 8e-3, 5e-3):
        st, rec = solver.advance(st, target)
        ost, (k, sigma, acc) = O.staggered_increment(oop, ost, target, rtol=1e-12, damage_kkt_tol=1e-9, **opts)
        print(target, rec.stagger_iters, k, rec.anderson_accepted, rec.anderson_rejected, acc)
        assert rec.stagger_iters == k
        if "anderson_depth" in opts:
            assert (rec.anderson_accepted, rec.anderson_rejected) == acc
        assert rel(st.d, ost["d"]) < RTOL and rel(st.x, ost["x"]) < RTOL
        assert rel(st.macro_stress, sigma) < RTOL and rel(st.H, ost["H"]) < RTOL
    assert [r["iteration"] for r in seen] == list(range(1, len(seen) + 1)) and len(seen) >= 2
    assert all(r["step"] == 2 and r["x"] is not None for r in seen)
    assert float(st.d.max()) > 0.05
    d = st.d
    assert abs(solver.fracture_energy(d) - O.fracture_energy_at2(oop, d.cpu().numpy())) \
        <= 1e-12 * solver.fracture_energy(d)
    if opts.get("irreversibility") == "history_only":
        assert solver.last_bound_violations is not None


@pytest.mark.parametrize("shape", [(16, 16, 16), (24, 16, 32)])
def test_multigrid_pcg_matches_jacobi_solution(dev, shape):
    """MG-preconditioned PCG converges to the Jacobi-PCG / oracle solution in far fewer iterations."""
    S, op, oop = build(shape, seed=9)
    rng = np.random.default_rng(10)
    d = T(rng.uniform(0, 0.8, op.n_red), dev)
    g = op.degradation(d)
    fixed, vals, diag = op.mech_fixed_mask(), op.mech_fixed_values(1e-3), op.mech_diagonal(g)
    xj, ij = op.solve_mechanics(d=d, fixed_mask=fixed, fixed_values=vals, diag=diag, rtol=1e-12, maxiter=20000)
    mg = S.VoxelMultigrid(op, min_coarse=4)
    assert len(mg.levels) >= 2
    xm, im = mg.solve(d=d, fixed_mask=fixed, fixed_values=vals, diag=diag, rtol=1e-12, maxiter=2000)
    assert ij.converged and im.converged
    print(f"jacobi-PCG {ij.iterations} iterations, MG-PCG {im.iterations} iterations, levels {mg.levels}, "
          f"lambda_max {mg.lambda_max}")
    assert im.iterations * 3 < ij.iterations, (im.iterations, ij.iterations)
    assert rel(xm, xj) < 1e-9
    # the V-cycle is symmetric: u.M v == v.M u
    u, v = (T(np.where(fixed.cpu().numpy(), 0.0, rng.normal(size=op.n_mech)), dev) for _ in range(2))
    a, b = float(torch.dot(u, mg.apply(v))), float(torch.dot(v, mg.apply(u)))
    assert abs(a - b) <= 1e-10 * max(abs(a), abs(b))
    assert float(torch.dot(u, mg.apply(u))) > 0


@pytest.mark.parametrize("shape", [(6, 5, 7), (16, 9, 10)])
def test_field_snapshot(dev, shape):
    """field_snapshot (solid3d/operator.py:1561-1652): element means against the oracle, the
    nodal fields through the periodic map, and the total displacement E (X - origin) + P u~."""
    S, op, oop = build(shape)
    rng = np.random.default_rng(21)
    x = rng.uniform(-1e-6, 1e-6, op.n_mech)
    x[-6:] = rng.uniform(-1e-3, 1e-3, 6)
    d = rng.uniform(0, 1, op.n_red)
    H = rng.uniform(0, 1, (op.n_elements, 8))
    snap = op.field_snapshot(3, 0.5, T(x, dev), T(d, dev), T(H, dev))
    g = oop.degradation(d)
    strain, stress, psi = O.field_snapshot_elements(oop, x, g)
    assert rel(snap.strain, strain) < 1e-12 and rel(snap.stress, stress) < 1e-12 and rel(snap.psi_elastic, psi) < 1e-12
    assert rel(snap.psi_driving, oop.driving_force(x).mean(axis=1)) < 1e-12
    assert rel(snap.macro_stress, oop.macro_stress(x, g)) < 1e-12
    np.testing.assert_array_equal(snap.history_max, H.max(axis=1))
    np.testing.assert_allclose(snap.history_mean, H.mean(axis=1), rtol=1e-15)
    nx, ny, nz = shape
    k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    red = ((i % nx) + nx * ((j % ny) + ny * (k % nz))).ravel()
    np.testing.assert_array_equal(snap.damage, d[red])
    u = x[:-6].reshape(-1, 3)
    np.testing.assert_array_equal(snap.fluctuation, u[red])
    e = x[-6:]
    X = np.stack([i.ravel() * op.grid.h[0], j.ravel() * op.grid.h[1], k.ravel() * op.grid.h[2]], axis=1)
    ux = e[0] * X[:, 0] + e[5] / 2 * X[:, 1] + e[4] / 2 * X[:, 2] + u[red, 0]
    uz = e[4] / 2 * X[:, 0] + e[3] / 2 * X[:, 1] + e[2] * X[:, 2] + u[red, 2]
    np.testing.assert_allclose(snap.displacement[:, 0], ux, rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(snap.displacement[:, 2], uz, rtol=1e-12, atol=1e-18)
    assert snap.step == 3 and snap.time_value == 0.5 and snap.psi_closure is None
