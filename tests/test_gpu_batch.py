"""The persistent batched explicit step (one kernel per step, optionally graph-captured)
against the per-kernel batched path and against each specimen run alone (GPU).

Every path shares the strip passes, the per-strip partial sums and the per-instance
scalar CG steps, so the results must agree bitwise, including the CG iteration counts.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def specimens(sizes, seed):
    from phast_amd import q4 as P
    mat = P.Material(E=32000.0, nu=0.2, rho=2.45e-9, Gc=3e-3, l0=0.5)
    rng = np.random.default_rng(seed)
    items, gcs, fext, fixed, states = [], [], [], [], []
    for nx in sizes:
        ny = max(nx // 2, 2)
        mesh = P.FEMMesh.rectangle(0.5 * nx, 0.5 * ny, nx, ny)
        items.append((mesh, mat))
        gcs.append(3e-3 * rng.lognormal(0.0, 0.2, mesh.n_elems))
        f = np.zeros((mesh.n_nodes, 2))
        top = np.arange(ny * (nx + 1), (ny + 1) * (nx + 1))
        f[top, 1] = 0.5
        f[np.arange(nx + 1), 1] = -0.5
        fext.append(f)
        jn = ny // 2
        a_n = int(rng.uniform(0.2, 0.5) * nx)
        fixed.append((jn * (nx + 1) + np.arange(a_n + 1), 1.0))
        n, E = mesh.n_nodes, mesh.n_elems
        states.append((rng.uniform(-1e-4, 1e-4, (n, 2)), rng.uniform(-1e-1, 1e-1, (n, 2)), np.zeros((n, 2)),
                       rng.uniform(0.0, 0.3, n), rng.uniform(0.0, 1e-2, (E, 4))))
    return items, gcs, fext, fixed, states


def make(items, gcs, fext, fixed, states, method, dev):
    from phast_amd import batch as B
    bat = B.BatchedExplicit(items, dt_factor=0.4, Gc_fields=gcs, f_ext=fext, fixed=fixed, bounds_method=method,
                            tol=1e-6)
    T = lambda x: torch.as_tensor(x, dtype=torch.float64, device=dev)
    loaded = []
    for k, (u, v, a, d, H) in enumerate(states):
        d = d.copy()
        ids, val = fixed[k]
        d[ids] = val
        loaded.append((T(u), T(v), T(a), T(d), T(H)))
    bat.load_states(loaded)
    return bat


def same(x, y):
    return all(torch.equal(p, q) for p, q in zip(x, y))


@pytest.mark.parametrize("method", ["projected_cg", "post_clamp"])
def test_persistent_step_is_bitwise_the_per_kernel_step(dev, method):
    data = specimens((40, 67, 9, 150), seed=3)
    steps = 6
    ref = make(*data, method, dev)
    iters_ref = []
    for _ in range(steps):
        ref.step(1)
        iters_ref.append(ref.last_iters.copy())
    per = make(*data, method, dev)
    grid = per.prepare_persistent()
    assert min(grid) >= per.n_inst
    iters_per = []
    for _ in range(steps):
        per.step_persistent(1)
        iters_per.append(per.iters_dev.cpu().numpy().astype(np.int64))
        per.check_persistent()
    for k in range(per.n_inst):
        assert same(per.instance_state(k), ref.instance_state(k)), k
    assert np.array_equal(np.array(iters_per), np.array(iters_ref))
    assert max(int(i.max()) for i in iters_ref) > 0  # the damage solves did work


def test_persistent_step_equals_each_specimen_alone(dev):
    data = specimens((33, 70, 12), seed=5)
    bat = make(*data, "projected_cg", dev)
    bat.step_persistent(4)
    bat.check_persistent()
    items, gcs, fext, fixed, states = data
    for k in range(len(items)):
        one = make([items[k]], [gcs[k]], [fext[k]], [fixed[k]], [states[k]], "projected_cg", dev)
        one.step(4)
        assert same(bat.instance_state(k), one.instance_state(0)), k


def test_graph_replay_matches_eager_persistent_steps(dev):
    data = specimens((48, 21, 90), seed=7)
    eager = make(*data, "projected_cg", dev)
    eager.step_persistent(8)
    eager.check_persistent()
    cap = make(*data, "projected_cg", dev)
    g = cap.capture(2)
    for _ in range(4):  # replays with eager work and synchronisation between them
        g.replay(1)
        torch.cuda.synchronize()
        cap.check_persistent(g.status[-1])
        _ = cap.u.sum().item()
    assert cap.steps == 8
    for k in range(cap.n_inst):
        assert same(cap.instance_state(k), eager.instance_state(k)), k


@pytest.mark.parametrize("method", ["projected_cg", "post_clamp"])
def test_several_steps_per_launch_are_bitwise_one_step_launches(dev, method):
    """Instances running ahead of each other inside one launch give the bits (and CG counts)
    of one launch per step: 7 steps as launches of 5 and 2 steps, and as 7 single launches."""
    from phast_amd import _lib
    data = specimens((40, 67, 9, 150), seed=11)
    one = make(*data, method, dev)
    per_step = []
    for _ in range(7):
        one.step_persistent(1)
        one.check_persistent()
        per_step.append(one.iters_dev.cpu().numpy().astype(np.int64))
    run = make(*data, method, dev)
    run.prepare_persistent()
    iters = torch.zeros((5, run.n_inst), dtype=torch.int32, device=dev)
    status = torch.zeros_like(iters)
    run._launch_run(5, iters, status)  # odd: the state moves to the second buffer set
    torch.cuda.synchronize()
    assert bool((status == _lib.CG_CONVERGED).all())
    assert np.array_equal(iters.cpu().numpy(), np.array(per_step[:5]))
    run.step_persistent(2, steps_per_launch=2)
    run.check_persistent()
    assert np.array_equal(run.iters_dev.cpu().numpy(), per_step[-1])
    for k in range(run.n_inst):
        assert same(run.instance_state(k), one.instance_state(k)), k
    assert max(int(i.max()) for i in per_step) > 0


def test_graph_of_several_step_launches_matches_eager_steps(dev):
    data = specimens((48, 21, 90, 130), seed=13)
    eager = make(*data, "projected_cg", dev)
    eager.step_persistent(12)
    eager.check_persistent()
    cap = make(*data, "projected_cg", dev)
    g = cap.capture(6, steps_per_launch=3)
    for _ in range(2):  # replays with eager work and synchronisation between them
        g.replay(1)
        torch.cuda.synchronize()
        cap.check_persistent(g.status)
        _ = cap.u.sum().item()
    assert cap.steps == 12
    for k in range(cap.n_inst):
        assert same(cap.instance_state(k), eager.instance_state(k)), k
