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
