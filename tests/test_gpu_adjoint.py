"""Adjoint of the voxel reaction curve (GPU): each adjoint kernel against the oracle's
finite-difference-pinned restatement, then the reverse sweep's gradient against central
finite differences of the forward solver itself."""
import numpy as np
import pytest
import torch

from oracle import hex8 as O
from tests.test_gpu_hex8 import T, build, rel

pytestmark = pytest.mark.gpu
RTOL = 1e-10


def test_adjoint_kernels_match_oracle(dev):
    S, op, oop = build((6, 5, 7), seed=3)
    rng = np.random.default_rng(4)
    x = rng.uniform(-1e-3, 1e-3, op.n_mech)
    l = rng.uniform(-1, 1, op.n_mech)
    d = rng.uniform(0, 0.9, op.n_red)
    ld = rng.uniform(-1, 1, op.n_red)
    t = rng.uniform(0.5, 1.5, (op.n_elements, 8))
    f = rng.uniform(-1, 1, (op.n_elements, 8))
    Sq = rng.uniform(-1, 1, (op.n_elements, 8, 6))
    got = op.qp_to_node(op.adj_coupling(T(x, dev), T(l, dev), T(d, dev)))
    assert rel(got, O.adjoint_coupling(oop, x, l, d)) < RTOL
    assert rel(op.qp_to_node(T(f, dev), 2.5), 2.5 * O.qp_to_node(oop, f)) < RTOL
    assert rel(op.qp_stress_apply(T(Sq, dev)), O.qp_stress_apply(oop, Sq)) < RTOL
    assert rel(op.qp_stress_apply(op.adj_driving(T(x, dev), T(t, dev))), O.adjoint_driving(oop, x, t)) < 1e-9
    assert rel(op.adj_damage_history(T(d, dev), T(ld, dev), 3.0), 3.0 * O.adjoint_damage_history(oop, d, ld)) < RTOL
    assert rel(op.adj_damage_gc(T(d, dev), T(ld, dev), 0.5), 0.5 * O.adjoint_damage_gc(oop, d, ld)) < RTOL


def _curve(shape=(8, 8, 8), steps=(1e-3, 2e-3, 3e-3, 4e-3), checkpoint_every=1):
    from phast_amd import adjoint3d as A
    from phast_amd import geometry as G
    from phast_amd import solid3d as S
    h = (1e-3,) * 3
    phase = G.sphere_pore_radius(shape, h, 2.5e-3)
    phases = {"mi_matrix": S.PhaseMaterial("mi_matrix", 3000.0, 0.35, 5e-4),
              "pore": S.PhaseMaterial("pore", 3000.0 * 1e-6, 0.35, 5e-4)}
    mats = S.build_element_arrays(phase, phases)
    load = S.LoadCase("xx", 0, steps, (1, 2, 3, 4, 5), ())
    op = S.PeriodicVoxelPhaseFieldOperator(S.VoxelGrid(shape, h), mats, load, kappa=1e-6, ell=2e-3)
    return A, op, A.ReactionCurve(op, checkpoint_every=checkpoint_every)


def test_gradient_matches_central_differences(dev):
    A, op, curve = _curve()
    rng = np.random.default_rng(11)
    E = op.n_elements
    Gc = T(5e-4 * (1 + 0.1 * rng.standard_normal(E)), dev)
    target = T(5e-4 * (1 + 0.1 * rng.standard_normal(E)), dev)
    R_hat, _ = curve.forward(target)
    R_hat = R_hat.detach().clone()

    def loss(g):
        R, _ = curve.forward(g)
        return float(((R - R_hat) ** 2).sum())

    Gc_var = Gc.clone().requires_grad_(True)
    R = A.reaction_curve(Gc_var, curve)
    L = ((R - R_hat) ** 2).sum()
    L.backward()
    grad = Gc_var.grad.detach()
    assert float(torch.linalg.vector_norm(grad)) > 0
    assert max(curve.stats.block_iterations) < curve.max_block
    # damage must have developed for the test to mean something
    R0, tape = curve.forward(Gc)
    assert float(tape["records"][len(R0) - 1].d.max()) > 0.05
    # central differences on the voxels of largest sensitivity and a few random ones
    order = torch.argsort(grad.abs(), descending=True).cpu().numpy()
    picks = list(order[:4]) + list(rng.choice(E, 4, replace=False))
    for e in picks:
        h = 1e-6 * float(Gc[e])
        gp, gm = Gc.clone(), Gc.clone()
        gp[e] += h
        gm[e] -= h
        fd = (loss(gp) - loss(gm)) / (2 * h)
        scale = max(abs(fd), 1e-6 * float(grad.abs().max()))
        assert abs(float(grad[e]) - fd) <= 1e-6 * scale + 1e-9 * float(grad.abs().max()), (e, float(grad[e]), fd)
    # a random direction over every voxel
    v = T(rng.standard_normal(E), dev) * Gc
    h = 1e-6
    fd = (loss(Gc + h * v) - loss(Gc - h * v)) / (2 * h)
    ad = float(torch.dot(grad, v))
    assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)


def test_checkpointed_sweep_equals_stored_sweep(dev):
    A, op, curve1 = _curve(steps=(1e-3, 2e-3, 3e-3))
    curve2 = A.ReactionCurve(op, checkpoint_every=2)
    rng = np.random.default_rng(5)
    Gc = T(5e-4 * (1 + 0.1 * rng.standard_normal(op.n_elements)), dev)
    w = T(rng.standard_normal(3), dev)
    R1, tape1 = curve1.forward(Gc)
    g1 = curve1.backward(tape1, w)
    R2, tape2 = curve2.forward(Gc)
    g2 = curve2.backward(tape2, w)
    assert torch.equal(R1, R2)
    assert torch.equal(g1, g2)
