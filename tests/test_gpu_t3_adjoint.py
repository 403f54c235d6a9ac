"""Parity of the T3 damage adjoint (HIP) with the oracle, and its gradients against central
differences of device solves (GPU).

The per-triangle adjoint terms round like the oracle and equal it bit for bit.  The
adjoint and unconstrained-damage PCG solves sum their dot products in another order, so
they are held to the oracle's iteration counts and to 1e-9 relative; the gradients built
from them likewise.  The reference's own full-solver test (a unit-square mesh, tol 1e-12,
loss sum d^2, directional central difference, rtol 1e-5) is restated on the device.
"""
import numpy as np
import pytest
import torch

from oracle import t3 as O
from oracle import t3_adjoint as OA
from oracle.q4 import Material as OMaterial
from tests.test_gpu_t3 import T, exact, rel, scrambled_mesh

pytestmark = pytest.mark.gpu
TOL = 1e-9


def setup(nodes, els, pf_model="AT2", **kw):
    from phast_amd import t3 as P
    mk = dict(E=32000.0, nu=0.2, rho=2.45e-9, Gc=2.7, l0=0.5, pf_model=pf_model)
    mk.update(kw)
    fem = P.FEMOperators(P.FEMMesh(nodes, els), P.Material(**mk))
    return P, fem, O.T3Mesh(nodes, els), OMaterial(**mk)


def pins(nodes, value=0.35):
    fixed = nodes[:, 0] < 1e-9
    return fixed, np.where(fixed, value, 0.0)


@pytest.mark.parametrize("with_pins", [False, True])
def test_adjoint_pcg_matches_oracle(dev, with_pins):
    nodes, els = scrambled_mesh(23, 17, seed=31)
    P, fem, om, omat = setup(nodes, els)
    rng = np.random.default_rng(32)
    H = rng.uniform(0.0, 5.0, len(els))
    rhs = rng.uniform(-1.0, 1.0, len(nodes))
    fixed, vals = pins(nodes) if with_pins else (None, None)
    dmg = P.PhaseFieldDamageSolver(fem, max_iter=5000)
    mask = None if fixed is None else torch.as_tensor(fixed, device=dev)
    vt = None if vals is None else T(vals, dev)
    op = O.T3Damage(om, omat)
    minv = 1.0 / (op.diagonal(H) + 1e-30)
    for r, v, rt in ((rhs, None, 1e-22), (None, vals, 1e-24)):
        x = dmg.adjoint_pcg(T(H, dev), None if r is None else T(r, dev), mask, None if v is None else vt, rt)
        ox, oit = OA.pcg_adjoint(lambda y: op.apply(y, H), op.rhs(H) if r is None else r, minv, fixed, v,
                                 None, rt, 1000)
        assert dmg.last_adjoint_iter == oit, (dmg.last_adjoint_iter, oit)
        assert rel(x, ox) < TOL
        if fixed is not None:
            assert np.all(x.cpu().numpy()[fixed] == (0.0 if v is None else v[fixed]))


def test_adjoint_terms_bitwise(dev):
    nodes, els = scrambled_mesh(19, 15, seed=33)
    P, fem, om, omat = setup(nodes, els)
    rng = np.random.default_rng(34)
    lam, dp, du = (rng.uniform(-1.0, 1.0, len(nodes)) for _ in range(3))
    ne = len(els)
    out = [torch.empty(ne, dtype=torch.float64, device=dev) for _ in range(4)]
    from phast_amd import _lib
    ins = [T(a, dev) for a in (lam, dp, du)]  # held until the kernel has run
    _lib.check(_lib.lib().phast_t3_adjoint_terms(fem.plan.handle, *[_lib.ptr(t) for t in ins],
                                                 *[_lib.ptr(o) for o in out], _lib.stream()))
    torch.cuda.synchronize()
    assert exact(out[0], OA.history_gradient(om, lam, dp))
    for o, ref in zip(out[1:], OA.quadratic_forms_perelem(om, lam, du)):
        assert exact(o, ref)


def _state(nodes, els, P, fem, om, omat, seed, fixed=None, vals=None, Gc_e=None):
    rng = np.random.default_rng(seed)
    H = rng.uniform(0.5, 6.0, len(els))
    d_prev = np.clip(rng.uniform(-0.4, 0.5, len(nodes)), 0.0, 1.0)
    w = rng.uniform(-1.0, 1.0, len(nodes))
    return H, d_prev, w


@pytest.mark.parametrize("bounds", ["post_clamp", "projected_cg"])
@pytest.mark.parametrize("pf_model", ["AT2", "AT1"])
@pytest.mark.parametrize("mode", ["H", "scalar", "field"])
@pytest.mark.parametrize("with_pins", [False, True])
def test_backward_matches_oracle(dev, pf_model, mode, with_pins, bounds):
    nodes, els = scrambled_mesh(21, 13, seed=35)
    P, fem, om, omat = setup(nodes, els, pf_model)
    H, d_prev, w = _state(nodes, els, P, fem, om, omat, 36)
    fixed, vals = pins(nodes) if with_pins else (None, None)
    mask = None if fixed is None else torch.as_tensor(fixed, device=dev)
    vt = None if vals is None else T(vals, dev)
    Gc_e = omat.Gc * (1.0 + 0.3 * np.random.default_rng(37).uniform(-1.0, 1.0, len(els)))
    dmg = P.PhaseFieldDamageSolver(fem, tol=1e-10, max_iter=3000, bounds_method=bounds)
    Ht = T(H, dev).requires_grad_(True)
    dp = T(d_prev, dev).requires_grad_(True)
    kw = {}
    if mode == "H":
        dmg.differentiable = True
    elif mode == "scalar":
        kw = dict(Gc=torch.tensor(2.9, dtype=torch.float64, requires_grad=True),
                  l0=torch.tensor(0.45, dtype=torch.float64, requires_grad=True))
    else:
        kw = dict(Gc_field=T(Gc_e, dev).requires_grad_(True),
                  l0=torch.tensor(0.45, dtype=torch.float64, requires_grad=True))
    d = dmg.solve(Ht, dp, pf_dirichlet_mask=mask, pf_dirichlet_values=vt, **kw)
    assert d.requires_grad
    (d * T(w, dev)).sum().backward()
    adj = OA.T3DamageAdjoint(om, omat, tol=1e-10, max_iter=3000, bounds_method=bounds)
    ok = dict(Gc=2.9, l0=0.45) if mode == "scalar" else dict(l0=0.45, Gc_e=Gc_e) if mode == "field" else {}
    od = adj.forward(H, d_prev, mode, fixed=fixed, vals=vals, **ok)
    assert rel(d, od) < 1e-10
    # the backward from the device's forward solution
    ref = adj.backward(H, d_prev, d.detach().cpu().numpy(), w, mode, fixed=fixed, vals=vals, **ok)
    assert rel(dmg.last_adjoint["d_unc"], ref["d_unc"]) < TOL
    assert rel(dmg.last_adjoint["lam"], ref["lam"]) < TOL
    assert rel(Ht.grad, ref["H"]) < TOL
    assert exact(dp.grad, ref["d_prev"])
    if mode == "scalar":
        assert abs(float(kw["Gc"].grad) - ref["Gc"]) <= TOL * abs(ref["Gc"])
        assert abs(float(kw["l0"].grad) - ref["l0"]) <= TOL * abs(ref["l0"])
    elif mode == "field":
        assert rel(kw["Gc_field"].grad, ref["Gc"]) < TOL
        assert abs(float(kw["l0"].grad) - ref["l0"]) <= TOL * abs(ref["l0"])
    # the plan's constants are restored: a plain solve equals the oracle's at the base material
    d2 = dmg.solve(T(H, dev), T(d_prev, dev), pf_dirichlet_mask=mask, pf_dirichlet_values=vt)
    od2 = O.T3DamageSolver(om, omat, tol=1e-10, max_iter=3000, bounds_method=bounds).solve(
        H, d_prev, fixed=fixed, vals=vals)
    assert rel(d2, od2) < 1e-10


def _unit_square(n=8):
    xs = np.linspace(0.0, 1.0, n + 1)
    yy, xx = np.meshgrid(xs, xs, indexing="ij")
    nodes = np.stack([xx.ravel(), yy.ravel()], axis=1)
    tris = []
    for j in range(n):
        for i in range(n):
            a = j * (n + 1) + i
            b, c, d = a + 1, a + n + 2, a + n + 1
            tris += [[a, b, c], [a, c, d]]
    return nodes, np.asarray(tris)


def test_history_gradient_is_not_detached_with_fixed_material(dev):
    """The reference's full-solver adjoint regression test on the device: Gc passed without a
    gradient still keeps the damage on the graph, and grad_H . v matches a central
    difference of two device solves."""
    nodes, els = _unit_square()
    P, fem, om, omat = setup(nodes, els, E=210000.0, nu=0.3, Gc=2.7, l0=0.1, rho=7.8e-9,
                             energy_split="spectral", eta_residual=1e-7)
    solver = P.PhaseFieldDamageSolver(fem, tol=1e-12, max_iter=2000)
    gen = torch.Generator().manual_seed(0)
    ne = len(els)
    H = (5.0 + 3.0 * torch.rand(ne, dtype=torch.float64, generator=gen)).to(dev).requires_grad_(True)
    d_prev = torch.zeros(len(nodes), dtype=torch.float64, device=dev)
    Gc = torch.tensor(2.7, dtype=torch.float64)
    out = solver.solve(H, d_prev, Gc=Gc)
    assert out.requires_grad
    grad_H = torch.autograd.grad((out ** 2).sum(), H)[0]
    assert float(grad_H.abs().max()) > 0.0
    direction = torch.randn(ne, dtype=torch.float64, generator=gen).to(dev)
    h = 1e-6
    with torch.no_grad():
        plus = (solver.solve(H.detach() + h * direction, d_prev, Gc=Gc) ** 2).sum()
        minus = (solver.solve(H.detach() - h * direction, d_prev, Gc=Gc) ** 2).sum()
    fd = float((plus - minus) / (2 * h))
    assert abs(float(grad_H @ direction) - fd) <= 1e-5 * abs(fd) + 1e-10


@pytest.mark.parametrize("pf_model", ["AT2", "AT1"])
def test_scalar_and_field_gradients_match_device_differences(dev, pf_model):
    nodes, els = scrambled_mesh(15, 11, seed=38)
    P, fem, om, omat = setup(nodes, els, pf_model)
    rng = np.random.default_rng(39)
    H = T(rng.uniform(2.0, 6.0, len(els)), dev)
    d_prev = torch.zeros(len(nodes), dtype=torch.float64, device=dev)
    w = T(rng.uniform(-1.0, 1.0, len(nodes)), dev)
    solver = P.PhaseFieldDamageSolver(fem, tol=1e-13, max_iter=5000)
    Gc = torch.tensor(2.7, dtype=torch.float64, requires_grad=True)
    l0 = torch.tensor(0.5, dtype=torch.float64, requires_grad=True)
    d = solver.solve(H, d_prev, Gc=Gc, l0=l0)
    assert bool(((d > 1e-9) & (d < 1.0 - 1e-9)).all())  # every node interior
    (w @ d).backward()

    def L(**k):
        with torch.no_grad():
            return float(w @ solver.solve(H, d_prev, **k))

    for name, t in (("Gc", Gc), ("l0", l0)):
        b = float(t.detach())
        h = 1e-6 * b
        other = {"l0": 0.5} if name == "Gc" else {"Gc": 2.7}
        fd = (L(**{name: b + h}, **other) - L(**{name: b - h}, **other)) / (2 * h)
        assert abs(float(t.grad) - fd) <= 1e-6 * abs(fd), (name, float(t.grad), fd)
    Ge = T(np.full(len(els), 2.7), dev).requires_grad_(True)
    d = solver.solve(H, d_prev, Gc_field=Ge)
    (w @ d).backward()
    assert abs(float(Ge.grad.sum()) - float(Gc.grad)) <= 1e-9 * abs(float(Gc.grad))


@pytest.mark.parametrize("pf_model", ["AT2", "AT1"])
def test_gamma_correction_solve_and_scalar_adjoint(dev, pf_model):
    """Gamma correction (damage_solver.py:1409-1433, 545-781): the solve runs with
    Gc_e = Gc gamma_e, and the scalar rule reduces the per-element Gc gradient through gamma_e
    with gamma_e(l0) held fixed in the l0 gradient."""
    from phast_amd import t3 as P
    nodes, els = scrambled_mesh(17, 12, seed=40)
    mk = dict(E=32000.0, nu=0.2, rho=2.45e-9, Gc=2.7, l0=0.5, pf_model=pf_model)
    fem = P.FEMOperators(P.FEMMesh(nodes, els), P.Material(gamma_correction=True, **mk))
    om, omat = O.T3Mesh(nodes, els), OMaterial(**mk)
    cw = 2.0 if pf_model == "AT2" else 8.0 / 3.0
    rng = np.random.default_rng(42)
    H = rng.uniform(2.0, 6.0, len(els))
    d_prev = np.clip(rng.uniform(-0.4, 0.5, len(nodes)), 0.0, 1.0)
    w = rng.uniform(-1.0, 1.0, len(nodes))
    dmg = P.PhaseFieldDamageSolver(fem, tol=1e-10, max_iter=3000)
    d = dmg.solve(T(H, dev), T(d_prev, dev))
    gam = 1.0 / (1.0 + om.elem_h / (cw * 0.5))
    od = O.T3DamageSolver(om, omat, tol=1e-10, max_iter=3000, bounds_method="post_clamp", Gc_e=2.7 * gam).solve(
        H, d_prev)
    assert rel(d, od) < 1e-10
    Ht = T(H, dev).requires_grad_(True)
    Gc = torch.tensor(2.9, dtype=torch.float64, requires_grad=True)
    l0 = torch.tensor(0.45, dtype=torch.float64, requires_grad=True)
    with pytest.warns(RuntimeWarning):
        d = dmg.solve(Ht, T(d_prev, dev), Gc=Gc, l0=l0)
    (d * T(w, dev)).sum().backward()
    gam2 = 1.0 / (1.0 + om.elem_h / (cw * 0.45))
    adj = OA.T3DamageAdjoint(om, omat, tol=1e-10, max_iter=3000)
    assert rel(d, adj.forward(H, d_prev, "field", l0=0.45, Gc_e=2.9 * gam2)) < 1e-10
    ref = adj.backward(H, d_prev, d.detach().cpu().numpy(), w, "field", l0=0.45, Gc_e=2.9 * gam2)
    gGc = float((gam2 * ref["Gc"]).sum())
    assert abs(float(Gc.grad) - gGc) <= TOL * abs(gGc)
    assert abs(float(l0.grad) - ref["l0"]) <= TOL * abs(ref["l0"])
    assert rel(Ht.grad, ref["H"]) < TOL
    # the same gradient through make_gamma_corrected_Gc_field and the field rule
    Gs = torch.tensor(2.7, dtype=torch.float64, device=dev, requires_grad=True)
    d = dmg.solve(T(H, dev), T(d_prev, dev), Gc_field=P.make_gamma_corrected_Gc_field(dmg, Gs))
    (d * T(w, dev)).sum().backward()
    ref = adj.backward(H, d_prev, d.detach().cpu().numpy(), w, "field", l0=0.5, Gc_e=2.7 * gam)
    g = float((gam * ref["Gc"]).sum())
    assert abs(float(Gs.grad) - g) <= TOL * abs(g)


def test_backward_edge_states(dev):
    """Edge states of the active-set split: every node held by irreversibility (the whole
    upstream gradient passes to d_prev, lam = 0, dL/dH = 0) and saturated nodes (d_prev = d = 1:
    upper-active, no gradient to either input)."""
    nodes, els = scrambled_mesh(9, 7, seed=43)
    P, fem, om, omat = setup(nodes, els)
    n, ne = len(nodes), len(els)
    rng = np.random.default_rng(44)
    w = rng.uniform(-1.0, 1.0, n)
    dmg = P.PhaseFieldDamageSolver(fem, tol=1e-10, max_iter=3000)
    dmg.differentiable = True
    H = T(np.full(ne, 1e-6), dev).requires_grad_(True)
    dp = T(np.full(n, 0.9), dev).requires_grad_(True)  # far above the tiny-history solution
    d = dmg.solve(H, dp)
    assert np.all(d.detach().cpu().numpy() == 0.9)
    (d * T(w, dev)).sum().backward()
    assert exact(dp.grad, w) and float(H.grad.abs().max()) == 0.0
    assert float(dmg.last_adjoint["lam"].abs().max()) == 0.0
    H2 = T(np.full(ne, 1e-6), dev).requires_grad_(True)
    sat = rng.uniform(0, 1, n) < 0.3
    dp2 = T(np.where(sat, 1.0, 0.9), dev).requires_grad_(True)
    d2 = dmg.solve(H2, dp2)
    (d2 * T(w, dev)).sum().backward()
    assert exact(dp2.grad, np.where(sat, 0.0, w))
    assert float(H2.grad.abs().max()) == 0.0


def test_adjoint_pcg_stopping_edges(dev):
    """The device PCG's edges match the oracle's: max_iter = 0 returns the start (pins at their
    values, 0 elsewhere) after no iteration; a zero right-hand side stops on p.Ap <= 0 in the
    first iteration with x = 0."""
    nodes, els = scrambled_mesh(9, 7, seed=45)
    P, fem, om, omat = setup(nodes, els)
    fixed, vals = pins(nodes)
    dmg = P.PhaseFieldDamageSolver(fem, max_iter=0)
    H = T(np.full(len(els), 1.0), dev)
    rhs = T(np.random.default_rng(46).uniform(-1.0, 1.0, len(nodes)), dev)
    x = dmg.adjoint_pcg(H, rhs, torch.as_tensor(fixed, device=dev), T(vals, dev), 1e-22)
    assert dmg.last_adjoint_iter == 0 and exact(x, np.where(fixed, vals, 0.0))
    dmg.max_iter = 50
    x = dmg.adjoint_pcg(H, torch.zeros_like(rhs), None, None, 1e-22)
    assert dmg.last_adjoint_iter == 1 and exact(x, np.zeros(len(nodes)))
