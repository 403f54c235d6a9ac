"""phast_q4_damage_solve_noalloc (static scratch, no allocation) against
phast_q4_damage_solve through the C ABI (GPU): two calls in a row, the second on a smaller
plate, give the allocating solve's damage, active set, iteration counts and status bit for
bit; a plate beyond the scratch is refused."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _case(dev, nx, ny, seed):
    from phast_amd import q4 as P
    mesh = P.FEMMesh.rectangle(0.25 * nx, 0.25 * ny, nx, ny)
    fem = P.FEMOperators(mesh, P.Material(E=32000.0, nu=0.2, rho=2.45e-9, Gc=3e-3, l0=0.5))
    rng = np.random.default_rng(seed)
    n, E = mesh.n_nodes, mesh.n_elems
    H = torch.as_tensor(rng.uniform(0, 2e-2, (E, 4)), dtype=torch.float64, device=dev)
    d_prev = torch.as_tensor(rng.uniform(0, 0.3, n), dtype=torch.float64, device=dev)
    fixed = np.zeros(n, dtype=np.uint8)
    fixed[(ny // 2) * (nx + 1) + np.arange(nx // 3)] = 1  # a notch held at d = 1
    fx = torch.as_tensor(fixed, device=dev)
    vals = fx.to(torch.float64)
    return fem, H, d_prev, fx, vals


def _solve(name, fem, method, H, d_prev, fx, vals, tol=1e-10):
    from phast_amd import _lib
    plan = fem.plan(0.0)
    n, k = fem.mesh.n_nodes, plan.n_inst
    d_out = torch.full_like(d_prev, float("nan"))
    active = torch.full((n,), 7, dtype=torch.uint8, device=d_prev.device)
    iters, status = (C.c_int32 * k)(), (C.c_int32 * k)()
    rc = getattr(_lib.lib(), name)(plan.handle, method, _lib.ptr(H), None, _lib.ptr(d_prev), None, _lib.ptr(fx),
                                   _lib.ptr(vals), tol, 5000, _lib.ptr(d_out), _lib.ptr(active), iters, status,
                                   _lib.stream())
    torch.cuda.synchronize()
    return rc, d_out, active, list(iters), list(status)


def test_noalloc_damage_solve_equals_the_allocating_solve(dev):
    from phast_amd import _lib
    cases = [_case(dev, 400, 160, 1), _case(dev, 90, 36, 2)]  # 64k nodes, then 3.4k
    for method in (0, 1):
        ref = [_solve("phast_q4_damage_solve", *c[:1], method, *c[1:]) for c in cases]
        got = [_solve("phast_q4_damage_solve_noalloc", *c[:1], method, *c[1:]) for c in cases]
        for (rc0, d0, a0, i0, s0), (rc1, d1, a1, i1, s1) in zip(ref, got):
            assert rc0 == _lib.PHAST_OK and rc1 == _lib.PHAST_OK
            assert s0 == s1 == [_lib.CG_CONVERGED] and i0 == i1 and i0[0] > 5
            assert torch.equal(d0, d1)
            if method == 0:
                assert torch.equal(a0, a1)


def test_noalloc_damage_solve_refuses_a_plate_beyond_its_scratch(dev):
    from phast_amd import _lib
    fem, H, d_prev, fx, vals = _case(dev, 1024, 520, 3)  # 534k nodes > 524288
    rc, *_ = _solve("phast_q4_damage_solve_noalloc", fem, 0, H, d_prev, fx, vals)
    assert rc == _lib.PHAST_ERR_ARG
    assert "too large for the static scratch" in _lib.lib().phast_last_error().decode()
