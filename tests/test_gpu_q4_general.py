"""General quadrilateral meshes on the GPU (the reference's Q4 branch on any quad mesh):
distorted, renumbered meshes against the oracle, and a renumbered rectangle against the
structured strip kernels."""
import numpy as np
import pytest
import torch

from oracle import q4 as O
from tests.test_oracle_q4 import distorted_q4_mesh

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    b = b.detach().cpu().numpy() if isinstance(b, torch.Tensor) else np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


T = lambda x, dev: torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device=dev)


def setup(split="amor", plane_stress=False, nx=12, ny=7, seed=0, **kw):
    from phast_amd import q4 as P
    nodes, quads = distorted_q4_mesh(nx=nx, ny=ny, W=float(nx), H=float(ny), seed=seed)
    mat = dict(E=32000.0, nu=0.2, rho=2.45e-9, Gc=3e-3, l0=0.5, energy_split=split, plane_stress=plane_stress, **kw)
    mesh = P.FEMMesh(nodes, quads)
    assert not mesh.structured
    fem = P.FEMOperators(mesh, P.Material(**mat))
    ofem = O.FEMOperators(O.Q4Mesh(nodes, quads), O.Material(**mat))
    return P, fem, ofem


@pytest.mark.parametrize("split,plane_stress", [("amor", False), ("amor", True), ("isotropic", False)])
def test_general_q4_mechanics(dev, split, plane_stress):
    P, fem, ofem = setup(split, plane_stress)
    n, E = fem.mesh.n_nodes, fem.mesh.n_elems
    rng = np.random.default_rng(1)
    u = rng.uniform(-1e-3, 1e-3, (n, 2))
    d = rng.uniform(0, 1, n)
    assert rel(fem.M_scalar, ofem.mesh.M_scalar) < 1e-14
    assert abs(fem.mesh.h_min - ofem.mesh.h_min) == 0.0
    assert rel(fem.internal_force(T(u, dev), T(d, dev)), ofem.internal_force(u, d)) < RTOL
    for got, want in zip(fem.compute_strain(T(u, dev)), ofem.compute_strain(u)):
        assert rel(got, want) < RTOL
    psi = ofem.compute_psi_plus(u)
    assert rel(fem.compute_psi_plus(T(u, dev)), psi) < RTOL
    H = rng.uniform(0, 1.5, (E, 4)) * psi.max()
    p_dev, H_dev = fem.driving_force_and_history(T(u, dev), T(H, dev))
    np.testing.assert_array_equal(H_dev.cpu().numpy() == H, H >= psi)   # the max takes one operand exactly
    assert rel(H_dev, np.maximum(H, psi)) < RTOL
    fq = rng.uniform(0, 1, (E, 4))
    assert rel(fem.elem_to_node(T(fq, dev)), ofem.mesh.elem_to_node(fq)) < RTOL


def test_general_q4_explicit_step_with_bc_and_damping(dev):
    P, fem, ofem = setup("amor")
    n, E = fem.mesh.n_nodes, fem.mesh.n_elems
    rng = np.random.default_rng(2)
    u, v, a = (rng.uniform(-1e-4, 1e-4, (n, 2)) for _ in range(3))
    d = rng.uniform(0, 0.6, n)
    f_ext = rng.uniform(-1, 1, (n, 2))
    bc = np.zeros((n, 2), dtype=bool)
    bc[fem.mesh.node_sets["bottom"], 1] = True
    bc[fem.mesh.node_sets["left"], 0] = True
    bcv = np.where(bc, 1e-5, 0.0)
    for zeta in (0.0, 0.3):
        dyn = P.ExplicitDynamics(fem, dt=None, damping_ratio_max=zeta)
        odyn = O.ExplicitDynamics(ofem, dt=dyn.dt, damping_ratio_max=zeta)
        u1, v1, a1, H1, psi = dyn.step_with_history(T(u, dev), T(v, dev), T(a, dev), T(d, dev), T(f_ext, dev),
                                                    torch.as_tensor(bc, device=dev), T(bcv, dev),
                                                    T(np.zeros((E, 4)), dev))
        ou, ov, oa = odyn.step(u, v, a, d, f_ext, bc, bcv)
        assert rel(u1, ou) < RTOL and rel(v1, ov) < RTOL and rel(a1, oa) < RTOL
        assert rel(psi, ofem.compute_psi_plus(ou)) < RTOL and rel(H1, ofem.compute_psi_plus(ou)) < RTOL
        u2, v2, a2 = dyn.step(T(u, dev), T(v, dev), T(a, dev), T(d, dev), T(f_ext, dev))
        ou2, ov2, oa2 = odyn.step(u, v, a, d, f_ext)
        assert rel(u2, ou2) < RTOL and rel(v2, ov2) < RTOL and rel(a2, oa2) < RTOL


@pytest.mark.parametrize("method", ["projected_cg", "post_clamp"])
@pytest.mark.parametrize("field", [False, True])
def test_general_q4_damage(dev, method, field):
    P, fem, ofem = setup("amor", seed=5)
    n, E = fem.mesh.n_nodes, fem.mesh.n_elems
    rng = np.random.default_rng(3)
    Gc_e = 3e-3 * rng.lognormal(0.0, 0.2, E) if field else None
    dmg = P.PhaseFieldDamageSolver(fem, tol=1e-12, max_iter=2000, bounds_method=method,
                                   Gc_field=None if Gc_e is None else T(Gc_e, dev))
    odmg = O.PhaseFieldDamageSolver(ofem, tol=1e-12, max_iter=2000, bounds_method=method, Gc_e=Gc_e)
    H = rng.uniform(0, 2e-2, (E, 4))
    d_prev = rng.uniform(0, 0.3, n)
    rho, b, M_inv = odmg.prepare(H, d_prev)
    gb, gdiag = dmg.damage_rhs_diag(T(H, dev))
    assert rel(gb, b) < RTOL and rel(1.0 / (gdiag + 1e-30), M_inv) < RTOL
    x = rng.uniform(0, 1, n)
    assert rel(dmg.damage_apply(T(H, dev), T(x, dev)), odmg.Ax(x, rho)) < RTOL
    fixed = np.zeros(n, dtype=bool)
    fixed[rng.choice(n, 5, replace=False)] = True
    vals = np.where(fixed, 1.0, 0.0)
    d = dmg.solve(T(H, dev), T(d_prev, dev), pf_dirichlet_mask=torch.as_tensor(fixed, device=dev),
                  pf_dirichlet_values=T(vals, dev))
    od = odmg.solve(H, d_prev, fixed=fixed, vals=vals)
    print(f"{method} field={field}: iterations {dmg.last_iter} (oracle {odmg.last_iter}), rel {rel(d, od):.2e}")
    assert dmg.last_converged and dmg.last_iter == odmg.last_iter
    assert rel(d, od) < 1e-10


def test_general_q4_staggered_steps_track_oracle(dev):
    """20 explicit staggered steps (projected damage solve each step) on a distorted mesh."""
    P, fem, ofem = setup("amor", nx=24, ny=10, seed=7)
    n, E = fem.mesh.n_nodes, fem.mesh.n_elems
    f_ext = np.zeros((n, 2))
    f_ext[fem.mesh.node_sets["top"], 1] = 40.0
    f_ext[fem.mesh.node_sets["bottom"], 1] = -40.0
    dt = 0.5 * fem.dt_cfl
    dyn = P.ExplicitDynamics(fem, dt=dt)
    dmg = P.PhaseFieldDamageSolver(fem, tol=1e-12, max_iter=2000, bounds_method="projected_cg")
    st = P.StaggeredSolver(fem, dyn, dmg, f_ext=T(f_ext, dev), damage_every=1)
    nodes = fem.mesh.nodes
    d0 = np.where((np.abs(nodes[:, 1] - 5.0) < 0.6) & (nodes[:, 0] <= 8.0), 1.0, 0.0)
    st.d = T(d0, dev)
    ost = O.ExplicitState(np.zeros((n, 2)), np.zeros((n, 2)), np.zeros((n, 2)), d0.copy(), np.zeros((E, 4)))
    odyn = O.ExplicitDynamics(ofem, dt=dt)
    odmg = O.PhaseFieldDamageSolver(ofem, tol=1e-12, max_iter=2000, bounds_method="projected_cg")
    for _ in range(20):
        st.step_full()
        O.step_full_explicit(odyn, odmg, ost, f_ext)
    print(f"20 steps: u {rel(st.u, ost.u):.1e} v {rel(st.v, ost.v):.1e} d {rel(st.d, ost.d):.1e} "
          f"H {rel(st.H_elem, ost.H):.1e}")
    assert rel(st.u, ost.u) < 1e-10 and rel(st.v, ost.v) < 1e-10 and rel(st.d, ost.d) < 1e-10
    assert rel(st.H_elem, ost.H) < 1e-10
    assert float(st.d.max()) > 0.0


def test_renumbered_rectangle_matches_the_strip_kernels(dev):
    """A rectangle whose nodes and cells are renumbered runs on the general plan; it gives the
    structured plate's force, step and damage solve (same mesh, other numbering)."""
    from phast_amd import q4 as P
    nodes, quads = O.structured_q4_mesh(10.0, 4.0, 20, 8)
    rng = np.random.default_rng(11)
    perm = rng.permutation(nodes.shape[0])
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    mat = P.Material(E=32000.0, nu=0.2, rho=2.45e-9, Gc=3e-3, l0=0.5)
    s_mesh = P.FEMMesh(nodes, quads)
    g_mesh = P.FEMMesh(nodes[perm], inv[quads])
    assert s_mesh.structured and not g_mesh.structured
    sf, gf = P.FEMOperators(s_mesh, mat), P.FEMOperators(g_mesh, mat)
    n = nodes.shape[0]
    u = rng.uniform(-1e-3, 1e-3, (n, 2))
    d = rng.uniform(0, 1, n)
    fs = sf.internal_force(T(u, dev), T(d, dev)).cpu().numpy()
    fg = gf.internal_force(T(u[perm], dev), T(d[perm], dev)).cpu().numpy()
    assert rel(fg, fs[perm]) < 1e-12
    assert rel(gf.M_scalar.cpu().numpy(), sf.M_scalar.cpu().numpy()[perm]) < 1e-14
    H = rng.uniform(0, 1e-2, (quads.shape[0], 4))
    ds = P.PhaseFieldDamageSolver(sf, tol=1e-12, bounds_method="projected_cg").solve(T(H, dev), T(d * 0.2, dev))
    dg = P.PhaseFieldDamageSolver(gf, tol=1e-12, bounds_method="projected_cg").solve(T(H, dev), T(d[perm] * 0.2, dev))
    assert rel(dg, ds.cpu().numpy()[perm]) < 1e-10
