"""Host-side loading set-up (CPU): Neumann traction assembly and damage preseeding."""
import numpy as np
import pytest
import torch

from phast_amd import loading as L
from phast_amd.q4 import FEMMesh as Q4Mesh
from phast_amd.t3 import FEMMesh as T3Mesh


def test_t3_top_traction_closed_form():
    W, H, nx, ny = 8.0, 3.0, 16, 6
    m = T3Mesh.sent_split_quad(W, H, nx, ny, device="cpu")
    f = L.neumann_forces(m, m.node_sets["top"], [0.0, 2.5]).numpy()
    h = W / nx
    top = m.node_sets["top"]
    expect = np.full(nx + 1, 2.5 * h)
    expect[[0, -1]] = 2.5 * h / 2
    np.testing.assert_allclose(f[top, 1], expect, rtol=1e-14)
    assert np.abs(f[:, 0]).max() == 0.0
    other = np.setdiff1d(np.arange(m.n_nodes), top)
    assert np.abs(f[other]).max() == 0.0
    assert f[:, 1].sum() == pytest.approx(2.5 * W, rel=1e-14)


def test_interior_edges_get_no_traction():
    """A node set along an interior grid line (the notch line) touches only interior edges."""
    m = T3Mesh.sent_split_quad(8.0, 4.0, 8, 4, a=8.0, device="cpu")
    f = L.neumann_forces(m, m.node_sets["crack"], [1.0, 1.0]).numpy()
    assert np.abs(f).max() == 0.0


def test_q4_all_edges_and_reference_edge_quirk():
    W, H, nx, ny = 6.0, 2.0, 12, 4
    q = Q4Mesh.rectangle(W, H, nx, ny, device="cpu")
    t = T3Mesh.sent_split_quad(W, H, nx, ny, device="cpu")
    for side in ("top", "bottom"):
        fq = L.neumann_forces(q, q.node_sets[side], [0.0, 1.0]).numpy()
        ft = L.neumann_forces(t, t.node_sets[side], [0.0, 1.0]).numpy()
        np.testing.assert_allclose(fq, ft, rtol=1e-14, atol=0)  # same node grid, same loads
    # The reference visits edges (0,1), (1,2), (2,0) of a quad: the bottom edge (0,1) is loaded,
    # the top edge (2,3) never is, and the left edge (3,0) is never visited either.
    fq_top = L.neumann_forces(q, q.node_sets["top"], [0.0, 1.0], reference_q4_edges=True).numpy()
    assert np.abs(fq_top).max() == 0.0
    fq_bot = L.neumann_forces(q, q.node_sets["bottom"], [0.0, 1.0], reference_q4_edges=True).numpy()
    assert fq_bot[:, 1].sum() == pytest.approx(W, rel=1e-14)
    fq_right = L.neumann_forces(q, q.node_sets["right"], [1.0, 0.0], reference_q4_edges=True).numpy()
    assert fq_right[:, 0].sum() == pytest.approx(H, rel=1e-14)  # edge (1,2) is the right edge
    fq_left = L.neumann_forces(q, q.node_sets["left"], [1.0, 0.0], reference_q4_edges=True).numpy()
    assert np.abs(fq_left).max() == 0.0


def test_load_and_ramp_factors_scale_linearly():
    m = T3Mesh.sent_split_quad(4.0, 2.0, 8, 4, device="cpu")
    f1 = L.neumann_forces(m, m.node_sets["right"], [3.0, -1.0]).numpy()
    f2 = L.neumann_forces(m, m.node_sets["right"], [3.0, -1.0], load_factor=0.5, ramp_factor=0.25).numpy()
    np.testing.assert_allclose(f2, 0.125 * f1, rtol=1e-15)
    assert f1[:, 0].sum() == pytest.approx(3.0 * 2.0, rel=1e-14)


def test_value_to_H_seed_inverts_the_equilibrium():
    Gc, l0 = 3e-3, 0.25
    for d in (0.1, 0.5, 0.9):
        H = L.value_to_H_seed(d, Gc, l0)
        assert 2 * H / (Gc / l0 + 2 * H) == pytest.approx(d, rel=1e-13)
    assert L.value_to_H_seed(1.0, Gc, l0) == 1e4 * Gc / l0
    assert L.value_to_H_seed(0.0, Gc, l0) == 0.0


def test_preseed_regions_and_one_ring():
    m = T3Mesh.sent_split_quad(8.0, 4.0, 16, 8, a=4.0, device="cpu")
    crack = m.node_sets["crack"]
    mask, vals = L.resolve_preseed_specs(m, L.normalise_legacy_preseed(["crack"]))
    assert np.array_equal(np.nonzero(mask)[0], np.sort(crack)) and np.all(vals[crack] == 1.0)
    seg = [{"region": {"type": "line_segment", "from": [0.0, 2.0], "to": [4.0, 2.0], "thickness": 1e-9}}]
    mask2, _ = L.resolve_preseed_specs(m, seg)
    assert np.array_equal(mask2, mask)
    rect = {"region": {"type": "rectangle", "origin": [0.0, 1.9], "size": [4.0, 0.2]}, "value": 0.5}
    poly = {"region": {"type": "polygon", "vertices": [[-1, 1.9], [4.1, 1.9], [4.1, 2.1], [-1, 2.1]]}, "value": 0.3}
    circ = {"region": {"type": "circle", "center": [8.0, 4.0], "radius": 0.6}, "value": 0.7}
    m3, v3 = L.resolve_preseed_specs(m, [rect, poly, circ])
    assert np.array_equal(m3[crack], np.ones(len(crack), bool)) and np.all(v3[crack] == 0.5)  # max of 0.5, 0.3
    corner = np.nonzero(np.linalg.norm(m.nodes - [8.0, 4.0], axis=1) <= 0.6)[0]
    assert np.all(v3[corner] == 0.7) and m3.sum() == len(crack) + len(corner)
    H = torch.zeros(m.n_elems, dtype=torch.float64)
    n = L.preseed_history(m, H, L.normalise_legacy_preseed(["crack"]), 3e-3, 0.25)
    touched = np.isin(m.elements, crack).any(axis=1)
    assert n == touched.sum()
    assert np.all(H.numpy()[touched] == 1e4 * 3e-3 / 0.25) and np.all(H.numpy()[~touched] == 0.0)
    Hq = torch.full((m.n_elems, 4), 5.0, dtype=torch.float64)  # an (E,4) history keeps larger values
    L.preseed_history(m, Hq, [{"nodes": "crack", "value": 0.5}], 3e-3, 0.25)
    assert torch.all(Hq == 5.0)
    with pytest.raises(ValueError):
        L.resolve_preseed_specs(m, [{"region": {"type": "ellipse"}}])
    with pytest.raises(RuntimeError):
        L.resolve_preseed_specs(m, [{"nodes": "nope"}])
