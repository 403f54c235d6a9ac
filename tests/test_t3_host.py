"""Host-side pieces of the T3 route (CPU): the structured split-quad mesh generator."""
import numpy as np
import pytest

from phast_amd.t3 import FEMMesh


def loop_split_quad(nx, ny):
    """Cell-by-cell construction (core/mesh_generator.py:1607-1626)."""
    node = lambda i, j: j * (nx + 1) + i
    tris = []
    for j in range(ny):
        js = min(j, ny - 1 - j)
        for i in range(nx):
            a, b, c, d = node(i, j), node(i + 1, j), node(i, j + 1), node(i + 1, j + 1)
            tris += [[a, b, d], [a, d, c]] if (i + js) % 2 == 0 else [[a, b, c], [b, d, c]]
    return np.asarray(tris)


def test_split_quad_matches_cell_loop():
    for nx, ny in ((2, 2), (5, 4), (8, 7)):
        m = FEMMesh.sent_split_quad(10.0, 4.0, nx, ny, a=5.0 if nx % 2 == 0 and ny % 2 == 0 else None, device="cpu")
        assert np.array_equal(m.elements, loop_split_quad(nx, ny))
        assert m.n_nodes == (nx + 1) * (ny + 1) and m.n_elems == 2 * nx * ny
        assert np.allclose(m.areas_np, 10.0 * 4.0 / (2 * nx * ny))
        assert np.all(m.nodes[m.node_sets["top"], 1] == 4.0) and np.all(m.nodes[m.node_sets["left"], 0] == 0.0)


def test_split_quad_diagonals_mirror_about_mid_height():
    """The diagonal choice of row j equals that of row ny-1-j, and alternates along a row."""
    nx, ny = 8, 6
    m = FEMMesh.sent_split_quad(8.0, 6.0, nx, ny, a=4.0, device="cpu")
    first = m.elements[0::2].reshape(ny, nx, 3)
    n00 = (np.arange(ny)[:, None] * (nx + 1) + np.arange(nx)[None, :])
    main = (first[..., 0] == n00) & (first[..., 2] == n00 + nx + 2)  # diagonal (i,j)-(i+1,j+1)
    assert np.array_equal(main, main[::-1])
    assert np.all(main[:, 1:] != main[:, :-1])
    crack = m.nodes[m.node_sets["crack"]]
    assert np.all(crack[:, 1] == 3.0) and crack[:, 0].max() == 4.0
    with pytest.raises(ValueError):
        FEMMesh.sent_split_quad(8.0, 5.0, 8, 5, a=4.0, device="cpu")  # H/2 off the grid


def test_gamma_factor_and_element_sizes():
    """Per-element h (T3: 4A / perimeter, the in-circle diameter; Q4: the shortest edge) and the
    gamma-convergence factor 1 / (1 + h / (c_w l0)) with c_w = 2 (AT2), 8/3 (AT1)."""
    from phast_amd.q4 import FEMMesh as Q4Mesh, gamma_factor, pf_c_alpha
    m = FEMMesh(np.array([[0.0, 0.0], [3.0, 0.0], [0.0, 4.0]]), np.array([[0, 1, 2]]), device="cpu")
    assert m.elem_h_np[0] == pytest.approx(4.0 * 6.0 / 12.0, rel=1e-15)  # in-circle diameter 2r = 2
    q = Q4Mesh.rectangle(4.0, 1.0, 2, 1, device="cpu")  # cells 2 x 1
    np.testing.assert_array_equal(q.elem_h_np, [1.0, 1.0])
    assert pf_c_alpha("AT2") == 2.0 and pf_c_alpha("AT1") == pytest.approx(8.0 / 3.0)
    np.testing.assert_allclose(gamma_factor(np.array([0.0, 1.0, 2.0]), 0.5, "AT2"), [1.0, 0.5, 1.0 / 3.0], rtol=1e-15)
    np.testing.assert_allclose(gamma_factor(np.array([4.0 / 3.0]), 0.5, "AT1"), [0.5], rtol=1e-15)
