"""Smeared-laminate interface material and its crack-closure rows (host set-up, CPU).

The laminate stiffness is checked against an independent solve of the layer problem, the
frames and rotations against tensor algebra, and the band closure rows and moduli against
the closed forms of an isotropic band and the energy bounds that make the closure sound."""
import math

import numpy as np
import pytest

from phast_amd import solid3d as S

VOIGT = ((0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1))


def tensor(v, shear):
    T = np.zeros((3, 3))
    for k, (i, j) in enumerate(VOIGT):
        T[i, j] = T[j, i] = v[k] * (1.0 if i == j else shear)
    return T


def eng(T):
    return np.array([T[i, j] * (1.0 if i == j else 2.0) for i, j in VOIGT])


LAYERS = ((0.10397175686880507, 20000.0, 0.2), (0.44801412156559745, 270000.0, 0.2),
          (0.44801412156559745, 345000.0, 0.2))


def test_identical_layers_give_the_isotropic_stiffness():
    C = S.laminate_stiffness([(0.3, 200.0, 0.3), (0.7, 200.0, 0.3)])
    np.testing.assert_allclose(C, S.isotropic_stiffness_3d(200.0, 0.3), rtol=0, atol=1e-12)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_laminate_equals_the_solved_layer_problem(seed):
    """Layers stacked along axis 1 under a macro strain E: the in-plane strains (yy, zz, yz)
    equal E's in every layer, the normal stresses (xx, xz, xy) are equal across layers and the
    layer normal strains average to E's.  Solve that for the layer strains; the mean layer
    stress must equal C E."""
    rng = np.random.default_rng(seed)
    layers = LAYERS if seed == 0 else tuple((f, rng.uniform(1e3, 5e5), rng.uniform(0.0, 0.45))
                                            for f in (0.2, 0.5, 0.3))
    C = S.laminate_stiffness(layers)
    E = rng.standard_normal(6)
    n, p = [0, 4, 5], [1, 2, 3]
    L = len(layers)
    Cl = [S.isotropic_stiffness_3d(Ey, nu) for _, Ey, nu in layers]
    f = [fr for fr, _, _ in layers]
    # unknowns: normal strains of each layer (3 L); equations: stresses equal (3 (L-1)), mean (3)
    A = np.zeros((3 * L, 3 * L))
    b = np.zeros(3 * L)
    for l in range(L - 1):
        for r, i in enumerate(n):
            row = 3 * l + r
            A[row, 3 * l:3 * l + 3] = Cl[l][i, n]
            A[row, 3 * (l + 1):3 * (l + 1) + 3] = -Cl[l + 1][i, n]
            b[row] = -(Cl[l][i, p] - Cl[l + 1][i, p]) @ E[p]
    for r in range(3):
        row = 3 * (L - 1) + r
        for l in range(L):
            A[row, 3 * l + r] = f[l]
        b[row] = E[n[r]]
    en = np.linalg.solve(A, b)
    mean = np.zeros(6)
    for l in range(L):
        eps = np.empty(6)
        eps[n] = en[3 * l:3 * l + 3]
        eps[p] = E[p]
        mean += f[l] * (Cl[l] @ eps)
    np.testing.assert_allclose(C @ E, mean, rtol=1e-11, atol=1e-9 * np.abs(mean).max())


def test_frames_and_rotation():
    rng = np.random.default_rng(3)
    t = rng.standard_normal((50, 3))
    t[0] = 0.0
    t[1] = (0.0, 0.0, 2.0)
    F = S.frame_from_tangent(t)
    np.testing.assert_allclose(np.einsum("nji,njk->nik", F, F), np.broadcast_to(np.eye(3), (50, 3, 3)), atol=1e-14)
    np.testing.assert_allclose(np.linalg.det(F), 1.0, atol=1e-14)
    np.testing.assert_allclose(F[2:, :, 0], t[2:] / np.linalg.norm(t[2:], axis=1, keepdims=True), atol=1e-15)
    np.testing.assert_array_equal(F[0, :, 0], [1.0, 0.0, 0.0])
    np.testing.assert_allclose(F[1, :, 0], [0.0, 0.0, 1.0])
    np.testing.assert_allclose(F[1, :, 1], np.cross([1.0, 0.0, 0.0], [0.0, 0.0, 1.0]))  # e_x is least aligned
    C_loc = S.laminate_stiffness(LAYERS)
    Cg = S.rotate_stiffness(C_loc, F)
    for k in (2, 7, 31):
        R = F[k]
        eg = rng.standard_normal(6)
        el = eng(R.T @ tensor(eg, 0.5) @ R)                 # local strain of the global one
        sl = C_loc @ el
        sg = R @ tensor(sl, 1.0) @ R.T                       # local stress back to global axes
        np.testing.assert_allclose(Cg[k] @ eg, [sg[i, j] for i, j in VOIGT], rtol=1e-12,
                                   atol=1e-10 * np.abs(sl).max())
    np.testing.assert_allclose(S.rotate_stiffness(S.isotropic_stiffness_3d(10.0, 0.25), F),
                               np.broadcast_to(S.isotropic_stiffness_3d(10.0, 0.25), (50, 6, 6)), atol=1e-13)


def test_band_rows_of_an_isotropic_band_have_closed_forms():
    E, nu = 345e3, 0.2
    lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    t = np.random.default_rng(4).standard_normal((20, 3))
    C = np.broadcast_to(S.isotropic_stiffness_3d(E, nu), (20, 6, 6)).copy()
    rows, k = S.band_closure_rows(C, t)
    np.testing.assert_allclose(k[:, 0], 1.0 / (lam + 2 * mu), rtol=1e-14)
    np.testing.assert_allclose(k[:, 1], E * nu / (1 - nu * nu), rtol=1e-11)   # plane-stress lam
    np.testing.assert_allclose(k[:, 2:], 2 * mu, rtol=1e-11)
    F = S.frame_from_tangent(t)
    eps = np.random.default_rng(5).standard_normal((20, 6))
    for m in range(20):
        n, e2, e3 = F[m][:, 0], F[m][:, 1], F[m][:, 2]
        Te = tensor(eps[m], 0.5)
        sig = tensor(C[m] @ eps[m], 1.0)
        np.testing.assert_allclose(rows[m, 0] @ eps[m], n @ sig @ n, rtol=1e-12, atol=1e-9)  # t11
        np.testing.assert_allclose(rows[m, 2] @ eps[m], e2 @ Te @ e2, rtol=1e-12, atol=1e-14)  # p2
        np.testing.assert_allclose(rows[m, 3] @ eps[m], e3 @ Te @ e3, rtol=1e-12, atol=1e-14)  # p3


def test_band_closure_energy_is_bounded_and_its_stress_is_its_gradient():
    """psi_c = sum_k k_k/2 <v_k . eps>-^2 lies in [0, psi0] for a laminate band (the closure
    never stiffens), with every bit active C - D is positive semidefinite, and the closure
    stress sum_k k_k <v_k . eps>- v_k is the gradient of psi_c."""
    rng = np.random.default_rng(6)
    t = rng.standard_normal((40, 3))
    C = S.rotate_stiffness(S.laminate_stiffness(LAYERS), S.frame_from_tangent(t))
    rows, k = S.band_closure_rows(C, t)
    assert float(S.closure_stiffening_margin(C, rows, k).min()) >= -1e-12
    for m in range(40):
        for _ in range(5):
            eps = rng.standard_normal(6) * 1e-3
            v = rows[m] @ eps
            psi_c = 0.5 * float((k[m] * np.minimum(v, 0.0) ** 2).sum())
            psi0 = 0.5 * float(eps @ C[m] @ eps)
            assert -1e-18 <= psi_c <= psi0 * (1 + 1e-12) + 1e-18
            sig = (k[m] * np.minimum(v, 0.0)) @ rows[m]
            h = 1e-7
            fd = np.empty(6)
            for i in range(6):
                ep, em = eps.copy(), eps.copy()
                ep[i] += h
                em[i] -= h
                vp, vm = rows[m] @ ep, rows[m] @ em
                fd[i] = (0.5 * (k[m] * np.minimum(vp, 0) ** 2).sum() - 0.5 * (k[m] * np.minimum(vm, 0) ** 2).sum()) / (2 * h)
            np.testing.assert_allclose(sig, fd, rtol=1e-6, atol=1e-6 * np.abs(C[m]).max() * 1e-3)


def test_element_arrays_with_a_band():
    from phast_amd import geometry as G
    shape, h = (12, 12, 12), (0.5e-3,) * 3
    phase, tangent, centres = G.particles_with_shell(shape, h, 2, 1.5e-3, 0.8e-3, 0.5e-3, seed=3)
    assert {0, 1, 3} <= set(np.unique(phase).tolist())
    phases = {"mi_matrix": S.PhaseMaterial("mi_matrix", 345e3, 0.2, 25e-3, tensile_strength=600.0),
              "fibre": S.PhaseMaterial("fibre", 270e3, 0.2, 0.518, tensile_strength=3136.0),
              "bn_band": S.PhaseMaterial("bn_band", 122601.74, 0.2, 6.46e-4, tensile_strength=63.0,
                                         kind="smeared_laminate", layers=LAYERS)}
    mats = S.build_element_arrays(phase, phases, tangent=tangent)
    band = phase.reshape(-1) == G.INTERFACE
    assert mats.laminate.sum() == band.sum() and not mats.all_isotropic
    assert mats.stiffness_bank.shape[0] == 2 + band.sum()
    # iso slots in phase-id order: fibre 0, matrix 1; band voxels own entries in voxel order
    assert set(mats.stiffness_index[~band].tolist()) == {0, 1}
    np.testing.assert_array_equal(mats.stiffness_index[band], 2 + np.arange(band.sum()))
    ids = np.nonzero(band)[0]
    Cl = S.laminate_stiffness(LAYERS)
    R = S.frame_from_tangent(tangent.reshape(-1, 3)[ids])
    np.testing.assert_allclose(mats.stiffness_bank[2:], S.rotate_stiffness(Cl, R), rtol=0, atol=1e-9)
    with pytest.raises(ValueError, match="tangent"):
        S.build_element_arrays(phase, phases)
    with pytest.raises(ValueError, match="layers"):
        S.PhaseMaterial("bn_band", 1.0, 0.2, 1.0, kind="smeared_laminate")


def test_particle_placement_is_reproducible_and_separated():
    from phast_amd import geometry as G
    c1 = G.place_particles((24e-3,) * 3, 4, 5.96e-3, 1e-3, 20261023)
    c2 = G.place_particles((24e-3,) * 3, 4, 5.96e-3, 1e-3, 20261023)
    np.testing.assert_array_equal(c1, c2)
    box = np.full(3, 24e-3)
    for a in range(4):
        for b in range(a):
            g = np.abs(c1[a] - c1[b])
            assert math.sqrt(float((np.minimum(g, box - g) ** 2).sum())) >= 2 * 5.96e-3 + 1e-3
