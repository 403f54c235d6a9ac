"""Synthetic inputs of the five benchmark workloads, built on the host with seeded numpy.

Every builder returns plain arrays and constants (meshes, phase arrays, loads, fields), so
the same inputs feed the HIP path and any CPU checker.  Device objects are made by the
caller.  Units are N, mm, s and MPa, as in the reference's configuration files.

* ``sent_plate``: 200 x 80 native-Q4 single-edge-notched plate (100 x 40 mm), the notch
  x in [0, 50] at y = 20 preseeded through the history of its one-ring, +-1 MPa traction on
  the top and bottom edges, dt = 0.05 us, 1000 explicit steps.
* ``pore_cube``: periodic voxel cube with one centred spherical pore (64^3, r = 12 voxels).
* ``particle_composite``: 128^3 periodic cube of non-overlapping spheres, radii U[6, 14]
  voxels up to 30 % volume fraction, particles E = 70 GPa in a 3 GPa matrix, a one-voxel
  interface shell with Gc x 0.1.
* ``specimen_batch``: 64 notched plates of uneven sizes with lognormal Gc fields.
* ``porous_gc_cube``: 96^3 cube with 20 seeded pores and a per-voxel Gc field.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from types import SimpleNamespace

import numpy as np

# Q4 plate material (MPa, t/mm^3, N/mm, mm): AT2, Amor split, plane strain
Q4_MATERIAL = dict(E=32000.0, nu=0.2, rho=2.45e-9, Gc=3e-3, l0=0.5)
VOXEL_MM = 1e-3          # 1 um voxels
MATRIX_GC = 5e-4         # 0.5 J/m^2 in N/mm


def structured_plate(W: float, H: float, nx: int, ny: int):
    """Nodes ``j (nx+1) + i`` and cells ``j nx + i`` with corners (i,j), (i+1,j), (i+1,j+1),
    (i,j+1), as ``core/mesh_generator.py:1695-1818`` numbers the structured SENT plate."""
    x = np.linspace(0.0, W, nx + 1)
    y = np.linspace(0.0, H, ny + 1)
    xx, yy = np.meshgrid(x, y, indexing="xy")
    nodes = np.column_stack([xx.ravel(), yy.ravel()])
    je, ie = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    n0 = (je * (nx + 1) + ie).ravel()
    cells = np.stack([n0, n0 + 1, n0 + nx + 2, n0 + nx + 1], axis=1)
    return nodes, cells


@dataclass
class SentPlate:
    """Inputs of the explicit SENT run.  ``H0`` is the preseeded Gauss-point history (E,4);
    the run starts from the damage solved at ``H0`` with ``d_prev = 0``."""
    nodes: np.ndarray
    elements: np.ndarray
    nx: int
    ny: int
    crack: np.ndarray
    top: np.ndarray
    bottom: np.ndarray
    f_ext: np.ndarray
    H0: np.ndarray
    dt: float
    steps: int
    material: dict = field(default_factory=lambda: dict(Q4_MATERIAL))


def sent_plate(W: float = 100.0, H: float = 40.0, a: float = 50.0, h: float = 0.5, traction: float = 1.0,
               dt: float = 5e-8, steps: int = 1000) -> SentPlate:
    """The ``crack`` node set is ``0 <= x <= a`` on ``y = H/2`` (mesh_generator.py:1776-1783).
    The history of every cell touching it is raised to ``1e4 Gc/l0`` at its four Gauss
    points (run_config.py:440-470, initial_conditions.py:206-226).  The traction is
    assembled on all four edges of the Q4 cells (each boundary edge gives half its length
    times the traction to both end nodes)."""
    from phast_amd import loading as L
    import torch
    nx, ny = int(round(W / h)), int(round(H / h))
    nodes, cells = structured_plate(W, H, nx, ny)
    j_mid, i_crack = int(round(ny / 2)), int(round(a / h))
    crack = j_mid * (nx + 1) + np.arange(i_crack + 1)
    top = ny * (nx + 1) + np.arange(nx + 1)
    bottom = np.arange(nx + 1)
    mesh = SimpleNamespace(nodes=nodes, elements=cells, element_type="Q4",
                           node_sets={"crack": crack, "top": top, "bottom": bottom})
    f = (L.neumann_forces(mesh, top, [0.0, traction], device="cpu")
         + L.neumann_forces(mesh, bottom, [0.0, -traction], device="cpu")).numpy()
    H0 = torch.zeros((cells.shape[0], 4), dtype=torch.float64)
    L.preseed_history(mesh, H0, [{"nodes": "crack", "value": 1.0}], Q4_MATERIAL["Gc"], Q4_MATERIAL["l0"])
    return SentPlate(nodes, cells, nx, ny, crack, top, bottom, f, H0.numpy(), dt, steps)


# ----------------------------------------------------------------------------------------
# voxel cells
# ----------------------------------------------------------------------------------------

def _phase_materials(**extra):
    from phast_amd import solid3d as S
    out = {"mi_matrix": S.PhaseMaterial("mi_matrix", 3000.0, 0.35, MATRIX_GC),
           "pore": S.PhaseMaterial("pore", 3000.0 * 1e-6, 0.35, MATRIX_GC)}
    out.update(extra)
    return out


def pore_cube(n: int = 64, r_vox: float = 12.0):
    """Phase array (n,n,n) and phase materials of the centred-pore cube: matrix E = 3 GPa,
    nu = 0.35, Gc = 0.5 J/m^2; pore stiffness E x 1e-6."""
    from phast_amd import geometry as G
    h = (VOXEL_MM,) * 3
    phase = G.sphere_pore_radius((n, n, n), h, r_vox * VOXEL_MM)
    return phase, _phase_materials(), h


def particle_composite(n: int = 128, seed: int = 128, r_range=(6.0, 14.0), fraction: float = 0.30):
    """Phase array, phase materials and placed spheres of the particle composite.
    Particles: E = 70 GPa, nu = 0.3; matrix: E = 3 GPa, nu = 0.3, Gc = 0.5 J/m^2; the
    one-voxel interface shell has the matrix stiffness and Gc x 0.1.  The particles keep the
    matrix Gc (the workload names no other value)."""
    from phast_amd import geometry as G
    from phast_amd import solid3d as S
    h = (VOXEL_MM,) * 3
    phase, centres, radii = G.polydisperse_particles((n, n, n), h, r_range, fraction, seed)
    phases = {"mi_matrix": S.PhaseMaterial("mi_matrix", 3000.0, 0.3, MATRIX_GC),
              "fibre": S.PhaseMaterial("fibre", 70000.0, 0.3, MATRIX_GC),
              "bn_band": S.PhaseMaterial("bn_band", 3000.0, 0.3, 0.1 * MATRIX_GC)}
    return phase, phases, h, centres, radii


def porous_gc_cube(n: int = 96, n_pores: int = 20, r_range=(4.0, 9.0), seed: int = 96, gc_seed: int = 4):
    """Phase array, phase materials and the per-voxel Gc field ``5e-4 (1 + 0.1 N(0,1))``
    (floored at 1e-5) of the porous cube; also the target field, a 5 % seeded perturbation."""
    from phast_amd import geometry as G
    h = (VOXEL_MM,) * 3
    phase = G.random_pores((n, n, n), h, n_pores, r_range, seed=seed)
    rng = np.random.default_rng(gc_seed)
    E = n ** 3
    Gc = np.maximum(MATRIX_GC * (1 + 0.1 * rng.standard_normal(E)), 1e-5)
    Gc_target = Gc * (1 + 0.05 * rng.standard_normal(E))
    return phase, _phase_materials(), h, Gc, Gc_target


# ----------------------------------------------------------------------------------------
# batched plates
# ----------------------------------------------------------------------------------------

def pore_in_cube_pfczm(n: int = 48, increments: int = 150, to: float = 3.0e-3):
    """The reference's PF-CZM pore example (examples/three_dimensional/pore_in_cube/run.py:19-27):
    a periodic cube of n^3 voxels of 0.5 um with a centred spherical pore of diameter 0.4 L,
    matrix E = 345 GPa, nu = 0.2, tensile strength 600 MPa, Gc = 25 J/m^2; pore stiffness
    E x 1e-6 (solid3d/api.py:34, 167-172); PF-CZM (linear softening, p = 2) with ell = 1.5 um;
    E_xx raised to ``to`` in ``increments`` equal steps, the other macro stresses zero.
    Returns (phase (n,n,n) uint8, phases, h, ell, load case)."""
    from phast_amd import geometry as G
    from phast_amd import solid3d as S
    h = (0.5e-3,) * 3
    phase = G.sphere_pore((n, n, n), h, 0.4)
    E, nu, f_t, Gc = 345e3, 0.2, 600.0, 25e-3
    phases = {"mi_matrix": S.PhaseMaterial("mi_matrix", E, nu, Gc, tensile_strength=f_t),
              "pore": S.PhaseMaterial("pore", E * 1.0e-6, nu, Gc, tensile_strength=f_t)}
    load = S.LoadCase.from_schedule("macro_xx", controlled=0, schedule=[(to, to / increments)], free=(1, 2, 3, 4, 5),
                                    fixed_zero=())
    return phase, phases, h, 1.5e-3, load


INTERFACE_LAYERS_PARTICLES = ((0.10397175686880507, 20000.0, 0.2), (0.44801412156559745, 270000.0, 0.2),
                              (0.44801412156559745, 345000.0, 0.2))


def particles_in_cube_pfczm(n: int = 48, increments: int = 150, to: float = 3.0e-3):
    """The reference's PF-CZM particle example (examples/three_dimensional/particles_in_cube/
    run.py:19-32): a periodic cube of n^3 voxels of 0.5 um with four spherical particles of
    radius 5 um placed by random sequential addition (seed 20261023, gap 1 um), each wrapped
    in a 1.92 um interface shell, a smeared laminate of (fraction, E, nu) layers across the
    shell (its isotropic equivalent E = 122.6 GPa, nu = 0.2 drives the damage); matrix
    E = 345 GPa, f_t = 600 MPa, Gc = 25 J/m^2; particles E = 270 GPa, f_t = 3136 MPa,
    Gc = 518 J/m^2; interface f_t = 63 MPa, Gc = 0.646 J/m^2; PF-CZM with ell = 1.5 um;
    E_xx raised to ``to`` in ``increments`` equal steps, the other macro stresses zero.
    Returns (phase (n,n,n) uint8, tangent (n,n,n,3), phases, h, ell, load case)."""
    from phast_amd import geometry as G
    from phast_amd import solid3d as S
    h = (0.5e-3,) * 3
    phase, tangent, _ = G.particles_with_shell((n, n, n), h, 4, 5.0e-3, 1.9235993121898142e-3, 1.0e-3, 20261023)
    phases = {"mi_matrix": S.PhaseMaterial("mi_matrix", 345e3, 0.2, 25e-3, tensile_strength=600.0),
              "fibre": S.PhaseMaterial("fibre", 270e3, 0.2, 0.5180310650205762, tensile_strength=3136.0),
              "bn_band": S.PhaseMaterial("bn_band", 122601.74043706217, 0.2, 0.00064633998586781,
                                         tensile_strength=63.0, kind="smeared_laminate",
                                         layers=INTERFACE_LAYERS_PARTICLES)}
    load = S.LoadCase.from_schedule("macro_xx", controlled=0, schedule=[(to, to / increments)], free=(1, 2, 3, 4, 5),
                                    fixed_zero=())
    return phase, tangent, phases, h, 1.5e-3, load


INTERFACE_LAYERS_FIBRES = ((0.10823084088511756, 20000.0, 0.2), (0.4458845795574412, 270000.0, 0.2),
                           (0.4458845795574412, 345000.0, 0.2))


def fibre_unit_cell_pfczm(increments: int = 150, to: float = 3.0e-3):
    """Materials and load of the reference's nine-fibre unit cell (examples/three_dimensional/
    fibre_unit_cell/run.py:19-32), whose microstructure (phase ids and interface normals of
    103 x 103 x 2 voxels of 0.5 um) is an imported image: matrix E = 345 GPa, f_t = 600 MPa,
    Gc = 25 J/m^2; fibres E = 270 GPa, f_t = 3136 MPa, Gc = 518 J/m^2; a smeared-laminate
    interface band (isotropic equivalent E = 119.7 GPa, nu = 0.2, f_t = 63 MPa,
    Gc = 0.646 J/m^2); PF-CZM with ell = 1.5 um; E_xx raised to ``to`` in ``increments``
    equal steps, the other macro stresses zero.  Returns (phases, h, ell, load case)."""
    from phast_amd import solid3d as S
    phases = {"mi_matrix": S.PhaseMaterial("mi_matrix", 345e3, 0.2, 25e-3, tensile_strength=600.0),
              "fibre": S.PhaseMaterial("fibre", 270e3, 0.2, 0.5180310650205762, tensile_strength=3136.0),
              "bn_band": S.PhaseMaterial("bn_band", 119683.28713197137, 0.2, 0.0006463399858678095,
                                         tensile_strength=63.0, kind="smeared_laminate",
                                         layers=INTERFACE_LAYERS_FIBRES)}
    load = S.LoadCase.from_schedule("macro_xx", controlled=0, schedule=[(to, to / increments)], free=(1, 2, 3, 4, 5),
                                    fixed_zero=())
    return phases, (0.5e-3,) * 3, 1.5e-3, load


def lognormal_field(rng, ny, nx, mean, cov, corr):
    """Lognormal per-cell field of the given mean and coefficient of variation whose log is a
    Gaussian random field with a Gaussian correlation of ``corr`` cells (white noise filtered
    in Fourier space)."""
    z = rng.standard_normal((ny, nx))
    ky = np.fft.fftfreq(ny)[:, None]
    kx = np.fft.fftfreq(nx)[None, :]
    z = np.real(np.fft.ifft2(np.fft.fft2(z) * np.exp(-2.0 * np.pi ** 2 * corr ** 2 * (kx ** 2 + ky ** 2))))
    z = (z - z.mean()) / max(z.std(), 1e-300)
    s2 = np.log1p(cov ** 2)
    return mean * np.exp(np.sqrt(s2) * z - 0.5 * s2)


@dataclass
class Specimen:
    nx: int
    ny: int
    h: float
    Gc_field: np.ndarray     # (E,)
    f_ext: np.ndarray        # (N,2)
    notch: np.ndarray        # node ids held at d = 1


def specimen_batch(n_spec: int = 64, seed: int = 2024, nx_range=(64, 512), h: float = 0.5):
    """Notched plates: nx ~ U{64..512}, ny = nx/2, 0.5 mm cells, lognormal Gc (mean 3e-3,
    CoV 0.2, correlation 4 cells), a notch of length U[0.2, 0.5] W held at d = 1 along the
    mid-height node row, traction +-1 MPa on the top/bottom edges."""
    rng = np.random.default_rng(seed)
    out = []
    for nx in rng.integers(nx_range[0], nx_range[1] + 1, n_spec):
        nx = int(nx)
        ny = nx // 2
        Gc = lognormal_field(rng, ny, nx, Q4_MATERIAL["Gc"], 0.2, 4.0).ravel()
        f = np.zeros(((nx + 1) * (ny + 1), 2))
        edge = np.full(nx + 1, h)
        edge[[0, -1]] = 0.5 * h
        f[ny * (nx + 1):, 1] = 1.0 * edge
        f[:nx + 1, 1] = -1.0 * edge
        a_len = rng.uniform(0.2, 0.5) * h * nx
        ids = (ny // 2) * (nx + 1) + np.arange(int(np.floor(a_len / h)) + 1)
        out.append(Specimen(nx, ny, h, Gc, f, ids))
    return out
