"""Voxel microstructure builders (inputs of the voxel path), after solid3d/geometry.py.

Phase ids follow solid3d/model.py:40-49: fibre 0 (inclusion), bn_band 1 (interface),
cvi_sic 2, mi_matrix 3 (matrix), pore 4.  Arrays are (nz, ny, nx) uint8, so that the
flattened index is the voxel index i + nx (j + ny k).
"""
from __future__ import annotations

import numpy as np

INCLUSION, INTERFACE, MATRIX, PORE = 0, 1, 3, 4


def voxel_centres(shape, h):
    """Voxel centre coordinates along x, y, z (geometry.py:12-16)."""
    return tuple((np.arange(n) + 0.5) * hh for n, hh in zip(shape, h))


def sphere_pore(shape, h, diameter_fraction: float) -> np.ndarray:
    """Matrix with one centred spherical pore of diameter ``fraction * Lx`` (geometry.py:19-27)."""
    nx, ny, nz = shape
    lx, ly, lz = nx * h[0], ny * h[1], nz * h[2]
    radius = 0.5 * float(diameter_fraction) * lx
    x, y, z = voxel_centres(shape, h)
    d = np.sqrt((x[None, None, :] - lx / 2) ** 2 + (y[None, :, None] - ly / 2) ** 2 + (z[:, None, None] - lz / 2) ** 2)
    phase = np.full((nz, ny, nx), MATRIX, dtype=np.uint8)
    phase[d < radius] = PORE
    return phase


def sphere_pore_radius(shape, h, radius: float, centre=None) -> np.ndarray:
    """Matrix with one spherical pore of the given radius (in length units)."""
    nx, ny, nz = shape
    lx, ly, lz = nx * h[0], ny * h[1], nz * h[2]
    cx, cy, cz = (lx / 2, ly / 2, lz / 2) if centre is None else centre
    x, y, z = voxel_centres(shape, h)
    d = np.sqrt((x[None, None, :] - cx) ** 2 + (y[None, :, None] - cy) ** 2 + (z[:, None, None] - cz) ** 2)
    phase = np.full((nz, ny, nx), MATRIX, dtype=np.uint8)
    phase[d < radius] = PORE
    return phase


def _min_image(dx, L):
    return dx - L * np.round(dx / L)


def random_pores(shape, h, n: int, r_range, seed: int, max_tries: int = 200000) -> np.ndarray:
    """``n`` non-overlapping periodic spherical pores, radii uniform in ``r_range`` (voxels)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    L = np.array([nx * h[0], ny * h[1], nz * h[2]])
    centres, radii = [], []
    for _ in range(max_tries):
        r = rng.uniform(*r_range) * h[0]
        c = rng.random(3) * L
        if all(np.sqrt((_min_image(c - o, L) ** 2).sum()) >= r + ro for o, ro in zip(centres, radii)):
            centres.append(c)
            radii.append(r)
            if len(centres) == n:
                break
    x, y, z = voxel_centres(shape, h)
    phase = np.full((nz, ny, nx), MATRIX, dtype=np.uint8)
    for c, r in zip(centres, radii):
        dx = _min_image(x - c[0], L[0])
        dy = _min_image(y - c[1], L[1])
        dz = _min_image(z - c[2], L[2])
        d = np.sqrt(dx[None, None, :] ** 2 + dy[None, :, None] ** 2 + dz[:, None, None] ** 2)
        phase[d < r] = PORE
    return phase


def polydisperse_particles(shape, h, r_range, max_fraction: float, seed: int, shell_voxels: float = 1.0,
                           max_tries: int = 200000):
    """Random sequential addition of non-overlapping periodic spheres with radii uniform in
    ``r_range`` (voxels) until ``max_fraction`` of the volume is reached (the polydisperse
    variant of geometry.py:30-42), with a one-voxel interface shell on each surface
    (the shell/core split of geometry.py:67-81)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    L = np.array([nx * h[0], ny * h[1], nz * h[2]])
    shell = shell_voxels * h[0]
    centres, radii, vol = [], [], 0.0
    for _ in range(max_tries):
        r = rng.uniform(*r_range) * h[0]
        v = 4.0 / 3.0 * np.pi * r ** 3
        if vol + v > max_fraction * L.prod():
            break
        c = rng.random(3) * L
        if all(np.sqrt((_min_image(c - o, L) ** 2).sum()) >= r + ro + shell for o, ro in zip(centres, radii)):
            centres.append(c)
            radii.append(r)
            vol += v
    x, y, z = voxel_centres(shape, h)
    phase = np.full((nz, ny, nx), MATRIX, dtype=np.uint8)
    for c, r in zip(centres, radii):
        dx = _min_image(x - c[0], L[0])
        dy = _min_image(y - c[1], L[1])
        dz = _min_image(z - c[2], L[2])
        d = np.sqrt(dx[None, None, :] ** 2 + dy[None, :, None] ** 2 + dz[:, None, None] ** 2)
        phase[(d < r + 0.5 * shell) & (phase == MATRIX)] = INTERFACE
        phase[d < r - 0.5 * shell] = INCLUSION
    return phase, np.array(centres), np.array(radii)


def place_particles(lengths, n: int, radius: float, min_gap: float, seed: int, max_tries: int = 200000):
    """Random sequential addition of ``n`` spheres of ``radius`` in a periodic box: draws
    ``default_rng(seed).random(3) * lengths`` and keeps a centre whose minimum-image distance
    to every kept centre is at least ``2 radius + min_gap`` (geometry.py:30-42).  Returns the
    (n, 3) centres."""
    rng = np.random.default_rng(seed)
    box = np.asarray(lengths, dtype=np.float64)
    d_min = 2.0 * radius + min_gap
    kept = []
    for _ in range(max_tries):
        c = rng.random(3) * box
        ok = True
        for o in kept:
            gap = np.abs(c - o)
            if np.sqrt((np.minimum(gap, box - gap) ** 2).sum()) < d_min:
                ok = False
                break
        if ok:
            kept.append(c)
            if len(kept) == n:
                return np.array(kept)
    raise RuntimeError(f"could not place {n} non-overlapping particles (placed {len(kept)}); lower n or the radius")


def nearest_centre(shape, h, centres):
    """Minimum-image distance from every voxel centre to the nearest of ``centres`` and the
    unit vector pointing away from it (0 at a centre), (nz, ny, nx) and (nz, ny, nx, 3)
    (geometry.py:45-63; the first centre wins a tie)."""
    nx, ny, nz = shape
    L = (nx * h[0], ny * h[1], nz * h[2])
    x, y, z = voxel_centres(shape, h)
    best = np.full((nz, ny, nx), np.inf)
    vec = np.zeros((nz, ny, nx, 3))
    for c in centres:
        comps = []
        for v, cc, LL, sl in ((x, c[0], L[0], (None, None, slice(None))), (y, c[1], L[1], (None, slice(None), None)),
                              (z, c[2], L[2], (slice(None), None, None))):
            dv = v - cc
            dv = dv - LL * np.round(dv / LL)
            comps.append(np.broadcast_to(dv[sl], (nz, ny, nx)))
        d = np.sqrt(comps[0] ** 2 + comps[1] ** 2 + comps[2] ** 2)
        closer = d < best
        best = np.where(closer, d, best)
        for k in range(3):
            vec[..., k] = np.where(closer, comps[k], vec[..., k])
    with np.errstate(invalid="ignore", divide="ignore"):
        unit = np.where(best[..., None] > 0, vec / best[..., None], 0.0)
    return best, unit


def particles_with_shell(shape, h, n: int, radius: float, shell: float, gap: float, seed: int):
    """``n`` spherical particles of radius ``radius`` wrapped in an interface shell of
    thickness ``shell`` straddling the surface, placed by :func:`place_particles` with the
    outer radius ``radius + shell / 2`` and the minimum gap ``gap`` (lengths in mm;
    geometry.py:67-81).  Returns (phase (nz, ny, nx) uint8, tangent (nz, ny, nx, 3): the outward
    radial direction on the shell, (0, 0, 1) elsewhere, centres (n, 3))."""
    nx, ny, nz = shape
    centres = place_particles((nx * h[0], ny * h[1], nz * h[2]), n, radius + 0.5 * shell, gap, seed)
    dist, unit = nearest_centre(shape, h, centres)
    phase = np.full((nz, ny, nx), MATRIX, dtype=np.uint8)
    phase[dist < radius + 0.5 * shell] = INTERFACE
    phase[dist < radius - 0.5 * shell] = INCLUSION
    tangent = np.zeros((nz, ny, nx, 3))
    tangent[..., 2] = 1.0
    on = phase == INTERFACE
    tangent[on] = unit[on]
    return phase, tangent, centres
