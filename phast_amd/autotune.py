"""Launch-shape autotuner: times the candidate tile shapes of an apply on the live device
and keeps the fastest.  Results are cached per (kind, problem size) for the process."""
from __future__ import annotations

import torch

# (16,16), (32,8), (16,14), (32,14), (32,6) and (64,6): voxel-parallel kernel with z slabs of kz layers; (16,8) and (8,8):
# node-column marching kernel walking kz node layers.
HEX8_TILES = [(16, 16, 8), (16, 16, 16), (32, 8, 2), (32, 8, 3), (32, 8, 4), (32, 8, 8), (32, 8, 16), (16, 14, 2),
              (16, 14, 3), (16, 14, 4), (16, 14, 6), (16, 14, 8), (16, 14, 12), (16, 14, 16), (32, 14, 2), (32, 14, 3),
              (32, 14, 4), (32, 14, 6), (32, 14, 8), (32, 14, 16), (16, 8, 8), (32, 6, 4), (32, 6, 8), (32, 6, 16),
              (64, 6, 4), (64, 6, 8), (64, 6, 16)]
Q4_TILES = [(128, 8), (128, 32), (256, 8), (256, 16), (256, 32), (256, 64), (64, 32), (256, 128)]
BATCH_TILES = [(64, 4), (64, 8), (128, 2), (128, 4), (128, 8), (256, 2), (256, 4), (256, 8), (256, 16)]

_cache: dict = {}


def _time(fn, reps: int) -> float:
    fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


def tune_hex8(op, x: torch.Tensor, d: torch.Tensor, reps: int = 10, candidates=HEX8_TILES):
    """Pick the launch shape (tx, ty, kz) of ``op`` for the fused apply."""
    key = ("hex8", op.grid.shape)
    if key in _cache:
        op.set_tiles(*_cache[key][0])
        return _cache[key]
    timings = {}
    for t in candidates:
        if t[2] > op.grid.shape[2]:
            continue
        op.set_tiles(*t)
        timings[t] = _time(lambda: op.mech_action_d(x, d), reps)
    best = min(timings, key=timings.get)
    op.set_tiles(*best)
    _cache[key] = (best, timings)
    return best, timings


def tune_q4(make_fem, u: torch.Tensor, d: torch.Tensor, reps: int = 10, candidates=Q4_TILES):
    """Pick the Q4 strip (threads, rows); ``make_fem(tile)`` builds an FEMOperators with that tile."""
    timings, fems = {}, {}
    for t in candidates:
        fem = make_fem(t)
        fems[t] = fem
        timings[t] = _time(lambda: fem.internal_force(u, d), reps)
    best = min(timings, key=timings.get)
    return best, timings, fems[best]


def tune_batch(make_batch, state=None, steps: int = 10, candidates=BATCH_TILES):
    """Pick the strip shape of the persistent batched step: ``make_batch(tile)`` builds a
    BatchedExplicit with that strip; each candidate runs ``steps`` persistent steps from the
    packed ``state`` (or its initial state) and the fastest per step wins."""
    import time
    timings = {}
    for t in candidates:
        b = make_batch(t)
        if state is not None:
            b.set_packed_state(state)
        b.prepare_persistent()
        b.step_persistent(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.step_persistent(steps)
        torch.cuda.synchronize()
        timings[t] = (time.perf_counter() - t0) / steps * 1e3
        b.check_persistent()
        del b
    best = min(timings, key=timings.get)
    return best, timings
