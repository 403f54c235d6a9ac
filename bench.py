#!/usr/bin/env python3
"""Benchmark of the matrix-free phase-field hot path on one MI355X.

Headline (one JSON line on stdout): element-apply GDOF/s of the periodic hex8 voxel
operator on the 128^3 particle composite (seeded non-overlapping spheres, radii U[6,14]
voxels, 30 % volume fraction, 70 GPa particles in a 3 GPa matrix, a one-voxel interface
shell with Gc x 0.1).  One step = one apply y = A(g(d)) x with the degradation formed
in-kernel from nodal damage, inputs resident in HBM.  GDOF = 3 n + 6 output dofs per apply.
``roofline`` prices the apply's algorithmic FLOPs (940 per voxel, sum-factorised) against
the 78.6 TFLOP/s FP64 peak -- its intensity (940 FLOP / 57 B) is above the FP64 ridge -- and
also reports the algorithmic bytes (57 B per node: u 24 + d 8 + phase 1 + r 24) against
8 TB/s; ``cpu_baseline`` times a torch-CPU restatement of the reference's batched apply on
the host cores.  ``secondary`` holds the Q4 apply and fused explicit step, the T3 force /
explicit step / damage apply on an 8.4 M-node split-quad plate, the other voxel sizes, the
200 x 80 SENT run, the reference's 48^3 PF-CZM pore example against its recorded history,
the 64-specimen batched run and the reaction-curve gradient, measured in the same run.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6         # MI355X FP64 (vector and matrix peaks coincide)
HEX8_BYTES_PER_NODE = 57.0      # u 24 + d 8 + phase 1 + r 24
HEX8_FLOPS_PER_VOXEL = 940.0    # sum-factorised apply
# HBM bytes per 128^3 apply (the apply kernel plus its seam/finalize kernel) measured with
# rocprofv3 (FETCH_SIZE and WRITE_SIZE in KiB per dispatch, scaled by the copy-kernel
# factors), per launch shape; see profiles/hex8_apply_128cube_particles.md
HEX8_128_TRAFFIC = {(16, 14, 8): ((40650.1 + 6200.7) * 2.0 + 55376.0 + 6144.1) * 1024.0}
Q4_APPLY_BYTES_PER_NODE = 40.0  # u 16 + d 8 + f 16
Q4_STEP_BYTES_PER_NODE = 192.0  # u,v,a r+w 96, f_ext 16, Minv 8, d 8, H r+w 64
# one projected damage-CG iteration without a bound hit or residual check (per node): A p with
# the previous direction update fused in (z 8, p_old 8, active 1, H 32, d 8, lb 8; Ap 8 and
# p 8 written), then the d/r update fused with the preconditioner (p, d, r, Ap, Minv, active
# read: 41; d, r, z written: 24); the residual check of every 10th iteration (about 66 B,
# then with the two halves as separate passes) adds 6.6 on average
Q4_CG_BYTES_PER_NODE = 81.0 + 65.0 + 6.6
# T3 split-quad plates (2 triangles per node; geometry recomputed from the coordinates)
T3_APPLY_BYTES_PER_NODE = 80.0  # u 16 + d 8 + f 16 + xy 16 + 2 triangles x 12 B connectivity
T3_STEP_BYTES_PER_NODE = 216.0  # u,v,a r+w 96, f_ext 16, Minv 8, d 8, xy 16, conn 24, H r+w + psi 48
T3_DAMAGE_BYTES_PER_NODE = 72.0  # x 8 + y 8 + H 16 + xy 16 + conn 24
# one projected damage-CG iteration: A p with step ratios (p 8, H 16, xy 16, conn 24, d 8, lb 8,
# Ap 8), d/r update with the preconditioner (41 read + 24 written), direction update (z, p,
# active 17 + p 8), the residual check of every 10th iteration (~90 B) 9 on average
T3_CG_BYTES_PER_NODE = 88.0 + 65.0 + 25.0 + 9.0
# one iteration of the adjoint PCG of the backward: A p (p 8, H 16, xy 16, conn 24; Ap 8
# written), x/r/z update (x, p, r, Ap, Minv read 40; x, r, z written 24), direction (z, p 16; p 8)
T3_ADJ_BYTES_PER_NODE = 72.0 + 64.0 + 24.0
# general quadrilateral meshes (one cell per node): the force reads u 16, d 8, the corner
# coordinates 16 and the connectivity 16 and writes f 16 (72 B/node algorithmic); the element
# pass also writes 64 B of corner values per cell that the node pass reads back with the
# incidence list (20 B), 148 B/node more
Q4G_APPLY_BYTES_PER_NODE = 72.0
Q4G_APPLY_COUNTED_PER_NODE = 72.0 + 148.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-secondary", action="store_true", help="skip the secondary measurements")
    ap.add_argument("--secondary", type=str, default="q4,q4g,hex8,t3,sent",
                    help="secondary parts to run: q4, q4g (a general quadrilateral mesh), hex8, t3, sent (the 200 x 80 "
                    "SENT plate run end to end)")
    ap.add_argument("--q4-tile", type=str, default="", help="fixed Q4 strip threads,rows (skips the autotuner)")
    ap.add_argument("--hex8-big-tiles", type=str, default="", help="fixed tx,ty,kz of the 128^3/256^3 runs")
    ap.add_argument("--batch-tile", type=str, default="", help="fixed strip of the batched run (skips its autotuner)")
    ap.add_argument("--no-sequential", action="store_true", help="skip the 64 single-specimen runs of the batched case")
    ap.add_argument("--tiles", type=str, default="", help="fixed hex8 tile tx,ty,kz (skips the autotuner)")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-specimen batched run")
    ap.add_argument("--no-pfczm", action="store_true", help="skip the 48^3 PF-CZM pore example")
    ap.add_argument("--no-c1", action="store_true", help="skip the 64^3 pore cube's 40-increment load path")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the voxel reaction-curve gradient run")
    ap.add_argument("--adjoint-size", type=int, default=96, help="voxels per edge of the gradient run")
    ap.add_argument("--adjoint-steps", type=int, default=30, help="load steps of the gradient run")
    ap.add_argument("--adjoint-fd", type=int, default=2, help="single-voxel central differences in the gradient run "
                    "on the most sensitive voxels")
    ap.add_argument("--adjoint-fd-random", type=int, default=0, help="also central differences on this many seeded "
                    "random voxels")
    ap.add_argument("--checkpoint-every", type=int, default=5, help="load-step checkpoint spacing of the gradient run")
    ap.add_argument("--sequential-window", type=int, default=0, help="steps of the batched and sequential runs "
                    "timed against each other (0: the whole batched run)")
    ap.add_argument("--batched-steps", type=int, default=2000, help="steps of the 64-specimen batched run")
    ap.add_argument("--batch-steps-per-launch", type=int, default=20, help="explicit steps per persistent launch "
                    "of the batched run (1: one launch per step, every specimen in lockstep)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run 5 device copies of 256 MiB (the reference kernel for HBM counter scaling)")
    return ap.parse_args()


def events_time(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    s.record()
    for _ in range(steps):
        fn()
    e.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return s.elapsed_time(e) / 1e3 / steps, wall


def pore_cube(n, r_vox, device):
    import numpy as np
    from phast_amd import geometry as G
    from phast_amd import solid3d as S
    h = (1e-3,) * 3
    phase = G.sphere_pore_radius((n, n, n), h, r_vox * 1e-3)
    phases = {"mi_matrix": S.PhaseMaterial("mi_matrix", 3000.0, 0.35, 5e-4),
              "pore": S.PhaseMaterial("pore", 3000.0 * 1e-6, 0.35, 5e-4)}
    mats = S.build_element_arrays(phase, phases)
    load = S.LoadCase("xx", 0, tuple(5e-4 * k for k in range(1, 41)), (1, 2, 3, 4, 5), ())
    op = S.PeriodicVoxelPhaseFieldOperator(S.VoxelGrid((n, n, n), h), mats, load, kappa=1e-6, ell=2e-3,
                                           device=device)
    return op, mats, phase


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported "
              "(each box has one GPU; see the sharding design note).", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from phast_amd import autotune

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    rng = np.random.default_rng(0)

    # ------------------------------------------------------------------ headline
    # 128^3 particle composite: the largest single-GPU voxel workload
    n = 128
    op, mats, phase_info = particle_composite_op(n, dev)
    x = torch.as_tensor(rng.uniform(-1e-6, 1e-6, op.n_mech), dtype=torch.float64, device=dev)
    x[-6:] = torch.tensor([2.5e-4, -7.5e-5, -7.5e-5, 0, 0, 0], dtype=torch.float64)
    d = torch.as_tensor(rng.uniform(0, 1, op.n_red), dtype=torch.float64, device=dev)
    if a.tiles:
        op.set_tiles(*(int(v) for v in a.tiles.split(",")))
        tune = {"fixed": list(op.tiles)}
    else:
        best, timings = autotune.tune_hex8(op, x, d)
        tune = {"best": list(best), "ms": {",".join(map(str, k)): round(v, 5) for k, v in timings.items()}}
    y = torch.empty_like(x)
    t_apply, wall = events_time(lambda: op.mech_action_d(x, d), a.steps, a.warmup)
    gdofs = op.n_mech / t_apply / 1e9
    hbm_gbs = HEX8_BYTES_PER_NODE * op.n_red / t_apply / 1e9
    tflops = HEX8_FLOPS_PER_VOXEL * op.n_elements / t_apply / 1e12

    if a.copy_calibration:  # a plain copy kernel of known size for FETCH_SIZE / WRITE_SIZE scaling
        src = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device=dev).uniform_()
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # ------------------------------------------------------------------ CPU baseline
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(op, mats, n, x, d, y)

    # ------------------------------------------------------------------ secondary measurements
    sec = {}
    if not a.no_secondary:
        sec = secondary(rng, dev, a)
    if not a.no_pfczm:
        try:
            sec["pfczm_pore48_recorded_history"] = pfczm_pore_history(dev)
        except Exception as exc:
            sec["pfczm_pore48_recorded_history"] = {"error": f"{type(exc).__name__}: {exc}"}
    if not a.no_c1:
        try:
            sec["c1_pore_cube_64_40_increments"] = pore_cube_load_path(dev)
        except Exception as exc:
            sec["c1_pore_cube_64_40_increments"] = {"error": f"{type(exc).__name__}: {exc}"}
    if not a.no_batched:
        try:
            sec["c3_batched_64_specimens"] = batched_specimens(dev, a.batched_steps, a.batch_tile,
                                                               not a.no_sequential, a.sequential_window,
                                                               a.batch_steps_per_launch)
        except Exception as exc:  # reported, so that the headline line is still printed
            sec["c3_batched_64_specimens"] = {"error": f"{type(exc).__name__}: {exc}"}
    if not a.no_adjoint:
        try:
            sec["c4_adjoint_gradient"] = adjoint_gradient(dev, a.adjoint_size, a.adjoint_steps, a.adjoint_fd,
                                                          a.adjoint_fd_random, a.checkpoint_every)
        except Exception as exc:
            sec["c4_adjoint_gradient"] = {"error": f"{type(exc).__name__}: {exc}"}

    out = {
        "metric": "element-apply GDOF/s (periodic hex8 voxel mech_action, 128^3 particle composite)",
        "value": gdofs,
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": t_apply * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded particle placement, random displacement fluctuation, macro strain 2.5e-4, "
                "damage U(0,1))",
        "config": {"workload": "hex8 periodic voxel element apply, 128^3 particle composite: seeded non-overlapping "
                               "spheres, radii U[6,14] voxels, 30% volume fraction, particles 70 GPa in a 3 GPa "
                               "matrix, one-voxel interface shell with Gc x 0.1",
                   "n_nodes": op.n_red, "n_dofs": op.n_mech, "particles": phase_info["particles"],
                   "tiles": list(op.tiles), "autotune": tune},
        "roofline": {"bound": "fp64", "achieved": tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": tflops / FP64_PEAK_TFLOPS, "traffic": HEX8_128_TRAFFIC.get(tuple(op.tiles)),
                     "flops_per_voxel": HEX8_FLOPS_PER_VOXEL, "hbm_algorithmic_GB_s": hbm_gbs,
                     "hbm_frac": hbm_gbs / HBM_PEAK_GBS,
                     "note": "FP64 vector arithmetic, no MFMA: 940 FLOP per voxel (sum-factorised) against "
                             "57 B per node is 16.5 FLOP/B, above the FP64 ridge (78.6 TFLOP/s / 8 TB/s = "
                             "9.8), so the FP64 vector peak is the roofline; one launch = apply kernel + "
                             "seam/finalize kernel, timed together with HIP events on the launch stream; "
                             "traffic = HBM bytes per launch from rocprofv3 FETCH_SIZE/WRITE_SIZE for the "
                             "profiled launch shape (null for another)"},
        "cpu_baseline": cpu,
        "secondary": sec,
        "wall_s_timed": wall,
    }
    print(json.dumps(out), flush=True)


def particle_composite_op(n, device):
    from phast_amd import solid3d as S
    from phast_amd import workloads as W
    phase, phases, h, centres, radii = W.particle_composite(n)
    mats = S.build_element_arrays(phase, phases)
    load = S.LoadCase("xx", 0, tuple(2.5e-4 * k for k in range(1, 61)), (1, 2, 3, 4, 5), ())
    op = S.PeriodicVoxelPhaseFieldOperator(S.VoxelGrid((n, n, n), h), mats, load, kappa=1e-6, ell=2e-3,
                                           device=device)
    return op, mats, {"particles": int(len(radii))}


def cpu_baseline(op, mats, n, x, d, y):
    """The torch-CPU restatement of the reference's batched apply (gather, (B,24)x(24,48)
    product, stiffness product, transposed product, index_add_) of the same state on the
    host cores this process may use, median of repeated applies over a bounded sample."""
    import numpy as np
    import torch
    from oracle import hex8 as OH
    from oracle import torch_cpu as TC
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    torch.set_num_threads(threads)
    oop = OH.VoxelOperator((n, n, n), op.grid.h, mats.stiffness_bank, mats.stiffness_index, mats.Gc, mats.lam,
                           mats.mu, mats.driver.astype(np.int64), kappa=1e-6, ell=2e-3)
    ref = TC.TorchCpuVoxelApply(oop)
    xc, dc = x.cpu(), d.cpu()
    del oop
    times = []
    t_end = time.perf_counter() + 20.0
    for k in range(40):
        t0 = time.perf_counter()
        yc = ref.apply(xc, dc)
        times.append(time.perf_counter() - t0)
        if k >= 5 and time.perf_counter() > t_end:
            break
    t_cpu = float(np.median(times[3:] if len(times) > 3 else times))  # 3 warm-ups
    err = float((y.copy_(op.mech_action_d(x, d)).cpu() - yc).abs().max() / yc.abs().max())
    return {"value": op.n_mech / t_cpu / 1e9, "unit": "GDOF/s", "cores": int(torch.get_num_threads()),
            "kind": "port", "sample": f"{len(times)} torch-CPU applies of the same 128^3 state (3 warm-ups, median "
                                      f"of {max(len(times) - 3, 1)}; batches of 65536 voxels: gather, matmul, bmm, "
                                      f"index_add_); max rel. diff to the GPU apply {err:.1e}",
            "ms_per_apply": t_cpu * 1e3, "cpu_model": _cpu_model()}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def secondary(rng, dev, a):
    import numpy as np
    import torch
    from phast_amd import autotune
    from phast_amd import q4 as P
    res = {}
    T = lambda z: torch.as_tensor(z, dtype=torch.float64, device=dev)
    mat = P.Material(E=32000.0, nu=0.2, rho=2.45e-9, Gc=3e-3, l0=0.5)
    parts = set(a.secondary.split(","))
    # the box's achievable HBM rate: a 1 GiB device-to-device copy (read + write), same run
    src = torch.empty(1 << 27, dtype=torch.float64, device=dev).uniform_()
    dst = torch.empty_like(src)
    tc, _ = events_time(lambda: dst.copy_(src), 20, 3)
    res["hbm_copy_GB_s"] = 2 * src.numel() * 8 / tc / 1e9
    del src, dst
    q4_cases = {"q4_c0_200x80": (200, 80, 100.0, 40.0), "q4_8192x4096": (8192, 4096, 4096.0, 2048.0)}
    for name, (nx, ny, W, H) in (q4_cases.items() if "q4" in parts else ()):
        mesh = P.FEMMesh.rectangle(W, H, nx, ny)
        n, E = mesh.n_nodes, mesh.n_elems
        u = T(rng.uniform(-1e-4, 1e-4, (n, 2)))
        d = T(rng.uniform(0, 1, n))
        timings = {}
        if a.q4_tile:
            best = tuple(int(v) for v in a.q4_tile.split(","))
            fem = P.FEMOperators(mesh, mat, tile=best)
        else:
            best, timings, fem = autotune.tune_q4(lambda t: P.FEMOperators(mesh, mat, tile=t), u, d)
        steps = max(10, a.steps // (10 if n > 1e6 else 1))
        t, _ = events_time(lambda: fem.internal_force(u, d), steps, 3)
        gbs = Q4_APPLY_BYTES_PER_NODE * n / t / 1e9
        entry = {"apply_GDOF_s": 2 * n / t / 1e9, "apply_GB_s": gbs, "apply_frac_hbm_peak": gbs / HBM_PEAK_GBS,
                 "apply_frac_of_copy": gbs / res["hbm_copy_GB_s"], "apply_ms": t * 1e3, "tile": list(best),
                 "autotune_ms": {f"{k[0]},{k[1]}": round(v, 4) for k, v in timings.items()}}
        dyn = P.ExplicitDynamics(fem, dt=0.5 * fem.dt_cfl)
        v, acc = T(rng.uniform(-1e-2, 1e-2, (n, 2))), T(np.zeros((n, 2)))
        fext = T(np.zeros((n, 2)))
        Hq = T(np.zeros((E, 4)))
        t2, _ = events_time(lambda: dyn._launch(u, v, acc, d, fext, None, None, H_old=Hq), steps, 3)
        entry.update({"explicit_step_ms": t2 * 1e3, "explicit_step_GB_s": Q4_STEP_BYTES_PER_NODE * n / t2 / 1e9})
        if n > 1e6:
            entry.update(q4_cg_iteration(P, fem, rng, T))
        res[name] = entry
        del fem, dyn, u, d, v, acc, fext, Hq
    if "q4g" in parts:
        res["q4_general_distorted_4096x2048"] = q4_general_secondary(rng, dev, a)
    if "t3" in parts:
        res["t3_split_quad_4096x2048"] = t3_secondary(rng, dev, a)
    if "sent" in parts:
        res["c0_sent_plate_1000_steps"] = sent_plate_run(dev)
    for nv in ((64, 256) if "hex8" in parts else ()):
        op, mats, _ = pore_cube(nv, 12 * nv // 64, dev)
        x = T(rng.uniform(-1e-6, 1e-6, op.n_mech))
        d = T(rng.uniform(0, 1, op.n_red))
        if a.hex8_big_tiles:
            best = tuple(int(v) for v in a.hex8_big_tiles.split(","))
            op.set_tiles(*best)
        else:
            best, _ = autotune.tune_hex8(op, x, d, reps=3)
        t, _ = events_time(lambda: op.mech_action_d(x, d), 20 if nv <= 64 else 10, 3)
        res[f"hex8_{nv}^3_pore_cube"] = {"apply_GDOF_s": op.n_mech / t / 1e9,
                                         "apply_GB_s": HEX8_BYTES_PER_NODE * op.n_red / t / 1e9,
                                         "apply_TFLOP_s": HEX8_FLOPS_PER_VOXEL * op.n_elements / t / 1e12,
                                         "apply_frac_fp64_peak": HEX8_FLOPS_PER_VOXEL * op.n_elements / t / 1e12
                                         / FP64_PEAK_TFLOPS,
                                         "apply_ms": t * 1e3, "tiles": list(best)}
        del op, x, d
        torch.cuda.empty_cache()
    return res


def sent_plate_run(dev, steps=1000, chunk=20):
    """The 200 x 80 SENT plate (preseeded notch, +-1 MPa traction, dt = 0.05 us, projected
    damage CG at tol 1e-5 every step) for 1000 steps, two ways from the same state: the
    per-kernel loop (StaggeredSolver.step_full: a fused explicit-step launch, then the damage
    CG's passes with a host read of its flags every 4 iterations) and the persistent kernel
    with one instance (one launch per step), replayed from a captured graph of ``chunk``
    steps.  Steps/s, mean damage CG iterations, and the largest relative difference between
    the two end states (the two organisations compute the same bits)."""
    import numpy as np
    import torch
    from phast_amd import batch as B
    from phast_amd import q4 as P
    from phast_amd import workloads as W
    p = W.sent_plate(steps=steps)
    mesh = P.FEMMesh(p.nodes, p.elements)
    mat = P.Material(**p.material)
    fem = P.FEMOperators(mesh, mat)
    T = lambda z: torch.as_tensor(np.ascontiguousarray(z), dtype=torch.float64, device=dev)
    dmg = P.PhaseFieldDamageSolver(fem, tol=1e-5, max_iter=5000, bounds_method="projected_cg")
    H0 = T(p.H0)
    d0 = dmg.solve(H0, torch.zeros(mesh.n_nodes, dtype=torch.float64, device=dev))
    fx = T(p.f_ext)

    def per_kernel():
        st = P.StaggeredSolver(fem, P.ExplicitDynamics(fem, dt=p.dt), dmg, f_ext=fx, damage_every=1)
        st.d, st.H_elem = d0.clone(), H0.clone()
        return st

    st = per_kernel()
    for _ in range(5):
        st.step_full()
    st = per_kernel()
    its = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        st.step_full()
        its += dmg.last_iter
    torch.cuda.synchronize()
    wall_pk = time.perf_counter() - t0
    z = torch.zeros((mesh.n_nodes, 2), dtype=torch.float64, device=dev)

    def persistent():
        bat = B.BatchedExplicit([(mesh, mat)], dts=[p.dt], f_ext=[p.f_ext], tol=1e-5, max_iter=5000,
                                bounds_method="projected_cg")
        bat.load_states([(z, z, z, d0, H0)])
        bat.prepare_persistent()
        return bat, bat.capture(chunk)

    bat, g = persistent()
    g.replay()
    del g, bat
    bat, g = persistent()
    its_p = torch.zeros((), dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps // chunk):
        g.replay()
        its_p += g.iters.sum()
    torch.cuda.synchronize()
    wall_p = time.perf_counter() - t0
    bat.check_persistent(g.status[-1])
    n_p = (steps // chunk) * chunk
    diff = 0.0
    for a_, b_ in zip(bat.instance_state(0), (st.u, st.v, st.a, st.d, st.H_elem)):
        diff = max(diff, float((a_ - b_).abs().max() / b_.abs().max().clamp_min(1e-300)))
    return {"nodes": mesh.n_nodes, "steps": steps, "per_kernel_steps_s": steps / wall_pk,
            "per_kernel_ms_per_step": wall_pk / steps * 1e3, "persistent_graph_steps_s": n_p / wall_p,
            "persistent_graph_ms_per_step": wall_p / n_p * 1e3, "speedup_persistent_vs_per_kernel": (n_p / wall_p)
            / (steps / wall_pk), "damage_cg_iters_mean": its / steps,
            "damage_cg_iters_mean_persistent": float(its_p.item()) / n_p,
            "max_rel_diff_persistent_vs_per_kernel": diff}


def q4_cg_iteration(P, fem, rng, T):
    """Per-iteration cost of the device-resident projected damage CG: the difference of two
    solves run to a fixed iteration count (tol 0), so set-up and finish cancel."""
    import time
    import torch
    E, n = fem.mesh.n_elems, fem.mesh.n_nodes
    H = T(rng.uniform(0.0, 2e-3, (E, 4)))
    d0 = torch.zeros(n, dtype=torch.float64, device=H.device)
    out = {}
    walls = {}
    for k in (20, 60):
        dmg = P.PhaseFieldDamageSolver(fem, tol=0.0, max_iter=k, bounds_method="projected_cg")
        dmg.solve(H, d0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dmg.solve(H, d0)
        torch.cuda.synchronize()
        walls[k] = (time.perf_counter() - t0, dmg.last_iter, dmg.last_status)
    (w1, i1, _), (w2, i2, st2) = walls[20], walls[60]
    if i2 > i1:
        t_it = (w2 - w1) / (i2 - i1)
        out = {"cg_iteration_ms": t_it * 1e3, "cg_iteration_GB_s": Q4_CG_BYTES_PER_NODE * n / t_it / 1e9,
               "cg_iterations_timed": [i1, i2], "cg_status": st2}
    return out


def q4_general_secondary(rng, dev, a):
    """Q4 on a general quadrilateral mesh: a 4096 x 2048-cell plate whose interior nodes are
    moved by up to 20 % of a cell (8.4 M nodes, 8.4 M cells; it no longer qualifies for the
    strip kernels): internal force, explicit step with psi+ and history, damage apply and one
    projected damage-CG iteration on the general-quadrilateral plan."""
    import time
    import numpy as np
    import torch
    from phast_amd import q4 as P
    from phast_amd import workloads as W
    T = lambda z: torch.as_tensor(z, dtype=torch.float64, device=dev)
    nx, ny, L = 4096, 2048, 0.125
    nodes, cells = W.structured_plate(nx * L, ny * L, nx, ny)
    inner = (nodes[:, 0] > 0) & (nodes[:, 0] < nx * L) & (nodes[:, 1] > 0) & (nodes[:, 1] < ny * L)
    nodes = nodes.copy()
    nodes[inner] += 0.2 * L * rng.uniform(-1, 1, (int(inner.sum()), 2))
    mesh = P.FEMMesh(nodes, cells, device=dev)
    fem = P.FEMOperators(mesh, P.Material(E=32000.0, nu=0.2, rho=2.45e-9, Gc=3e-3, l0=0.25))
    n, E = mesh.n_nodes, mesh.n_elems
    u = T(rng.uniform(-1e-4, 1e-4, (n, 2)))
    d = T(rng.uniform(0, 1, n))
    steps = max(10, a.steps // 10)
    t, _ = events_time(lambda: fem.internal_force(u, d), steps, 3)
    entry = {"nodes": n, "cells": E, "structured": mesh.structured, "apply_GDOF_s": 2 * n / t / 1e9,
             "apply_GB_s": Q4G_APPLY_BYTES_PER_NODE * n / t / 1e9,
             "apply_design_GB_s": Q4G_APPLY_COUNTED_PER_NODE * n / t / 1e9, "apply_ms": t * 1e3}
    dyn = P.ExplicitDynamics(fem, dt=0.5 * fem.dt_cfl)
    v, acc, fext = T(rng.uniform(-1e-2, 1e-2, (n, 2))), T(np.zeros((n, 2))), T(np.zeros((n, 2)))
    Hq = T(np.zeros((E, 4)))
    t2, _ = events_time(lambda: dyn._launch(u, v, acc, d, fext, None, None, H_old=Hq, want_psi=True), steps, 3)
    dmg = P.PhaseFieldDamageSolver(fem)
    Hd = T(rng.uniform(0, 1e-3, (E, 4)))
    t3, _ = events_time(lambda: dmg.damage_apply(Hd, d), steps, 3)
    entry.update({"explicit_step_ms": t2 * 1e3, "damage_apply_ms": t3 * 1e3})
    walls = {}
    for k in (20, 60):
        solver = P.PhaseFieldDamageSolver(fem, tol=0.0, max_iter=k, bounds_method="projected_cg")
        d0 = torch.zeros(n, dtype=torch.float64, device=dev)
        solver.solve(Hd, d0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        solver.solve(Hd, d0)
        torch.cuda.synchronize()
        walls[k] = (time.perf_counter() - t0, solver.last_iter)
    if walls[60][1] > walls[20][1]:
        entry["cg_iteration_ms"] = (walls[60][0] - walls[20][0]) / (walls[60][1] - walls[20][1]) * 1e3
    del fem, dyn, dmg, u, d, v, acc, fext, Hq, Hd
    torch.cuda.empty_cache()
    return entry


def t3_secondary(rng, dev, a):
    """T3 route on a 4096 x 2048-cell split-quad plate (8.4 M nodes, 16.8 M triangles; the
    mesh of core/mesh_generator.py:1552-1633 at h = 0.125): internal force, explicit step
    with psi+ and history, damage apply, one damage-CG iteration and one adjoint-PCG iteration."""
    import numpy as np
    import torch
    from phast_amd import t3 as P3
    T = lambda z: torch.as_tensor(z, dtype=torch.float64, device=dev)
    mesh = P3.FEMMesh.sent_split_quad(512.0, 256.0, 4096, 2048, device=dev)
    mat = P3.Material(E=32000.0, nu=0.2, rho=2.45e-9, Gc=3e-3, l0=0.25)
    fem = P3.FEMOperators(mesh, mat)
    n, E = mesh.n_nodes, mesh.n_elems
    u = T(rng.uniform(-1e-4, 1e-4, (n, 2)))
    d = T(rng.uniform(0, 1, n))
    steps = max(10, a.steps // 10)
    t, _ = events_time(lambda: fem.internal_force(u, d), steps, 3)
    entry = {"nodes": n, "triangles": E, "apply_GDOF_s": 2 * n / t / 1e9,
             "apply_GB_s": T3_APPLY_BYTES_PER_NODE * n / t / 1e9, "apply_ms": t * 1e3}
    dyn = P3.ExplicitDynamics(fem, dt=0.5 * fem.dt_cfl)
    v, acc, fext = T(rng.uniform(-1e-2, 1e-2, (n, 2))), T(np.zeros((n, 2))), T(np.zeros((n, 2)))
    H = T(np.zeros(E))
    t2, _ = events_time(lambda: dyn._launch(u, v, acc, d, fext, None, None, H_old=H, want_psi=True), steps, 3)
    dmg = P3.PhaseFieldDamageSolver(fem)
    Hd = T(rng.uniform(0, 1e-3, E))
    t3, _ = events_time(lambda: dmg.damage_apply(Hd, d), steps, 3)
    entry.update({"explicit_step_ms": t2 * 1e3, "explicit_step_GB_s": T3_STEP_BYTES_PER_NODE * n / t2 / 1e9,
                  "damage_apply_ms": t3 * 1e3, "damage_apply_GB_s": T3_DAMAGE_BYTES_PER_NODE * n / t3 / 1e9})
    # per-iteration cost of the projected damage CG: difference of 60- and 20-iteration solves
    import time
    walls = {}
    for k in (20, 60):
        solver = P3.PhaseFieldDamageSolver(fem, tol=0.0, max_iter=k, bounds_method="projected_cg")
        d0 = torch.zeros(n, dtype=torch.float64, device=dev)
        solver.solve(Hd, d0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        solver.solve(Hd, d0)
        torch.cuda.synchronize()
        walls[k] = (time.perf_counter() - t0, solver.last_iter)
    if walls[60][1] > walls[20][1]:
        t_it = (walls[60][0] - walls[20][0]) / (walls[60][1] - walls[20][1])
        entry.update({"cg_iteration_ms": t_it * 1e3, "cg_iteration_GB_s": T3_CG_BYTES_PER_NODE * n / t_it / 1e9})
    # per-iteration cost of the adjoint PCG that the backward runs twice (rel_tol 0: fixed count)
    g = T(rng.uniform(-1.0, 1.0, n))
    adj = P3.PhaseFieldDamageSolver(fem)
    walls = {}
    for k in (20, 60):
        adj.max_iter = k
        adj.adjoint_pcg(Hd, g, rel_tol=0.0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        adj.adjoint_pcg(Hd, g, rel_tol=0.0)
        torch.cuda.synchronize()
        walls[k] = (time.perf_counter() - t0, adj.last_adjoint_iter)
    if walls[60][1] > walls[20][1]:
        t_it = (walls[60][0] - walls[20][0]) / (walls[60][1] - walls[20][1])
        entry.update({"adjoint_pcg_iteration_ms": t_it * 1e3,
                      "adjoint_pcg_iteration_GB_s": T3_ADJ_BYTES_PER_NODE * n / t_it / 1e9})
    del fem, dyn, dmg, adj, u, d, v, acc, fext, H, Hd, g
    torch.cuda.empty_cache()
    return entry


def c3_batch_inputs(n_spec=64, seed=2024):
    """The 64 notched plates of phast_amd.workloads.specimen_batch as batch inputs."""
    from phast_amd import q4 as P
    from phast_amd import workloads as W
    mat = P.Material(**W.Q4_MATERIAL)
    specs = W.specimen_batch(n_spec, seed)
    items = [(P.FEMMesh.rectangle(sp.h * sp.nx, sp.h * sp.ny, sp.nx, sp.ny), mat) for sp in specs]
    return items, [sp.Gc_field for sp in specs], [sp.f_ext for sp in specs], [(sp.notch, 1.0) for sp in specs]


def pfczm_pore_history(dev):
    """The reference's PF-CZM pore example (48^3 voxels of 0.5 um, pore diameter 0.4 L, E_xx
    to 3e-3 in 150 increments, spectral driver, crack closure, stagger tol 1e-4, PCG rtol
    1e-8, up to 5 cutbacks): wall time of the whole load path and the largest relative
    difference of the macro stress from the reference's recorded history."""
    import torch
    from phast_amd import solid3d as S
    from phast_amd import staggered3d as G3
    from phast_amd import workloads as W
    gold = json.loads((ROOT / "tests" / "golden" / "pore_in_cube_pfczm_history.json").read_text())
    c = gold["cell"]
    n = c["shape"][0]
    phase, phases, h, ell, load = W.pore_in_cube_pfczm(n, c["increments"], c["to"])
    mats = S.build_element_arrays(phase, phases)
    op = S.PeriodicVoxelPhaseFieldOperator(S.VoxelGrid((n, n, n), h), mats, load, kappa=1e-6, ell=ell,
                                           damage_law="pfczm", crack_closure="axes")
    cfg = G3.SolverConfig(stagger_tol=c["stagger_tol"], pcg_rtol_mech=c["pcg_rtol_mech"],
                          damage_kkt_tol=c["damage_kkt_tol"], max_cutbacks=c["max_cutbacks"])
    solver = G3.StaggeredSolver(op, load, cfg)
    st = solver.initial_state()
    worst, stagger, pcg, newton = 0.0, 0, 0, 0
    peak = (0.0, 0.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for row in gold["rows"]:
        k = int(row["step"])
        st, recs = solver.advance_increment(st, load.targets[k - 1])
        rec = recs[-1]
        worst = max(worst, abs(rec.macro_stress[0] - row["macro_stress_xx"]) / abs(row["macro_stress_xx"]))
        stagger += sum(r.stagger_iters for r in recs)
        pcg += sum(r.pcg_iters_mech for r in recs)
        newton += sum(r.newton_iters_pf or 0 for r in recs)
        if rec.macro_stress[0] > peak[0]:
            peak = (rec.macro_stress[0], rec.macro_strain[0])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return {"voxels": n ** 3, "increments": len(gold["rows"]), "wall_s": wall,
            "increments_per_s": len(gold["rows"]) / wall,
            "max_rel_diff_macro_stress_vs_recorded": worst, "peak_stress_MPa": peak[0],
            "peak_strain": peak[1], "d_max_end": float(st.d.max().item()), "stagger_iterations": stagger,
            "mech_pcg_iterations": pcg, "newton_iterations": newton}


def pore_cube_load_path(dev, n=64, increments=40, active_set_cap=1024):
    """The 64^3 pore cube (centred pore of radius 12 voxels, matrix E = 3 GPa, nu = 0.35,
    Gc = 0.5 J/m^2, l0 = 2 voxels) under 40 E_xx increments of 5e-4, the other five macro
    strains free: AT2 with the spectral driver, multigrid-preconditioned mechanics PCG to rel
    1e-8, box-KKT damage solves with d_prev <= d <= 1.  The reference's default cap of 64
    active-set iterations stops the damage solve of increment 27 (after the cell has broken,
    the bound releases spread by a few voxels per active-set iteration), so the cap is
    raised; everything else is the reference's default SolverConfig."""
    import torch
    from phast_amd import solid3d as S
    from phast_amd import staggered3d as G3
    from phast_amd import workloads as W
    phase, phases, h = W.pore_cube(n, 12.0 * n / 64)
    load = S.LoadCase("xx", 0, tuple(5e-4 * k for k in range(1, increments + 1)), (1, 2, 3, 4, 5), ())
    mats = S.build_element_arrays(phase, phases)
    op = S.PeriodicVoxelPhaseFieldOperator(S.VoxelGrid((n, n, n), h), mats, load, kappa=1e-6, ell=2 * h[0],
                                           device=dev)
    solver = G3.StaggeredSolver(op, cfg=G3.SolverConfig(damage_max_active_set=active_set_cap),
                                preconditioner=S.VoxelMultigrid(op))
    st = solver.initial_state()
    tot = {"stagger": 0, "mech_pcg": 0, "damage_pcg": 0, "active_set": 0, "cutbacks": 0}
    peak, hm, bal, kkt = (0.0, 0.0), 0.0, 0.0, True
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(increments):
        st, recs = solver.advance_increment(st, load.targets[k])
        for r in recs:
            tot["stagger"] += r.stagger_iters
            tot["mech_pcg"] += r.pcg_iters_mech
            tot["damage_pcg"] += r.pcg_iters_pf
            tot["active_set"] += r.active_set_iters
            tot["cutbacks"] = max(tot["cutbacks"], r.cutbacks)
            hm = max(hm, r.hill_mandel_rel)
            kkt = kkt and r.kkt_ok
        rec = recs[-1]
        if rec.macro_stress[0] > peak[0]:
            peak = (rec.macro_stress[0], rec.macro_strain[0])
        bal = max(bal, abs(rec.balance_residual_rel))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return {"voxels": n ** 3, "increments": increments, "wall_s": wall, "increments_per_s": increments / wall,
            "damage_max_active_set": active_set_cap, "peak_stress_MPa": peak[0], "peak_strain": peak[1],
            "stress_end_MPa": float(st.macro_stress[0]), "d_max_end": float(st.d.max().item()),
            "max_hill_mandel_rel": hm, "all_damage_solves_kkt": bool(kkt),
            **{f"{k}_iterations" if k != "cutbacks" else "max_cutback_level": v for k, v in tot.items()}}


def batched_specimens(dev, steps, fixed_tile="", sequential=True, window=0, steps_per_launch=20):
    """Specimen-steps/s of the 64-plate batch: the persistent batched step replayed from a
    captured graph over the whole run; the per-kernel batched path; and, over the same
    steps (the whole run, or the first ``window`` steps when ``window`` is given), the
    batched run against 64 sequential single-specimen runs, with their agreement at the
    last of those steps."""
    import numpy as np
    import torch
    from phast_amd import batch as B
    from phast_amd import _lib
    from phast_amd import autotune
    items, gcs, fext, fixed = c3_batch_inputs()
    n_spec = len(items)
    t_start = time.perf_counter()
    log = lambda msg: print(f"[batched {time.perf_counter() - t_start:7.1f}s] {msg}", file=sys.stderr, flush=True)
    mk = lambda it, g, f, x, tile=(128, 8): B.BatchedExplicit(it, dt_factor=0.4, Gc_fields=g, f_ext=f, fixed=x,
                                                              bounds_method="projected_cg", tol=1e-5, max_iter=5000,
                                                              tile=tile)
    res = {"specimens": n_spec, "total_nodes": int(sum(m.n_nodes for m, _ in items)),
           "nx_range": [int(min(m.nx for m, _ in items)), int(max(m.nx for m, _ in items))]}
    if fixed_tile:
        best = tuple(int(v) for v in fixed_tile.split(","))
        res["tile"] = list(best)
    else:  # strip shape of the persistent step: tuned on a state 100 steps into the run
        warm = mk(items, gcs, fext, fixed)
        warm.step_persistent(100)
        warm.check_persistent()
        state = warm.packed_state()
        del warm
        best, timings = autotune.tune_batch(lambda t: mk(items, gcs, fext, fixed, t), state=state, steps=10)
        res["tile"] = list(best)
        res["autotune_ms_per_step"] = {f"{k[0]},{k[1]}": round(v, 3) for k, v in timings.items()}
        del state
    log(f"strip {best}")
    # persistent kernel + graph over the full run: `steps_per_launch` steps per launch (each
    # specimen runs ahead on its own blocks within a launch), and one launch per step (every
    # specimen in lockstep)
    chunk = 20
    reps = max(steps // chunk, 1)
    n_steps = reps * chunk

    def full(spl):
        bat = mk(items, gcs, fext, fixed, best)
        grid = list(bat.prepare_persistent())
        g = bat.capture(chunk, steps_per_launch=spl)
        bad = torch.zeros((), dtype=torch.int64, device=dev)
        its = torch.zeros(n_spec, dtype=torch.int64, device=dev)
        its_max = torch.zeros((), dtype=torch.int64, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for rep_i in range(reps):
            g.replay()
            if rep_i % 20 == 19:
                log(f"{spl} steps per launch: replayed {(rep_i + 1) * chunk} steps")
            bad += (g.status != _lib.CG_CONVERGED).sum()
            its += g.iters.sum(0)
            its_max = torch.maximum(its_max, g.iters.max().to(torch.int64))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        bat.check_persistent(g.status)
        return bat, g, wall, grid, int(bad.item()), its, int(its_max.item())

    bat, g, wall, grid, bad, its, its_max = full(steps_per_launch)
    res.update({"steps": n_steps, "graph_chunk": chunk, "steps_per_launch": steps_per_launch, "grid_blocks": grid,
                "persistent_graph_wall_s": wall,
                "persistent_graph_specimen_steps_s": n_spec * n_steps / wall,
                "unconverged_solves": bad,
                "damage_cg_iters_per_step_mean": float(its.sum().item()) / (n_spec * n_steps),
                "damage_cg_iters_max": its_max,
                "ms_per_step": wall / n_steps * 1e3,
                "d_max_end": float(bat.d.max().item())})
    full_run = (bat, g, wall, n_steps)
    if steps_per_launch != 1:
        lock, g1, wall1, grid1, bad1, _, _ = full(1)
        same = all(torch.equal(x, y) for x, y in zip(lock.packed_state(), bat.packed_state()))
        res.update({"lockstep_grid_blocks": grid1, "lockstep_wall_s": wall1,
                    "lockstep_specimen_steps_s": n_spec * n_steps / wall1,
                    "lockstep_unconverged_solves": bad1,
                    "speedup_run_ahead_vs_lockstep": wall1 / wall,
                    "lockstep_bitwise_equal_at_end": bool(same)})
        del lock, g1
    # per-kernel batched path (one launch per kernel over all specimens, host polls the CG)
    short = max(min(100, steps // 10), 10)
    bat = mk(items, gcs, fext, fixed)
    bat.step(2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bat.step(short)
    torch.cuda.synchronize()
    res["per_kernel_batched_specimen_steps_s"] = n_spec * short / (time.perf_counter() - t0)
    del bat
    if not sequential:
        return res
    # the same steps, batched (persistent kernel, graph) and as 64 sequential runs of the
    # per-kernel single-specimen path; agreement at the last of those steps
    if not window or window >= n_steps:
        bat, g, wall_bat, window = full_run
    else:
        del full_run
        window = max((window // chunk) * chunk, chunk)
        bat = mk(items, gcs, fext, fixed, best)
        bat.prepare_persistent()
        g = bat.capture(chunk, steps_per_launch=steps_per_launch)
        g.replay()  # warm the graph, then restart from the initial state
        del g, bat
        bat = mk(items, gcs, fext, fixed, best)
        bat.prepare_persistent()
        g = bat.capture(chunk, steps_per_launch=steps_per_launch)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.replay(window // chunk)
        torch.cuda.synchronize()
        wall_bat = time.perf_counter() - t0
        bat.check_persistent(g.status)
    singles = [mk([items[k]], [gcs[k]], [fext[k]], [fixed[k]], best) for k in range(n_spec)]
    singles[0].step(1)  # warm the kernels, then rebuild the first specimen
    singles[0] = mk([items[0]], [gcs[0]], [fext[0]], [fixed[0]], best)
    log(f"sequential single-specimen runs over {window} steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k, one in enumerate(singles):
        one.step(window)
        if k % 16 == 15:
            log(f"{k + 1} sequential specimens done")
    torch.cuda.synchronize()
    wall_seq = time.perf_counter() - t0
    worst = 0.0
    for k, one in enumerate(singles):
        for xb, ys in zip(bat.instance_state(k), one.instance_state(0)):
            worst = max(worst, float((xb - ys).abs().max().item()) / max(float(ys.abs().max().item()), 1e-300))
    res.update({"window_steps": window,
                "window_batched_specimen_steps_s": n_spec * window / wall_bat,
                "window_sequential_specimen_steps_s": n_spec * window / wall_seq,
                "speedup_batched_vs_sequential_same_window": wall_seq / wall_bat,
                "max_rel_diff_batched_vs_sequential_at_window_end": worst})
    del singles, bat, g
    torch.cuda.empty_cache()
    return res


def adjoint_gradient(dev, n, steps, n_fd, n_fd_random=0, checkpoint_every=5):
    """Gradient of L = sum_k (R_k(Gc) - R_k(Gc_target))^2 over a load path of the n^3 cell
    with 20 seeded periodic pores (radii U[4, 9] voxels), Gc = 5e-4 (1 + 0.1 N(0,1)) per voxel
    (phast_amd.workloads.porous_gc_cube): reverse-sweep wall (load-step checkpoints every
    ``checkpoint_every`` steps, the rest recomputed) against forward wall; central
    differences (step 1e-6 N/mm) on the ``n_fd`` most sensitive voxels and on ``n_fd_random``
    seeded random voxels."""
    import numpy as np
    import torch
    from phast_amd import adjoint3d as A
    from phast_amd import solid3d as S
    from phast_amd import workloads as W
    phase, phases, h, Gc_np, Gc_t_np = W.porous_gc_cube(n, 20, (4.0, 9.0), seed=96, gc_seed=4)
    mats = S.build_element_arrays(phase, phases)
    targets = tuple(1e-4 * k for k in range(1, steps + 1))  # up to 3e-3: before the peak
    load = S.LoadCase("xx", 0, targets, (1, 2, 3, 4, 5), ())
    op = S.PeriodicVoxelPhaseFieldOperator(S.VoxelGrid((n, n, n), h), mats, load, kappa=1e-6, ell=2e-3, device=dev)
    mg = S.VoxelMultigrid(op)
    t_start = time.perf_counter()
    log = lambda msg: print(f"[adjoint {time.perf_counter() - t_start:7.1f}s] {msg}", file=sys.stderr, flush=True)
    curve = A.ReactionCurve(op, preconditioner=mg, progress=log, checkpoint_every=checkpoint_every)
    E = op.n_elements
    Gc = torch.as_tensor(Gc_np, dtype=torch.float64, device=dev)
    Gc_t = torch.as_tensor(Gc_t_np, dtype=torch.float64, device=dev)
    R_hat, _ = curve.forward(Gc_t)
    R_hat = R_hat.detach().clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    R, tape = curve.forward(Gc)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    L = float(((R - R_hat) ** 2).sum())
    t0 = time.perf_counter()
    grad = curve.backward(tape, 2.0 * (R - R_hat))
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t0
    stats = curve.stats
    del tape

    def loss(g):
        Rg, _ = curve.forward(g)
        return float(((Rg - R_hat) ** 2).sum())

    def fd(e):
        step = 1e-6
        gp, gm = Gc.clone(), Gc.clone()
        gp[e] += step
        gm[e] -= step
        return (loss(gp) - loss(gm)) / (2 * step)

    gmax = float(grad.abs().max())
    top = torch.argsort(grad.abs(), descending=True)[:n_fd].tolist()
    errs = []
    for e in top:
        f1 = fd(e)
        errs.append(abs(float(grad[e]) - f1) / max(abs(f1), 1e-300))
    rnd = np.random.default_rng(32).choice(E, n_fd_random, replace=False).tolist() if n_fd_random else []
    rnd_errs = []
    for e in rnd:  # |adjoint - FD| / max(|FD|, 1e-2 max|grad|)
        f1 = fd(e)
        rnd_errs.append(abs(float(grad[e]) - f1) / max(abs(f1), 1e-2 * gmax))
    out = {"cells": n, "voxels": E, "load_steps": steps, "pores": 20, "checkpoint_every": checkpoint_every,
           "forward_wall_s": t_fwd, "gradient_wall_s": t_bwd, "gradient_over_forward": t_bwd / t_fwd, "loss": L,
           "fd_step": 1e-6, "fd_sensitive_voxels": n_fd, "voxel_fd_rel_errs": errs,
           "voxel_fd_rel_err_max": max(errs) if errs else None,
           "adjoint_block_iterations": stats.block_iterations[::-1],
           "adjoint_mech_pcg_iterations": stats.mech_pcg_iterations,
           "adjoint_damage_pcg_iterations": stats.damage_pcg_iterations,
           "preconditioner": "voxel multigrid V-cycle"}
    if rnd:
        out.update({"fd_random_voxels": len(rnd), "fd_random_err_max": max(rnd_errs),
                    "fd_random_criterion": "|adjoint - FD| / max(|FD|, 1e-2 max|grad|)"})
    return out


if __name__ == "__main__":
    main()
