"""Many explicit phase-field plates of uneven sizes advanced together.

All plates of a batch live packed in one set of HBM arrays with per-instance offsets.
Two ways to advance them:

* ``step``: every kernel launch of a step (the fused mechanics/psi+/history kernel and
  each kernel of the damage CG) covers the whole batch through the plan's
  block -> (instance, strip) table; the host polls the CG every few iterations.
* ``step_persistent`` / ``capture``: ONE persistent kernel per step runs the mechanics
  and the whole damage solve of every instance, with barriers among the blocks of each
  instance instead of kernel boundaries and no host round trip, so a sequence of steps
  can be captured into a graph and replayed.  With ``steps_per_launch > 1`` one launch
  runs several steps, and each instance advances through them on its own blocks without
  waiting for the other instances between steps.

Each instance keeps its own CG scalars and its own partial sums, so the result of an
instance is bitwise the result of running it alone, on either path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr, stream
from .q4 import F64, FEMMesh, Material, Q4Plan, q4_desc


class BatchedExplicit:
    """Explicit staggered steps (staggered_solver.py:1115-1165 with damage_every = 1) for a
    batch of plates.  ``items`` is a list of ``(FEMMesh, Material)``; ``Gc_fields`` an
    optional list of per-cell Gc arrays (E_k,); ``f_ext`` an optional list of (N_k,2)."""

    def __init__(self, items, *, dt_factor: float = 0.5, dts=None, Gc_fields=None, f_ext=None, tol: float = 1e-5,
                 max_iter: int = 5000, bounds_method: str = "projected_cg", tile=(128, 8), fixed=None):
        dev = _lib.require_device()
        self.device = dev
        self.meshes = [m for m, _ in items]
        self.materials = [mat for _, mat in items]
        cfl = [m.h_min / mat.c_p for m, mat in items]
        self.dt = list(dts) if dts is not None else [dt_factor * c for c in cfl]
        self.plan = Q4Plan([q4_desc(m, mat, dt) for (m, mat), dt in zip(items, self.dt)], tile=tile)
        self.tol, self.max_iter = float(tol), int(max_iter)
        self.method = 0 if bounds_method == "projected_cg" else 1
        N, E = self.plan.n_nodes, self.plan.n_cells
        z2 = lambda: torch.zeros((N, 2), dtype=F64, device=dev)
        self.u, self.v, self.a = z2(), z2(), z2()
        self.u1, self.v1, self.a1 = z2(), z2(), z2()
        self.d = torch.zeros(N, dtype=F64, device=dev)
        self.d1 = torch.zeros_like(self.d)
        self.H = torch.zeros((E, 4), dtype=F64, device=dev)
        self.H1 = torch.zeros_like(self.H)
        self.psi = torch.zeros_like(self.H)
        self.f_ext = z2()
        M = torch.empty(N, dtype=F64, device=dev)
        check(_lib.lib().phast_q4_lumped_mass(self.plan.handle, ptr(M), stream()), "lumped mass")
        rho = torch.empty(N, dtype=F64, device=dev)
        for k, mat in enumerate(self.materials):
            rho[self._nodes(k)] = mat.rho
        self.Minv = 1.0 / (rho * M + 1e-30)
        self.Gc = None
        if Gc_fields is not None:
            self.Gc = torch.empty(E, dtype=F64, device=dev)
            for k, g in enumerate(Gc_fields):
                self.Gc[self._cells(k)] = torch.as_tensor(np.asarray(g, dtype=np.float64), device=dev)
        if f_ext is not None:
            for k, f in enumerate(f_ext):
                if f is not None:
                    self.f_ext[self._nodes(k)] = torch.as_tensor(np.asarray(f, dtype=np.float64), device=dev)
        # pf_dirichlet pins (e.g. a notch held at d = 1): fixed[k] = (node ids, values) or None
        self.fixed = self.fixed_vals = None
        if fixed is not None and any(f is not None for f in fixed):
            fm = np.zeros(N, dtype=np.uint8)
            fv = np.zeros(N, dtype=np.float64)
            for k, f in enumerate(fixed):
                if f is None:
                    continue
                ids, vals = f
                ids = np.asarray(ids, dtype=np.int64) + self.plan.node_off[k]
                fm[ids] = 1
                fv[ids] = vals
            self.fixed = torch.as_tensor(fm, device=dev)
            self.fixed_vals = torch.as_tensor(fv, device=dev)
            self.d[self.fixed.bool()] = self.fixed_vals[self.fixed.bool()]
        self.iters_dev = torch.zeros(self.n_inst, dtype=torch.int32, device=dev)
        self.status_dev = torch.zeros(self.n_inst, dtype=torch.int32, device=dev)
        self.grid = None
        self.iters = (C.c_int32 * self.plan.n_inst)()
        self.status = (C.c_int32 * self.plan.n_inst)()
        self.last_iters = np.zeros(self.plan.n_inst, dtype=np.int64)
        self.steps = 0

    @property
    def n_inst(self) -> int:
        return self.plan.n_inst

    def _nodes(self, k):
        s = self.plan.node_off[k]
        return slice(s, s + self.meshes[k].n_nodes)

    def _cells(self, k):
        s = self.plan.cell_off[k]
        return slice(s, s + self.meshes[k].n_elems)

    def load_states(self, states):
        """states[k] = (u, v, a, d, H) device tensors of instance k."""
        for k, (u, v, a, d, H) in enumerate(states):
            n, c = self._nodes(k), self._cells(k)
            self.u[n], self.v[n], self.a[n], self.d[n], self.H[c] = u, v, a, d, H

    def set_damage(self, k: int, d):
        self.d[self._nodes(k)] = d

    def instance_state(self, k: int):
        n, c = self._nodes(k), self._cells(k)
        return self.u[n], self.v[n], self.a[n], self.d[n], self.H[c]

    def step(self, n_steps: int = 1) -> None:
        L = _lib.lib()
        s = stream()
        for _ in range(n_steps):
            check(L.phast_q4_explicit_step(self.plan.handle, ptr(self.u), ptr(self.v), ptr(self.a), ptr(self.d),
                                           ptr(self.f_ext), None, ptr(self.Minv), None, None, ptr(self.H), ptr(self.u1),
                                           ptr(self.v1), ptr(self.a1), ptr(self.H1), ptr(self.psi), s),
                  "batched explicit step")
            self.u, self.u1 = self.u1, self.u
            self.v, self.v1 = self.v1, self.v
            self.a, self.a1 = self.a1, self.a
            self.H, self.H1 = self.H1, self.H
            check(L.phast_q4_damage_solve(self.plan.handle, self.method, ptr(self.H), ptr(self.Gc), ptr(self.d), None,
                                          ptr(self.fixed), ptr(self.fixed_vals), self.tol, self.max_iter,
                                          ptr(self.d1), None, self.iters, self.status, s), "batched damage solve")
            self.d, self.d1 = self.d1, self.d
            self.last_iters = np.frombuffer(self.iters, dtype=np.int32).astype(np.int64)
            self.steps += 1

    def packed_state(self):
        """Clones of the packed (u, v, a, d, H) of the whole batch."""
        return tuple(x.clone() for x in (self.u, self.v, self.a, self.d, self.H))

    def set_packed_state(self, state) -> None:
        """Load a packed state taken from a batch of the same instances (any strip shape)."""
        for dst, src in zip((self.u, self.v, self.a, self.d, self.H), state):
            dst.copy_(src)

    # ------------------------------------------------------------------ persistent path
    def prepare_persistent(self):
        """Size the co-resident grid of the persistent step kernel (once, outside capture)."""
        g = (C.c_int32 * 2)()
        check(_lib.lib().phast_q4_batch_prepare(self.plan.handle, g), "batch prepare")
        self.grid = tuple(g)
        return self.grid

    def _launch_persistent(self, iters: torch.Tensor, status: torch.Tensor) -> None:
        check(_lib.lib().phast_q4_batch_step(
            self.plan.handle, self.method, ptr(self.u), ptr(self.v), ptr(self.a), ptr(self.d), ptr(self.f_ext),
            ptr(self.Minv), None, None, ptr(self.H), ptr(self.Gc), ptr(self.fixed), ptr(self.fixed_vals), self.tol,
            self.max_iter, ptr(self.u1), ptr(self.v1), ptr(self.a1), ptr(self.H1), ptr(self.psi), ptr(self.d1),
            ptr(iters), ptr(status), stream()), "persistent batched step")
        self.u, self.u1 = self.u1, self.u
        self.v, self.v1 = self.v1, self.v
        self.a, self.a1 = self.a1, self.a
        self.H, self.H1 = self.H1, self.H
        self.d, self.d1 = self.d1, self.d

    def _launch_run(self, n_steps: int, iters: torch.Tensor, status: torch.Tensor) -> None:
        check(_lib.lib().phast_q4_batch_run(
            self.plan.handle, self.method, int(n_steps), ptr(self.u), ptr(self.v), ptr(self.a), ptr(self.d),
            ptr(self.f_ext), ptr(self.Minv), None, None, ptr(self.H), ptr(self.Gc), ptr(self.fixed),
            ptr(self.fixed_vals), self.tol, self.max_iter, ptr(self.u1), ptr(self.v1), ptr(self.a1), ptr(self.H1),
            ptr(self.psi), ptr(self.d1), ptr(iters), ptr(status), stream()), "persistent batched run")
        if n_steps % 2:  # an odd number of steps leaves the state in the second buffer set
            self.u, self.u1 = self.u1, self.u
            self.v, self.v1 = self.v1, self.v
            self.a, self.a1 = self.a1, self.a
            self.H, self.H1 = self.H1, self.H
            self.d, self.d1 = self.d1, self.d

    def step_persistent(self, n_steps: int = 1, steps_per_launch: int = 1) -> None:
        """Advance every instance ``n_steps`` explicit steps.  ``steps_per_launch = 1``: one
        persistent kernel per step, every instance in lockstep.  Larger: that many steps per
        launch, each instance running ahead on its own blocks within a launch."""
        if self.grid is None:
            self.prepare_persistent()
        spl = max(1, int(steps_per_launch))
        done = 0
        while done < n_steps:
            k = min(spl, n_steps - done)
            if spl == 1:
                self._launch_persistent(self.iters_dev, self.status_dev)
            else:
                iters = torch.zeros((k, self.n_inst), dtype=torch.int32, device=self.device)
                status = torch.zeros_like(iters)
                self._launch_run(k, iters, status)
                # per instance: the last step's iterations; the first failed step's status, if any
                bad = status != _lib.CG_CONVERGED
                first_bad = status.gather(0, bad.to(torch.int32).argmax(dim=0, keepdim=True))[0]
                self.iters_dev.copy_(iters[-1])
                self.status_dev.copy_(torch.where(bad.any(dim=0), first_bad, status[-1]))
            done += k
            self.steps += k

    def capture(self, n_steps: int = 2, steps_per_launch: int = 1) -> "BatchGraph":
        """Capture ``n_steps`` (even, so the ping-pong buffers come back) persistent steps into
        a graph, ``steps_per_launch`` (a divisor of ``n_steps``) steps per kernel launch; each
        replay advances the batch by ``n_steps``."""
        spl = int(steps_per_launch)
        if n_steps <= 0 or n_steps % 2:
            raise ValueError("n_steps must be a positive even number")
        if spl <= 0 or n_steps % spl:
            raise ValueError("steps_per_launch must be a positive divisor of n_steps")
        if self.grid is None:
            self.prepare_persistent()
        iters = torch.zeros((n_steps, self.n_inst), dtype=torch.int32, device=self.device)
        status = torch.zeros_like(iters)
        graph = torch.cuda.CUDAGraph()
        torch.cuda.synchronize()
        with torch.cuda.graph(graph):
            for k in range(0, n_steps, spl):
                if spl == 1:
                    self._launch_persistent(iters[k], status[k])
                else:
                    self._launch_run(spl, iters[k:k + spl], status[k:k + spl])
        return BatchGraph(self, graph, iters, status, n_steps)

    def check_persistent(self, status: torch.Tensor | None = None) -> np.ndarray:
        """Raise if a barrier wait ran out or an instance's damage solve did not converge;
        return the device status array as numpy."""
        err = C.c_int32()
        check(_lib.lib().phast_q4_batch_error(self.plan.handle, C.byref(err), 1), "batch error flag")
        if err.value:
            raise _lib.PhastLibraryError("persistent batched step: a barrier wait timed out (results invalid)")
        st = (self.status_dev if status is None else status).cpu().numpy()
        if np.any(st != _lib.CG_CONVERGED):
            bad = sorted(set(int(v) for v in np.unique(st)) - {_lib.CG_CONVERGED})
            raise RuntimeError("damage solve did not converge on every instance: status "
                               + ", ".join(_lib.CG_STATUS_NAMES.get(b, str(b)) for b in bad))
        return st


class BatchGraph:
    """A captured sequence of persistent batched steps."""

    def __init__(self, batch: BatchedExplicit, graph, iters, status, n_steps: int):
        self.batch, self.graph, self.iters, self.status, self.n_steps = batch, graph, iters, status, n_steps

    def replay(self, times: int = 1) -> None:
        for _ in range(times):
            self.graph.replay()
            self.batch.steps += self.n_steps
