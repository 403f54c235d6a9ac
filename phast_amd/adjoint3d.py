"""Gradient of a reaction curve of the voxel cell with respect to the per-voxel Gc field.

The forward is the staggered AT2 solver over the load targets (``staggered3d``, which
mirrors ``solid3d/staggered.py``); the quantity of interest is the controlled macro stress
R_k after each increment (the reaction, ``solid3d/staggered.py:1487``).  The reference has
no adjoint for this path; the reverse sweep here differentiates the converged discrete
states exactly (up to the linear-solver tolerances):

at each accepted state k (x, d, H) with history H = max(H_prev, Y(x)) and the box-KKT
active sets (lower L, upper U) of its damage solve,

    F(x, d)   = A(g(d)) x                      = 0   on the free mechanics dofs
    G(d, H)   = s (A_Gc + c_h M_H) d - s c_h b_H = 0 on the inactive damage nodes I
    d_L = d_prev_L,  d_U = 1,  H_Q = Y(x),  H_notQ = H_prev  (Q: Y(x) > H_prev)

the adjoint pair (l_x, l_d) solves, by block Gauss-Seidel (the transpose of the forward's
alternate minimisation),

    K l_x      = -g_x - Y'(x)^T [Q (g_H + G_H^T l_d)]     (mechanics, same constraints)
    G_II l_d_I = -g_d_I - (F_d^T l_x)_I,  l_d = 0 on L u U (damage, exact Dirichlet)

and then dL/dGc += G_Gc^T l_d, the previous state's seeds are
g_d_prev = [g_d + F_d^T l_x + G_dd l_d] on L and g_H_prev = [g_H + G_H^T l_d] off Q.
Every operator above is a HIP kernel of the library; the solves are the device-resident
constrained PCG and exact-Dirichlet PCG of the forward.

``checkpoint_every = m`` keeps the converged state of every m-th load step only and
recomputes the others segment by segment during the reverse sweep.
"""
from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass

import torch

from .solid3d import F64, LoadCase, PeriodicVoxelPhaseFieldOperator, solve_h8_scalar_dirichlet_pcg
from .staggered3d import SolverConfig, SolverState, StaggeredSolver

TIGHT = SolverConfig(stagger_tol=1e-11, min_stagger=2, pcg_rtol_mech=1e-12, pcg_maxit_mech=50000,
                     damage_linear_tol=1e-13, damage_kkt_tol=1e-11, damage_maxit=20000,
                     on_stagger_failure="raise")


@dataclass
class StepRecord:
    """Converged state of one load step and what the reverse sweep needs from it."""
    x: torch.Tensor
    d: torch.Tensor
    H: torch.Tensor
    H_prev: torch.Tensor
    lower: torch.Tensor
    upper: torch.Tensor
    target: float


@dataclass
class AdjointStats:
    block_iterations: list
    mech_pcg_iterations: int = 0
    damage_pcg_iterations: int = 0


class ReactionCurve:
    """R(Gc) = controlled macro stress after each load target, and its reverse sweep."""

    def __init__(self, op: PeriodicVoxelPhaseFieldOperator, load: LoadCase | None = None,
                 cfg: SolverConfig = TIGHT, *, adjoint_tol: float = 1e-11, max_block_iterations: int = 100,
                 mech_rtol: float = 1e-12, damage_tol: float = 1e-13, checkpoint_every: int = 1,
                 preconditioner=None, progress=None):
        if checkpoint_every < 1:
            raise ValueError("checkpoint_every must be >= 1")
        self.op = op
        self.load = op.load if load is None else load
        self.cfg = cfg
        self.adjoint_tol, self.max_block = float(adjoint_tol), int(max_block_iterations)
        self.mech_rtol, self.damage_tol = float(mech_rtol), float(damage_tol)
        self.checkpoint_every = int(checkpoint_every)
        self.preconditioner = preconditioner
        self.progress = progress  # optional callable(str) told of every finished step
        self.stats: AdjointStats | None = None

    # ------------------------------------------------------------------ forward
    def _solver(self) -> StaggeredSolver:
        return StaggeredSolver(self.op, self.load, self.cfg, preconditioner=self.preconditioner)

    def _step(self, solver, state: SolverState, k: int):
        new, _ = solver.advance(state, self.load.targets[k], increment=k + 1)
        lower, upper = new.active
        rec = StepRecord(new.x, new.d, new.H, state.H, lower.clone(), upper.clone(), float(self.load.targets[k]))
        return new, rec

    def forward(self, Gc: torch.Tensor):
        """Run the load path for the Gc field; return (R (K,), tape)."""
        self.op.set_gc(Gc)
        solver = self._solver()
        state = solver.initial_state()
        c = int(self.load.controlled)
        R, tape = [], {"checkpoints": {0: state.clone()}, "records": {}, "Gc": Gc.detach().clone()}
        for k in range(len(self.load.targets)):
            state, rec = self._step(solver, state, k)
            R.append(state.macro_stress[c])
            if self.progress is not None:
                self.progress(f"forward step {k + 1}/{len(self.load.targets)}")
            if self.checkpoint_every == 1:
                tape["records"][k] = rec
            elif (k + 1) % self.checkpoint_every == 0:
                tape["checkpoints"][k + 1] = state.clone()
        return torch.stack(R), tape

    def _records(self, tape, k: int) -> StepRecord:
        """Record of step k, recomputing its segment from the nearest checkpoint if needed."""
        recs = tape["records"]
        if k in recs:
            return recs.pop(k) if self.checkpoint_every > 1 else recs[k]
        start = max(s for s in tape["checkpoints"] if s <= k)
        solver = self._solver()
        state = tape["checkpoints"][start].clone()
        for j in range(start, k + 1):
            state, rec = self._step(solver, state, j)
            recs[j] = rec
        return recs.pop(k)

    # ------------------------------------------------------------------ reverse sweep
    def backward(self, tape, grad_R: torch.Tensor) -> torch.Tensor:
        op = self.op
        dev = op.device
        n, E = op.n_red, op.n_elements
        K = len(self.load.targets)
        c = int(self.load.controlled)
        grad_R = grad_R.to(device=dev, dtype=F64)
        op.set_gc(tape["Gc"])  # the operator may have seen other fields since the forward
        mask = op.mech_fixed_mask()
        zeros_x = torch.zeros(op.n_mech, dtype=F64, device=dev)
        zeros_n = torch.zeros(n, dtype=F64, device=dev)
        ec = zeros_x.clone()
        ec[3 * n + c] = 1.0
        gd_next = zeros_n.clone()
        gH_next = torch.zeros((E, 8), dtype=F64, device=dev)
        grad = torch.zeros(E, dtype=F64, device=dev)
        stats = AdjointStats(block_iterations=[])
        for k in range(K - 1, -1, -1):
            rec = self._records(tape, k)
            rho = grad_R[k] / op.volume
            g = op.degradation(rec.d)
            gx = rho * op.mech_action(ec, g)                                   # dR/dx
            gd = rho * op.qp_to_node(op.adj_coupling(rec.x, ec, rec.d)) + gd_next  # dR/dd + carry
            gH = gH_next
            Q = rec.H > rec.H_prev
            act = rec.lower | rec.upper
            system = op.damage_system(rec.H)
            s = float(system._scale)
            diag_x = op.mech_diagonal(g)
            lam_x, lam_d = zeros_x.clone(), zeros_n.clone()
            fd = zeros_n.clone()
            fresh = True  # the multigrid hierarchy is rebuilt once per state
            it = 0
            for it in range(1, self.max_block + 1):
                t = torch.where(Q, gH + op.adj_damage_history(rec.d, lam_d, s), 0.0)
                rhs_x = -(gx + op.qp_stress_apply(op.adj_driving(rec.x, t)))
                if bool((rhs_x[~mask] != 0).any()):
                    kw = dict(d=rec.d, fixed_mask=mask, fixed_values=zeros_x, diag=diag_x, rhs=rhs_x, x0=lam_x,
                              rtol=self.mech_rtol, atol=1e-300, maxiter=self.cfg.pcg_maxit_mech,
                              check_every=self.cfg.check_every)
                    if self.preconditioner is not None:
                        lam_x, info = self.preconditioner.solve(refresh=fresh, **kw)
                        fresh = False
                    else:
                        lam_x, info = op.solve_mechanics(**kw)
                    if not info.converged:
                        raise RuntimeError(f"adjoint mechanics PCG did not converge at step {k + 1} "
                                           f"(relative residual {info.relative_residual:.3e})")
                    stats.mech_pcg_iterations += info.iterations
                else:
                    lam_x = zeros_x.clone()
                fd = op.qp_to_node(op.adj_coupling(rec.x, lam_x, rec.d))
                rhs_d = -(gd + fd)
                if bool((rhs_d[~act] != 0).any()):
                    res = solve_h8_scalar_dirichlet_pcg(dataclasses.replace(system, rhs=rhs_d), fixed=act,
                                                        values=zeros_n, initial=lam_d, tolerance=self.damage_tol,
                                                        max_iterations=self.cfg.damage_maxit)
                    if not res.converged:
                        raise RuntimeError(f"adjoint damage PCG did not converge at step {k + 1}")
                    new = res.solution
                    stats.damage_pcg_iterations += res.krylov.iterations
                else:
                    new = zeros_n.clone()
                num = float(torch.linalg.vector_norm(new - lam_d))
                den = float(torch.linalg.vector_norm(new))
                lam_d = new
                if num <= self.adjoint_tol * den or den == 0.0:
                    break
            else:
                raise RuntimeError(f"adjoint block iteration did not converge at step {k + 1}")
            stats.block_iterations.append(it)
            if self.progress is not None:
                self.progress(f"reverse step {k + 1} ({it} block iterations)")
            grad += op.adj_damage_gc(rec.d, lam_d, s)
            hist = gH + op.adj_damage_history(rec.d, lam_d, s)
            gd_next = torch.where(rec.lower, gd + fd + system.apply(lam_d), 0.0)
            gH_next = torch.where(Q, 0.0, hist)
        self.stats = stats
        return grad


class _ReactionFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, Gc, curve):
        R, tape = curve.forward(Gc.detach().contiguous())
        ctx.curve, ctx.tape = curve, tape
        return R

    @staticmethod
    def backward(ctx, grad_R):
        grad = ctx.curve.backward(ctx.tape, grad_R.detach())
        ctx.tape = None
        return grad, None


def reaction_curve(Gc: torch.Tensor, curve: ReactionCurve) -> torch.Tensor:
    """Differentiable R(Gc): ``torch.autograd`` calls the reverse sweep of ``curve``."""
    return _ReactionFunction.apply(Gc, curve)
