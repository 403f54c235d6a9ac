"""Staggered alternate minimisation per macro-strain increment on the voxel cell.

Mirrors ``solid3d/staggered.py`` (StaggeredSolver: ``_attempt`` 1259-1517,
``_damage`` 753-802, ``_change`` 1097-1121, ``_predictor`` 1207-1213, ``advance`` 1215-1257,
``advance_increment`` 1523-1582, projected Anderson mixing 290-374) without interface
stress, for the AT2 law and for the PF-CZM law (``_damage_pfczm`` 843-1096, the projected Newton solve), with or
without the crack closure (``_closure_mechanics`` 698-751 with ``_energy_line_search``
589-626).  The linear mechanics solves are the device-resident constrained PCG (for AT2
without closure with the degradation formed in-kernel from the nodal damage); the damage
subproblems are the device box-KKT active set.
"""
from __future__ import annotations

import dataclasses
import math
from collections import deque
from dataclasses import dataclass, field
from time import perf_counter

import numpy as np
import torch

from .solid3d import (F64, H8ScalarBoxKKTConvergenceError, LoadCase, PCGBreakdownError,
                      PeriodicVoxelPhaseFieldOperator, solve_h8_scalar_box_kkt, solve_h8_scalar_dirichlet_pcg)

_MIN_LINEAR_TOLERANCE = 4.0 * float(np.finfo(np.float64).eps)  # staggered.py:97
_TINY = float(np.finfo(np.float64).tiny)                        # staggered.py:96
INTERFACE_MAX_SOLVES = 50           # staggered.py:99
CLOSURE_PATTERN_TOLERANCE = 1.0e-8  # staggered.py:105
LINE_SEARCH_ARMIJO = 1.0e-4         # staggered.py:109
LINE_SEARCH_ENERGY_RTOL = 1.0e-12   # staggered.py:112
LINE_SEARCH_MAX_HALVINGS = 30       # staggered.py:115
PFCZM_ARMIJO = 1.0e-4               # staggered.py:118
PFCZM_ENERGY_RTOL = 1.0e-12         # staggered.py:121
PFCZM_MAX_HALVINGS = 30             # staggered.py:124
PFCZM_MAX_DOUBLINGS = 6             # staggered.py:127
PFCZM_INNER_KKT_FRACTION = 0.1      # staggered.py:130
ANDERSON_REGULARISATION = 1.0e-10   # staggered.py:133
ANDERSON_ENERGY_RTOL = 1.0e-12      # staggered.py:136


@dataclass(frozen=True)
class SolverConfig:
    """The settings of ``solid3d/model.py:SolverConfig`` this route reads."""
    stagger_tol: float = 1.0e-4
    max_stagger: int = 2000
    min_stagger: int = 2
    pcg_rtol_mech: float = 1.0e-8
    pcg_maxit_mech: int = 20000
    check_every: int = 10
    damage_linear_tol: float = 1.0e-10
    damage_kkt_tol: float = 1.0e-8
    damage_maxit: int = 4000
    damage_max_active_set: int = 64
    on_stagger_failure: str = "cutback"
    max_cutbacks: int = 1
    final_mechanics_resolve: bool = True
    damaged_threshold: float = 0.95
    preconditioner: str = "jacobi"
    pfczm_max_newton: int = 100
    irreversibility: str = "box_kkt"  # or "history_only": unconstrained damage solve, bound violations counted
    history: str = "max"              # H = max(H_prev_step, Y), or "current": H = Y
    anderson_depth: int = 0           # projected Anderson mixing of the damage fixed point when > 0
    anderson_start: int = 2           # plain staggered iterations before the first mix

    def __post_init__(self):
        """The checks of model.py:830-857 for the settings read here."""
        for key, choices in (("irreversibility", ("box_kkt", "history_only")), ("history", ("max", "current")),
                             ("on_stagger_failure", ("cutback", "raise"))):
            if getattr(self, key) not in choices:
                raise ValueError(f"{key} must be one of {choices}, got {getattr(self, key)!r}")
        depth, start = self.anderson_depth, self.anderson_start
        if isinstance(depth, bool) or int(depth) != depth or int(depth) < 0:
            raise ValueError(f"anderson_depth must be an integer >= 0, got {depth!r}")
        if isinstance(start, bool) or int(start) != start or int(start) < 2:
            raise ValueError(f"anderson_start must be an integer >= 2 (a mix needs two staggered residuals), "
                             f"got {start!r}")
        object.__setattr__(self, "anderson_depth", int(depth))
        object.__setattr__(self, "anderson_start", int(start))


@dataclass(eq=False)
class SolverState:
    """Accepted solution (solid3d/staggered.py:167-218)."""
    step: int
    x: torch.Tensor
    d: torch.Tensor
    H: torch.Tensor
    macro_strain: torch.Tensor
    macro_stress: torch.Tensor
    external_work: float
    d_prev_step: torch.Tensor
    increment: int = 0
    active: tuple | None = None  # (lower, upper) active masks of the accepted damage solve

    def clone(self) -> "SolverState":
        act = None if self.active is None else tuple(a.clone() for a in self.active)
        return SolverState(self.step, self.x.clone(), self.d.clone(), self.H.clone(), self.macro_strain.clone(),
                           self.macro_stress.clone(), float(self.external_work), self.d_prev_step.clone(),
                           self.increment, act)


@dataclass
class IncrementRecord:
    step: int = 0
    increment: int = 0
    macro_strain: tuple = (0.0,) * 6
    macro_stress: tuple = (0.0,) * 6
    reaction_force: float = 0.0
    applied_disp: float = 0.0
    elastic: float = 0.0
    fracture: float = 0.0
    external: float = 0.0
    balance_residual_rel: float = 0.0
    max_damage: float = 0.0
    mean_damage: float = 0.0
    max_H: float = 0.0
    stagger_iters: int = 0
    pcg_iters_mech: int = 0
    pcg_iters_pf: int = 0
    active_set_iters: int = 0
    cutbacks: int = 0
    stagger_change: float = 0.0
    hill_mandel_rel: float = 0.0
    wall_s: float = 0.0
    mech_s: float = 0.0
    damage_s: float = 0.0
    history_s: float = 0.0
    newton_iters_pf: int | None = None   # PF-CZM: projected Newton iterations of the increment
    closure_solves: int | None = None    # crack closure: linear mechanics solves of the increment
    kkt_ok: bool = True                  # the last damage solve met its bounds (history_only: no violation)
    anderson_accepted: int | None = None  # Anderson mixing: mixed iterates kept / discarded
    anderson_rejected: int | None = None


class StaggerConvergenceError(RuntimeError):
    def __init__(self, message, *, state=None, target=None, increment=None, cutbacks=0, reason="", records=()):
        super().__init__(message)
        self.state, self.target, self.increment = state, target, increment
        self.cutbacks, self.reason, self.records = cutbacks, reason, records


class _AttemptFailure(RuntimeError):
    """staggered.py:273-285: a failed attempt and the iterations its solves had spent."""

    def __init__(self, message, *, pcg_mech: int = 0, pcg_pf: int = 0, active_set: int = 0):
        super().__init__(message)
        self.pcg_mech, self.pcg_pf, self.active_set = int(pcg_mech), int(pcg_pf), int(active_set)


class ProjectedAnderson:
    """Type II Anderson mixing of the damage fixed point d = G(d), projected onto the box of
    the damage solve (staggered.py:290-374; Walker and Ni 2011).

    The last depth + 1 pairs (d_i, G(d_i)) are kept.  ``mix`` takes the residuals
    f_i = G(d_i) - d_i, the differences dF (of residuals) and dG (of outputs) of neighbouring
    pairs, scales each column of dF to unit length (columns of zero length are dropped),
    solves (A^T A + rho I) beta = A^T f_k for the scaled A with rho = ANDERSON_REGULARISATION,
    and returns clamp(G(d_k) - dG gamma, lower, upper) with gamma = beta / scale.  None when
    fewer than two pairs are stored, when no column remains, or when the solve fails or gives
    a non-finite iterate."""

    def __init__(self, depth: int, *, regularisation: float = ANDERSON_REGULARISATION):
        if isinstance(depth, bool) or int(depth) != depth or int(depth) < 1:
            raise ValueError(f"depth must be an integer >= 1, got {depth!r}")
        if not (math.isfinite(regularisation) and regularisation >= 0.0):
            raise ValueError(f"regularisation must be finite and >= 0, got {regularisation!r}")
        self.depth, self.regularisation = int(depth), float(regularisation)
        self._pairs: deque = deque(maxlen=self.depth + 1)

    def __len__(self):
        return len(self._pairs)

    def reset(self):
        self._pairs.clear()

    def push(self, d_in: torch.Tensor, d_out: torch.Tensor):
        if d_in.shape != d_out.shape or d_in.dim() != 1:
            raise ValueError("d_in and d_out must be vectors of the same length")
        self._pairs.append((d_in.detach().clone(), d_out.detach().clone()))

    def mix(self, lower, upper=1.0):
        if len(self._pairs) < 2:
            return None
        outs = [o for _, o in self._pairs]
        res = [o - i for i, o in self._pairs]
        dF = torch.stack([b - a for a, b in zip(res[:-1], res[1:])], dim=1)
        dG = torch.stack([b - a for a, b in zip(outs[:-1], outs[1:])], dim=1)
        scale = torch.linalg.vector_norm(dF, dim=0)
        keep = scale > 0.0
        if not bool(keep.any()):
            return None
        dF, dG, scale = dF[:, keep], dG[:, keep], scale[keep]
        A = dF / scale
        normal = A.T @ A
        normal.diagonal().add_(self.regularisation)
        try:
            beta = torch.linalg.solve(normal, A.T @ res[-1])
        except RuntimeError:
            return None
        mixed = outs[-1] - dG @ (beta / scale)
        if not bool(torch.isfinite(mixed).all()):
            return None
        mixed = torch.maximum(mixed, lower) if isinstance(lower, torch.Tensor) else mixed.clamp(min=float(lower))
        mixed = torch.minimum(mixed, upper) if isinstance(upper, torch.Tensor) else mixed.clamp(max=float(upper))
        return mixed


class StaggeredSolver:
    """AT2 / PF-CZM staggered solver of one voxel cell under mixed macro control.

    ``iterate_observer`` (a callable, or None) receives one dict per staggered iterate of
    the attempts whose step is in ``observe_steps`` (staggered.py:1180-1195): step, attempt,
    increment, cutbacks, target, iteration, x, g, d, change, input_mixed, discarded."""

    def __init__(self, op: PeriodicVoxelPhaseFieldOperator, load: LoadCase | None = None,
                 cfg: SolverConfig | None = None, preconditioner=None):
        self.op = op
        self.load = op.load if load is None else load
        self.cfg = SolverConfig() if cfg is None else cfg
        self.device = op.device
        self._mask = op.mech_fixed_mask()
        self._n_u = 3 * op.n_red
        self._controlled = int(self.load.controlled)
        self._max_abs_target = max(abs(t) for t in self.load.targets)
        self._u_floor = min(op.grid.h) * math.sqrt(self._n_u)
        lengths = op.grid.lengths
        axis = self._controlled if self._controlled < 3 else 0
        others = tuple(lengths[a] for a in range(3) if a != axis)
        self._load_axis, self._load_length, self._face_lengths = axis, lengths[axis], others
        if op.has_closure and self.cfg.anderson_depth > 0:  # staggered.py:445-446
            raise ValueError("Anderson mixing (anderson_depth > 0) is not available with crack_closure 'axes'")
        self.preconditioner = preconditioner  # optional callable building an MG preconditioner
        self.last_state: SolverState | None = None
        self.n_attempts = 0
        self.iterate_observer = None
        self.observe_steps: frozenset = frozenset()
        self.last_bound_violations: dict | None = None  # history_only: the last solve's violations

    def initial_state(self) -> SolverState:
        op = self.op
        z = lambda *s: torch.zeros(s, dtype=F64, device=self.device)
        return SolverState(0, z(op.n_mech), z(op.n_red), z(op.n_elements, 8), z(6), z(6), 0.0, z(op.n_red), 0)

    # -- pieces ----------------------------------------------------------------------
    def _predictor(self, state, previous, target):
        """staggered.py:1207-1213."""
        x = state.x * (target / previous) if previous != 0.0 else torch.zeros_like(state.x)
        x[self._n_u + self._controlled] = target
        return x

    def _linear_mechanics(self, g, values, x0, closure=None):
        """staggered.py:550-587: one constrained PCG solve at fixed g (and closure pattern)."""
        op, cfg = self.op, self.cfg
        diag = op.mech_diagonal(g, closure=closure)
        try:
            x, info = op.solve_mechanics(g=g, closure=closure, fixed_mask=self._mask, fixed_values=values, diag=diag,
                                         x0=x0, rtol=cfg.pcg_rtol_mech, maxiter=cfg.pcg_maxit_mech,
                                         check_every=cfg.check_every)
        except PCGBreakdownError as exc:
            raise _AttemptFailure(f"mechanics PCG breakdown: {exc}") from exc
        if not info.converged:
            raise _AttemptFailure(f"mechanics PCG did not converge in {cfg.pcg_maxit_mech} iterations "
                                  f"(relative residual {info.relative_residual:.3e} > {cfg.pcg_rtol_mech:.3e})",
                                  pcg_mech=info.iterations)
        return x, info

    def _energy_line_search(self, x, y, g):
        """staggered.py:589-626: halve the step from x towards the pattern solution y until the
        degraded energy (each point at its own pattern) passes the Armijo test with the
        round-off allowance; alpha = 1 returns y itself."""
        op = self.op
        p = y - x
        energy0 = float(op.elastic_energy(x, g))
        force = op.mech_action(x, g)
        free = ~self._mask
        slope = min(float(torch.dot(force[free], p[free])), 0.0)
        allowance = LINE_SEARCH_ENERGY_RTOL * abs(energy0)
        alpha = 1.0
        for _ in range(LINE_SEARCH_MAX_HALVINGS + 1):
            trial = y if alpha == 1.0 else x + alpha * p
            energy = float(op.elastic_energy(trial, g))
            if not math.isfinite(energy):
                raise _AttemptFailure("mechanics energy is not finite in the line search")
            if energy <= energy0 + LINE_SEARCH_ARMIJO * alpha * slope + allowance:
                return trial, alpha
            alpha *= 0.5
        raise _AttemptFailure(f"the energy line search of the closure mechanics solve found no decrease in "
                              f"{LINE_SEARCH_MAX_HALVINGS} halvings")

    def _closure_mechanics(self, g, values, x0):
        """staggered.py:698-751: linear solves at the fixed closure pattern, each followed by
        the energy line search, until no bit flips materially (CLOSURE_PATTERN_TOLERANCE)
        after a full step."""
        op = self.op
        closure = op.closure_pattern(x0)
        iterations = damped = 0
        for solve in range(1, INTERFACE_MAX_SOLVES + 1):
            try:
                y, info = self._linear_mechanics(g, values, x0, closure)
                iterations += info.iterations
                x, alpha = self._energy_line_search(x0, y, g)
                damped += int(alpha < 1.0)
            except _AttemptFailure as exc:
                exc.pcg_mech += iterations
                raise
            closure, changed = op.closure_pattern_change(x, closure, g, tolerance=CLOSURE_PATTERN_TOLERANCE)
            if not changed and alpha == 1.0:
                self.last_interface_solves = solve
                self.last_line_search_steps = damped
                return x, dataclasses.replace(info, iterations=iterations)
            x0 = x
        raise _AttemptFailure(f"the closure pattern (crack_closure 'axes') did not settle in {INTERFACE_MAX_SOLVES} "
                              f"mechanics solves", pcg_mech=iterations)

    def _mechanics(self, d, values, x0):
        """staggered.py:628-696 (mechanics at fixed damage): one linear solve, or the closure
        pattern iteration."""
        op, cfg = self.op, self.cfg
        if op.has_closure or op.is_pfczm:
            g = op.degradation(d)
            if op.has_closure:
                x, info = self._closure_mechanics(g, values, x0)
            else:
                x, info = self._linear_mechanics(g, values, x0)
            return x, info, g
        g = op.degradation(d)
        diag = op.mech_diagonal(g)
        try:
            if self.preconditioner is not None:
                x, info = self.preconditioner.solve(d=d, g=g, fixed_mask=self._mask, fixed_values=values, diag=diag,
                                                    x0=x0, rtol=cfg.pcg_rtol_mech, maxiter=cfg.pcg_maxit_mech,
                                                    check_every=cfg.check_every)
            else:
                x, info = op.solve_mechanics(d=d, fixed_mask=self._mask, fixed_values=values, diag=diag, x0=x0,
                                             rtol=cfg.pcg_rtol_mech, maxiter=cfg.pcg_maxit_mech,
                                             check_every=cfg.check_every)
        except PCGBreakdownError as exc:
            raise _AttemptFailure(f"mechanics PCG breakdown: {exc}") from exc
        if not info.converged:
            raise _AttemptFailure(f"mechanics PCG did not converge in {cfg.pcg_maxit_mech} iterations "
                                  f"(relative residual {info.relative_residual:.3e} > {cfg.pcg_rtol_mech:.3e})")
        return x, info, g

    def _damage(self, H, lower, initial):
        """staggered.py:753-802 (box_kkt irreversibility)."""
        if self.op.is_pfczm:
            return self._damage_pfczm(H, lower, initial)
        cfg = self.cfg
        system = self.op.damage_system(H)
        r0 = float(torch.linalg.vector_norm(system.rhs - system.apply(initial)))
        if not math.isfinite(r0):
            raise _AttemptFailure("damage residual is not finite")
        tol = min(cfg.damage_linear_tol, 0.1 * cfg.damage_kkt_tol / max(r0, 1.0))
        tol = max(tol, _MIN_LINEAR_TOLERANCE)
        if cfg.irreversibility == "history_only":
            return self._damage_unconstrained(system, lower, initial, tol)
        try:
            res = solve_h8_scalar_box_kkt(system, lower_bound=lower, upper_bound=1.0, initial=initial,
                                          linear_tolerance=tol, kkt_tolerance=cfg.damage_kkt_tol,
                                          max_linear_iterations=cfg.damage_maxit,
                                          max_active_set_iterations=cfg.damage_max_active_set)
        except H8ScalarBoxKKTConvergenceError as exc:
            raise _AttemptFailure(f"damage box-KKT solve failed: {exc}") from exc
        except (RuntimeError, ValueError) as exc:
            raise _AttemptFailure(f"damage PCG breakdown: {exc}") from exc
        self._last_active = (res.lower_active, res.upper_active)
        return res.solution, res.total_pcg_iterations, res.active_set_iterations, res.kkt.satisfies_kkt

    def _damage_unconstrained(self, system, lower, initial, tol):
        """irreversibility='history_only' (staggered.py:804-839): one PCG solve with no bound;
        the violations of d_prev <= d <= 1 are counted, not enforced."""
        fixed = torch.zeros(self.op.n_red, dtype=torch.bool, device=self.device)
        try:
            solve = solve_h8_scalar_dirichlet_pcg(system, fixed=fixed, values=torch.zeros_like(initial),
                                                  initial=initial, tolerance=tol,
                                                  max_iterations=self.cfg.damage_maxit)
        except (RuntimeError, ValueError) as exc:
            raise _AttemptFailure(f"damage PCG breakdown: {exc}") from exc
        if not solve.krylov.converged:
            raise _AttemptFailure(f"damage PCG did not converge in {self.cfg.damage_maxit} iterations",
                                  pcg_pf=solve.krylov.iterations)
        d = solve.solution
        stats = torch.stack(((d < lower).sum().to(F64), (d < 0.0).sum().to(F64), (d > 1.0).sum().to(F64),
                             (lower - d).clamp_min(0.0).max(), (d - 1.0).clamp_min(0.0).max())).cpu().tolist()
        self.last_bound_violations = {"below_previous": int(stats[0]), "below_zero": int(stats[1]),
                                      "above_one": int(stats[2]), "max_decrease": stats[3],
                                      "max_excess_over_one": stats[4]}
        self._last_active = None
        ok = int(stats[0]) == 0 and int(stats[1]) == 0 and int(stats[2]) == 0
        return d, int(solve.krylov.iterations), 0, ok

    def _damage_pfczm(self, H, lower, initial):
        """staggered.py:843-1096 (absolute-field QP): projected Newton on the PF-CZM damage
        energy over lower <= d <= 1.  Each iteration: stop when the scaled projected gradient
        is at most damage_kkt_tol in the maximum norm; else solve the box-QP of the Newton
        tangent holding the bound nodes whose gradient points outward, take the Armijo step
        with halving, and after a full step that is not yet converged double it along the
        projected path while the energy decreases."""
        op, cfg = self.op, self.cfg
        tol = cfg.damage_kkt_tol
        inner_tol = PFCZM_INNER_KKT_FRACTION * tol
        scale = op.damage_scale

        def projected_residual(field):
            grad = op.pfczm_damage_gradient(field, H, scale)
            pg = torch.where(field <= lower, grad.clamp(max=0.0), grad)
            pg = torch.where(field >= 1.0, pg.clamp(min=0.0), pg)
            return float(pg.abs().max())

        d = initial
        energy = float(op.pfczm_damage_energy(d, H))
        if not math.isfinite(energy):
            raise _AttemptFailure("damage PF-CZM energy is not finite")
        pcg = sets = 0
        ones = torch.ones_like(d)
        residual = math.inf
        iteration = 0
        for iteration in range(cfg.pfczm_max_newton + 1):
            system, gradient = op.pfczm_newton_system(d, H)
            at_lower = d <= lower
            at_upper = d >= 1.0
            projected = torch.where(at_lower, gradient.clamp(max=0.0), gradient)
            projected = torch.where(at_upper, projected.clamp(min=0.0), projected)
            residual = float(projected.abs().max())
            if not math.isfinite(residual):
                raise _AttemptFailure("damage PF-CZM gradient is not finite", pcg_pf=pcg, active_set=sets)
            if residual <= tol:
                self.last_newton_iters = iteration
                self._newton_count += iteration
                self._last_active = None
                return d, pcg, sets, True
            if iteration == cfg.pfczm_max_newton:
                break
            hold_lower = at_lower & (gradient > tol)
            hold_upper = at_upper & (gradient < -tol)
            upper_k = torch.where(hold_lower, lower, ones)
            lower_k = torch.where(hold_upper, ones, lower)
            r0 = float(torch.linalg.vector_norm(gradient))
            tolerance = max(min(cfg.damage_linear_tol, 0.1 * inner_tol / max(r0, 1.0)), _MIN_LINEAR_TOLERANCE)
            try:
                result = solve_h8_scalar_box_kkt(system, lower_bound=lower_k, upper_bound=upper_k, initial=d,
                                                 linear_tolerance=tolerance, kkt_tolerance=inner_tol,
                                                 max_linear_iterations=cfg.damage_maxit,
                                                 max_active_set_iterations=cfg.damage_max_active_set)
            except H8ScalarBoxKKTConvergenceError as exc:
                raise _AttemptFailure(f"damage PF-CZM Newton subproblem failed at Newton iteration "
                                      f"{iteration + 1}: {exc}", pcg_pf=pcg, active_set=sets) from exc
            except (RuntimeError, ValueError) as exc:
                raise _AttemptFailure(f"damage PF-CZM Newton subproblem PCG breakdown: {exc}") from exc
            pcg += int(result.total_pcg_iterations)
            sets += int(result.active_set_iterations)
            target = result.solution
            step = target - d
            slope = float((gradient * step).sum()) / scale
            t = 1.0
            for _ in range(PFCZM_MAX_HALVINGS + 1):
                trial = target if t == 1.0 else torch.minimum(torch.maximum(d + t * step, lower), ones)
                trial_energy = float(op.pfczm_damage_energy(trial, H))
                if math.isfinite(trial_energy) and trial_energy <= (
                        energy + PFCZM_ARMIJO * t * slope + PFCZM_ENERGY_RTOL * abs(energy)):
                    break
                t *= 0.5
            else:
                raise _AttemptFailure(f"damage PF-CZM line search failed at Newton iteration {iteration + 1} "
                                      f"(no decrease after {PFCZM_MAX_HALVINGS} halvings)", pcg_pf=pcg,
                                      active_set=sets)
            if t == 1.0 and not projected_residual(trial) <= tol:
                for _ in range(PFCZM_MAX_DOUBLINGS):
                    t *= 2.0
                    further = torch.minimum(torch.maximum(d + t * step, lower), ones)
                    if torch.equal(further, trial):
                        break
                    further_energy = float(op.pfczm_damage_energy(further, H))
                    if not (math.isfinite(further_energy) and further_energy < trial_energy):
                        break
                    trial, trial_energy = further, further_energy
            d = trial
            energy = trial_energy
        raise _AttemptFailure(f"damage PF-CZM Newton did not converge in {cfg.pfczm_max_newton} iterations "
                              f"(projected gradient {residual:.3e} > damage_kkt_tol {tol:.3e})", pcg_pf=pcg,
                              active_set=sets)

    def _change(self, x_new, x_old, d_new, d_old, step_size):
        """staggered.py:1097-1121."""
        n = self._n_u
        parts = torch.stack([torch.linalg.vector_norm(x_new[:n] - x_old[:n]), torch.linalg.vector_norm(x_new[:n]),
                             (x_new[n:] - x_old[n:]).abs().max(), x_new[n:].abs().max(),
                             (d_new - d_old).abs().max()]).cpu()
        du, u, dE, E_inf, dd = (float(v) for v in parts)
        if not all(math.isfinite(v) for v in (du, u, dE, E_inf, dd)):
            raise _AttemptFailure("staggered iterate is not finite")
        scale_E = max(E_inf, step_size)
        u_ref = max(u, scale_E * self._u_floor)
        du_rel = du / u_ref if u_ref > 0.0 else 0.0
        dE_rel = dE / scale_E if scale_E > 0.0 else 0.0
        return max(du_rel, dE_rel, dd)

    def snapshot(self, state: SolverState):
        """CPU fields of ``state`` (solid3d/staggered.py:517-522); the time value is the
        controlled macro strain."""
        return self.op.field_snapshot(state.step, float(state.macro_strain[self._controlled]), state.x, state.d,
                                      state.H)

    def _observe(self, state, **row):
        """Hand one staggered iterate to ``iterate_observer`` when its step is observed."""
        step = int(state.step) + 1
        if self.iterate_observer is None or step not in self.observe_steps:
            return
        self.iterate_observer({"step": step, "attempt": int(self.n_attempts), **row})

    def _total_energy(self, x, d, g=None):
        """(E_el(x, g(d)), Gamma(d)) (staggered.py:1201-1205)."""
        g = self.op.degradation(d) if g is None else g
        return float(self.op._stress_sums(x, g)[6]), self.fracture_energy(d)

    def fracture_energy(self, d) -> float:
        """Gamma(d) (operator.py:1511-1544, total): 1/2 sum_e Gc d_e.(M/ell + ell L) d_e for AT2,
        the PF-CZM surface functional for PF-CZM (operator.py:611-631)."""
        return self.op.fracture_energy(d)["total"]

    # -- one attempt -----------------------------------------------------------------
    def _attempt(self, state, target, previous, increment, cutbacks, start):
        op, cfg = self.op, self.cfg
        values = op.mech_fixed_values(target)
        step_size = abs(target - previous)
        x_old = self._predictor(state, previous, target)
        d_old = state.d.clone()
        lower = state.d
        mech_s = damage_s = history_s = 0.0
        pcg_mech = pcg_pf = active_set = 0
        converged, change, H, info = False, math.inf, state.H, None
        closure_solves = 0
        self._newton_count = 0
        kkt_ok = True
        # Anderson mixing: a fresh history for every attempt (staggered.py:1283-1289)
        anderson = ProjectedAnderson(cfg.anderson_depth) if cfg.anderson_depth > 0 else None
        fallback = None  # (d, x, residual) of the plain iterate a pending mixed input replaced
        best_energy = math.inf
        n_accepted = n_rejected = 0
        where = dict(increment=increment, cutbacks=cutbacks, target=target)
        k = 0
        for k in range(1, cfg.max_stagger + 1):
            t0 = perf_counter()
            try:
                x_new, info, g = self._mechanics(d_old, values, x_old)
                pcg_mech += info.iterations
                if op.has_closure:
                    closure_solves += self.last_interface_solves
                t1 = perf_counter()
                Y = op.driving_force(x_new)
                H = torch.maximum(state.H, Y) if cfg.history == "max" else Y
                t2 = perf_counter()
                d_new, iters, sets, kkt_ok = self._damage(H, lower, d_old)
                pcg_pf += iters
                active_set += sets
                t3 = perf_counter()
                change = self._change(x_new, x_old, d_new, d_old, step_size)
            except _AttemptFailure as exc:
                if fallback is None:
                    raise
                # the iteration from a mixed input failed: back to the plain iterate; the
                # iterations its solves had spent count (staggered.py:1311-1341)
                n_rejected += 1
                pcg_mech += exc.pcg_mech
                pcg_pf += exc.pcg_pf
                active_set += exc.active_set
                mech_s += perf_counter() - t0
                self._observe(state, **where, iteration=k, x=None, g=None, d=d_old, change=float("nan"),
                              input_mixed=True, discarded=True)
                d_old, x_old, _ = fallback
                fallback = None
                anderson.reset()
                continue
            mech_s += t1 - t0
            history_s += t2 - t1
            damage_s += t3 - t2
            rejected = False
            if anderson is not None:
                residual = float(torch.linalg.vector_norm(d_new - d_old))
                elastic_k, fracture_k = self._total_energy(x_new, d_new)
                energy = elastic_k + fracture_k
                if fallback is not None:
                    # the mixed input is judged by its own iteration: its end energy must not
                    # exceed the lowest of the attempt, nor its residual that of the iteration
                    # that produced the replaced plain iterate (staggered.py:1354-1378)
                    rejected = (energy > best_energy + ANDERSON_ENERGY_RTOL * abs(best_energy)
                                or residual > fallback[2])
                    if rejected:
                        n_rejected += 1
                    else:
                        n_accepted += 1
                if not rejected:
                    best_energy = min(best_energy, energy)
            self._observe(state, **where, iteration=k, x=x_new, g=g, d=d_new, change=float(change),
                          input_mixed=fallback is not None, discarded=bool(rejected))
            if rejected:
                d_old, x_old, _ = fallback
                fallback = None
                anderson.reset()
                continue
            fallback = None
            if change <= cfg.stagger_tol and k >= cfg.min_stagger:
                x_old, d_old = x_new, d_new
                converged = True
                break
            if anderson is not None:
                anderson.push(d_old, d_new)
                mixed = anderson.mix(lower) if k >= cfg.anderson_start else None
                if mixed is not None and not torch.equal(mixed, d_new):
                    fallback = (d_new, x_new, residual)
                    x_old, d_old = x_new, mixed
                    continue
            x_old, d_old = x_new, d_new
        if not converged:
            raise _AttemptFailure(f"staggered iteration did not converge in {cfg.max_stagger} iterations "
                                  f"(change {change:.3e} > stagger_tol {cfg.stagger_tol:.3e})")
        d = d_old
        t0 = perf_counter()
        if cfg.final_mechanics_resolve:
            x, info, g = self._mechanics(d, values, x_old)
            pcg_mech += info.iterations
            if op.has_closure:
                closure_solves += self.last_interface_solves
        else:
            x, g = x_old, op.degradation(d)
        mech_s += perf_counter() - t0
        sums = op._stress_sums(x, g)
        sigma = sums[:6] / op.volume
        energy = float(sums[6])
        E = x[self._n_u:]
        mean_work = 2.0 * energy / op.volume
        macro_work = float(torch.dot(sigma, E))
        scale = max(abs(mean_work), abs(macro_work))
        hill_rel = abs(mean_work - macro_work) / scale if scale > 0.0 else 0.0
        fracture = self.fracture_energy(d)
        elastic = 0.5 * op.volume * mean_work
        stats = torch.stack([d.max(), d.mean(), H.max()]).cpu().numpy()
        macro_strain = tuple(float(v) for v in E.cpu())
        macro_stress = tuple(float(v) for v in sigma.cpu())
        if not np.isfinite(macro_stress).all() or not math.isfinite(elastic) or not math.isfinite(fracture):
            raise _AttemptFailure("accepted state has non-finite stress, strain or energy")
        c = self._controlled
        sigma_prev = float(state.macro_stress[c])
        external = state.external_work + op.volume * 0.5 * (macro_stress[c] + sigma_prev) * (target - previous)
        residual = external - (elastic + fracture)
        completes = 1 <= increment <= len(self.load.targets) and target == self.load.targets[increment - 1]
        new_state = SolverState(state.step + 1, x, d, H, E.clone(), sigma, external, state.d,
                                increment if completes else state.increment, self._last_active)
        fa, fb = self._face_lengths
        rec = IncrementRecord(step=new_state.step, increment=increment, macro_strain=macro_strain,
                              macro_stress=macro_stress, reaction_force=macro_stress[self._load_axis] * fa * fb,
                              applied_disp=macro_strain[self._load_axis] * self._load_length, elastic=elastic,
                              fracture=fracture, external=external,
                              balance_residual_rel=residual / max(abs(external), _TINY),
                              max_damage=float(stats[0]), mean_damage=float(stats[1]), max_H=float(stats[2]),
                              stagger_iters=k, pcg_iters_mech=pcg_mech, pcg_iters_pf=pcg_pf,
                              active_set_iters=active_set, cutbacks=cutbacks, stagger_change=change,
                              hill_mandel_rel=hill_rel, wall_s=perf_counter() - start, mech_s=mech_s,
                              damage_s=damage_s, history_s=history_s,
                              newton_iters_pf=self._newton_count if op.is_pfczm else None,
                              closure_solves=closure_solves if op.has_closure else None, kkt_ok=bool(kkt_ok),
                              anderson_accepted=n_accepted if anderson is not None else None,
                              anderson_rejected=n_rejected if anderson is not None else None)
        return new_state, rec

    def advance(self, state, target, *, increment=None, cutbacks=0):
        """staggered.py:1215-1257."""
        target = float(target)
        previous = float(state.macro_strain[self._controlled])
        if target == previous:
            raise ValueError(f"target {target!r} equals the current controlled strain")
        increment = state.increment + 1 if increment is None else int(increment)
        self.n_attempts += 1
        start = perf_counter()
        try:
            new_state, rec = self._attempt(state, target, previous, increment, int(cutbacks), start)
        except _AttemptFailure as exc:
            raise StaggerConvergenceError(f"increment {increment} (cutback level {cutbacks}) failed: {exc}",
                                          state=state, target=target, increment=increment, cutbacks=cutbacks,
                                          reason=str(exc)) from exc
        self.last_state = new_state
        return new_state, rec

    def advance_increment(self, state, target, *, increment=None, on_increment=None):
        """Recursive half-step cutback (staggered.py:1523-1582)."""
        increment = state.increment + 1 if increment is None else int(increment)
        accepted, latest = [], [state]

        def solve(current, end, depth):
            try:
                new_state, rec = self.advance(current, end, increment=increment, cutbacks=depth)
            except StaggerConvergenceError as exc:
                if self.cfg.on_stagger_failure == "raise" or depth >= self.cfg.max_cutbacks:
                    raise StaggerConvergenceError(f"increment {increment} failed at cutback level {depth}: "
                                                  f"{exc.reason}", state=latest[0], target=end,
                                                  increment=increment, cutbacks=depth, reason=exc.reason,
                                                  records=tuple(accepted)) from exc
                begin = float(current.macro_strain[self._controlled])
                middle = 0.5 * (begin + end)
                current = solve(current, middle, depth + 1)
                return solve(current, end, depth + 1)
            accepted.append(rec)
            latest[0] = new_state
            if on_increment is not None:
                on_increment(new_state, rec)
            return new_state

        final = solve(state, float(target), 0)
        if final.increment != increment:
            final = dataclasses.replace(final, increment=increment)
            self.last_state = final
        return final, accepted

    def run(self, *, start_state=None, max_increments=None, on_increment=None):
        """Follow the load path (staggered.py:1584-1627)."""
        state = self.initial_state() if start_state is None else start_state
        records = []
        stop = len(self.load.targets) if max_increments is None else min(len(self.load.targets),
                                                                          state.increment + max_increments)
        while state.increment < stop:
            state, recs = self.advance_increment(state, self.load.targets[state.increment],
                                                 on_increment=on_increment)
            records.extend(recs)
        self.last_state = state
        return records
