"""Oracle (numpy, float64) for the PF-CZM damage law and the crack closure of the voxel path.

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

Restates (paths relative to ``src/phast``):
* ``physics/pfczm.py:17-165``: ``a1``, the softening parameters, Wu's rational degradation
  and its first two derivatives;
* ``solid3d/splits.py:424-560``: the isotropic closure rows ``(1, e_x, e_y, e_z)``, their
  moduli ``(lam, 2 mu, 2 mu, 2 mu)`` and the closure stress of a sign pattern;
* ``solid3d/operator.py``: the PF-CZM damage energy, gradient, Newton tangent weights,
  action and diagonal (272-583), the closure set-up of the isotropic group (631-690), the
  closure values, pattern and pattern change (755-876), the stress pass with the closure
  add-back (1039-1105) and the closure part of the Jacobi diagonal (1255-1300);
* ``solid3d/staggered.py``: the energy line search (589-626), the closure pattern
  iteration (698-751), the projected Newton damage solve (843-1096) and the staggered
  attempt (1259-1517).

The voxel geometry, the strain, the PCG, the Dirichlet PCG and the box-KKT active set are
those of :mod:`oracle.hex8`.
"""
from __future__ import annotations

import math

import numpy as np

from .hex8 import VoxelOperator, box_kkt, reference_voxel, scatter, solve_constrained

PFCZM_C_ALPHA = math.pi                 # physics/pfczm.py:6
TANGENT_FLOOR = 0.25                    # solid3d/operator.py:124
PATTERN_TOLERANCE = 1.0e-8              # solid3d/staggered.py:105
MAX_PATTERN_SOLVES = 50                 # staggered.py:99
LINE_SEARCH_ARMIJO = 1.0e-4             # staggered.py:109
LINE_SEARCH_ENERGY_RTOL = 1.0e-12       # staggered.py:112
LINE_SEARCH_MAX_HALVINGS = 30           # staggered.py:115
NEWTON_ARMIJO = 1.0e-4                  # staggered.py:118
NEWTON_ENERGY_RTOL = 1.0e-12            # staggered.py:121
NEWTON_MAX_HALVINGS = 30                # staggered.py:124
NEWTON_MAX_DOUBLINGS = 6                # staggered.py:127
INNER_KKT_FRACTION = 0.1                # staggered.py:130
MIN_LINEAR_TOLERANCE = 4.0 * float(np.finfo(np.float64).eps)  # staggered.py:97

CLOSURE_ROWS = np.array([[1.0, 1.0, 1.0, 0.0, 0.0, 0.0],
                         [1.0, 0.0, 0.0, 0.0, 0.0, 0.0],
                         [0.0, 1.0, 0.0, 0.0, 0.0, 0.0],
                         [0.0, 0.0, 1.0, 0.0, 0.0, 0.0]])  # splits.py:489-495


def softening_parameters(softening: str):
    """``(a2, a3)`` of a Wu softening law (physics/pfczm.py:55-63)."""
    if softening == "linear":
        return -0.5, 0.0
    if softening == "exponential":
        return 2.0 ** (5.0 / 3.0) - 3.0, 0.0
    raise ValueError(softening)


def pfczm_a1(E, Gc, ell, f_t, eta=0.0):
    """``4 E Gc / (pi ell f_t^2 (1 - eta))`` (physics/pfczm.py:17-32)."""
    return 4.0 * E * Gc / (math.pi * ell * f_t ** 2 * max(1.0 - float(eta), 1.0e-30))


def degradation_derivatives(d, a1, a2, a3, p=2, eta=0.0):
    """``g, g', g''`` of ``(1 - eta) A/(A + Q) + eta`` with ``A = (1 - d)^p`` and
    ``Q = a1 d (1 + a2 d + a2 a3 d^2)`` (physics/pfczm.py:79-130).  ``a1`` is broadcast
    over the trailing axis of ``d`` when it has one dimension fewer."""
    d = np.asarray(d, dtype=np.float64)
    a1 = np.asarray(a1, dtype=np.float64)
    while a1.ndim < d.ndim:
        a1 = a1[..., None]
    p = int(p) if float(p).is_integer() else float(p)
    omd = np.maximum(1.0 - d, 0.0)
    A = omd ** p
    Ap = -float(p) * omd ** (p - 1)
    App = np.full_like(d, 2.0) if p == 2 else float(p * (p - 1)) * omd ** (p - 2)
    Q = a1 * d * (1.0 + a2 * d + a2 * a3 * d * d)
    Qp = a1 * (1.0 + 2.0 * a2 * d + 3.0 * a2 * a3 * d * d)
    Qpp = a1 * (2.0 * a2 + 6.0 * a2 * a3 * d)
    S = np.maximum(A + Q, 1.0e-30)
    Nn = Ap * Q - A * Qp
    Np = App * Q - A * Qpp
    scale = 1.0 - eta
    return (scale * (A / S) + eta, scale * (Nn / (S * S)),
            scale * (Np / (S * S) - 2.0 * Nn * (Ap + Qp) / (S * S * S)))


class PfczmOperator:
    """The voxel operator with the PF-CZM law and (optionally) the crack closure.

    ``base`` carries the geometry, the stiffness bank, ``lam``/``mu``/``Gc``/``driver`` per
    voxel and ``kappa``, ``ell``.  ``E_ref`` and ``f_t`` per voxel give ``a1``
    (operator.py:317-383).  The closure has two groups (operator.py:631-690): the isotropic
    voxels with a damage driver (rows trace, e_x, e_y, e_z; moduli lam, 2 mu, 2 mu, 2 mu) and,
    when ``band = (voxel ids, rows (nb, 4, 6), moduli (nb, 4))`` is given, the smeared-laminate
    voxels with their own rows and moduli; pores and driver 'none' add nothing back."""

    def __init__(self, base: VoxelOperator, E_ref, f_t, *, softening="linear", p=2, closure=True, band=None):
        self.b = base
        self.kappa, self.ell = base.kappa, base.ell
        self.a2, self.a3 = softening_parameters(softening)
        self.p = p
        self.a1 = pfczm_a1(np.asarray(E_ref, dtype=np.float64), base.Gc, base.ell, np.asarray(f_t, dtype=np.float64),
                           base.kappa)
        bound = float(p) + self.a2                                      # physics/pfczm.py:66-76
        active = base.driver != 0
        if active.any() and float(self.a1[active].min()) < bound:
            raise ValueError("onset condition a1 >= p + a2 fails")
        _, _, _, B, M, L = reference_voxel(base.h)
        self.B = B
        self.M, self.L = M, L
        self.m = M.sum(axis=1)                                          # operator.py:321
        self.K_t = (2.0 / PFCZM_C_ALPHA) * self.ell * L                 # operator.py:322
        self.closure = bool(closure)
        E = base.E
        group = np.where(active, 1, 0).astype(np.int8)
        rows = np.broadcast_to(CLOSURE_ROWS, (E, 4, 6)).copy()
        lam = np.where(active, base.lam, 0.0)
        mu = np.where(active, base.mu, 0.0)
        k = np.stack((lam, 2.0 * mu, 2.0 * mu, 2.0 * mu), axis=-1)     # splits.py:498-503
        if band is not None:
            ids, b_rows, b_k = band
            ids = np.asarray(ids, dtype=np.int64)
            on = active[ids]
            group[ids] = np.where(on, 2, 0)
            rows[ids] = b_rows
            k[ids] = np.where(on[:, None], b_k, 0.0)
        self.cl_group = group
        self.cl_on = group != 0
        self.cl_rows = rows
        self.cl_k = k

    # -- law ------------------------------------------------------------------------
    def derivatives(self, d):
        d_q = d[self.b.nodes] @ self.b.N.T
        return degradation_derivatives(d_q, self.a1, self.a2, self.a3, self.p, self.kappa)

    def degradation(self, d):
        """operator.py:939-946 with the PF-CZM law."""
        return self.derivatives(d)[0]

    # -- mechanics with the closure add-back ----------------------------------------
    def closure_values(self, x):
        """``v . eps`` (E, 8, 4) with each voxel's rows (operator.py:786-816)."""
        return np.einsum("eqm,ekm->eqk", self.b.qp_strain(x), self.cl_rows)

    def stress_pass(self, x, g=None, pattern=None, assemble=True):
        """operator.py:1039-1105: ``g C eps + (1 - g) sigma_c`` (own pattern when ``pattern``
        is None); ``g`` None is undamaged and adds nothing back."""
        b = self.b
        eps = b.qp_strain(x)
        sigma = np.einsum("eqm,ekm->eqk", eps, b.bank[b.index])
        if g is not None:
            sigma = g[..., None] * sigma
            if self.closure:
                values = np.einsum("eqm,ekm->eqk", eps, self.cl_rows)
                active = values < 0.0 if pattern is None else pattern
                coef = np.where(active, values, 0.0) * self.cl_k[:, None, :]
                sigma = sigma + (1.0 - g)[..., None] * np.einsum("eqk,ekm->eqm", coef, self.cl_rows)
        weighted = sigma * b.w[None, :, None]
        r_E = weighted.sum(axis=(0, 1))
        energy = 0.5 * (weighted * eps).sum()
        r_u = scatter(b.dofs, weighted.reshape(-1, 48) @ b.B_flat.T, 3 * b.n_red) if assemble else None
        return r_u, r_E, energy

    def mech_action(self, x, g=None, pattern=None):
        r_u, r_E, _ = self.stress_pass(x, g, pattern)
        return np.concatenate((r_u, r_E))

    def macro_stress(self, x, g=None, pattern=None):
        return self.stress_pass(x, g, pattern, assemble=False)[1] / self.b.volume

    def elastic_energy(self, x, g=None, pattern=None):
        return self.stress_pass(x, g, pattern, assemble=False)[2]

    def mech_diagonal(self, g=None, pattern=None):
        """operator.py:1196-1300: the plain Jacobi diagonal plus, for every closure bit
        (all bits when ``pattern`` is None), ``sum_q w (1 - g) k (B_q^T v)^2`` and, in the
        macro rows, ``sum_q w (1 - g) k v_i^2``."""
        diag = self.b.mech_diagonal(g)
        if g is None or not self.closure:
            return diag
        n3 = 3 * self.b.n_red
        coef = (1.0 - g)[..., None] * self.cl_k[:, None, :] * self.b.w[None, :, None]
        if pattern is not None:
            coef = coef * pattern
        local = np.empty((self.b.E, 24))
        for s0 in range(0, self.b.E, 4096):
            sl = slice(s0, s0 + 4096)
            BtR = np.einsum("qia,eki->eqka", self.B, self.cl_rows[sl])
            local[sl] = np.einsum("eqk,eqka->ea", coef[sl], BtR ** 2)
        diag[:n3] += scatter(self.b.dofs, local, n3)
        diag[n3:] += np.einsum("eqk,eki->i", coef, self.cl_rows ** 2)
        return diag

    def closure_pattern(self, x):
        """operator.py:792-806."""
        return (self.closure_values(x) < 0.0) & self.cl_on[:, None, None]

    def closure_pattern_change(self, x, pattern, g, tolerance=PATTERN_TOLERANCE):
        """operator.py:808-841: the new pattern and whether, in either group, a flipped bit
        moves the stress by more than ``tolerance`` times that group's largest
        ``k |v . eps|``."""
        values = self.closure_values(x)
        new = (values < 0.0) & self.cl_on[:, None, None]
        strength = self.cl_k[:, None, :] * np.abs(values)
        changed = False
        for grp in (1, 2):
            sel = self.cl_group == grp
            if grp == 1:
                sel = np.ones_like(sel) if not (self.cl_group == 2).any() else self.cl_group != 2
            if not sel.any():
                continue
            flipped = new[sel] != pattern[sel]
            if not changed and flipped.any():
                jump = (1.0 - g[sel])[..., None] * strength[sel]
                changed = bool(float(jump[flipped].max()) > tolerance * float(strength[sel].max()))
        return new, changed

    # -- PF-CZM damage --------------------------------------------------------------
    def damage_energy(self, d, H=None):
        """operator.py:461-488."""
        b = self.b
        d_e = d[b.nodes]
        local = 2.0 * (d_e @ self.m) - ((d_e @ self.M) * d_e).sum(axis=-1)
        grad = ((d_e @ self.L) * d_e).sum(axis=-1)
        total = (b.Gc * (local / (PFCZM_C_ALPHA * self.ell) + self.ell / PFCZM_C_ALPHA * grad)).sum()
        if H is not None:
            total += ((H * b.w) * self.derivatives(d)[0]).sum()
        return float(total)

    def damage_gradient(self, d, H=None):
        """operator.py:490-512."""
        b = self.b
        d_e = d[b.nodes]
        y = b.Gc[:, None] * ((2.0 / (PFCZM_C_ALPHA * self.ell)) * (self.m - d_e @ self.M)
                             + (2.0 * self.ell / PFCZM_C_ALPHA) * (d_e @ self.L))
        if H is not None:
            y = y + ((H * b.w) * self.derivatives(d)[1]) @ b.N
        return scatter(b.nodes, y, b.n_red)

    def tangent_weights(self, d, H=None):
        """operator.py:514-538."""
        concave = (2.0 / (PFCZM_C_ALPHA * self.ell) * self.b.Gc)[:, None]
        floor = TANGENT_FLOOR * concave
        if H is None:
            return np.broadcast_to(floor, (self.b.E, 8)).copy()
        return np.maximum(H * self.derivatives(d)[2] - concave, floor)

    def tangent_action(self, f, W):
        """operator.py:540-552."""
        b = self.b
        f_e = f[b.nodes]
        y = b.Gc[:, None] * (f_e @ self.K_t) + ((W * b.w) * (f_e @ b.N.T)) @ b.N
        return scatter(b.nodes, y, b.n_red)

    def tangent_diagonal(self, W):
        """operator.py:554-564."""
        b = self.b
        y = b.Gc[:, None] * np.diagonal(self.K_t)[None, :] + (W * b.w) @ (b.N ** 2)
        return scatter(b.nodes, y, b.n_red)

    def fracture_energy(self, d):
        """operator.py:611-631: ``Gc/(pi ell) int alpha(d) + Gc ell/pi int |grad d|^2``."""
        return self.damage_energy(d, None)


# ---------------------------------------------------------------------------------------
# solves
# ---------------------------------------------------------------------------------------

def energy_line_search(op: PfczmOperator, fixed, x, y, g):
    """staggered.py:589-626: halve from ``y`` towards ``x`` until the degraded energy (own
    pattern) passes the Armijo test with the round-off allowance."""
    p = y - x
    e0 = op.elastic_energy(x, g)
    force = op.mech_action(x, g)
    free = ~fixed
    slope = min(float(force[free] @ p[free]), 0.0)
    allowance = LINE_SEARCH_ENERGY_RTOL * abs(e0)
    alpha = 1.0
    for _ in range(LINE_SEARCH_MAX_HALVINGS + 1):
        trial = y if alpha == 1.0 else x + alpha * p
        if op.elastic_energy(trial, g) <= e0 + LINE_SEARCH_ARMIJO * alpha * slope + allowance:
            return trial, alpha
        alpha *= 0.5
    raise RuntimeError("energy line search found no decrease")


def closure_mechanics(op: PfczmOperator, g, fixed, values, x0, *, rtol=1e-8, check_every=10, maxiter=20000):
    """staggered.py:698-751: linear solves at fixed patterns, each followed by the energy
    line search, until no bit flips materially after a full step.  Returns (x, pcg
    iterations, solves)."""
    pattern = op.closure_pattern(x0)
    its = 0
    for solve in range(1, MAX_PATTERN_SOLVES + 1):
        y, info = solve_constrained(lambda v: op.mech_action(v, g, pattern), op.mech_diagonal(g, pattern), fixed,
                                    values, x0=x0, rtol=rtol, maxiter=maxiter, check_every=check_every)
        if not info[1]:
            raise RuntimeError("mechanics PCG did not converge")
        its += info[0]
        x, alpha = energy_line_search(op, fixed, x0, y, g)
        pattern, changed = op.closure_pattern_change(x, pattern, g)
        if not changed and alpha == 1.0:
            return x, its, solve
        x0 = x
    raise RuntimeError("closure pattern did not settle")


def mechanics(op: PfczmOperator, g, fixed, values, x0, **kw):
    if op.closure:
        return closure_mechanics(op, g, fixed, values, x0, **kw)
    y, info = solve_constrained(lambda v: op.mech_action(v, g), op.mech_diagonal(g), fixed, values, x0=x0,
                                rtol=kw.get("rtol", 1e-8), maxiter=kw.get("maxiter", 20000),
                                check_every=kw.get("check_every", 10))
    return y, info[0], 1


def damage_pfczm(op: PfczmOperator, H, lower, initial, *, tol=1e-8, linear_tol=1e-10, max_newton=100,
                 max_lin=4000, max_as=64):
    """staggered.py:843-1096 (absolute-field QP): projected Newton over ``lower <= d <= 1``
    with Armijo halving and projected step doubling.  Returns (d, newton iterations, pcg
    iterations, active-set iterations)."""
    scale = op.b.damage_scale
    inner = INNER_KKT_FRACTION * tol
    ones = np.ones_like(initial)

    def projected_residual(field):
        grad = scale * op.damage_gradient(field, H)
        pg = np.where(field <= lower, np.minimum(grad, 0.0), grad)
        pg = np.where(field >= 1.0, np.maximum(pg, 0.0), pg)
        return float(np.abs(pg).max())

    d = initial.copy()
    energy = op.damage_energy(d, H)
    pcg = sets = 0
    for iteration in range(max_newton + 1):
        W = op.tangent_weights(d, H)
        gradient = scale * op.damage_gradient(d, H)
        action = lambda f: scale * op.tangent_action(f, W)     # noqa: E731
        diagonal = scale * op.tangent_diagonal(W)
        at_lower, at_upper = d <= lower, d >= 1.0
        pg = np.where(at_lower, np.minimum(gradient, 0.0), gradient)
        pg = np.where(at_upper, np.maximum(pg, 0.0), pg)
        residual = float(np.abs(pg).max())
        if residual <= tol:
            return d, iteration, pcg, sets
        if iteration == max_newton:
            break
        hold_lower = at_lower & (gradient > tol)
        hold_upper = at_upper & (gradient < -tol)
        upper_k = np.where(hold_lower, lower, ones)
        lower_k = np.where(hold_upper, ones, lower)
        r0 = float(np.linalg.norm(gradient))
        lin = max(min(linear_tol, 0.1 * inner / max(r0, 1.0)), MIN_LINEAR_TOLERANCE)
        rhs = action(d) - gradient
        target, _, _, it_as, total, _ = box_kkt(action, rhs, diagonal, lower_k, upper_k, initial=d, lin_tol=lin,
                                                kkt_tol=inner, max_lin=max_lin, max_as=max_as)
        pcg += total
        sets += it_as
        step = target - d
        slope = float((gradient * step).sum()) / scale
        t = 1.0
        for _ in range(NEWTON_MAX_HALVINGS + 1):
            trial = target if t == 1.0 else np.minimum(np.maximum(d + t * step, lower), ones)
            trial_energy = op.damage_energy(trial, H)
            if math.isfinite(trial_energy) and trial_energy <= (
                    energy + NEWTON_ARMIJO * t * slope + NEWTON_ENERGY_RTOL * abs(energy)):
                break
            t *= 0.5
        else:
            raise RuntimeError("PF-CZM line search failed")
        if t == 1.0 and not projected_residual(trial) <= tol:
            for _ in range(NEWTON_MAX_DOUBLINGS):
                t *= 2.0
                further = np.minimum(np.maximum(d + t * step, lower), ones)
                if np.array_equal(further, trial):
                    break
                further_energy = op.damage_energy(further, H)
                if not (math.isfinite(further_energy) and further_energy < trial_energy):
                    break
                trial, trial_energy = further, further_energy
        d, energy = trial, trial_energy
    raise RuntimeError("PF-CZM Newton did not converge")


def staggered_increment(op: PfczmOperator, state: dict, target: float, *, stagger_tol=1e-4, min_stagger=2,
                        max_stagger=2000, rtol=1e-8, damage_linear_tol=1e-10, damage_kkt_tol=1e-8,
                        check_every=10, max_newton=100):
    """One accepted increment of staggered.py:_attempt (1259-1517) with the PF-CZM law and the
    closure: predictor, alternate mechanics (closure pattern iteration) and projected Newton
    damage until the change measure (1097-1121) is below ``stagger_tol``, final mechanics
    re-solve.  ``state``: x, d, H, sigma (macro stress), external (work).  Returns the new
    state and a record dict."""
    b = op.b
    c = b.controlled
    n = 3 * b.n_red
    previous = float(state["x"][n + c])
    x_old = state["x"] * (target / previous) if previous != 0.0 else np.zeros_like(state["x"])
    x_old[n + c] = target
    d_old, lower = state["d"].copy(), state["d"]
    fixed = b.mech_fixed_mask()
    vals = np.zeros(b.n_mech)
    vals[n + c] = target
    step_size = abs(target - previous)
    u_floor = min(b.h) * math.sqrt(n)
    H = state["H"]
    kw = dict(rtol=rtol, check_every=check_every)
    newton = solves = 0
    change = math.inf
    for k in range(1, max_stagger + 1):
        g = op.degradation(d_old)
        x_new, _, ns = mechanics(op, g, fixed, vals, x_old, **kw)
        solves += ns
        H = np.maximum(state["H"], b.driving_force(x_new))
        d_new, nit, _, _ = damage_pfczm(op, H, lower, d_old, tol=damage_kkt_tol, linear_tol=damage_linear_tol,
                                        max_newton=max_newton)
        newton += nit
        du = np.linalg.norm(x_new[:n] - x_old[:n])
        u = np.linalg.norm(x_new[:n])
        dE = np.abs(x_new[n:] - x_old[n:]).max()
        E_inf = np.abs(x_new[n:]).max()
        dd = np.abs(d_new - d_old).max()
        scale_E = max(E_inf, step_size)
        u_ref = max(u, scale_E * u_floor)
        change = max(du / u_ref if u_ref > 0 else 0.0, dE / scale_E if scale_E > 0 else 0.0, dd)
        x_old, d_old = x_new, d_new
        if change <= stagger_tol and k >= min_stagger:
            break
    else:
        raise RuntimeError("staggered iteration did not converge")
    g = op.degradation(d_old)
    x, _, ns = mechanics(op, g, fixed, vals, x_old, **kw)
    solves += ns
    _, r_E, elastic = op.stress_pass(x, g, assemble=False)
    sigma = r_E / b.volume
    fracture = op.fracture_energy(d_old)
    external = state["external"] + b.volume * 0.5 * (sigma[c] + state["sigma"][c]) * (target - previous)
    rec = {"stagger_iters": k, "newton": newton, "closure_solves": solves, "sigma": sigma, "elastic": elastic,
           "fracture": fracture, "external": external,
           "balance_residual_rel": (external - (elastic + fracture)) / max(abs(external), np.finfo(float).tiny),
           "change": change}
    return {"x": x, "d": d_old, "H": H, "sigma": sigma, "external": external}, rec


def initial_state(op: PfczmOperator):
    b = op.b
    return {"x": np.zeros(b.n_mech), "d": np.zeros(b.n_red), "H": np.zeros((b.E, 8)), "sigma": np.zeros(6),
            "external": 0.0}
