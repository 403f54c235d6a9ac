"""Oracle (numpy, float64) for the implicit differentiation of the T3 damage solve.

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.  Imported by ``tests/``; never by
the product.

Restates, op for op, ``solvers/damage_solver.py`` (paths relative to the reference's
``src/phast``): the active-set split of the upstream gradient (:70-105, :400-421), the
adjoint Jacobi-PCG (:423-478), the history pullback (:269-281), the per-element quadratic
forms (:284-331) and the backward passes of ``_AdjointDamageSolve`` (:1038-1152),
``_AdjointDamageSolveScalar`` (:482-856, without gamma correction) and
``_AdjointDamageSolveField`` (:859-1035).
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .q4 import Material
from .t3 import T3Damage, T3DamageSolver, T3Mesh


def classify_active_set(d_new, d_prev, fixed=None, upper_bound=1.0, lower_atol=1e-12, upper_atol=1e-12):
    """(lower, upper, fixed, interior) masks of classify_damage_active_set (damage_solver.py:70-105):
    the upper bound and pins take precedence over the lower (irreversibility) bound."""
    upper = d_new >= (upper_bound - upper_atol)
    fx = np.zeros(d_new.shape, dtype=bool) if fixed is None else np.asarray(fixed, dtype=bool)
    lower = (d_new <= d_prev + lower_atol) & ~upper & ~fx
    upper = upper & ~fx
    interior = ~(lower | upper | fx)
    return lower, upper, fx, interior


def masked_adjoint_inputs(grad_output, d_sol, d_prev, fixed=None):
    """(g, grad_d_prev) of _masked_adjoint_inputs (damage_solver.py:400-421): the interior part
    drives the adjoint system, the lower-active part passes through to d_prev."""
    lower, _, _, interior = classify_active_set(d_sol, d_prev, fixed)
    return grad_output * interior.astype(np.float64), grad_output * lower.astype(np.float64)


def pcg_adjoint(Ax, rhs, M_inv, fixed=None, vals=None, initial=None, rel_tol=1e-22, max_iter=1000):
    """_pcg_solve_adjoint_system (damage_solver.py:423-478): Jacobi-PCG with pinned DOFs
    eliminated (held at ``vals``, zero residual, zero preconditioner).  Returns (x, iters)."""
    x = np.zeros_like(rhs) if initial is None else initial.copy()
    M_eff = M_inv
    if fixed is not None:
        if vals is None:
            vals = np.zeros_like(rhs)
        x = np.where(fixed, vals, x)
        M_eff = M_inv.copy()
        M_eff[fixed] = 0.0
    r = rhs - Ax(x)
    if fixed is not None:
        r[fixed] = 0.0
    z = M_eff * r
    p = z.copy()
    rz = float(r @ z)
    rhs_norm = max(float(rhs @ rhs), 1e-30)
    tol_sq = rel_tol * rhs_norm
    it = 0
    for _ in range(max_iter):
        it += 1
        Ap = Ax(p)
        if fixed is not None:
            Ap[fixed] = 0.0
        pAp = float(p @ Ap)
        if pAp <= 0:
            break
        alpha = rz / (pAp + 1e-30)
        x = x + alpha * p
        if fixed is not None:
            x = np.where(fixed, vals, x)
        r = r - alpha * Ap
        if fixed is not None:
            r[fixed] = 0.0
        if float(r @ r) < tol_sq:
            break
        z = M_eff * r
        rz_new = float(r @ z)
        beta = rz_new / (rz + 1e-30)
        p = z + beta * p
        if fixed is not None:
            p[fixed] = 0.0
        rz = rz_new
    if fixed is not None:
        x = np.where(fixed, vals, x)
    return x, it


def quadratic_forms_perelem(mesh: T3Mesh, lam, d):
    """(lam K d, lam M d, lam M 1) per element (damage_solver.py:284-331)."""
    gp, areas = mesh.grad_phi, mesh.areas
    lam_e, d_e = lam[mesh.elements], d[mesh.elements]
    g_lam_x, g_lam_y = (gp[:, :, 0] * lam_e).sum(1), (gp[:, :, 1] * lam_e).sum(1)
    g_d_x, g_d_y = (gp[:, :, 0] * d_e).sum(1), (gp[:, :, 1] * d_e).sum(1)
    lKd = areas * (g_lam_x * g_d_x + g_lam_y * g_d_y)
    lam_sum = lam_e.sum(1)
    lMd = (areas / 12.0) * (lam_sum * d_e.sum(1) + (lam_e * d_e).sum(1))
    lM1 = (areas / 3.0) * lam_sum
    return lKd, lMd, lM1


def history_gradient(mesh: T3Mesh, lam, d_pull):
    """dL/dH_e = sum_a lam_a 2 A (1/3 - (d_a + sum d)/12) (damage_solver.py:1140-1149)."""
    d_e = d_pull[mesh.elements]
    d_sum = d_e.sum(1, keepdims=True)
    per_node = 2.0 * mesh.areas[:, None] * (1.0 / 3.0 - (d_e + d_sum) / 12.0)
    return (lam[mesh.elements] * per_node).sum(1)


def override_material(material: Material, Gc=None, l0=None) -> Material:
    """The constants the scalar rule installs around its solves (damage_solver.py:574-585):
    AT1 then takes S_H = 3 Gc/(8 l0) whatever the threshold setting."""
    return dataclasses.replace(material, Gc=material.Gc if Gc is None else float(Gc),
                               l0=material.l0 if l0 is None else float(l0), at1_threshold="auto")


class T3DamageAdjoint:
    """Forward and backward of the three autograd rules of the T3 damage solve.

    ``mode`` 'H' is _AdjointDamageSolve (gradients for H and d_prev), 'scalar'
    _AdjointDamageSolveScalar (adds scalar Gc and l0), 'field' _AdjointDamageSolveField
    (adds a per-element Gc field and l0)."""

    def __init__(self, mesh: T3Mesh, material: Material, tol=1e-5, max_iter=5000, bounds_method="post_clamp"):
        self.mesh, self.material = mesh, material
        self.tol, self.max_iter, self.bounds_method = tol, max_iter, bounds_method

    def _material(self, mode, Gc, l0):
        return override_material(self.material, Gc, l0) if mode == "scalar" else \
            override_material(self.material, None, l0) if mode == "field" else self.material

    def forward(self, H, d_prev, mode="H", Gc=None, l0=None, Gc_e=None, fixed=None, vals=None):
        mat = self._material(mode, Gc, l0)
        s = T3DamageSolver(self.mesh, mat, self.tol, self.max_iter, self.bounds_method,
                           Gc_e=Gc_e if mode == "field" else None)
        return s.solve(H, d_prev, fixed=fixed, vals=vals)

    def backward(self, H, d_prev, d_sol, grad_output, mode="H", Gc=None, l0=None, Gc_e=None, fixed=None,
                 vals=None):
        """Gradients as a dict: 'H', 'd_prev', and 'Gc' / 'l0' for the scalar and field rules;
        also the adjoint 'lam' and the unconstrained damage 'd_unc'."""
        mat = self._material(mode, Gc, l0)
        op = T3Damage(self.mesh, mat, Gc_e if mode == "field" else None)
        b = op.rhs(H)
        M_inv = 1.0 / (op.diagonal(H) + 1e-30)
        g, grad_prev = masked_adjoint_inputs(grad_output, d_sol, d_prev, fixed)
        cap = min(self.max_iter, 1000)
        init = None
        if fixed is not None:
            init = np.zeros_like(b)
            init[fixed] = vals[fixed]
        d_unc, _ = pcg_adjoint(lambda x: op.apply(x, H), b, M_inv, fixed, vals, init, 1e-24, cap)
        lam, _ = pcg_adjoint(lambda x: op.apply(x, H), g, M_inv, fixed, None, None,
                             1e-10 if mode == "H" else 1e-22, cap)
        d_pull = d_unc if self.bounds_method in ("post_clamp", "direct") else d_sol  # :269-281
        out = {"H": history_gradient(self.mesh, lam, d_pull), "d_prev": grad_prev, "lam": lam, "d_unc": d_unc}
        if mode == "H":
            return out
        lKd, lMd, lM1 = quadratic_forms_perelem(self.mesh, lam, d_unc)
        l0v = mat.l0
        at1 = mat.pf_model == "AT1"
        if mode == "scalar":  # damage_solver.py:782-797
            Gcv = mat.Gc
            sKd, sMd, sM1 = lKd.sum(), lMd.sum(), lM1.sum()
            if at1:
                out["Gc"] = -3.0 / (8.0 * l0v) * sM1 - 0.75 * l0v * sKd
                out["l0"] = 3.0 * Gcv / (8.0 * l0v ** 2) * sM1 - 0.75 * Gcv * sKd
            else:
                out["Gc"] = -(l0v * sKd + (1.0 / l0v) * sMd)
                out["l0"] = -(Gcv * sKd - (Gcv / l0v ** 2) * sMd)
        else:  # damage_solver.py:982-1001
            if at1:
                out["Gc"] = -3.0 / (8.0 * l0v) * lM1 - 0.75 * l0v * lKd
                out["l0"] = 3.0 / (8.0 * l0v ** 2) * (Gc_e * lM1).sum() - 0.75 * (Gc_e * lKd).sum()
            else:
                out["Gc"] = -(l0v * lKd + (1.0 / l0v) * lMd)
                out["l0"] = -((Gc_e * lKd).sum() - (Gc_e / l0v ** 2 * lMd).sum())
        return out
