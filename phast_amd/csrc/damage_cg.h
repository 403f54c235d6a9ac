// Device-resident damage CG shared by the Q4 strip kernels and the T3 node kernels: the
// per-instance state and the scalar steps that turn per-block partial sums into the CG
// scalars and stop flags (solvers/damage_solver.py:3377-3775, projected and post-clamp).
#pragma once
#include "common.h"

#include "../../include/phast_mi.h"

namespace phast {

// Per-instance state of the device-resident damage CG.
struct DamageCG {
    double rz, rz_new, pAp, alpha, rr0, bb, tol_sq, rfn, beta, tol;
    int32_t iter, max_iter, done, limited, check, restart, status, changed;
};

// ---------------------------------------------------------------------------------------
// per-instance scalar steps (one block per instance reduces part[off .. off+cnt) of each
// partial row; row k of the partials starts at k * n_tiles)
// ---------------------------------------------------------------------------------------

enum ScalarOp : int { S_INIT = 0, S_A = 1, S_B = 2, S_C = 3, S_PC_INIT = 4, S_PC_A = 5, S_PC_C = 6,
                      // plain Jacobi-PCG of the adjoint recomputation (damage_solver.py:423-478)
                      S_AJ_INIT = 7, S_AJ_A = 8, S_AJ_C = 9 };

// Sum of part[off .. off+cnt) in one fixed order whatever the block width NT: 256 virtual
// lanes, lane k summing part[k], part[k+256], ...; each virtual wave of 64 lanes reduced by
// butterfly, then the four wave totals in order.  A block of NT threads runs 256/NT
// virtual passes, so every block width (and every block that repeats the reduction)
// produces the same bits.  red holds 4 doubles.
template <int NT>
__device__ __forceinline__ double reduce_sum_parts(const double *part, int off, int cnt, double *red) {
    constexpr int V = 256 / NT;
    double w[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        // lane sums part[k], part[k+256], ... in that order; up to sixteen loads in flight
        double s = 0.0;
        int k = threadIdx.x + v * NT;
        for (; k + 15 * 256 < cnt; k += 16 * 256) {
            double a[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) a[j] = part[off + k + j * 256];
#pragma unroll
            for (int j = 0; j < 16; ++j) s += a[j];
        }
        for (; k + 3 * 256 < cnt; k += 4 * 256) {
            const double a = part[off + k], b = part[off + k + 256], c = part[off + k + 512],
                         e = part[off + k + 768];
            s += a;
            s += b;
            s += c;
            s += e;
        }
        for (; k < cnt; k += 256) s += part[off + k];
        w[v] = wave_sum(s);
    }
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) {
#pragma unroll
        for (int v = 0; v < V; ++v) red[v * (NT / kWave) + threadIdx.x / kWave] = w[v];
    }
    __syncthreads();
    return ((red[0] + red[1]) + red[2]) + red[3];
}
template <int NT>
__device__ __forceinline__ double reduce_min_parts(const double *part, int off, int cnt, double *red) {
    double s = __builtin_huge_val(), s2 = s, s3 = s, s4 = s;  // the minimum is exact in any order
    int k = threadIdx.x;
    for (; k + 15 * NT < cnt; k += 16 * NT) {
        double a[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) a[j] = part[off + k + j * NT];
#pragma unroll
        for (int j = 0; j < 16; j += 4) {
            s = fmin(s, a[j]);
            s2 = fmin(s2, a[j + 1]);
            s3 = fmin(s3, a[j + 2]);
            s4 = fmin(s4, a[j + 3]);
        }
    }
    for (; k + 3 * NT < cnt; k += 4 * NT) {
        s = fmin(s, part[off + k]);
        s2 = fmin(s2, part[off + k + NT]);
        s3 = fmin(s3, part[off + k + 2 * NT]);
        s4 = fmin(s4, part[off + k + 3 * NT]);
    }
    for (; k < cnt; k += NT) s = fmin(s, part[off + k]);
    return block_min<NT>(fmin(fmin(s, s2), fmin(s3, s4)), red);
}

// reduce_sum_parts of two rows (part + off and part + off2) in one pass, each in
// reduce_sum_parts' order (same bits), the loads of both rows in flight together.
template <int NT>
__device__ __forceinline__ void reduce_sum2_parts(const double *part, int off, int off2, int cnt, double *red,
                                                  double &s1, double &s2) {
    constexpr int V = 256 / NT;
    const double *p0 = part + off, *p1 = part + off2;
    double w0[V], w1[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        double s = 0.0, t = 0.0;
        int k = threadIdx.x + v * NT;
        for (; k + 7 * 256 < cnt; k += 8 * 256) {
            double a[8], b[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                a[j] = p0[k + j * 256];
                b[j] = p1[k + j * 256];
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                s += a[j];
                t += b[j];
            }
        }
        for (; k < cnt; k += 256) {
            s += p0[k];
            t += p1[k];
        }
        w0[v] = wave_sum(s);
        w1[v] = wave_sum(t);
    }
    constexpr int NW = NT / kWave;
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            red[v * NW + threadIdx.x / kWave] = w0[v];
            red[4 + v * NW + threadIdx.x / kWave] = w1[v];
        }
    }
    __syncthreads();
    s1 = ((red[0] + red[1]) + red[2]) + red[3];
    s2 = ((red[4] + red[5]) + red[6]) + red[7];
}

// reduce_sum_parts of row 0 and reduce_min_parts of rows 1 and 2 (each n_tiles apart) in one
// pass: the three rows' loads are in flight together; the sum keeps reduce_sum_parts' order
// (same bits) and the minima are exact in any order.
template <int NT>
__device__ __forceinline__ void reduce_sum_min2_parts(const double *part, int off, int cnt, int n_tiles,
                                                      double *red, double &sum, double &mn1, double &mn2) {
    constexpr int V = 256 / NT;
    const double *p0 = part + off, *p1 = part + n_tiles + off, *p2 = part + 2 * n_tiles + off;
    double w[V], a1 = __builtin_huge_val(), a2 = a1;
#pragma unroll
    for (int v = 0; v < V; ++v) {
        double s = 0.0;
        int k = threadIdx.x + v * NT;
        for (; k + 7 * 256 < cnt; k += 8 * 256) {
            double a[8], b[8], c[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                a[j] = p0[k + j * 256];
                b[j] = p1[k + j * 256];
                c[j] = p2[k + j * 256];
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                s += a[j];
                a1 = fmin(a1, b[j]);
                a2 = fmin(a2, c[j]);
            }
        }
        for (; k < cnt; k += 256) {
            s += p0[k];
            a1 = fmin(a1, p1[k]);
            a2 = fmin(a2, p2[k]);
        }
        w[v] = wave_sum(s);
    }
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) {
#pragma unroll
        for (int v = 0; v < V; ++v) red[v * (NT / kWave) + threadIdx.x / kWave] = w[v];
    }
    __syncthreads();
    sum = ((red[0] + red[1]) + red[2]) + red[3];
    mn1 = block_min<NT>(a1, red);
    mn2 = block_min<NT>(a2, red);
}

// One per-instance scalar step of the CG (thread 0 writes S; red holds 8 doubles).  Every
// thread of the block takes part in the reductions; all return paths are block-uniform.
template <int OP, int NT>
__device__ __forceinline__ void scalar_step(int32_t off, int32_t cnt, DamageCG &S, const double *part,
                                            int32_t n_tiles, int32_t check_every, double *red) {
    if (OP == S_INIT || OP == S_PC_INIT) {
        const double rr = reduce_sum_parts<NT>(part, off, cnt, red);
        const double bb = reduce_sum_parts<NT>(part, n_tiles + off, cnt, red);
        const double rz = reduce_sum_parts<NT>(part, 2 * n_tiles + off, cnt, red);
        if (threadIdx.x == 0) {
            S.rr0 = rr;
            S.bb = bb;
            S.rz = rz;
            const double ref = fmax(fmax(sqrt(bb), sqrt(rr)), 1e-30);
            S.tol_sq = (S.tol * ref) * (S.tol * ref);
            S.iter = 0;
            S.done = S.max_iter <= 0 ? 1 : 0;
            S.status = S.max_iter <= 0 ? PHAST_CG_MAXITER : PHAST_CG_RUNNING;
            S.limited = S.check = S.restart = 0;
        }
    } else if (OP == S_A || OP == S_PC_A) {
        double pAp, up = 0.0, dn = 0.0;
        if (OP == S_A) {
            reduce_sum_min2_parts<NT>(part, off, cnt, n_tiles, red, pAp, up, dn);
        } else {
            pAp = reduce_sum_parts<NT>(part, off, cnt, red);
        }
        if (threadIdx.x == 0) {
            S.iter += 1;
            S.pAp = pAp;
            if (!(pAp > 0.0)) {  // damage_solver.py:3560-3575 (pAp <= 0 -> break)
                S.done = 1;
                S.status = PHAST_CG_NONPOSITIVE;
                S.limited = 0;
                return;
            }
            const double alpha_cg = S.rz / (pAp + 1e-30);
            if (OP == S_A) {
                double amax = alpha_cg;  // damage_solver.py:3579-3605
                if (up < amax) amax = up;
                if (dn < amax) amax = dn;
                S.limited = (amax < alpha_cg && amax >= 0.0) ? 1 : 0;
                S.alpha = S.limited ? amax : alpha_cg;
                S.check = (!S.limited && (S.iter % check_every) == 0) ? 1 : 0;
            } else {
                S.alpha = alpha_cg;
                S.check = (S.iter % check_every) == 0 ? 1 : 0;
            }
        }
    } else if (OP == S_B) {
        if (!(S.limited || S.check)) {
            if (threadIdx.x == 0) S.restart = 0;
            return;
        }
        double rfn, chg;
        reduce_sum2_parts<NT>(part, off, n_tiles + off, cnt, red, rfn, chg);
        if (threadIdx.x == 0) {
            S.rfn = rfn;
            if (rfn < S.tol_sq) {
                S.done = 1;
                S.status = PHAST_CG_CONVERGED;
            } else if (rfn > 1e12 * fmax(S.rr0, 1e-30)) {
                S.done = 1;
                S.status = PHAST_CG_DIVERGED;
            } else if (!isfinite(rfn)) {
                S.done = 1;
                S.status = PHAST_CG_NONFINITE;
            }
            S.changed = chg > 0.0 ? 1 : 0;
            S.restart = (S.limited || (S.check && chg > 0.0)) ? 1 : 0;
        }
    } else if (OP == S_C) {
        const double rz_new = reduce_sum_parts<NT>(part, off, cnt, red);
        if (threadIdx.x == 0) {
            S.beta = S.restart ? 0.0 : rz_new / (S.rz + 1e-30);
            S.rz = rz_new;
            if (S.iter >= S.max_iter) {
                S.done = 1;
                S.status = PHAST_CG_MAXITER;
            }
        }
    } else if (OP == S_AJ_INIT) {
        // rows: r.r (unused), rhs.rhs, r.z; tol_sq = rel_tol * max(rhs.rhs, 1e-30) with S.tol = rel_tol
        double bb, rz;
        reduce_sum2_parts<NT>(part, n_tiles + off, 2 * n_tiles + off, cnt, red, bb, rz);
        if (threadIdx.x == 0) {
            S.bb = bb;
            S.rz = rz;
            S.tol_sq = S.tol * fmax(bb, 1e-30);
            S.iter = 0;
            S.done = S.max_iter <= 0 ? 1 : 0;
            S.status = S.max_iter <= 0 ? PHAST_CG_MAXITER : PHAST_CG_RUNNING;
            S.limited = S.check = S.restart = 0;
        }
    } else if (OP == S_AJ_A) {
        const double pAp = reduce_sum_parts<NT>(part, off, cnt, red);
        if (threadIdx.x == 0) {
            S.iter += 1;
            S.pAp = pAp;
            if (pAp <= 0.0) {  // `if pAp.item() <= 0: break`, before the update
                S.done = 1;
                S.status = PHAST_CG_NONPOSITIVE;
                return;
            }
            S.alpha = S.rz / (pAp + 1e-30);
        }
    } else if (OP == S_AJ_C) {
        double rr, rz_new;
        reduce_sum2_parts<NT>(part, off, n_tiles + off, cnt, red, rr, rz_new);
        if (threadIdx.x == 0) {
            S.rfn = rr;
            if (rr < S.tol_sq) {  // converged after the update, before the direction
                S.done = 1;
                S.status = PHAST_CG_CONVERGED;
                return;
            }
            S.beta = rz_new / (S.rz + 1e-30);
            S.rz = rz_new;
            if (S.iter >= S.max_iter) {
                S.done = 1;
                S.status = PHAST_CG_MAXITER;
            }
        }
    } else if (OP == S_PC_C) {
        double rr, rz_new;
        reduce_sum2_parts<NT>(part, off, n_tiles + off, cnt, red, rr, rz_new);
        if (threadIdx.x == 0) {
            S.rfn = rr;
            if (rr < S.tol_sq) {  // damage_solver.py:3496-3506
                S.done = 1;
                S.status = PHAST_CG_CONVERGED;
                return;
            }
            if (rr > 1e12 * fmax(S.rr0, 1e-30)) {
                S.done = 1;
                S.status = PHAST_CG_DIVERGED;
                return;
            }
            S.beta = rz_new / (S.rz + 1e-30);
            S.rz = rz_new;
            if (S.iter >= S.max_iter) {
                S.done = 1;
                S.status = PHAST_CG_MAXITER;
            }
        }
    }
}

}  // namespace phast
