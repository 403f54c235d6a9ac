// Device-resident Jacobi PCG engines and the box-KKT active-set kernels (FP64).
//
// Two Krylov loops mirror the reference's two PCG flavours exactly:
//   * pcg_flags  -- solid3d/linalg.py:99-220: alpha/beta masked by device flags, the host
//                   reads (nonfinite, breakdown, active, rr, done) every check_every
//                   iterations; iterates after convergence leave x unchanged.
//   * pcg_stop   -- solvers/box_kkt_h8.py:256-347: stops at the first iteration with
//                   ||r|| <= tol max(||r0||, 1); breakdowns raise.
// All scalars live on the device; every reduction is a fixed-order two-pass sum
// (per-block wave64 butterflies into partials, then one block), so runs are bitwise
// reproducible.  Operators are supplied as host callbacks that enqueue an apply kernel.
#include "common.h"
#include "hex8_internal.h"

#include <cstring>
#include <vector>

namespace phast {

constexpr int kVecBlocks = 1024;  // grid of the vector kernels (grid-stride)

struct PcgState {
    double rz, rz_new, pAp, alpha, beta, rr, thr2, b_norm;
    int32_t iter, done, active, breakdown, nonfinite, steps, status;
    int32_t phase;  // fused linalg.pcg passes: the latest r.z is rz (phase even) or rz_new (odd)
};

enum : int { PS_RUNNING = 0, PS_CONVERGED = 1, PS_NOT_PD = 2, PS_NONPOS_RZ = 3, PS_NONFINITE = 4, PS_MAXITER = 5 };

// Block partials of up to 5 sums, reduced together (block_sum's arithmetic per value).
// `red` holds kRed doubles.
constexpr int kRed = 5 * 4;
__device__ __forceinline__ void write_parts(double *part, double a, double b, double c, double *red) {
    double v[3] = {a, b, c};
    block_sum_n<256, 3>(v, red);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = v[0];
        part[kVecBlocks + blockIdx.x] = v[1];
        part[2 * kVecBlocks + blockIdx.x] = v[2];
    }
}

__device__ __forceinline__ void write_parts5(double *part, const double v[5], double *red) {
    double t[5] = {v[0], v[1], v[2], v[3], v[4]};
    block_sum_n<256, 5>(t, red);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 5; ++k) part[k * kVecBlocks + blockIdx.x] = t[k];
}

// Sums of partial rows k0 .. k0+K-1 (each in block order, as one block_sum per row), with
// every row's loads in flight together.
template <int K>
__device__ __forceinline__ void sum_parts_n(const double *part, int k0, double out[K], double *red) {
    static_assert(kVecBlocks % 256 == 0, "whole rows of partials per thread");
    constexpr int R = kVecBlocks / 256;
    double v[R][K];  // every load issued before the first add
#pragma unroll
    for (int j = 0; j < R; ++j)
#pragma unroll
        for (int k = 0; k < K; ++k) v[j][k] = part[(k0 + k) * kVecBlocks + threadIdx.x + 256 * j];
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = 0.0;
#pragma unroll
    for (int j = 0; j < R; ++j)
#pragma unroll
        for (int k = 0; k < K; ++k) out[k] += v[j][k];
    block_sum_n<256, K>(out, red);
}

#define GRID_STRIDE(i, n) for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < (n); i += (int64_t)kVecBlocks * 256)

// y <- free ? y : 0 ; partials: x.y, count(!finite(y))
__global__ __launch_bounds__(256) void k_mask_dot(int64_t n, const double *x, double *y, const uint8_t *fixed,
                                                  const PcgState *S, double *part) {
    __shared__ double red[kRed];
    if (S && S->done) return;
    double s = 0.0, bad = 0.0;
    GRID_STRIDE(i, n) {
        double v = y[i];
        if (fixed && fixed[i]) {
            v = 0.0;
            y[i] = 0.0;
        }
        if (!isfinite(v)) bad += 1.0;
        s += x[i] * v;
    }
    write_parts(part, s, bad, 0.0, red);
}

// generic partials: x.x, x.y, y.y over masked vectors
__global__ __launch_bounds__(256) void k_dots(int64_t n, const double *x, const double *y, double *part) {
    __shared__ double red[kRed];
    double a = 0.0, b = 0.0, c = 0.0, bad = 0.0;
    GRID_STRIDE(i, n) {
        const double xv = x[i];
        a += xv * xv;
        if (!isfinite(xv)) bad += 1.0;
        if (y) {
            const double yv = y[i];
            b += xv * yv;
            c += yv * yv;
            if (!isfinite(yv)) bad += 1.0;
        }
    }
    (void)c;
    write_parts(part, a, b, bad, red);
}

// The box-KKT PCG's start: z = r / diag; partials r.z, r.r (k_update<1>'s sums with a step
// of zero) and the count of non-finite residual entries (k_dots'), in one pass.
__global__ __launch_bounds__(256) void k_start_stop(int64_t n, const double *r, const double *dg, double *z,
                                                    double *part) {
    __shared__ double red[kRed];
    double rz = 0.0, rr = 0.0, bad = 0.0;
    GRID_STRIDE(i, n) {
        const double ri = r[i];
        const double zi = ri / dg[i];
        z[i] = zi;
        rz += ri * zi;
        rr += ri * ri;
        if (!isfinite(ri)) bad += 1.0;
    }
    write_parts(part, rz, rr, bad, red);
}

// linalg.py:194-203: x += a p, r -= a Ap, z = r * inv_diag; partials r.z, r.r
// (mode 1, box-KKT: z = r / diag; partials r.r, r.z).
template <int MODE>
__global__ __launch_bounds__(256) void k_update(int64_t n, double *x, double *r, const double *p, const double *Ap,
                                                const double *dg, double *z, const PcgState *S, double *part) {
    __shared__ double red[kRed];
    if (S->done) return;
    const double a = S->alpha;
    double rz = 0.0, rr = 0.0;
    GRID_STRIDE(i, n) {
        x[i] = x[i] + a * p[i];
        const double ri = r[i] - a * Ap[i];
        r[i] = ri;
        if (MODE != 2) {
            const double zi = MODE == 0 ? ri * dg[i] : ri / dg[i];
            z[i] = zi;
            rz += ri * zi;
        }
        rr += ri * ri;
    }
    write_parts(part, rz, rr, 0.0, red);
}

__global__ __launch_bounds__(256) void k_pupdate(int64_t n, double *p, const double *z, const PcgState *S) {
    if (S->done) return;
    const double b = S->beta;
    GRID_STRIDE(i, n) p[i] = z[i] + b * p[i];
}

// linalg.pcg scalar steps
__global__ __launch_bounds__(256) void k_scalar_flags_a(PcgState *S, const double *part) {
    __shared__ double red[kRed];
    double t[2];
    sum_parts_n<2>(part, 0, t, red);
    const double pAp = t[0], bad = t[1];
    if (threadIdx.x) return;
    const bool positive = pAp > 0.0;
    if (S->active && !positive) S->breakdown = 1;
    if (!isfinite(pAp) || bad > 0.0) S->nonfinite = 1;
    const bool step = S->active && positive;
    S->pAp = pAp;
    S->alpha = step ? S->rz / pAp : 0.0;
    S->steps += step ? 1 : 0;
    S->iter += 1;
}
__global__ __launch_bounds__(256) void k_scalar_flags_b(PcgState *S, const double *part) {
    __shared__ double red[kRed];
    double t[2];
    sum_parts_n<2>(part, 0, t, red);
    const double rz_new = t[0], rr = t[1];
    if (threadIdx.x) return;
    if (!isfinite(rz_new) || !isfinite(rr)) S->nonfinite = 1;
    S->rr = rr;
    S->active = S->active && (rr > S->thr2);
    S->beta = (S->active && S->rz > 0.0) ? rz_new / S->rz : 0.0;
    S->rz = rz_new;
}

// linalg.pcg's scalar steps fused into its vector passes, as k_update_stop / k_pupdate_stop
// do for the box-KKT PCG: every block sums the partials in k_scalar_flags_a/b's order and
// block 0 records the flags.  The flags block 0 writes are not read by the other blocks of
// the same pass, except `active`, which only ever falls to 0 and then falls for every
// block's own reading too.  S->phase counts the direction passes and equals S->iter at the
// start of an iteration.  MODE 0: z = r * inv_diag with partials r.z, r.r to rows 4, 5;
// MODE 2: r only (an external preconditioner and k_rz follow).
template <int MODE>
__global__ __launch_bounds__(256) void k_update_flags(int64_t n, double *x, double *r, const double *p,
                                                      const double *Ap, const double *dg, double *z, PcgState *S,
                                                      double *part) {
    __shared__ double red[kRed];
    double t[2];
    sum_parts_n<2>(part, 0, t, red);  // k_scalar_flags_a's sums
    const double pAp = t[0], bad = t[1];
    const int active = S->active;
    const double rz = (S->phase & 1) ? S->rz_new : S->rz;
    const bool positive = pAp > 0.0;
    const bool step = active && positive;
    const double a = step ? rz / pAp : 0.0;
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (active && !positive) S->breakdown = 1;
        if (!isfinite(pAp) || bad > 0.0) S->nonfinite = 1;
        S->pAp = pAp;
        S->alpha = a;
        S->steps += step ? 1 : 0;
        S->iter += 1;
    }
    double rzn = 0.0, rr = 0.0;
    GRID_STRIDE(i, n) {
        x[i] = x[i] + a * p[i];
        const double ri = r[i] - a * Ap[i];
        r[i] = ri;
        if (MODE == 0) {
            const double zi = ri * dg[i];
            z[i] = zi;
            rzn += ri * zi;
        }
        rr += ri * ri;
    }
    if (MODE == 0) write_parts(part + 4 * kVecBlocks, rzn, rr, 0.0, red);
}
// k_scalar_flags_b fused into the direction update p = z + beta p; the r.z, r.r partials
// are in rows `row`, `row + 1`.
__global__ __launch_bounds__(256) void k_pupdate_flags(int64_t n, double *p, const double *z, PcgState *S,
                                                       const double *part, int row) {
    __shared__ double red[kRed];
    double t[2];
    sum_parts_n<2>(part, row, t, red);  // k_scalar_flags_b's sums
    const double rz_new = t[0], rr = t[1];
    // this iteration's phase, from iter (which only the update pass writes): block 0 writes
    // S->phase here, so no block may read it in this pass
    const int ph = S->iter - 1;
    const double rz_old = (ph & 1) ? S->rz_new : S->rz;
    const bool active = S->active && (rr > S->thr2);
    const double b = (active && rz_old > 0.0) ? rz_new / rz_old : 0.0;
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (!isfinite(rz_new) || !isfinite(rr)) S->nonfinite = 1;
        S->rr = rr;
        S->active = active;
        S->beta = b;
        if (ph & 1) S->rz = rz_new;
        else S->rz_new = rz_new;
        S->phase = ph + 1;
    }
    GRID_STRIDE(i, n) p[i] = z[i] + b * p[i];
}

// box_kkt_h8._pcg scalar steps (stop at first convergence; breakdowns end the loop)
__global__ __launch_bounds__(256) void k_scalar_stop_a(PcgState *S, const double *part) {
    __shared__ double red[kRed];
    if (S->done) return;
    double t[2];
    sum_parts_n<2>(part, 0, t, red);
    const double den = t[0], bad = t[1];
    if (threadIdx.x) return;
    S->iter += 1;
    if (bad > 0.0) {
        S->done = 1;
        S->status = PS_NONFINITE;
        return;
    }
    if (!isfinite(den) || den <= 0.0) {
        S->done = 1;
        S->status = PS_NOT_PD;
        return;
    }
    S->pAp = den;
    S->alpha = S->rz / den;
}
__global__ __launch_bounds__(256) void k_scalar_stop_b(PcgState *S, const double *part, int32_t max_iter) {
    __shared__ double red[kRed];
    if (S->done) return;
    double t[2];
    sum_parts_n<2>(part, 0, t, red);
    const double rz_new = t[0], rr = t[1];
    if (threadIdx.x) return;
    S->rr = rr;
    if (sqrt(rr) <= S->thr2) {  // thr2 holds tol * reference_norm here
        S->done = 1;
        S->status = PS_CONVERGED;
        return;
    }
    if (!isfinite(rz_new) || rz_new <= 0.0) {
        S->done = 1;
        S->status = PS_NONPOS_RZ;
        return;
    }
    S->beta = rz_new / S->rz;
    S->rz = rz_new;
    if (S->iter >= max_iter) {
        S->done = 1;
        S->status = PS_MAXITER;
    }
}

// The box-KKT PCG's scalar steps fused into its two vector passes.  Every block sums the
// partials of the previous pass in the scalar kernels' order (so it takes the same decision
// from the same bits) and block 0 records the state; a field block 0 writes in a pass is
// not one the other blocks read there (S->done aside: a block that already sees it set has
// the same decision to follow).  The latest r.z alternates between S->rz and S->rz_new:
// iteration k (k = S->steps in the update pass, S->iter - 1 in the direction pass) reads
// slot k & 1 and the direction pass writes slot (k + 1) & 1.  The update pass writes its
// partials to rows 4, 5 of `part`, since rows 0, 1 are still being read.
__global__ __launch_bounds__(256) void k_update_stop(int64_t n, double *x, double *r, const double *p,
                                                     const double *Ap, const double *dg, double *z, PcgState *S,
                                                     double *part) {
    __shared__ double red[kRed];
    __shared__ int sdone;  // one reading per block, so that the whole block takes one branch
    if (threadIdx.x == 0) sdone = S->done;
    __syncthreads();
    if (sdone) return;
    const double rz = (S->steps & 1) ? S->rz_new : S->rz;
    double t[2];
    sum_parts_n<2>(part, 0, t, red);  // k_scalar_stop_a's sums
    const double den = t[0], bad = t[1];
    const bool fail_bad = bad > 0.0, fail_pd = !isfinite(den) || den <= 0.0;
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        S->iter += 1;
        if (fail_bad) {
            S->done = 1;
            S->status = PS_NONFINITE;
        } else if (fail_pd) {
            S->done = 1;
            S->status = PS_NOT_PD;
        } else {
            S->pAp = den;
            S->alpha = rz / den;
        }
    }
    if (fail_bad || fail_pd) return;
    const double a = rz / den;
    double rzn = 0.0, rr = 0.0;
    GRID_STRIDE(i, n) {
        x[i] = x[i] + a * p[i];
        const double ri = r[i] - a * Ap[i];
        r[i] = ri;
        const double zi = ri / dg[i];
        z[i] = zi;
        rzn += ri * zi;
        rr += ri * ri;
    }
    write_parts(part + 4 * kVecBlocks, rzn, rr, 0.0, red);
}
__global__ __launch_bounds__(256) void k_pupdate_stop(int64_t n, double *p, const double *z, PcgState *S,
                                                      const double *part, int32_t max_iter) {
    __shared__ double red[kRed];
    __shared__ int sdone;
    if (threadIdx.x == 0) sdone = S->done;
    __syncthreads();
    if (sdone) return;
    const int k = S->iter - 1;
    const double rz_old = (k & 1) ? S->rz_new : S->rz;
    double t[2];
    sum_parts_n<2>(part, 4, t, red);  // k_scalar_stop_b's sums
    const double rz_new = t[0], rr = t[1];
    const bool conv = sqrt(rr) <= S->thr2;  // thr2 holds tol * reference_norm here
    const bool nonpos = !conv && (!isfinite(rz_new) || rz_new <= 0.0);
    const double b = rz_new / rz_old;
    const bool maxed = !conv && !nonpos && S->iter >= max_iter;
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        S->rr = rr;
        if (conv) {
            S->done = 1;
            S->status = PS_CONVERGED;
        } else if (nonpos) {
            S->done = 1;
            S->status = PS_NONPOS_RZ;
        } else {
            S->beta = b;
            if (k & 1) S->rz = rz_new;
            else S->rz_new = rz_new;
            S->steps = k + 1;
            if (maxed) {
                S->done = 1;
                S->status = PS_MAXITER;
            }
        }
    }
    if (conv || nonpos || maxed) return;
    GRID_STRIDE(i, n) p[i] = z[i] + b * p[i];
}

// partials r.z, r.r, count(!finite(z)) after an external preconditioner
__global__ __launch_bounds__(256) void k_rz(int64_t n, const double *r, const double *z, const PcgState *S,
                                            double *part) {
    __shared__ double red[kRed];
    if (S && S->done) return;
    double rz = 0.0, rr = 0.0, bad = 0.0;
    GRID_STRIDE(i, n) {
        const double ri = r[i], zi = z[i];
        rz += ri * zi;
        rr += ri * ri;
        if (!isfinite(zi)) bad += 1.0;
    }
    write_parts(part, rz, rr, bad, red);
}

// elementwise helpers
__global__ __launch_bounds__(256) void k_axpby_mask(int64_t n, double alpha, const double *x, double beta,
                                                    const double *y, const uint8_t *fixed, double fixed_val,
                                                    double *out) {
    GRID_STRIDE(i, n) {
        double v = alpha * x[i] + (y ? beta * y[i] : 0.0);
        if (fixed && fixed[i]) v = fixed_val;
        out[i] = v;
    }
}
__global__ __launch_bounds__(256) void k_where(int64_t n, const uint8_t *mask, const double *a, const double *b,
                                               double bval, double *out) {
    GRID_STRIDE(i, n) out[i] = mask[i] ? a[i] : (b ? b[i] : bval);
}
__global__ __launch_bounds__(256) void k_zero_fixed(int64_t n, const uint8_t *fixed, double *out) {
    GRID_STRIDE(i, n) if (fixed[i]) out[i] = 0.0;
}
__global__ __launch_bounds__(256) void k_invert(int64_t n, const double *d, double *out) {
    GRID_STRIDE(i, n) out[i] = 1.0 / d[i];
}
__global__ __launch_bounds__(256) void k_diag_stats(int64_t n, const double *dg, const double *b, double *part) {
    __shared__ double red[kRed];
    double nonpos = 0.0, bad = 0.0, bb = 0.0;
    GRID_STRIDE(i, n) {
        const double v = dg[i];
        if (!(v > 0.0)) nonpos += 1.0;
        if (!isfinite(v)) bad += 1.0;
        if (b) {
            const double bv = b[i];
            bb += bv * bv;
            if (!isfinite(bv)) bad += 1.0;
        }
    }
    write_parts(part, nonpos, bad, bb, red);
}
__global__ __launch_bounds__(256) void k_final_parts(const double *part, double *out3, int nv) {
    __shared__ double red[kRed];
    double t[5];  // nv <= 5 (checked by the host)
    sum_parts_n<5>(part, 0, t, red);
    if (threadIdx.x == 0)
        for (int k = 0; k < nv; ++k) out3[k] = t[k];
}
__global__ __launch_bounds__(256) void k_final_parts8(const double *part, double *out8) {
    __shared__ double red[8 * 4];
    double t[8];
    sum_parts_n<8>(part, 0, t, red);
    if (threadIdx.x < 8) {
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (k == (int)threadIdx.x) v = t[k];
        out8[threadIdx.x] = v;
    }
}

// box-KKT active-set bookkeeping (box_kkt_h8.py:672-685, 731-733):
//   grad = A cand - b ; add/release masks ; counts of each.
// states: lower_active, upper_active, prescribed (lb == ub).
// counts: added lower, added upper, released lower, released upper, free, active lower,
// active upper (the last three of the masks the pass reads), free after the update.
__global__ __launch_bounds__(256) void k_box_update(int64_t n, const double *Acand, const double *b, const double *cand,
                                                    const double *lo, const double *up, uint8_t *lact, uint8_t *uact,
                                                    const uint8_t *presc, double kkt_tol, int apply, double *grad,
                                                    double *part) {
    __shared__ double red[8 * 4];
    double cnt[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    GRID_STRIDE(i, n) {
        const double g = Acand[i] - b[i];
        grad[i] = g;
        const bool la = lact[i], ua = uact[i], pr = presc[i], fixed = la || ua || pr;
        const bool add_l = !fixed && cand[i] < lo[i];
        const bool add_u = !fixed && cand[i] > up[i];
        const bool rel_l = la && g < -kkt_tol;
        const bool rel_u = ua && g > kkt_tol;
        const bool la1 = (la || add_l) && !rel_l, ua1 = (ua || add_u) && !rel_u;
        cnt[0] += add_l ? 1.0 : 0.0;
        cnt[1] += add_u ? 1.0 : 0.0;
        cnt[2] += rel_l ? 1.0 : 0.0;
        cnt[3] += rel_u ? 1.0 : 0.0;
        cnt[4] += fixed ? 0.0 : 1.0;
        cnt[5] += la ? 1.0 : 0.0;
        cnt[6] += ua ? 1.0 : 0.0;
        cnt[7] += (apply ? (la1 || ua1 || pr) : fixed) ? 0.0 : 1.0;
        if (apply) {
            lact[i] = la1;
            uact[i] = ua1;
        }
    }
    block_sum_n<256, 8>(cnt, red);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 8; ++k) part[k * kVecBlocks + blockIdx.x] = cnt[k];
}

// KKT audit of a box solution in one pass (box_kkt_h8.py:459-558): per block the maxima of
// |grad| on free entries, of the projected gradient and of the two complementarity products,
// and the counts of sign violations and bound violations.  The maxima propagate NaN, as
// torch's max does.
__device__ __forceinline__ double nan_max(double a, double b) { return (a != a || b != b) ? NAN : fmax(a, b); }

__global__ __launch_bounds__(256) void k_kkt_stats(int64_t n, const double *grad, const double *sol,
                                                   const double *lo, const double *up, const uint8_t *lact,
                                                   const uint8_t *uact, const uint8_t *presc, double kkt_tol,
                                                   double *part) {
    __shared__ double red[8 * 4];
    double mx[3] = {0.0, 0.0, 0.0}, cnt[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    GRID_STRIDE(i, n) {
        const double g = grad[i], x = sol[i], l = lo[i], u = up[i];
        const bool la = lact[i], ua = uact[i], fr = !(la || ua || presc[i]);
        double proj = fr ? g : 0.0;
        if (la) proj = fmin(g, 0.0);
        if (ua) proj = fmax(g, 0.0);
        if (la && g != g) proj = g;  // keep a NaN gradient visible
        if (ua && g != g) proj = g;
        const double lm = la ? (g != g ? g : fmax(g, 0.0)) : 0.0;
        const double um = ua ? (g != g ? g : fmax(-g, 0.0)) : 0.0;
        mx[0] = nan_max(mx[0], fr ? fabs(g) : 0.0);
        mx[1] = nan_max(mx[1], fabs(proj));
        mx[2] = nan_max(mx[2], nan_max(fabs(lm * (x - l)), fabs(um * (u - x))));
        cnt[0] += (la && g < -kkt_tol) ? 1.0 : 0.0;
        cnt[1] += (ua && -g < -kkt_tol) ? 1.0 : 0.0;
        cnt[2] += x < l ? 1.0 : 0.0;
        cnt[3] += x > u ? 1.0 : 0.0;
        cnt[4] += l > u ? 1.0 : 0.0;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int k = 0; k < 3; ++k) mx[k] = nan_max(mx[k], __shfl_xor(mx[k], o, kWave));
    block_sum_n<256, 5>(cnt, red);  // its barriers also order the LDS use below
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 3; ++k) red[k * 4 + w] = mx[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k)
            part[k * kVecBlocks + blockIdx.x] = nan_max(nan_max(red[k * 4], red[k * 4 + 1]),
                                                        nan_max(red[k * 4 + 2], red[k * 4 + 3]));
#pragma unroll
        for (int k = 0; k < 5; ++k) part[(3 + k) * kVecBlocks + blockIdx.x] = cnt[k];
    }
}
__global__ __launch_bounds__(256) void k_kkt_final(const double *part, int nb, double *out8) {
    if (threadIdx.x >= 8) return;
    const int k = threadIdx.x;
    double v = 0.0;
    for (int b = 0; b < nb; ++b) v = k < 3 ? nan_max(v, part[k * kVecBlocks + b]) : v + part[k * kVecBlocks + b];
    out8[k] = v;
}

// ---------------------------------------------------------------------------------------
// host drivers
// ---------------------------------------------------------------------------------------

typedef int (*ApplyFn)(void *ctx, const double *in, double *out, hipStream_t st);
typedef ApplyFn PrecondFn;  // z = M^-1 r (zero on the fixed set)

// Scratch of the solves on one plan: five Krylov vectors, the constrained-solve vectors
// (x_c, masked rhs and diagonal, x0, result, free mask), the flag state and its pinned
// readback.  Sized for n_mech = 3n + 6 (the damage solves use the first n entries).
struct Work {
    double *part = nullptr;
    PcgState *S = nullptr;
    PcgState *hS = nullptr;
    double *h3 = nullptr, *d3 = nullptr;
    double *r = nullptr, *p = nullptr, *z = nullptr, *Ap = nullptr, *dg = nullptr;
    double *xc = nullptr, *bf = nullptr, *df = nullptr, *x0f = nullptr, *y = nullptr;
    uint8_t *freem = nullptr;
    int64_t cap = 0;

    int ensure(int64_t n) {
        if (!part) {
            PHAST_HIP(hipMalloc(&part, sizeof(double) * 8 * kVecBlocks));
            PHAST_HIP(hipMalloc(&S, sizeof(PcgState)));
            PHAST_HIP(hipHostMalloc(&hS, sizeof(PcgState), hipHostMallocDefault));
            PHAST_HIP(hipHostMalloc(&h3, sizeof(double) * 8, hipHostMallocDefault));
            PHAST_HIP(hipMalloc(&d3, sizeof(double) * 8));
        }
        if (n > cap) {
            double **vec[] = {&r, &p, &z, &Ap, &dg, &xc, &bf, &df, &x0f, &y};
            for (double **v : vec) {
                (void)hipFree(*v);
                *v = nullptr;
            }
            (void)hipFree(freem);
            freem = nullptr;
            cap = 0;
            for (double **v : vec) PHAST_HIP(hipMalloc(v, sizeof(double) * n));
            PHAST_HIP(hipMalloc(&freem, (size_t)n));
            cap = n;
        }
        return PHAST_OK;
    }
    void release() {
        (void)hipFree(part);
        (void)hipFree(S);
        if (hS) (void)hipHostFree(hS);
        if (h3) (void)hipHostFree(h3);
        (void)hipFree(d3);
        double *vec[] = {r, p, z, Ap, dg, xc, bf, df, x0f, y};
        for (double *v : vec) (void)hipFree(v);
        (void)hipFree(freem);
        *this = Work();
    }
};

// The plan's scratch, created on first use and grown to n doubles per vector.
static int plan_work(const phast_hex8_plan *pl, int64_t n, Work **out) {
    if (!pl->solver_work) pl->solver_work = new Work();
    Work *W = static_cast<Work *>(pl->solver_work);
    if (int rc = W->ensure(n)) return rc;
    *out = W;
    return PHAST_OK;
}

void release_solver_work(const phast_hex8_plan *pl) {
    if (!pl->solver_work) return;
    Work *W = static_cast<Work *>(pl->solver_work);
    W->release();
    delete W;
    pl->solver_work = nullptr;
}

static inline int nblk(int64_t n) { return (int)std::min<int64_t>(kVecBlocks, ceil_div(std::max<int64_t>(n, 1), 256)); }

// nv (<= 5) fixed-order sums of the current partials -> host
static int readn(Work &W, hipStream_t st, double *out, int nv) {
    PHAST_REQUIRE(nv >= 1 && nv <= 5, "internal: %d partial rows requested (at most 5)", nv);
    hipLaunchKernelGGL(k_final_parts, dim3(1), dim3(256), 0, st, W.part, W.d3, nv);
    PHAST_LAUNCHED();
    PHAST_HIP(hipMemcpyAsync(W.h3, W.d3, sizeof(double) * nv, hipMemcpyDeviceToHost, st));
    PHAST_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < nv; ++k) out[k] = W.h3[k];
    return PHAST_OK;
}
static int read3(Work &W, hipStream_t st, double out[3]) { return readn(W, st, out, 3); }

static int zero_parts(Work &W, hipStream_t st) {
    PHAST_HIP(hipMemsetAsync(W.part, 0, sizeof(double) * 8 * kVecBlocks, st));
    return PHAST_OK;
}

// linalg.pcg on the free set: `free_apply` must return the masked action.
// info: [iterations, converged, rhs_norm, residual_norm, relative_residual]; returns
// PHAST_ERR_NONFINITE / PHAST_ERR_NONPOSITIVE on breakdown (messages as in the reference).
static int pcg_flags(Work &W, ApplyFn apply, void *ctx, int64_t n, const uint8_t *fixed, const double *b,
                     const double *diag, const double *x0, double *x, double rtol, double atol, int maxiter,
                     int check_every, double *info, hipStream_t st, PrecondFn precond = nullptr,
                     void *pctx = nullptr) {
    const int nb = nblk(n);
    // stats: ||b||, finite(b), diag > 0, finite(diag)   (linalg.py:139-155)
    PHAST_TRY0(zero_parts(W, st));
    hipLaunchKernelGGL(k_diag_stats, dim3(nb), dim3(256), 0, st, n, diag, b, W.part);
    PHAST_LAUNCHED();
    double s3[3];
    if (int rc = read3(W, st, s3)) return rc;
    const double b_norm = std::sqrt(s3[2]);
    if (!std::isfinite(b_norm)) {
        set_error("PCG right-hand side contains a non-finite value");
        return PHAST_ERR_NONFINITE;
    }
    if (s3[1] > 0.0) {
        set_error("PCG diagonal contains a non-finite value");
        return PHAST_ERR_NONFINITE;
    }
    if (s3[0] > 0.0) {
        set_error("PCG diagonal has a non-positive entry (non-positive curvature)");
        return PHAST_ERR_NONPOSITIVE;
    }
    if (b_norm == 0.0) {
        hipLaunchKernelGGL(k_axpby_mask, dim3(nb), dim3(256), 0, st, n, 0.0, b, 0.0, (const double *)nullptr,
                           (const uint8_t *)nullptr, 0.0, x);
        PHAST_LAUNCHED();
        info[0] = 0; info[1] = 1; info[2] = 0; info[3] = 0; info[4] = 0;
        return PHAST_OK;
    }
    const double threshold = std::max(rtol * b_norm, atol);
    double r_norm = b_norm;
    if (x0) {
        PHAST_HIP(hipMemcpyAsync(x, x0, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
        if (int rc = apply(ctx, x, W.Ap, st)) return rc;
        hipLaunchKernelGGL(k_axpby_mask, dim3(nb), dim3(256), 0, st, n, 1.0, b, -1.0, W.Ap, fixed, 0.0, W.r);
        PHAST_LAUNCHED();
        PHAST_TRY0(zero_parts(W, st));
        hipLaunchKernelGGL(k_dots, dim3(nb), dim3(256), 0, st, n, W.r, (const double *)nullptr, W.part);
        PHAST_LAUNCHED();
        if (int rc = read3(W, st, s3)) return rc;
        r_norm = std::sqrt(s3[0]);
        if (!std::isfinite(r_norm)) {
            set_error("PCG initial residual is not finite");
            return PHAST_ERR_NONFINITE;
        }
    } else {
        PHAST_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, st));
        PHAST_HIP(hipMemcpyAsync(W.r, b, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    }
    if (r_norm <= threshold) {
        info[0] = 0; info[1] = 1; info[2] = b_norm; info[3] = r_norm; info[4] = r_norm / b_norm;
        return PHAST_OK;
    }
    // inv_diag, z = r inv_diag, p = z, rz = r.z, rr = r.r
    hipLaunchKernelGGL(k_invert, dim3(nb), dim3(256), 0, st, n, diag, W.dg);
    PHAST_LAUNCHED();
    {
        PcgState h;
        std::memset(&h, 0, sizeof(h));
        h.active = 1;
        h.thr2 = threshold * threshold;
        h.b_norm = b_norm;
        h.alpha = 0.0;
        *W.hS = h;
        PHAST_HIP(hipMemcpyAsync(W.S, W.hS, sizeof(PcgState), hipMemcpyHostToDevice, st));
        // z = r * inv_diag via an update with alpha = 0 (or z = M^-1 r)
        if (precond) {
            if (int rc = precond(pctx, W.r, W.z, st)) return rc;
            hipLaunchKernelGGL(k_rz, dim3(nb), dim3(256), 0, st, n, W.r, W.z, (const PcgState *)nullptr, W.part);
        } else {
            hipLaunchKernelGGL((k_update<0>), dim3(nb), dim3(256), 0, st, n, x, W.r, W.r, W.r, W.dg, W.z, W.S, W.part);
        }
        PHAST_LAUNCHED();
        PHAST_HIP(hipMemcpyAsync(W.p, W.z, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_scalar_flags_b, dim3(1), dim3(256), 0, st, W.S, W.part);  // sets rz, rr (active stays)
        PHAST_LAUNCHED();
    }
    // restore beta-free state: k_scalar_flags_b may clear active if r is already tiny (it is not:
    // r_norm > threshold was checked above).
    for (int it = 1; it <= maxiter; ++it) {
        if (int rc = apply(ctx, W.p, W.Ap, st)) return rc;
        hipLaunchKernelGGL(k_mask_dot, dim3(nb), dim3(256), 0, st, n, W.p, W.Ap, fixed, (const PcgState *)nullptr,
                           W.part);
        PHAST_LAUNCHED();
        // x, r, z update with the step the partials give (alpha = 0 for inactive iterations
        // leaves x and r unchanged), then the direction
        if (precond) {
            hipLaunchKernelGGL((k_update_flags<2>), dim3(nb), dim3(256), 0, st, n, x, W.r, W.p, W.Ap, W.dg, W.z, W.S,
                               W.part);
            PHAST_LAUNCHED();
            if (int rc = precond(pctx, W.r, W.z, st)) return rc;
            hipLaunchKernelGGL(k_rz, dim3(nb), dim3(256), 0, st, n, W.r, W.z, (const PcgState *)nullptr, W.part);
            PHAST_LAUNCHED();
            hipLaunchKernelGGL(k_pupdate_flags, dim3(nb), dim3(256), 0, st, n, W.p, W.z, W.S, W.part, 0);
        } else {
            hipLaunchKernelGGL((k_update_flags<0>), dim3(nb), dim3(256), 0, st, n, x, W.r, W.p, W.Ap, W.dg, W.z, W.S,
                               W.part);
            PHAST_LAUNCHED();
            hipLaunchKernelGGL(k_pupdate_flags, dim3(nb), dim3(256), 0, st, n, W.p, W.z, W.S, W.part, 4);
        }
        PHAST_LAUNCHED();
        if (it % check_every == 0 || it == maxiter) {
            PHAST_HIP(hipMemcpyAsync(W.hS, W.S, sizeof(PcgState), hipMemcpyDeviceToHost, st));
            PHAST_HIP(hipStreamSynchronize(st));
            const PcgState &h = *W.hS;
            if (h.nonfinite) {
                set_error("PCG produced a non-finite value by iteration %d", it);
                return PHAST_ERR_NONFINITE;
            }
            if (h.breakdown) {
                set_error("PCG met non-positive curvature p.A p <= 0 by iteration %d; the operator is not positive "
                          "definite on the free set", it);
                return PHAST_ERR_NONPOSITIVE;
            }
            r_norm = std::sqrt(h.rr);
            if (!h.active) {
                info[0] = h.steps; info[1] = 1; info[2] = b_norm; info[3] = r_norm; info[4] = r_norm / b_norm;
                return PHAST_OK;
            }
        }
    }
    info[0] = maxiter; info[1] = 0; info[2] = b_norm; info[3] = r_norm; info[4] = r_norm / b_norm;
    return PHAST_OK;
}

// box_kkt_h8._pcg on the free set (vectors zero on `fixed`).  rhs must already be masked.
// info: [iterations, converged, initial_norm, residual_norm, relative_residual]
static int pcg_stop(Work &W, ApplyFn apply, void *ctx, int64_t n, const uint8_t *fixed, const double *rhs,
                    const double *diag, const double *x0, double *x, double tol, int max_iter, int poll, double *info,
                    hipStream_t st) {
    const int nb = nblk(n);
    // diag: the reduced diagonal; on fixed entries use 1 (vectors are zero there anyway)
    hipLaunchKernelGGL(k_axpby_mask, dim3(nb), dim3(256), 0, st, n, 1.0, diag, 0.0, (const double *)nullptr, fixed,
                       1.0, W.dg);
    PHAST_LAUNCHED();
    if (x0) {
        hipLaunchKernelGGL(k_axpby_mask, dim3(nb), dim3(256), 0, st, n, 1.0, x0, 0.0, (const double *)nullptr, fixed,
                           0.0, x);
        PHAST_LAUNCHED();
    } else {
        PHAST_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, st));
    }
    if (int rc = apply(ctx, x, W.Ap, st)) return rc;
    hipLaunchKernelGGL(k_axpby_mask, dim3(nb), dim3(256), 0, st, n, 1.0, rhs, -1.0, W.Ap, fixed, 0.0, W.r);
    PHAST_LAUNCHED();
    // residual norm, its finiteness and the preconditioned residual (z = r / diag, r.z) in one
    // pass and one read
    PHAST_TRY0(zero_parts(W, st));
    hipLaunchKernelGGL(k_start_stop, dim3(nb), dim3(256), 0, st, n, W.r, W.dg, W.z, W.part);
    PHAST_LAUNCHED();
    double s3[3];
    if (int rc = read3(W, st, s3)) return rc;  // [r.z, r.r, non-finite count]
    if (s3[2] > 0.0) {
        set_error("PCG action returned a non-finite vector");
        return PHAST_ERR_NONFINITE;
    }
    const double initial = std::sqrt(s3[1]);
    const double ref = std::max(initial, 1.0);
    if (initial <= tol * ref) {
        info[0] = 0; info[1] = 1; info[2] = initial; info[3] = initial; info[4] = initial / ref;
        return PHAST_OK;
    }
    // direction = preconditioned residual; rz = r.preconditioned
    {
        if (!std::isfinite(s3[0]) || s3[0] <= 0.0) {
            set_error("PCG encountered a non-positive initial residual product");
            return PHAST_ERR_NONPOSITIVE;
        }
        PcgState h;
        std::memset(&h, 0, sizeof(h));
        h.thr2 = tol * ref;
        h.rz = s3[0];
        *W.hS = h;
        PHAST_HIP(hipMemcpyAsync(W.S, W.hS, sizeof(PcgState), hipMemcpyHostToDevice, st));
        PHAST_HIP(hipMemcpyAsync(W.p, W.z, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    }
    int it = 0;
    while (it < max_iter) {
        if (int rc = apply(ctx, W.p, W.Ap, st)) return rc;
        hipLaunchKernelGGL(k_mask_dot, dim3(nb), dim3(256), 0, st, n, W.p, W.Ap, fixed, W.S, W.part);
        PHAST_LAUNCHED();
        hipLaunchKernelGGL(k_update_stop, dim3(nb), dim3(256), 0, st, n, x, W.r, W.p, W.Ap, W.dg, W.z, W.S, W.part);
        PHAST_LAUNCHED();
        hipLaunchKernelGGL(k_pupdate_stop, dim3(nb), dim3(256), 0, st, n, W.p, W.z, W.S, W.part, max_iter);
        PHAST_LAUNCHED();
        ++it;
        if (it % poll == 0 || it == max_iter) {
            PHAST_HIP(hipMemcpyAsync(W.hS, W.S, sizeof(PcgState), hipMemcpyDeviceToHost, st));
            PHAST_HIP(hipStreamSynchronize(st));
            if (W.hS->done) break;
        }
    }
    const PcgState &h = *W.hS;
    const double rn = std::sqrt(h.rr);
    if (h.status == PS_NONFINITE) {
        set_error("PCG action returned a non-finite vector");
        return PHAST_ERR_NONFINITE;
    }
    if (h.status == PS_NOT_PD) {
        set_error("PCG operator is not positive definite on the free space");
        return PHAST_ERR_NONPOSITIVE;
    }
    if (h.status == PS_NONPOS_RZ) {
        set_error("PCG encountered a non-positive residual product");
        return PHAST_ERR_NONPOSITIVE;
    }
    info[0] = h.iter;
    info[1] = h.status == PS_CONVERGED ? 1 : 0;
    info[2] = initial;
    info[3] = rn;
    info[4] = rn / ref;
    return PHAST_OK;
}

}  // namespace phast

using namespace phast;

// ---------------------------------------------------------------------------------------
// operator adaptors and C entry points
// ---------------------------------------------------------------------------------------

struct MechCtx {
    const phast_hex8_plan *plan;
    const double *g, *d;
    const uint8_t *fixed;
    int64_t n;
    const uint32_t *pat;  // crack closure: the fixed sign pattern (null: no closure)
};

static int mech_apply_cb(void *c, const double *in, double *out, hipStream_t st) {
    auto *m = static_cast<MechCtx *>(c);
    int rc = m->pat  ? phast_hex8_mech_action_cl(m->plan, in, m->g, m->pat, 2, out, st)
             : m->d ? phast_hex8_mech_action_d(m->plan, in, m->d, out, st)
                    : phast_hex8_mech_action(m->plan, in, m->g, out, st);
    if (rc) return rc;
    if (m->fixed) {
        hipLaunchKernelGGL(k_zero_fixed, dim3(nblk(m->n)), dim3(256), 0, st, m->n, m->fixed, out);
        PHAST_LAUNCHED();
    }
    return PHAST_OK;
}

// The scalar damage systems: form 0 the AT2 operator with history H, form 1 the PF-CZM
// Newton tangent with the curvatures W in H's place.
struct DmgCtx {
    const phast_hex8_plan *plan;
    const double *H;
    double scale;
    int form;
};

static int dmg_apply_cb(void *c, const double *in, double *out, hipStream_t st) {
    auto *m = static_cast<DmgCtx *>(c);
    if (m->form == 1) return phast_hex8_pfczm_tangent_action(m->plan, in, m->H, m->scale, out, st);
    return phast_hex8_damage_action(m->plan, in, m->H, m->scale, out, st);
}

__global__ void k_not_mask(int64_t n, const uint8_t *fixed, uint8_t *freem) {
    GRID_STRIDE(i, n) freem[i] = fixed[i] ? 0 : 1;
}

// solid3d/linalg.py:223-287 solve_constrained with the hex8 mechanics action A(g)
// (d given: g formed in-kernel from nodal damage).  info[5] as in pcg_flags.
struct phast_hex8_mg;
extern "C" int phast_hex8_mg_apply(phast_hex8_mg *mg, const double *r, double *z, hipStream_t st);

static int mg_precond_cb(void *c, const double *in, double *out, hipStream_t st) {
    return phast_hex8_mg_apply(static_cast<phast_hex8_mg *>(c), in, out, st);
}

static int mech_solve_impl(const phast_hex8_plan *plan, phast_hex8_mg *mg, const double *g, const double *d,
                           const uint8_t *fixed, const double *fixed_values, const double *rhs, const double *x0,
                           const double *diag, double rtol, double atol, int32_t maxiter, int32_t check_every,
                           int64_t n_mech, double *x_out, double *info, hipStream_t st,
                           const uint32_t *pat = nullptr);

extern "C" int phast_hex8_mech_solve(const phast_hex8_plan *plan, const double *g, const double *d,
                                     const uint8_t *fixed, const double *fixed_values, const double *rhs,
                                     const double *x0, const double *diag, double rtol, double atol, int32_t maxiter,
                                     int32_t check_every, int64_t n_mech, double *x_out, double *info,
                                     hipStream_t st) {
    return mech_solve_impl(plan, nullptr, g, d, fixed, fixed_values, rhs, x0, diag, rtol, atol, maxiter, check_every,
                           n_mech, x_out, info, st);
}

// The same solve with the crack closure at the fixed sign pattern `pat` (A_s(g) of
// operator.py:1039-1105; g required).
extern "C" int phast_hex8_mech_solve_cl(const phast_hex8_plan *plan, const double *g, const uint32_t *pat,
                                        const uint8_t *fixed, const double *fixed_values, const double *rhs,
                                        const double *x0, const double *diag, double rtol, double atol,
                                        int32_t maxiter, int32_t check_every, int64_t n_mech, double *x_out,
                                        double *info, hipStream_t st) {
    PHAST_REQUIRE(g && pat, "phast_hex8_mech_solve_cl: g and the pattern are required");
    return mech_solve_impl(plan, nullptr, g, nullptr, fixed, fixed_values, rhs, x0, diag, rtol, atol, maxiter,
                           check_every, n_mech, x_out, info, st, pat);
}

// The same solve preconditioned by the voxel multigrid V-cycle (phast_hex8_mg_update must
// have been called with the same g / d).
extern "C" int phast_hex8_mech_solve_mg(const phast_hex8_plan *plan, phast_hex8_mg *mg, const double *g,
                                        const double *d, const uint8_t *fixed, const double *fixed_values,
                                        const double *rhs, const double *x0, const double *diag, double rtol,
                                        double atol, int32_t maxiter, int32_t check_every, int64_t n_mech,
                                        double *x_out, double *info, hipStream_t st) {
    PHAST_REQUIRE(mg, "phast_hex8_mech_solve_mg: null multigrid");
    return mech_solve_impl(plan, mg, g, d, fixed, fixed_values, rhs, x0, diag, rtol, atol, maxiter, check_every,
                           n_mech, x_out, info, st);
}

static int mech_solve_impl(const phast_hex8_plan *plan, phast_hex8_mg *mg, const double *g, const double *d,
                           const uint8_t *fixed, const double *fixed_values, const double *rhs, const double *x0,
                           const double *diag, double rtol, double atol, int32_t maxiter, int32_t check_every,
                           int64_t n_mech, double *x_out, double *info, hipStream_t st, const uint32_t *pat) {
    PHAST_REQUIRE(plan && fixed && fixed_values && diag && x_out && info, "phast_hex8_mech_solve: null argument");
    PHAST_REQUIRE(maxiter >= 1 && check_every >= 1, "maxiter and check_every must be >= 1");
    PHAST_REQUIRE(!(rtol == 0.0 && atol == 0.0) && rtol >= 0.0 && atol >= 0.0, "rtol and atol must not both be zero");
    Work *Wp = nullptr;
    if (int rc = plan_work(plan, n_mech, &Wp)) return rc;
    Work &W = *Wp;
    const int64_t n = n_mech;
    const int nb = nblk(n);
    double *xc = W.xc, *bf = W.bf, *df = W.df, *x0f = x0 ? W.x0f : nullptr, *y = W.y;
    uint8_t *freem = W.freem;
    hipLaunchKernelGGL(k_not_mask, dim3(nb), dim3(256), 0, st, n, fixed, freem);
    // x_c = where(fixed, values, 0); image = A x_c; b_f = where(free, rhs - image, 0)
    hipLaunchKernelGGL(k_where, dim3(nb), dim3(256), 0, st, n, fixed, fixed_values, (const double *)nullptr, 0.0, xc);
    PHAST_LAUNCHED();
    MechCtx raw{plan, g, d, nullptr, n, pat};
    int rc = mech_apply_cb(&raw, xc, W.Ap, st);
    if (rc == PHAST_OK) {
        hipLaunchKernelGGL(k_axpby_mask, dim3(nb), dim3(256), 0, st, n, rhs ? 1.0 : 0.0, rhs ? rhs : W.Ap, -1.0, W.Ap,
                           fixed, 0.0, bf);
        hipLaunchKernelGGL(k_where, dim3(nb), dim3(256), 0, st, n, freem, diag, (const double *)nullptr, 1.0, df);
        if (x0) hipLaunchKernelGGL(k_where, dim3(nb), dim3(256), 0, st, n, freem, x0, (const double *)nullptr, 0.0, x0f);
        // the apply leaves the fixed entries of A p as they are: k_mask_dot zeroes them before
        // anything reads them
        MechCtx ctx{plan, g, d, nullptr, n, pat};
        rc = pcg_flags(W, mech_apply_cb, &ctx, n, fixed, bf, df, x0f, y, rtol, atol, maxiter, check_every, info, st,
                       mg ? mg_precond_cb : nullptr, mg);
        // x = where(free, y, x_c)
        if (rc == PHAST_OK)
            hipLaunchKernelGGL(k_where, dim3(nb), dim3(256), 0, st, n, freem, y, xc, 0.0, x_out);
    }
    return rc;
}

// box_kkt_h8.py:350-432 solve_h8_scalar_dirichlet_pcg for the scaled AT2 damage system
// (operator.py:1496-1509).  b and diag are the scaled rhs / diagonal.  info[5] as pcg_stop.
static int dirichlet_pcg_impl(const phast_hex8_plan *plan, int form, const double *H, double scale, const double *b,
                              const double *diag, const uint8_t *fixed, const double *values, const double *initial,
                              double tol, int32_t max_iter, int64_t n, double *out, double *info, hipStream_t st);

extern "C" int phast_hex8_damage_dirichlet_pcg(const phast_hex8_plan *plan, const double *H, double scale,
                                               const double *b, const double *diag, const uint8_t *fixed,
                                               const double *values, const double *initial, double tol,
                                               int32_t max_iter, int64_t n, double *out, double *info,
                                               hipStream_t st) {
    return dirichlet_pcg_impl(plan, 0, H, scale, b, diag, fixed, values, initial, tol, max_iter, n, out, info, st);
}

// The same solve on the PF-CZM Newton tangent with curvatures W (operator.py:566-609).
extern "C" int phast_hex8_pfczm_dirichlet_pcg(const phast_hex8_plan *plan, const double *W, double scale,
                                              const double *b, const double *diag, const uint8_t *fixed,
                                              const double *values, const double *initial, double tol,
                                              int32_t max_iter, int64_t n, double *out, double *info,
                                              hipStream_t st) {
    PHAST_REQUIRE(W, "phast_hex8_pfczm_dirichlet_pcg: null curvatures");
    return dirichlet_pcg_impl(plan, 1, W, scale, b, diag, fixed, values, initial, tol, max_iter, n, out, info, st);
}

static int dirichlet_pcg_impl(const phast_hex8_plan *plan, int form, const double *H, double scale, const double *b,
                              const double *diag, const uint8_t *fixed, const double *values, const double *initial,
                              double tol, int32_t max_iter, int64_t n, double *out, double *info, hipStream_t st) {
    PHAST_REQUIRE(plan && b && diag && fixed && values && out && info, "phast_hex8_damage_dirichlet_pcg: null argument");
    PHAST_REQUIRE(tol > 0.0 && max_iter >= 1, "tolerance must be positive and max_iterations >= 1");
    Work *Wp = nullptr;
    if (int rc = plan_work(plan, 3 * n + 6, &Wp)) return rc;
    Work &W = *Wp;
    const int nb = nblk(n);
    double *known = W.xc, *rr = W.bf, *y = W.y;
    hipLaunchKernelGGL(k_where, dim3(nb), dim3(256), 0, st, n, fixed, values, (const double *)nullptr, 0.0, known);
    DmgCtx ctx{plan, H, scale, form};
    int rc = dmg_apply_cb(&ctx, known, W.Ap, st);
    if (rc == PHAST_OK) {
        // reduced rhs = (b - A known)[free]
        hipLaunchKernelGGL(k_axpby_mask, dim3(nb), dim3(256), 0, st, n, 1.0, b, -1.0, W.Ap, fixed, 0.0, rr);
        // pcg_stop masks the fixed entries of every image it reads (k_axpby_mask, k_mask_dot),
        // so the plain action serves as the masked one
        rc = pcg_stop(W, dmg_apply_cb, &ctx, n, fixed, rr, diag, initial, y, tol, max_iter, 4, info, st);
        // solution = known with free entries from y
        if (rc == PHAST_OK)
            hipLaunchKernelGGL(k_where, dim3(nb), dim3(256), 0, st, n, fixed, known, y, 0.0, out);
    }
    return rc;
}

// One active-set bookkeeping pass: grad = A(cand) - b (scaled system), add/release counts
// and, with apply != 0, the updated masks.  counts: [added_lower, added_upper,
// released_lower, released_upper, free]; grad (n) is written.
static int box_update_impl(const phast_hex8_plan *plan, int form, const double *H, double scale, const double *b,
                           const double *cand, const double *lower, const double *upper, uint8_t *lower_active,
                           uint8_t *upper_active, const uint8_t *prescribed, double kkt_tol, int32_t apply, int64_t n,
                           double *grad, int64_t *counts, hipStream_t st);

extern "C" int phast_hex8_box_update(const phast_hex8_plan *plan, const double *H, double scale, const double *b,
                                     const double *cand, const double *lower, const double *upper,
                                     uint8_t *lower_active, uint8_t *upper_active, const uint8_t *prescribed,
                                     double kkt_tol, int32_t apply, int64_t n, double *grad, int64_t *counts,
                                     hipStream_t st) {
    return box_update_impl(plan, 0, H, scale, b, cand, lower, upper, lower_active, upper_active, prescribed, kkt_tol,
                           apply, n, grad, counts, st);
}

// The same pass on the PF-CZM Newton tangent with curvatures W.
extern "C" int phast_hex8_pfczm_box_update(const phast_hex8_plan *plan, const double *W, double scale,
                                           const double *b, const double *cand, const double *lower,
                                           const double *upper, uint8_t *lower_active, uint8_t *upper_active,
                                           const uint8_t *prescribed, double kkt_tol, int32_t apply, int64_t n,
                                           double *grad, int64_t *counts, hipStream_t st) {
    PHAST_REQUIRE(W, "phast_hex8_pfczm_box_update: null curvatures");
    return box_update_impl(plan, 1, W, scale, b, cand, lower, upper, lower_active, upper_active, prescribed, kkt_tol,
                           apply, n, grad, counts, st);
}

static int box_update_impl(const phast_hex8_plan *plan, int form, const double *H, double scale, const double *b,
                           const double *cand, const double *lower, const double *upper, uint8_t *lower_active,
                           uint8_t *upper_active, const uint8_t *prescribed, double kkt_tol, int32_t apply, int64_t n,
                           double *grad, int64_t *counts, hipStream_t st) {
    PHAST_REQUIRE(plan && b && cand && lower && upper && lower_active && upper_active && prescribed && grad && counts,
                  "phast_hex8_box_update: null argument");
    Work *Wp = nullptr;
    if (int rc = plan_work(plan, 3 * n + 6, &Wp)) return rc;
    Work &W = *Wp;
    DmgCtx ctx{plan, H, scale, form};
    if (int rc = dmg_apply_cb(&ctx, cand, W.Ap, st)) return rc;
    PHAST_TRY0(zero_parts(W, st));
    hipLaunchKernelGGL(k_box_update, dim3(nblk(n)), dim3(256), 0, st, n, W.Ap, b, cand, lower, upper, lower_active,
                       upper_active, prescribed, kkt_tol, apply, grad, W.part);
    PHAST_LAUNCHED();
    hipLaunchKernelGGL(k_final_parts8, dim3(1), dim3(256), 0, st, W.part, W.d3);
    PHAST_LAUNCHED();
    PHAST_HIP(hipMemcpyAsync(W.h3, W.d3, sizeof(double) * 8, hipMemcpyDeviceToHost, st));
    PHAST_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < 8; ++k) counts[k] = (int64_t)W.h3[k];
    return PHAST_OK;
}

// The KKT audit of a box solution: out8 (host) = max |grad| on free entries, max projected
// gradient, max complementarity, lower / upper sign violations, entries below lower, above
// upper, bounds with lower > upper (box_kkt_h8.py:459-558).
extern "C" int phast_hex8_kkt_stats(const phast_hex8_plan *plan, const double *grad, const double *sol,
                                    const double *lower, const double *upper, const uint8_t *lower_active,
                                    const uint8_t *upper_active, const uint8_t *prescribed, double kkt_tol, int64_t n,
                                    double *out8, hipStream_t st) {
    PHAST_REQUIRE(plan && grad && sol && lower && upper && lower_active && upper_active && prescribed && out8 && n >= 1,
                  "phast_hex8_kkt_stats: null argument");
    Work *Wp = nullptr;
    if (int rc = plan_work(plan, 3 * n + 6, &Wp)) return rc;
    Work &W = *Wp;
    const int nb = nblk(n);
    hipLaunchKernelGGL(k_kkt_stats, dim3(nb), dim3(256), 0, st, n, grad, sol, lower, upper, lower_active, upper_active,
                       prescribed, kkt_tol, W.part);
    PHAST_LAUNCHED();
    hipLaunchKernelGGL(k_kkt_final, dim3(1), dim3(64), 0, st, W.part, nb, W.d3);
    PHAST_LAUNCHED();
    PHAST_HIP(hipMemcpyAsync(W.h3, W.d3, sizeof(double) * 8, hipMemcpyDeviceToHost, st));
    PHAST_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < 8; ++k) out8[k] = W.h3[k];
    return PHAST_OK;
}

// Allocate the plan's solver scratch now (set-up time), so that no later solve on the plan
// allocates.
extern "C" int phast_hex8_solver_reserve(const phast_hex8_plan *plan) {
    PHAST_REQUIRE(plan, "phast_hex8_solver_reserve: null plan");
    Work *W = nullptr;
    return plan_work(plan, 3 * plan->n + 6, &W);
}

extern "C" int phast_hex8_solver_release(const phast_hex8_plan *plan) {
    PHAST_REQUIRE(plan, "phast_hex8_solver_release: null plan");
    release_solver_work(plan);
    return PHAST_OK;
}

// linalg.pcg (solid3d/linalg.py:99-220) of an operator the caller evaluates: before each
// apply the library writes the vector into `in_buf`, calls apply(ctx), which must enqueue
// out_buf = A(in_buf) on `st` and return 0, and reads `out_buf`; entries flagged in `fixed`
// (optional) are zeroed after each apply.  Scratch is allocated for the call and freed.
extern "C" int phast_pcg_callback(phast_apply_callback apply, void *ctx, int64_t n, double *in_buf,
                                  const double *out_buf, const uint8_t *fixed, const double *b, const double *diag,
                                  const double *x0, double *x, double rtol, double atol, int32_t maxiter,
                                  int32_t check_every, double *info, hipStream_t st) {
    PHAST_REQUIRE(apply && in_buf && out_buf && b && diag && x && info && n >= 1, "phast_pcg_callback: null argument");
    PHAST_REQUIRE(maxiter >= 1 && check_every >= 1, "maxiter and check_every must be >= 1");
    PHAST_REQUIRE(!(rtol == 0.0 && atol == 0.0) && rtol >= 0.0 && atol >= 0.0, "rtol and atol must not both be zero");
    struct Tramp {
        phast_apply_callback fn;
        void *ctx;
        double *in;
        const double *out;
        const uint8_t *fixed;
        int64_t n;
    } t{apply, ctx, in_buf, out_buf, fixed, n};
    auto cb = [](void *c, const double *in, double *out, hipStream_t s) -> int {
        auto *T = static_cast<Tramp *>(c);
        PHAST_HIP(hipMemcpyAsync(T->in, in, sizeof(double) * T->n, hipMemcpyDeviceToDevice, s));
        if (int rc = T->fn(T->ctx)) {
            set_error("phast_pcg_callback: the apply callback failed (%d)", rc);
            return PHAST_ERR_ARG;
        }
        PHAST_HIP(hipMemcpyAsync(out, T->out, sizeof(double) * T->n, hipMemcpyDeviceToDevice, s));
        if (T->fixed) {
            hipLaunchKernelGGL(k_zero_fixed, dim3(nblk(T->n)), dim3(256), 0, s, T->n, T->fixed, out);
            PHAST_LAUNCHED();
        }
        return PHAST_OK;
    };
    Work W;
    int rc = W.ensure(n);
    if (rc == PHAST_OK) rc = pcg_flags(W, cb, &t, n, fixed, b, diag, x0, x, rtol, atol, maxiter, check_every, info, st);
    (void)hipStreamSynchronize(st);
    W.release();
    return rc;
}
