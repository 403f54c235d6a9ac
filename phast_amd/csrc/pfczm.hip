// PF-CZM damage law and crack closure of the periodic voxel cell, gfx950, FP64.
//
// Wu's rational degradation with a per-voxel a1 (physics/pfczm.py:17-130), the PF-CZM damage
// energy, its gradient and the quadrature-point curvatures of the projected Newton tangent
// (solid3d/operator.py:461-583), and the sign-pattern scan of the isotropic crack closure
// (operator.py:755-876).  The Newton tangent itself, Gc (2 ell/pi) L + sum_q w c_q N_q N_q^T,
// has the form of the AT2 damage operator, so it runs on those kernels with the constants
// geo_t (K = (2 ell/pi) L, ch = 1) and the curvatures c_q in place of H; the closure stress
// is fused into the voxel apply (hex8.hip).
//
// Per-voxel kernels gather their 8 corner values straight from HBM: each runs once per Newton
// iteration or pattern solve, not per Krylov iteration, and on the cells of this path (48^3 to
// 128^3) that is far below the apply's cost.  Reductions are block sums in a fixed order, then
// one block over the partials: the results are deterministic.
#include "common.h"

#include <cstring>
#include <vector>

#include "hex8_device.h"
#include "hex8_internal.h"

namespace phast {

// g, g', g'' at d (physics/pfczm.py:79-130), with the reference's operation order.
__device__ __forceinline__ void pfczm_eval(const PfczmLaw &L, double a1, double d, double &g, double &gp,
                                           double &gpp) {
    const double omd = fmax(1.0 - d, 0.0);
    double A, Ap, App;
    if (L.p_int) {
        A = omd * omd;
        Ap = -2.0 * omd;
        App = 2.0;
    } else {
        A = pow(omd, L.p);
        Ap = -L.p * pow(omd, L.p - 1.0);
        App = (L.p * (L.p - 1.0)) * pow(omd, L.p - 2.0);
    }
    const double a2 = L.a2, a23 = L.a2 * L.a3;
    const double Q = (a1 * d) * ((1.0 + a2 * d) + (a23 * d) * d);
    const double Qp = a1 * ((1.0 + (2.0 * a2) * d) + ((3.0 * a23) * d) * d);
    const double Qpp = a1 * (2.0 * a2 + (6.0 * a23) * d);
    const double S = fmax(A + Q, 1.0e-30);
    const double N = Ap * Q - A * Qp;
    const double Np = App * Q - A * Qpp;
    const double S2 = S * S;
    const double scale = 1.0 - L.eta;
    g = scale * (A / S) + L.eta;
    gp = scale * (N / S2);
    gpp = scale * (Np / S2 - (2.0 * N) * (Ap + Qp) / (S2 * S));
}

__device__ __forceinline__ void gather8(const Hex8Geo &G, const double *f, int64_t e, double out[8]) {
    const int i = (int)(e % G.nx), j = (int)((e / G.nx) % G.ny), k = (int)(e / ((int64_t)G.nx * G.ny));
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const int ii = wrap(i + (a & 1), G.nx), jj = wrap(j + ((a >> 1) & 1), G.ny), kk = wrap(k + (a >> 2), G.nz);
        out[a] = f[ii + (int64_t)G.nx * (jj + (int64_t)G.ny * kk)];
    }
}

// g, g', g'' (E,8) at the Gauss points of nodal d; null outputs are skipped.
__global__ __launch_bounds__(256) void pfczm_deriv_kernel(const Hex8Geo G, const PfczmLaw L, const double *a1,
                                                          const double *d, double *g, double *gp, double *gpp) {
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    double dc[8], dq[8];
    gather8(G, d, e, dc);
    voxel_interp(G, dc, dq);
    const double ae = a1[e];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        double v0, v1, v2;
        pfczm_eval(L, ae, dq[q], v0, v1, v2);
        if (g) g[e * 8 + q] = v0;
        if (gp) gp[e * 8 + q] = v1;
        if (gpp) gpp[e * 8 + q] = v2;
    }
}

// Per-voxel energy parts (operator.py:461-488, 611-631): Gc (2 m.d - d.M d)/(pi ell),
// Gc ell/pi d.L d and sum_q w H_q g(d_q); block partials to part[8 b + 0..2].
__global__ __launch_bounds__(256) void pfczm_energy_kernel(const Hex8Geo G, const PfczmLaw L, const double *a1,
                                                           const double *Gc, const double *d, const double *H,
                                                           double *part) {
    __shared__ double red[3 * 4];
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double v[3] = {0.0, 0.0, 0.0};
    if (e < n) {
        double dc[8];
        gather8(G, d, e, dc);
        double md = 0.0, dMd = 0.0, dLd = 0.0;
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            double tm = 0.0, tl = 0.0;
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                tm += dc[b] * L.M[b * 8 + a];
                tl += dc[b] * L.L[b * 8 + a];
            }
            md += dc[a] * L.m[a];
            dMd += tm * dc[a];
            dLd += tl * dc[a];
        }
        const double gc = Gc[e];
        v[0] = gc * (L.c_local * (2.0 * md - dMd));
        v[1] = gc * (L.c_grad * dLd);
        if (H) {
            double dq[8];
            voxel_interp(G, dc, dq);
            const double ae = a1[e];
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                double g, gp, gpp;
                pfczm_eval(L, ae, dq[q], g, gp, gpp);
                s += (H[e * 8 + q] * G.w) * g;
            }
            v[2] = s;
        }
    }
    block_sum_n<256, 3>(v, red);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 3; ++k) part[8 * blockIdx.x + k] = v[k];
}

// Unscaled-times-scale gradient (operator.py:490-512): per node, over its 8 voxels in
// increasing (dz, dy, dx) order, Gc (2/(pi ell)(m - M d_e) + 2 ell/pi L d_e)_a +
// sum_q w H_q g'(d_q) N_qa.
__global__ __launch_bounds__(256) void pfczm_gradient_kernel(const Hex8Geo G, const PfczmLaw L, const double *a1,
                                                             const double *Gc, const double *d, const double *H,
                                                             double scale, double *out) {
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    const int64_t node = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (node >= n) return;
    const int i = (int)(node % G.nx), j = (int)((node / G.nx) % G.ny), k = (int)(node / ((int64_t)G.nx * G.ny));
    const double cl2 = 2.0 * L.c_local, cg2 = 2.0 * L.c_grad;
    double acc = 0.0;
    for (int v = 0; v < 8; ++v) {
        const int dx = v & 1, dy = (v >> 1) & 1, dz = v >> 2;
        const int64_t e = wrap(i - 1 + dx, G.nx) + (int64_t)G.nx * (wrap(j - 1 + dy, G.ny) +
                                                                  (int64_t)G.ny * wrap(k - 1 + dz, G.nz));
        const int a = (1 - dx) + 2 * (1 - dy) + 4 * (1 - dz);
        double dc[8];
        gather8(G, d, e, dc);
        double tm = 0.0, tl = 0.0;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            tm += dc[b] * L.M[b * 8 + a];
            tl += dc[b] * L.L[b * 8 + a];
        }
        double y = Gc[e] * (cl2 * (L.m[a] - tm) + cg2 * tl);
        if (H) {
            double dq[8];
            voxel_interp(G, dc, dq);
            const double ae = a1[e];
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                double g, gp, gpp;
                pfczm_eval(L, ae, dq[q], g, gp, gpp);
                s += ((H[e * 8 + q] * G.w) * gp) * G.Nq[q * 8 + a];
            }
            y += s;
        }
        acc += y;
    }
    out[node] = scale * acc;
}

// Newton tangent curvatures (operator.py:514-538): max(H g''(d_q) - c, 0.25 c) with
// c = 2 Gc/(pi ell); the floor everywhere without H.
__global__ __launch_bounds__(256) void pfczm_weights_kernel(const Hex8Geo G, const PfczmLaw L, const double *a1,
                                                            const double *Gc, const double *d, const double *H,
                                                            double *W) {
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const double concave = L.concave * Gc[e], floor_e = L.floor * concave;
    if (!H) {
#pragma unroll
        for (int q = 0; q < 8; ++q) W[e * 8 + q] = floor_e;
        return;
    }
    double dc[8], dq[8];
    gather8(G, d, e, dc);
    voxel_interp(G, dc, dq);
    const double ae = a1[e];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        double g, gp, gpp;
        pfczm_eval(L, ae, dq[q], g, gp, gpp);
        W[e * 8 + q] = fmax(H[e * 8 + q] * gpp - concave, floor_e);
    }
}

// Closure sign pattern of state x (operator.py:786-841): bit 4 q + k of newp[e] is
// [v_k . eps_q < 0] on closure voxels -- rows trace, e_x, e_y, e_z with moduli (lam, 2 mu,
// 2 mu, 2 mu) on isotropic cells, or the rows and moduli of the voxel's bank entry in
// `tab` (28 doubles each) on cells with a general bank.  With `old` (and g): per closure
// group (clm 1: isotropic, 2: laminate band), block maxima of k_k |v_k . eps| (all bits)
// and of (1 - g) k_k |v_k . eps| over the bits that differ from `old` (-1 when none), to
// part[8 b + 2 (group - 1) + 0..1].
__global__ __launch_bounds__(256) void closure_scan_kernel(const Hex8Geo G, const double *x, const double *g,
                                                           const double *lam, const double *mu, const uint8_t *clm,
                                                           const double *tab, const int32_t *bank_index,
                                                           const uint32_t *old, uint32_t *newp, double *part) {
    __shared__ double red[4][4];
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double smax[2] = {0.0, 0.0}, jmax[2] = {-1.0, -1.0};
    if (e < n) {
        const int i = (int)(e % G.nx), j = (int)((e / G.nx) % G.ny), k = (int)(e / ((int64_t)G.nx * G.ny));
        double u[3][8], E[6];
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const int64_t node = wrap(i + (a & 1), G.nx) + (int64_t)G.nx * (wrap(j + ((a >> 1) & 1), G.ny) +
                                                                           (int64_t)G.ny * wrap(k + (a >> 2), G.nz));
            u[0][a] = x[3 * node];
            u[1][a] = x[3 * node + 1];
            u[2][a] = x[3 * node + 2];
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) E[c] = x[3 * n + c];
        double gx[3][2][2], gy[3][2][2], gz[3][2][2];
        voxel_grad(G, u, gx, gy, gz);
        fold_macro(E, gx, gy, gz);
        const int grp = clm[e];
        const bool on = grp != 0;
        const int gi = grp == 2 ? 1 : 0;
        const double *T = tab ? tab + 28 * (int64_t)bank_index[e] : nullptr;
        double kk[4];
        if (T) {
#pragma unroll
            for (int b = 0; b < 4; ++b) kk[b] = on ? T[24 + b] : 0.0;
        } else {
            const double kl = on ? lam[e] : 0.0, km = on ? 2.0 * mu[e] : 0.0;
            kk[0] = kl;
            kk[1] = kk[2] = kk[3] = km;
        }
        const uint32_t ob = old ? old[e] : 0u;
        uint32_t bits = 0u;
        double sm = 0.0, jm = -1.0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int ix = q & 1, iy = (q >> 1) & 1, iz = q >> 2;
            double s[6];
            qp_strain(gx, gy, gz, ix, iy, iz, s);
            double v[4];
            if (T) {
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    double t = 0.0;
#pragma unroll
                    for (int m = 0; m < 6; ++m) t += T[6 * b + m] * s[m];
                    v[b] = t;
                }
            } else {
                v[0] = (s[0] + s[1]) + s[2];
                v[1] = s[0];
                v[2] = s[1];
                v[3] = s[2];
            }
            const double omg = old ? 1.0 - g[e * 8 + q] : 0.0;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const uint32_t bit = (on && v[b] < 0.0) ? 1u : 0u;
                bits |= bit << (4 * q + b);
                const double strength = kk[b] * fabs(v[b]);
                sm = fmax(sm, strength);
                if (old && (((ob >> (4 * q + b)) & 1u) != bit)) jm = fmax(jm, omg * strength);
            }
        }
        smax[gi] = sm;
        jmax[gi] = jm;
        newp[e] = bits;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x / 64;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const double a = wave_max(smax[c]), b = wave_max(jmax[c]);
        if (lane == 0) {
            red[2 * c][w] = a;
            red[2 * c + 1][w] = b;
        }
    }
    __syncthreads();
    if (threadIdx.x < 4)
        part[8 * blockIdx.x + threadIdx.x] = fmax(fmax(red[threadIdx.x][0], red[threadIdx.x][1]),
                                                  fmax(red[threadIdx.x][2], red[threadIdx.x][3]));
}

// One block: out[k] = sum (op 0) or max (op 1) of part[8 b + k] over b < nb, in a fixed order.
__global__ __launch_bounds__(256) void partials_kernel(const double *part, int nb, int nvals, int op, double *out) {
    __shared__ double red[4];
    for (int k = 0; k < nvals; ++k) {
        double v = op == 0 ? 0.0 : -1.0;
        for (int b = threadIdx.x; b < nb; b += 256) v = op == 0 ? v + part[8 * b + k] : fmax(v, part[8 * b + k]);
        if (op == 0) {
            v = block_sum<256>(v, red);
        } else {
            v = wave_max(v);
            __syncthreads();
            if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = v;
            __syncthreads();
            v = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        }
        if (threadIdx.x == 0) out[k] = v;
        __syncthreads();
    }
}

static int require_pfczm(const phast_hex8_plan *pl, const char *who) {
    if (!pl) {
        set_error("%s: null plan", who);
        return PHAST_ERR_ARG;
    }
    if (!pl->a1) {
        set_error("%s: the plan has no PF-CZM law (phast_hex8_pfczm_setup)", who);
        return PHAST_ERR_ARG;
    }
    return PHAST_OK;
}

}  // namespace phast

using namespace phast;

extern "C" int phast_hex8_pfczm_setup(phast_hex8_plan *pl, const double *a1, double a2, double a3, double p) {
    PHAST_REQUIRE(pl && a1, "phast_hex8_pfczm_setup: null argument");
    PHAST_REQUIRE(p >= 2.0 && std::isfinite(a2) && std::isfinite(a3), "phast_hex8_pfczm_setup: p must be >= 2");
    for (int64_t e = 0; e < pl->n; ++e)
        PHAST_REQUIRE(std::isfinite(a1[e]) && a1[e] > 0.0, "phast_hex8_pfczm_setup: a1[%lld] must be finite and > 0",
                      (long long)e);
    PfczmLaw L;
    std::memset(&L, 0, sizeof(L));
    L.a2 = a2;
    L.a3 = a3;
    L.p = p;
    L.p_int = p == 2.0 ? 1 : 0;
    L.eta = pl->geo.kappa;
    L.c_local = 1.0 / (M_PI * pl->ell);
    L.c_grad = pl->ell / M_PI;
    L.concave = 2.0 / (M_PI * pl->ell);
    L.floor = 0.25;
    Hex8Geo Gt = pl->geo;
    voxel_constants(Gt, pl->h[0], pl->h[1], pl->h[2], pl->ell, L.M, L.L);
    for (int a = 0; a < 8; ++a) {
        double s = 0.0;
        for (int b = 0; b < 8; ++b) s += L.M[a * 8 + b];
        L.m[a] = s;
    }
    for (int k = 0; k < 64; ++k) Gt.K[k] = (2.0 / M_PI) * pl->ell * L.L[k];
    Gt.ch = 1.0;
    double *dev = pl->a1;
    if (!dev) PHAST_HIP(hipMalloc(&dev, sizeof(double) * (size_t)pl->n));
    pl->a1 = dev;
    PHAST_HIP(hipMemcpy(dev, a1, sizeof(double) * (size_t)pl->n, hipMemcpyHostToDevice));
    pl->law = L;
    pl->geo_t = Gt;
    return PHAST_OK;
}

extern "C" int phast_hex8_closure_setup(phast_hex8_plan *pl, const uint8_t *cl_on) {
    PHAST_REQUIRE(pl && cl_on, "phast_hex8_closure_setup: null argument");
    PHAST_REQUIRE(pl->iso, "phast_hex8_closure_setup is for cells whose every stiffness is isotropic; a general "
                  "bank needs phast_hex8_closure_setup_rows");
    for (int64_t e = 0; e < pl->n; ++e)
        PHAST_REQUIRE(cl_on[e] <= 1, "phast_hex8_closure_setup: cl_on[%lld] must be 0 or 1", (long long)e);
    uint8_t *dev = pl->cl_on;
    if (!dev) PHAST_HIP(hipMalloc(&dev, (size_t)pl->n));
    pl->cl_on = dev;
    PHAST_HIP(hipMemcpy(dev, cl_on, (size_t)pl->n, hipMemcpyHostToDevice));
    pl->cl_band = false;
    return PHAST_OK;
}

extern "C" int phast_hex8_closure_setup_rows(phast_hex8_plan *pl, const uint8_t *cl_group, const double *rows) {
    PHAST_REQUIRE(pl && cl_group && rows, "phast_hex8_closure_setup_rows: null argument");
    bool band = false;
    for (int64_t e = 0; e < pl->n; ++e) {
        PHAST_REQUIRE(cl_group[e] <= 2, "phast_hex8_closure_setup_rows: cl_group[%lld] must be 0, 1 or 2",
                      (long long)e);
        band = band || cl_group[e] == 2;
    }
    for (int64_t k = 0; k < 28 * (int64_t)pl->n_bank; ++k)
        PHAST_REQUIRE(std::isfinite(rows[k]), "phast_hex8_closure_setup_rows: rows[%lld] is not finite", (long long)k);
    for (int32_t b = 0; b < pl->n_bank; ++b)
        for (int k = 0; k < 4; ++k)
            PHAST_REQUIRE(rows[28 * (int64_t)b + 24 + k] >= 0.0,
                          "phast_hex8_closure_setup_rows: modulus %d of bank entry %d is negative", k, b);
    uint8_t *dev = pl->cl_on;
    if (!dev) PHAST_HIP(hipMalloc(&dev, (size_t)pl->n));
    pl->cl_on = dev;
    PHAST_HIP(hipMemcpy(dev, cl_group, (size_t)pl->n, hipMemcpyHostToDevice));
    double *tab = pl->cl_tab;
    if (!tab) PHAST_HIP(hipMalloc(&tab, sizeof(double) * 28 * (size_t)pl->n_bank));
    pl->cl_tab = tab;
    PHAST_HIP(hipMemcpy(tab, rows, sizeof(double) * 28 * (size_t)pl->n_bank, hipMemcpyHostToDevice));
    pl->cl_band = band;
    return PHAST_OK;
}

extern "C" int phast_hex8_pfczm_degradation(const phast_hex8_plan *pl, const double *d, double *g, double *gp,
                                            double *gpp, hipStream_t st) {
    PHAST_TRY0(require_pfczm(pl, "phast_hex8_pfczm_degradation"));
    PHAST_REQUIRE(d && (g || gp || gpp), "phast_hex8_pfczm_degradation: null argument");
    hipLaunchKernelGGL(pfczm_deriv_kernel, dim3((unsigned)ceil_div(pl->n, 256)), dim3(256), 0, st, pl->geo, pl->law,
                       pl->a1, d, g, gp, gpp);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

extern "C" int phast_hex8_pfczm_energy(const phast_hex8_plan *pl, const double *d, const double *H, double *out3,
                                       hipStream_t st) {
    PHAST_TRY0(require_pfczm(pl, "phast_hex8_pfczm_energy"));
    PHAST_REQUIRE(d && out3, "phast_hex8_pfczm_energy: null argument");
    const int nb = (int)ceil_div(pl->n, 256);
    hipLaunchKernelGGL(pfczm_energy_kernel, dim3(nb), dim3(256), 0, st, pl->geo, pl->law, pl->a1, pl->Gc, d, H,
                       pl->part);
    PHAST_LAUNCHED();
    hipLaunchKernelGGL(partials_kernel, dim3(1), dim3(256), 0, st, pl->part, nb, 3, 0, out3);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

extern "C" int phast_hex8_pfczm_gradient(const phast_hex8_plan *pl, const double *d, const double *H, double scale,
                                         double *out, hipStream_t st) {
    PHAST_TRY0(require_pfczm(pl, "phast_hex8_pfczm_gradient"));
    PHAST_REQUIRE(d && out && d != out, "phast_hex8_pfczm_gradient: bad argument");
    hipLaunchKernelGGL(pfczm_gradient_kernel, dim3((unsigned)ceil_div(pl->n, 256)), dim3(256), 0, st, pl->geo,
                       pl->law, pl->a1, pl->Gc, d, H, scale, out);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

extern "C" int phast_hex8_pfczm_tangent_weights(const phast_hex8_plan *pl, const double *d, const double *H,
                                                double *W, hipStream_t st) {
    PHAST_TRY0(require_pfczm(pl, "phast_hex8_pfczm_tangent_weights"));
    PHAST_REQUIRE(d && W, "phast_hex8_pfczm_tangent_weights: null argument");
    hipLaunchKernelGGL(pfczm_weights_kernel, dim3((unsigned)ceil_div(pl->n, 256)), dim3(256), 0, st, pl->geo,
                       pl->law, pl->a1, pl->Gc, d, H, W);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

extern "C" int phast_hex8_pfczm_tangent_action(const phast_hex8_plan *pl, const double *f, const double *W,
                                               double scale, double *y, hipStream_t st) {
    PHAST_TRY0(require_pfczm(pl, "phast_hex8_pfczm_tangent_action"));
    PHAST_REQUIRE(f && W && y && f != y, "phast_hex8_pfczm_tangent_action: bad argument");
    return hex8_damage_launch(pl, pl->geo_t, f, W, scale, y, st);
}

extern "C" int phast_hex8_pfczm_tangent_diagonal(const phast_hex8_plan *pl, const double *W, double scale,
                                                 double *diag, hipStream_t st) {
    PHAST_TRY0(require_pfczm(pl, "phast_hex8_pfczm_tangent_diagonal"));
    PHAST_REQUIRE(W && diag, "phast_hex8_pfczm_tangent_diagonal: null argument");
    return hex8_damage_rd_launch(pl, pl->geo_t, W, scale, nullptr, diag, st);
}

extern "C" int phast_hex8_closure_pattern(const phast_hex8_plan *pl, const double *x, const double *g,
                                          const uint32_t *old, uint32_t *newp, double *out2, hipStream_t st) {
    PHAST_REQUIRE(pl && x && newp, "phast_hex8_closure_pattern: null argument");
    PHAST_REQUIRE(pl->cl_on, "phast_hex8_closure_pattern: the plan has no crack closure (phast_hex8_closure_setup)");
    PHAST_REQUIRE(!old || (g && out2), "phast_hex8_closure_pattern: a pattern change needs g and out2");
    const int nb = (int)ceil_div(pl->n, 256);
    const double *tab = pl->iso ? nullptr : pl->cl_tab;
    PHAST_REQUIRE(pl->iso || tab, "phast_hex8_closure_pattern: a general bank needs phast_hex8_closure_setup_rows");
    hipLaunchKernelGGL(closure_scan_kernel, dim3(nb), dim3(256), 0, st, pl->geo, x, g, pl->lam, pl->mu, pl->cl_on,
                       tab, pl->bank_index, old, newp, pl->part);
    PHAST_LAUNCHED();
    if (out2) {
        hipLaunchKernelGGL(partials_kernel, dim3(1), dim3(256), 0, st, pl->part, nb, pl->cl_band ? 4 : 2, 1, out2);
        PHAST_LAUNCHED();
    }
    return PHAST_OK;
}

extern "C" int phast_hex8_mech_action_cl(const phast_hex8_plan *pl, const double *x, const double *g,
                                         const uint32_t *pat, int32_t mode, double *y, hipStream_t st) {
    PHAST_REQUIRE(pl && x && g && y && x != y && (mode == 1 || mode == 2), "phast_hex8_mech_action_cl: bad argument");
    return hex8_apply_closure(pl, x, g, pat, mode, y, nullptr, false, st);
}

extern "C" int phast_hex8_stress_sums_cl(const phast_hex8_plan *pl, const double *x, const double *g,
                                         const uint32_t *pat, int32_t mode, double *out7, hipStream_t st) {
    PHAST_REQUIRE(pl && x && g && out7 && (mode == 1 || mode == 2), "phast_hex8_stress_sums_cl: bad argument");
    double *scratch = nullptr;
    PHAST_HIP(hipMallocAsync((void **)&scratch, sizeof(double) * (3 * pl->n + 6), st));
    const int rc = hex8_apply_closure(pl, x, g, pat, mode, scratch, out7, true, st);
    PHAST_HIP(hipFreeAsync(scratch, st));
    return rc;
}
