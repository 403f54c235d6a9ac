// Adjoint pieces of the converged staggered voxel state (FP64, gfx950).
//
// The reference has no adjoint for the voxel path.  These kernels supply the transposed
// derivatives of its residuals that a reverse sweep over load steps needs:
//   coupling   (dF/dd)^T l     for F = A(g(d)) x  (solid3d/operator.py:939-946, 1108-1194)
//   driving    (dY/dx)^T t     for the driving energy Y (operator.py:1415-1444, splits.py:68-161)
//   history    (dG/dH)^T l     for G = A_H d - b_H   (operator.py:1450-1480)
//   Gc         (dG/dGc)^T l    per voxel             (operator.py:1450-1466)
// plus the two assembly shapes they feed: Gauss-point scalars to nodes (sum_q N_a f_q) and
// Gauss-point stress-like tensors to the mechanics space (sum_q B_q^T S_q and the macro
// rows sum_q S_q).  Node sums gather the 8 voxels of a node in a fixed order (no atomics).
#include "common.h"
#include "hex8_device.h"
#include "hex8_internal.h"

namespace phast {

__device__ __forceinline__ void voxel_ijk(const Hex8Geo &G, int64_t e, int &i, int &j, int &k) {
    i = (int)(e % G.nx);
    j = (int)((e / G.nx) % G.ny);
    k = (int)(e / ((int64_t)G.nx * G.ny));
}

__device__ __forceinline__ int64_t corner_node(const Hex8Geo &G, int i, int j, int k, int a) {
    return wrap(i + (a & 1), G.nx) + (int64_t)G.nx * (wrap(j + ((a >> 1) & 1), G.ny) +
                                                      (int64_t)G.ny * wrap(k + (a >> 2), G.nz));
}

// engineering strains at the 8 Gauss points of voxel (i,j,k) of the mechanics vector x
__device__ __forceinline__ void voxel_strains(const Hex8Geo &G, const double *x, int64_t n, int i, int j, int k,
                                              double eps[8][6]) {
    double u[3][8], E[6];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const int64_t node = corner_node(G, i, j, k, a);
        u[0][a] = x[3 * node];
        u[1][a] = x[3 * node + 1];
        u[2][a] = x[3 * node + 2];
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) E[c] = x[3 * n + c];
    double gx[3][2][2], gy[3][2][2], gz[3][2][2];
    voxel_grad(G, u, gx, gy, gz);
    fold_macro(E, gx, gy, gz);
#pragma unroll
    for (int q = 0; q < 8; ++q) qp_strain(gx, gy, gz, q & 1, (q >> 1) & 1, q >> 2, eps[q]);
}

// (A - a I)(A - b I) for the symmetric 3x3 A = [[m0, m5, m4], [m5, m1, m3], [m4, m3, m2]]
// (the two factors commute, so the product is symmetric); result in the same order.
__device__ __forceinline__ void sym_prod_shifted(const double m[6], double a, double b, double p[6]) {
    const double x0 = m[0] - a, x1 = m[1] - a, x2 = m[2] - a;
    const double y0 = m[0] - b, y1 = m[1] - b, y2 = m[2] - b;
    p[0] = x0 * y0 + m[5] * m[5] + m[4] * m[4];
    p[1] = m[5] * m[5] + x1 * y1 + m[3] * m[3];
    p[2] = m[4] * m[4] + m[3] * m[3] + x2 * y2;
    p[3] = m[5] * m[4] + x1 * m[3] + m[3] * y2;  // (1,2)
    p[4] = x0 * m[4] + m[5] * m[3] + m[4] * y2;  // (0,2)
    p[5] = x0 * m[5] + m[5] * y1 + m[4] * m[3];  // (0,1)
}

// dY/d(engineering strain) of the spectral tensile energy
// Y = lam/2 <tr eps>+^2 + mu sum_a <eps_a>+^2: the tensor lam <tr>+ I + 2 mu eps+, with
// eps+ = sum over positive eigenvalues of eps_a P_a formed from spectral projectors of the
// sign groups (so a repeated eigenvalue inside one group needs no eigenvectors).
__device__ __forceinline__ void spectral_grad(const double s[6], double lam, double mu, double T[6]) {
    const double m[6] = {s[0], s[1], s[2], 0.5 * s[3], 0.5 * s[4], 0.5 * s[5]};
    double ev[3];
    eig3(m[0], m[1], m[2], m[5], m[4], m[3], ev);  // ascending
    double ep[6] = {0, 0, 0, 0, 0, 0};
    const int npos = (ev[0] > 0.0) + (ev[1] > 0.0) + (ev[2] > 0.0);
    if (npos == 3) {
#pragma unroll
        for (int k = 0; k < 6; ++k) ep[k] = m[k];
    } else if (npos == 1) {  // ev2 P2
        double p[6];
        sym_prod_shifted(m, ev[0], ev[1], p);
        const double den = (ev[2] - ev[0]) * (ev[2] - ev[1]);
        const double f = den > 0.0 ? ev[2] / den : 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) ep[k] = f * p[k];
    } else if (npos == 2) {  // A - ev0 P0
        double p[6];
        sym_prod_shifted(m, ev[1], ev[2], p);
        const double den = (ev[0] - ev[1]) * (ev[0] - ev[2]);
        const double f = den > 0.0 ? ev[0] / den : 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) ep[k] = m[k] - f * p[k];
    }
    const double trp = fmax((s[0] + s[1]) + s[2], 0.0);
    T[0] = lam * trp + 2.0 * mu * ep[0];
    T[1] = lam * trp + 2.0 * mu * ep[1];
    T[2] = lam * trp + 2.0 * mu * ep[2];
    T[3] = 2.0 * mu * ep[3];
    T[4] = 2.0 * mu * ep[4];
    T[5] = 2.0 * mu * ep[5];
}

// out[e,q] = w g'(d_q) eps(l)_q . C_e eps(x)_q, g'(d) = -2 (1 - kappa)(1 - d)
__global__ __launch_bounds__(256) void adj_coupling_kernel(const Hex8Geo G, const double *x, const double *l,
                                                           const double *d, const double *bank,
                                                           const int32_t *bank_index, double *out) {
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    int i, j, k;
    voxel_ijk(G, e, i, j, k);
    double dc[8], dq[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) dc[a] = d[corner_node(G, i, j, k, a)];
    voxel_interp(G, dc, dq);
    const double *C = bank + 36 * (int64_t)bank_index[e];
    double ex[8][6];
    voxel_strains(G, x, n, i, j, k, ex);
    double el[8][6];
    voxel_strains(G, l, n, i, j, k, el);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        double t = 0.0;
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            double c6 = 0.0;
#pragma unroll
            for (int c = 0; c < 6; ++c) c6 += C[r * 6 + c] * ex[q][c];
            t += el[q][r] * c6;
        }
        out[e * 8 + q] = G.w * (-2.0 * G.omk * (1.0 - dq[q])) * t;
    }
}

// S[(e*8+q)*6 + k] = t[e,q] dY/deps_k at (e,q), by driver code (0 none, 1 full, 2 spectral)
__global__ __launch_bounds__(256) void adj_driving_kernel(const Hex8Geo G, const double *x, const double *t,
                                                          const double *bank, const int32_t *bank_index,
                                                          const double *lam, const double *mu, const int8_t *driver,
                                                          double *S) {
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const int code = driver[e];
    if (code != 1 && code != 2) {
        for (int q = 0; q < 48; ++q) S[e * 48 + q] = 0.0;
        return;
    }
    int i, j, k;
    voxel_ijk(G, e, i, j, k);
    double ex[8][6];
    voxel_strains(G, x, n, i, j, k, ex);
    const double *C = bank + 36 * (int64_t)bank_index[e];
    const double la = lam[e], m = mu[e];
    for (int q = 0; q < 8; ++q) {
        double T[6];
        if (code == 1) {
            for (int r = 0; r < 6; ++r) {
                double c6 = 0.0;
                for (int c = 0; c < 6; ++c) c6 += C[r * 6 + c] * ex[q][c];
                T[r] = c6;
            }
        } else {
            spectral_grad(ex[q], la, m, T);
        }
        const double tq = t[e * 8 + q];
        for (int r = 0; r < 6; ++r) S[(e * 8 + q) * 6 + r] = tq * T[r];
    }
}

// Element fields of a field snapshot (solid3d/operator.py:1561-1608, no interface drivers or
// closure): Gauss-point means of the total strain, of the degraded stress g C eps and of the
// degraded energy 0.5 g eps . C eps.  g null = 1.
__global__ __launch_bounds__(256) void snapshot_kernel(const Hex8Geo G, const double *x, const double *g,
                                                       const double *bank, const int32_t *bank_index,
                                                       double *strain, double *stress, double *psi) {
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    int i, j, k;
    voxel_ijk(G, e, i, j, k);
    double ex[8][6];
    voxel_strains(G, x, n, i, j, k, ex);
    const double *C = bank + 36 * (int64_t)bank_index[e];
    double se[6] = {0, 0, 0, 0, 0, 0}, ss[6] = {0, 0, 0, 0, 0, 0}, sp = 0.0;
    for (int q = 0; q < 8; ++q) {
        const double gq = g ? g[e * 8 + q] : 1.0;
        double en = 0.0;
        for (int r = 0; r < 6; ++r) {
            double c6 = 0.0;
            for (int c = 0; c < 6; ++c) c6 += C[r * 6 + c] * ex[q][c];
            se[r] += ex[q][r];
            ss[r] += gq * c6;
            en += ex[q][r] * c6;
        }
        sp += 0.5 * gq * en;
    }
    for (int r = 0; r < 6; ++r) {
        strain[e * 6 + r] = se[r] / 8.0;
        stress[e * 6 + r] = ss[r] / 8.0;
    }
    psi[e] = sp / 8.0;
}

// out[node] = scale * sum over the 8 voxels of the node of sum_q N_a(q) f[e,q]
__global__ __launch_bounds__(256) void qp_to_node_kernel(const Hex8Geo G, const double *f, double scale,
                                                         double *out) {
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    const int64_t node = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (node >= n) return;
    int i, j, k;
    voxel_ijk(G, node, i, j, k);
    double s = 0.0;
    for (int v = 0; v < 8; ++v) {  // voxels in increasing (dz,dy,dx) order
        const int dx = v & 1, dy = (v >> 1) & 1, dz = v >> 2;
        const int64_t e = wrap(i - 1 + dx, G.nx) +
                          (int64_t)G.nx * (wrap(j - 1 + dy, G.ny) + (int64_t)G.ny * wrap(k - 1 + dz, G.nz));
        const int a = (1 - dx) + 2 * (1 - dy) + 4 * (1 - dz);
        double t = 0.0;
#pragma unroll
        for (int q = 0; q < 8; ++q) t += G.Nq[q * 8 + a] * f[e * 8 + q];
        s += t;
    }
    out[node] = scale * s;
}

// r = sum_e sum_q B_q^T S_q (3n rows, gathered per node) and partial sums of sum_q S_q
__global__ __launch_bounds__(256) void qp_stress_apply_kernel(const Hex8Geo G, const double *S, double *out,
                                                              double *part) {
    __shared__ double red[24];
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    const int64_t node = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double sE[6] = {0, 0, 0, 0, 0, 0};
    if (node < n) {
        int i, j, k;
        voxel_ijk(G, node, i, j, k);
        double acc[3] = {0.0, 0.0, 0.0};
        for (int v = 0; v < 8; ++v) {
            const int dx = v & 1, dy = (v >> 1) & 1, dz = v >> 2;
            const int64_t e = wrap(i - 1 + dx, G.nx) +
                              (int64_t)G.nx * (wrap(j - 1 + dy, G.ny) + (int64_t)G.ny * wrap(k - 1 + dz, G.nz));
            const int lx = 1 - dx, ly = 1 - dy, lz = 1 - dz;
            const double sx = lx ? 1.0 : -1.0, sy = ly ? 1.0 : -1.0, sz = lz ? 1.0 : -1.0;
            for (int q = 0; q < 8; ++q) {
                const int ix = q & 1, iy = (q >> 1) & 1, iz = q >> 2;
                const double Wx = (lx == ix) ? G.al : G.be, Wy = (ly == iy) ? G.al : G.be,
                             Wz = (lz == iz) ? G.al : G.be;
                const double bx = sx * G.ihx * Wy * Wz, by = sy * G.ihy * Wx * Wz, bz = sz * G.ihz * Wx * Wy;
                const double *s = S + (e * 8 + q) * 6;
                // B rows of the corner's dofs: x (0 bx, 4 bz, 5 by); y (1 by, 3 bz, 5 bx); z (2 bz, 3 by, 4 bx)
                acc[0] += (bx * s[0] + bz * s[4]) + by * s[5];
                acc[1] += (by * s[1] + bz * s[3]) + bx * s[5];
                acc[2] += (bz * s[2] + by * s[3]) + bx * s[4];
            }
        }
        out[3 * node] = acc[0];
        out[3 * node + 1] = acc[1];
        out[3 * node + 2] = acc[2];
        // the node's own voxel contributes its Gauss-point sums to the macro rows
        for (int q = 0; q < 8; ++q)
            for (int c = 0; c < 6; ++c) sE[c] += S[(node * 8 + q) * 6 + c];
    }
    block_sum_n<256, 6>(sE, red);
    if (threadIdx.x == 0)
        for (int c = 0; c < 6; ++c) part[8 * blockIdx.x + c] = sE[c];
}

__global__ __launch_bounds__(256) void sum_parts6_kernel(const double *part, int nblocks, double *out6) {
    __shared__ double red[24];
    double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    sum_block_partials<256, 6>(part, nblocks, 6, s);
    block_sum_n<256, 6>(s, red);
    if (threadIdx.x == 0)
        for (int c = 0; c < 6; ++c) out6[c] = s[c];
}

// out[e,q] = scale ch w (N_q.d_e - 1)(N_q.l_e)
__global__ __launch_bounds__(256) void adj_history_kernel(const Hex8Geo G, const double *d, const double *l,
                                                          double scale, double *out) {
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    int i, j, k;
    voxel_ijk(G, e, i, j, k);
    double dc[8], lc[8], dq[8], lq[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const int64_t node = corner_node(G, i, j, k, a);
        dc[a] = d[node];
        lc[a] = l[node];
    }
    voxel_interp(G, dc, dq);
    voxel_interp(G, lc, lq);
#pragma unroll
    for (int q = 0; q < 8; ++q) out[e * 8 + q] = scale * G.ch * G.w * ((dq[q] - 1.0) * lq[q]);
}

// out[e] = scale l_e . (M/ell + ell L) d_e
__global__ __launch_bounds__(256) void adj_gc_kernel(const Hex8Geo G, const double *d, const double *l,
                                                     double scale, double *out) {
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    int i, j, k;
    voxel_ijk(G, e, i, j, k);
    double dc[8], lc[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const int64_t node = corner_node(G, i, j, k, a);
        dc[a] = d[node];
        lc[a] = l[node];
    }
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        double t = 0.0;
#pragma unroll
        for (int b = 0; b < 8; ++b) t += G.K[a * 8 + b] * dc[b];
        s += lc[a] * t;
    }
    out[e] = scale * s;
}

}  // namespace phast

using namespace phast;

static unsigned voxel_blocks(const phast_hex8_plan *pl) { return (unsigned)ceil_div(pl->n, 256); }

extern "C" int phast_hex8_set_gc(phast_hex8_plan *pl, const double *Gc, hipStream_t st) {
    PHAST_REQUIRE(pl && Gc, "phast_hex8_set_gc: null argument");
    PHAST_HIP(hipMemcpyAsync(pl->Gc, Gc, sizeof(double) * pl->n, hipMemcpyDeviceToDevice, st));
    return PHAST_OK;
}

extern "C" int phast_hex8_adj_coupling(const phast_hex8_plan *pl, const double *x, const double *l, const double *d,
                                       double *out_qp, hipStream_t st) {
    PHAST_REQUIRE(pl && x && l && d && out_qp, "phast_hex8_adj_coupling: null argument");
    hipLaunchKernelGGL(adj_coupling_kernel, dim3(voxel_blocks(pl)), dim3(256), 0, st, pl->geo, x, l, d, pl->bank,
                       pl->bank_index, out_qp);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

extern "C" int phast_hex8_adj_driving(const phast_hex8_plan *pl, const double *x, const double *t_qp,
                                      double *S_qp6, hipStream_t st) {
    PHAST_REQUIRE(pl && x && t_qp && S_qp6, "phast_hex8_adj_driving: null argument");
    hipLaunchKernelGGL(adj_driving_kernel, dim3(voxel_blocks(pl)), dim3(256), 0, st, pl->geo, x, t_qp, pl->bank,
                       pl->bank_index, pl->lam, pl->mu, pl->driver, S_qp6);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

extern "C" int phast_hex8_qp_to_node(const phast_hex8_plan *pl, const double *f_qp, double scale, double *out_n,
                                     hipStream_t st) {
    PHAST_REQUIRE(pl && f_qp && out_n, "phast_hex8_qp_to_node: null argument");
    hipLaunchKernelGGL(qp_to_node_kernel, dim3(voxel_blocks(pl)), dim3(256), 0, st, pl->geo, f_qp, scale, out_n);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

extern "C" int phast_hex8_qp_stress_apply(const phast_hex8_plan *pl, const double *S_qp6, double *out_mech,
                                          hipStream_t st) {
    PHAST_REQUIRE(pl && S_qp6 && out_mech, "phast_hex8_qp_stress_apply: null argument");
    const unsigned nb = voxel_blocks(pl);  // the plan's partial buffer holds ceil(n/256) blocks
    hipLaunchKernelGGL(qp_stress_apply_kernel, dim3(nb), dim3(256), 0, st, pl->geo, S_qp6, out_mech, pl->part);
    PHAST_LAUNCHED();
    hipLaunchKernelGGL(sum_parts6_kernel, dim3(1), dim3(256), 0, st, pl->part, (int)nb, out_mech + 3 * pl->n);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

extern "C" int phast_hex8_adj_damage_history(const phast_hex8_plan *pl, const double *d, const double *l,
                                             double scale, double *out_qp, hipStream_t st) {
    PHAST_REQUIRE(pl && d && l && out_qp, "phast_hex8_adj_damage_history: null argument");
    hipLaunchKernelGGL(adj_history_kernel, dim3(voxel_blocks(pl)), dim3(256), 0, st, pl->geo, d, l, scale, out_qp);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

extern "C" int phast_hex8_adj_damage_gc(const phast_hex8_plan *pl, const double *d, const double *l, double scale,
                                        double *out_e, hipStream_t st) {
    PHAST_REQUIRE(pl && d && l && out_e, "phast_hex8_adj_damage_gc: null argument");
    hipLaunchKernelGGL(adj_gc_kernel, dim3(voxel_blocks(pl)), dim3(256), 0, st, pl->geo, d, l, scale, out_e);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

extern "C" int phast_hex8_field_snapshot(const phast_hex8_plan *pl, const double *x, const double *g_qp,
                                         double *strain, double *stress, double *psi, hipStream_t st) {
    PHAST_REQUIRE(pl && x && strain && stress && psi, "phast_hex8_field_snapshot: null argument");
    hipLaunchKernelGGL(snapshot_kernel, dim3(voxel_blocks(pl)), dim3(256), 0, st, pl->geo, x, g_qp, pl->bank,
                       pl->bank_index, strain, stress, psi);
    PHAST_LAUNCHED();
    return PHAST_OK;
}
