// Triply periodic hex8 voxel kernels for gfx950 (FP64 throughout).
//
// HBM layout: reduced node n = i + nx (j + ny k) (solid3d/periodic.py:61-108); the mechanics
// vector is x = [u~ (n,3) interleaved, E (6)]; per-voxel quadrature fields (g, H, Y) are
// (E,8) rows with Gauss points ordered xi fastest (core/hex8.py:157-179); material data
// is per voxel (bank index int32, Gc, lam, mu, driver int8).
//
// The element applies walk z slabs: a block owns a column of nodes and a slab of kz node
// layers.  At each voxel layer it stages the two bounding node layers (with a one-node x/y
// halo) in LDS, evaluates every voxel of the layer once with sum-factorised gradients (the
// trilinear field's derivative along x is a difference in x interpolated in y and z, and
// the transpose for the force), stages the corner results in LDS, and gathers each owned
// node's sum from the 4 voxels below and the 4 above it.  Two organisations:
//   * hex8_mech_kernel ("marching"): TX x TY threads = owned node columns, a strided loop
//     over the (TX+1) x (TY+1) voxels, one extra voxel layer below the slab (the carry);
//   * hex8_apply_kernel (voxel-parallel, the default): one thread per voxel, whole-wave
//     blocks, node layers and material prefetched a layer ahead, and no carry layer -- the
//     top faces of a slab's last layer go to a seam buffer that the follow-on kernel adds.
// There are no atomics: the result is deterministic, and the 6 macro-strain rows are
// reduced per block then summed by one block of a follow-on kernel in a fixed order.
#include "common.h"

#include <cstring>
#include <vector>

#include "hex8_device.h"
#include "hex8_internal.h"

namespace phast {

struct MechArgs3 {
    const double *x;       // (3n+6)
    const double *g;       // (E,8) or null
    const double *d;       // (n,) when g is formed in-kernel
    const double *bank;
    const int32_t *bank_index;
    const double *lam, *mu;
    double *y;             // (3n+6)
    double *part;          // 8 per block
    double *y6;            // macro rows destination (y + 3n) or null
    double *out7;          // stress sums destination or null
    double *seam;          // voxel-parallel apply: top-face sums of each slab's last voxel layer
    const uint8_t *clm;    // crack closure: 1 where the voxel adds (1 - g) sigma_c back
    const uint32_t *pat;   // crack closure: fixed sign pattern, bit 4 q + k per voxel
    const double *cltab;   // crack closure, general banks: rows and moduli per bank entry (28)
    const uint8_t *phase;  // isotropic cells of few phases: phase of each voxel, or null
    const double *ptab;    // (lam, mu) of each phase
    int32_t kz, tiles_x, tiles_y, nvals, n_phase;
};

// End of a mechanics apply: the block's macro-row (and energy) partials, all reduced
// together, go to part[8 b + k]; hex8_finalize_kernel, launched right after on the same
// stream, sums them in block order (a fixed order, so the macro rows are deterministic).
// The kernel boundary orders the partials before that sum: no fence, counter or atomic.
// `red` holds 7 * NT/64 doubles.
template <int NT, bool ENERGY>
__device__ __forceinline__ void macro_epilogue(const MechArgs3 &A, const double sE[6], double en, double *red,
                                               int blk) {
    constexpr int K = ENERGY ? 7 : 6;
    double v[K];
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = sE[k];
    if (ENERGY) v[K - 1] = en;
    block_sum_n<NT, K>(v, red);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) A.part[8 * blk + k] = v[k];
}

template <int TX, int TY, int GMODE, bool ISO, bool ENERGY>  // GMODE 0: g=1, 1: g array, 2: g from d
__global__ __launch_bounds__(TX *TY) void hex8_mech_kernel(const Hex8Geo G, MechArgs3 A) {
    constexpr int NT = TX * TY;
    constexpr int NXH = TX + 2, NYH = TY + 2, NH = NXH * NYH;
    constexpr int NXV = TX + 1, NYV = TY + 1, NV = NXV * NYV;
    constexpr int NF = (GMODE == 2) ? 4 : 3;
    __shared__ double su[2][NF][NH];
    __shared__ double sr[24][NV];
    __shared__ double red[7 * (NT / 64)];

    const int nx = G.nx, ny = G.ny, nz = G.nz;
    const int bx = blockIdx.x % A.tiles_x, by = (blockIdx.x / A.tiles_x) % A.tiles_y;
    const int bz = blockIdx.x / (A.tiles_x * A.tiles_y);
    const int i0 = bx * TX, j0 = by * TY, k0 = bz * A.kz;
    const int k1 = min(k0 + A.kz, nz);
    const int64_t n = (int64_t)nx * ny * nz;
    double E[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) E[k] = A.x[3 * n + k];

    auto load_layer = [&](int buf, int k) {
        const int kk = wrap(k, nz);
        for (int s = threadIdx.x; s < NH; s += NT) {
            const int i = wrapm(i0 + s % NXH - 1, nx), j = wrapm(j0 + s / NXH - 1, ny);
            const int64_t node = i + (int64_t)nx * (j + (int64_t)ny * kk);
            su[buf][0][s] = A.x[3 * node];
            su[buf][1][s] = A.x[3 * node + 1];
            su[buf][2][s] = A.x[3 * node + 2];
            if (GMODE == 2) su[buf][3][s] = A.d[node];
        }
    };

    double sE[6] = {0, 0, 0, 0, 0, 0}, en = 0.0;
    double acc[3] = {0.0, 0.0, 0.0};
    int cur = 0;
    load_layer(0, k0 - 1);
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    for (int kv = k0 - 1; kv < k1; ++kv) {
        load_layer(cur ^ 1, kv + 1);
        __syncthreads();
        const int kw = wrap(kv, nz);
        for (int s = threadIdx.x; s < NV; s += NT) {
            const int li = s % NXV - 1, lj = s / NXV - 1;
            const int iv = i0 + li, jv = j0 + lj;
            const int64_t e = wrapm(iv, nx) + (int64_t)nx * (wrapm(jv, ny) + (int64_t)ny * kw);
            const int b0 = (lj + 1) * NXH + (li + 1);
            double u[3][8], g[8];
#pragma unroll
            for (int lz = 0; lz < 2; ++lz) {
                const int bb = lz ? (cur ^ 1) : cur;
#pragma unroll
                for (int ly = 0; ly < 2; ++ly)
#pragma unroll
                    for (int lx = 0; lx < 2; ++lx) {
                        const int slot = b0 + ly * NXH + lx, a = lx + 2 * ly + 4 * lz;
                        u[0][a] = su[bb][0][slot];
                        u[1][a] = su[bb][1][slot];
                        u[2][a] = su[bb][2][slot];
                        if (GMODE == 2) g[a] = su[bb][NF - 1][slot];
                    }
            }
            if (GMODE == 2) {
                double dq[8];
                voxel_interp(G, g, dq);
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const double omd = 1.0 - dq[q];
                    g[q] = G.omk * (omd * omd) + G.kappa;  // splits.py:56-59
                }
            } else if (GMODE == 1) {
                load_qp8(A.g, e, g);
            } else {
#pragma unroll
                for (int q = 0; q < 8; ++q) g[q] = 1.0;
            }
            const bool own = li >= 0 && lj >= 0 && kv >= k0 && iv < nx && jv < ny;
            double r[3][8];
            if (ISO) {
                voxel_mech<true, ENERGY>(G, u, E, g, nullptr, A.lam[e], A.mu[e], own, r, sE, en);
            } else {
                voxel_mech<false, ENERGY>(G, u, E, g, A.bank + 36 * (int64_t)A.bank_index[e], 0.0, 0.0, own, r,
                                          sE, en);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int a = 0; a < 8; ++a) sr[c * 8 + a][s] = r[c][a];
        }
        __syncthreads();
        {
            // voxels around node (tx,ty): (tx-1,ty-1) corner (1,1), (tx,ty-1) corner (0,1),
            // (tx-1,ty) corner (1,0), (tx,ty) corner (0,0); bottom faces -> node level kv,
            // top faces -> node level kv+1.
            const int s00 = ty * NXV + tx, s10 = s00 + 1, s01 = s00 + NXV, s11 = s01 + 1;
            const int i = i0 + tx, j = j0 + ty;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double bot = ((sr[c * 8 + 3][s00] + sr[c * 8 + 2][s10]) + sr[c * 8 + 1][s01]) +
                                   sr[c * 8 + 0][s11];
                const double top = ((sr[c * 8 + 7][s00] + sr[c * 8 + 6][s10]) + sr[c * 8 + 5][s01]) +
                                   sr[c * 8 + 4][s11];
                if (kv >= k0 && i < nx && j < ny) {
                    const int64_t node = i + (int64_t)nx * (j + (int64_t)ny * kv);
                    A.y[3 * node + c] = acc[c] + bot;
                }
                acc[c] = top;
            }
        }
        cur ^= 1;
    }
    macro_epilogue<NT, ENERGY>(A, sE, en, red, blockIdx.x);
}


// ---------------------------------------------------------------------------------------
// Voxel-parallel apply: one thread per voxel of a (OX+1) x (OY+1) voxel tile (whole waves,
// no second pass for the halo voxels), walking a z slab of kz node layers.
// ---------------------------------------------------------------------------------------
//
// Thread t holds voxel (vx, vy) = (t % (OX+1), t / (OX+1)) of the tile, i.e. voxel
// (i0-1+vx, j0-1+vy); owned node (ox, oy) = (vx-1, vy-1) is corner (0,0) of that voxel, so
// its owner keeps one corner of each face in registers and reads the other three from the
// voxels to its left, below and diagonally below-left through LDS -- in the same order as
// the marching kernel, face by face (bottom face, then top face).  The top-face sum of
// voxel layer kv is carried in registers to node layer kv+1.  A slab evaluates only its
// own voxel layers [k0, k1): its first node layer gets the bottom-face sums here, and the
// top-face sums of its last voxel layer go to a seam buffer; hex8_seam_finalize_kernel,
// launched next on the stream, adds each seam into the first node layer of the slab
// above as (top + bottom) -- the marching kernel's sum, operand for operand.
// Values loaded from memory are made ordinary register values at this point: the empty
// asm reads and redefines them, so the compiler waits for their loads here rather than at
// a later use, where its conservative count would also wait for loads issued since.
__device__ __forceinline__ void settle(double &a, double &b, int32_t &c) {
    asm volatile("" : "+v"(a), "+v"(b), "+v"(c));
}
__device__ __forceinline__ void settle2(double &a, double &b) { asm volatile("" : "+v"(a), "+v"(b)); }
__device__ __forceinline__ void settle1(double &a) { asm volatile("" : "+v"(a)); }
__device__ __forceinline__ void settle_u(uint32_t &a) { asm volatile("" : "+v"(a)); }

// threads of the voxel-parallel block: the voxel tile rounded up to whole waves
template <int OX, int OY>
constexpr int apply_threads() {
    return ((OX + 1) * (OY + 1) + 63) / 64 * 64;
}

// Waves per SIMD the register budget must allow: a block of four waves puts one wave on each
// SIMD, so two resident blocks per CU need the registers of two waves per SIMD (at most 256
// per lane); larger blocks place two waves on a SIMD by themselves.  Only the isotropic
// variants without the energy sum fit that budget without spilling; the others keep one wave.
template <int OX, int OY, bool ISO, bool ENERGY>
constexpr int apply_min_waves() {
    return (apply_threads<OX, OY>() <= 256 && ISO && !ENERGY) ? 2 : 1;
}

// CL: crack closure off (0), at each point's own sign pattern (1) or at A.pat (2).
// PH: the voxel's (lam, mu) from its 1-byte phase and the plan's phase table (held in LDS)
// instead of the two per-voxel arrays (isotropic cells without closure).
template <int OX, int OY, int GMODE, bool ISO, bool ENERGY, int CL = 0, bool PH = false>
__global__ __launch_bounds__((apply_threads<OX, OY>()), (apply_min_waves<OX, OY, ISO, ENERGY>())) void hex8_apply_kernel(
    const Hex8Geo G, MechArgs3 A) {
    constexpr int NT = apply_threads<OX, OY>();
    constexpr int NVX = OX + 1, NVY = OY + 1, NV = NVX * NVY;
    constexpr int NNX = OX + 2, NNY = OY + 2, NN = NNX * NNY;
    constexpr int NF = (GMODE == 2) ? 4 : 3;
    constexpr int NS = (NN + NT - 1) / NT;  // node slots loaded per thread
    static_assert(NV <= NT, "voxel tile larger than the block");
    static_assert(!PH || (ISO && CL == 0), "the phase table serves the isotropic apply without closure");
    // node layers: NN slots and one dummy slot that lanes past the tile's nodes write to
    __shared__ double su[2][NF][NN + 1];
    __shared__ double sph[PH ? 2 * kHex8MaxPhases : 1];  // (lam, mu) per phase
    __shared__ double st[18][NV];  // corners (1,0), (0,1), (1,1) of both faces, 3 components each
    __shared__ double red[7 * (NT / 64)];

    const int nx = G.nx, ny = G.ny, nz = G.nz;
    const int tiles = A.tiles_x * A.tiles_y;
    const int blk = xcd_tile(blockIdx.x, gridDim.x);  // tile index; partials go to part[8 blk + k]
    const int bxy = blk % tiles, bz = blk / tiles;
    const int i0 = (bxy % A.tiles_x) * OX, j0 = (bxy / A.tiles_x) * OY;
    const int k0 = bz * A.kz, k1 = min(k0 + A.kz, nz);
    const int t = threadIdx.x;
    const int vx = t % NVX, vy = t / NVX;
    const bool has_voxel = t < NV;
    const int iv = i0 - 1 + vx, jv = j0 - 1 + vy;
    const bool owner = has_voxel && vx >= 1 && vy >= 1 && iv < nx && jv < ny;  // node (iv, jv)
    const int e_ij = has_voxel ? wrapm(iv, nx) + nx * wrapm(jv, ny) : 0;
    const int64_t n = (int64_t)nx * ny * nz;
    const int64_t nxy = (int64_t)nx * ny;
    double E[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) E[k] = A.x[3 * n + k];

    // in-layer offsets of this thread's node slots; a slot past the tile's NN nodes loads
    // node (0,0) of the layer into LDS padding, so every lane loads and stores without
    // branching (which keeps the compiler's wait counts exact across the loop)
    // Addresses are a uniform layer base (scalar registers) plus a 32-bit byte offset per
    // lane within the layer (24 nx ny < 2^32, checked at plan creation), so the loads and
    // stores need no 64-bit vector address arithmetic per layer.
    uint32_t slot_ij[NS];
#pragma unroll
    for (int m = 0; m < NS; ++m) {
        const int s = t + m * NT;
        slot_ij[m] = s < NN ? (uint32_t)(wrapm(i0 - 1 + s % NNX, nx) + nx * wrapm(j0 - 1 + s / NNX, ny)) : 0u;
    }
    // node layer k of this thread's slots, fetched into registers (issued one layer ahead)
    double pf[NS][NF];
    auto fetch = [&](int k, double (&dst)[NS][NF]) {
        const int64_t base = nxy * wrap(k, nz);
        const char *xb = reinterpret_cast<const char *>(A.x + 3 * base);
        const char *db = reinterpret_cast<const char *>(A.d + base);
#pragma unroll
        for (int m = 0; m < NS; ++m) {
            const double *px = reinterpret_cast<const double *>(xb + 24u * slot_ij[m]);
            dst[m][0] = px[0];
            dst[m][1] = px[1];
            dst[m][2] = px[2];
            if (GMODE == 2) dst[m][NF - 1] = *reinterpret_cast<const double *>(db + 8u * slot_ij[m]);
        }
    };
    int slot_lds[NS];  // LDS slot of each node slot (NN for the padding lanes)
#pragma unroll
    for (int m = 0; m < NS; ++m) slot_lds[m] = min(t + m * NT, NN);
    auto stage = [&](int buf, const double (&src)[NS][NF]) {
#pragma unroll
        for (int m = 0; m < NS; ++m)
#pragma unroll
            for (int f = 0; f < NF; ++f) su[buf][f][slot_lds[m]] = src[m][f];
    };

    double sE[6] = {0, 0, 0, 0, 0, 0}, en = 0.0;
    double acc[3] = {0.0, 0.0, 0.0};
    const int s00 = t - NVX - 1, s10 = t - NVX, s01 = t - 1;  // voxels left-below, below, left
    const uint32_t y_off = owner ? (uint32_t)(iv + nx * jv) : 0u;  // owned node within its layer
    // voxel material of layer k (lam, mu or the bank index; with the closure, the closure
    // flag in place of the bank index and the pattern word), fetched one layer ahead
    double ml0 = 0.0, ml1 = 0.0;
    int32_t mb = 0;
    uint32_t mp = 0u;
    const uint32_t e_off = (uint32_t)e_ij;
    auto fetch_mat = [&](int k, double &l0, double &l1, int32_t &b, uint32_t &pw) {
        const int64_t base = nxy * wrap(k, nz);
        if (PH) {
            b = (A.phase + base)[e_off];
        } else if (ISO) {
            l0 = (A.lam + base)[e_off];
            l1 = (A.mu + base)[e_off];
            if (CL != 0) b = (A.clm + base)[e_off];
            if (CL == 2) pw = (A.pat + base)[e_off];
        } else {
            b = (A.bank_index + base)[e_off];
            if (CL != 0) l0 = (A.clm + base)[e_off] != 0 ? 1.0 : 0.0;
            if (CL == 2) pw = (A.pat + base)[e_off];
        }
    };
    if (PH)  // visible after the first barrier of the layer loop
        for (int i = t; i < 2 * A.n_phase; i += NT) sph[i] = A.ptab[i];
    fetch_mat(k0, ml0, ml1, mb, mp);
    {  // the first two node layers, both in flight before either is staged
        double pf1[NS][NF];
        fetch(k0, pf);
        fetch(k0 + 1, pf1);
        stage(0, pf);
        stage(1, pf1);
    }
    settle(ml0, ml1, mb);
    if (CL == 2) settle_u(mp);
#pragma unroll
    for (int k = 0; k < 6; k += 2) settle2(E[k], E[k + 1]);
    int cur = 0;
    for (int kv = k0; kv < k1; ++kv) {
        __syncthreads();
        if (kv + 2 <= k1) fetch(kv + 2, pf);  // lands in su[cur] once this layer's voxels are done
        double nl0 = 0.0, nl1 = 0.0;
        int32_t nb = 0;
        uint32_t np = 0u;
        if (kv + 1 < k1) fetch_mat(kv + 1, nl0, nl1, nb, np);
        double r[3][8];
        if (PH) {
            ml0 = sph[2 * mb];
            ml1 = sph[2 * mb + 1];
        }
        if (has_voxel) {
            const int64_t e = e_ij + nxy * wrap(kv, nz);
            double u[3][8], g[8];
#pragma unroll
            for (int lz = 0; lz < 2; ++lz) {
                const int buf = lz ? (cur ^ 1) : cur;
#pragma unroll
                for (int ly = 0; ly < 2; ++ly)
#pragma unroll
                    for (int lx = 0; lx < 2; ++lx) {
                        const int slot = (vy + ly) * NNX + (vx + lx), a = lx + 2 * ly + 4 * lz;
                        u[0][a] = su[buf][0][slot];
                        u[1][a] = su[buf][1][slot];
                        u[2][a] = su[buf][2][slot];
                        if (GMODE == 2) g[a] = su[buf][NF - 1][slot];
                    }
            }
            if (GMODE == 2) {
                double dq[8];
                voxel_interp(G, g, dq);
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const double omd = 1.0 - dq[q];
                    g[q] = G.omk * (omd * omd) + G.kappa;  // splits.py:56-59
                }
            } else if (GMODE == 1) {
                load_qp8(A.g, e, g);
            } else {
#pragma unroll
                for (int q = 0; q < 8; ++q) g[q] = 1.0;
            }
            if (ISO) {
                const double lc = (CL != 0 && mb) ? ml0 : 0.0, mc = (CL != 0 && mb) ? ml1 : 0.0;
                voxel_mech<true, ENERGY, CL>(G, u, E, g, nullptr, ml0, ml1, owner, r, sE, en, lc, mc, mp);
            } else {
                voxel_mech<false, ENERGY, CL>(G, u, E, g, A.bank + 36 * (int64_t)mb, 0.0, 0.0, owner, r, sE, en,
                                              0.0, 0.0, mp, CL != 0 ? A.cltab + 28 * (int64_t)mb : nullptr,
                                              ml0 != 0.0);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                st[c][t] = r[c][1];
                st[3 + c][t] = r[c][2];
                st[6 + c][t] = r[c][3];
                st[9 + c][t] = r[c][5];
                st[12 + c][t] = r[c][6];
                st[15 + c][t] = r[c][7];
            }
        }
        __syncthreads();
        if (kv + 2 <= k1) stage(cur, pf);  // node layer kv+2 replaces layer kv
        // next layer's material: wait for it here, before this layer's stores are issued
        ml0 = nl0;
        ml1 = nl1;
        mb = nb;
        mp = np;
        settle(ml0, ml1, mb);
        if (CL == 2) settle_u(mp);
        if (owner) {
            double *yl = A.y + 3 * (nxy * kv) + 3u * y_off;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double bot = ((st[6 + c][s00] + st[3 + c][s10]) + st[c][s01]) + r[c][0];
                yl[c] = kv == k0 ? bot : acc[c] + bot;  // first layer: seam added later
                acc[c] = ((st[15 + c][s00] + st[12 + c][s10]) + st[9 + c][s01]) + r[c][4];
            }
        }
        cur ^= 1;
    }
    if (owner) {  // top faces of voxel layer k1-1 -> seam of the slab holding node layer k1 (mod nz)
        const int nslabs = (int)gridDim.x / tiles;
        const int64_t at = ((int64_t)((bz + 1) % nslabs) * nxy + iv + (int64_t)nx * jv) * 3;
#pragma unroll
        for (int c = 0; c < 3; ++c) A.seam[at + c] = acc[c];
    }
    macro_epilogue<NT, ENERGY>(A, sE, en, red, blk);
}

// Sum the block partials of the macro rows in a fixed order (block order per row, all
// rows' loads in flight together); write into y[3n + k] and/or out7.
__global__ __launch_bounds__(256) void hex8_finalize_kernel(const double *part, int nblocks, int nvals,
                                                            double *y6, double *out7) {
    __shared__ double red[7 * 4];
    double s[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    sum_block_partials<256, 7>(part, nblocks, nvals, s);
    block_sum_n<256, 7>(s, red);
    if (threadIdx.x == 0)
        for (int k = 0; k < nvals; ++k) {
            if (y6 && k < 6) y6[k] = s[k];
            if (out7) out7[k] = s[k];
        }
}

// After the voxel-parallel apply: block (0, 0) sums the macro-row partials (as
// hex8_finalize_kernel); the blocks (x > 0, slab) add that slab's seam (top-face sums of the
// voxel layer below its first node layer, computed by the slab below) to that node layer:
// y = seam + bottom, the same sum the marching kernel forms.  Grid: (1 + ceil(3 nx ny / 256),
// slabs); the slab is the block row, so no thread divides an index.
__global__ __launch_bounds__(256) void hex8_seam_finalize_kernel(const Hex8Geo G, const double *seam, int kz,
                                                                 double *y, const double *part, int nblocks,
                                                                 int nvals, double *y6, double *out7) {
    if (blockIdx.x == 0) {
        if (blockIdx.y != 0) return;
        __shared__ double red[7 * 4];
        double s[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        sum_block_partials<256, 7>(part, nblocks, nvals, s);
        block_sum_n<256, 7>(s, red);
        if (threadIdx.x == 0)
            for (int k = 0; k < nvals; ++k) {
                if (y6 && k < 6) y6[k] = s[k];
                if (out7) out7[k] = s[k];
            }
        return;
    }
    const int64_t m = 3 * (int64_t)G.nx * G.ny;  // doubles per node layer
    const int64_t i = (int64_t)(blockIdx.x - 1) * 256 + threadIdx.x;
    if (i >= m) return;
    const int64_t sl = blockIdx.y, at = m * (sl * kz) + i;
    y[at] = seam[m * sl + i] + y[at];
}

// A multigrid smoothing step's finalize: the macro rows (block (0, 0)) and, after the
// voxel-parallel apply, the seams, fused with the damped Jacobi update
// x += omega (b - A x) / diag on free dofs (fixed dofs stay 0) -- k_jacobi's arithmetic.
// Grid: (1 + ceil(3 nx ny / 256), nz), one block row per node layer.
__global__ __launch_bounds__(256) void hex8_finalize_jacobi_kernel(const Hex8Geo G, const double *seam, int kz,
                                                                   double *Ax, const double *part, int nblocks,
                                                                   const double *b, const double *diag,
                                                                   const uint8_t *fixed, double omega, double *x) {
    const int64_t m = 3 * (int64_t)G.nx * G.ny, n3 = m * G.nz;
    if (blockIdx.x == 0) {
        if (blockIdx.y != 0) return;
        __shared__ double red[6 * 4];
        double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        sum_block_partials<256, 6>(part, nblocks, 6, s);
        block_sum_n<256, 6>(s, red);
        const int k = threadIdx.x;
        if (k < 6) {
            const double v = k == 0 ? s[0] : k == 1 ? s[1] : k == 2 ? s[2] : k == 3 ? s[3] : k == 4 ? s[4] : s[5];
            const int64_t i = n3 + k;
            Ax[i] = v;
            if (fixed[i]) {
                x[i] = 0.0;
            } else {
                const double r = b[i] - v;
                x[i] = x[i] + omega * r / diag[i];
            }
        }
        return;
    }
    const int64_t il = (int64_t)(blockIdx.x - 1) * 256 + threadIdx.x;  // within the node layer
    if (il >= m) return;
    const int layer = blockIdx.y;
    const int64_t i = m * layer + il;
    double v = Ax[i];
    if (seam && layer % kz == 0) {
        v = seam[m * (layer / kz) + il] + v;
        Ax[i] = v;
    }
    if (fixed[i]) {
        x[i] = 0.0;
        return;
    }
    const double r = b[i] - v;
    x[i] = x[i] + omega * r / diag[i];
}

__global__ __launch_bounds__(256) void hex8_degradation_kernel(const Hex8Geo G, const double *d, double *g) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    if (e >= n) return;
    const int i = (int)(e % G.nx), j = (int)((e / G.nx) % G.ny), k = (int)(e / ((int64_t)G.nx * G.ny));
    double dc[8], dq[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const int ii = wrap(i + (a & 1), G.nx), jj = wrap(j + ((a >> 1) & 1), G.ny),
                  kk = wrap(k + (a >> 2), G.nz);
        dc[a] = d[ii + (int64_t)G.nx * (jj + (int64_t)G.ny * kk)];
    }
    voxel_interp(G, dc, dq);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const double omd = 1.0 - dq[q];
        g[e * 8 + q] = G.omk * (omd * omd) + G.kappa;
    }
}

// Jacobi diagonal: one thread per node gathers its 8 voxels (operator.py:1196-1254); the
// thread of node n also forms voxel n's contribution to the macro block.  With the crack
// closure (clmode 1: every bit, 2: the bits of `pat`; isotropic voxels) each closure bit k of
// Gauss point q adds w (1 - g_q) k_k (B_q^T v_k)^2 (operator.py:1255-1300): for dof c of a
// node, (d N / d x_c)^2 (s_0 lam + s_{1+c} 2 mu), and s_0 lam + s_{1+i} 2 mu to macro row i < 3.
//
// General banks (cltab): the rows v_k and moduli k_k of the voxel's bank entry, with
// (B_q^T v)_x = bx v_0 + bz v_4 + by v_5, (B_q^T v)_y = by v_1 + bz v_3 + bx v_5,
// (B_q^T v)_z = bz v_2 + by v_3 + bx v_4, and k_k v_i^2 for each macro row i.
__global__ __launch_bounds__(256) void hex8_diag_kernel(const Hex8Geo G, const double *g, const double *bank,
                                                        const int32_t *bank_index, const double *lam,
                                                        const double *mu, int iso, double *diag, double *part,
                                                        const uint8_t *clm = nullptr, const uint32_t *pat = nullptr,
                                                        int clmode = 0, const double *cltab = nullptr) {
    __shared__ double red[4];
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    const int64_t node = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double dE[6] = {0, 0, 0, 0, 0, 0};
    if (node < n) {
        const int i = (int)(node % G.nx), j = (int)((node / G.nx) % G.ny), k = (int)(node / ((int64_t)G.nx * G.ny));
        double acc[3] = {0.0, 0.0, 0.0};
        for (int v = 0; v < 8; ++v) {  // voxels in increasing (dz,dy,dx) order
            const int dx = v & 1, dy = (v >> 1) & 1, dz = v >> 2;
            const int64_t e = wrap(i - 1 + dx, G.nx) + (int64_t)G.nx * (wrap(j - 1 + dy, G.ny) +
                                                                      (int64_t)G.ny * wrap(k - 1 + dz, G.nz));
            const int lx = 1 - dx, ly = 1 - dy, lz = 1 - dz;
            const double sx = lx ? 1.0 : -1.0, sy = ly ? 1.0 : -1.0, sz = lz ? 1.0 : -1.0;
            double C[36];
            voxel_C(iso, bank, bank_index, lam, mu, e, C);
            for (int q = 0; q < 8; ++q) {
                const int ix = q & 1, iy = (q >> 1) & 1, iz = q >> 2;
                const double Wx = (lx == ix) ? G.al : G.be, Wy = (ly == iy) ? G.al : G.be,
                             Wz = (lz == iz) ? G.al : G.be;
                const double bx = sx * G.ihx * Wy * Wz, by = sy * G.ihy * Wx * Wz, bz = sz * G.ihz * Wx * Wy;
                const double gw = (g ? g[e * 8 + q] : 1.0) * G.w;
                // B columns of the three dofs: x: rows (0 bx, 4 bz, 5 by); y: (1 by, 3 bz, 5 bx); z: (2 bz, 3 by, 4 bx)
                const int rx[3] = {0, 4, 5}, ry[3] = {1, 3, 5}, rz[3] = {2, 3, 4};
                const double vx[3] = {bx, bz, by}, vy[3] = {by, bz, bx}, vz[3] = {bz, by, bx};
                double tx = 0.0, ty = 0.0, tz = 0.0;
#pragma unroll
                for (int p = 0; p < 3; ++p)
#pragma unroll
                    for (int m = 0; m < 3; ++m) {
                        tx += C[rx[p] * 6 + rx[m]] * vx[p] * vx[m];
                        ty += C[ry[p] * 6 + ry[m]] * vy[p] * vy[m];
                        tz += C[rz[p] * 6 + rz[m]] * vz[p] * vz[m];
                    }
                acc[0] += gw * tx;
                acc[1] += gw * ty;
                acc[2] += gw * tz;
                if (clmode != 0 && clm[e]) {
                    const uint32_t bits = clmode == 2 ? (pat[e] >> (4 * q)) & 15u : 15u;
                    const double cw = (1.0 - g[e * 8 + q]) * G.w;
                    if (cltab) {
                        const double *T = cltab + 28 * (int64_t)bank_index[e];
                        for (int b = 0; b < 4; ++b) {
                            if (!((bits >> b) & 1u)) continue;
                            const double *v = T + 6 * b;
                            const double px = (bx * v[0] + bz * v[4]) + by * v[5];
                            const double py = (by * v[1] + bz * v[3]) + bx * v[5];
                            const double pz = (bz * v[2] + by * v[3]) + bx * v[4];
                            const double ck = cw * T[24 + b];
                            acc[0] += ck * (px * px);
                            acc[1] += ck * (py * py);
                            acc[2] += ck * (pz * pz);
                        }
                    } else {
                        const double l0 = (bits & 1u) ? lam[e] : 0.0, m2 = 2.0 * mu[e];
                        acc[0] += cw * (bx * bx) * (l0 + ((bits & 2u) ? m2 : 0.0));
                        acc[1] += cw * (by * by) * (l0 + ((bits & 4u) ? m2 : 0.0));
                        acc[2] += cw * (bz * bz) * (l0 + ((bits & 8u) ? m2 : 0.0));
                    }
                }
            }
        }
        diag[3 * node] = acc[0];
        diag[3 * node + 1] = acc[1];
        diag[3 * node + 2] = acc[2];
        // macro block of voxel `node`: (sum_q g w) C_kk
        double gws = 0.0;
        for (int q = 0; q < 8; ++q) gws += (g ? g[node * 8 + q] : 1.0) * G.w;
        double C[36];
        voxel_C(iso, bank, bank_index, lam, mu, node, C);
#pragma unroll
        for (int k2 = 0; k2 < 6; ++k2) dE[k2] = gws * C[k2 * 7];
        if (clmode != 0 && clm[node]) {
            for (int q = 0; q < 8; ++q) {
                const uint32_t bits = clmode == 2 ? (pat[node] >> (4 * q)) & 15u : 15u;
                const double cw = (1.0 - g[node * 8 + q]) * G.w;
                if (cltab) {
                    const double *T = cltab + 28 * (int64_t)bank_index[node];
                    for (int b = 0; b < 4; ++b) {
                        if (!((bits >> b) & 1u)) continue;
                        const double ck = cw * T[24 + b];
#pragma unroll
                        for (int k2 = 0; k2 < 6; ++k2) dE[k2] += ck * (T[6 * b + k2] * T[6 * b + k2]);
                    }
                } else {
                    const double l0 = (bits & 1u) ? lam[node] : 0.0, m2 = 2.0 * mu[node];
                    dE[0] += cw * (l0 + ((bits & 2u) ? m2 : 0.0));
                    dE[1] += cw * (l0 + ((bits & 4u) ? m2 : 0.0));
                    dE[2] += cw * (l0 + ((bits & 8u) ? m2 : 0.0));
                }
            }
        }
    }
#pragma unroll
    for (int k2 = 0; k2 < 6; ++k2) {
        const double t = block_sum<256>(dE[k2], red);
        if (threadIdx.x == 0) part[8 * blockIdx.x + k2] = t;
    }
}

// Undamaged driving energy per Gauss point by driver code (operator.py:1415-1444).
__global__ __launch_bounds__(256) void hex8_driving_kernel(const Hex8Geo G, const double *x, const double *bank,
                                                           const int32_t *bank_index, const double *lam,
                                                           const double *mu, const int8_t *driver, double *Y) {
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const int code = driver[e];
    if (code != 1 && code != 2) {
#pragma unroll
        for (int q = 0; q < 8; ++q) Y[e * 8 + q] = 0.0;
        return;
    }
    const int i = (int)(e % G.nx), j = (int)((e / G.nx) % G.ny), k = (int)(e / ((int64_t)G.nx * G.ny));
    double u[3][8], E[6];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const int64_t node = wrap(i + (a & 1), G.nx) +
                             (int64_t)G.nx * (wrap(j + ((a >> 1) & 1), G.ny) + (int64_t)G.ny * wrap(k + (a >> 2), G.nz));
        u[0][a] = x[3 * node];
        u[1][a] = x[3 * node + 1];
        u[2][a] = x[3 * node + 2];
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) E[c] = x[3 * n + c];
    double gx[3][2][2], gy[3][2][2], gz[3][2][2];
    voxel_grad(G, u, gx, gy, gz);
    fold_macro(E, gx, gy, gz);
    const double *C = bank + 36 * (int64_t)bank_index[e];
    const double la = lam[e], m = mu[e];
    for (int q = 0; q < 8; ++q) {
        const int ix = q & 1, iy = (q >> 1) & 1, iz = q >> 2;
        double s[6];
        qp_strain(gx, gy, gz, ix, iy, iz, s);
        double val;
        if (code == 1) {  // 0.5 eps.C.eps (splits.py:68-72)
            double t = 0.0;
            for (int kk = 0; kk < 6; ++kk) {
                double c6 = 0.0;
                for (int mm = 0; mm < 6; ++mm) c6 += C[kk * 6 + mm] * s[mm];
                t += s[kk] * c6;
            }
            val = 0.5 * t;
        } else {  // Miehe spectral (splits.py:146-161)
            double ev[3];
            eig3(s[0], s[1], s[2], 0.5 * s[5], 0.5 * s[4], 0.5 * s[3], ev);
            const double trp = fmax(s[0] + s[1] + s[2], 0.0);
            const double e0 = fmax(ev[0], 0.0), e1 = fmax(ev[1], 0.0), e2 = fmax(ev[2], 0.0);
            val = 0.5 * la * (trp * trp) + m * ((e0 * e0 + e1 * e1) + e2 * e2);
        }
        Y[e * 8 + q] = val;
    }
}

// ---------------------------------------------------------------------------------------
// AT2 damage operator on the voxel grid (operator.py:1450-1494)
// ---------------------------------------------------------------------------------------

struct DmgArgs3 {
    const double *d, *H, *Gc;
    double *y;
    double scale;
    int32_t kz, tiles_x, tiles_y;
};

// The voxel matrix K and the Gauss-point shape values Nq (64 doubles each) staged in LDS at
// block start for the voxel-parallel damage apply: read from the kernel argument inside the
// unrolled 8x8 products, they need 256 scalar registers, and the compiler spills those to
// vector lanes (v_readlane / v_writelane on every use).  Same values in the same order, so
// the same bits as the marching kernel, which keeps the kernel-argument reads (staged there,
// the compiler holds both tables in vector registers).
__device__ __forceinline__ void stage_voxel_tables(const Hex8Geo &G, double *sK, double *sN) {
    for (int i = threadIdx.x; i < 128; i += blockDim.x) {
        if (i < 64) sK[i] = G.K[i];
        else sN[i - 64] = G.Nq[i - 64];
    }
}

template <int TX, int TY, bool HIST>
__global__ __launch_bounds__(TX *TY) void hex8_damage_kernel(const Hex8Geo G, DmgArgs3 A) {
    constexpr int NT = TX * TY;
    constexpr int NXH = TX + 2, NYH = TY + 2, NH = NXH * NYH;
    constexpr int NXV = TX + 1, NYV = TY + 1, NV = NXV * NYV;
    __shared__ double su[2][NH];
    __shared__ double sr[8][NV];
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    const int bx = blockIdx.x % A.tiles_x, by = (blockIdx.x / A.tiles_x) % A.tiles_y;
    const int bz = blockIdx.x / (A.tiles_x * A.tiles_y);
    const int i0 = bx * TX, j0 = by * TY, k0 = bz * A.kz;
    const int k1 = min(k0 + A.kz, nz);
    auto load_layer = [&](int buf, int k) {
        const int kk = wrap(k, nz);
        for (int s = threadIdx.x; s < NH; s += NT) {
            const int i = wrapm(i0 + s % NXH - 1, nx), j = wrapm(j0 + s / NXH - 1, ny);
            su[buf][s] = A.d[i + (int64_t)nx * (j + (int64_t)ny * kk)];
        }
    };
    double acc = 0.0;
    int cur = 0;
    load_layer(0, k0 - 1);
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    for (int kv = k0 - 1; kv < k1; ++kv) {
        load_layer(cur ^ 1, kv + 1);
        __syncthreads();
        const int kw = wrap(kv, nz);
        for (int s = threadIdx.x; s < NV; s += NT) {
            const int li = s % NXV - 1, lj = s / NXV - 1;
            const int64_t e = wrapm(i0 + li, nx) + (int64_t)nx * (wrapm(j0 + lj, ny) + (int64_t)ny * kw);
            const int b0 = (lj + 1) * NXH + (li + 1);
            double dc[8];
#pragma unroll
            for (int a = 0; a < 8; ++a) dc[a] = su[(a >> 2) ? (cur ^ 1) : cur][b0 + ((a >> 1) & 1) * NXH + (a & 1)];
            const double gc = A.Gc[e];
            double y[8];
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                double t = 0.0;
#pragma unroll
                for (int b = 0; b < 8; ++b) t += dc[b] * G.K[b * 8 + a];
                y[a] = gc * t;
            }
            if (HIST) {
                double dq[8];
                voxel_interp(G, dc, dq);
                double hq[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) hq[q] = (A.H[e * 8 + q] * G.w) * dq[q];
#pragma unroll
                for (int a = 0; a < 8; ++a) {
                    double t = 0.0;
#pragma unroll
                    for (int q = 0; q < 8; ++q) t += hq[q] * G.Nq[q * 8 + a];
                    y[a] += G.ch * t;
                }
            }
#pragma unroll
            for (int a = 0; a < 8; ++a) sr[a][s] = y[a];
        }
        __syncthreads();
        const int s00 = ty * NXV + tx, s10 = s00 + 1, s01 = s00 + NXV, s11 = s01 + 1;
        const double bot = ((sr[3][s00] + sr[2][s10]) + sr[1][s01]) + sr[0][s11];
        const double top = ((sr[7][s00] + sr[6][s10]) + sr[5][s01]) + sr[4][s11];
        const int i = i0 + tx, j = j0 + ty;
        if (kv >= k0 && i < nx && j < ny) A.y[i + (int64_t)nx * (j + (int64_t)ny * kv)] = A.scale * (acc + bot);
        acc = top;
        cur ^= 1;
    }
}

// rhs / diagonal of the damage operator: thread per node over its 8 voxels.
__global__ __launch_bounds__(256) void hex8_damage_rd_kernel(const Hex8Geo G, const double *H, const double *Gc,
                                                             double scale, double *b, double *diag) {
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    const int64_t node = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (node >= n) return;
    const int i = (int)(node % G.nx), j = (int)((node / G.nx) % G.ny), k = (int)(node / ((int64_t)G.nx * G.ny));
    double sb = 0.0, sd = 0.0;
    for (int v = 0; v < 8; ++v) {
        const int dx = v & 1, dy = (v >> 1) & 1, dz = v >> 2;
        const int64_t e = wrap(i - 1 + dx, G.nx) +
                          (int64_t)G.nx * (wrap(j - 1 + dy, G.ny) + (int64_t)G.ny * wrap(k - 1 + dz, G.nz));
        const int a = (1 - dx) + 2 * (1 - dy) + 4 * (1 - dz);
        double hb = 0.0, hd = 0.0;
        if (H) {
            for (int q = 0; q < 8; ++q) {
                const double hw = H[e * 8 + q] * G.w;
                hb += hw * G.Nq[q * 8 + a];
                hd += hw * (G.Nq[q * 8 + a] * G.Nq[q * 8 + a]);
            }
        }
        sb += G.ch * hb;
        sd += Gc[e] * G.K[a * 8 + a] + G.ch * hd;
    }
    if (b) b[node] = scale * sb;
    if (diag) diag[node] = scale * sd;
}

// Damped Jacobi update fused into the apply's finalize (multigrid smoothing).
struct JacobiSpec {
    const double *b, *diag;
    const uint8_t *fixed;
    double omega;
    double *x;
};

template <int TX, int TY>
static int launch_mech_t(const phast_hex8_plan *pl, MechArgs3 A, int gmode, bool energy, hipStream_t st,
                         int *nblocks, const JacobiSpec *jac) {
    A.tiles_x = (int)ceil_div(pl->geo.nx, TX);
    A.tiles_y = (int)ceil_div(pl->geo.ny, TY);
    const int tz = (int)ceil_div(pl->geo.nz, A.kz);
    const int nb = A.tiles_x * A.tiles_y * tz;
    *nblocks = nb;
    if (nb > pl->max_blocks) {
        set_error("hex8 apply needs %d blocks, plan holds partials for %d", nb, pl->max_blocks);
        return PHAST_ERR_ARG;
    }
    dim3 grid(nb), block(TX * TY);
#define PHAST_L(GM, ISO, EN) \
    hipLaunchKernelGGL((hex8_mech_kernel<TX, TY, GM, ISO, EN>), grid, block, 0, st, pl->geo, A)
    if (pl->iso) {
        if (energy) {
            if (gmode == 0) PHAST_L(0, true, true);
            else if (gmode == 1) PHAST_L(1, true, true);
            else PHAST_L(2, true, true);
        } else {
            if (gmode == 0) PHAST_L(0, true, false);
            else if (gmode == 1) PHAST_L(1, true, false);
            else PHAST_L(2, true, false);
        }
    } else {
        if (energy) {
            if (gmode == 0) PHAST_L(0, false, true);
            else if (gmode == 1) PHAST_L(1, false, true);
            else PHAST_L(2, false, true);
        } else {
            if (gmode == 0) PHAST_L(0, false, false);
            else if (gmode == 1) PHAST_L(1, false, false);
            else PHAST_L(2, false, false);
        }
    }
#undef PHAST_L
    PHAST_LAUNCHED();
    if (jac)
        hipLaunchKernelGGL(hex8_finalize_jacobi_kernel,
                           dim3((unsigned)(1 + ceil_div(3 * (int64_t)pl->geo.nx * pl->geo.ny, 256)), pl->geo.nz),
                           dim3(256), 0, st, pl->geo, (const double *)nullptr, 1, A.y, A.part, nb, jac->b, jac->diag,
                           jac->fixed, jac->omega, jac->x);
    else
        hipLaunchKernelGGL(hex8_finalize_kernel, dim3(1), dim3(256), 0, st, A.part, nb, A.nvals, A.y6, A.out7);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

template <int OX, int OY>
static int launch_apply_t(const phast_hex8_plan *pl, MechArgs3 A, int gmode, bool energy, hipStream_t st,
                          const JacobiSpec *jac, int cl = 0) {
    A.tiles_x = (int)ceil_div(pl->geo.nx, OX);
    A.tiles_y = (int)ceil_div(pl->geo.ny, OY);
    const int nbz = (int)ceil_div(pl->geo.nz, A.kz);
    const int nb = A.tiles_x * A.tiles_y * nbz;
    if (nb > pl->max_blocks) {
        set_error("hex8 apply needs %d blocks, plan holds partials for %d", nb, pl->max_blocks);
        return PHAST_ERR_ARG;
    }
    if (!pl->seam || pl->seam_cap < 3 * (int64_t)pl->geo.nx * pl->geo.ny * nbz) {
        set_error("internal: hex8 seam buffer too small for kz = %d", A.kz);
        return PHAST_ERR_ARG;
    }
    A.seam = pl->seam;
    dim3 grid(nb), block(apply_threads<OX, OY>());
#define PHAST_L(GM, ISO, EN) hipLaunchKernelGGL((hex8_apply_kernel<OX, OY, GM, ISO, EN>), grid, block, 0, st, pl->geo, A)
#define PHAST_LP(GM) \
    hipLaunchKernelGGL((hex8_apply_kernel<OX, OY, GM, true, false, 0, true>), grid, block, 0, st, pl->geo, A)
    const bool ph = A.phase != nullptr && !energy && cl == 0;
#define PHAST_LC(ISO, EN, CL) \
    hipLaunchKernelGGL((hex8_apply_kernel<OX, OY, 1, ISO, EN, CL>), grid, block, 0, st, pl->geo, A)
    if constexpr (apply_threads<OX, OY>() > 256) {
        // large blocks carry only the isotropic apply without energy or closure (the caller
        // sends the other variants to the 16x14 tile)
        if (!(pl->iso && !energy && cl == 0)) {
            set_error("internal: the %dx%d tile runs only the isotropic apply", OX, OY);
            return PHAST_ERR_ARG;
        }
        if (ph) {
            if (gmode == 0) PHAST_LP(0);
            else if (gmode == 1) PHAST_LP(1);
            else PHAST_LP(2);
        } else if (gmode == 0) PHAST_L(0, true, false);
        else if (gmode == 1) PHAST_L(1, true, false);
        else PHAST_L(2, true, false);
    } else if (cl != 0 && pl->iso) {  // crack closure: g array (checked by the caller)
        if (energy) {
            if (cl == 1) PHAST_LC(true, true, 1);
            else PHAST_LC(true, true, 2);
        } else {
            if (cl == 1) PHAST_LC(true, false, 1);
            else PHAST_LC(true, false, 2);
        }
    } else if (cl != 0) {  // general bank: the closure rows of each bank entry
        if (energy) {
            if (cl == 1) PHAST_LC(false, true, 1);
            else PHAST_LC(false, true, 2);
        } else {
            if (cl == 1) PHAST_LC(false, false, 1);
            else PHAST_LC(false, false, 2);
        }
    } else if (pl->iso) {
        if (energy) {
            if (gmode == 0) PHAST_L(0, true, true);
            else if (gmode == 1) PHAST_L(1, true, true);
            else PHAST_L(2, true, true);
        } else if (ph) {
            if (gmode == 0) PHAST_LP(0);
            else if (gmode == 1) PHAST_LP(1);
            else PHAST_LP(2);
        } else {
            if (gmode == 0) PHAST_L(0, true, false);
            else if (gmode == 1) PHAST_L(1, true, false);
            else PHAST_L(2, true, false);
        }
    } else {
        if (energy) {
            if (gmode == 0) PHAST_L(0, false, true);
            else if (gmode == 1) PHAST_L(1, false, true);
            else PHAST_L(2, false, true);
        } else {
            if (gmode == 0) PHAST_L(0, false, false);
            else if (gmode == 1) PHAST_L(1, false, false);
            else PHAST_L(2, false, false);
        }
    }
#undef PHAST_L
#undef PHAST_LP
#undef PHAST_LC
    PHAST_LAUNCHED();
    const unsigned layer_blocks = (unsigned)(1 + ceil_div(3 * (int64_t)pl->geo.nx * pl->geo.ny, 256));
    if (jac)
        hipLaunchKernelGGL(hex8_finalize_jacobi_kernel, dim3(layer_blocks, pl->geo.nz), dim3(256), 0, st, pl->geo,
                           A.seam, A.kz, A.y, A.part, nb, jac->b, jac->diag, jac->fixed, jac->omega, jac->x);
    else
        hipLaunchKernelGGL(hex8_seam_finalize_kernel, dim3(layer_blocks, nbz), dim3(256), 0, st, pl->geo, A.seam,
                           A.kz, A.y, A.part, nb, A.nvals, A.y6, A.out7);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

static int hex8_apply_any(const phast_hex8_plan *pl, const double *x, const double *g, const double *d, double *y,
                          double *out7, bool energy, hipStream_t st, const JacobiSpec *jac, int cl = 0,
                          const uint32_t *pat = nullptr) {
    MechArgs3 A;
    std::memset(&A, 0, sizeof(A));
    A.x = x;
    A.g = g;
    A.d = d;
    A.bank = pl->bank;
    A.bank_index = pl->bank_index;
    A.lam = pl->lam;
    A.mu = pl->mu;
    A.phase = pl->phase;
    A.ptab = pl->ptab;
    A.n_phase = pl->n_phase;
    A.y = y;
    A.part = pl->part;
    A.y6 = out7 ? nullptr : y + 3 * pl->n;
    A.out7 = out7;
    A.nvals = energy ? 7 : 6;
    A.kz = pl->kz;
    A.clm = pl->cl_on;
    A.pat = pat;
    A.cltab = pl->cl_tab;
    const int gmode = d ? 2 : (g ? 1 : 0);
    int nb = 0, rc;
    if (!y) {  // stress sums only: still needs a write target for the nodal rows
        set_error("internal: hex8 apply without output");
        return PHAST_ERR_ARG;
    }
    if (cl != 0) {
        if (!pl->cl_on || gmode != 1 || (!pl->iso && !pl->cl_tab) || (cl == 2 && !pat) || jac) {
            set_error("internal: the closure apply needs the closure set-up (with the rows of a general bank), "
                      "g (E,8) and a pattern when fixed");
            return PHAST_ERR_ARG;
        }
        if (!((pl->tx == 16 && (pl->ty == 16 || pl->ty == 14)) || (pl->tx == 32 && (pl->ty == 8 || pl->ty == 14 ||
                                                                                     pl->ty == 6)) ||
              (pl->tx == 64 && pl->ty == 6))) {
            set_error("the crack-closure apply runs on the voxel-parallel tiles; the plan has %dx%d", pl->tx, pl->ty);
            return PHAST_ERR_UNSUPPORTED;
        }
    }
    // Blocks of more than 256 threads (16x16, 32x8: 320; 32x14, 64x6: 512) put two waves on a
    // SIMD, so their kernels get at most 256 registers; only the isotropic apply without the
    // energy sum and the closure fits that (the others spill 240-670 B per lane).  Those run
    // on the 16x14 tile with the same slab depth: the nodal rows are the same bits on every
    // tile, the macro rows differ in rounding only.
    const bool lean = pl->iso && !energy && cl == 0;
    const bool big = (pl->tx == 16 && pl->ty == 16) || (pl->tx == 32 && (pl->ty == 8 || pl->ty == 14)) ||
                     (pl->tx == 64 && pl->ty == 6);
    if (big && !lean) rc = launch_apply_t<16, 14>(pl, A, gmode, energy, st, jac, cl);
    else if (pl->tx == 16 && pl->ty == 16) rc = launch_apply_t<16, 16>(pl, A, gmode, energy, st, jac, cl);
    else if (pl->tx == 32 && pl->ty == 8) rc = launch_apply_t<32, 8>(pl, A, gmode, energy, st, jac, cl);
    else if (pl->tx == 16 && pl->ty == 14) rc = launch_apply_t<16, 14>(pl, A, gmode, energy, st, jac, cl);
    else if (pl->tx == 32 && pl->ty == 14) rc = launch_apply_t<32, 14>(pl, A, gmode, energy, st, jac, cl);
    else if (pl->tx == 32 && pl->ty == 6) rc = launch_apply_t<32, 6>(pl, A, gmode, energy, st, jac, cl);
    else if (pl->tx == 64 && pl->ty == 6) rc = launch_apply_t<64, 6>(pl, A, gmode, energy, st, jac, cl);
    else if (pl->tx == 16 && pl->ty == 8) rc = launch_mech_t<16, 8>(pl, A, gmode, energy, st, &nb, jac);
    else if (pl->tx == 8 && pl->ty == 8) rc = launch_mech_t<8, 8>(pl, A, gmode, energy, st, &nb, jac);
    else {
        set_error("unsupported hex8 tile %dx%d", pl->tx, pl->ty);
        return PHAST_ERR_UNSUPPORTED;
    }
    return rc;
}

int hex8_apply_impl(const phast_hex8_plan *pl, const double *x, const double *g, const double *d, double *y,
                    double *out7, bool energy, hipStream_t st) {
    return hex8_apply_any(pl, x, g, d, y, out7, energy, st, nullptr);
}

int hex8_apply_closure(const phast_hex8_plan *pl, const double *x, const double *g, const uint32_t *pat, int mode,
                       double *y, double *out7, bool energy, hipStream_t st) {
    return hex8_apply_any(pl, x, g, nullptr, y, out7, energy, st, nullptr, mode, pat);
}

int hex8_apply_jacobi(const phast_hex8_plan *pl, double *x, const double *g, const double *d, double *Ax,
                      const double *b, const double *diag, const uint8_t *fixed, double omega, hipStream_t st) {
    const JacobiSpec jac{b, diag, fixed, omega, x};
    return hex8_apply_any(pl, x, g, d, Ax, nullptr, false, st, &jac);
}

// Voxel-parallel damage operator y = scale (Gc K d + ch N^T (w H) N d) (operator.py:1450-1509),
// organised as hex8_apply_kernel: one thread per voxel of an (OX+1) x (OY+1) tile, node
// layers of d and the voxel operands (Gc, H) fetched one layer ahead with branch-free loads,
// 6 corner values per voxel staged in LDS, and the owner of node (ox, oy) gathering its 8
// voxels in the marching kernel's order -- so y is bitwise the marching kernel's.  The slab
// starts one voxel layer early (k0-1) to form the carry into its first node layer.
template <int OX, int OY, bool HIST>
__global__ __launch_bounds__((apply_threads<OX, OY>())) void hex8_damage_vp_kernel(const Hex8Geo G, DmgArgs3 A) {
    constexpr int NT = apply_threads<OX, OY>();
    constexpr int NVX = OX + 1, NVY = OY + 1, NV = NVX * NVY;
    constexpr int NNX = OX + 2, NNY = OY + 2, NN = NNX * NNY;
    constexpr int NS = (NN + NT - 1) / NT;
    __shared__ double su[2][NS * NT];  // node layers of d; slots past NN are padding
    __shared__ double st[6][NV];       // corners 1, 2, 3, 5, 6, 7 of each voxel
    __shared__ double sK[64], sN[64];
    stage_voxel_tables(G, sK, sN);
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    const int tiles = A.tiles_x * A.tiles_y;
    const int blk = xcd_tile(blockIdx.x, gridDim.x);  // tile index of this block
    const int bxy = blk % tiles, bz = blk / tiles;
    const int i0 = (bxy % A.tiles_x) * OX, j0 = (bxy / A.tiles_x) * OY;
    const int k0 = bz * A.kz, k1 = min(k0 + A.kz, nz);
    const int t = threadIdx.x;
    const int vx = t % NVX, vy = t / NVX;
    const bool has_voxel = t < NV;
    const int iv = i0 - 1 + vx, jv = j0 - 1 + vy;
    const bool owner = has_voxel && vx >= 1 && vy >= 1 && iv < nx && jv < ny;
    const int e_ij = has_voxel ? wrapm(iv, nx) + nx * wrapm(jv, ny) : 0;
    const int64_t nxy = (int64_t)nx * ny;
    int64_t slot_ij[NS];
#pragma unroll
    for (int m = 0; m < NS; ++m) {
        const int s = t + m * NT;
        slot_ij[m] = s < NN ? wrapm(i0 - 1 + s % NNX, nx) + (int64_t)nx * wrapm(j0 - 1 + s / NNX, ny) : 0;
    }
    double pf[NS];
    auto fetch = [&](int k, double (&dst)[NS]) {
        const int64_t base = nxy * wrap(k, nz);
#pragma unroll
        for (int m = 0; m < NS; ++m) dst[m] = A.d[base + slot_ij[m]];
    };
    auto stage = [&](int buf, const double (&src)[NS]) {
#pragma unroll
        for (int m = 0; m < NS; ++m) su[buf][t + m * NT] = src[m];
    };
    auto fetch_vox = [&](int k, double &gc, double (&h)[8]) {
        const int64_t e = e_ij + nxy * wrap(k, nz);
        gc = A.Gc[e];
        if (HIST) load_qp8(A.H, e, h);
    };
    double gc0 = 0.0, h0[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    fetch_vox(k0 - 1, gc0, h0);
    {
        double pf1[NS];
        fetch(k0 - 1, pf);
        fetch(k0, pf1);
        stage(0, pf);
        stage(1, pf1);
    }
    settle1(gc0);
#pragma unroll
    for (int q = 0; q < 8; ++q) settle1(h0[q]);
    const int s00 = t - NVX - 1, s10 = t - NVX, s01 = t - 1;
    double acc = 0.0;
    int cur = 0;
    for (int kv = k0 - 1; kv < k1; ++kv) {
        __syncthreads();
        if (kv + 2 <= k1) fetch(kv + 2, pf);
        double gc1 = 0.0, h1[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (kv + 1 < k1) fetch_vox(kv + 1, gc1, h1);
        double y[8];
        if (has_voxel) {
            double dc[8];
#pragma unroll
            for (int a = 0; a < 8; ++a)
                dc[a] = su[(a >> 2) ? (cur ^ 1) : cur][(vy + ((a >> 1) & 1)) * NNX + (vx + (a & 1))];
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                double tt = 0.0;
#pragma unroll
                for (int b = 0; b < 8; ++b) tt += dc[b] * sK[b * 8 + a];
                y[a] = gc0 * tt;
            }
            if (HIST) {
                double dq[8];
                voxel_interp(G, dc, dq);
                double hq[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) hq[q] = (h0[q] * G.w) * dq[q];
#pragma unroll
                for (int a = 0; a < 8; ++a) {
                    double tt = 0.0;
#pragma unroll
                    for (int q = 0; q < 8; ++q) tt += hq[q] * sN[q * 8 + a];
                    y[a] += G.ch * tt;
                }
            }
            st[0][t] = y[1];
            st[1][t] = y[2];
            st[2][t] = y[3];
            st[3][t] = y[5];
            st[4][t] = y[6];
            st[5][t] = y[7];
        }
        __syncthreads();
        if (kv + 2 <= k1) stage(cur, pf);  // node layer kv+2 replaces layer kv
        gc0 = gc1;
        settle1(gc0);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            h0[q] = h1[q];
            settle1(h0[q]);
        }
        if (owner) {
            const double bot = ((st[2][s00] + st[1][s10]) + st[0][s01]) + y[0];
            const double top = ((st[5][s00] + st[4][s10]) + st[3][s01]) + y[4];
            if (kv >= k0) A.y[iv + (int64_t)nx * (jv + (int64_t)ny * kv)] = A.scale * (acc + bot);
            acc = top;
        }
        cur ^= 1;
    }
}

template <int OX, int OY>
static int launch_damage_vp_t(const phast_hex8_plan *pl, const Hex8Geo &G, DmgArgs3 A, hipStream_t st) {
    A.tiles_x = (int)ceil_div(pl->geo.nx, OX);
    A.tiles_y = (int)ceil_div(pl->geo.ny, OY);
    const int tz = (int)ceil_div(pl->geo.nz, A.kz);
    dim3 grid(A.tiles_x * A.tiles_y * tz), block(apply_threads<OX, OY>());
    if (A.H) hipLaunchKernelGGL((hex8_damage_vp_kernel<OX, OY, true>), grid, block, 0, st, G, A);
    else hipLaunchKernelGGL((hex8_damage_vp_kernel<OX, OY, false>), grid, block, 0, st, G, A);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

template <int TX, int TY>
static int launch_damage_t(const phast_hex8_plan *pl, const Hex8Geo &G, DmgArgs3 A, hipStream_t st) {
    A.tiles_x = (int)ceil_div(pl->geo.nx, TX);
    A.tiles_y = (int)ceil_div(pl->geo.ny, TY);
    const int tz = (int)ceil_div(pl->geo.nz, A.kz);
    dim3 grid(A.tiles_x * A.tiles_y * tz), block(TX * TY);
    if (A.H) hipLaunchKernelGGL((hex8_damage_kernel<TX, TY, true>), grid, block, 0, st, G, A);
    else hipLaunchKernelGGL((hex8_damage_kernel<TX, TY, false>), grid, block, 0, st, G, A);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

}  // namespace phast

using namespace phast;

// Host construction of the lexicographic-corner voxel matrices (solid3d/periodic.py:193-221;
// operator.py:242-247): N, gradN at the 2x2x2 Gauss points, M = sum_q w N N^T,
// L = sum_q w gradN gradN^T, both symmetrised, K = M/ell + ell L.
void phast::build_voxel_constants(Hex8Geo &G, double hx, double hy, double hz, double ell) {
    voxel_constants(G, hx, hy, hz, ell, nullptr, nullptr);
}

void phast::voxel_constants(Hex8Geo &G, double hx, double hy, double hz, double ell, double *Mo, double *Lo) {
    const double a = 1.0 / std::sqrt(3.0);
    const double pt[2] = {-a, a};
    double N[8][8], dN[8][8][3];
    for (int q = 0; q < 8; ++q) {
        const double xi = pt[q & 1], eta = pt[(q >> 1) & 1], zeta = pt[q >> 2];
        for (int c = 0; c < 8; ++c) {
            const double sx = (c & 1) ? 1.0 : -1.0, sy = ((c >> 1) & 1) ? 1.0 : -1.0, sz = (c >> 2) ? 1.0 : -1.0;
            N[q][c] = 0.125 * (1.0 + sx * xi) * (1.0 + sy * eta) * (1.0 + sz * zeta);
            dN[q][c][0] = 0.125 * sx * (1.0 + sy * eta) * (1.0 + sz * zeta) * (2.0 / hx);
            dN[q][c][1] = 0.125 * (1.0 + sx * xi) * sy * (1.0 + sz * zeta) * (2.0 / hy);
            dN[q][c][2] = 0.125 * (1.0 + sx * xi) * (1.0 + sy * eta) * sz * (2.0 / hz);
        }
    }
    const double w = hx * hy * hz / 8.0;
    double M[8][8] = {}, L[8][8] = {};
    for (int q = 0; q < 8; ++q)
        for (int i = 0; i < 8; ++i)
            for (int j = 0; j < 8; ++j) {
                M[i][j] += w * N[q][i] * N[q][j];
                L[i][j] += w * (dN[q][i][0] * dN[q][j][0] + dN[q][i][1] * dN[q][j][1] + dN[q][i][2] * dN[q][j][2]);
            }
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j) {
            const double Ms = 0.5 * (M[i][j] + M[j][i]), Ls = 0.5 * (L[i][j] + L[j][i]);
            G.K[i * 8 + j] = Ms / ell + ell * Ls;
            if (Mo) Mo[i * 8 + j] = Ms;
            if (Lo) Lo[i * 8 + j] = Ls;
        }
    for (int q = 0; q < 8; ++q)
        for (int c = 0; c < 8; ++c) G.Nq[q * 8 + c] = N[q][c];
    G.w = w;
    G.al = 0.5 * (1.0 + a);
    G.be = 0.5 * (1.0 - a);
    G.ihx = 1.0 / hx;
    G.ihy = 1.0 / hy;
    G.ihz = 1.0 / hz;
    G.alx = G.al * G.ihx;
    G.bex = G.be * G.ihx;
    G.aly = G.al * G.ihy;
    G.bey = G.be * G.ihy;
    G.alz = G.al * G.ihz;
    G.bez = G.be * G.ihz;
    G.c3 = a;
    G.kx = 0.25 / hx;
    G.ky = 0.25 / hy;
    G.kz = 0.25 / hz;
    G.kcx = G.kx * a;
    G.kcy = G.ky * a;
    G.kcz = G.kz * a;
}


// equal bit patterns (a phase groups voxels whose moduli are the same doubles)
static bool same_bits(double a, double b) {
    uint64_t x, y;
    std::memcpy(&x, &a, sizeof x);
    std::memcpy(&y, &b, sizeof y);
    return x == y;
}

extern "C" int phast_hex8_plan_create(int32_t nx, int32_t ny, int32_t nz, double hx, double hy, double hz,
                                      const double *bank, int32_t n_bank, const int32_t *bank_index,
                                      const double *Gc, const double *lam, const double *mu, const int8_t *driver,
                                      double kappa, double ell, phast_hex8_plan **out) {
    PHAST_REQUIRE(out && bank && bank_index && Gc && lam && mu && driver, "phast_hex8_plan_create: null argument");
    PHAST_REQUIRE(nx >= 2 && ny >= 2 && nz >= 2, "periodic voxel cell needs at least 2 voxels per axis");
    PHAST_REQUIRE(nz <= 65535, "periodic voxel cell: at most 65535 voxel layers along z (one grid row per layer)");
    PHAST_REQUIRE(24 * (int64_t)nx * ny < ((int64_t)1 << 32),
                  "periodic voxel cell: 24 nx ny bytes must fit 32-bit offsets within a node layer");
    PHAST_REQUIRE(hx > 0 && hy > 0 && hz > 0 && ell > 0 && n_bank >= 1, "bad geometry / bank size");
    auto *pl = new phast_hex8_plan();
    std::memset(&pl->geo, 0, sizeof(pl->geo));
    pl->geo.nx = nx;
    pl->geo.ny = ny;
    pl->geo.nz = nz;
    pl->geo.kappa = kappa;
    pl->geo.omk = 1.0 - kappa;
    pl->geo.ch = 2.0 * (1.0 - kappa);
    build_voxel_constants(pl->geo, hx, hy, hz, ell);
    pl->h[0] = hx;
    pl->h[1] = hy;
    pl->h[2] = hz;
    pl->ell = ell;
    pl->n = (int64_t)nx * ny * nz;
    pl->n_bank = n_bank;
    for (int64_t e = 0; e < pl->n; ++e) {
        if (bank_index[e] < 0 || bank_index[e] >= n_bank) {
            delete pl;
            set_error("bank_index[%lld] out of range", (long long)e);
            return PHAST_ERR_ARG;
        }
    }
    // isotropic bank: every entry has the (lam, mu) structure of core/hex8.py:296-320 and the
    // per-voxel lam/mu arrays reproduce it exactly
    bool iso = true;
    for (int64_t e = 0; e < pl->n && iso; ++e) {
        const double *C = bank + 36 * (int64_t)bank_index[e];
        const double l = lam[e], m = mu[e];
        for (int r = 0; r < 6 && iso; ++r)
            for (int c = 0; c < 6 && iso; ++c) {
                double ref = 0.0;
                if (r < 3 && c < 3) ref = l + (r == c ? 2.0 * m : 0.0);
                else if (r == c) ref = m;
                if (C[r * 6 + c] != ref) iso = false;
            }
    }
    pl->iso = iso;
    pl->max_blocks = (int32_t)(ceil_div(nx, 8) * ceil_div(ny, 6) * nz);  // the narrowest tile: 8 x 6
    const size_t nb = sizeof(double) * (size_t)pl->n;
    if (hipMalloc(&pl->bank, sizeof(double) * 36 * n_bank) != hipSuccess ||
        hipMalloc(&pl->bank_index, sizeof(int32_t) * pl->n) != hipSuccess || hipMalloc(&pl->Gc, nb) != hipSuccess ||
        hipMalloc(&pl->lam, nb) != hipSuccess || hipMalloc(&pl->mu, nb) != hipSuccess ||
        hipMalloc(&pl->driver, (size_t)pl->n) != hipSuccess ||
        hipMalloc(&pl->part, sizeof(double) * 8 * (size_t)std::max<int64_t>(pl->max_blocks, (pl->n + 255) / 256)) !=
            hipSuccess ||
        hipMemcpy(pl->bank, bank, sizeof(double) * 36 * n_bank, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pl->bank_index, bank_index, sizeof(int32_t) * pl->n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pl->Gc, Gc, nb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pl->lam, lam, nb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pl->mu, mu, nb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pl->driver, driver, (size_t)pl->n, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("phast_hex8_plan_create: device allocation/copy failed");
        phast_hex8_plan_destroy(pl);
        return PHAST_ERR_HIP;
    }
    if (int rc = hex8_ensure_seam(pl, pl->kz)) {
        phast_hex8_plan_destroy(pl);
        return rc;
    }
    if (iso) {  // few distinct (lam, mu) pairs (bitwise): a phase index per voxel and a table
        std::vector<uint8_t> ph((size_t)pl->n);
        std::vector<double> tab;
        int last = -1;
        for (int64_t e = 0; e < pl->n; ++e) {
            const double l = lam[e], m = mu[e];
            int p = -1;
            if (last >= 0 && same_bits(tab[2 * last], l) && same_bits(tab[2 * last + 1], m)) p = last;
            for (int q = 0; p < 0 && q < (int)tab.size() / 2; ++q)
                if (same_bits(tab[2 * q], l) && same_bits(tab[2 * q + 1], m)) p = q;
            if (p < 0) {
                if ((int)tab.size() / 2 == kHex8MaxPhases) {
                    tab.clear();
                    break;
                }
                p = (int)tab.size() / 2;
                tab.push_back(l);
                tab.push_back(m);
            }
            ph[(size_t)e] = (uint8_t)p;
            last = p;
        }
        if (!tab.empty()) {
            if (hipMalloc(&pl->phase, (size_t)pl->n) != hipSuccess ||
                hipMalloc(&pl->ptab, sizeof(double) * tab.size()) != hipSuccess ||
                hipMemcpy(pl->phase, ph.data(), (size_t)pl->n, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(pl->ptab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice) != hipSuccess) {
                set_error("phast_hex8_plan_create: device allocation/copy failed");
                phast_hex8_plan_destroy(pl);
                return PHAST_ERR_HIP;
            }
            pl->n_phase = (int32_t)(tab.size() / 2);
        }
    }
    *out = pl;
    return PHAST_OK;
}

extern "C" int phast_hex8_plan_destroy(phast_hex8_plan *pl) {
    if (!pl) return PHAST_OK;
    if (pl->owns_material) {
        (void)hipFree(pl->bank);
        (void)hipFree(pl->bank_index);
        (void)hipFree(pl->Gc);
        (void)hipFree(pl->driver);
    }
    (void)hipFree(pl->lam);
    (void)hipFree(pl->mu);
    (void)hipFree(pl->phase);
    (void)hipFree(pl->ptab);
    (void)hipFree(pl->part);
    (void)hipFree(pl->seam);
    (void)hipFree(pl->a1);
    (void)hipFree(pl->cl_on);
    (void)hipFree(pl->cl_tab);
    release_solver_work(pl);
    delete pl;
    return PHAST_OK;
}

// Seam buffer of the voxel-parallel apply for slabs of kz node layers: 3 doubles per node
// of each slab's first layer.  Grown here (set-up time), never inside an apply.
int phast::hex8_ensure_seam(phast_hex8_plan *pl, int32_t kz) {
    const int64_t need = 3 * (int64_t)pl->geo.nx * pl->geo.ny * ceil_div(pl->geo.nz, kz);
    if (pl->seam && pl->seam_cap >= need) return PHAST_OK;
    (void)hipFree(pl->seam);
    pl->seam = nullptr;
    pl->seam_cap = 0;
    PHAST_HIP(hipMalloc(&pl->seam, sizeof(double) * (size_t)need));
    pl->seam_cap = need;
    return PHAST_OK;
}

extern "C" int phast_hex8_set_tiles(phast_hex8_plan *pl, int32_t tx, int32_t ty, int32_t kz) {
    PHAST_REQUIRE(pl && kz >= 1, "phast_hex8_set_tiles: bad argument");
    PHAST_REQUIRE((tx == 16 && ty == 16) || (tx == 16 && ty == 8) || (tx == 8 && ty == 8) || (tx == 32 && ty == 8) ||
                      (tx == 16 && ty == 14) || (tx == 32 && ty == 14) || (tx == 32 && ty == 6) ||
                      (tx == 64 && ty == 6),
                  "unsupported hex8 tile %dx%d", tx, ty);
    if (int rc = hex8_ensure_seam(pl, kz)) return rc;
    pl->tx = tx;
    pl->ty = ty;
    pl->kz = kz;
    return PHAST_OK;
}

extern "C" int phast_hex8_degradation(const phast_hex8_plan *pl, const double *d, double *g, hipStream_t st) {
    PHAST_REQUIRE(pl && d && g, "phast_hex8_degradation: null argument");
    hipLaunchKernelGGL(hex8_degradation_kernel, dim3((unsigned)ceil_div(pl->n, 256)), dim3(256), 0, st, pl->geo, d, g);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

extern "C" int phast_hex8_mech_action(const phast_hex8_plan *pl, const double *x, const double *g, double *y,
                                      hipStream_t st) {
    PHAST_REQUIRE(pl && x && y && x != y, "phast_hex8_mech_action: bad argument");
    return hex8_apply_impl(pl, x, g, nullptr, y, nullptr, false, st);
}

extern "C" int phast_hex8_mech_action_d(const phast_hex8_plan *pl, const double *x, const double *d, double *y,
                                        hipStream_t st) {
    PHAST_REQUIRE(pl && x && d && y && x != y, "phast_hex8_mech_action_d: bad argument");
    return hex8_apply_impl(pl, x, nullptr, d, y, nullptr, false, st);
}

extern "C" int phast_hex8_stress_sums(const phast_hex8_plan *pl, const double *x, const double *g, double *out7,
                                      hipStream_t st) {
    PHAST_REQUIRE(pl && x && out7, "phast_hex8_stress_sums: null argument");
    // the nodal rows are computed and discarded into the solver scratch of the plan
    double *scratch = nullptr;
    PHAST_HIP(hipMallocAsync((void **)&scratch, sizeof(double) * (3 * pl->n + 6), st));
    const int rc = hex8_apply_impl(pl, x, g, nullptr, scratch, out7, true, st);
    PHAST_HIP(hipFreeAsync(scratch, st));
    return rc;
}

extern "C" int phast_hex8_mech_diagonal(const phast_hex8_plan *pl, const double *g, double *diag, hipStream_t st) {
    PHAST_REQUIRE(pl && diag, "phast_hex8_mech_diagonal: null argument");
    const int nb = (int)ceil_div(pl->n, 256);
    hipLaunchKernelGGL(hex8_diag_kernel, dim3(nb), dim3(256), 0, st, pl->geo, g, pl->bank, pl->bank_index, pl->lam,
                       pl->mu, pl->iso ? 1 : 0, diag, pl->part, (const uint8_t *)nullptr, (const uint32_t *)nullptr, 0);
    PHAST_LAUNCHED();
    hipLaunchKernelGGL(hex8_finalize_kernel, dim3(1), dim3(256), 0, st, pl->part, nb, 6, diag + 3 * pl->n,
                       (double *)nullptr);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

// The Jacobi diagonal with the crack closure: every closure bit (mode 1) or those of `pat`
// (mode 2) adds its (1 - g) term (operator.py:1196-1300).
extern "C" int phast_hex8_mech_diagonal_cl(const phast_hex8_plan *pl, const double *g, const uint32_t *pat,
                                           int32_t mode, double *diag, hipStream_t st) {
    PHAST_REQUIRE(pl && g && diag && (mode == 1 || (mode == 2 && pat)), "phast_hex8_mech_diagonal_cl: bad argument");
    PHAST_REQUIRE(pl->cl_on && (pl->iso || pl->cl_tab), "phast_hex8_mech_diagonal_cl: the plan has no crack closure");
    const int nb = (int)ceil_div(pl->n, 256);
    hipLaunchKernelGGL(hex8_diag_kernel, dim3(nb), dim3(256), 0, st, pl->geo, g, pl->bank, pl->bank_index, pl->lam,
                       pl->mu, pl->iso ? 1 : 0, diag, pl->part, (const uint8_t *)pl->cl_on, pat, (int)mode,
                       pl->iso ? (const double *)nullptr : (const double *)pl->cl_tab);
    PHAST_LAUNCHED();
    hipLaunchKernelGGL(hex8_finalize_kernel, dim3(1), dim3(256), 0, st, pl->part, nb, 6, diag + 3 * pl->n,
                       (double *)nullptr);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

extern "C" int phast_hex8_driving_force(const phast_hex8_plan *pl, const double *x, double *Y, hipStream_t st) {
    PHAST_REQUIRE(pl && x && Y, "phast_hex8_driving_force: null argument");
    hipLaunchKernelGGL(hex8_driving_kernel, dim3((unsigned)ceil_div(pl->n, 256)), dim3(256), 0, st, pl->geo, x,
                       pl->bank, pl->bank_index, pl->lam, pl->mu, pl->driver, Y);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

// y = scale (Gc_e K d_e + ch sum_q w H_q N_q N_q^T d_e) with the constants of G: the AT2
// damage operator with the plan's own constants, the PF-CZM Newton tangent with geo_t.
int phast::hex8_damage_launch(const phast_hex8_plan *pl, const Hex8Geo &G, const double *d, const double *H,
                              double scale, double *y, hipStream_t st) {
    DmgArgs3 A;
    std::memset(&A, 0, sizeof(A));
    A.d = d;
    A.H = H;
    A.Gc = pl->Gc;
    A.y = y;
    A.scale = scale;
    A.kz = pl->kz;
    if (pl->tx == 16 && pl->ty == 16) return launch_damage_vp_t<16, 16>(pl, G, A, st);
    if (pl->tx == 16 && pl->ty == 8) return launch_damage_t<16, 8>(pl, G, A, st);
    if (pl->tx == 8 && pl->ty == 8) return launch_damage_t<8, 8>(pl, G, A, st);
    if (pl->tx == 32 && pl->ty == 8) return launch_damage_vp_t<32, 8>(pl, G, A, st);
    if (pl->tx == 16 && pl->ty == 14) return launch_damage_vp_t<16, 14>(pl, G, A, st);
    if (pl->tx == 32 && pl->ty == 14) return launch_damage_vp_t<32, 14>(pl, G, A, st);
    if (pl->tx == 32 && pl->ty == 6) return launch_damage_vp_t<32, 6>(pl, G, A, st);
    if (pl->tx == 64 && pl->ty == 6) return launch_damage_vp_t<64, 6>(pl, G, A, st);
    set_error("unsupported hex8 tile");
    return PHAST_ERR_UNSUPPORTED;
}

int phast::hex8_damage_rd_launch(const phast_hex8_plan *pl, const Hex8Geo &G, const double *H, double scale,
                                 double *b, double *diag, hipStream_t st) {
    hipLaunchKernelGGL(hex8_damage_rd_kernel, dim3((unsigned)ceil_div(pl->n, 256)), dim3(256), 0, st, G, H, pl->Gc,
                       scale, b, diag);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

extern "C" int phast_hex8_damage_action(const phast_hex8_plan *pl, const double *d, const double *H, double scale,
                                        double *y, hipStream_t st) {
    PHAST_REQUIRE(pl && d && y && d != y, "phast_hex8_damage_action: bad argument");
    return hex8_damage_launch(pl, pl->geo, d, H, scale, y, st);
}

extern "C" int phast_hex8_damage_rhs(const phast_hex8_plan *pl, const double *H, double scale, double *b,
                                     hipStream_t st) {
    PHAST_REQUIRE(pl && b, "phast_hex8_damage_rhs: null argument");
    return hex8_damage_rd_launch(pl, pl->geo, H, scale, b, nullptr, st);
}

extern "C" int phast_hex8_damage_diagonal(const phast_hex8_plan *pl, const double *H, double scale, double *diag,
                                          hipStream_t st) {
    PHAST_REQUIRE(pl && diag, "phast_hex8_damage_diagonal: null argument");
    return hex8_damage_rd_launch(pl, pl->geo, H, scale, nullptr, diag, st);
}

int phast::hex8_apply(const phast_hex8_plan *pl, const double *x, const double *g, const double *d, double *y,
                      double *out7, bool energy, hipStream_t st) {
    return hex8_apply_impl(pl, x, g, d, y, out7, energy, st);
}
