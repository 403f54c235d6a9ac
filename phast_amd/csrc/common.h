// Shared helpers of the gfx950 kernels: error reporting, wave64 reductions, launch checks.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <cmath>

#include "../../include/phast_mi.h"

namespace phast {

void set_error(const char *fmt, ...);

#define PHAST_HIP(expr)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            ::phast::set_error("%s:%d: %s -> %s", __FILE__, __LINE__, #expr,             \
                               hipGetErrorString(e_));                                   \
            return PHAST_ERR_HIP;                                                        \
        }                                                                                \
    } while (0)

#define PHAST_REQUIRE(cond, ...)                                                         \
    do {                                                                                 \
        if (!(cond)) {                                                                   \
            ::phast::set_error(__VA_ARGS__);                                             \
            return PHAST_ERR_ARG;                                                        \
        }                                                                                \
    } while (0)

// Launch error check after hipLaunchKernelGGL / <<<>>>.
#define PHAST_LAUNCHED() PHAST_HIP(hipGetLastError())

#define PHAST_TRY0(x)                  \
    do {                               \
        int rc0_ = (x);                \
        if (rc0_ != PHAST_OK) return rc0_; \
    } while (0)

constexpr int kWave = 64;

// Deterministic wave64 butterfly reductions (fixed pairing order on every run).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}

// Block-wide reductions for NT threads (NT a multiple of 64).  `sh` holds NT/64 doubles.
// The result is valid in every thread.  Caller must not reuse `sh` before a barrier.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double *sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NT / kWave; ++i) s += sh[i];
    return s;
}
// K block sums at once, each with exactly block_sum's arithmetic (same bits), sharing one
// pair of barriers; the K butterflies interleave so their shuffle latencies overlap.
// `sh` holds K * NT/64 doubles.  v[k] is replaced by its block sum in every thread.
template <int NT, int K>
__device__ __forceinline__ void block_sum_n(double v[K], double *sh) {
    constexpr int NW = NT / kWave;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] += __shfl_xor(v[k], o, kWave);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) sh[k * NW + w] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < NW; ++i) s += sh[k * NW + i];
        v[k] = s;
    }
}
// Per-thread part of a fixed-order sum of block partials: s[k] += part[8 b + k] for
// b = threadIdx.x, threadIdx.x + NT, ... < nblocks and k < nvals, in that order.  Four rows
// of loads are in flight before their adds (lanes past the end load the last row and skip
// its adds), so the sum waits for one memory latency per four rows instead of one per row.
template <int NT, int K>
__device__ __forceinline__ void sum_block_partials(const double *part, int nblocks, int nvals, double s[K]) {
    for (int b0 = threadIdx.x; b0 < nblocks; b0 += 4 * NT) {
        double v[4][K];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int b = min(b0 + u * NT, nblocks - 1);
#pragma unroll
            for (int k = 0; k < K; ++k) v[u][k] = part[8 * b + k];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (b0 + u * NT < nblocks)
#pragma unroll
                for (int k = 0; k < K; ++k)
                    if (k < nvals) s[k] += v[u][k];
    }
}
// Tile worked by block b of an nblocks-block launch, in an XCD-aware order.  The dispatcher
// hands consecutive workgroups to the 8 XCDs in turn (block b on XCD b % 8), each XCD with
// its own L2.  Block b takes tile (b % 8) q + b / 8 with q = nblocks / 8, so each XCD works
// through one contiguous run of tiles and the halo lines neighbouring tiles share are
// fetched into one L2; the last nblocks % 8 blocks keep their own index.  The map is a
// bijection whatever the dispatch order: that order only sets the speed.
__device__ __forceinline__ int xcd_tile(int b, int nblocks) {
    const int q = nblocks >> 3;
    return b < (q << 3) ? (b & 7) * q + (b >> 3) : b;
}
// One block sum (block_sum's arithmetic, same bits) and two block minima sharing one pair of
// barriers.  `sh` holds 3 * NT/64 doubles; results are valid in every thread.
template <int NT>
__device__ __forceinline__ void block_sum_min2(double &s, double &m1, double &m2, double *sh) {
    constexpr int NW = NT / kWave;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o, kWave);
        m1 = fmin(m1, __shfl_xor(m1, o, kWave));
        m2 = fmin(m2, __shfl_xor(m2, o, kWave));
    }
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) {
        sh[w] = s;
        sh[NW + w] = m1;
        sh[2 * NW + w] = m2;
    }
    __syncthreads();
    double a = 0.0, b = sh[NW], c = sh[2 * NW];
#pragma unroll
    for (int i = 0; i < NW; ++i) a += sh[i];
#pragma unroll
    for (int i = 1; i < NW; ++i) {
        b = fmin(b, sh[NW + i]);
        c = fmin(c, sh[2 * NW + i]);
    }
    s = a;
    m1 = b;
    m2 = c;
}
template <int NT>
__device__ __forceinline__ double block_min(double v, double *sh) {
    v = wave_min(v);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = sh[0];
#pragma unroll
    for (int i = 1; i < NT / kWave; ++i) s = fmin(s, sh[i]);
    return s;
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace phast
