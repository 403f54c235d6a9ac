// Geometric multigrid preconditioner for the periodic voxel mechanics operator (FP64).
//
// Hierarchy: the voxel cell is coarsened 2:1 per axis while every dimension stays even
// and >= min_coarse.  Coarse operators are rediscretisations with voxel-averaged
// isotropic moduli: lam_c = mean over the 8 children (and their 8 Gauss points) of
// g * lam, likewise mu; for constant coefficients this equals the Galerkin product
// P^T A P because trilinear interpolation is exact on the coarse space.  P is periodic
// trilinear interpolation of the nodal displacement with the 6 macro-strain dofs
// carried over unchanged; R = P^T.  The smoother is damped Jacobi (nu sweeps before and
// after, same weight), the coarsest level takes a fixed number of Jacobi sweeps, so the
// V-cycle is a fixed symmetric positive operator and plain PCG applies.  Every level
// pins the node above the fine pinned node and the same macro components.
#include "hex8_device.h"
#include "hex8_internal.h"

#include <cstring>
#include <vector>

namespace phast {

constexpr int kMaxLevels = 8;

// nodal periodic index helpers
__device__ __forceinline__ int64_t nid(int i, int j, int k, int nx, int ny) {
    return i + (int64_t)nx * (j + (int64_t)ny * k);
}

// fine -> coarse full weighting (P^T), 3 components per node; macro rows copied.
__global__ __launch_bounds__(256) void k_restrict(int nxc, int nyc, int nzc, const double *rf, double *bc,
                                                  const uint8_t *fixed_c) {
    const int64_t nc = (int64_t)nxc * nyc * nzc;
    const int64_t I = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int nxf = 2 * nxc, nyf = 2 * nyc, nzf = 2 * nzc;
    const int64_t nf = (int64_t)nxf * nyf * nzf;
    if (I < 6) bc[3 * nc + I] = fixed_c[3 * nc + I] ? 0.0 : rf[3 * nf + I];
    if (I >= nc) return;
    const int i = (int)(I % nxc), j = (int)((I / nxc) % nyc), k = (int)(I / ((int64_t)nxc * nyc));
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int c = -1; c <= 1; ++c) {
        const double wc = c ? 0.5 : 1.0;
        const int fk = (2 * k + c + nzf) % nzf;
        for (int b = -1; b <= 1; ++b) {
            const double wb = wc * (b ? 0.5 : 1.0);
            const int fj = (2 * j + b + nyf) % nyf;
            for (int a = -1; a <= 1; ++a) {
                const double w = wb * (a ? 0.5 : 1.0);
                const int fi = (2 * i + a + nxf) % nxf;
                const int64_t f = nid(fi, fj, fk, nxf, nyf);
                s0 += w * rf[3 * f];
                s1 += w * rf[3 * f + 1];
                s2 += w * rf[3 * f + 2];
            }
        }
    }
    const bool fx0 = fixed_c[3 * I], fx1 = fixed_c[3 * I + 1], fx2 = fixed_c[3 * I + 2];
    bc[3 * I] = fx0 ? 0.0 : s0;
    bc[3 * I + 1] = fx1 ? 0.0 : s1;
    bc[3 * I + 2] = fx2 ? 0.0 : s2;
}

// xf += P xc (masked on the fine fixed set)
__global__ __launch_bounds__(256) void k_prolong_add(int nxc, int nyc, int nzc, const double *xc, double *xf,
                                                     const uint8_t *fixed_f) {
    const int nxf = 2 * nxc, nyf = 2 * nyc, nzf = 2 * nzc;
    const int64_t nf = (int64_t)nxf * nyf * nzf, nc = (int64_t)nxc * nyc * nzc;
    const int64_t F = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (F < 6 && !fixed_f[3 * nf + F]) xf[3 * nf + F] += xc[3 * nc + F];
    if (F >= nf) return;
    const int fi = (int)(F % nxf), fj = (int)((F / nxf) % nyf), fk = (int)(F / ((int64_t)nxf * nyf));
    int ci[2], cj[2], ck[2];
    double wi[2], wj[2], wk[2];
    int ni = 1, nj = 1, nk = 1;
    ci[0] = fi / 2; wi[0] = 1.0;
    if (fi & 1) { ni = 2; wi[0] = wi[1] = 0.5; ci[1] = (fi / 2 + 1) % nxc; }
    cj[0] = fj / 2; wj[0] = 1.0;
    if (fj & 1) { nj = 2; wj[0] = wj[1] = 0.5; cj[1] = (fj / 2 + 1) % nyc; }
    ck[0] = fk / 2; wk[0] = 1.0;
    if (fk & 1) { nk = 2; wk[0] = wk[1] = 0.5; ck[1] = (fk / 2 + 1) % nzc; }
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int c = 0; c < nk; ++c)
        for (int b = 0; b < nj; ++b)
            for (int a = 0; a < ni; ++a) {
                const double w = wi[a] * wj[b] * wk[c];
                const int64_t C = nid(ci[a], cj[b], ck[c], nxc, nyc);
                s0 += w * xc[3 * C];
                s1 += w * xc[3 * C + 1];
                s2 += w * xc[3 * C + 2];
            }
    if (!fixed_f[3 * F]) xf[3 * F] += s0;
    if (!fixed_f[3 * F + 1]) xf[3 * F + 1] += s1;
    if (!fixed_f[3 * F + 2]) xf[3 * F + 2] += s2;
}

// coarse moduli: mean over the 8 children of (mean_q g) * lam, (mean_q g) * mu; g (E,8) is
// the fine degradation at level 1 and 1 (null) deeper down.
__global__ __launch_bounds__(256) void k_coarsen_moduli(const Hex8Geo Gf, int nxc, int nyc, int nzc, const double *g,
                                                        const double *lamf, const double *muf, double *lamc,
                                                        double *muc) {
    const int64_t nc = (int64_t)nxc * nyc * nzc;
    const int64_t I = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (I >= nc) return;
    const int i = (int)(I % nxc), j = (int)((I / nxc) % nyc), k = (int)(I / ((int64_t)nxc * nyc));
    const int nxf = Gf.nx, nyf = Gf.ny, nzf = Gf.nz;
    double sl = 0.0, sm = 0.0;
    for (int ch = 0; ch < 8; ++ch) {
        const int fi = 2 * i + (ch & 1), fj = 2 * j + ((ch >> 1) & 1), fk = 2 * k + (ch >> 2);
        const int64_t e = nid(fi, fj, fk, nxf, nyf);
        double gm = 1.0;
        if (g) {
            gm = 0.0;
            for (int q = 0; q < 8; ++q) gm += g[e * 8 + q];
            gm *= 0.125;
        }
        sl += gm * lamf[e];
        sm += gm * muf[e];
    }
    lamc[I] = 0.125 * sl;
    muc[I] = 0.125 * sm;
}

// x += omega (b - Ax) / diag on free dofs; fixed dofs stay 0.
__global__ __launch_bounds__(256) void k_jacobi(int64_t n, double *x, const double *b, const double *Ax,
                                                const double *diag, const uint8_t *fixed, double omega, int first) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    if (fixed[i]) {
        x[i] = 0.0;
        return;
    }
    const double r = b[i] - (first ? 0.0 : Ax[i]);
    x[i] = (first ? 0.0 : x[i]) + omega * r / diag[i];
}

// Every damped-Jacobi sweep of a small coarsest level in one block (one thread per voxel,
// at most kCoarseMax voxels): x, the per-voxel corner forces and the macro rows of A x live
// in LDS, so a sweep costs two barriers instead of three kernel launches.  Per sweep the
// same arithmetic as hex8_apply (g = 1, isotropic moduli) followed by k_jacobi; the nodal
// gather visits a node's 8 voxels in a fixed order, and the macro rows are block sums.
constexpr int kCoarseMax = 512;
__global__ __launch_bounds__(kCoarseMax) void k_coarse_sweeps(const Hex8Geo G, const double *lam, const double *mu,
                                                              const double *b, const double *diag,
                                                              const uint8_t *fixed, double omega, int sweeps,
                                                              double *x_out) {
    constexpr int NT = kCoarseMax;
    __shared__ double sx[3 * NT + 6];
    __shared__ double sr[24][NT];
    __shared__ double red[6 * (NT / 64)];
    const int nx = G.nx, ny = G.ny, nz = G.nz, n = nx * ny * nz;
    const int t = threadIdx.x;
    const bool mine = t < n;
    const int i = mine ? t % nx : 0, j = mine ? (t / nx) % ny : 0, k = mine ? t / (nx * ny) : 0;
    // this thread's dofs: node t (3 components) and, for t < 6, macro dof 3n + t
    double bb[4] = {0.0, 0.0, 0.0, 0.0}, dg[4] = {1.0, 1.0, 1.0, 1.0};
    bool fx[4] = {true, true, true, true};
    if (mine)
        for (int c = 0; c < 3; ++c) {
            bb[c] = b[3 * t + c];
            dg[c] = diag[3 * t + c];
            fx[c] = fixed[3 * t + c] != 0;
        }
    if (t < 6) {
        bb[3] = b[3 * n + t];
        dg[3] = diag[3 * n + t];
        fx[3] = fixed[3 * n + t] != 0;
    }
    const double l0 = mine ? lam[t] : 0.0, m0 = mine ? mu[t] : 0.0;
    // first sweep: x = omega D^-1 b on free dofs, 0 on fixed
    if (mine)
        for (int c = 0; c < 3; ++c) sx[3 * t + c] = fx[c] ? 0.0 : omega * bb[c] / dg[c];
    if (t < 6) sx[3 * n + t] = fx[3] ? 0.0 : omega * bb[3] / dg[3];
    __syncthreads();
    const double g1[8] = {1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0};
    for (int s = 1; s < sweeps; ++s) {
        double sE[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, en = 0.0;
        if (mine) {
            double u[3][8], E[6], r[3][8];
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                const int node = wrap(i + (a & 1), nx) + nx * (wrap(j + ((a >> 1) & 1), ny) + ny * wrap(k + (a >> 2), nz));
#pragma unroll
                for (int c = 0; c < 3; ++c) u[c][a] = sx[3 * node + c];
            }
#pragma unroll
            for (int q = 0; q < 6; ++q) E[q] = sx[3 * n + q];
            voxel_mech<true, false>(G, u, E, g1, nullptr, l0, m0, true, r, sE, en);
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int a = 0; a < 8; ++a) sr[c * 8 + a][t] = r[c][a];
        }
        block_sum_n<NT, 6>(sE, red);  // also the barrier after the corner forces
        double ax[3] = {0.0, 0.0, 0.0};
        if (mine)
#pragma unroll
            for (int a = 0; a < 8; ++a) {  // node t is corner a of voxel (i-lx, j-ly, k-lz)
                const int v = wrap(i - (a & 1), nx) + nx * (wrap(j - ((a >> 1) & 1), ny) + ny * wrap(k - (a >> 2), nz));
#pragma unroll
                for (int c = 0; c < 3; ++c) ax[c] += sr[c * 8 + a][v];
            }
        __syncthreads();  // every thread has read x for this sweep's apply
        if (mine)
            for (int c = 0; c < 3; ++c)
                if (!fx[c]) sx[3 * t + c] += omega * (bb[c] - ax[c]) / dg[c];
        if (t < 6 && !fx[3]) sx[3 * n + t] += omega * (bb[3] - sE[t]) / dg[3];
        __syncthreads();
    }
    for (int q = t; q < 3 * n + 6; q += NT) x_out[q] = sx[q];
}

// r = b - Ax on free dofs, 0 on fixed.
__global__ __launch_bounds__(256) void k_residual(int64_t n, const double *b, const double *Ax, const uint8_t *fixed,
                                                  double *r) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    r[i] = fixed[i] ? 0.0 : b[i] - Ax[i];
}

// w = D^-1 A v on free dofs (0 on fixed); partials |w|^2, |v|^2 per block
__global__ __launch_bounds__(256) void k_power_step(int64_t n, const double *Av, const double *diag, const uint8_t *fixed,
                                                    const double *v, double *w, double *part) {
    __shared__ double red[4];
    double sw = 0.0, sv = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double wi = fixed[i] ? 0.0 : Av[i] / diag[i];
        w[i] = wi;
        sw += wi * wi;
        sv += v[i] * v[i];
    }
    sw = block_sum<256>(sw, red);
    sv = block_sum<256>(sv, red);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sw;
        part[2 * blockIdx.x + 1] = sv;
    }
}

__global__ void k_scale_copy(int64_t n, const double *w, double s, double *v) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) v[i] = s * w[i];
}

// One power-iteration normalisation on the device: |w|^2 and |v|^2 summed over the gb
// block partials in block order by one thread; out[0] = 1/|w| (the next v = out[0] w),
// out[1] = |w|/|v| (the estimate); out[2] counts breakdowns (a sum that is not positive or
// not finite; the next v is then zero and every later step breaks down too), restarted at
// the first step of a level.
__global__ void k_power_norm(unsigned gb, const double *part, int first, double *out) {
    if (threadIdx.x != 0) return;
    double sw = 0.0, sv = 0.0;
    for (unsigned k = 0; k < gb; ++k) {
        sw += part[2 * k];
        sv += part[2 * k + 1];
    }
    const bool bad = !(sw > 0.0) || !(sv > 0.0) || !isfinite(sw);
    out[0] = bad ? 0.0 : 1.0 / sqrt(sw);
    out[1] = bad ? 0.0 : sqrt(sw / sv);
    out[2] = (first ? 0.0 : out[2]) + (bad ? 1.0 : 0.0);
}

__global__ void k_scale_copy_dev(int64_t n, const double *w, const double *s, double *v) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) v[i] = s[0] * w[i];
}

// deterministic start vector for the power iteration (zero on the fixed set)
__global__ void k_power_start(int64_t n, const uint8_t *fixed, double *v) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const uint64_t h = (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull;
        v[i] = fixed[i] ? 0.0 : ((double)(h >> 11) * (1.0 / 9007199254740992.0) - 0.5);
    }
}

__global__ void k_set_fixed(int64_t n_nodes, int64_t pinned, const uint8_t *macro_fixed6, uint8_t *fixed) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < 3 * n_nodes + 6) {
        uint8_t v = 0;
        if (i >= 3 * pinned && i < 3 * pinned + 3) v = 1;
        if (i >= 3 * n_nodes) v = macro_fixed6[i - 3 * n_nodes];
        fixed[i] = v;
    }
}

}  // namespace phast

using namespace phast;

struct phast_hex8_mg {
    int levels = 0;
    const phast_hex8_plan *fine = nullptr;
    phast_hex8_plan *lv[kMaxLevels] = {};   // lv[0] = fine (not owned)
    int64_t n[kMaxLevels] = {};            // nodes per level
    double *diag[kMaxLevels] = {}, *x[kMaxLevels] = {}, *b[kMaxLevels] = {}, *Ax[kMaxLevels] = {},
           *r[kMaxLevels] = {};
    uint8_t *fixed[kMaxLevels] = {};
    uint8_t *macro6 = nullptr;
    const double *g = nullptr, *d = nullptr;  // current fine degradation source
    int nu = 2, coarse_sweeps = 30;
    double omega = 1.2;                  // weight factor: omega_l = omega / lambda_max(D^-1 A)_l
    double omega_l[kMaxLevels] = {};
    double lam_max[kMaxLevels] = {};
    double *hpart = nullptr;             // pinned host partials
    double *pw = nullptr;                // device power-iteration state: 3 doubles per level
};

static inline unsigned nb256(int64_t n) { return (unsigned)ceil_div(std::max<int64_t>(n, 1), 256); }

extern "C" int phast_hex8_mg_destroy(phast_hex8_mg *mg) {
    if (!mg) return PHAST_OK;
    for (int l = 0; l < mg->levels; ++l) {
        if (l > 0 && mg->lv[l]) {
            (void)hipFree(mg->lv[l]->lam);
            (void)hipFree(mg->lv[l]->mu);
            (void)hipFree(mg->lv[l]->part);
            (void)hipFree(mg->lv[l]->seam);
            (void)hipFree(mg->lv[l]->bank);
            (void)hipFree(mg->lv[l]->bank_index);
            delete mg->lv[l];
        }
        (void)hipFree(mg->diag[l]);
        (void)hipFree(mg->x[l]);
        (void)hipFree(mg->b[l]);
        (void)hipFree(mg->Ax[l]);
        (void)hipFree(mg->r[l]);
        (void)hipFree(mg->fixed[l]);
    }
    (void)hipFree(mg->macro6);
    (void)hipFree(mg->pw);
    if (mg->hpart) (void)hipHostFree(mg->hpart);
    delete mg;
    return PHAST_OK;
}

// Build the hierarchy under `fine` (which must be isotropic).  pinned: fine pinned node;
// macro_fixed6: host 6 flags of the fixed macro components.
extern "C" int phast_hex8_mg_create(const phast_hex8_plan *fine, int64_t pinned, const uint8_t *macro_fixed6,
                                    int32_t max_levels, int32_t min_coarse, int32_t nu, int32_t coarse_sweeps,
                                    double omega, phast_hex8_mg **out) {
    PHAST_REQUIRE(fine && macro_fixed6 && out, "phast_hex8_mg_create: null argument");
    PHAST_REQUIRE(fine->iso, "the multigrid needs isotropic voxel moduli");
    PHAST_REQUIRE(max_levels >= 1 && max_levels <= kMaxLevels && nu >= 1 && coarse_sweeps >= 1 && omega > 0,
                  "bad multigrid parameters");
    auto *mg = new phast_hex8_mg();
    mg->fine = fine;
    mg->nu = nu;
    mg->coarse_sweeps = coarse_sweeps;
    mg->omega = omega;
    int nx = fine->geo.nx, ny = fine->geo.ny, nz = fine->geo.nz;
    double hx = 1.0 / fine->geo.ihx, hy = 1.0 / fine->geo.ihy, hz = 1.0 / fine->geo.ihz;
    mg->lv[0] = const_cast<phast_hex8_plan *>(fine);
    mg->levels = 1;
    auto fail = [&](const char *what) {
        set_error("phast_hex8_mg_create: %s", what);
        phast_hex8_mg_destroy(mg);
        return PHAST_ERR_HIP;
    };
    if (hipHostMalloc(&mg->hpart, sizeof(double) * 2 * 512, hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&mg->pw, sizeof(double) * 3 * kMaxLevels) != hipSuccess || hipMalloc(&mg->macro6, 6) != hipSuccess ||
        hipMemcpy(mg->macro6, macro_fixed6, 6, hipMemcpyHostToDevice) != hipSuccess)
        return fail("alloc");
    int64_t pin = pinned;
    while (mg->levels < max_levels && nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0 && nx / 2 >= min_coarse &&
           ny / 2 >= min_coarse && nz / 2 >= min_coarse) {
        const int l = mg->levels;
        const int pnx = nx, pny = ny;
        nx /= 2; ny /= 2; nz /= 2;
        hx *= 2; hy *= 2; hz *= 2;
        auto *pl = new phast_hex8_plan();
        pl->geo = fine->geo;
        pl->geo.nx = nx; pl->geo.ny = ny; pl->geo.nz = nz;
        build_voxel_constants(pl->geo, hx, hy, hz, 1.0);
        pl->n = (int64_t)nx * ny * nz;
        pl->iso = true;
        pl->owns_material = false;
        pl->n_bank = 1;
        // voxel-parallel apply with slab seams; thin slabs so that small levels still
        // spread over many blocks
        pl->tx = 16; pl->ty = 14; pl->kz = nz >= 32 ? 3 : (nz >= 16 ? 2 : 1);
        pl->max_blocks = (int32_t)(ceil_div(nx, 8) * ceil_div(ny, 6) * nz);  // the narrowest tile: 8 x 6
        mg->lv[l] = pl;
        mg->levels = l + 1;
        const size_t nb = sizeof(double) * (size_t)pl->n;
        if (hipMalloc(&pl->lam, nb) != hipSuccess || hipMalloc(&pl->mu, nb) != hipSuccess ||
            hipMalloc(&pl->part, sizeof(double) * 8 * (size_t)std::max<int64_t>(pl->max_blocks, (pl->n + 255) / 256)) !=
                hipSuccess ||
            hipMalloc(&pl->bank, sizeof(double) * 36) != hipSuccess ||
            hipMemset(pl->bank, 0, sizeof(double) * 36) != hipSuccess ||
            hipMalloc(&pl->bank_index, sizeof(int32_t) * pl->n) != hipSuccess ||
            hipMemset(pl->bank_index, 0, sizeof(int32_t) * pl->n) != hipSuccess || hex8_ensure_seam(pl, pl->kz) != PHAST_OK)
            return fail("alloc");
        const int pi = (int)(pin % pnx), pj = (int)((pin / pnx) % pny), pk = (int)(pin / ((int64_t)pnx * pny));
        pin = (pi / 2) + (int64_t)nx * ((pj / 2) + (int64_t)ny * (pk / 2));
    }
    pin = pinned;
    for (int l = 0; l < mg->levels; ++l) {
        const int64_t nn = mg->lv[l]->n, nm = 3 * nn + 6;
        mg->n[l] = nn;
        const size_t b = sizeof(double) * (size_t)nm;
        if (hipMalloc(&mg->diag[l], b) != hipSuccess || hipMalloc(&mg->x[l], b) != hipSuccess ||
            hipMalloc(&mg->b[l], b) != hipSuccess || hipMalloc(&mg->Ax[l], b) != hipSuccess ||
            hipMalloc(&mg->r[l], b) != hipSuccess || hipMalloc(&mg->fixed[l], (size_t)nm) != hipSuccess)
            return fail("alloc");
        if (l > 0) {
            const phast_hex8_plan *up = mg->lv[l - 1];
            const int pi = (int)(pin % up->geo.nx), pj = (int)((pin / up->geo.nx) % up->geo.ny),
                      pk = (int)(pin / ((int64_t)up->geo.nx * up->geo.ny));
            pin = (pi / 2) + (int64_t)mg->lv[l]->geo.nx * ((pj / 2) + (int64_t)mg->lv[l]->geo.ny * (pk / 2));
        }
        hipLaunchKernelGGL(k_set_fixed, dim3(nb256(nm)), dim3(256), 0, 0, nn, pin, mg->macro6, mg->fixed[l]);
    }
    if (hipDeviceSynchronize() != hipSuccess) return fail("setup");
    *out = mg;
    return PHAST_OK;
}

// Recompute coarse moduli and all Jacobi diagonals for the fine degradation (g or d).
extern "C" int phast_hex8_mg_update(phast_hex8_mg *mg, const double *g, const double *d, hipStream_t st) {
    PHAST_REQUIRE(mg && !(g && d), "phast_hex8_mg_update: pass g or d (or neither)");
    mg->g = g;
    mg->d = d;
    const double *gq = g;
    double *gtmp = nullptr;
    if (d) {  // the fine diagonal needs g at the Gauss points
        PHAST_HIP(hipMallocAsync((void **)&gtmp, sizeof(double) * 8 * (size_t)mg->n[0], st));
        PHAST_TRY0(phast_hex8_degradation(mg->lv[0], d, gtmp, st));
        gq = gtmp;
    }
    int rc = phast_hex8_mech_diagonal(mg->lv[0], gq, mg->diag[0], st);
    for (int l = 1; l < mg->levels && rc == PHAST_OK; ++l) {
        const phast_hex8_plan *f = mg->lv[l - 1];
        phast_hex8_plan *c = mg->lv[l];
        hipLaunchKernelGGL(k_coarsen_moduli, dim3(nb256(c->n)), dim3(256), 0, st, f->geo, c->geo.nx, c->geo.ny,
                           c->geo.nz, l == 1 ? gq : (const double *)nullptr, f->lam, f->mu, c->lam, c->mu);
        rc = phast_hex8_mech_diagonal(c, nullptr, mg->diag[l], st);
    }
    if (gtmp) (void)hipFreeAsync(gtmp, st);
    // lambda_max(D^-1 A) per level by 12 power iterations -> Jacobi weights; the
    // normalisations run on the device and the estimates are read once at the end
    for (int l = 0; l < mg->levels && rc == PHAST_OK; ++l) {
        const int64_t nm = 3 * mg->n[l] + 6;
        const unsigned gb = (unsigned)std::min<int64_t>(512, nb256(nm));
        const double *g = l == 0 ? mg->g : nullptr;
        const double *d = l == 0 ? mg->d : nullptr;
        double *pw = mg->pw + 3 * l;
        hipLaunchKernelGGL(k_power_start, dim3(nb256(nm)), dim3(256), 0, st, nm, mg->fixed[l], mg->x[l]);
        for (int it = 0; it < 12 && rc == PHAST_OK; ++it) {
            rc = hex8_apply(mg->lv[l], mg->x[l], g, d, mg->Ax[l], nullptr, false, st);
            if (rc) break;
            hipLaunchKernelGGL(k_power_step, dim3(gb), dim3(256), 0, st, nm, mg->Ax[l], mg->diag[l], mg->fixed[l],
                               mg->x[l], mg->r[l], mg->b[l]);
            hipLaunchKernelGGL(k_power_norm, dim3(1), dim3(64), 0, st, gb, mg->b[l], it == 0 ? 1 : 0, pw);
            hipLaunchKernelGGL(k_scale_copy_dev, dim3(nb256(nm)), dim3(256), 0, st, nm, mg->r[l], pw, mg->x[l]);
            PHAST_LAUNCHED();
        }
    }
    if (rc != PHAST_OK) return rc;
    PHAST_HIP(hipMemcpyAsync(mg->hpart, mg->pw, sizeof(double) * 3 * mg->levels, hipMemcpyDeviceToHost, st));
    PHAST_HIP(hipStreamSynchronize(st));
    for (int l = 0; l < mg->levels; ++l) {
        if (mg->hpart[3 * l + 2] != 0.0) {
            set_error("multigrid power iteration broke down on level %d", l);
            return PHAST_ERR_NONFINITE;
        }
        const double lam = mg->hpart[3 * l + 1];
        mg->lam_max[l] = lam;
        mg->omega_l[l] = mg->omega / (1.05 * lam);
    }
    return rc;
}

static int vcycle(phast_hex8_mg *mg, int l, hipStream_t st) {
    phast_hex8_plan *pl = mg->lv[l];
    const int64_t nm = 3 * mg->n[l] + 6;
    const unsigned gb = nb256(nm);
    const double *g = l == 0 ? mg->g : nullptr;
    const double *d = l == 0 ? mg->d : nullptr;
    const bool coarsest = l == mg->levels - 1;
    const int pre = coarsest ? mg->coarse_sweeps : mg->nu;
    if (coarsest && mg->n[l] <= kCoarseMax && l > 0) {  // all sweeps in one block
        hipLaunchKernelGGL(k_coarse_sweeps, dim3(1), dim3(kCoarseMax), 0, st, pl->geo, pl->lam, pl->mu, mg->b[l],
                           mg->diag[l], mg->fixed[l], mg->omega_l[l], pre, mg->x[l]);
        PHAST_LAUNCHED();
        return PHAST_OK;
    }
    // x = omega D^-1 b, then pre - 1 more sweeps
    hipLaunchKernelGGL(k_jacobi, dim3(gb), dim3(256), 0, st, nm, mg->x[l], mg->b[l], mg->Ax[l], mg->diag[l],
                       mg->fixed[l], mg->omega_l[l], 1);
    for (int s = 1; s < pre; ++s)
        PHAST_TRY0(hex8_apply_jacobi(pl, mg->x[l], g, d, mg->Ax[l], mg->b[l], mg->diag[l], mg->fixed[l],
                                     mg->omega_l[l], st));
    if (coarsest) return PHAST_OK;
    PHAST_TRY0(hex8_apply(pl, mg->x[l], g, d, mg->Ax[l], nullptr, false, st));
    hipLaunchKernelGGL(k_residual, dim3(gb), dim3(256), 0, st, nm, mg->b[l], mg->Ax[l], mg->fixed[l], mg->r[l]);
    const phast_hex8_plan *c = mg->lv[l + 1];
    hipLaunchKernelGGL(k_restrict, dim3(nb256(c->n)), dim3(256), 0, st, c->geo.nx, c->geo.ny, c->geo.nz, mg->r[l],
                       mg->b[l + 1], mg->fixed[l + 1]);
    PHAST_TRY0(vcycle(mg, l + 1, st));
    hipLaunchKernelGGL(k_prolong_add, dim3(nb256(mg->n[l])), dim3(256), 0, st, c->geo.nx, c->geo.ny, c->geo.nz,
                       mg->x[l + 1], mg->x[l], mg->fixed[l]);
    for (int s = 0; s < mg->nu; ++s)
        PHAST_TRY0(hex8_apply_jacobi(pl, mg->x[l], g, d, mg->Ax[l], mg->b[l], mg->diag[l], mg->fixed[l],
                                     mg->omega_l[l], st));
    PHAST_LAUNCHED();
    return PHAST_OK;
}

// z = V-cycle(r) on the fine level (r, z: 3n+6 device vectors).
extern "C" int phast_hex8_mg_apply(phast_hex8_mg *mg, const double *r, double *z, hipStream_t st) {
    PHAST_REQUIRE(mg && r && z, "phast_hex8_mg_apply: null argument");
    const int64_t nm = 3 * mg->n[0] + 6;
    PHAST_HIP(hipMemcpyAsync(mg->b[0], r, sizeof(double) * nm, hipMemcpyDeviceToDevice, st));
    PHAST_TRY0(vcycle(mg, 0, st));
    PHAST_HIP(hipMemcpyAsync(z, mg->x[0], sizeof(double) * nm, hipMemcpyDeviceToDevice, st));
    return PHAST_OK;
}

extern "C" int phast_hex8_mg_weights(const phast_hex8_mg *mg, double *lam_max, double *omega) {
    PHAST_REQUIRE(mg, "phast_hex8_mg_weights: null argument");
    for (int l = 0; l < mg->levels; ++l) {
        if (lam_max) lam_max[l] = mg->lam_max[l];
        if (omega) omega[l] = mg->omega_l[l];
    }
    return PHAST_OK;
}

extern "C" int phast_hex8_mg_levels(const phast_hex8_mg *mg, int32_t *levels, int32_t *dims) {
    PHAST_REQUIRE(mg && levels, "phast_hex8_mg_levels: null argument");
    *levels = mg->levels;
    if (dims)
        for (int l = 0; l < mg->levels; ++l) {
            dims[3 * l] = mg->lv[l]->geo.nx;
            dims[3 * l + 1] = mg->lv[l]->geo.ny;
            dims[3 * l + 2] = mg->lv[l]->geo.nz;
        }
    return PHAST_OK;
}
