// Native-Q4 structured-plate kernels for gfx950 (FP64 throughout).
//
// Layout in HBM (packed over all instances of a plan):
//   nodes  : u, v, a, f as (N,2) interleaved float64 (16 B per node, double2 loads);
//            d, M, b, r, p, z, Ap, lb as (N,) float64; masks as (N,) uint8.
//   cells  : H, psi as (E,4) float64 (one 32 B row per cell, Gauss points in
//            (xi,eta) = (-,-),(+,-),(-,+),(+,+) order).
//
// Work decomposition: one block of NT threads (64, 128 or 256) per strip of NT-1 node
// columns and KY node rows of one instance.  The block marches up its strip one cell row
// at a time: each node row is staged once in LDS (plus a one-node x halo on either side),
// each thread evaluates one cell of the row, the corners going to the right-hand
// neighbour node pass through LDS and the corners going up stay in registers, and every
// owned node sums the four cells around it in the reference's element order.  No
// atomics: results are deterministic and race-free; each cell is owned (for per-cell
// outputs such as H) by the strip of its lower-left node.  The block -> (instance, strip)
// table makes one launch cover a whole batch of plates of uneven sizes.
#include "common.h"
#include "damage_cg.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

namespace phast {

struct Q4Inst {
    int32_t nx, ny, tiles_x, tiles_y;
    int64_t node_off, cell_off;
    int32_t tile_off, n_tiles;
    double gx[16], gy[16], Nq[16], w[4];  // [q*4 + a]
    double dlap[4];                       // sum_q w (gx^2 + gy^2) of corner a
    // sum-factorised rectangle constants: W(l,g) in {al, be} weights a corner row/column
    // at a Gauss point; xa = al/hx ... fold 1/h, and fxa = w al/hx ... also fold w.
    double al, be, xa, xb, ya, yb, fxa, fxb, fya, fyb, wal, wbe, wq, ome;
    double lam, mu, kappa, ezz_coef, eta;
    double Gc_l0, Gc_over_l0, l0, dt;
    int32_t split, ky;
    int32_t ssplit;  // split of the stress (isotropic under stress_degradation 'full')
};


using Q4CG = DamageCG;

constexpr int kParts = 4;  // per-tile partial sums kept between kernels

// Instance descriptors and the block -> instance tables are read-only while any kernel
// runs.  Read through the constant address space with a wave-uniform index, every field
// is a scalar load (SGPRs, scalar cache): the descriptor never enters the vector-memory
// wait counts that the strips' row prefetch relies on.
typedef const __attribute__((address_space(4))) Q4Inst QI;
__device__ __forceinline__ QI &inst_at(const Q4Inst *inst, int k) {  // k wave-uniform
    return *(QI *)(inst + __builtin_amdgcn_readfirstlane(k));
}
__device__ __forceinline__ int uniform_at(const int32_t *a, int i) {  // i wave-uniform
    typedef const __attribute__((address_space(4))) int32_t CI;
    return __builtin_amdgcn_readfirstlane(*(CI *)(a + __builtin_amdgcn_readfirstlane(i)));
}

}  // namespace phast

struct phast_q4_plan {
    int32_t n_inst = 0, tx = 128, ty = 8, n_tiles = 0;  // strip width (threads) and rows
    int64_t n_nodes = 0, n_cells = 0;
    std::vector<phast::Q4Inst> h_inst;
    phast::Q4Inst *d_inst = nullptr;
    int32_t *d_block_inst = nullptr;
    // solver work (allocated lazily on first solve)
    double *r = nullptr, *p = nullptr, *p2 = nullptr, *z = nullptr, *Ap = nullptr, *b = nullptr,
           *minv = nullptr, *lb = nullptr, *part = nullptr;  // p2: second direction buffer
    uint8_t *active = nullptr;
    phast::Q4CG *cg = nullptr;
    phast::Q4CG *h_cg = nullptr;  // pinned mirror
    int32_t last_iters = 0;
    // persistent batched step (set up by phast_q4_batch_prepare)
    // grid and block tables per kernel: [method] for the one-step kernel, [2 + method] for
    // the several-step kernel
    int32_t b_grid[4] = {0, 0, 0, 0};
    int32_t *b_pb_inst[4] = {}, *b_pb_rank[4] = {}, *b_nblk[4] = {};
    int32_t b_nb_max[4] = {0, 0, 0, 0};  // most blocks of one instance
    uint32_t *b_bar = nullptr;
    int32_t *b_err = nullptr;
    uint64_t b_spin_ticks = 0;
};

namespace phast {

// ---------------------------------------------------------------------------------------
// per-cell arithmetic
// ---------------------------------------------------------------------------------------

__device__ __forceinline__ double degrade(double dq, double eta) {
    const double omd = 1.0 - dq;
    return omd * omd * (1.0 - eta) + eta;  // physics/material.py:529
}

// Internal force contributions of one rectangular cell at its four corners (c0..c3 =
// (0,0),(1,0),(1,1),(0,1)), and optionally psi+ at its Gauss points q = ix + 2 iy.
// Sum-factorised: on a rectangle d/dx of the bilinear field is the row difference
// interpolated in y (and d/dy the column difference interpolated in x), and the force is
// the transpose.  Same arithmetic as fem_operators.py:244-254 (strain), 626-630 (g at the
// Gauss points), 301-333 (Amor) / 277-290 (isotropic), 681-699 (assembly), 983-1006 (psi+).
template <bool PSI>
__device__ __forceinline__ void cell_mech(QI &P, const double ux[4], const double uy[4],
                                          const double dd[4], double fx[4], double fy[4], double psi[4]) {
    const double al = P.al, be = P.be;
    const double dx0x = ux[1] - ux[0], dx1x = ux[2] - ux[3], dx0y = uy[1] - uy[0], dx1y = uy[2] - uy[3];
    const double dy0x = ux[3] - ux[0], dy1x = ux[2] - ux[1], dy0y = uy[3] - uy[0], dy1y = uy[2] - uy[1];
    const double exx_[2] = {P.xa * dx0x + P.xb * dx1x, P.xb * dx0x + P.xa * dx1x};  // [iy]
    const double vyx_[2] = {P.xa * dx0y + P.xb * dx1y, P.xb * dx0y + P.xa * dx1y};  // d uy/dx [iy]
    const double eyy_[2] = {P.ya * dy0y + P.yb * dy1y, P.yb * dy0y + P.ya * dy1y};  // [ix]
    const double uxy_[2] = {P.ya * dy0x + P.yb * dy1x, P.yb * dy0x + P.ya * dy1x};  // d ux/dy [ix]
    const double r00 = al * dd[0] + be * dd[1], r01 = be * dd[0] + al * dd[1];      // bottom row at ix
    const double r10 = al * dd[3] + be * dd[2], r11 = be * dd[3] + al * dd[2];      // top row at ix
    const double dq[4] = {al * r00 + be * r10, al * r01 + be * r11, be * r00 + al * r10, be * r01 + al * r11};
    double sxx[4], syy[4], sxy[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int ix = q & 1, iy = q >> 1;
        const double exx = exx_[iy], eyy = eyy_[ix], gxy = uxy_[ix] + vyx_[iy];
        const double omd = 1.0 - dq[q];
        const double g = omd * omd * P.ome + P.eta;  // physics/material.py:529
        // stress by the stress split, psi+ by the energy split (the same unless
        // stress_degradation is 'full')
        const double tr2 = exx + eyy;
        const double ezz = P.ezz_coef * tr2;
        const double tr = tr2 + ezz;
        const double trp = fmax(tr, 0.0), trm = tr - trp;
        const double t3 = tr * (1.0 / 3.0);  // within one ulp of tr / 3 (no IEEE division)
        const double dxx = exx - t3, dyy = eyy - t3;
        if (P.ssplit == PHAST_SPLIT_AMOR) {
            const double kp = P.kappa * trp, km = P.kappa * trm;
            sxx[q] = g * (kp + 2.0 * P.mu * dxx) + km;
            syy[q] = g * (kp + 2.0 * P.mu * dyy) + km;
            sxy[q] = g * P.mu * gxy;
        } else {
            const double c11 = P.lam + 2.0 * P.mu;
            sxx[q] = g * (c11 * exx + P.lam * eyy);
            syy[q] = g * (P.lam * exx + c11 * eyy);
            sxy[q] = g * (P.mu * gxy);
        }
        if (PSI) {
            const double exy = gxy / 2.0;
            if (P.split == PHAST_SPLIT_AMOR) {
                const double dzz = ezz - t3;
                psi[q] = 0.5 * P.kappa * trp * trp + P.mu * (dxx * dxx + dyy * dyy + dzz * dzz + 2.0 * exy * exy);
            } else {
                const double t = exx + eyy;
                psi[q] = 0.5 * P.lam * t * t + P.mu * (exx * exx + eyy * eyy + 2.0 * exy * exy);
            }
        }
    }
    const double sxx0 = sxx[0] + sxx[1], sxx1 = sxx[2] + sxx[3];      // [iy]
    const double syy0 = syy[0] + syy[2], syy1 = syy[1] + syy[3];      // [ix]
    const double sxr0 = sxy[0] + sxy[1], sxr1 = sxy[2] + sxy[3];      // [iy]
    const double sxc0 = sxy[0] + sxy[2], sxc1 = sxy[1] + sxy[3];      // [ix]
    const double Rxx0 = P.fxa * sxx0 + P.fxb * sxx1, Rxx1 = P.fxb * sxx0 + P.fxa * sxx1;  // [ly]
    const double Rxr0 = P.fxa * sxr0 + P.fxb * sxr1, Rxr1 = P.fxb * sxr0 + P.fxa * sxr1;  // [ly]
    const double Ryy0 = P.fya * syy0 + P.fyb * syy1, Ryy1 = P.fyb * syy0 + P.fya * syy1;  // [lx]
    const double Rxc0 = P.fya * sxc0 + P.fyb * sxc1, Rxc1 = P.fyb * sxc0 + P.fya * sxc1;  // [lx]
    fx[0] = -Rxx0 - Rxc0;
    fx[1] = Rxx0 - Rxc1;
    fx[2] = Rxx1 + Rxc1;
    fx[3] = -Rxx1 + Rxc0;
    fy[0] = -Ryy0 - Rxr0;
    fy[1] = -Ryy1 + Rxr0;
    fy[2] = Ryy1 + Rxr1;
    fy[3] = Ryy0 - Rxr1;
}

// AT2 damage operator of one rectangular cell (damage_solver.py:1797-1838):
// y_a = Gc l0 sum_q w gradN_a.grad x_q + sum_q N_a rho_q x_q w, rho_q = 2 H_q + Gc/l0,
// sum-factorised as cell_mech.
__device__ __forceinline__ void cell_damage(QI &P, const double x[4], const double rho[4],
                                            double gcl0, double y[4]) {
    const double al = P.al, be = P.be;
    const double dxr0 = x[1] - x[0], dxr1 = x[2] - x[3], dyc0 = x[3] - x[0], dyc1 = x[2] - x[1];
    const double gdx0 = P.xa * dxr0 + P.xb * dxr1, gdx1 = P.xb * dxr0 + P.xa * dxr1;  // [iy]
    const double gdy0 = P.ya * dyc0 + P.yb * dyc1, gdy1 = P.yb * dyc0 + P.ya * dyc1;  // [ix]
    // sum over the two Gauss points along the constant direction: factor 2
    const double Lx0 = 2.0 * (P.fxa * gdx0 + P.fxb * gdx1), Lx1 = 2.0 * (P.fxb * gdx0 + P.fxa * gdx1);  // [ly]
    const double Ly0 = 2.0 * (P.fya * gdy0 + P.fyb * gdy1), Ly1 = 2.0 * (P.fyb * gdy0 + P.fya * gdy1);  // [lx]
    const double lap[4] = {-Lx0 - Ly0, Lx0 - Ly1, Lx1 + Ly1, -Lx1 + Ly0};
    const double r00 = al * x[0] + be * x[1], r01 = be * x[0] + al * x[1];
    const double r10 = al * x[3] + be * x[2], r11 = be * x[3] + al * x[2];
    const double t0 = rho[0] * (al * r00 + be * r10), t1 = rho[1] * (al * r01 + be * r11);
    const double t2 = rho[2] * (be * r00 + al * r10), t3 = rho[3] * (be * r01 + al * r11);
    // transpose of the interpolation, w folded: rows (ly) then columns (lx)
    const double b0 = P.wal * t0 + P.wbe * t2, b1 = P.wal * t1 + P.wbe * t3;  // ly = 0, [ix]
    const double u0 = P.wbe * t0 + P.wal * t2, u1 = P.wbe * t1 + P.wal * t3;  // ly = 1, [ix]
    y[0] = gcl0 * lap[0] + (al * b0 + be * b1);
    y[1] = gcl0 * lap[1] + (be * b0 + al * b1);
    y[2] = gcl0 * lap[2] + (be * u0 + al * u1);
    y[3] = gcl0 * lap[3] + (al * u0 + be * u1);
}

// ---------------------------------------------------------------------------------------
// strip geometry
// ---------------------------------------------------------------------------------------
//
// A block of NT threads owns a strip of NT-2 node columns [i0, i0+NT-2) and KY node rows
// [j0, j0+KY) of one instance.  Thread t holds node slot t <-> column i0-1+t (slots 0 and
// NT-1 are the x halo).  The block marches up the strip one cell row at a time: thread t
// (t <= NT-2) evaluates cell column i0-1+t of the current cell row from two node rows
// staged in LDS, hands the two corners that belong to its right neighbour node through
// LDS, and completes node (i0-1+t) of the lower row from its left cell (corners c1, c2),
// its own cell (c0, c3) and the upper-corner sums of the previous cell row kept in
// registers -- in the reference's element order, with no atomics.  Every node row is read
// once per strip (plus the 2-node x halo), so the only redundancy is one extra cell row
// per strip.

struct Tile {
    int inst, i0, j0, nxn, nyn, jend;
};

template <int NT>
__device__ __forceinline__ Tile tile_of(const Q4Inst *inst, const int32_t *block_inst, int strip) {
    Tile t;
    t.inst = uniform_at(block_inst, strip);
    QI &P = inst_at(inst, t.inst);
    const int k = strip - P.tile_off;
    t.i0 = (k % P.tiles_x) * (NT - 2);
    t.j0 = (k / P.tiles_x) * P.ky;
    t.nxn = P.nx + 1;
    t.nyn = P.ny + 1;
    t.jend = min(t.j0 + P.ky, t.nyn);
    return t;
}

// ---------------------------------------------------------------------------------------
// mechanics: internal force / explicit step (+ psi+, history)
// ---------------------------------------------------------------------------------------

struct MechArgs {
    const Q4Inst *inst;
    const int32_t *block_inst;
    const double2 *u, *v, *a;
    const double *d;
    const double2 *fext;
    const double *minv;
    const uint8_t *bcm;
    const double *bcv;
    const double *Hold;
    double2 *u1, *v1, *a1, *f;
    double *Hnew, *psi;
    const double2 *fdamp;  // explicit step: Kelvin-Voigt damping force beta f_int(v, d), or null
    double fscale;         // internal force: f = fscale f_int (1 for the plain force)
};

// LDS of one strip pass: three node rows (u_x, u_y, d) double-buffered, plus the corners
// c1 (x,y) and c2 (x,y) of each cell of the current row.
template <int NT>
constexpr int kMechLds = 10 * NT;

// Node-row operands of one thread's slot, fetched ahead of use.
struct RowRegs {
    double2 u, v, a;
    double d;
    uint32_t m;  // bc mask bytes of the node (x in the low byte, y in the high byte)
    bool ok;     // slot inside the plate (the loads of other slots read a clamped node)
};

// Results of the strips are written with non-temporal stores: the kernel does not read them
// again, and keeping them out of the caches leaves room for the node rows that neighbouring
// strips read (explicit step at 8192x4096: 1.39 -> 1.32 ms; the force pass is unchanged).
typedef double nd2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void st_nt(double2 *p, double2 v) {
    const nd2_t w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<nd2_t *>(p));
}

template <int NT, bool STEP, bool HIST>
__device__ __forceinline__ void mech_strip(const MechArgs &A, int strip, double *sm) {
    double(*sux)[NT] = reinterpret_cast<double(*)[NT]>(sm);
    double(*suy)[NT] = reinterpret_cast<double(*)[NT]>(sm + 2 * NT);
    double(*sd)[NT] = reinterpret_cast<double(*)[NT]>(sm + 4 * NT);
    double(*sf)[NT] = reinterpret_cast<double(*)[NT]>(sm + 6 * NT);
    __syncthreads();  // the LDS may still be read by the previous strip of this block

    const Tile T = tile_of<NT>(A.inst, A.block_inst, strip);
    QI &P = inst_at(A.inst, T.inst);
    const double dt = P.dt, c2 = 0.5 * P.dt * P.dt;
    const int t = threadIdx.x;
    const int ie = T.i0 - 1 + t;  // node column of slot t, cell column of thread t
    const bool col_ok = ie >= 0 && ie < T.nxn;
    const bool owner = t >= 1 && t <= NT - 2 && ie < T.nxn;  // slot t holds an owned node column
    const bool cell_col = t <= NT - 2 && ie >= 0 && ie < P.nx;
    const bool bc = STEP && A.bcm != nullptr;

    // Slot t of node row j, fetched two rows ahead (two register sets taking turns) so that
    // its latency overlaps two rows of cell arithmetic.  Every lane loads (out-of-range slots
    // read a clamped node and keep zeros), and these loads are the last memory operations
    // of a row, after its stores and operand loads: the compiler's wait for the row staged
    // next then leaves the newest prefetch in flight.
    const int ic = min(max(ie, 0), T.nxn - 1);
    const uint8_t *bcm = bc ? A.bcm : reinterpret_cast<const uint8_t *>(A.d);  // any readable bytes
    auto fetch = [&](RowRegs &R, int j) {
        R.ok = col_ok && j >= 0 && j < T.nyn && j <= T.jend;
        // rows outside [0, min(jend, nyn-1)] read a row of the strip that is read anyway
        const int64_t n = P.node_off + (int64_t)min(max(j, 0), min(T.jend, T.nyn - 1)) * T.nxn + ic;
        R.u = A.u[n];
        R.d = A.d[n];
        if (STEP) {
            R.v = A.v[n];
            R.a = A.a[n];
            R.m = *reinterpret_cast<const uint16_t *>(bcm + 2 * n);  // both bytes, decoded when staged
        }
    };
    double2 gv = make_double2(0.0, 0.0), ga = gv, nv = gv, na = gv;  // v, a of rows jr and jr+1
    uint32_t gm = 0, nm = 0;
    // stage a fetched row into LDS (slots outside the plate hold zeros); the values of the
    // other operands are masked here too, where the row's loads are waited for anyway
    auto store = [&](const RowRegs &R, int buf, int j) {
        const double2 z2 = make_double2(0.0, 0.0);
        const double2 u = R.ok ? R.u : z2;
        double ux = u.x, uy = u.y;
        if (STEP) {
            const double2 v = R.ok ? R.v : z2, a = R.ok ? R.a : z2;
            const uint32_t m = (R.ok && bc) ? ((R.m & 0xffu) | ((R.m >> 8) << 1)) : 0u;
            ux = (u.x + dt * v.x) + c2 * a.x;  // mechanics_solver.py:160
            uy = (u.y + dt * v.y) + c2 * a.y;
            if (m) {
                const int64_t n = P.node_off + (int64_t)j * T.nxn + ie;
                if (m & 1u) ux = A.bcv[2 * n];
                if (m & 2u) uy = A.bcv[2 * n + 1];
            }
            nv = v;
            na = a;
            nm = m;
        }
        sux[buf][t] = ux;
        suy[buf][t] = uy;
        sd[buf][t] = R.ok ? R.d : 0.0;
    };

    // results of a row, written at the start of the next row (before its loads)
    bool w_node = false, w_cell = false;
    int64_t w_n = 0, w_c = 0;
    double2 w_f = make_double2(0.0, 0.0), w_v = w_f, w_a = w_f, w_h01 = w_f, w_h23 = w_f, w_p01 = w_f,
            w_p23 = w_f;
    auto flush = [&]() {
        if (w_node) {
            if (!STEP) {
                st_nt(&A.f[w_n], w_f);
            } else {
                st_nt(&A.u1[w_n], w_f);
                st_nt(&A.v1[w_n], w_v);
                st_nt(&A.a1[w_n], w_a);
            }
        }
        if (HIST && w_cell) {
            reinterpret_cast<double2 *>(A.Hnew)[2 * w_c] = w_h01;
            reinterpret_cast<double2 *>(A.Hnew)[2 * w_c + 1] = w_h23;
            if (A.psi) {
                reinterpret_cast<double2 *>(A.psi)[2 * w_c] = w_p01;
                reinterpret_cast<double2 *>(A.psi)[2 * w_c + 1] = w_p23;
            }
        }
    };

    // D register sets take turns: node rows are fetched D rows ahead of their staging
    constexpr int D = 2;
    RowRegs R[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        R[k].v = R[k].a = make_double2(0.0, 0.0);
        R[k].m = 0u;
        R[k].ok = false;
    }
    fetch(R[0], T.j0 - 1);
    store(R[0], 0, T.j0 - 1);
#pragma unroll
    for (int k = 0; k < D; ++k) {
        fetch(R[k], T.j0 + k);
        __builtin_amdgcn_sched_barrier(0);  // rows are issued in order
    }
    double accx = 0.0, accy = 0.0;
    int cur = 0;
    const int cc = min(max(ie, 0), P.nx - 1);
    // one cell row: Q holds node row jr+1 on entry and row jr+1+D on exit
    auto row = [&](int jr, RowRegs &Q) {
        const int up = cur ^ 1;
        store(Q, up, jr + 1);  // also sets nv, na, nm (v, a, bc bits of row jr+1)
        flush();
        // operands consumed later in this row (every lane loads a clamped index)
        const bool gat = owner && jr >= T.j0;
        const int64_t ng = P.node_off + (int64_t)jr * T.nxn + ie;
        double2 fe = make_double2(0.0, 0.0), fdm = fe;
        double mi = 0.0;
        if (STEP) {
            const int64_t nc = P.node_off + (int64_t)max(jr, 0) * T.nxn + ic;
            if (A.fext) fe = A.fext[nc];
            if (A.fdamp) fdm = A.fdamp[nc];
            mi = A.minv[nc];
        }
        const bool cell_ok = cell_col && jr >= 0 && jr < P.ny;
        const bool own_cell = HIST && cell_ok && t >= 1 && jr >= T.j0;
        const int64_t c = P.cell_off + (int64_t)jr * P.nx + ie;
        double2 h01 = make_double2(0.0, 0.0), h23 = h01;
        if (HIST) {
            const int64_t ccl = P.cell_off + (int64_t)min(max(jr, 0), P.ny - 1) * P.nx + cc;
            h01 = reinterpret_cast<const double2 *>(A.Hold)[2 * ccl];
            h23 = reinterpret_cast<const double2 *>(A.Hold)[2 * ccl + 1];
        }
        __builtin_amdgcn_sched_barrier(0);  // the prefetch issues after the operand loads
        fetch(Q, jr + 1 + D);
        __syncthreads();
        double fx[4] = {0.0, 0.0, 0.0, 0.0}, fy[4] = {0.0, 0.0, 0.0, 0.0};
        w_cell = false;
        if (cell_ok) {
            const double ux[4] = {sux[cur][t], sux[cur][t + 1], sux[up][t + 1], sux[up][t]};
            const double uy[4] = {suy[cur][t], suy[cur][t + 1], suy[up][t + 1], suy[up][t]};
            const double dd[4] = {sd[cur][t], sd[cur][t + 1], sd[up][t + 1], sd[up][t]};
            double ps[4];
            if (own_cell) {  // cells owned by this strip: lower-left node inside
                cell_mech<true>(P, ux, uy, dd, fx, fy, ps);
                // max returns one operand bit-exactly (no NaN inputs)
                w_h01.x = h01.x >= ps[0] ? h01.x : ps[0];
                w_h01.y = h01.y >= ps[1] ? h01.y : ps[1];
                w_h23.x = h23.x >= ps[2] ? h23.x : ps[2];
                w_h23.y = h23.y >= ps[3] ? h23.y : ps[3];
                w_p01 = make_double2(ps[0], ps[1]);
                w_p23 = make_double2(ps[2], ps[3]);
                w_c = c;
                w_cell = true;
            } else {
                cell_mech<false>(P, ux, uy, dd, fx, fy, ps);
            }
        }
        sf[0][t] = fx[1];
        sf[1][t] = fy[1];
        sf[2][t] = fx[2];
        sf[3][t] = fy[2];
        __syncthreads();
        w_node = false;
        if (t >= 1 && t <= NT - 2) {
            if (gat) {
                // cells (i-1,j-1) c2, (i,j-1) c3 [acc], (i-1,j) c1, (i,j) c0
                const double fxn = (accx + sf[0][t - 1]) + fx[0];
                const double fyn = (accy + sf[1][t - 1]) + fy[0];
                if (!STEP) {
                    w_f = make_double2(A.fscale * fxn, A.fscale * fyn);  // x 1 is exact
                } else {
                    // mechanics_solver.py:192-198: f_net = (f_ext - f_int) [- beta f_int(v, d)]
                    double nx = fe.x - fxn, ny = fe.y - fyn;
                    if (A.fdamp) {
                        nx = nx - fdm.x;
                        ny = ny - fdm.y;
                    }
                    double ax = mi * nx, ay = mi * ny;
                    double vx = gv.x + 0.5 * dt * (ga.x + ax), vy = gv.y + 0.5 * dt * (ga.y + ay);
                    if (gm & 1u) ax = vx = 0.0;
                    if (gm & 2u) ay = vy = 0.0;
                    w_f = make_double2(sux[cur][t], suy[cur][t]);
                    w_v = make_double2(vx, vy);
                    w_a = make_double2(ax, ay);
                }
                w_n = ng;
                w_node = true;
            }
            accx = sf[2][t - 1] + fx[3];
            accy = sf[3][t - 1] + fy[3];
        }
        if (STEP) {
            gv = nv;
            ga = na;
            gm = nm;
        }
        cur = up;
    };
    // all D rows unconditional inside the loop (the last < D rows run after it), so the
    // compiler sees the other sets' loads issued between two stagings of one set on
    // every path
    int jr = T.j0 - 1;
    for (; jr + D - 1 < T.jend; jr += D) {
#pragma unroll
        for (int k = 0; k < D; ++k) row(jr + k, R[k]);
    }
#pragma unroll
    for (int k = 0; k < D - 1; ++k)
        if (jr + k < T.jend) row(jr + k, R[k]);
    flush();
}

template <int NT, bool STEP, bool HIST>
__global__ __launch_bounds__(NT) void q4_mech_kernel(MechArgs A) {
    __shared__ double sm[kMechLds<NT>];
    mech_strip<NT, STEP, HIST>(A, blockIdx.x, sm);
}

// Element-parallel strain / psi+ (one thread per cell, nodes read from global memory).
struct CellArgs {
    const Q4Inst *inst;
    int n_inst;
    const double2 *u;
    double *eps;  // [3][E][4]
    double *psi;
    const double *Hold;
    double *Hnew;
    int64_t n_cells;
};

__device__ __forceinline__ int find_inst_by_cell(const Q4Inst *inst, int n_inst, int64_t c) {
    int lo = 0, hi = n_inst - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (inst[mid].cell_off <= c) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void q4_cell_kernel(CellArgs A) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= A.n_cells) return;
    QI &P = *(QI *)(A.inst + find_inst_by_cell(A.inst, A.n_inst, c));
    const int64_t lc = c - P.cell_off;
    const int ie = (int)(lc % P.nx), je = (int)(lc / P.nx);
    const int64_t n0 = P.node_off + (int64_t)je * (P.nx + 1) + ie;
    const int64_t id[4] = {n0, n0 + 1, n0 + P.nx + 2, n0 + P.nx + 1};
    double ux[4], uy[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double2 uu = A.u[id[k]];
        ux[k] = uu.x;
        uy[k] = uu.y;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double exx = 0.0, eyy = 0.0, gxa = 0.0, gxb = 0.0;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            exx += P.gx[q * 4 + a] * ux[a];
            eyy += P.gy[q * 4 + a] * uy[a];
            gxa += P.gy[q * 4 + a] * ux[a];
            gxb += P.gx[q * 4 + a] * uy[a];
        }
        const double gxy = gxa + gxb;
        if (A.eps) {
            A.eps[c * 4 + q] = exx;
            A.eps[(A.n_cells + c) * 4 + q] = eyy;
            A.eps[(2 * A.n_cells + c) * 4 + q] = gxy;
        }
        if (A.psi || A.Hnew) {
            const double exy = gxy / 2.0;
            double ps;
            if (P.split == PHAST_SPLIT_AMOR) {
                const double tr2 = exx + eyy, ezz = P.ezz_coef * tr2, tr = tr2 + ezz;
                const double trp = fmax(tr, 0.0);
                const double t3 = tr * (1.0 / 3.0);
                const double dxx = exx - t3, dyy = eyy - t3, dzz = ezz - t3;
                ps = 0.5 * P.kappa * trp * trp + P.mu * (dxx * dxx + dyy * dyy + dzz * dzz + 2.0 * exy * exy);
            } else {
                const double t = exx + eyy;
                ps = 0.5 * P.lam * t * t + P.mu * (exx * exx + eyy * eyy + 2.0 * exy * exy);
            }
            if (A.psi) A.psi[c * 4 + q] = ps;
            if (A.Hnew) {
                const double h = A.Hold[c * 4 + q];
                A.Hnew[c * 4 + q] = h >= ps ? h : ps;
            }
        }
    }
}

__global__ __launch_bounds__(256) void history_max_kernel(int64_t n, const double *H, const double *psi,
                                                          double *out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const double h = H[i], p = psi[i];
        out[i] = h >= p ? h : p;  // torch.maximum for non-NaN inputs (h_update.py:47-49)
    }
}

// ---------------------------------------------------------------------------------------
// nodal gathers of per-cell scalar quantities (lumped mass, elem_to_node, damage operator)
// ---------------------------------------------------------------------------------------

enum DmgOp : int {
    OP_MASS = 0,       // out = sum_q N_a w                          (M_scalar)
    OP_E2N = 1,        // out = sum_q N_a f_q w / (M + 1e-30)         (elem_to_node)
    OP_RHSDIAG = 2,    // b, diag
    OP_APPLY = 3,      // y = A x
    OP_CG_INIT = 4,    // r = b - A d, active set, z, p; partials r.r, b.b, rf.z
    OP_CG_AP = 5,      // Ap = A p; partials p.Ap, min step ratios
    OP_CG_RESID = 6,   // r = b - A d; new active set; partials rf.rf, changed
    OP_PC_INIT = 7,    // post-clamp: r = b - A d, z = r Minv, p = z; partials r.r, b.b, r.z
    OP_PC_AP = 8,      // post-clamp: Ap = A p; partial p.Ap
    OP_PC_RESID = 9,   // post-clamp true residual r = b - A d
    // A p with the previous iteration's direction update fused in: the staged p of every node
    // is (active ? 0 :) (restart ? z : beta p_old + z), the arithmetic of V_UPDATE_P /
    // V_PC_UPDATE_P; the owner writes it to pout.  x holds p_old and pout is another buffer,
    // since neighbouring strips still stage p_old for their halo rows.
    OP_CG_APF = 10,
    OP_PC_APF = 11,
};

struct DmgArgs {
    const Q4Inst *inst;
    const int32_t *block_inst;
    const double *H;      // (E,4)
    const double *Gc;     // (E,) or null
    const double *x;      // gathered field (d, p, field_q for E2N)
    const double *M;      // lumped mass for E2N
    double *out, *out2;   // outputs (y / b, diag)
    // CG vectors
    double *d, *r, *z, *p, *Ap;
    double *pout;  // OP_CG_APF / OP_PC_APF: the updated direction
    const double *b, *minv, *lb;
    const uint8_t *fixed;
    uint8_t *active;
    Q4CG *cg;
    double *part;
    int32_t n_tiles;
};

template <int NT>
constexpr int kDmgLds = 6 * NT;

// One strip of a damage-operator pass; per-strip partial sums land in part[slot] with
// slot = strip (and n_tiles + slot, 2 n_tiles + slot for the second and third sums).
// For the CG passes the gathered field x is the iterate the node-wise update reads (d for
// the residual passes, p for the A p passes), so the owned node's value is taken from the
// staged row.
template <int NT, int OP>
__device__ __forceinline__ void dmg_strip(const DmgArgs &A, int strip, double *sm, double *red) {
    constexpr bool FUSEP = (OP == OP_CG_APF || OP == OP_PC_APF);
    constexpr bool AP_CG = (OP == OP_CG_AP || OP == OP_CG_APF), AP_PC = (OP == OP_PC_AP || OP == OP_PC_APF);
    constexpr bool GATHER_X = (OP == OP_APPLY || OP == OP_CG_INIT || AP_CG || OP == OP_CG_RESID ||
                               OP == OP_PC_INIT || AP_PC || OP == OP_PC_RESID);
    constexpr bool CELL_H = (OP != OP_MASS);  // per-cell (E,4) operand: H, or the field for E2N
    constexpr int NOUT = (OP == OP_RHSDIAG) ? 2 : 1;
    double(*sx)[NT] = reinterpret_cast<double(*)[NT]>(sm);       // gathered field rows
    double(*sy)[NT] = reinterpret_cast<double(*)[NT]>(sm + 2 * NT);  // corners c1, c2 per output
    __syncthreads();  // the LDS may still be read by the previous strip of this block

    const Tile T = tile_of<NT>(A.inst, A.block_inst, strip);
    QI &P = inst_at(A.inst, T.inst);
    const int t = threadIdx.x;
    const int ie = T.i0 - 1 + t;
    const bool col_ok = ie >= 0 && ie < T.nxn;
    const bool owner = t >= 1 && t <= NT - 2 && ie < T.nxn;
    const bool cell_col = t <= NT - 2 && ie >= 0 && ie < P.nx;
    const double *Hc = (OP == OP_E2N) ? A.x : A.H;

    // next node row's x (slot t) and next cell row's (E,4) operand, fetched one row ahead
    double px = 0.0, pgc = 0.0;
    double2 ph01 = make_double2(0.0, 0.0), ph23 = ph01;
    double beta = 0.0;
    int32_t restart = 0;
    if (FUSEP) {
        const Q4CG &S = A.cg[T.inst];
        beta = S.beta;
        restart = S.restart;
    }
    auto fetch_x = [&](int j) {
        if (!GATHER_X) return;
        const bool in = col_ok && j >= 0 && j < T.nyn && j <= T.jend;
        if (!FUSEP) {
            px = in ? A.x[P.node_off + (int64_t)j * T.nxn + ie] : 0.0;
        } else {
            // every lane loads (a clamped node for slots outside the strip), then masks
            const int jc = min(max(j, 0), min(T.jend, T.nyn - 1)), icl = min(max(ie, 0), T.nxn - 1);
            const int64_t q = P.node_off + (int64_t)jc * T.nxn + icl;
            const double zq = A.z[q], pq = A.x[q];
            const bool act = OP == OP_CG_APF && A.active[q];
            double pn = (OP == OP_CG_APF && restart) ? zq : pq * beta + zq;  // q4_vec_kernel V_UPDATE_P
            if (act) pn = 0.0;
            px = in ? pn : 0.0;
        }
    };
    auto fetch_cell = [&](int jc) {
        if (CELL_H && cell_col && jc >= 0 && jc < P.ny && jc < T.jend) {
            const int64_t c = P.cell_off + (int64_t)jc * P.nx + ie;
            ph01 = reinterpret_cast<const double2 *>(Hc)[2 * c];
            ph23 = reinterpret_cast<const double2 *>(Hc)[2 * c + 1];
            if (OP != OP_E2N && A.Gc) pgc = A.Gc[c];
        } else {
            ph01 = ph23 = make_double2(0.0, 0.0);
            pgc = 0.0;
        }
    };

    constexpr double INF = __builtin_huge_val();
    double p0 = 0.0, p1 = AP_CG ? INF : 0.0, p2 = AP_CG ? INF : 0.0;
    double acc0 = 0.0, acc1 = 0.0;
    int cur = 0;
    fetch_x(T.j0 - 1);
    if (GATHER_X) sx[0][t] = px;
    fetch_x(T.j0);
    fetch_cell(T.j0 - 1);
    for (int jr = T.j0 - 1; jr < T.jend; ++jr) {
        const int up = cur ^ 1;
        if (GATHER_X) sx[up][t] = px;
        const double2 h01 = ph01, h23 = ph23;
        const double gc = pgc;
        // node operands of the gather, issued before the prefetch
        const bool gat = owner && jr >= T.j0;
        const int64_t n = P.node_off + (int64_t)jr * T.nxn + ie;
        double o_b = 0.0, o_lb = 0.0, o_mi = 0.0, o_d = 0.0, o_M = 0.0;
        bool o_fx = false, o_act = false;
        if (gat) {
            if (OP == OP_E2N) o_M = A.M[n];
            if (OP == OP_CG_INIT || OP == OP_CG_RESID || OP == OP_PC_INIT || OP == OP_PC_RESID) {
                o_b = A.b[n];
                o_fx = A.fixed && A.fixed[n];
            }
            if (OP == OP_CG_INIT || OP == OP_CG_RESID || AP_CG) o_lb = A.lb[n];
            if (OP == OP_CG_INIT || OP == OP_PC_INIT) o_mi = A.minv[n];
            if (OP == OP_CG_RESID) o_act = A.active[n] != 0;
            if (AP_CG) o_d = A.d[n];
        }
        fetch_x(jr + 2);
        fetch_cell(jr + 1);
        __syncthreads();
        double y[NOUT * 4];
#pragma unroll
        for (int k = 0; k < NOUT * 4; ++k) y[k] = 0.0;
        if (cell_col && jr >= 0 && jr < P.ny) {
            if (OP == OP_MASS) {
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    double m = 0.0;
#pragma unroll
                    for (int q = 0; q < 4; ++q) m += P.Nq[q * 4 + a] * P.w[q];
                    y[a] = m;
                }
            } else if (OP == OP_E2N) {
                const double fq[4] = {h01.x, h01.y, h23.x, h23.y};
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    double m = 0.0;
#pragma unroll
                    for (int q = 0; q < 4; ++q) m += P.Nq[q * 4 + a] * fq[q] * P.w[q];
                    y[a] = m;
                }
            } else {
                const double gcl0 = A.Gc ? gc * P.l0 : P.Gc_l0;
                const double ratio = A.Gc ? gc / P.l0 : P.Gc_over_l0;
                const double Hq[4] = {h01.x, h01.y, h23.x, h23.y};
                double rho[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) rho[q] = 2.0 * Hq[q] + ratio;
                if (OP == OP_RHSDIAG) {
#pragma unroll
                    for (int a = 0; a < 4; ++a) {
                        double bb = 0.0, dg = 0.0;
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            bb += P.Nq[q * 4 + a] * (2.0 * Hq[q] * P.w[q]);
                            dg += (P.Nq[q * 4 + a] * P.Nq[q * 4 + a]) * (rho[q] * P.w[q]);
                        }
                        y[a] = bb;
                        y[4 + a] = gcl0 * P.dlap[a] + dg;
                    }
                } else {
                    const double xe[4] = {sx[cur][t], sx[cur][t + 1], sx[up][t + 1], sx[up][t]};
                    cell_damage(P, xe, rho, gcl0, y);
                }
            }
        }
        sy[0][t] = y[1];
        sy[1][t] = y[2];
        if (NOUT == 2) {
            sy[2][t] = y[5];
            sy[3][t] = y[6];
        }
        __syncthreads();
        if (t >= 1 && t <= NT - 2) {
            if (gat) {
                const double g0 = (acc0 + sy[0][t - 1]) + y[0];
                const double g1 = (NOUT == 2) ? (acc1 + sy[2][t - 1]) + y[4] : 0.0;
                const double xn = GATHER_X ? sx[cur][t] : 0.0;  // x at the owned node
                if (OP == OP_MASS || OP == OP_APPLY) {
                    A.out[n] = g0;
                } else if (OP == OP_E2N) {
                    A.out[n] = g0 / (o_M + 1e-30);
                } else if (OP == OP_RHSDIAG) {
                    A.out[n] = g0;
                    A.out2[n] = g1;
                } else if (OP == OP_CG_INIT || OP == OP_CG_RESID) {
                    const double dn = xn;
                    double r = o_b - g0;
                    if (o_fx) r = 0.0;
                    const bool act =
                        ((dn <= o_lb + 1e-14) && (r < 0.0)) || ((dn >= 1.0 - 1e-14) && (r > 0.0)) || o_fx;
                    const double rf = act ? 0.0 : r;
                    A.r[n] = r;
                    if (OP == OP_CG_RESID) {
                        p1 += (act != o_act) ? 1.0 : 0.0;
                        p0 += rf * rf;
                    } else {
                        double zz = rf * o_mi;
                        if (act) zz = 0.0;
                        A.z[n] = zz;
                        A.p[n] = zz;
                        p0 += r * r;
                        p1 += o_b * o_b;
                        p2 += rf * zz;
                    }
                    A.active[n] = act ? 1 : 0;
                } else if (AP_CG) {
                    const double pn = xn;
                    if (FUSEP) A.pout[n] = pn;
                    A.Ap[n] = g0;
                    p0 += pn * g0;
                    if (pn > 0.0) p1 = fmin(p1, (1.0 - o_d) / pn);
                    if (pn < 0.0) p2 = fmin(p2, (o_lb - o_d) / pn);
                } else if (OP == OP_PC_INIT) {
                    double r = o_b - g0;
                    if (o_fx) r = 0.0;
                    const double zz = r * (o_fx ? 0.0 : o_mi);
                    A.r[n] = r;
                    A.z[n] = zz;
                    A.p[n] = zz;
                    p0 += r * r;
                    p1 += o_b * o_b;
                    p2 += r * zz;
                } else if (AP_PC) {
                    if (FUSEP) A.pout[n] = xn;
                    A.Ap[n] = g0;
                    p0 += xn * g0;
                } else if (OP == OP_PC_RESID) {
                    double r = o_b - g0;
                    if (o_fx) r = 0.0;
                    A.r[n] = r;
                }
            }
            acc0 = sy[1][t - 1] + y[3];
            if (NOUT == 2) acc1 = sy[3][t - 1] + y[7];
        }
        cur = up;
    }
    if (OP == OP_CG_INIT || OP == OP_CG_RESID || AP_CG || OP == OP_PC_INIT || AP_PC) {
        const int64_t slot = strip;
        const double s0 = block_sum<NT>(p0, red);
        if (threadIdx.x == 0) A.part[slot] = s0;
        if (AP_CG) {
            const double s1 = block_min<NT>(p1, red);
            const double s2 = block_min<NT>(p2, red);
            if (threadIdx.x == 0) {
                A.part[A.n_tiles + slot] = s1;
                A.part[2 * A.n_tiles + slot] = s2;
            }
        } else if (!AP_PC) {
            const double s1 = block_sum<NT>(p1, red);
            if (threadIdx.x == 0) A.part[A.n_tiles + slot] = s1;
            if (OP == OP_CG_INIT || OP == OP_PC_INIT) {
                const double s2 = block_sum<NT>(p2, red);
                if (threadIdx.x == 0) A.part[2 * A.n_tiles + slot] = s2;
            }
        }
    }
}

template <int NT, int OP>
__global__ __launch_bounds__(NT) void q4_dmg_kernel(DmgArgs A) {
    __shared__ double sm[kDmgLds<NT>];
    __shared__ double red[NT / 64];
    const int strip = blockIdx.x;
    const int inst = uniform_at(A.block_inst, strip);
    if (OP >= OP_CG_AP && OP != OP_PC_INIT && A.cg[inst].done) return;
    if ((OP == OP_CG_RESID) && !(A.cg[inst].limited || A.cg[inst].check)) return;
    if ((OP == OP_PC_RESID) && !A.cg[inst].check) return;
    dmg_strip<NT, OP>(A, strip, sm, red);
}

// ---------------------------------------------------------------------------------------
// node-wise CG vector kernels (same strips, so the partial slots line up)
// ---------------------------------------------------------------------------------------

enum VecOp : int {
    V_SETUP = 0,     // d = fixed ? vals : init ; lb = fixed ? vals : d_prev
    V_UPDATE_D = 1,  // projected: limited ? clamp(d + a p) : d += a p, r -= a Ap
    V_PRECOND = 2,   // projected: z = active ? 0 : r Minv ; partial rf.z
    V_UPDATE_P = 3,  // projected: p = active ? 0 : (restart ? z : beta p + z)
    V_FINISH = 4,    // d = min(max(d, lb), 1), fixed -> vals ; copy active
    V_PC_UPDATE = 5, // post-clamp: d += a p, r -= a Ap
    V_PC_PRECOND = 6,// post-clamp: z = r Minv ; partials r.r, r.z
    V_PC_UPDATE_P = 7,
    V_PC_FINISH = 8, // d = clamp(max(d, lb), 0, 1), fixed -> vals
    // iterations without a residual recomputation (neither a limited step nor a check):
    // V_UPDATE_D and V_PRECOND in one pass, the latter then skipped (V_PRECOND_CHK); the same
    // for the post-clamp pair.  The arithmetic and the partial sums are those of the separate
    // passes.
    V_UPDATE_D_PRE = 9,
    V_PRECOND_CHK = 10,
    V_PC_UPDATE_PRE = 11,
    V_PC_PRECOND_CHK = 12,
};

struct VecArgs {
    const Q4Inst *inst;
    const int32_t *block_inst;
    double *d, *r, *z, *p, *lb;
    const double *Ap, *minv, *init, *dprev, *vals;
    const uint8_t *fixed;
    uint8_t *active, *active_out;
    double *d_out;
    Q4CG *cg;
    double *part;
    int32_t n_tiles;
};

// Node-wise pass over the owned nodes of one strip.  Rows are taken U at a time: all the
// operands of the U rows are loaded before any of them is used, so each thread keeps U
// rows of loads in flight; partial sums still accumulate row by row.
template <int NT, int OP0>
__device__ __forceinline__ void vec_strip(const VecArgs &A, int strip, const Q4CG &S, double *red) {
    // the fused passes run as the update alone when the residual is recomputed in between
    constexpr bool FUSED = (OP0 == V_UPDATE_D_PRE || OP0 == V_PC_UPDATE_PRE);
    const bool fuse = FUSED && !(OP0 == V_UPDATE_D_PRE ? (S.limited || S.check) : S.check);
    constexpr int OP = OP0 == V_PRECOND_CHK ? V_PRECOND
                       : OP0 == V_PC_PRECOND_CHK ? V_PC_PRECOND
                       : OP0 == V_UPDATE_D_PRE ? V_UPDATE_D
                       : OP0 == V_PC_UPDATE_PRE ? V_PC_UPDATE
                                                : OP0;
    constexpr int U = 4;
    const Tile T = tile_of<NT>(A.inst, A.block_inst, strip);
    QI &P = inst_at(A.inst, T.inst);
    const int t = threadIdx.x;
    const int i = T.i0 - 1 + t;
    double p0 = 0.0, p1 = 0.0;
    if (t >= 1 && t <= NT - 2 && i < T.nxn) {
        for (int j0 = T.j0; j0 < T.jend; j0 += U) {
            int64_t n[U];
            bool ok[U];
            double x0[U], x1[U], x2[U], x3[U], xm[U];
            bool m0[U], m1[U], ma[U];
#pragma unroll
            for (int k = 0; k < U; ++k) {
                ok[k] = j0 + k < T.jend;
                n[k] = P.node_off + (int64_t)(j0 + k) * T.nxn + i;
                x0[k] = x1[k] = x2[k] = x3[k] = xm[k] = 0.0;
                m0[k] = m1[k] = ma[k] = false;
                if (!ok[k]) continue;
                const int64_t q = n[k];
                if (FUSED && fuse) {  // operands of the preconditioner half
                    xm[k] = A.minv[q];
                    ma[k] = (OP0 == V_UPDATE_D_PRE) ? A.active[q] != 0 : (A.fixed && A.fixed[q]);
                }
                if (OP == V_SETUP) {
                    m0[k] = A.fixed && A.fixed[q];
                    x0[k] = A.dprev[q];
                    if (A.fixed) x1[k] = A.vals[q];
                    if (A.init) x2[k] = A.init[q];
                } else if (OP == V_UPDATE_D || OP == V_PC_UPDATE) {
                    x0[k] = A.p[q];
                    x1[k] = A.d[q];
                    if (OP == V_UPDATE_D && S.limited) {
                        x2[k] = A.lb[q];
                    } else {
                        x2[k] = A.r[q];
                        x3[k] = A.Ap[q];
                    }
                } else if (OP == V_PRECOND) {
                    m0[k] = A.active[q] != 0;
                    x0[k] = A.r[q];
                    x1[k] = A.minv[q];
                } else if (OP == V_UPDATE_P || OP == V_PC_UPDATE_P) {
                    if (OP == V_UPDATE_P) m0[k] = A.active[q] != 0;
                    x0[k] = A.z[q];
                    x1[k] = A.p[q];
                } else if (OP == V_FINISH || OP == V_PC_FINISH) {
                    x0[k] = A.d[q];
                    x1[k] = A.lb[q];
                    m0[k] = A.fixed && A.fixed[q];
                    if (A.fixed) x2[k] = A.vals[q];
                    if (OP == V_FINISH && A.active_out) m1[k] = A.active[q] != 0;
                } else if (OP == V_PC_PRECOND) {
                    x0[k] = A.r[q];
                    m0[k] = A.fixed && A.fixed[q];
                    x1[k] = A.minv[q];
                }
            }
#pragma unroll
            for (int k = 0; k < U; ++k) {
                if (!ok[k]) continue;
                const int64_t q = n[k];
                if (OP == V_SETUP) {
                    A.d[q] = m0[k] ? x1[k] : (A.init ? x2[k] : x0[k]);
                    A.lb[q] = m0[k] ? x1[k] : x0[k];
                } else if (OP == V_UPDATE_D) {
                    if (S.limited) {
                        double dn = x1[k] + S.alpha * x0[k];
                        dn = dn >= x2[k] ? dn : x2[k];  // torch.maximum(d, lb)
                        dn = dn <= 1.0 ? dn : 1.0;      // clamp(max=1)
                        A.d[q] = dn;
                    } else {
                        A.d[q] = x1[k] + S.alpha * x0[k];
                        const double rn = x2[k] - S.alpha * x3[k];
                        A.r[q] = rn;
                        if (FUSED && fuse) {  // as V_PRECOND
                            const double rf = ma[k] ? 0.0 : rn;
                            double zz = rf * xm[k];
                            if (ma[k]) zz = 0.0;
                            A.z[q] = zz;
                            p0 += rf * zz;
                        }
                    }
                } else if (OP == V_PRECOND) {
                    const double rf = m0[k] ? 0.0 : x0[k];
                    double zz = rf * x1[k];
                    if (m0[k]) zz = 0.0;
                    A.z[q] = zz;
                    p0 += rf * zz;
                } else if (OP == V_UPDATE_P) {
                    double pn = S.restart ? x0[k] : x1[k] * S.beta + x0[k];
                    if (m0[k]) pn = 0.0;
                    A.p[q] = pn;
                } else if (OP == V_FINISH) {
                    double dn = x0[k];
                    dn = dn >= x1[k] ? dn : x1[k];
                    dn = dn <= 1.0 ? dn : 1.0;
                    if (m0[k]) dn = x2[k];
                    A.d_out[q] = dn;
                    if (A.active_out) A.active_out[q] = m1[k] ? 1 : 0;
                } else if (OP == V_PC_UPDATE) {
                    A.d[q] = x1[k] + S.alpha * x0[k];
                    const double rn = x2[k] - S.alpha * x3[k];
                    A.r[q] = rn;
                    if (FUSED && fuse) {  // as V_PC_PRECOND
                        const double zz = rn * (ma[k] ? 0.0 : xm[k]);
                        A.z[q] = zz;
                        p0 += rn * rn;
                        p1 += rn * zz;
                    }
                } else if (OP == V_PC_PRECOND) {
                    const double rr = x0[k];
                    const double zz = rr * (m0[k] ? 0.0 : x1[k]);
                    A.z[q] = zz;
                    p0 += rr * rr;
                    p1 += rr * zz;
                } else if (OP == V_PC_UPDATE_P) {
                    A.p[q] = x1[k] * S.beta + x0[k];
                } else if (OP == V_PC_FINISH) {
                    double dn = x0[k];
                    dn = dn >= x1[k] ? dn : x1[k];
                    dn = fmin(fmax(dn, 0.0), 1.0);
                    if (m0[k]) dn = x2[k];
                    A.d_out[q] = dn;
                }
            }
        }
    }
    if (OP == V_PRECOND || OP == V_PC_PRECOND || (FUSED && fuse)) {
        const double s0 = block_sum<NT>(p0, red);
        if (threadIdx.x == 0) A.part[strip] = s0;
        if (OP == V_PC_PRECOND || (OP0 == V_PC_UPDATE_PRE && fuse)) {
            const double s1 = block_sum<NT>(p1, red);
            if (threadIdx.x == 0) A.part[A.n_tiles + strip] = s1;
        }
    }
}

template <int NT, int OP>
__global__ __launch_bounds__(NT) void q4_vec_kernel(VecArgs A) {
    __shared__ double red[NT / 64];
    const int strip = blockIdx.x;
    const Q4CG &S = A.cg[uniform_at(A.block_inst, strip)];
    if (OP != V_SETUP && OP != V_FINISH && OP != V_PC_FINISH && S.done) return;
    if (OP == V_PRECOND_CHK && !(S.limited || S.check)) return;  // done by V_UPDATE_D_PRE
    if (OP == V_PC_PRECOND_CHK && !S.check) return;             // done by V_PC_UPDATE_PRE
    vec_strip<NT, OP>(A, strip, S, red);
}

template <int OP>
__global__ __launch_bounds__(256) void q4_scalar_kernel(const Q4Inst *inst, Q4CG *cg, const double *part,
                                                        int32_t n_tiles, int32_t check_every) {
    __shared__ double red[8];
    Q4CG &S = cg[blockIdx.x];
    if (OP != S_INIT && OP != S_PC_INIT && S.done) return;
    {
        QI &P = inst_at(inst, blockIdx.x);
        scalar_step<OP, 256>(P.tile_off, P.n_tiles, S, part, n_tiles, check_every, red);
    }
}

// ---------------------------------------------------------------------------------------
// persistent batched damage solve: every instance's whole CG in one launch
// ---------------------------------------------------------------------------------------
//
// A batched step is two launches: the fused explicit-step kernel over all instances, then
// this persistent kernel.  Its grid is one wave of co-resident blocks (sized from the
// occupancy query).  Instance k
// owns nblk[k] of them; block rank r of instance k runs strips r, r + nblk[k], ... of that
// instance, so the strip passes, per-strip partial sums and per-instance scalar steps are
// exactly those of the separate kernels, and the result is bitwise the same.  Kernel
// boundaries become barriers among the blocks of one instance (arrival counter in device
// memory, agent-scope fences); each block repeats the instance's scalar CG step itself on
// an LDS copy of the state, from the same partials in the same order.  Every wait is
// bounded by a wall-clock limit: a wait that runs out raises the error flag and falls
// through, so the launch always ends.

struct BatchArgs {
    MechArgs M;
    DmgArgs R;   // b, diag
    DmgArgs Dx;  // x = d (initial and true residuals)
    DmgArgs Dp;  // x = p (A p)
    VecArgs V;
    const int32_t *pb_inst, *pb_rank, *nblk;
    uint32_t *bar;  // [2 * n_inst]: arrivals, departures
    int32_t *err;
    int32_t *iters, *status;
    double *minv;
    double tol;
    int32_t max_iter, method;
    uint64_t spin_ticks;
};

__device__ __forceinline__ void inst_barrier(uint32_t *arrive, uint32_t target, int32_t *err, uint64_t spin_ticks) {
    // Every wave waits for its own stores and writes the L2 back (a fence by one thread
    // would only wait for that wave's stores), then one thread arrives and spins.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(arrive, 1u);
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
            const uint64_t t0 = wall_clock64();
            while (__hip_atomic_load(arrive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                __builtin_amdgcn_s_sleep(2);
                if (wall_clock64() - t0 > spin_ticks) {
                    atomicOr(reinterpret_cast<uint32_t *>(err), 1u);
                    break;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // drop stale L1/L2 lines of this CU/XCD
    }
    __syncthreads();
}

template <int NT>
__device__ __forceinline__ void minv_strip(const Q4Inst *inst, const int32_t *block_inst, int strip, double *minv) {
    const Tile T = tile_of<NT>(inst, block_inst, strip);
    QI &P = inst_at(inst, T.inst);
    const int t = threadIdx.x, i = T.i0 - 1 + t;
    if (t >= 1 && t <= NT - 2 && i < T.nxn) {
        for (int j = T.j0; j < T.jend; ++j) {
            const int64_t n = P.node_off + (int64_t)j * T.nxn + i;
            minv[n] = 1.0 / (minv[n] + 1e-30);  // damage_solver.py:2299
        }
    }
}

// The damage solve of instance k on its nb blocks (this block has rank `rank`): the strip
// passes and scalar steps of phast_q4_damage_solve in the same order, kernel boundaries
// replaced by `barrier`.  Ends with the solve's final pass; no barrier after it.
template <int NT, int METHOD, class BAR>
__device__ __forceinline__ void batch_damage(const BatchArgs &A, int k, int rank, int nb, double *sm, double *red,
                                             Q4CG &S, BAR &&barrier) {
    QI &P = inst_at(A.M.inst, k);
    const int n_tiles = A.V.n_tiles;
    const int check_every = METHOD == 0 ? 10 : 50;
#define PHAST_STRIPS for (int s_ = rank; s_ < P.n_tiles; s_ += nb)
    const int base = P.tile_off;

    // b, M_inv, initial iterate and bounds (the explicit step ran in the preceding launch)
    PHAST_STRIPS {
        dmg_strip<NT, OP_RHSDIAG>(A.R, base + s_, sm, red);
        minv_strip<NT>(A.V.inst, A.V.block_inst, base + s_, A.minv);
        vec_strip<NT, V_SETUP>(A.V, base + s_, S, red);
    }
    barrier();
    if (threadIdx.x == 0) {
        S.tol = A.tol;
        S.max_iter = A.max_iter;
    }
    VecArgs V3 = A.V;  // node-wise partials go to their own slots (no barrier before them)
    V3.part = A.V.part + (METHOD == 0 ? 3 : 2) * (int64_t)n_tiles;
    if (METHOD == 0) {
        PHAST_STRIPS dmg_strip<NT, OP_CG_INIT>(A.Dx, base + s_, sm, red);
        barrier();
        scalar_step<S_INIT, NT>(P.tile_off, P.n_tiles, S, A.V.part, n_tiles, check_every, red);
        barrier();
        while (!S.done) {
            PHAST_STRIPS dmg_strip<NT, OP_CG_AP>(A.Dp, base + s_, sm, red);
            barrier();
            scalar_step<S_A, NT>(P.tile_off, P.n_tiles, S, A.V.part, n_tiles, check_every, red);
            __syncthreads();
            if (S.done) break;
            PHAST_STRIPS vec_strip<NT, V_UPDATE_D>(A.V, base + s_, S, red);
            if (S.limited || S.check) {
                barrier();
                PHAST_STRIPS dmg_strip<NT, OP_CG_RESID>(A.Dx, base + s_, sm, red);
                barrier();
                scalar_step<S_B, NT>(P.tile_off, P.n_tiles, S, A.V.part, n_tiles, check_every, red);
                __syncthreads();
                if (S.done) break;
            } else {
                __syncthreads();
                if (threadIdx.x == 0) S.restart = 0;
                __syncthreads();
            }
            PHAST_STRIPS vec_strip<NT, V_PRECOND>(V3, base + s_, S, red);
            barrier();
            scalar_step<S_C, NT>(P.tile_off, P.n_tiles, S, V3.part, n_tiles, check_every, red);
            __syncthreads();
            if (S.done) break;
            PHAST_STRIPS vec_strip<NT, V_UPDATE_P>(A.V, base + s_, S, red);
            barrier();
        }
        PHAST_STRIPS vec_strip<NT, V_FINISH>(A.V, base + s_, S, red);
    } else {
        PHAST_STRIPS dmg_strip<NT, OP_PC_INIT>(A.Dx, base + s_, sm, red);
        barrier();
        scalar_step<S_PC_INIT, NT>(P.tile_off, P.n_tiles, S, A.V.part, n_tiles, check_every, red);
        barrier();
        while (!S.done) {
            PHAST_STRIPS dmg_strip<NT, OP_PC_AP>(A.Dp, base + s_, sm, red);
            barrier();
            scalar_step<S_PC_A, NT>(P.tile_off, P.n_tiles, S, A.V.part, n_tiles, check_every, red);
            __syncthreads();
            if (S.done) break;
            PHAST_STRIPS vec_strip<NT, V_PC_UPDATE>(A.V, base + s_, S, red);
            if (S.check) {
                barrier();
                PHAST_STRIPS dmg_strip<NT, OP_PC_RESID>(A.Dx, base + s_, sm, red);
            }
            PHAST_STRIPS vec_strip<NT, V_PC_PRECOND>(V3, base + s_, S, red);
            barrier();
            scalar_step<S_PC_C, NT>(P.tile_off, P.n_tiles, S, V3.part, n_tiles, check_every, red);
            __syncthreads();
            if (S.done) break;
            PHAST_STRIPS vec_strip<NT, V_PC_UPDATE_P>(A.V, base + s_, S, red);
            barrier();
        }
        PHAST_STRIPS vec_strip<NT, V_PC_FINISH>(A.V, base + s_, S, red);
    }
#undef PHAST_STRIPS
}

// Leaving a persistent launch: the last block of an instance out resets its counters, so
// the next launch (or graph replay) starts from zero.
__device__ __forceinline__ void batch_depart(uint32_t *arrive, int nb) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(arrive + 1, 1u) == (uint32_t)nb - 1) {  // last block of the instance out
            atomicExch(arrive, 0u);
            atomicExch(arrive + 1, 0u);
        }
    }
}

template <int NT, int METHOD>
__global__ __launch_bounds__(NT) void q4_batch_kernel(BatchArgs A) {
    __shared__ double sm[kDmgLds<NT>];
    __shared__ double red[8];
    __shared__ Q4CG S;
    const int k = uniform_at(A.pb_inst, blockIdx.x), rank = uniform_at(A.pb_rank, blockIdx.x),
              nb = uniform_at(A.nblk, k);
    uint32_t *arrive = A.bar + 2 * k;
    uint32_t nbar = 0;
    auto barrier = [&]() {
        ++nbar;
        inst_barrier(arrive, nbar * (uint32_t)nb, A.err, A.spin_ticks);
    };
    batch_damage<NT, METHOD>(A, k, rank, nb, sm, red, S, barrier);
    __syncthreads();
    if (threadIdx.x == 0 && rank == 0) {
        A.iters[k] = S.iter;
        A.status[k] = S.status;
    }
    batch_depart(arrive, nb);
}

// Several explicit steps in ONE launch.  Each instance runs its steps on its own blocks,
// joined only by the instance's barriers, so an instance whose damage solve needs few
// iterations goes on to its next step instead of waiting for the slowest instance; the
// launch ends when every instance has done n_steps.  Step s reads the buffers of A[s & 1]
// and writes the other set (A[1] is A[0] with inputs and outputs swapped), and every step
// is the explicit-step pass followed by the damage solve, exactly as one launch of
// q4_mech_kernel<STEP, HIST> and one of q4_batch_kernel: the same strips, partial sums and
// scalar steps, so the bits are those of n_steps separate steps.
struct BatchRunArgs {
    BatchArgs A[2];
    int32_t n_steps, n_inst;
    int32_t *iters, *status;  // [n_steps][n_inst]
};

template <int NT, int METHOD>
__global__ __launch_bounds__(NT) void q4_batch_run_kernel(BatchRunArgs R) {
    __shared__ double sm[kMechLds<NT> > kDmgLds<NT> ? kMechLds<NT> : kDmgLds<NT>];
    __shared__ double red[8];
    __shared__ Q4CG S;
    const BatchArgs &A0 = R.A[0];
    const int k = uniform_at(A0.pb_inst, blockIdx.x), rank = uniform_at(A0.pb_rank, blockIdx.x),
              nb = uniform_at(A0.nblk, k);
    QI &P = inst_at(A0.M.inst, k);
    uint32_t *arrive = A0.bar + 2 * k;
    uint32_t nbar = 0;
    auto barrier = [&]() {
        ++nbar;
        inst_barrier(arrive, nbar * (uint32_t)nb, A0.err, A0.spin_ticks);
    };
    for (int step = 0; step < R.n_steps; ++step) {
        const BatchArgs &A = R.A[step & 1];
        for (int s_ = rank; s_ < P.n_tiles; s_ += nb) mech_strip<NT, true, true>(A.M, P.tile_off + s_, sm);
        barrier();  // u1, v1, a1 and H_new of every strip before the damage passes read them
        batch_damage<NT, METHOD>(A, k, rank, nb, sm, red, S, barrier);
        __syncthreads();
        if (threadIdx.x == 0 && rank == 0) {
            R.iters[(int64_t)step * R.n_inst + k] = S.iter;
            R.status[(int64_t)step * R.n_inst + k] = S.status;
        }
        barrier();  // d_out of every strip before the next step's rows are staged
    }
    batch_depart(arrive, nb);
}

// ---------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------

static bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

template <int NT>
static int launch_mech(const phast_q4_plan *pl, MechArgs A, bool step, bool hist, hipStream_t st) {
    dim3 g(pl->n_tiles), b(NT);
    if (!step && !hist) hipLaunchKernelGGL((q4_mech_kernel<NT, false, false>), g, b, 0, st, A);
    else if (!step && hist) hipLaunchKernelGGL((q4_mech_kernel<NT, false, true>), g, b, 0, st, A);
    else if (step && !hist) hipLaunchKernelGGL((q4_mech_kernel<NT, true, false>), g, b, 0, st, A);
    else hipLaunchKernelGGL((q4_mech_kernel<NT, true, true>), g, b, 0, st, A);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

static int mech_dispatch(const phast_q4_plan *pl, MechArgs A, bool step, bool hist, hipStream_t st) {
    switch (pl->tx) {
        case 256: return launch_mech<256>(pl, A, step, hist, st);
        case 128: return launch_mech<128>(pl, A, step, hist, st);
        case 64: return launch_mech<64>(pl, A, step, hist, st);
    }
    set_error("unsupported strip width %d", pl->tx);
    return PHAST_ERR_UNSUPPORTED;
}

template <int OP>
static int launch_dmg(const phast_q4_plan *pl, const DmgArgs &A, hipStream_t st) {
    switch (pl->tx) {
        case 256: hipLaunchKernelGGL((q4_dmg_kernel<256, OP>), dim3(pl->n_tiles), dim3(256), 0, st, A); break;
        case 128: hipLaunchKernelGGL((q4_dmg_kernel<128, OP>), dim3(pl->n_tiles), dim3(128), 0, st, A); break;
        case 64: hipLaunchKernelGGL((q4_dmg_kernel<64, OP>), dim3(pl->n_tiles), dim3(64), 0, st, A); break;
        default: set_error("unsupported strip width %d", pl->tx); return PHAST_ERR_UNSUPPORTED;
    }
    PHAST_LAUNCHED();
    return PHAST_OK;
}
template <int OP>
static int launch_vec(const phast_q4_plan *pl, const VecArgs &A, hipStream_t st) {
    switch (pl->tx) {
        case 256: hipLaunchKernelGGL((q4_vec_kernel<256, OP>), dim3(pl->n_tiles), dim3(256), 0, st, A); break;
        case 128: hipLaunchKernelGGL((q4_vec_kernel<128, OP>), dim3(pl->n_tiles), dim3(128), 0, st, A); break;
        case 64: hipLaunchKernelGGL((q4_vec_kernel<64, OP>), dim3(pl->n_tiles), dim3(64), 0, st, A); break;
        default: set_error("unsupported strip width %d", pl->tx); return PHAST_ERR_UNSUPPORTED;
    }
    PHAST_LAUNCHED();
    return PHAST_OK;
}
template <int OP>
static int launch_scalar(const phast_q4_plan *pl, Q4CG *cg, double *part, int check_every, hipStream_t st) {
    hipLaunchKernelGGL((q4_scalar_kernel<OP>), dim3(pl->n_inst), dim3(256), 0, st, pl->d_inst, cg, part,
                       pl->n_tiles, check_every);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

#define PHAST_TRY(x)                 \
    do {                             \
        int rc_ = (x);               \
        if (rc_ != PHAST_OK) return rc_; \
    } while (0)

}  // namespace phast

using namespace phast;

extern "C" int phast_q4_plan_create(const phast_q4_desc *desc, int32_t n_inst, int32_t tile_x, int32_t tile_y,
                                    phast_q4_plan **out) {
    PHAST_REQUIRE(desc && out && n_inst >= 1, "phast_q4_plan_create: null argument or n_inst < 1");
    if (tile_x == 0 && tile_y == 0) {
        tile_x = 128;
        tile_y = 8;
    }
    PHAST_REQUIRE((tile_x == 64 || tile_x == 128 || tile_x == 256) && tile_y >= 1 && tile_y <= 4096,
                  "unsupported strip %dx%d (width 64, 128 or 256 threads; 1..4096 rows)", tile_x, tile_y);
    auto *pl = new phast_q4_plan();
    pl->n_inst = n_inst;
    pl->tx = tile_x;
    pl->ty = tile_y;
    std::vector<int32_t> block_inst;
    int64_t node_off = 0, cell_off = 0;
    for (int k = 0; k < n_inst; ++k) {
        const phast_q4_desc &D = desc[k];
        if (D.nx < 1 || D.ny < 1 || D.split < 0 || D.split > 1 || D.stress_split < 0 || D.stress_split > 1 ||
            !(D.hx > 0) || !(D.hy > 0)) {
            delete pl;
            set_error("instance %d: bad nx/ny/split", k);
            return PHAST_ERR_ARG;
        }
        Q4Inst I;
        std::memset(&I, 0, sizeof(I));
        I.nx = D.nx;
        I.ny = D.ny;
        I.tiles_x = (int32_t)ceil_div(D.nx + 1, tile_x - 2);
        I.tiles_y = (int32_t)ceil_div(D.ny + 1, tile_y);
        I.ky = tile_y;
        I.node_off = node_off;
        I.cell_off = cell_off;
        I.tile_off = (int32_t)block_inst.size();
        I.n_tiles = I.tiles_x * I.tiles_y;
        for (int q = 0; q < 4; ++q) {
            double dl = 0.0;
            for (int a = 0; a < 4; ++a) {
                I.gx[q * 4 + a] = D.gradN[q][a][0];
                I.gy[q * 4 + a] = D.gradN[q][a][1];
                I.Nq[q * 4 + a] = D.N[q][a];
            }
            I.w[q] = D.wdetJ[q];
            (void)dl;
        }
        for (int a = 0; a < 4; ++a) {  // damage_solver.py:1397-1402
            double s = 0.0;
            for (int q = 0; q < 4; ++q)
                s += D.wdetJ[q] * (D.gradN[q][a][0] * D.gradN[q][a][0] + D.gradN[q][a][1] * D.gradN[q][a][1]);
            I.dlap[a] = s;
        }
        {  // rectangle constants (wdetJ is uniform: hx hy / 4)
            const double a = 1.0 / std::sqrt(3.0);
            I.al = 0.5 * (1.0 + a);
            I.be = 0.5 * (1.0 - a);
            const double ihx = 1.0 / D.hx, ihy = 1.0 / D.hy, w = D.wdetJ[0];
            I.xa = I.al * ihx;
            I.xb = I.be * ihx;
            I.ya = I.al * ihy;
            I.yb = I.be * ihy;
            I.fxa = w * I.xa;
            I.fxb = w * I.xb;
            I.fya = w * I.ya;
            I.fyb = w * I.yb;
            I.wal = w * I.al;
            I.wbe = w * I.be;
            I.wq = w;
            I.ome = 1.0 - D.eta_residual;
        }
        I.lam = D.lam;
        I.mu = D.mu;
        I.kappa = D.kappa;
        I.ezz_coef = D.plane_stress ? (-D.nu / (1.0 - D.nu)) : 0.0;
        I.eta = D.eta_residual;
        I.Gc_l0 = D.Gc * D.l0;
        I.Gc_over_l0 = D.Gc / D.l0;
        I.l0 = D.l0;
        I.dt = D.dt;
        I.split = D.split;
        I.ssplit = D.stress_split;
        for (int t = 0; t < I.n_tiles; ++t) block_inst.push_back(k);
        node_off += (int64_t)(D.nx + 1) * (D.ny + 1);
        cell_off += (int64_t)D.nx * D.ny;
        pl->h_inst.push_back(I);
    }
    pl->n_tiles = (int32_t)block_inst.size();
    pl->n_nodes = node_off;
    pl->n_cells = cell_off;
    if (hipMalloc(&pl->d_inst, sizeof(Q4Inst) * n_inst) != hipSuccess ||
        hipMalloc(&pl->d_block_inst, sizeof(int32_t) * block_inst.size()) != hipSuccess ||
        hipMemcpy(pl->d_inst, pl->h_inst.data(), sizeof(Q4Inst) * n_inst, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pl->d_block_inst, block_inst.data(), sizeof(int32_t) * block_inst.size(),
                  hipMemcpyHostToDevice) != hipSuccess) {
        set_error("phast_q4_plan_create: device allocation failed");
        phast_q4_plan_destroy(pl);
        return PHAST_ERR_HIP;
    }
    *out = pl;
    return PHAST_OK;
}

extern "C" int phast_q4_plan_destroy(phast_q4_plan *pl) {
    if (!pl) return PHAST_OK;
    (void)hipFree(pl->d_inst);
    (void)hipFree(pl->d_block_inst);
    (void)hipFree(pl->r);
    (void)hipFree(pl->p);
    (void)hipFree(pl->p2);
    (void)hipFree(pl->z);
    (void)hipFree(pl->Ap);
    (void)hipFree(pl->b);
    (void)hipFree(pl->minv);
    (void)hipFree(pl->lb);
    (void)hipFree(pl->part);
    (void)hipFree(pl->active);
    (void)hipFree(pl->cg);
    if (pl->h_cg) (void)hipHostFree(pl->h_cg);
    for (int m = 0; m < 4; ++m) {
        (void)hipFree(pl->b_pb_inst[m]);
        (void)hipFree(pl->b_pb_rank[m]);
        (void)hipFree(pl->b_nblk[m]);
    }
    (void)hipFree(pl->b_bar);
    (void)hipFree(pl->b_err);
    delete pl;
    return PHAST_OK;
}

extern "C" int phast_q4_plan_sizes(const phast_q4_plan *pl, int64_t *n_nodes, int64_t *n_cells) {
    PHAST_REQUIRE(pl, "null plan");
    if (n_nodes) *n_nodes = pl->n_nodes;
    if (n_cells) *n_cells = pl->n_cells;
    return PHAST_OK;
}

extern "C" int phast_q4_plan_offsets(const phast_q4_plan *pl, int64_t *node_off, int64_t *cell_off) {
    PHAST_REQUIRE(pl, "null plan");
    for (int k = 0; k < pl->n_inst; ++k) {
        if (node_off) node_off[k] = pl->h_inst[k].node_off;
        if (cell_off) cell_off[k] = pl->h_inst[k].cell_off;
    }
    return PHAST_OK;
}

static DmgArgs dmg_base(const phast_q4_plan *pl) {
    DmgArgs A;
    std::memset(&A, 0, sizeof(A));
    A.inst = pl->d_inst;
    A.block_inst = pl->d_block_inst;
    A.n_tiles = pl->n_tiles;
    return A;
}

extern "C" int phast_q4_lumped_mass(const phast_q4_plan *pl, double *M, hipStream_t st) {
    PHAST_REQUIRE(pl && M, "phast_q4_lumped_mass: null argument");
    DmgArgs A = dmg_base(pl);
    A.out = M;
    return launch_dmg<OP_MASS>(pl, A, st);
}

extern "C" int phast_q4_elem_to_node(const phast_q4_plan *pl, const double *field_q, const double *M,
                                     double *out, hipStream_t st) {
    PHAST_REQUIRE(pl && field_q && M && out, "phast_q4_elem_to_node: null argument");
    DmgArgs A = dmg_base(pl);
    A.x = field_q;
    A.M = M;
    A.out = out;
    return launch_dmg<OP_E2N>(pl, A, st);
}

extern "C" int phast_q4_internal_force(const phast_q4_plan *pl, const double *u, const double *d, double *f,
                                       hipStream_t st) {
    PHAST_REQUIRE(pl && u && d && f, "phast_q4_internal_force: null argument");
    PHAST_REQUIRE(aligned16(u) && aligned16(f), "u and f must be 16-byte aligned (N,2) float64 arrays");
    MechArgs A;
    std::memset(&A, 0, sizeof(A));
    A.inst = pl->d_inst;
    A.block_inst = pl->d_block_inst;
    A.u = reinterpret_cast<const double2 *>(u);
    A.d = d;
    A.f = reinterpret_cast<double2 *>(f);
    A.fscale = 1.0;
    return mech_dispatch(pl, A, false, false, st);
}

extern "C" int phast_q4_damping_force(const phast_q4_plan *pl, const double *v, const double *d, double beta,
                                      double *f, hipStream_t st) {
    PHAST_REQUIRE(pl && v && d && f, "phast_q4_damping_force: null argument");
    PHAST_REQUIRE(aligned16(v) && aligned16(f), "v and f must be 16-byte aligned (N,2) float64 arrays");
    MechArgs A;
    std::memset(&A, 0, sizeof(A));
    A.inst = pl->d_inst;
    A.block_inst = pl->d_block_inst;
    A.u = reinterpret_cast<const double2 *>(v);
    A.d = d;
    A.f = reinterpret_cast<double2 *>(f);
    A.fscale = beta;
    return mech_dispatch(pl, A, false, false, st);
}

extern "C" int phast_q4_explicit_step(const phast_q4_plan *pl, const double *u, const double *v, const double *a,
                                      const double *d, const double *f_ext, const double *f_damp, const double *Minv,
                                      const uint8_t *bc_mask, const double *bc_vals, const double *H_old,
                                      double *u1, double *v1, double *a1, double *H_new, double *psi,
                                      hipStream_t st) {
    PHAST_REQUIRE(pl && u && v && a && d && Minv && u1 && v1 && a1, "phast_q4_explicit_step: null argument");
    PHAST_REQUIRE(!bc_mask == !bc_vals, "bc_mask and bc_vals must be given together");
    PHAST_REQUIRE(!H_old == !H_new, "H_old and H_new must be given together");
    PHAST_REQUIRE(aligned16(u) && aligned16(v) && aligned16(a) && aligned16(u1) && aligned16(v1) &&
                      aligned16(a1) && aligned16(f_ext) && aligned16(f_damp) && aligned16(H_old) &&
                      aligned16(H_new) && aligned16(psi),
                  "state arrays must be 16-byte aligned");
    PHAST_REQUIRE(u1 != u && v1 != v && a1 != a && u1 != v && u1 != a && (!H_new || H_new != H_old),
                  "explicit step outputs must not alias its inputs");
    PHAST_REQUIRE(!psi || H_old, "psi output requires the history arrays");
    MechArgs A;
    std::memset(&A, 0, sizeof(A));
    A.inst = pl->d_inst;
    A.block_inst = pl->d_block_inst;
    A.u = reinterpret_cast<const double2 *>(u);
    A.v = reinterpret_cast<const double2 *>(v);
    A.a = reinterpret_cast<const double2 *>(a);
    A.d = d;
    A.fext = reinterpret_cast<const double2 *>(f_ext);
    A.fdamp = reinterpret_cast<const double2 *>(f_damp);
    A.minv = Minv;
    A.bcm = bc_mask;
    A.bcv = bc_vals;
    A.Hold = H_old;
    A.Hnew = H_new;
    A.psi = psi;
    A.u1 = reinterpret_cast<double2 *>(u1);
    A.v1 = reinterpret_cast<double2 *>(v1);
    A.a1 = reinterpret_cast<double2 *>(a1);
    return mech_dispatch(pl, A, true, H_old != nullptr, st);
}

extern "C" int phast_q4_strain(const phast_q4_plan *pl, const double *u, double *eps, hipStream_t st) {
    PHAST_REQUIRE(pl && u && eps && aligned16(u), "phast_q4_strain: bad argument");
    CellArgs A;
    std::memset(&A, 0, sizeof(A));
    A.inst = pl->d_inst;
    A.n_inst = pl->n_inst;
    A.u = reinterpret_cast<const double2 *>(u);
    A.eps = eps;
    A.n_cells = pl->n_cells;
    hipLaunchKernelGGL(q4_cell_kernel, dim3((unsigned)ceil_div(pl->n_cells, 256)), dim3(256), 0, st, A);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

extern "C" int phast_q4_driving_force(const phast_q4_plan *pl, const double *u, double *psi, const double *H_old,
                                      double *H_new, hipStream_t st) {
    PHAST_REQUIRE(pl && u && aligned16(u) && (psi || H_new), "phast_q4_driving_force: bad argument");
    PHAST_REQUIRE(!H_old == !H_new, "H_old and H_new must be given together");
    CellArgs A;
    std::memset(&A, 0, sizeof(A));
    A.inst = pl->d_inst;
    A.n_inst = pl->n_inst;
    A.u = reinterpret_cast<const double2 *>(u);
    A.psi = psi;
    A.Hold = H_old;
    A.Hnew = H_new;
    A.n_cells = pl->n_cells;
    hipLaunchKernelGGL(q4_cell_kernel, dim3((unsigned)ceil_div(pl->n_cells, 256)), dim3(256), 0, st, A);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

extern "C" int phast_history_max(int64_t n, const double *H_old, const double *psi, double *H_new,
                                 hipStream_t st) {
    PHAST_REQUIRE(n >= 0 && (n == 0 || (H_old && psi && H_new)), "phast_history_max: bad argument");
    if (n == 0) return PHAST_OK;
    hipLaunchKernelGGL(history_max_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, n, H_old, psi, H_new);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

extern "C" int phast_q4_damage_rhs_diag(const phast_q4_plan *pl, const double *H, const double *Gc_cell,
                                        double *b, double *diag, hipStream_t st) {
    PHAST_REQUIRE(pl && H && b && diag, "phast_q4_damage_rhs_diag: null argument");
    DmgArgs A = dmg_base(pl);
    A.H = H;
    A.Gc = Gc_cell;
    A.out = b;
    A.out2 = diag;
    return launch_dmg<OP_RHSDIAG>(pl, A, st);
}

extern "C" int phast_q4_damage_apply(const phast_q4_plan *pl, const double *H, const double *Gc_cell,
                                     const double *x, double *y, hipStream_t st) {
    PHAST_REQUIRE(pl && H && x && y, "phast_q4_damage_apply: null argument");
    DmgArgs A = dmg_base(pl);
    A.H = H;
    A.Gc = Gc_cell;
    A.x = x;
    A.out = y;
    return launch_dmg<OP_APPLY>(pl, A, st);
}

__global__ void minv_kernel(int64_t n, double *diag) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) diag[i] = 1.0 / (diag[i] + 1e-30);  // damage_solver.py:2299
}

static int ensure_work(phast_q4_plan *pl) {
    if (pl->r) return PHAST_OK;
    const size_t nb = sizeof(double) * (size_t)pl->n_nodes;
    PHAST_HIP(hipMalloc(&pl->r, nb));
    PHAST_HIP(hipMalloc(&pl->p, nb));
    PHAST_HIP(hipMalloc(&pl->p2, nb));
    PHAST_HIP(hipMalloc(&pl->z, nb));
    PHAST_HIP(hipMalloc(&pl->Ap, nb));
    PHAST_HIP(hipMalloc(&pl->b, nb));
    PHAST_HIP(hipMalloc(&pl->minv, nb));
    PHAST_HIP(hipMalloc(&pl->lb, nb));
    PHAST_HIP(hipMalloc(&pl->part, sizeof(double) * kParts * (size_t)pl->n_tiles));
    PHAST_HIP(hipMalloc(&pl->active, (size_t)pl->n_nodes));
    PHAST_HIP(hipMalloc(&pl->cg, sizeof(Q4CG) * pl->n_inst));
    PHAST_HIP(hipHostMalloc(&pl->h_cg, sizeof(Q4CG) * pl->n_inst, hipHostMallocDefault));
    return PHAST_OK;
}

// Scratch of one damage solve: the CG vectors, the strip partials, the active mask, the
// per-instance CG state on the device and its host copy.
struct Q4SolveScratch {
    double *r, *p, *p2, *z, *Ap, *b, *minv, *lb, *part;
    uint8_t *active;
    Q4CG *cg, *h_cg;
};

static int damage_solve_impl(phast_q4_plan *pl, const Q4SolveScratch &W, int32_t method, const double *H,
                             const double *Gc_cell, const double *d_prev, const double *d_init, const uint8_t *fixed,
                             const double *fixed_vals, double tol, int32_t max_iter, double *d_out,
                             uint8_t *active_out, int32_t *iters, int32_t *status, hipStream_t st);

extern "C" int phast_q4_damage_solve(phast_q4_plan *pl, int32_t method, const double *H, const double *Gc_cell,
                                     const double *d_prev, const double *d_init, const uint8_t *fixed,
                                     const double *fixed_vals, double tol, int32_t max_iter, double *d_out,
                                     uint8_t *active_out, int32_t *iters, int32_t *status, hipStream_t st) {
    PHAST_REQUIRE(pl && H && d_prev && d_out, "phast_q4_damage_solve: null argument");
    PHAST_REQUIRE(method == 0 || method == 1, "method must be 0 (projected) or 1 (post-clamp)");
    PHAST_REQUIRE(!fixed == !fixed_vals, "fixed and fixed_vals must be given together");
    PHAST_REQUIRE(tol >= 0.0 && max_iter >= 0, "tol and max_iter must be non-negative");
    PHAST_TRY(ensure_work(pl));
    const Q4SolveScratch W{pl->r, pl->p, pl->p2, pl->z, pl->Ap, pl->b, pl->minv, pl->lb, pl->part,
                           pl->active, pl->cg, pl->h_cg};
    return damage_solve_impl(pl, W, method, H, Gc_cell, d_prev, d_init, fixed, fixed_vals, tol, max_iter, d_out,
                             active_out, iters, status, st);
}

// ---------------------------------------------------------------------------------------
// the damage solve that allocates nothing: its scratch is the file-scope array `scratch`
// ---------------------------------------------------------------------------------------

constexpr int64_t kNoAllocNodes = int64_t(1) << 19;  // nodes of all instances of the plan
constexpr int64_t kNoAllocTiles = int64_t(1) << 16;  // strips of the plan
constexpr int kNoAllocInst = 64;                     // instances of the plan
constexpr int64_t kNoAllocCgWords = (int64_t)((sizeof(Q4CG) + 7) / 8);
constexpr int64_t kNoAllocWords =
    8 * kNoAllocNodes + kParts * kNoAllocTiles + kNoAllocNodes / 8 + kNoAllocInst * kNoAllocCgWords;
constexpr int kNoAllocDevices = 64;

// 8 node vectors | kParts x strips partials | active mask (1 byte per node) | CG states
static __device__ __align__(256) double scratch[kNoAllocWords];

static std::mutex g_scratch_mutex;
static double *g_scratch_ptr[kNoAllocDevices];           // device address of `scratch`, per device
static Q4CG g_scratch_host_cg[kNoAllocDevices][kNoAllocInst];  // host copy of the CG states, per device

extern "C" int phast_q4_damage_solve_noalloc(phast_q4_plan *pl, int32_t method, const double *H,
                                             const double *Gc_cell, const double *d_prev, const double *d_init,
                                             const uint8_t *fixed, const double *fixed_vals, double tol,
                                             int32_t max_iter, double *d_out, uint8_t *active_out, int32_t *iters,
                                             int32_t *status, hipStream_t st) {
    PHAST_REQUIRE(pl && H && d_prev && d_out, "phast_q4_damage_solve_noalloc: null argument");
    PHAST_REQUIRE(method == 0 || method == 1, "method must be 0 (projected) or 1 (post-clamp)");
    PHAST_REQUIRE(!fixed == !fixed_vals, "fixed and fixed_vals must be given together");
    PHAST_REQUIRE(tol >= 0.0 && max_iter >= 0, "tol and max_iter must be non-negative");
    PHAST_REQUIRE(pl->n_nodes <= kNoAllocNodes && pl->n_tiles <= kNoAllocTiles && pl->n_inst <= kNoAllocInst,
                  "phast_q4_damage_solve_noalloc: plan too large for the static scratch (%lld nodes, %d strips, "
                  "%d instances; at most %lld, %lld, %d)", (long long)pl->n_nodes, pl->n_tiles, pl->n_inst,
                  (long long)kNoAllocNodes, (long long)kNoAllocTiles, kNoAllocInst);
    int dev = 0;
    PHAST_HIP(hipGetDevice(&dev));
    PHAST_REQUIRE(dev >= 0 && dev < kNoAllocDevices, "phast_q4_damage_solve_noalloc: device %d out of range", dev);
    double *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mutex);
        if (!g_scratch_ptr[dev]) {  // first call on this device (it is current): resolve its copy
            void *ptr = nullptr;
            PHAST_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
            g_scratch_ptr[dev] = static_cast<double *>(ptr);
        }
        base = g_scratch_ptr[dev];
    }
    const int64_t N = kNoAllocNodes;
    double *parts = base + 8 * N;
    uint8_t *active = reinterpret_cast<uint8_t *>(parts + kParts * kNoAllocTiles);
    Q4CG *cg = reinterpret_cast<Q4CG *>(parts + kParts * kNoAllocTiles + N / 8);
    const Q4SolveScratch W{base, base + N, base + 2 * N, base + 3 * N, base + 4 * N, base + 5 * N, base + 6 * N,
                           base + 7 * N, parts, active, cg, g_scratch_host_cg[dev]};
    return damage_solve_impl(pl, W, method, H, Gc_cell, d_prev, d_init, fixed, fixed_vals, tol, max_iter, d_out,
                             active_out, iters, status, st);
}

static int damage_solve_impl(phast_q4_plan *pl, const Q4SolveScratch &W, int32_t method, const double *H,
                             const double *Gc_cell, const double *d_prev, const double *d_init, const uint8_t *fixed,
                             const double *fixed_vals, double tol, int32_t max_iter, double *d_out,
                             uint8_t *active_out, int32_t *iters, int32_t *status, hipStream_t st) {
    // per-instance initial state: tolerance and iteration cap
    for (int k = 0; k < pl->n_inst; ++k) {
        Q4CG &S = W.h_cg[k];
        std::memset(&S, 0, sizeof(S));
        S.tol = tol;
        S.max_iter = max_iter;
    }
    PHAST_HIP(hipMemcpyAsync(W.cg, W.h_cg, sizeof(Q4CG) * pl->n_inst, hipMemcpyHostToDevice, st));

    DmgArgs D = dmg_base(pl);
    D.H = H;
    D.Gc = Gc_cell;
    D.d = d_out;
    D.r = W.r;
    D.z = W.z;
    D.p = W.p;
    D.Ap = W.Ap;
    D.b = W.b;
    D.minv = W.minv;
    D.lb = W.lb;
    D.fixed = fixed;
    D.active = W.active;
    D.cg = W.cg;
    D.part = W.part;

    VecArgs V;
    std::memset(&V, 0, sizeof(V));
    V.inst = pl->d_inst;
    V.block_inst = pl->d_block_inst;
    V.d = d_out;
    V.r = W.r;
    V.z = W.z;
    V.p = W.p;
    V.lb = W.lb;
    V.Ap = W.Ap;
    V.minv = W.minv;
    V.init = d_init;
    V.dprev = d_prev;
    V.vals = fixed_vals;
    V.fixed = fixed;
    V.active = W.active;
    V.active_out = active_out;
    V.d_out = d_out;
    V.cg = W.cg;
    V.part = W.part;
    V.n_tiles = pl->n_tiles;

    // b, diag -> M_inv; d = init, lb = d_prev (with pf_dirichlet pins)
    {
        DmgArgs R = dmg_base(pl);
        R.H = H;
        R.Gc = Gc_cell;
        R.out = W.b;
        R.out2 = W.minv;
        PHAST_TRY(launch_dmg<OP_RHSDIAG>(pl, R, st));
        hipLaunchKernelGGL(minv_kernel, dim3((unsigned)ceil_div(pl->n_nodes, 256)), dim3(256), 0, st,
                           pl->n_nodes, W.minv);
        PHAST_LAUNCHED();
    }
    PHAST_TRY(launch_vec<V_SETUP>(pl, V, st));
    const int poll = 4;
    int it = 0;
    if (method == 0) {
        D.x = d_out;
        PHAST_TRY(launch_dmg<OP_CG_INIT>(pl, D, st));
        PHAST_TRY(launch_scalar<S_INIT>(pl, W.cg, W.part, 10, st));
        DmgArgs Dr = D;
        Dr.x = d_out;
        // the direction update of iteration k runs inside the A p pass of iteration k+1
        // (OP_CG_APF), alternating between the two direction buffers
        double *pc = W.p, *pn = W.p2;
        while (it < max_iter) {
            D.x = pc;
            if (it == 0) {
                PHAST_TRY(launch_dmg<OP_CG_AP>(pl, D, st));
            } else {
                D.pout = pn;
                PHAST_TRY(launch_dmg<OP_CG_APF>(pl, D, st));
                std::swap(pc, pn);
            }
            PHAST_TRY(launch_scalar<S_A>(pl, W.cg, W.part, 10, st));
            V.p = pc;
            PHAST_TRY(launch_vec<V_UPDATE_D_PRE>(pl, V, st));
            PHAST_TRY(launch_dmg<OP_CG_RESID>(pl, Dr, st));
            PHAST_TRY(launch_scalar<S_B>(pl, W.cg, W.part, 10, st));
            PHAST_TRY(launch_vec<V_PRECOND_CHK>(pl, V, st));
            PHAST_TRY(launch_scalar<S_C>(pl, W.cg, W.part, 10, st));
            ++it;
            if (it % poll == 0 || it == max_iter) {
                PHAST_HIP(hipMemcpyAsync(W.h_cg, W.cg, sizeof(Q4CG) * pl->n_inst, hipMemcpyDeviceToHost, st));
                PHAST_HIP(hipStreamSynchronize(st));
                bool all = true;
                for (int k = 0; k < pl->n_inst; ++k) all = all && W.h_cg[k].done;
                if (all) break;
            }
        }
        PHAST_TRY(launch_vec<V_FINISH>(pl, V, st));
    } else {
        D.x = d_out;
        PHAST_TRY(launch_dmg<OP_PC_INIT>(pl, D, st));
        PHAST_TRY(launch_scalar<S_PC_INIT>(pl, W.cg, W.part, 50, st));
        DmgArgs Dr = D;
        Dr.x = d_out;
        double *pc = W.p, *pn = W.p2;
        while (it < max_iter) {
            D.x = pc;
            if (it == 0) {
                PHAST_TRY(launch_dmg<OP_PC_AP>(pl, D, st));
            } else {
                D.pout = pn;
                PHAST_TRY(launch_dmg<OP_PC_APF>(pl, D, st));
                std::swap(pc, pn);
            }
            PHAST_TRY(launch_scalar<S_PC_A>(pl, W.cg, W.part, 50, st));
            V.p = pc;
            PHAST_TRY(launch_vec<V_PC_UPDATE_PRE>(pl, V, st));
            PHAST_TRY(launch_dmg<OP_PC_RESID>(pl, Dr, st));
            PHAST_TRY(launch_vec<V_PC_PRECOND_CHK>(pl, V, st));
            PHAST_TRY(launch_scalar<S_PC_C>(pl, W.cg, W.part, 50, st));
            ++it;
            if (it % poll == 0 || it == max_iter) {
                PHAST_HIP(hipMemcpyAsync(W.h_cg, W.cg, sizeof(Q4CG) * pl->n_inst, hipMemcpyDeviceToHost, st));
                PHAST_HIP(hipStreamSynchronize(st));
                bool all = true;
                for (int k = 0; k < pl->n_inst; ++k) all = all && W.h_cg[k].done;
                if (all) break;
            }
        }
        PHAST_TRY(launch_vec<V_PC_FINISH>(pl, V, st));
    }
    PHAST_HIP(hipMemcpyAsync(W.h_cg, W.cg, sizeof(Q4CG) * pl->n_inst, hipMemcpyDeviceToHost, st));
    PHAST_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < pl->n_inst; ++k) {
        const Q4CG &S = W.h_cg[k];
        int32_t its = S.iter, stt = S.status;
        if (!S.done) {  // loop ended by the cap without the flag (max_iter == 0 handled in init)
            stt = PHAST_CG_MAXITER;
        }
        if (iters) iters[k] = its;
        if (status) status[k] = stt;
    }
    pl->last_iters = it;
    return PHAST_OK;
}

// ---------------------------------------------------------------------------------------
// persistent batched step (host side)
// ---------------------------------------------------------------------------------------

namespace phast {

template <int NT, int M>
static int batch_occupancy(int *blocks_per_cu) {
    const hipError_t e = M < 2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, q4_batch_kernel<NT, M & 1>,
                                                                              NT, 0)
                               : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                     blocks_per_cu, q4_batch_run_kernel<NT, M & 1>, NT, 0);
    return e == hipSuccess ? PHAST_OK : PHAST_ERR_HIP;
}

template <int NT>
static int batch_occupancy_nt(int m, int *bpc) {
    switch (m) {
        case 0: return batch_occupancy<NT, 0>(bpc);
        case 1: return batch_occupancy<NT, 1>(bpc);
        case 2: return batch_occupancy<NT, 2>(bpc);
        case 3: return batch_occupancy<NT, 3>(bpc);
    }
    return PHAST_ERR_UNSUPPORTED;
}

static int batch_occupancy_of(int nt, int m, int *bpc) {
    switch (nt) {
        case 256: return batch_occupancy_nt<256>(m, bpc);
        case 128: return batch_occupancy_nt<128>(m, bpc);
        case 64: return batch_occupancy_nt<64>(m, bpc);
    }
    return PHAST_ERR_UNSUPPORTED;
}

template <int NT>
static void batch_launch(const phast_q4_plan *pl, int m, const BatchArgs &A, hipStream_t st) {
    if (m == 0) hipLaunchKernelGGL((q4_batch_kernel<NT, 0>), dim3(pl->b_grid[0]), dim3(NT), 0, st, A);
    else hipLaunchKernelGGL((q4_batch_kernel<NT, 1>), dim3(pl->b_grid[1]), dim3(NT), 0, st, A);
}

template <int NT>
static void batch_run_launch(const phast_q4_plan *pl, int m, const BatchRunArgs &R, hipStream_t st) {
    if (m == 0) hipLaunchKernelGGL((q4_batch_run_kernel<NT, 0>), dim3(pl->b_grid[2]), dim3(NT), 0, st, R);
    else hipLaunchKernelGGL((q4_batch_run_kernel<NT, 1>), dim3(pl->b_grid[3]), dim3(NT), 0, st, R);
}

}  // namespace phast

using namespace phast;

extern "C" int phast_q4_batch_prepare(phast_q4_plan *pl, int32_t *grid_out) {
    PHAST_REQUIRE(pl, "phast_q4_batch_prepare: null plan");
    PHAST_TRY(ensure_work(pl));
    int dev = 0, cus = 0, rate_khz = 0;
    PHAST_HIP(hipGetDevice(&dev));
    PHAST_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    PHAST_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev));
    pl->b_spin_ticks = (uint64_t)(rate_khz > 0 ? rate_khz : 100000) * 1000ull * 2ull;  // 2 s per wait
    for (int m = 0; m < 4; ++m) {
        int bpc = 0;
        const int rc = batch_occupancy_of(pl->tx, m, &bpc);
        if (rc != PHAST_OK) {
            set_error("occupancy query failed for the batched step kernel (strip width %d)", pl->tx);
            return rc;
        }
        // 7/8 of the occupancy bound: the grid stays co-resident even when a few CUs are
        // unavailable to this queue
        const int64_t gmax = std::max<int64_t>((int64_t)bpc * cus * 7 / 8, 1);
        PHAST_REQUIRE(gmax >= pl->n_inst, "%d instances exceed the %lld co-resident blocks of the batched step",
                      pl->n_inst, (long long)gmax);
        // blocks per instance in proportion to its strips, at least one, at most one per strip
        std::vector<int32_t> nb(pl->n_inst);
        int64_t total = 0;
        for (int k = 0; k < pl->n_inst; ++k) {
            const int64_t t = pl->h_inst[k].n_tiles;
            nb[k] = (int32_t)std::max<int64_t>(1, std::min<int64_t>(t, gmax * t / pl->n_tiles));
            total += nb[k];
        }
        while (total > gmax) {  // trim the largest shares until the grid is co-resident
            const int kmax = (int)(std::max_element(nb.begin(), nb.end()) - nb.begin());
            nb[kmax] -= 1;
            total -= 1;
        }
        std::vector<int32_t> pb_inst, pb_rank;
        pb_inst.reserve(total);
        pb_rank.reserve(total);
        for (int k = 0; k < pl->n_inst; ++k)
            for (int r = 0; r < nb[k]; ++r) {
                pb_inst.push_back(k);
                pb_rank.push_back(r);
            }
        (void)hipFree(pl->b_pb_inst[m]);
        (void)hipFree(pl->b_pb_rank[m]);
        (void)hipFree(pl->b_nblk[m]);
        pl->b_pb_inst[m] = pl->b_pb_rank[m] = pl->b_nblk[m] = nullptr;
        PHAST_HIP(hipMalloc(&pl->b_pb_inst[m], sizeof(int32_t) * total));
        PHAST_HIP(hipMalloc(&pl->b_pb_rank[m], sizeof(int32_t) * total));
        PHAST_HIP(hipMalloc(&pl->b_nblk[m], sizeof(int32_t) * pl->n_inst));
        PHAST_HIP(hipMemcpy(pl->b_pb_inst[m], pb_inst.data(), sizeof(int32_t) * total, hipMemcpyHostToDevice));
        PHAST_HIP(hipMemcpy(pl->b_pb_rank[m], pb_rank.data(), sizeof(int32_t) * total, hipMemcpyHostToDevice));
        PHAST_HIP(hipMemcpy(pl->b_nblk[m], nb.data(), sizeof(int32_t) * pl->n_inst, hipMemcpyHostToDevice));
        pl->b_grid[m] = (int32_t)total;
        pl->b_nb_max[m] = *std::max_element(nb.begin(), nb.end());
    }
    if (!pl->b_bar) {
        PHAST_HIP(hipMalloc(&pl->b_bar, sizeof(uint32_t) * 2 * pl->n_inst));
        PHAST_HIP(hipMemset(pl->b_bar, 0, sizeof(uint32_t) * 2 * pl->n_inst));
        PHAST_HIP(hipMalloc(&pl->b_err, sizeof(int32_t)));
        PHAST_HIP(hipMemset(pl->b_err, 0, sizeof(int32_t)));
        PHAST_HIP(hipDeviceSynchronize());
    }
    if (grid_out) {
        grid_out[0] = pl->b_grid[0];
        grid_out[1] = pl->b_grid[1];
    }
    return PHAST_OK;
}

namespace phast {

// The arguments of one batched step: inputs u, v, a, d (= d_prev), H_old; outputs u1, v1,
// a1, H_new, psi, d_out; the plan's CG work and the block tables of kernel `m`.
static int batch_args(phast_q4_plan *pl, int32_t method, int m, const double *u, const double *v, const double *a,
                      const double *d, const double *f_ext, const double *M_inv, const uint8_t *bc_mask,
                      const double *bc_vals, const double *H_old, const double *Gc_cell, const uint8_t *fixed,
                      const double *fixed_vals, double tol, int32_t max_iter, double *u1, double *v1, double *a1,
                      double *H_new, double *psi, double *d_out, BatchArgs &A) {
    PHAST_REQUIRE(pl && u && v && a && d && M_inv && H_old && u1 && v1 && a1 && H_new && d_out,
                  "batched step: null argument");
    PHAST_REQUIRE(method == 0 || method == 1, "method must be 0 (projected) or 1 (post-clamp)");
    PHAST_REQUIRE(pl->b_grid[m] > 0, "batched step: call phast_q4_batch_prepare first");
    PHAST_REQUIRE(!bc_mask == !bc_vals, "bc_mask and bc_vals must be given together");
    PHAST_REQUIRE(!fixed == !fixed_vals, "fixed and fixed_vals must be given together");
    PHAST_REQUIRE(tol >= 0.0 && max_iter >= 0, "tol and max_iter must be non-negative");
    PHAST_REQUIRE(u1 != u && v1 != v && a1 != a && H_new != H_old && d_out != d,
                  "batched step: outputs must not alias inputs");
    for (const void *q : {(const void *)u, (const void *)v, (const void *)a, (const void *)f_ext,
                          (const void *)u1, (const void *)v1, (const void *)a1, (const void *)H_old,
                          (const void *)H_new, (const void *)psi, (const void *)bc_vals})
        PHAST_REQUIRE(!q || aligned16(q), "vector arguments must be 16-byte aligned");
    std::memset(&A, 0, sizeof(A));
    A.M.inst = pl->d_inst;
    A.M.block_inst = pl->d_block_inst;
    A.M.u = reinterpret_cast<const double2 *>(u);
    A.M.v = reinterpret_cast<const double2 *>(v);
    A.M.a = reinterpret_cast<const double2 *>(a);
    A.M.d = d;
    A.M.fext = reinterpret_cast<const double2 *>(f_ext);
    A.M.minv = M_inv;
    A.M.bcm = bc_mask;
    A.M.bcv = bc_vals;
    A.M.Hold = H_old;
    A.M.u1 = reinterpret_cast<double2 *>(u1);
    A.M.v1 = reinterpret_cast<double2 *>(v1);
    A.M.a1 = reinterpret_cast<double2 *>(a1);
    A.M.Hnew = H_new;
    A.M.psi = psi;

    A.R = dmg_base(pl);
    A.R.H = H_new;
    A.R.Gc = Gc_cell;
    A.R.out = pl->b;
    A.R.out2 = pl->minv;

    DmgArgs D = dmg_base(pl);
    D.H = H_new;
    D.Gc = Gc_cell;
    D.d = d_out;
    D.r = pl->r;
    D.z = pl->z;
    D.p = pl->p;
    D.Ap = pl->Ap;
    D.b = pl->b;
    D.minv = pl->minv;
    D.lb = pl->lb;
    D.fixed = fixed;
    D.active = pl->active;
    D.part = pl->part;
    A.Dx = D;
    A.Dx.x = d_out;
    A.Dp = D;
    A.Dp.x = pl->p;

    VecArgs &V = A.V;
    V.inst = pl->d_inst;
    V.block_inst = pl->d_block_inst;
    V.d = d_out;
    V.r = pl->r;
    V.z = pl->z;
    V.p = pl->p;
    V.lb = pl->lb;
    V.Ap = pl->Ap;
    V.minv = pl->minv;
    V.init = nullptr;
    V.dprev = d;
    V.vals = fixed_vals;
    V.fixed = fixed;
    V.active = pl->active;
    V.active_out = nullptr;
    V.d_out = d_out;
    V.part = pl->part;
    V.n_tiles = pl->n_tiles;

    A.pb_inst = pl->b_pb_inst[m];
    A.pb_rank = pl->b_pb_rank[m];
    A.nblk = pl->b_nblk[m];
    A.bar = pl->b_bar;
    A.err = pl->b_err;
    A.minv = pl->minv;
    A.tol = tol;
    A.max_iter = max_iter;
    A.method = method;
    A.spin_ticks = pl->b_spin_ticks;
    return PHAST_OK;
}

}  // namespace phast

extern "C" int phast_q4_batch_step(phast_q4_plan *pl, int32_t method, const double *u, const double *v,
                                   const double *a, const double *d, const double *f_ext, const double *M_inv,
                                   const uint8_t *bc_mask, const double *bc_vals, const double *H_old,
                                   const double *Gc_cell, const uint8_t *fixed, const double *fixed_vals, double tol,
                                   int32_t max_iter, double *u1, double *v1, double *a1, double *H_new, double *psi,
                                   double *d_out, int32_t *iters, int32_t *status, hipStream_t st) {
    PHAST_REQUIRE(pl && iters && status, "phast_q4_batch_step: null argument");
    BatchArgs A;
    PHAST_TRY(batch_args(pl, method, method, u, v, a, d, f_ext, M_inv, bc_mask, bc_vals, H_old, Gc_cell, fixed,
                         fixed_vals, tol, max_iter, u1, v1, a1, H_new, psi, d_out, A));
    A.iters = iters;
    A.status = status;
    PHAST_TRY(mech_dispatch(pl, A.M, true, true, st));
    switch (pl->tx) {
        case 256: batch_launch<256>(pl, method, A, st); break;
        case 128: batch_launch<128>(pl, method, A, st); break;
        case 64: batch_launch<64>(pl, method, A, st); break;
        default: set_error("unsupported strip width %d", pl->tx); return PHAST_ERR_UNSUPPORTED;
    }
    PHAST_LAUNCHED();
    return PHAST_OK;
}


extern "C" int phast_q4_batch_run(phast_q4_plan *pl, int32_t method, int32_t n_steps, double *u, double *v,
                                  double *a, double *d, const double *f_ext, const double *M_inv,
                                  const uint8_t *bc_mask, const double *bc_vals, double *H, const double *Gc_cell,
                                  const uint8_t *fixed, const double *fixed_vals, double tol, int32_t max_iter,
                                  double *u1, double *v1, double *a1, double *H1, double *psi, double *d1,
                                  int32_t *iters, int32_t *status, hipStream_t st) {
    PHAST_REQUIRE(pl && iters && status, "phast_q4_batch_run: null argument");
    PHAST_REQUIRE(n_steps >= 1, "phast_q4_batch_run: n_steps must be at least 1");
    PHAST_REQUIRE(method == 0 || method == 1, "method must be 0 (projected) or 1 (post-clamp)");
    // the arrival counter of an instance counts (barriers passed) x (its blocks) within one
    // launch: at most 5 barriers per CG iteration plus 6 per step
    const int32_t nb_max = std::max(pl->b_nb_max[2 + method], 1);
    const double arrivals = (double)n_steps * (5.0 * ((double)max_iter + 1.0) + 6.0) * (double)nb_max;
    PHAST_REQUIRE(arrivals < 4.0e9, "phast_q4_batch_run: %d steps of up to %d CG iterations overflow the barrier "
                  "counters; use fewer steps per launch", n_steps, max_iter);
    BatchRunArgs R;
    std::memset(&R, 0, sizeof(R));
    PHAST_TRY(batch_args(pl, method, 2 + method, u, v, a, d, f_ext, M_inv, bc_mask, bc_vals, H, Gc_cell, fixed,
                         fixed_vals, tol, max_iter, u1, v1, a1, H1, psi, d1, R.A[0]));
    PHAST_TRY(batch_args(pl, method, 2 + method, u1, v1, a1, d1, f_ext, M_inv, bc_mask, bc_vals, H1, Gc_cell, fixed,
                         fixed_vals, tol, max_iter, u, v, a, H, psi, d, R.A[1]));
    R.n_steps = n_steps;
    R.n_inst = pl->n_inst;
    R.iters = iters;
    R.status = status;
    switch (pl->tx) {
        case 256: batch_run_launch<256>(pl, method, R, st); break;
        case 128: batch_run_launch<128>(pl, method, R, st); break;
        case 64: batch_run_launch<64>(pl, method, R, st); break;
        default: set_error("unsupported strip width %d", pl->tx); return PHAST_ERR_UNSUPPORTED;
    }
    PHAST_LAUNCHED();
    return PHAST_OK;
}

extern "C" int phast_q4_batch_error(phast_q4_plan *pl, int32_t *err, int32_t clear) {
    PHAST_REQUIRE(pl && err, "phast_q4_batch_error: null argument");
    *err = 0;
    if (!pl->b_err) return PHAST_OK;
    PHAST_HIP(hipMemcpy(err, pl->b_err, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (clear && *err) PHAST_HIP(hipMemset(pl->b_err, 0, sizeof(int32_t)));
    return PHAST_OK;
}
