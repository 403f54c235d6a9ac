// Linear-triangle (T3) kernels for gfx950 (FP64 throughout).
//
// Layout: nodes (N,2) float64 interleaved, elements (E,3) int32, per-element fields (E,)
// (psi, H), nodal u/v/a/f (N,2), d/Minv (N,).  The plan holds a CSR of element corners per
// node: inc[ptr[n] .. ptr[n+1]) = e*3 + a for every element e with node n at corner a, in
// increasing e*3 + a -- the order in which the reference's scatter_add over elem_flat visits
// the node (core/fem_operators.py:661-679 with _elem_flat = elements.flatten()).
//
// Every pass is a gather: one thread per node sums its corners' contributions in that order,
// so there are no atomics and results are deterministic.  Two organisations produce the same
// bits:
//  * patches (default): the nodes are grouped into patches of 256 in the Morton order of
//    their coordinates; a block first evaluates each triangle touching its patch once
//    (geometry from the corner coordinates, strain, degradation, the split's stress) into
//    LDS, then each thread sums its node's corners from LDS;
//  * recompute (meshes whose patch lists exceed the LDS budget, or on request): each
//    incidence re-evaluates its triangle.
// Element-wise outputs (psi+, history) are written once per triangle: by the patch that owns
// its corner 0, or by the thread of corner 0.
#include "common.h"
#include "damage_cg.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>

#include "../../include/phast_mi.h"

// No FMA contraction in this file: every product and sum rounds on its own, as in the
// reference's tensor expressions, and the two pass organisations (which the compiler would
// otherwise contract differently) give the same bits.
#pragma clang fp contract(off)

namespace phast {

enum : int { T3_AMOR = PHAST_SPLIT_AMOR, T3_ISO = PHAST_SPLIT_ISOTROPIC, T3_SPECTRAL = PHAST_SPLIT_SPECTRAL };
constexpr double kSpectralEps = 1e-12;  // core/fem_operators.py:126 (float64)

struct T3Mat {
    double lam, mu, kappa, ezz_coef, eta, Gc_l0, Gc_over_l0, l0;
    int32_t split, plane_stress;
    int32_t ssplit;  // split of the stress (isotropic under stress_degradation 'full')
    int32_t at1;     // AT1 damage law (else AT2)
    double src;      // AT1 source S_H of the right-hand side (0 for AT2)
};

struct T3Mesh {
    const double2 *xy;
    const int32_t *el;   // (E,3) triangles, (E,4) quadrilaterals
    const int32_t *ptr;  // (N+1)
    const int32_t *inc;  // incidences e*K + a (K corners per element)
    int64_t n_nodes, n_elems;
    double *ce = nullptr;  // quadrilaterals: two values per corner (E,4,2) from the element pass
};

// Where a node pass takes its element corner values: SRC_RECOMPUTE evaluates the triangle of
// each incidence, SRC_PATCH reads the patch's triangles from LDS, SRC_BUFFER reads the corner
// values an element pass wrote to m.ce (quadrilateral meshes; the incidence e*4 + a indexes
// them directly).
enum : int { SRC_RECOMPUTE = 0, SRC_PATCH = 1, SRC_BUFFER = 2 };

// Area and shape-function gradients of one triangle (core/mesh.py:236-268).
__device__ __forceinline__ void t3_geometry(double2 p0, double2 p1, double2 p2, double &area, double gp[3][2]) {
    const double v01x = p1.x - p0.x, v01y = p1.y - p0.y, v02x = p2.x - p0.x, v02y = p2.y - p0.y;
    const double s2a = v01x * v02y - v01y * v02x;
    area = 0.5 * fabs(s2a);
    const double sgn = s2a > 0.0 ? 1.0 : (s2a < 0.0 ? -1.0 : 1.0);
    const double inv = 1.0 / (fabs(s2a) > 1e-10 ? s2a : sgn * 1e-10);
    gp[0][0] = (p1.y - p2.y) * inv;
    gp[0][1] = (p2.x - p1.x) * inv;
    gp[1][0] = (p2.y - p0.y) * inv;
    gp[1][1] = (p0.x - p2.x) * inv;
    gp[2][0] = (p0.y - p1.y) * inv;
    gp[2][1] = (p1.x - p0.x) * inv;
}

// Degraded stress of one element for the stress split (fem_operators.py:277-290 isotropic,
// 301-333 amor, 335-390 spectral; chosen as :24-33, 641-653), g = g(mean of the corner damages).
__device__ __forceinline__ void t3_stress(const T3Mat &M, double exx, double eyy, double gxy, double g, double &sxx,
                                          double &syy, double &sxy) {
    const double lam = M.lam, mu = M.mu;
    if (M.ssplit == T3_ISO) {
        const double c11 = lam + 2.0 * mu;
        sxx = g * (c11 * exx + lam * eyy);
        syy = g * (lam * exx + c11 * eyy);
        sxy = g * (mu * gxy);
    } else if (M.ssplit == T3_AMOR) {
        const double tr2 = exx + eyy;
        const double ezz = M.plane_stress ? M.ezz_coef * tr2 : 0.0;
        const double tr = tr2 + ezz;
        const double trp = tr > 0.0 ? tr : 0.0;
        const double trm = tr - trp;
        const double dxx = exx - tr / 3.0, dyy = eyy - tr / 3.0;
        sxx = g * (M.kappa * trp + 2 * mu * dxx) + M.kappa * trm;
        syy = g * (M.kappa * trp + 2 * mu * dyy) + M.kappa * trm;
        sxy = g * mu * gxy;
    } else {
        const double exy = gxy / 2.0, trace = exx + eyy, hd = (exx - eyy) / 2.0;
        const double delta = sqrt(hd * hd + exy * exy + kSpectralEps);
        const double e1 = trace / 2.0 + delta, e2 = trace / 2.0 - delta;
        const double e1p = e1 > 0.0 ? e1 : 0.0, e2p = e2 > 0.0 ? e2 : 0.0;
        const double trp = trace > 0.0 ? trace : 0.0, trm = trace - trp;
        const double inv = 1.0 / (2.0 * delta + 1e-30);
        const double pxx = (hd + delta) * inv, pyy = (-hd + delta) * inv, pxy = exy * inv;
        const double exx_p = e1p * pxx + e2p * (1.0 - pxx), eyy_p = e1p * pyy + e2p * (1.0 - pyy);
        const double exy_p = (e1p - e2p) * pxy;
        const double exx_m = exx - exx_p, eyy_m = eyy - eyy_p, exy_m = exy - exy_p;
        sxx = g * (lam * trp + 2 * mu * exx_p) + (lam * trm + 2 * mu * exx_m);
        syy = g * (lam * trp + 2 * mu * eyy_p) + (lam * trm + 2 * mu * eyy_m);
        sxy = g * (2 * mu * exy_p) + (2 * mu * exy_m);
    }
}

// Tensile energy density (fem_operators.py:983-1030).
__device__ __forceinline__ double t3_psi_plus(const T3Mat &M, double exx, double eyy, double gxy) {
    const double lam = M.lam, mu = M.mu, exy = gxy / 2.0;
    if (M.split == T3_ISO) {
        const double tr = exx + eyy;
        return 0.5 * lam * (tr * tr) + mu * ((exx * exx + eyy * eyy) + 2 * (exy * exy));
    }
    if (M.split == T3_AMOR) {
        const double tr2 = exx + eyy;
        const double ezz = M.plane_stress ? M.ezz_coef * tr2 : 0.0;
        const double tr = tr2 + ezz, trp = tr > 0.0 ? tr : 0.0;
        const double dxx = exx - tr / 3.0, dyy = eyy - tr / 3.0, dzz = ezz - tr / 3.0;
        return 0.5 * M.kappa * (trp * trp) + mu * (((dxx * dxx + dyy * dyy) + dzz * dzz) + 2.0 * (exy * exy));
    }
    const double trace = exx + eyy, hd = (exx - eyy) / 2.0;
    const double delta = sqrt(hd * hd + exy * exy + kSpectralEps);
    const double e1 = trace / 2.0 + delta, e2 = trace / 2.0 - delta;
    const double e1p = e1 > 0.0 ? e1 : 0.0, e2p = e2 > 0.0 ? e2 : 0.0, trp = trace > 0.0 ? trace : 0.0;
    return 0.5 * lam * (trp * trp) + mu * (e1p * e1p + e2p * e2p);
}

// One element: geometry, strain, degradation.
struct T3Elem {
    int32_t n[3];
    double area, gp[3][2], exx, eyy, gxy, g;
};

// ux(n) returns the displacement of node n (read, or predicted by the explicit step).
template <class UX>
__device__ __forceinline__ void t3_load(const T3Mesh &m, const T3Mat &M, int4 t, UX ux, const double *d, T3Elem &E) {
    E.n[0] = t.y;
    E.n[1] = t.z;
    E.n[2] = t.w;
    t3_geometry(m.xy[E.n[0]], m.xy[E.n[1]], m.xy[E.n[2]], E.area, E.gp);
    const double2 u0 = ux(E.n[0]), u1 = ux(E.n[1]), u2 = ux(E.n[2]);
    E.exx = (E.gp[0][0] * u0.x + E.gp[1][0] * u1.x) + E.gp[2][0] * u2.x;
    E.eyy = (E.gp[0][1] * u0.y + E.gp[1][1] * u1.y) + E.gp[2][1] * u2.y;
    E.gxy = ((E.gp[0][1] * u0.x + E.gp[1][1] * u1.x) + E.gp[2][1] * u2.x) +
            ((E.gp[0][0] * u0.y + E.gp[1][0] * u1.y) + E.gp[2][0] * u2.y);
    if (d) {
        const double dm = ((d[E.n[0]] + d[E.n[1]]) + d[E.n[2]]) / 3.0;
        const double omd = 1.0 - dm;
        E.g = (omd * omd) * (1.0 - M.eta) + M.eta;  // physics/material.py:529
    } else {
        E.g = 1.0;
    }
}

// Per-node incidence walk over the element patches: the plan groups the nodes into patches
// of 256 (a Morton order of the coordinates, so a patch is a compact piece of the mesh), and
// lists for each patch the triangles touching it.  A block evaluates each listed triangle
// once into LDS (2 values per corner), then each thread sums its node's corners from LDS in
// the node's e*3+a order.  Without patches every incidence re-evaluates its triangle.
// A triangle is passed around as int4 {e, n0, n1, n2}.  Everything a block reads for its
// patch is contiguous: the triangle table, the slot -> node map and, per slot in patch order,
// the LDS slots of the node's corners.
struct T3Patch {
    const int32_t *pnode;    // (N) node of each patch slot, patch b = slots [256 b, 256 b + 256)
    const int32_t *pel_ptr;  // (npatch + 1)
    const int4 *pel;         // {e, n0, n1, n2} of each patch's triangles, e ascending; the sign bit
                             // of e marks the patch that owns corner 0
    const int32_t *sptr;     // (N + 1) incidences of slot s: sloc[sptr[s] .. sptr[s+1])
    const uint16_t *sloc;    // (3E) LDS slot (list position * 3 + corner), in the node's e*3+a order
};

__device__ __forceinline__ int4 tri_of(const T3Mesh &m, int64_t e) {
    return make_int4((int32_t)e, m.el[3 * e], m.el[3 * e + 1], m.el[3 * e + 2]);
}

// Two values per corner of one triangle, as named scalars: selecting a corner from an array
// with a run-time index would put the array in scratch memory.
struct Corners {
    double c00, c01, c10, c11, c20, c21;
    __device__ __forceinline__ void set(int a, double v0, double v1) {
        if (a == 0) {
            c00 = v0;
            c01 = v1;
        } else if (a == 1) {
            c10 = v0;
            c11 = v1;
        } else {
            c20 = v0;
            c21 = v1;
        }
    }
};

// Fill the block's LDS with the corner values of its patch's triangles; corners(t, c, owner)
// fills the Corners of triangle t.  Every thread of the block must call it.
template <class F>
__device__ __forceinline__ void t3_patch_fill(const T3Patch &P, double *lds, F corners, int blk) {
    const int32_t j0 = P.pel_ptr[blk], nel = P.pel_ptr[blk + 1] - j0;
    for (int32_t j = threadIdx.x; j < nel; j += 256) {
        int4 t = P.pel[j0 + j];
        const bool owner = t.x < 0;
        t.x &= 0x7fffffff;
        Corners c;
        corners(t, c, owner);
        lds[6 * j] = c.c00;
        lds[6 * j + 1] = c.c01;
        lds[6 * j + 2] = c.c10;
        lds[6 * j + 3] = c.c11;
        lds[6 * j + 4] = c.c20;
        lds[6 * j + 5] = c.c21;
    }
    __syncthreads();
}

// Sums over node n's incidences, in e*3+a order, of the two corner values.  __dadd_rn keeps
// each addition a separate rounding in both organisations, so they agree bitwise.
template <int SRC, class F>
__device__ __forceinline__ void t3_node_sum(const T3Mesh &m, const T3Patch &P, const double *lds, int64_t slot,
                                            int64_t n, F corners, double &s0, double &s1) {
    s0 = 0.0;
    s1 = 0.0;
    const int32_t kb = SRC == SRC_PATCH ? P.sptr[slot] : m.ptr[n];
    const int32_t ke = SRC == SRC_PATCH ? P.sptr[slot + 1] : m.ptr[n + 1];
    for (int32_t k = kb; k < ke; ++k) {
        double v0, v1;
        if (SRC == SRC_BUFFER) {
            const int64_t ic = m.inc[k];
            v0 = m.ce[2 * ic];
            v1 = m.ce[2 * ic + 1];
        } else if (SRC == SRC_PATCH) {
            const int sl = P.sloc[k];
            v0 = lds[2 * sl];
            v1 = lds[2 * sl + 1];
        } else {
            const int32_t ic = m.inc[k];
            const int64_t e = ic / 3;
            const int a_ = ic - 3 * (int)e;
            Corners c;
            corners(tri_of(m, e), c, a_ == 0);
            v0 = a_ == 0 ? c.c00 : (a_ == 1 ? c.c10 : c.c20);
            v1 = a_ == 0 ? c.c01 : (a_ == 1 ? c.c11 : c.c21);
        }
        s0 = __dadd_rn(s0, v0);
        s1 = __dadd_rn(s1, v1);
    }
}

// Corner forces A (s . grad phi_a) of one triangle (fem_operators.py:661-679); with HIST the
// owner also writes psi+ and H = max(H_old, psi+) (fem_operators.py:951-1030,
// staggered_solver.py:804-845).
template <bool HIST, class UX>
__device__ __forceinline__ void t3_force_corners(const T3Mesh &m, const T3Mat &M, UX ux, const double *d,
                                                 const double *H_old, double *H_new, double *psi, int4 t,
                                                 Corners &c, bool owner) {
    const int64_t e = t.x;
    T3Elem E;
    t3_load(m, M, t, ux, d, E);
    double sxx, syy, sxy;
    t3_stress(M, E.exx, E.eyy, E.gxy, E.g, sxx, syy, sxy);
#pragma unroll
    for (int a = 0; a < 3; ++a)
        c.set(a, E.area * (sxx * E.gp[a][0] + sxy * E.gp[a][1]), E.area * (sxy * E.gp[a][0] + syy * E.gp[a][1]));
    if (HIST && owner) {
        const double p = t3_psi_plus(M, E.exx, E.eyy, E.gxy);
        if (psi) psi[e] = p;
        const double h = H_old[e];
        H_new[e] = h >= p ? h : p;  // max returns one operand bit-exactly
    }
}

// Operands of the explicit step (solvers/mechanics_solver.py:153-239).
struct T3Step {
    double dt, c2;  // c2 = dt^2 / 2
    const double2 *v, *a, *fext;
    const double2 *fdamp;  // Kelvin-Voigt damping force beta f_int(v, d), or null
    double fscale;         // MODE 0: f = fscale f_int (1 for the plain force)
    const double *minv;
    const uint8_t *bcm;  // (N,2) prescribed components, or null
    const double *bcv;
    double2 *u1, *v1, *a1;
};

// Predictor u1 = u + dt v + dt^2/2 a with prescribed components overridden
// (mechanics_solver.py:160-170).
__device__ __forceinline__ double2 t3_predict(const T3Step &S, const double2 *u, int64_t n) {
    const double2 u0 = u[n], v0 = S.v[n], a0 = S.a[n];
    double ux = (u0.x + S.dt * v0.x) + S.c2 * a0.x, uy = (u0.y + S.dt * v0.y) + S.c2 * a0.y;
    if (S.bcm) {
        if (S.bcm[2 * n]) ux = S.bcv[2 * n];
        if (S.bcm[2 * n + 1]) uy = S.bcv[2 * n + 1];
    }
    return make_double2(ux, uy);
}

// The predictor as its own pass.  Evaluating it inside the force pass, for the corners of
// every listed triangle, was measured slower (8.4 M nodes: 0.675 ms against 0.625 ms): the
// triangles then gather u, v and a instead of u1 alone.
__global__ __launch_bounds__(256) void t3_predictor_kernel(int64_t nn, const double2 *u, const T3Step S) {
    const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (n < nn) S.u1[n] = t3_predict(S, u, n);
}

// MODE 0: f = internal force; 1: + psi+ and H = max(H_old, psi+) per element;
// 2: explicit-step corrector on top of 1 (u is the predicted u1; writes v1, a1).
template <int MODE, int SRC>
__global__ __launch_bounds__(256) void t3_force_kernel(const T3Mesh m, const T3Patch P, const T3Mat M, const double2 *u,
                                                       const double *d, double2 *f, const double *H_old, double *H_new,
                                                       double *psi, const T3Step S) {
    extern __shared__ double lds[];
    auto ux = [&](int64_t n) { return u[n]; };
    auto corners = [&](int4 t, Corners &c, bool owner) {
        t3_force_corners<(MODE >= 1)>(m, M, ux, d, H_old, H_new, psi, t, c, owner);
    };
    const int blk = xcd_tile(blockIdx.x, gridDim.x);  // patch of this block
    if (SRC == SRC_PATCH) t3_patch_fill(P, lds, corners, blk);
    const int64_t slot = (int64_t)blk * 256 + threadIdx.x;
    if (slot >= m.n_nodes) return;
    const int64_t n = SRC == SRC_PATCH ? (int64_t)P.pnode[slot] : slot;
    double fx, fy;
    t3_node_sum<SRC>(m, P, lds, slot, n, corners, fx, fy);
    if (MODE == 0) {
        f[n] = make_double2(S.fscale * fx, S.fscale * fy);  // x 1 is exact
        return;
    }
    if (MODE == 1) {
        f[n] = make_double2(fx, fy);
        return;
    }
    // mechanics_solver.py:194-198: a1 = Minv (f_ext - f_int), v1 = v + dt/2 (a + a1); BC zero
    const double dt = S.dt;
    const double mi = S.minv[n];
    const double2 fe = S.fext ? S.fext[n] : make_double2(0.0, 0.0);
    const double2 v0 = S.v[n], a0 = S.a[n];
    // mechanics_solver.py:192-198: f_net = (f_ext - f_int) [- beta f_int(v, d)]
    double nx = fe.x - fx, ny = fe.y - fy;
    if (S.fdamp) {
        const double2 fd = S.fdamp[n];
        nx = nx - fd.x;
        ny = ny - fd.y;
    }
    double ax = mi * nx, ay = mi * ny;
    double vx = v0.x + 0.5 * dt * (a0.x + ax), vy = v0.y + 0.5 * dt * (a0.y + ay);
    if (S.bcm) {
        if (S.bcm[2 * n]) ax = vx = 0.0;
        if (S.bcm[2 * n + 1]) ay = vy = 0.0;
    }
    S.v1[n] = make_double2(vx, vy);
    S.a1[n] = make_double2(ax, ay);
}

// Damage operator of the element-H AT2 path (solvers/damage_solver.py), corner values of
// one triangle:
// OP 0: c[a] = (A grad phi_a . grad x_e [x Gc_e l0], rc (x_a + sum_b x_b)),
//       rc = (2H + Gc/l0) A/12: the two scatters of :1719-1755, summed separately and
//       combined as Gc l0 s0 + s1 (s0 + s1 with a per-element Gc field);
// OP 1: c[a] = (2H A/3 (:2249-2258), Gc l0 A |grad phi_a|^2 + 2 rc (:2112-2119)).
template <int OP>
__device__ __forceinline__ void t3_damage_corners(const T3Mesh &m, const T3Mat &M, const double *H, const double *Gc,
                                                  const double *x, int4 t, Corners &c) {
    const int64_t e = t.x;
    const int32_t nd[3] = {t.y, t.z, t.w};
    double area, gp[3][2];
    t3_geometry(m.xy[nd[0]], m.xy[nd[1]], m.xy[nd[2]], area, gp);
    const double h = H[e];
    const double ge = Gc ? Gc[e] : 0.0;
    // per-element Gc: damage_solver.py:591-611 (AT1: 0.75 Gc l0, no Gc/l0, S_H = 3 Gc/(8 l0))
    const double ratio = Gc ? (M.at1 ? 0.0 : ge / M.l0) : M.Gc_over_l0;  // damage_solver.py:2168-2177
    const double gl = Gc ? (M.at1 ? 0.75 * ge : ge) * M.l0 : M.Gc_l0;
    const double rc = (2.0 * h + ratio) * area / 12.0;
    if (OP == 0) {
        const double d0 = x[nd[0]], d1 = x[nd[1]], d2 = x[nd[2]];
        const double gdx = (gp[0][0] * d0 + gp[1][0] * d1) + gp[2][0] * d2;
        const double gdy = (gp[0][1] * d0 + gp[1][1] * d1) + gp[2][1] * d2;
        const double ds = (d0 + d1) + d2;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double lap = area * (gp[a][0] * gdx + gp[a][1] * gdy);
            c.set(a, Gc ? gl * lap : lap, rc * ((a == 0 ? d0 : (a == 1 ? d1 : d2)) + ds));
        }
    } else {
        const double src = Gc ? (M.at1 ? 3.0 * ge / (8.0 * M.l0) : 0.0) : M.src;
#pragma unroll
        for (int a = 0; a < 3; ++a)
            c.set(a, (2.0 * h - src) * (area / 3.0),
                  gl * (area * (gp[a][0] * gp[a][0] + gp[a][1] * gp[a][1])) + 2.0 * rc);
    }
}

// OP 0: y = A x; OP 1: y = b, y2 = diag; OP 2: y = b, y2 = M_inv = 1/(diag + 1e-30)
// (damage_solver.py:2265)
template <int OP, int SRC>
__global__ __launch_bounds__(256) void t3_damage_kernel(const T3Mesh m, const T3Patch P, const T3Mat M, const double *H,
                                                        const double *Gc, const double *x, double *y, double *y2) {
    extern __shared__ double lds[];
    auto corners = [&](int4 t, Corners &c, bool) { t3_damage_corners<(OP == 0 ? 0 : 1)>(m, M, H, Gc, x, t, c); };
    const int blk = xcd_tile(blockIdx.x, gridDim.x);  // patch of this block
    if (SRC == SRC_PATCH) t3_patch_fill(P, lds, corners, blk);
    const int64_t slot = (int64_t)blk * 256 + threadIdx.x;
    if (slot >= m.n_nodes) return;
    const int64_t n = SRC == SRC_PATCH ? (int64_t)P.pnode[slot] : slot;
    double s0, s1;
    t3_node_sum<SRC>(m, P, lds, slot, n, corners, s0, s1);
    if (OP == 0) {
        y[n] = Gc ? s0 + s1 : M.Gc_l0 * s0 + s1;
    } else {
        if (y) y[n] = s0;
        if (y2) y2[n] = OP == 2 ? 1.0 / (s1 + 1e-30) : s1;
    }
}

// Element fields, one thread per triangle: centroid strain (fem_operators.py:223-242), the
// split's degraded stress (:614-653) and psi+ (:951-1030); each output optional.
__global__ __launch_bounds__(256) void t3_elem_fields_kernel(const T3Mesh m, const T3Mat M, const double2 *u,
                                                             const double *d, double *eps, double *sig, double *psi) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= m.n_elems) return;
    T3Elem E;
    t3_load(m, M, tri_of(m, e), [&](int64_t n) { return u[n]; }, d, E);
    const int64_t ne = m.n_elems;
    if (eps) {
        eps[e] = E.exx;
        eps[ne + e] = E.eyy;
        eps[2 * ne + e] = E.gxy;
    }
    if (sig) {
        double sxx, syy, sxy;
        t3_stress(M, E.exx, E.eyy, E.gxy, E.g, sxx, syy, sxy);
        sig[e] = sxx;
        sig[ne + e] = syy;
        sig[2 * ne + e] = sxy;
    }
    if (psi) psi[e] = t3_psi_plus(M, E.exx, E.eyy, E.gxy);
}

// Area-weighted projection of an element field (core/mesh.py:430-433):
// out_n = sum_e f_e A_e / (3 M_n + 1e-30).
__global__ __launch_bounds__(256) void t3_e2n_kernel(const T3Mesh m, const double *fe, const double *Mn, double *out) {
    const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (n >= m.n_nodes) return;
    double s = 0.0;
    for (int32_t k = m.ptr[n]; k < m.ptr[n + 1]; ++k) {
        const int64_t e = m.inc[k] / 3;
        const double2 p0 = m.xy[m.el[3 * e]], p1 = m.xy[m.el[3 * e + 1]], p2 = m.xy[m.el[3 * e + 2]];
        const double s2a = (p1.x - p0.x) * (p2.y - p0.y) - (p1.y - p0.y) * (p2.x - p0.x);
        s += fe[e] * (0.5 * fabs(s2a));
    }
    out[n] = s / (3.0 * Mn[n] + 1e-30);
}

// Adjoint terms of the damage solve, one thread per triangle (damage_solver.py:284-331 and the
// history pullback of :800-810, :1140-1149).  With lam, dp (the damage of the history
// pullback) and du (the unconstrained damage of the quadratic forms) at the corners:
//   gH  = sum_a lam_a 2A (1/3 - (dp_a + sum dp) / 12)
//   lKd = A (grad lam . grad du),  lMd = A/12 (sum lam sum du + lam . du),  lM1 = A/3 sum lam
__global__ __launch_bounds__(256) void t3_adjoint_terms_kernel(const T3Mesh m, const double *lam, const double *dp,
                                                               const double *du, double *gH, double *lKd,
                                                               double *lMd, double *lM1) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= m.n_elems) return;
    const int32_t n0 = m.el[3 * e], n1 = m.el[3 * e + 1], n2 = m.el[3 * e + 2];
    double area, gp[3][2];
    t3_geometry(m.xy[n0], m.xy[n1], m.xy[n2], area, gp);
    const double l0 = lam[n0], l1 = lam[n1], l2 = lam[n2];
    const double ls = (l0 + l1) + l2;
    if (gH) {
        const double a0 = dp[n0], a1 = dp[n1], a2 = dp[n2];
        const double ds = (a0 + a1) + a2, w = 2.0 * area;
        gH[e] = (l0 * (w * (1.0 / 3.0 - (a0 + ds) / 12.0)) + l1 * (w * (1.0 / 3.0 - (a1 + ds) / 12.0))) +
                l2 * (w * (1.0 / 3.0 - (a2 + ds) / 12.0));
    }
    if (lKd || lMd) {
        const double d0 = du[n0], d1 = du[n1], d2 = du[n2];
        if (lKd) {
            const double glx = (gp[0][0] * l0 + gp[1][0] * l1) + gp[2][0] * l2;
            const double gly = (gp[0][1] * l0 + gp[1][1] * l1) + gp[2][1] * l2;
            const double gdx = (gp[0][0] * d0 + gp[1][0] * d1) + gp[2][0] * d2;
            const double gdy = (gp[0][1] * d0 + gp[1][1] * d1) + gp[2][1] * d2;
            lKd[e] = area * (glx * gdx + gly * gdy);
        }
        if (lMd) lMd[e] = (area / 12.0) * (ls * ((d0 + d1) + d2) + ((l0 * d0 + l1 * d1) + l2 * d2));
    }
    if (lM1) lM1[e] = (area / 3.0) * ls;
}

// ---------------------------------------------------------------------------------------
// device-resident projected / post-clamp Jacobi CG on the T3 operator (the Q4 solve's state
// machine and scalar steps; one thread per node, per-block partial sums in part[k*nblk+b])
// ---------------------------------------------------------------------------------------

enum T3CgOp : int {
    T_CG_INIT = 0,   // r = b - A d, active set, z, p; partials r.r, b.b, rf.z
    T_CG_AP = 1,     // Ap = A p; partials p.Ap, min step ratios to the bounds
    T_CG_RESID = 2,  // r = b - A d; new active set; partials rf.rf, changed
    T_PC_INIT = 3,   // post-clamp: r = b - A d, z = r Minv, p = z; partials r.r, b.b, r.z
    T_PC_AP = 4,     // post-clamp: Ap = A p; partial p.Ap
    T_PC_RESID = 5,  // post-clamp true residual r = b - A d
};

enum T3VecOp : int {
    TV_SETUP = 0,       // d = fixed ? vals : init ; lb = fixed ? vals : d_prev
    TV_UPDATE_D = 1,    // projected: limited ? clamp(d + a p) : d += a p, r -= a Ap
    TV_PRECOND = 2,     // projected: z = active ? 0 : r Minv ; partial rf.z
    TV_UPDATE_P = 3,    // projected: p = active ? 0 : (restart ? z : beta p + z)
    TV_FINISH = 4,      // d = min(max(d, lb), 1), fixed -> vals ; copy active
    TV_PC_UPDATE = 5,   // post-clamp: d += a p, r -= a Ap
    TV_PC_PRECOND = 6,  // post-clamp: z = r Minv ; partials r.r, r.z
    TV_PC_UPDATE_P = 7,
    TV_PC_FINISH = 8,   // d = clamp(max(d, lb), 0, 1), fixed -> vals
    // without a residual recomputation in between (no limited step, no check), the d/r update
    // and the preconditioner share one pass and the separate preconditioner pass returns at
    // once; arithmetic and partial sums are those of the separate passes
    TV_UPDATE_D_PRE = 9,
    TV_PRECOND_CHK = 10,
    TV_PC_UPDATE_PRE = 11,
    TV_PC_PRECOND_CHK = 12,
    // plain Jacobi-PCG of the adjoint recomputation (damage_solver.py:423-478); pinned DOFs
    // are eliminated: held at vals (0 when no values are given), r = z = p = 0 there
    TA_SETUP = 13,     // x = fixed ? vals : (init or 0)
    TA_UPDATE = 14,    // x += a p, r -= a Ap, r = 0 at pins, z = M_eff r ; partials r.r, r.z
    TA_UPDATE_P = 15,  // p = z + beta p, 0 at pins
    TA_FINISH = 16,    // x = fixed ? vals : x
};


// ---------------------------------------------------------------------------------------
// General quadrilateral meshes (the Q4 branch of the reference on any quad mesh)
// ---------------------------------------------------------------------------------------
// An element pass (one thread per quadrilateral) evaluates the element at its 2x2 Gauss
// points and writes two values per corner to m.ce; the node passes above (SRC_BUFFER) sum
// them in each node's e*4 + a order, the order of the reference's scatter_add over
// elements.flatten().  Geometry is recomputed from the corner coordinates as
// core/quad_elements.py:315-349 does once: Gauss points (xi, eta) = (-a,-a), (a,-a), (-a,a),
// (a,a), a = 1/sqrt3, weights 1; J_ij = sum_a dN_a/dxi_i x_aj, grad N_a = J^-T dN_a/dxi,
// wdetJ = det J.  History and psi+ are per Gauss point (E,4).

struct Q4uElem {
    int32_t n[4];
    double N[4][4];      // [q][a]
    double gp[4][4][2];  // [q][a][x|y]
    double wd[4];        // [q]
};

__device__ __forceinline__ void q4u_geometry(const T3Mesh &m, int64_t e, Q4uElem &Q) {
    double2 X[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        Q.n[a] = m.el[4 * e + a];
        X[a] = m.xy[Q.n[a]];
    }
    const double g = 1.0 / sqrt(3.0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double xi = (q & 1) ? g : -g, eta = (q & 2) ? g : -g;
        Q.N[q][0] = 0.25 * ((1.0 - xi) * (1.0 - eta));
        Q.N[q][1] = 0.25 * ((1.0 + xi) * (1.0 - eta));
        Q.N[q][2] = 0.25 * ((1.0 + xi) * (1.0 + eta));
        Q.N[q][3] = 0.25 * ((1.0 - xi) * (1.0 + eta));
        const double dxi[4] = {0.25 * -(1.0 - eta), 0.25 * (1.0 - eta), 0.25 * (1.0 + eta), 0.25 * -(1.0 + eta)};
        const double det_[4] = {0.25 * -(1.0 - xi), 0.25 * -(1.0 + xi), 0.25 * (1.0 + xi), 0.25 * (1.0 - xi)};
        double J00 = 0.0, J01 = 0.0, J10 = 0.0, J11 = 0.0;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            J00 += dxi[a] * X[a].x;
            J01 += dxi[a] * X[a].y;
            J10 += det_[a] * X[a].x;
            J11 += det_[a] * X[a].y;
        }
        const double det = J00 * J11 - J01 * J10;
        const double i00 = J11 / det, i01 = -J01 / det, i10 = -J10 / det, i11 = J00 / det;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            Q.gp[q][a][0] = dxi[a] * i00 + det_[a] * i01;
            Q.gp[q][a][1] = dxi[a] * i10 + det_[a] * i11;
        }
        Q.wd[q] = det;
    }
}

// Forces of one quadrilateral at its corners (fem_operators.py:244-254 strain, 614-653
// stress with g at the Gauss points from N d, 681-699 assembly); with HIST also psi+ and
// H = max(H_old, psi+) at each Gauss point (fem_operators.py:951-1006, staggered_solver.py:
// 804-845).  MODE 2 reads the predicted displacement u1.
template <bool HIST>
__global__ __launch_bounds__(256) void q4u_force_elem_kernel(const T3Mesh m, const T3Mat M, const double2 *u,
                                                             const double *d, const double *H_old, double *H_new,
                                                             double *psi, double *ce) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= m.n_elems) return;
    Q4uElem Q;
    q4u_geometry(m, e, Q);
    double2 U[4];
    double D[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        U[a] = u[Q.n[a]];
        D[a] = d ? d[Q.n[a]] : 0.0;
    }
    double fx[4] = {0.0, 0.0, 0.0, 0.0}, fy[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double exx = 0.0, eyy = 0.0, gx = 0.0, gy = 0.0, dq = 0.0;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            exx += Q.gp[q][a][0] * U[a].x;
            eyy += Q.gp[q][a][1] * U[a].y;
            gx += Q.gp[q][a][1] * U[a].x;
            gy += Q.gp[q][a][0] * U[a].y;
            dq += Q.N[q][a] * D[a];
        }
        const double gxy = gx + gy;
        double g = 1.0;
        if (d) {
            const double omd = 1.0 - dq;
            g = (omd * omd) * (1.0 - M.eta) + M.eta;  // physics/material.py:529
        }
        double sxx, syy, sxy;
        t3_stress(M, exx, eyy, gxy, g, sxx, syy, sxy);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            fx[a] += (Q.gp[q][a][0] * sxx + Q.gp[q][a][1] * sxy) * Q.wd[q];
            fy[a] += (Q.gp[q][a][1] * syy + Q.gp[q][a][0] * sxy) * Q.wd[q];
        }
        if (HIST) {
            const double p = t3_psi_plus(M, exx, eyy, gxy);
            if (psi) psi[4 * e + q] = p;
            const double h = H_old[4 * e + q];
            H_new[4 * e + q] = h >= p ? h : p;  // max returns one operand bit-exactly
        }
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        double *c = m.ce + 8 * e + 2 * a;
        c[0] = fx[a];
        c[1] = fy[a];
    }
}

// Gauss-point strain (3,E,4) and psi+ (E,4) of one quadrilateral (fem_operators.py:244-254,
// 951-1006); each output optional.
__global__ __launch_bounds__(256) void q4u_fields_kernel(const T3Mesh m, const T3Mat M, const double2 *u, double *eps,
                                                         double *psi) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= m.n_elems) return;
    Q4uElem Q;
    q4u_geometry(m, e, Q);
    double2 U[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) U[a] = u[Q.n[a]];
    const int64_t ne = m.n_elems;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double exx = 0.0, eyy = 0.0, gx = 0.0, gy = 0.0;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            exx += Q.gp[q][a][0] * U[a].x;
            eyy += Q.gp[q][a][1] * U[a].y;
            gx += Q.gp[q][a][1] * U[a].x;
            gy += Q.gp[q][a][0] * U[a].y;
        }
        const double gxy = gx + gy;
        if (eps) {
            eps[4 * e + q] = exx;
            eps[4 * (ne + e) + q] = eyy;
            eps[4 * (2 * ne + e) + q] = gxy;
        }
        if (psi) psi[4 * e + q] = t3_psi_plus(M, exx, eyy, gxy);
    }
}

// Damage operator of one quadrilateral with Gauss-point history H (E,4) (damage_solver.py:
// 1797-1838 _Ax_q4, 2270-2301 _prepare_cg_q4), corner values:
// OP 0: (sum_q wdet grad N_a . grad x_q [x Gc_e l0], sum_q N_a (2 H_q + Gc/l0) x_q wdet);
//       the node pass forms Gc l0 s0 + s1 (s0 + s1 with a per-element Gc field);
// OP 1: (sum_q N_a 2 H_q wdet, Gc l0 sum_q wdet |grad N_a|^2 + sum_q N_a^2 (2 H_q + Gc/l0) wdet);
// OP 2: x |-> the Gauss-point projection weights sum_q N_a f_q wdet of a field f (E,4)
//       (core/mesh.py:407-428), one value per corner (the second is 0).
// CGOP (a T_* pass of the damage CG, or -1): the pass is skipped as the node pass is.
template <int OP, int CGOP>
__global__ __launch_bounds__(256) void q4u_damage_elem_kernel(const T3Mesh m, const T3Mat M, const double *H,
                                                              const double *Gc, const double *x, const DamageCG *cg,
                                                              double *ce) {
    if (CGOP >= 0) {
        const DamageCG &S = cg[0];
        if (CGOP != T_CG_INIT && CGOP != T_PC_INIT && S.done) return;
        if (CGOP == T_CG_RESID && !(S.limited || S.check)) return;
        if (CGOP == T_PC_RESID && !S.check) return;
    }
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= m.n_elems) return;
    Q4uElem Q;
    q4u_geometry(m, e, Q);
    double c0[4] = {0.0, 0.0, 0.0, 0.0}, c1[4] = {0.0, 0.0, 0.0, 0.0};
    if (OP == 2) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double fw = H[4 * e + q] * Q.wd[q];
#pragma unroll
            for (int a = 0; a < 4; ++a) c0[a] += Q.N[q][a] * fw;
        }
    } else {
        const double ge = Gc ? Gc[e] : 0.0;
        const double ratio = Gc ? ge / M.l0 : M.Gc_over_l0;  // damage_solver.py:2274-2279
        const double gl = Gc ? ge * M.l0 : M.Gc_l0;
        double X[4];
        if (OP == 0) {
#pragma unroll
            for (int a = 0; a < 4; ++a) X[a] = x[Q.n[a]];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double rd = 2.0 * H[4 * e + q] + ratio;
            if (OP == 0) {
                double gdx = 0.0, gdy = 0.0, xq = 0.0;
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    gdx += Q.gp[q][a][0] * X[a];
                    gdy += Q.gp[q][a][1] * X[a];
                    xq += Q.N[q][a] * X[a];
                }
                const double mq = (rd * xq) * Q.wd[q];
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    c0[a] += Q.wd[q] * (Q.gp[q][a][0] * gdx + Q.gp[q][a][1] * gdy);
                    c1[a] += Q.N[q][a] * mq;
                }
            } else {
                const double bw = (2.0 * H[4 * e + q]) * Q.wd[q], rw = rd * Q.wd[q];
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    c0[a] += Q.N[q][a] * bw;
                    c1[a] += (Q.N[q][a] * Q.N[q][a]) * rw;
                }
            }
        }
        if (OP == 0 && Gc) {
#pragma unroll
            for (int a = 0; a < 4; ++a) c0[a] = gl * c0[a];
        }
        if (OP == 1) {  // the Laplacian diagonal first, then the mass diagonal added
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                double lap = 0.0;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    lap += Q.wd[q] * (Q.gp[q][a][0] * Q.gp[q][a][0] + Q.gp[q][a][1] * Q.gp[q][a][1]);
                c1[a] = gl * lap + c1[a];
            }
        }
    }
    double *c = m.ce + 8 * e;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        c[2 * a] = c0[a];
        c[2 * a + 1] = c1[a];
    }
}

// Nodal projection of a Gauss-point field on quadrilaterals: sum of the corner weights over
// the node's incidences / (M + 1e-30) (core/mesh.py:407-428).
__global__ __launch_bounds__(256) void q4u_e2n_kernel(const T3Mesh m, const double *Mn, double *out) {
    const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (n >= m.n_nodes) return;
    double s = 0.0;
    for (int32_t k = m.ptr[n]; k < m.ptr[n + 1]; ++k) s = __dadd_rn(s, m.ce[2 * (int64_t)m.inc[k]]);
    out[n] = s / (Mn[n] + 1e-30);
}

struct T3Cg {
    const double *H, *Gc;
    double *d, *r, *z, *p, *Ap, *lb;
    const double *b, *minv, *init, *dprev, *vals;
    const uint8_t *fixed;
    uint8_t *active, *active_out;
    double *d_out;
    DamageCG *cg;
    double *part;
    int32_t nblk;
};

template <int OP, int SRC>
__global__ __launch_bounds__(256) void t3_cg_kernel(const T3Mesh m, const T3Patch P, const T3Mat M, T3Cg A) {
    extern __shared__ double lds[];
    __shared__ double red[12];
    const DamageCG &S = A.cg[0];
    if (OP != T_CG_INIT && OP != T_PC_INIT && S.done) return;
    if (OP == T_CG_RESID && !(S.limited || S.check)) return;
    if (OP == T_PC_RESID && !S.check) return;
    constexpr double INF = __builtin_huge_val();
    const double *x = (OP == T_CG_AP || OP == T_PC_AP) ? A.p : A.d;
    auto corners = [&](int4 t, Corners &c, bool) { t3_damage_corners<0>(m, M, A.H, A.Gc, x, t, c); };
    const int blk = xcd_tile(blockIdx.x, gridDim.x);  // patch of this block
    if (SRC == SRC_PATCH) t3_patch_fill(P, lds, corners, blk);
    const int64_t slot = (int64_t)blk * 256 + threadIdx.x;
    double p0 = 0.0, p1 = (OP == T_CG_AP) ? INF : 0.0, p2 = (OP == T_CG_AP) ? INF : 0.0;
    if (slot < m.n_nodes) {
        const int64_t n = SRC == SRC_PATCH ? (int64_t)P.pnode[slot] : slot;
        double s0, s1;
        t3_node_sum<SRC>(m, P, lds, slot, n, corners, s0, s1);
        const double g0 = A.Gc ? s0 + s1 : M.Gc_l0 * s0 + s1;
        const double xn = x[n];
        if (OP == T_CG_INIT || OP == T_CG_RESID) {
            const double ob = A.b[n], olb = A.lb[n];
            const bool ofx = A.fixed && A.fixed[n];
            double r = ob - g0;
            if (ofx) r = 0.0;
            // damage_solver.py:109-124 (projected_damage_active_mask)
            const bool act = ((xn <= olb + 1e-14) && (r < 0.0)) || ((xn >= 1.0 - 1e-14) && (r > 0.0)) || ofx;
            const double rf = act ? 0.0 : r;
            A.r[n] = r;
            if (OP == T_CG_RESID) {
                p1 = (act != (A.active[n] != 0)) ? 1.0 : 0.0;
                p0 = rf * rf;
            } else {
                double zz = rf * A.minv[n];
                if (act) zz = 0.0;
                A.z[n] = zz;
                A.p[n] = zz;
                p0 = r * r;
                p1 = ob * ob;
                p2 = rf * zz;
            }
            A.active[n] = act ? 1 : 0;
        } else if (OP == T_CG_AP) {
            const double od = A.d[n], olb = A.lb[n];
            A.Ap[n] = g0;
            p0 = xn * g0;
            if (xn > 0.0) p1 = (1.0 - od) / xn;
            if (xn < 0.0) p2 = (olb - od) / xn;
        } else if (OP == T_PC_INIT) {
            const double ob = A.b[n];
            const bool ofx = A.fixed && A.fixed[n];
            double r = ob - g0;
            if (ofx) r = 0.0;
            const double zz = r * (ofx ? 0.0 : A.minv[n]);
            A.r[n] = r;
            A.z[n] = zz;
            A.p[n] = zz;
            p0 = r * r;
            p1 = ob * ob;
            p2 = r * zz;
        } else if (OP == T_PC_AP) {
            A.Ap[n] = g0;
            p0 = xn * g0;
        } else {
            const bool ofx = A.fixed && A.fixed[n];
            double r = A.b[n] - g0;
            if (ofx) r = 0.0;
            A.r[n] = r;
        }
    }
    if (OP == T_PC_RESID) return;
    // per-block partials; the reductions of one pass share their barriers (same bits as one
    // block_sum / block_min each), indexed by patch
    const int b = blk;
    if (OP == T_CG_AP) {
        block_sum_min2<256>(p0, p1, p2, red);
        if (threadIdx.x == 0) {
            A.part[b] = p0;
            A.part[A.nblk + b] = p1;
            A.part[2 * A.nblk + b] = p2;
        }
    } else if (OP == T_PC_AP) {
        const double s0 = block_sum<256>(p0, red);
        if (threadIdx.x == 0) A.part[b] = s0;
    } else if (OP == T_CG_INIT || OP == T_PC_INIT) {
        double v[3] = {p0, p1, p2};
        block_sum_n<256, 3>(v, red);
        if (threadIdx.x == 0) {
            A.part[b] = v[0];
            A.part[A.nblk + b] = v[1];
            A.part[2 * A.nblk + b] = v[2];
        }
    } else {
        double v[2] = {p0, p1};
        block_sum_n<256, 2>(v, red);
        if (threadIdx.x == 0) {
            A.part[b] = v[0];
            A.part[A.nblk + b] = v[1];
        }
    }
}

template <int OP0>
__global__ __launch_bounds__(256) void t3_vec_kernel(int64_t nn, T3Cg A) {
    __shared__ double red[12];
    const DamageCG &S = A.cg[0];
    if (OP0 != TV_SETUP && OP0 != TV_FINISH && OP0 != TV_PC_FINISH && OP0 != TA_SETUP && OP0 != TA_FINISH && S.done)
        return;
    if (OP0 == TV_PRECOND_CHK && !(S.limited || S.check)) return;  // done by TV_UPDATE_D_PRE
    if (OP0 == TV_PC_PRECOND_CHK && !S.check) return;             // done by TV_PC_UPDATE_PRE
    constexpr int OP = OP0 == TV_PRECOND_CHK ? TV_PRECOND
                       : OP0 == TV_PC_PRECOND_CHK ? TV_PC_PRECOND
                       : OP0 == TV_UPDATE_D_PRE ? TV_UPDATE_D
                       : OP0 == TV_PC_UPDATE_PRE ? TV_PC_UPDATE
                                                 : OP0;
    const bool fuse = (OP0 == TV_UPDATE_D_PRE && !(S.limited || S.check)) || (OP0 == TV_PC_UPDATE_PRE && !S.check);
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double p0 = 0.0, p1 = 0.0;
    if (q < nn) {
        if (OP == TV_SETUP) {
            const bool fx = A.fixed && A.fixed[q];
            const double dp = A.dprev[q];
            const double v = A.fixed ? A.vals[q] : 0.0;
            A.d[q] = fx ? v : (A.init ? A.init[q] : dp);
            A.lb[q] = fx ? v : dp;
        } else if (OP == TV_UPDATE_D) {
            if (S.limited) {
                double dn = A.d[q] + S.alpha * A.p[q];
                const double l = A.lb[q];
                dn = dn >= l ? dn : l;      // torch.maximum(d, lb)
                dn = dn <= 1.0 ? dn : 1.0;  // clamp(max=1)
                A.d[q] = dn;
            } else {
                A.d[q] = A.d[q] + S.alpha * A.p[q];
                const double rn = A.r[q] - S.alpha * A.Ap[q];
                A.r[q] = rn;
                if (fuse) {  // as TV_PRECOND
                    const bool act = A.active[q] != 0;
                    const double rf = act ? 0.0 : rn;
                    double zz = rf * A.minv[q];
                    if (act) zz = 0.0;
                    A.z[q] = zz;
                    p0 = rf * zz;
                }
            }
        } else if (OP == TV_PRECOND) {
            const bool act = A.active[q] != 0;
            const double rf = act ? 0.0 : A.r[q];
            double zz = rf * A.minv[q];
            if (act) zz = 0.0;
            A.z[q] = zz;
            p0 = rf * zz;
        } else if (OP == TV_UPDATE_P) {
            double pn = S.restart ? A.z[q] : A.p[q] * S.beta + A.z[q];
            if (A.active[q]) pn = 0.0;
            A.p[q] = pn;
        } else if (OP == TV_FINISH || OP == TV_PC_FINISH) {
            double dn = A.d[q];
            const double l = A.lb[q];
            dn = dn >= l ? dn : l;
            if (OP == TV_FINISH) {
                dn = dn <= 1.0 ? dn : 1.0;
            } else {
                dn = fmin(fmax(dn, 0.0), 1.0);
            }
            if (A.fixed && A.fixed[q]) dn = A.vals[q];
            A.d_out[q] = dn;
            if (OP == TV_FINISH && A.active_out) A.active_out[q] = A.active[q] ? 1 : 0;
        } else if (OP == TV_PC_UPDATE) {
            A.d[q] = A.d[q] + S.alpha * A.p[q];
            const double rn = A.r[q] - S.alpha * A.Ap[q];
            A.r[q] = rn;
            if (fuse) {  // as TV_PC_PRECOND
                const double zz = rn * ((A.fixed && A.fixed[q]) ? 0.0 : A.minv[q]);
                A.z[q] = zz;
                p0 = rn * rn;
                p1 = rn * zz;
            }
        } else if (OP == TV_PC_PRECOND) {
            const double rr = A.r[q];
            const double zz = rr * ((A.fixed && A.fixed[q]) ? 0.0 : A.minv[q]);
            A.z[q] = zz;
            p0 = rr * rr;
            p1 = rr * zz;
        } else if (OP == TV_PC_UPDATE_P) {
            A.p[q] = A.p[q] * S.beta + A.z[q];
        } else if (OP == TA_SETUP || OP == TA_FINISH) {
            const bool fx = A.fixed && A.fixed[q];
            const double v = A.vals ? A.vals[q] : 0.0;
            const double x = OP == TA_FINISH ? A.d[q] : (A.init ? A.init[q] : 0.0);
            A.d[q] = fx ? v : x;
        } else if (OP == TA_UPDATE) {
            const bool fx = A.fixed && A.fixed[q];
            A.d[q] = A.d[q] + S.alpha * A.p[q];
            double rn = A.r[q] - S.alpha * A.Ap[q];
            if (fx) rn = 0.0;
            A.r[q] = rn;
            const double zz = rn * (fx ? 0.0 : A.minv[q]);
            A.z[q] = zz;
            p0 = rn * rn;
            p1 = rn * zz;
        } else if (OP == TA_UPDATE_P) {
            double pn = A.z[q] + S.beta * A.p[q];
            if (A.fixed && A.fixed[q]) pn = 0.0;
            A.p[q] = pn;
        }
    }
    if (OP == TV_PC_PRECOND || OP == TA_UPDATE || (OP0 == TV_PC_UPDATE_PRE && fuse)) {
        double v[2] = {p0, p1};
        block_sum_n<256, 2>(v, red);
        if (threadIdx.x == 0) {
            A.part[blockIdx.x] = v[0];
            A.part[A.nblk + blockIdx.x] = v[1];
        }
    } else if (OP == TV_PRECOND || fuse) {
        const double s0 = block_sum<256>(p0, red);
        if (threadIdx.x == 0) A.part[blockIdx.x] = s0;
    }
}

template <int OP>
__global__ __launch_bounds__(256) void t3_scalar_kernel(DamageCG *cg, const double *part, int32_t nblk,
                                                        int32_t check_every) {
    __shared__ double red[8];
    DamageCG &S = cg[0];
    if (OP != S_INIT && OP != S_PC_INIT && OP != S_AJ_INIT && S.done) return;
    scalar_step<OP, 256>(0, nblk, S, part, nblk, check_every, red);
}

}  // namespace phast

using namespace phast;

struct phast_t3_plan {
    T3Mesh mesh;
    T3Mat mat;
    double2 *xy = nullptr;
    int32_t *el = nullptr, *ptr = nullptr, *inc = nullptr;
    std::vector<double> M_scalar;  // lumped mass per node (area/3 scatter), host copy
    // damage-solve workspace, allocated on the first solve
    double *r = nullptr, *p = nullptr, *z = nullptr, *Ap = nullptr, *b = nullptr, *minv = nullptr, *lb = nullptr;
    double *part = nullptr;
    uint8_t *active = nullptr;
    DamageCG *cg = nullptr, *h_cg = nullptr;
    int32_t nblk = 0;
    double h_min = 0.0, total_area = 0.0;
    // element patches (see T3Patch); use_patch selects them, lds_bytes sizes a block's LDS
    int32_t *pnode = nullptr, *pel_ptr = nullptr, *sptr = nullptr;
    int4 *pel = nullptr;
    uint16_t *sloc = nullptr;
    T3Patch patch{};
    bool has_patch = false, use_patch = false;
    int32_t max_list = 0;
    size_t lds_bytes = 0;
    int32_t K = 3;          // corners per element: 3 triangles, 4 quadrilaterals
    double *ce = nullptr;   // quadrilaterals: corner values of the element passes (E,4,2)
};

namespace {

constexpr int32_t kPatch = 256;                 // nodes per patch (= threads per block)
constexpr size_t kPatchLdsMax = 60 * 1024;      // 6 doubles per listed triangle (+ static LDS < 64 KiB)
constexpr int32_t kPatchListMax = (int32_t)(kPatchLdsMax / 48);

inline uint64_t spread_bits(uint64_t v) {  // bit i -> bit 2i (Morton interleave)
    v &= 0xffffffffull;
    v = (v | (v << 16)) & 0x0000ffff0000ffffull;
    v = (v | (v << 8)) & 0x00ff00ff00ff00ffull;
    v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0full;
    v = (v | (v << 2)) & 0x3333333333333333ull;
    v = (v | (v << 1)) & 0x5555555555555555ull;
    return v;
}

// Patches of kPatch nodes in the Morton order of the coordinates, each with the ascending
// list of triangles touching it and, per CSR entry, the LDS slot of its corner.  Leaves the
// plan without patches when a list would not fit the LDS budget.
int t3_build_patches(phast_t3_plan *pl, int64_t nn, int64_t ne, const double *nodes, const int32_t *els,
                     const std::vector<int32_t> &ptr, const std::vector<int32_t> &inc) {
    double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
    for (int64_t i = 0; i < nn; ++i) {
        x0 = std::min(x0, nodes[2 * i]);
        x1 = std::max(x1, nodes[2 * i]);
        y0 = std::min(y0, nodes[2 * i + 1]);
        y1 = std::max(y1, nodes[2 * i + 1]);
    }
    const double sx = x1 > x0 ? 2147483647.0 / (x1 - x0) : 0.0, sy = y1 > y0 ? 2147483647.0 / (y1 - y0) : 0.0;
    std::vector<std::pair<uint64_t, int32_t>> key((size_t)nn);
    for (int64_t i = 0; i < nn; ++i) {
        const uint64_t qx = (uint64_t)((nodes[2 * i] - x0) * sx), qy = (uint64_t)((nodes[2 * i + 1] - y0) * sy);
        key[(size_t)i] = {spread_bits(qx) | (spread_bits(qy) << 1), (int32_t)i};
    }
    std::sort(key.begin(), key.end());
    const int64_t npatch = ceil_div(nn, kPatch);
    std::vector<int32_t> pnode((size_t)nn), owner_patch((size_t)nn);
    for (int64_t s = 0; s < nn; ++s) {
        pnode[(size_t)s] = key[(size_t)s].second;
        owner_patch[(size_t)key[(size_t)s].second] = (int32_t)(s / kPatch);
    }
    std::vector<int32_t> pel_ptr((size_t)npatch + 1, 0), sptr((size_t)nn + 1, 0);
    std::vector<int4> pel;
    pel.reserve((size_t)(ne * 3 / 2));
    std::vector<uint16_t> sloc((size_t)(3 * ne));
    std::vector<int32_t> list;
    int32_t max_list = 0, nk = 0;
    for (int64_t b = 0; b < npatch; ++b) {
        const int64_t s0 = b * kPatch, s1 = std::min<int64_t>(nn, s0 + kPatch);
        list.clear();
        for (int64_t s = s0; s < s1; ++s) {
            const int32_t n = pnode[(size_t)s];
            for (int32_t k = ptr[(size_t)n]; k < ptr[(size_t)n + 1]; ++k) list.push_back(inc[(size_t)k] / 3);
        }
        std::sort(list.begin(), list.end());
        list.erase(std::unique(list.begin(), list.end()), list.end());
        max_list = std::max<int32_t>(max_list, (int32_t)list.size());
        if ((int32_t)list.size() > kPatchListMax) {  // no patches for this mesh
            pl->max_list = max_list;
            return PHAST_OK;
        }
        for (int64_t s = s0; s < s1; ++s) {
            const int32_t n = pnode[(size_t)s];
            for (int32_t k = ptr[(size_t)n]; k < ptr[(size_t)n + 1]; ++k) {
                const int32_t e = inc[(size_t)k] / 3, a = inc[(size_t)k] - 3 * e;
                const int32_t j = (int32_t)(std::lower_bound(list.begin(), list.end(), e) - list.begin());
                sloc[(size_t)nk++] = (uint16_t)(3 * j + a);
            }
            sptr[(size_t)s + 1] = nk;
        }
        for (int32_t e : list) {
            const int32_t *q = els + 3 * (int64_t)e;
            pel.push_back(make_int4(owner_patch[(size_t)q[0]] == b ? (e | INT32_MIN) : e, q[0], q[1], q[2]));
        }
        pel_ptr[(size_t)b + 1] = (int32_t)pel.size();
    }
    PHAST_HIP(hipMalloc(&pl->pnode, sizeof(int32_t) * pnode.size()));
    PHAST_HIP(hipMalloc(&pl->pel_ptr, sizeof(int32_t) * pel_ptr.size()));
    PHAST_HIP(hipMalloc(&pl->pel, sizeof(int4) * pel.size()));
    PHAST_HIP(hipMalloc(&pl->sptr, sizeof(int32_t) * sptr.size()));
    PHAST_HIP(hipMalloc(&pl->sloc, sizeof(uint16_t) * sloc.size()));
    PHAST_HIP(hipMemcpy(pl->pnode, pnode.data(), sizeof(int32_t) * pnode.size(), hipMemcpyHostToDevice));
    PHAST_HIP(hipMemcpy(pl->pel_ptr, pel_ptr.data(), sizeof(int32_t) * pel_ptr.size(), hipMemcpyHostToDevice));
    PHAST_HIP(hipMemcpy(pl->pel, pel.data(), sizeof(int4) * pel.size(), hipMemcpyHostToDevice));
    PHAST_HIP(hipMemcpy(pl->sptr, sptr.data(), sizeof(int32_t) * sptr.size(), hipMemcpyHostToDevice));
    PHAST_HIP(hipMemcpy(pl->sloc, sloc.data(), sizeof(uint16_t) * sloc.size(), hipMemcpyHostToDevice));
    pl->patch = T3Patch{pl->pnode, pl->pel_ptr, pl->pel, pl->sptr, pl->sloc};
    pl->max_list = max_list;
    pl->lds_bytes = 48 * (size_t)max_list;
    pl->has_patch = pl->use_patch = true;
    return PHAST_OK;
}

}  // namespace

extern "C" int phast_t3_plan_create(int64_t n_nodes, int64_t n_elems, const double *nodes, const int32_t *elements,
                                    double E, double nu, double Gc, double l0, double eta_residual,
                                    int32_t plane_stress, int32_t split, int32_t stress_split, int32_t pf_model,
                                    double at1_source, phast_t3_plan **out) {
    PHAST_REQUIRE(out && nodes && elements && n_nodes >= 3 && n_elems >= 1, "phast_t3_plan_create: bad argument");
    PHAST_REQUIRE(3 * n_elems < ((int64_t)1 << 31) && n_nodes < ((int64_t)1 << 31),
                  "phast_t3_plan_create: mesh too large for 32-bit incidence indices");
    PHAST_REQUIRE(split == T3_ISO || split == T3_AMOR || split == T3_SPECTRAL, "unknown T3 energy split %d", split);
    PHAST_REQUIRE(stress_split == T3_ISO || stress_split == T3_AMOR || stress_split == T3_SPECTRAL,
                  "unknown T3 stress split %d", stress_split);
    PHAST_REQUIRE(pf_model == PHAST_PF_AT2 || pf_model == PHAST_PF_AT1, "unknown phase-field model %d", pf_model);
    for (int64_t k = 0; k < 3 * n_elems; ++k)
        PHAST_REQUIRE(elements[k] >= 0 && elements[k] < n_nodes, "element %lld refers to node %d outside the mesh",
                      (long long)(k / 3), elements[k]);
    auto *pl = new phast_t3_plan();
    T3Mat &M = pl->mat;
    M.lam = plane_stress ? E * nu / (1.0 - nu * nu) : E * nu / ((1 + nu) * (1 - 2 * nu));  // material.py:379-388
    M.mu = E / (2 * (1 + nu));
    M.kappa = E / (3.0 * (1.0 - 2.0 * nu));
    M.ezz_coef = -nu / (1.0 - nu);
    M.eta = eta_residual;
    // damage_solver.py:1321-1336: AT1 puts 3/4 Gc l0 on the Laplacian and no Gc/l0 in the reaction
    M.at1 = pf_model == PHAST_PF_AT1;
    M.Gc_l0 = M.at1 ? 0.75 * Gc * l0 : Gc * l0;
    M.Gc_over_l0 = M.at1 ? 0.0 : Gc / l0;
    M.src = M.at1 ? at1_source : 0.0;
    M.l0 = l0;
    M.split = split;
    M.ssplit = stress_split;
    M.plane_stress = plane_stress;
    // host geometry: lumped mass in elem_flat order and h_min (core/mesh.py:270-286)
    pl->M_scalar.assign((size_t)n_nodes, 0.0);
    double hmin = INFINITY, tot = 0.0;
    std::vector<int32_t> cnt((size_t)n_nodes + 1, 0);
    for (int64_t e = 0; e < n_elems; ++e) {
        const int32_t *q = elements + 3 * e;
        const double *p0 = nodes + 2 * q[0], *p1 = nodes + 2 * q[1], *p2 = nodes + 2 * q[2];
        const double v01x = p1[0] - p0[0], v01y = p1[1] - p0[1], v02x = p2[0] - p0[0], v02y = p2[1] - p0[1];
        const double v12x = p2[0] - p1[0], v12y = p2[1] - p1[1];
        const double area = 0.5 * std::fabs(v01x * v02y - v01y * v02x);
        const double per = std::sqrt(v01x * v01x + v01y * v01y) + std::sqrt(v12x * v12x + v12y * v12y) +
                           std::sqrt(v02x * v02x + v02y * v02y);
        hmin = std::min(hmin, 4.0 * area / (per + 1e-30));
        tot += area;
        for (int k = 0; k < 3; ++k) {
            pl->M_scalar[(size_t)q[k]] += area / 3.0;
            ++cnt[(size_t)q[k] + 1];
        }
    }
    pl->h_min = hmin;
    pl->total_area = tot;
    for (int64_t i = 0; i < n_nodes; ++i) cnt[(size_t)i + 1] += cnt[(size_t)i];
    std::vector<int32_t> inc((size_t)(3 * n_elems)), fill(cnt.begin(), cnt.end() - 1);
    for (int64_t k = 0; k < 3 * n_elems; ++k) inc[(size_t)fill[(size_t)elements[k]]++] = (int32_t)k;
    if (hipMalloc(&pl->xy, sizeof(double2) * (size_t)n_nodes) != hipSuccess ||
        hipMalloc(&pl->el, sizeof(int32_t) * 3 * (size_t)n_elems) != hipSuccess ||
        hipMalloc(&pl->ptr, sizeof(int32_t) * (size_t)(n_nodes + 1)) != hipSuccess ||
        hipMalloc(&pl->inc, sizeof(int32_t) * 3 * (size_t)n_elems) != hipSuccess ||
        hipMemcpy(pl->xy, nodes, sizeof(double2) * (size_t)n_nodes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pl->el, elements, sizeof(int32_t) * 3 * (size_t)n_elems, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pl->ptr, cnt.data(), sizeof(int32_t) * (size_t)(n_nodes + 1), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pl->inc, inc.data(), sizeof(int32_t) * 3 * (size_t)n_elems, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("phast_t3_plan_create: device allocation/copy failed");
        phast_t3_plan_destroy(pl);
        return PHAST_ERR_HIP;
    }
    pl->mesh = T3Mesh{pl->xy, pl->el, pl->ptr, pl->inc, n_nodes, n_elems};
    if (t3_build_patches(pl, n_nodes, n_elems, nodes, elements, cnt, inc) != PHAST_OK) {
        phast_t3_plan_destroy(pl);
        return PHAST_ERR_HIP;
    }
    *out = pl;
    return PHAST_OK;
}

// Plan of a general quadrilateral mesh (the reference's Q4 branch on any quad mesh): the same
// plan type, K = 4 corners per element, no patches (every pass is an element pass writing
// corner values and a node pass summing them).  Host set-up mirrors core/mesh.py:305-352
// _precompute_q4: Gauss-point Jacobians (every one must be positive, quad_elements.py:
// 331-333), the lumped mass sum_q N_a wdetJ scattered per node, h_min = the smallest edge.
extern "C" int phast_q4u_plan_create(int64_t n_nodes, int64_t n_elems, const double *nodes, const int32_t *elements,
                                     double E, double nu, double Gc, double l0, double eta_residual,
                                     int32_t plane_stress, int32_t split, int32_t stress_split, phast_t3_plan **out) {
    PHAST_REQUIRE(out && nodes && elements && n_nodes >= 4 && n_elems >= 1, "phast_q4u_plan_create: bad argument");
    PHAST_REQUIRE(4 * n_elems < ((int64_t)1 << 31) && n_nodes < ((int64_t)1 << 31),
                  "phast_q4u_plan_create: mesh too large for 32-bit incidence indices");
    PHAST_REQUIRE(split == T3_ISO || split == T3_AMOR, "the quadrilateral path implements the amor and isotropic "
                  "splits, got %d", split);
    PHAST_REQUIRE(stress_split == T3_ISO || stress_split == T3_AMOR, "unknown stress split %d", stress_split);
    for (int64_t k = 0; k < 4 * n_elems; ++k)
        PHAST_REQUIRE(elements[k] >= 0 && elements[k] < n_nodes, "element %lld refers to node %d outside the mesh",
                      (long long)(k / 4), elements[k]);
    const double g = 1.0 / std::sqrt(3.0);
    std::vector<double> M_scalar((size_t)n_nodes, 0.0);
    double hmin = INFINITY, tot = 0.0;
    std::vector<int32_t> cnt((size_t)n_nodes + 1, 0);
    for (int64_t e = 0; e < n_elems; ++e) {
        const int32_t *q4 = elements + 4 * e;
        double wd[4], Nq[4][4];
        for (int q = 0; q < 4; ++q) {
            const double xi = (q & 1) ? g : -g, eta = (q & 2) ? g : -g;
            const double dxi[4] = {0.25 * -(1.0 - eta), 0.25 * (1.0 - eta), 0.25 * (1.0 + eta), 0.25 * -(1.0 + eta)};
            const double det_[4] = {0.25 * -(1.0 - xi), 0.25 * -(1.0 + xi), 0.25 * (1.0 + xi), 0.25 * (1.0 - xi)};
            Nq[q][0] = 0.25 * ((1.0 - xi) * (1.0 - eta));
            Nq[q][1] = 0.25 * ((1.0 + xi) * (1.0 - eta));
            Nq[q][2] = 0.25 * ((1.0 + xi) * (1.0 + eta));
            Nq[q][3] = 0.25 * ((1.0 - xi) * (1.0 + eta));
            double J00 = 0.0, J01 = 0.0, J10 = 0.0, J11 = 0.0;
            for (int a = 0; a < 4; ++a) {
                const double *X = nodes + 2 * q4[a];
                J00 += dxi[a] * X[0];
                J01 += dxi[a] * X[1];
                J10 += det_[a] * X[0];
                J11 += det_[a] * X[1];
            }
            wd[q] = J00 * J11 - J01 * J10;
            PHAST_REQUIRE(wd[q] > 0.0, "Q4 element %lld has a non-positive Jacobian at Gauss point %d (check the "
                          "counter-clockwise corner order)", (long long)e, q);
        }
        double hmin_e = INFINITY;
        for (int k = 0; k < 4; ++k) {
            const double *A = nodes + 2 * q4[k], *B = nodes + 2 * q4[(k + 1) % 4];
            hmin_e = std::min(hmin_e, std::sqrt((B[0] - A[0]) * (B[0] - A[0]) + (B[1] - A[1]) * (B[1] - A[1])));
        }
        hmin = std::min(hmin, hmin_e);
        for (int a = 0; a < 4; ++a) {
            double m = 0.0;
            for (int q = 0; q < 4; ++q) m += Nq[q][a] * wd[q];
            M_scalar[(size_t)q4[a]] += m;
            ++cnt[(size_t)q4[a] + 1];
        }
        tot += ((wd[0] + wd[1]) + wd[2]) + wd[3];
    }
    auto *pl = new phast_t3_plan();
    pl->K = 4;
    T3Mat &M = pl->mat;
    M.lam = plane_stress ? E * nu / (1.0 - nu * nu) : E * nu / ((1 + nu) * (1 - 2 * nu));  // material.py:379-388
    M.mu = E / (2 * (1 + nu));
    M.kappa = E / (3.0 * (1.0 - 2.0 * nu));
    M.ezz_coef = -nu / (1.0 - nu);
    M.eta = eta_residual;
    M.at1 = 0;
    M.Gc_l0 = Gc * l0;
    M.Gc_over_l0 = Gc / l0;
    M.src = 0.0;
    M.l0 = l0;
    M.split = split;
    M.ssplit = stress_split;
    M.plane_stress = plane_stress;
    pl->M_scalar = std::move(M_scalar);
    pl->h_min = hmin;
    pl->total_area = tot;
    for (int64_t i = 0; i < n_nodes; ++i) cnt[(size_t)i + 1] += cnt[(size_t)i];
    std::vector<int32_t> inc((size_t)(4 * n_elems)), fill(cnt.begin(), cnt.end() - 1);
    for (int64_t k = 0; k < 4 * n_elems; ++k) inc[(size_t)fill[(size_t)elements[k]]++] = (int32_t)k;
    if (hipMalloc(&pl->xy, sizeof(double2) * (size_t)n_nodes) != hipSuccess ||
        hipMalloc(&pl->el, sizeof(int32_t) * 4 * (size_t)n_elems) != hipSuccess ||
        hipMalloc(&pl->ptr, sizeof(int32_t) * (size_t)(n_nodes + 1)) != hipSuccess ||
        hipMalloc(&pl->inc, sizeof(int32_t) * 4 * (size_t)n_elems) != hipSuccess ||
        hipMalloc(&pl->ce, sizeof(double) * 8 * (size_t)n_elems) != hipSuccess ||
        hipMemcpy(pl->xy, nodes, sizeof(double2) * (size_t)n_nodes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pl->el, elements, sizeof(int32_t) * 4 * (size_t)n_elems, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pl->ptr, cnt.data(), sizeof(int32_t) * (size_t)(n_nodes + 1), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pl->inc, inc.data(), sizeof(int32_t) * 4 * (size_t)n_elems, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("phast_q4u_plan_create: device allocation/copy failed");
        phast_t3_plan_destroy(pl);
        return PHAST_ERR_HIP;
    }
    pl->mesh = T3Mesh{pl->xy, pl->el, pl->ptr, pl->inc, n_nodes, n_elems, pl->ce};
    *out = pl;
    return PHAST_OK;
}

extern "C" int phast_t3_plan_destroy(phast_t3_plan *pl) {
    if (!pl) return PHAST_OK;
    (void)hipFree(pl->xy);
    (void)hipFree(pl->el);
    (void)hipFree(pl->ptr);
    (void)hipFree(pl->inc);
    (void)hipFree(pl->pnode);
    (void)hipFree(pl->pel_ptr);
    (void)hipFree(pl->pel);
    (void)hipFree(pl->sptr);
    (void)hipFree(pl->sloc);
    (void)hipFree(pl->ce);
    for (double *q : {pl->r, pl->p, pl->z, pl->Ap, pl->b, pl->minv, pl->lb, pl->part}) (void)hipFree(q);
    (void)hipFree(pl->active);
    (void)hipFree(pl->cg);
    if (pl->h_cg) (void)hipHostFree(pl->h_cg);
    delete pl;
    return PHAST_OK;
}

extern "C" int phast_t3_mesh_info(const phast_t3_plan *pl, double *M_scalar, double *h_min, double *total_area) {
    PHAST_REQUIRE(pl, "phast_t3_mesh_info: null plan");
    if (M_scalar) std::memcpy(M_scalar, pl->M_scalar.data(), sizeof(double) * pl->M_scalar.size());
    if (h_min) *h_min = pl->h_min;
    if (total_area) *total_area = pl->total_area;
    return PHAST_OK;
}

static inline dim3 node_grid(const phast_t3_plan *pl) { return dim3((unsigned)ceil_div(pl->mesh.n_nodes, 256)); }

extern "C" int phast_t3_set_path(phast_t3_plan *pl, int32_t path, int32_t *active_path, int32_t *max_list) {
    PHAST_REQUIRE(pl && (path == 0 || path == 1), "phast_t3_set_path: path must be 0 (patches) or 1 (recompute)");
    pl->use_patch = path == 0 && pl->has_patch;
    if (active_path) *active_path = pl->use_patch ? 0 : 1;
    if (max_list) *max_list = pl->max_list;
    return PHAST_OK;
}

static inline dim3 elem_grid(const phast_t3_plan *pl) { return dim3((unsigned)ceil_div(pl->mesh.n_elems, 256)); }

template <int MODE>
static int launch_force(const phast_t3_plan *pl, const double *u, const double *d, double *f, const double *H_old,
                        double *H_new, double *psi, const T3Step &S, hipStream_t st) {
    if (pl->K == 4) {  // element pass, then the node pass over the corner values
        if (MODE >= 1)
            hipLaunchKernelGGL(q4u_force_elem_kernel<true>, elem_grid(pl), dim3(256), 0, st, pl->mesh, pl->mat,
                               (const double2 *)u, d, H_old, H_new, psi, pl->ce);
        else
            hipLaunchKernelGGL(q4u_force_elem_kernel<false>, elem_grid(pl), dim3(256), 0, st, pl->mesh, pl->mat,
                               (const double2 *)u, d, H_old, H_new, psi, pl->ce);
        PHAST_LAUNCHED();
        hipLaunchKernelGGL((t3_force_kernel<MODE, SRC_BUFFER>), node_grid(pl), dim3(256), 0, st, pl->mesh, pl->patch,
                           pl->mat, (const double2 *)u, d, (double2 *)f, H_old, H_new, psi, S);
    } else if (pl->use_patch) {
        hipLaunchKernelGGL((t3_force_kernel<MODE, SRC_PATCH>), node_grid(pl), dim3(256), pl->lds_bytes, st, pl->mesh,
                           pl->patch, pl->mat, (const double2 *)u, d, (double2 *)f, H_old, H_new, psi, S);
    } else {
        hipLaunchKernelGGL((t3_force_kernel<MODE, SRC_RECOMPUTE>), node_grid(pl), dim3(256), 0, st, pl->mesh,
                           pl->patch, pl->mat, (const double2 *)u, d, (double2 *)f, H_old, H_new, psi, S);
    }
    PHAST_LAUNCHED();
    return PHAST_OK;
}

template <int OP>
static int launch_damage(const phast_t3_plan *pl, const double *H, const double *Gc, const double *x, double *y,
                         double *y2, hipStream_t st) {
    if (pl->K == 4) {
        hipLaunchKernelGGL((q4u_damage_elem_kernel<(OP == 0 ? 0 : 1), -1>), elem_grid(pl), dim3(256), 0, st, pl->mesh,
                           pl->mat, H, Gc, x, (const DamageCG *)nullptr, pl->ce);
        PHAST_LAUNCHED();
        hipLaunchKernelGGL((t3_damage_kernel<OP, SRC_BUFFER>), node_grid(pl), dim3(256), 0, st, pl->mesh, pl->patch,
                           pl->mat, H, Gc, x, y, y2);
    } else if (pl->use_patch) {
        hipLaunchKernelGGL((t3_damage_kernel<OP, SRC_PATCH>), node_grid(pl), dim3(256), pl->lds_bytes, st, pl->mesh,
                           pl->patch, pl->mat, H, Gc, x, y, y2);
    } else {
        hipLaunchKernelGGL((t3_damage_kernel<OP, SRC_RECOMPUTE>), node_grid(pl), dim3(256), 0, st, pl->mesh, pl->patch,
                           pl->mat, H, Gc, x, y, y2);
    }
    PHAST_LAUNCHED();
    return PHAST_OK;
}

extern "C" int phast_t3_internal_force(const phast_t3_plan *pl, const double *u, const double *d, double *f,
                                       hipStream_t st) {
    PHAST_REQUIRE(pl && u && d && f, "phast_t3_internal_force: null argument");
    T3Step S{};
    S.fscale = 1.0;
    return launch_force<0>(pl, u, d, f, nullptr, nullptr, nullptr, S, st);
}

extern "C" int phast_t3_damping_force(const phast_t3_plan *pl, const double *v, const double *d, double beta,
                                      double *f, hipStream_t st) {
    PHAST_REQUIRE(pl && v && d && f, "phast_t3_damping_force: null argument");
    T3Step S{};
    S.fscale = beta;
    return launch_force<0>(pl, v, d, f, nullptr, nullptr, nullptr, S, st);
}

extern "C" int phast_t3_force_history(const phast_t3_plan *pl, const double *u, const double *d, const double *H_old,
                                      double *f, double *H_new, double *psi, hipStream_t st) {
    PHAST_REQUIRE(pl && u && d && H_old && f && H_new, "phast_t3_force_history: null argument");
    return launch_force<1>(pl, u, d, f, H_old, H_new, psi, T3Step{}, st);
}

extern "C" int phast_t3_explicit_step(const phast_t3_plan *pl, double dt, const double *u, const double *v,
                                      const double *a, const double *d, const double *f_ext, const double *f_damp,
                                      const double *Minv,
                                      const uint8_t *bc_mask, const double *bc_vals, const double *H_old, double *u1,
                                      double *v1, double *a1, double *H_new, double *psi, hipStream_t st) {
    PHAST_REQUIRE(pl && u && v && a && d && Minv && H_old && u1 && v1 && a1 && H_new,
                  "phast_t3_explicit_step: null argument");
    PHAST_REQUIRE(!bc_mask == !bc_vals, "phast_t3_explicit_step: bc_mask and bc_vals go together");
    PHAST_REQUIRE(u1 != u && u1 != v && u1 != a && v1 != u && v1 != v && v1 != a && a1 != u && a1 != v && a1 != a,
                  "phast_t3_explicit_step: outputs alias inputs");
    T3Step S{};
    S.dt = dt;
    S.c2 = 0.5 * dt * dt;
    S.v = (const double2 *)v;
    S.a = (const double2 *)a;
    S.fext = (const double2 *)f_ext;
    S.fdamp = (const double2 *)f_damp;
    S.minv = Minv;
    S.bcm = bc_mask;
    S.bcv = bc_vals;
    S.u1 = (double2 *)u1;
    S.v1 = (double2 *)v1;
    S.a1 = (double2 *)a1;
    hipLaunchKernelGGL(t3_predictor_kernel, node_grid(pl), dim3(256), 0, st, pl->mesh.n_nodes, (const double2 *)u, S);
    PHAST_LAUNCHED();
    return launch_force<2>(pl, u1, d, nullptr, H_old, H_new, psi, S, st);
}

extern "C" int phast_t3_element_fields(const phast_t3_plan *pl, const double *u, const double *d, double *eps,
                                       double *sig, double *psi, hipStream_t st) {
    PHAST_REQUIRE(pl && u && (eps || sig || psi), "phast_t3_element_fields: null argument");
    PHAST_REQUIRE(d || !sig, "phast_t3_element_fields: the stress needs the damage d");
    if (pl->K == 4) {
        PHAST_REQUIRE(!sig, "phast_t3_element_fields: quadrilateral plans give strain and psi+ only");
        hipLaunchKernelGGL(q4u_fields_kernel, elem_grid(pl), dim3(256), 0, st, pl->mesh, pl->mat, (const double2 *)u,
                           eps, psi);
        PHAST_LAUNCHED();
        return PHAST_OK;
    }
    hipLaunchKernelGGL(t3_elem_fields_kernel, dim3((unsigned)ceil_div(pl->mesh.n_elems, 256)), dim3(256), 0, st,
                       pl->mesh, pl->mat, (const double2 *)u, d, eps, sig, psi);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

extern "C" int phast_t3_elem_to_node(const phast_t3_plan *pl, const double *field_e, const double *M, double *out,
                                     hipStream_t st) {
    PHAST_REQUIRE(pl && field_e && M && out, "phast_t3_elem_to_node: null argument");
    if (pl->K == 4) {  // Gauss-point field (E,4)
        hipLaunchKernelGGL((q4u_damage_elem_kernel<2, -1>), elem_grid(pl), dim3(256), 0, st, pl->mesh, pl->mat,
                           field_e, (const double *)nullptr, (const double *)nullptr, (const DamageCG *)nullptr,
                           pl->ce);
        PHAST_LAUNCHED();
        hipLaunchKernelGGL(q4u_e2n_kernel, node_grid(pl), dim3(256), 0, st, pl->mesh, M, out);
        PHAST_LAUNCHED();
        return PHAST_OK;
    }
    hipLaunchKernelGGL(t3_e2n_kernel, node_grid(pl), dim3(256), 0, st, pl->mesh, field_e, M, out);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

extern "C" int phast_t3_damage_apply(const phast_t3_plan *pl, const double *H, const double *Gc_e, const double *x,
                                     double *y, hipStream_t st) {
    PHAST_REQUIRE(pl && H && x && y && x != y, "phast_t3_damage_apply: bad argument");
    return launch_damage<0>(pl, H, Gc_e, x, y, nullptr, st);
}

extern "C" int phast_t3_damage_rhs_diag(const phast_t3_plan *pl, const double *H, const double *Gc_e, double *b,
                                        double *diag, hipStream_t st) {
    PHAST_REQUIRE(pl && H && (b || diag), "phast_t3_damage_rhs_diag: null argument");
    return launch_damage<1>(pl, H, Gc_e, nullptr, b, diag, st);
}

#define T3_TRY(x)                         \
    do {                                  \
        const int rc_ = (x);              \
        if (rc_ != PHAST_OK) return rc_;  \
    } while (0)

static int t3_ensure_work(phast_t3_plan *pl) {
    if (pl->r) return PHAST_OK;
    const size_t nb = sizeof(double) * (size_t)pl->mesh.n_nodes;
    pl->nblk = (int32_t)ceil_div(pl->mesh.n_nodes, 256);
    for (double **q : {&pl->r, &pl->p, &pl->z, &pl->Ap, &pl->b, &pl->minv, &pl->lb}) PHAST_HIP(hipMalloc(q, nb));
    PHAST_HIP(hipMalloc(&pl->part, sizeof(double) * 3 * (size_t)pl->nblk));
    PHAST_HIP(hipMalloc(&pl->active, (size_t)pl->mesh.n_nodes));
    PHAST_HIP(hipMalloc(&pl->cg, sizeof(DamageCG)));
    PHAST_HIP(hipHostMalloc(&pl->h_cg, sizeof(DamageCG), hipHostMallocDefault));
    return PHAST_OK;
}

template <int OP>
static int t3_cg(const phast_t3_plan *pl, const T3Cg &A, hipStream_t st) {
    if (pl->K == 4) {  // the operator's element pass on the pass's field, skipped as the node pass is
        const double *x = (OP == T_CG_AP || OP == T_PC_AP) ? A.p : A.d;
        hipLaunchKernelGGL((q4u_damage_elem_kernel<0, OP>), elem_grid(pl), dim3(256), 0, st, pl->mesh, pl->mat, A.H,
                           A.Gc, x, (const DamageCG *)A.cg, pl->ce);
        PHAST_LAUNCHED();
        hipLaunchKernelGGL((t3_cg_kernel<OP, SRC_BUFFER>), node_grid(pl), dim3(256), 0, st, pl->mesh, pl->patch,
                           pl->mat, A);
    } else if (pl->use_patch) {
        hipLaunchKernelGGL((t3_cg_kernel<OP, SRC_PATCH>), node_grid(pl), dim3(256), pl->lds_bytes, st, pl->mesh,
                           pl->patch, pl->mat, A);
    } else {
        hipLaunchKernelGGL((t3_cg_kernel<OP, SRC_RECOMPUTE>), node_grid(pl), dim3(256), 0, st, pl->mesh, pl->patch,
                           pl->mat, A);
    }
    PHAST_LAUNCHED();
    return PHAST_OK;
}
template <int OP>
static int t3_vec(const phast_t3_plan *pl, const T3Cg &A, hipStream_t st) {
    hipLaunchKernelGGL(t3_vec_kernel<OP>, node_grid(pl), dim3(256), 0, st, pl->mesh.n_nodes, A);
    PHAST_LAUNCHED();
    return PHAST_OK;
}
template <int OP>
static int t3_scalar(const phast_t3_plan *pl, int check_every, hipStream_t st) {
    hipLaunchKernelGGL(t3_scalar_kernel<OP>, dim3(1), dim3(256), 0, st, pl->cg, (const double *)pl->part, pl->nblk,
                       check_every);
    PHAST_LAUNCHED();
    return PHAST_OK;
}

extern "C" int phast_t3_damage_solve(phast_t3_plan *pl, int32_t method, const double *H, const double *Gc_e,
                                     const double *d_prev, const double *d_init, const uint8_t *fixed,
                                     const double *fixed_vals, double tol, int32_t max_iter, double *d_out,
                                     uint8_t *active_out, int32_t *iters, int32_t *status, hipStream_t st) {
    PHAST_REQUIRE(pl && H && d_prev && d_out, "phast_t3_damage_solve: null argument");
    PHAST_REQUIRE(method == 0 || method == 1, "method must be 0 (projected) or 1 (post-clamp)");
    PHAST_REQUIRE(!fixed == !fixed_vals, "fixed and fixed_vals must be given together");
    PHAST_REQUIRE(tol >= 0.0 && max_iter >= 0, "tol and max_iter must be non-negative");
    PHAST_REQUIRE(d_out != d_prev && d_out != d_init, "d_out must not alias d_prev or d_init");
    T3_TRY(t3_ensure_work(pl));
    DamageCG &h = *pl->h_cg;
    std::memset(&h, 0, sizeof(h));
    h.tol = tol;
    h.max_iter = max_iter;
    PHAST_HIP(hipMemcpyAsync(pl->cg, pl->h_cg, sizeof(DamageCG), hipMemcpyHostToDevice, st));

    T3Cg A;
    std::memset(&A, 0, sizeof(A));
    A.H = H;
    A.Gc = Gc_e;
    A.d = d_out;
    A.r = pl->r;
    A.z = pl->z;
    A.p = pl->p;
    A.Ap = pl->Ap;
    A.lb = pl->lb;
    A.b = pl->b;
    A.minv = pl->minv;
    A.init = d_init;
    A.dprev = d_prev;
    A.vals = fixed_vals;
    A.fixed = fixed;
    A.active = pl->active;
    A.active_out = active_out;
    A.d_out = d_out;
    A.cg = pl->cg;
    A.part = pl->part;
    A.nblk = pl->nblk;

    // b and M_inv (damage_solver.py:2249-2265); d = init or d_prev, lb = d_prev (pins applied)
    T3_TRY(launch_damage<2>(pl, H, Gc_e, nullptr, pl->b, pl->minv, st));
    T3_TRY(t3_vec<TV_SETUP>(pl, A, st));
    const int poll = 4;
    int it = 0;
    auto all_done = [&]() -> int {
        PHAST_HIP(hipMemcpyAsync(pl->h_cg, pl->cg, sizeof(DamageCG), hipMemcpyDeviceToHost, st));
        PHAST_HIP(hipStreamSynchronize(st));
        return PHAST_OK;
    };
    if (method == 0) {  // damage_solver.py:3527-3775
        T3_TRY(t3_cg<T_CG_INIT>(pl, A, st));
        T3_TRY(t3_scalar<S_INIT>(pl, 10, st));
        while (it < max_iter) {
            T3_TRY(t3_cg<T_CG_AP>(pl, A, st));
            T3_TRY(t3_scalar<S_A>(pl, 10, st));
            T3_TRY(t3_vec<TV_UPDATE_D_PRE>(pl, A, st));
            T3_TRY(t3_cg<T_CG_RESID>(pl, A, st));
            T3_TRY(t3_scalar<S_B>(pl, 10, st));
            T3_TRY(t3_vec<TV_PRECOND_CHK>(pl, A, st));
            T3_TRY(t3_scalar<S_C>(pl, 10, st));
            T3_TRY(t3_vec<TV_UPDATE_P>(pl, A, st));
            ++it;
            if (it % poll == 0 || it == max_iter) {
                T3_TRY(all_done());
                if (h.done) break;
            }
        }
        T3_TRY(t3_vec<TV_FINISH>(pl, A, st));
    } else {  // damage_solver.py:3377-3521
        T3_TRY(t3_cg<T_PC_INIT>(pl, A, st));
        T3_TRY(t3_scalar<S_PC_INIT>(pl, 50, st));
        while (it < max_iter) {
            T3_TRY(t3_cg<T_PC_AP>(pl, A, st));
            T3_TRY(t3_scalar<S_PC_A>(pl, 50, st));
            T3_TRY(t3_vec<TV_PC_UPDATE_PRE>(pl, A, st));
            T3_TRY(t3_cg<T_PC_RESID>(pl, A, st));
            T3_TRY(t3_vec<TV_PC_PRECOND_CHK>(pl, A, st));
            T3_TRY(t3_scalar<S_PC_C>(pl, 50, st));
            T3_TRY(t3_vec<TV_PC_UPDATE_P>(pl, A, st));
            ++it;
            if (it % poll == 0 || it == max_iter) {
                T3_TRY(all_done());
                if (h.done) break;
            }
        }
        T3_TRY(t3_vec<TV_PC_FINISH>(pl, A, st));
    }
    T3_TRY(all_done());
    if (iters) *iters = h.iter;
    if (status) *status = h.done ? h.status : PHAST_CG_MAXITER;
    return PHAST_OK;
}

extern "C" int phast_t3_set_fracture(phast_t3_plan *pl, double Gc, double l0, double at1_source) {
    PHAST_REQUIRE(pl, "phast_t3_set_fracture: null plan");
    PHAST_REQUIRE(Gc > 0.0 && l0 > 0.0 && std::isfinite(Gc) && std::isfinite(l0), "Gc and l0 must be positive");
    PHAST_REQUIRE(std::isfinite(at1_source), "at1_source must be finite");
    T3Mat &M = pl->mat;  // as in phast_t3_plan_create (damage_solver.py:574-585)
    M.Gc_l0 = M.at1 ? 0.75 * Gc * l0 : Gc * l0;
    M.Gc_over_l0 = M.at1 ? 0.0 : Gc / l0;
    M.src = M.at1 ? at1_source : 0.0;
    M.l0 = l0;
    return PHAST_OK;
}

extern "C" int phast_t3_adjoint_solve(phast_t3_plan *pl, const double *H, const double *Gc_e, const double *rhs,
                                      const uint8_t *fixed, const double *fixed_vals, const double *x_init,
                                      double rel_tol, int32_t max_iter, double *x_out, int32_t *iters,
                                      int32_t *status, hipStream_t st) {
    PHAST_REQUIRE(pl && H && x_out, "phast_t3_adjoint_solve: null argument");
    PHAST_REQUIRE(pl->K == 3, "phast_t3_adjoint_solve: the damage adjoint is built for triangle meshes");
    PHAST_REQUIRE(rel_tol >= 0.0 && max_iter >= 0, "rel_tol and max_iter must be non-negative");
    PHAST_REQUIRE(fixed || !fixed_vals, "fixed_vals needs fixed");
    PHAST_REQUIRE(x_out != rhs && x_out != H && x_out != Gc_e, "x_out must not alias rhs, H or Gc_e");
    T3_TRY(t3_ensure_work(pl));
    DamageCG &h = *pl->h_cg;
    std::memset(&h, 0, sizeof(h));
    h.tol = rel_tol;
    h.max_iter = max_iter;
    PHAST_HIP(hipMemcpyAsync(pl->cg, pl->h_cg, sizeof(DamageCG), hipMemcpyHostToDevice, st));

    T3Cg A;
    std::memset(&A, 0, sizeof(A));
    A.H = H;
    A.Gc = Gc_e;
    A.d = x_out;
    A.d_out = x_out;
    A.r = pl->r;
    A.z = pl->z;
    A.p = pl->p;
    A.Ap = pl->Ap;
    A.b = rhs ? rhs : pl->b;
    A.minv = pl->minv;
    A.init = x_init;
    A.vals = fixed_vals;
    A.fixed = fixed;
    A.cg = pl->cg;
    A.part = pl->part;
    A.nblk = pl->nblk;

    // M_inv = 1/(diag + 1e-30) of the forward system, and its b when no right-hand side is given
    T3_TRY(launch_damage<2>(pl, H, Gc_e, nullptr, pl->b, pl->minv, st));
    T3_TRY(t3_vec<TA_SETUP>(pl, A, st));
    T3_TRY(t3_cg<T_PC_INIT>(pl, A, st));  // r = rhs - A x (0 at pins), z = M_eff r, p = z
    T3_TRY(t3_scalar<S_AJ_INIT>(pl, 0, st));
    const int poll = 4;
    int it = 0;
    auto sync_state = [&]() -> int {
        PHAST_HIP(hipMemcpyAsync(pl->h_cg, pl->cg, sizeof(DamageCG), hipMemcpyDeviceToHost, st));
        PHAST_HIP(hipStreamSynchronize(st));
        return PHAST_OK;
    };
    while (it < max_iter) {
        T3_TRY(t3_cg<T_PC_AP>(pl, A, st));
        T3_TRY(t3_scalar<S_AJ_A>(pl, 0, st));
        T3_TRY(t3_vec<TA_UPDATE>(pl, A, st));
        T3_TRY(t3_scalar<S_AJ_C>(pl, 0, st));
        T3_TRY(t3_vec<TA_UPDATE_P>(pl, A, st));
        ++it;
        if (it % poll == 0 || it == max_iter) {
            T3_TRY(sync_state());
            if (h.done) break;
        }
    }
    T3_TRY(t3_vec<TA_FINISH>(pl, A, st));
    T3_TRY(sync_state());
    if (iters) *iters = h.iter;
    if (status) *status = h.done ? h.status : PHAST_CG_MAXITER;
    return PHAST_OK;
}

extern "C" int phast_t3_adjoint_terms(const phast_t3_plan *pl, const double *lam, const double *d_pull,
                                      const double *d_unc, double *gH, double *lKd, double *lMd, double *lM1,
                                      hipStream_t st) {
    PHAST_REQUIRE(pl && lam, "phast_t3_adjoint_terms: null argument");
    PHAST_REQUIRE(pl->K == 3, "phast_t3_adjoint_terms: the damage adjoint is built for triangle meshes");
    PHAST_REQUIRE(!gH || d_pull, "gH needs d_pull");
    PHAST_REQUIRE(!(lKd || lMd) || d_unc, "lKd and lMd need d_unc");
    const int64_t ne = pl->mesh.n_elems;
    hipLaunchKernelGGL(t3_adjoint_terms_kernel, dim3((unsigned)ceil_div(ne, 256)), dim3(256), 0, st, pl->mesh, lam,
                       d_pull, d_unc, gH, lKd, lMd, lM1);
    PHAST_LAUNCHED();
    return PHAST_OK;
}
