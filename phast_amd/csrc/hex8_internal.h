// Internal declarations shared by the voxel kernels (hex8.hip) and the multigrid (mg.hip).
#pragma once

#include "common.h"

namespace phast {

// Most distinct (lam, mu) pairs the voxel-parallel apply reads through a 1-byte phase index.
constexpr int kHex8MaxPhases = 256;

struct Hex8Geo {
    int32_t nx, ny, nz, pad;
    double ihx, ihy, ihz;   // 1/h
    double w;               // Gauss weight * det J = hx hy hz / 8
    double al, be;          // linear interpolation weights (1 +- 1/sqrt3)/2
    double alx, bex, aly, bey, alz, bez;  // al/h, be/h per axis (interpolation + derivative scale)
    double c3;              // Gauss point in centred coordinates, 1/sqrt3
    double kx, ky, kz;      // 1/(4h): derivative of the Walsh-Hadamard modes per axis
    double kcx, kcy, kcz;   // k c3 per axis
    double kappa, omk, ch;  // residual stiffness, 1-kappa, 2(1-kappa)
    double K[64];           // AT2 voxel matrix M/ell + ell L, lexicographic corners [a*8+b]
    double Nq[64];          // N[q*8 + a], lexicographic corners, xi-fastest Gauss points
};

// Wu's PF-CZM law on the voxel cell (physics/pfczm.py; solid3d/operator.py:272-583):
// g = (1 - eta) A/(A + Q) + eta, A = (1 - d)^p, Q = a1 d (1 + a2 d + a2 a3 d^2) with a1 per
// voxel; the damage energy Gc/(pi ell)(2 m.d - d.M d) + Gc ell/pi d.L d + sum_q w H g(d_q).
struct PfczmLaw {
    double a2, a3, p, eta;
    int32_t p_int;          // p == 2: the integer-power forms of the reference
    int32_t pad;
    double c_local, c_grad; // 1/(pi ell), ell/pi
    double concave, floor;  // 2/(pi ell) (times Gc per voxel) and the tangent floor fraction 0.25
    double M[64], L[64];    // consistent mass and stiffness of the voxel, symmetrised [a*8+b]
    double m[8];            // row sums of M
};

}  // namespace phast

struct phast_hex8_plan {
    phast::Hex8Geo geo;
    int64_t n = 0;  // reduced nodes == voxels
    double *bank = nullptr;
    int32_t *bank_index = nullptr;
    double *Gc = nullptr, *lam = nullptr, *mu = nullptr;
    int8_t *driver = nullptr;
    int32_t n_bank = 0;
    bool iso = false;   // every bank entry isotropic: stress from (lam, mu) per voxel
    // isotropic cells with at most kHex8MaxPhases distinct (lam, mu) pairs: the phase of each
    // voxel (1 byte) and the pairs, read by the voxel-parallel apply instead of lam and mu
    uint8_t *phase = nullptr;
    double *ptab = nullptr;
    int32_t n_phase = 0;
    bool owns_material = true;
    double *part = nullptr;  // block partials (8 per block)
    double *seam = nullptr;  // voxel-parallel apply: slab seams (seam_cap doubles)
    int64_t seam_cap = 0;
    int32_t max_blocks = 0;
    int32_t tx = 16, ty = 8, kz = 8;  // launch shape of the apply kernels (phast_hex8_set_tiles)
    // scratch of the solves on this plan (linalg.hip): vectors of 3n+6, the flag state and the
    // pinned readback; made by phast_hex8_solver_reserve or the first solve, freed with the plan
    mutable void *solver_work = nullptr;
    double h[3] = {0.0, 0.0, 0.0};
    double ell = 0.0;
    // PF-CZM damage law (phast_hex8_pfczm_setup): a1 per voxel, the law and the Newton
    // tangent's constants (K = (2 ell/pi) L, ch = 1) for the damage operator kernels
    double *a1 = nullptr;
    phast::PfczmLaw law{};
    phast::Hex8Geo geo_t{};
    // crack closure (phast_hex8_closure_setup / _rows): the group of each voxel (0: adds
    // nothing back, 1: isotropic group, 2: laminate band group) and, for cells with a general
    // stiffness bank, the closure rows of each bank entry (28 doubles: rows v_k (4 x 6),
    // moduli k_k (4))
    uint8_t *cl_on = nullptr;
    double *cl_tab = nullptr;
    bool cl_band = false;  // some voxel is in the band group
};

namespace phast {
// Enqueue y = A(g) x on `st`; g from `g` (E,8), or from nodal `d`, or 1 when both are null.
// With out7 the macro rows and the energy go to out7 instead of y[3n..3n+5].
int hex8_apply(const phast_hex8_plan *pl, const double *x, const double *g, const double *d, double *y,
               double *out7, bool energy, hipStream_t st);
void build_voxel_constants(Hex8Geo &G, double hx, double hy, double hz, double ell);
// One multigrid smoothing step: Ax = A(g) x, then x += omega (b - Ax) / diag on free dofs
// (fixed dofs set to 0), the update fused into the apply's finalize kernel.
int hex8_apply_jacobi(const phast_hex8_plan *pl, double *x, const double *g, const double *d, double *Ax,
                      const double *b, const double *diag, const uint8_t *fixed, double omega, hipStream_t st);
// Grow the seam buffer of the voxel-parallel apply for slabs of kz node layers (set-up time).
int hex8_ensure_seam(phast_hex8_plan *pl, int32_t kz);
// Free the solver scratch of a plan (linalg.hip; called by phast_hex8_plan_destroy).
void release_solver_work(const phast_hex8_plan *pl);
// y = A(g) x with the crack closure at each point's own pattern (mode 1) or at `pat` (mode 2).
int hex8_apply_closure(const phast_hex8_plan *pl, const double *x, const double *g, const uint32_t *pat, int mode,
                       double *y, double *out7, bool energy, hipStream_t st);
// y = scale (Gc K d + ch sum_q w H N N^T d) and its rhs / diagonal with the constants of G.
int hex8_damage_launch(const phast_hex8_plan *pl, const Hex8Geo &G, const double *d, const double *H, double scale,
                       double *y, hipStream_t st);
int hex8_damage_rd_launch(const phast_hex8_plan *pl, const Hex8Geo &G, const double *H, double scale, double *b,
                          double *diag, hipStream_t st);
// voxel N, M, L (symmetrised), K = M/ell + ell L and the interpolation constants into G;
// M and L also into Mo, Lo when given
void voxel_constants(Hex8Geo &G, double hx, double hy, double hz, double ell, double *Mo, double *Lo);
}  // namespace phast
