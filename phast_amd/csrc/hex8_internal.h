// Internal declarations shared by the voxel kernels (hex8.hip) and the multigrid (mg.hip).
#pragma once

#include "common.h"

namespace phast {

struct Hex8Geo {
    int32_t nx, ny, nz, pad;
    double ihx, ihy, ihz;   // 1/h
    double w;               // Gauss weight * det J = hx hy hz / 8
    double al, be;          // linear interpolation weights (1 +- 1/sqrt3)/2
    double alx, bex, aly, bey, alz, bez;  // al/h, be/h per axis (interpolation + derivative scale)
    double kappa, omk, ch;  // residual stiffness, 1-kappa, 2(1-kappa)
    double K[64];           // AT2 voxel matrix M/ell + ell L, lexicographic corners [a*8+b]
    double Nq[64];          // N[q*8 + a], lexicographic corners, xi-fastest Gauss points
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
    bool owns_material = true;
    double *part = nullptr;  // block partials (8 per block)
    double *seam = nullptr;  // voxel-parallel apply: slab seams (seam_cap doubles)
    int64_t seam_cap = 0;
    int32_t max_blocks = 0;
    int32_t tx = 16, ty = 8, kz = 8;  // launch shape of the apply kernels (phast_hex8_set_tiles)
    // scratch of the solves on this plan (linalg.hip): vectors of 3n+6, the flag state and the
    // pinned readback; made by phast_hex8_solver_reserve or the first solve, freed with the plan
    mutable void *solver_work = nullptr;
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
}  // namespace phast
