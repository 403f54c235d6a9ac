/*
 * phast_mi.h -- C ABI of the MI355X (gfx950) matrix-free phase-field hot path.
 *
 * Conventions
 *   - Every entry point returns int: PHAST_OK (0) or a negative PHAST_ERR_* code;
 *     phast_last_error() returns a thread-local message for the last failure.
 *   - Array arguments are DEVICE pointers (float64 unless stated) that the caller owns;
 *     outputs are written into caller-allocated buffers.  Scalars and descriptor
 *     structs are host values.  Each call enqueues work on `stream` and returns
 *     without synchronising unless the comment says otherwise (solvers synchronise to
 *     read their convergence flags).
 *   - Plans are opaque handles holding immutable per-problem geometry constants,
 *     launch tables and solver work buffers.  One plan must not be used from two
 *     streams at the same time.
 *
 * The reference (CEMS-Lab/PhAST) is pure Python/PyTorch; it has no FFI.  Each entry
 * point names the reference method whose arithmetic it replaces (paths relative to
 * the reference's src/phast).  The Python binding is phast_amd/_lib.py (ctypes).
 */
#ifndef PHAST_MI_H
#define PHAST_MI_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PHAST_OK 0
#define PHAST_ERR_ARG (-1)               /* shape / value / pointer violation -> ValueError   */
#define PHAST_ERR_HIP (-2)               /* HIP runtime failure -> RuntimeError               */
#define PHAST_ERR_NONPOSITIVE (-3)       /* p.Ap <= 0 -> PCGBreakdownError                    */
#define PHAST_ERR_NONFINITE (-4)         /* non-finite value in a solve -> PCGBreakdownError  */
#define PHAST_ERR_NOT_CONVERGED (-5)     /* iteration cap reached (info carries the state)    */
#define PHAST_ERR_UNSUPPORTED (-6)       /* configuration outside the built path              */

/* solver termination codes written into the per-instance status arrays */
#define PHAST_CG_RUNNING 0
#define PHAST_CG_CONVERGED 1
#define PHAST_CG_MAXITER 2
#define PHAST_CG_NONPOSITIVE 3
#define PHAST_CG_DIVERGED 4
#define PHAST_CG_NONFINITE 5

const char *phast_version(void);
const char *phast_last_error(void);

/* =====================================================================================
 * 2D native-Q4 structured plates (one or many instances packed in one plan)
 * ===================================================================================== */

#define PHAST_SPLIT_AMOR 0
#define PHAST_SPLIT_ISOTROPIC 1
#define PHAST_SPLIT_SPECTRAL 2 /* T3 only (core/fem_operators.py:335-390) */
#define PHAST_PF_AT2 0
#define PHAST_PF_AT1 1         /* T3 only (solvers/damage_solver.py:1321-1329) */

/* Host descriptor of one structured Q4 plate: (nx+1)(ny+1) nodes numbered
 * j*(nx+1)+i and nx*ny cells numbered j*nx+i with corners (i,j),(i+1,j),(i+1,j+1),(i,j+1)
 * (core/mesh_generator.py:1735-1752).  Geometry constants (gradN, N, wdetJ of the
 * congruent cell, core/quad_elements.py:315-349) are supplied by the caller so that the
 * library uses the same quadrature data as the host mesh precompute. */
typedef struct phast_q4_desc {
    int32_t nx, ny;          /* cells along x and y                                       */
    double gradN[4][4][2];   /* [q][a][x|y] physical shape-function gradients              */
    double N[4][4];          /* [q][a] shape values at the 2x2 Gauss points                */
    double wdetJ[4];         /* [q] weight * det J                                         */
    double lam, mu, kappa;   /* Lame lambda (split-dependent), shear modulus, 3D bulk modulus */
    double nu;               /* Poisson ratio (plane-stress eps_zz)                        */
    int32_t plane_stress;    /* 0 plane strain, 1 plane stress                             */
    int32_t split;           /* PHAST_SPLIT_*                                              */
    double eta_residual;     /* g(d) = (1-d)^2 (1-eta) + eta  (physics/material.py:529)     */
    double Gc, l0;           /* AT2 damage constants                                       */
    double dt;               /* explicit step size                                         */
    double hx, hy;           /* cell edge lengths (cells are congruent axis-aligned rectangles) */
    int32_t stress_split;    /* PHAST_SPLIT_* of the stress: ISOTROPIC under stress_degradation
                                'full' (physics/material.py:367-376), else the same as split  */
} phast_q4_desc;

typedef struct phast_q4_plan phast_q4_plan;

/* Work decomposition: strips of tile_x threads (64, 128 or 256) covering tile_x - 1 node
 * columns and tile_y node rows each; (0,0) = default (128, 8). */
int phast_q4_plan_create(const phast_q4_desc *desc, int32_t n_inst, int32_t tile_x, int32_t tile_y,
                         phast_q4_plan **out);
int phast_q4_plan_destroy(phast_q4_plan *plan);
/* total packed nodes / cells of the plan, and the node/cell offset of instance k */
int phast_q4_plan_sizes(const phast_q4_plan *plan, int64_t *n_nodes, int64_t *n_cells);
int phast_q4_plan_offsets(const phast_q4_plan *plan, int64_t *node_off, int64_t *cell_off);

/* Scalar lumped mass M_a = sum_q N_a(q) wdetJ(q) gathered per node (core/mesh.py:348-352). */
int phast_q4_lumped_mass(const phast_q4_plan *plan, double *M, hipStream_t stream);

/* f_int (N,2) = internal force of u (N,2) at damage d (N,)
 * (core/fem_operators.py:598-612 internal_force, 244-254, 614-653, 301-333, 681-699). */
int phast_q4_internal_force(const phast_q4_plan *plan, const double *u, const double *d, double *f,
                            hipStream_t stream);

/* f (N,2) = beta f_int(v, d): the Kelvin-Voigt damping force of ExplicitDynamics with
 * damping_ratio_max > 0, beta = zeta dt_cfl (mechanics_solver.py:110-148, 192-195). */
int phast_q4_damping_force(const phast_q4_plan *plan, const double *v, const double *d, double beta, double *f,
                           hipStream_t stream);

/* Gauss-point strain, three (E,4) arrays packed as eps[3][E][4] (fem_operators.py:244-254). */
int phast_q4_strain(const phast_q4_plan *plan, const double *u, double *eps, hipStream_t stream);

/* psi+ (E,4) of u (fem_operators.py:951-1006 compute_psi_plus / 1173 compute_driving_force);
 * when H_old and H_new are non-null also H_new = max(H_old, psi+)
 * (solvers/staggered_solver.py:804-845, physics/h_update.py:47-49). */
int phast_q4_driving_force(const phast_q4_plan *plan, const double *u, double *psi,
                           const double *H_old, double *H_new, hipStream_t stream);

/* H_new = max(H_old, psi) elementwise over n entries (physics/h_update.py:47-49). */
int phast_history_max(int64_t n, const double *H_old, const double *psi, double *H_new,
                      hipStream_t stream);

/* Nodal projection of a Gauss-point field (E,4): sum_e N w f / (M + 1e-30) (core/mesh.py:407-428). */
int phast_q4_elem_to_node(const phast_q4_plan *plan, const double *field_q, const double *M,
                          double *out, hipStream_t stream);

/* One central-difference step with lumped mass, fused with psi+ and the history update
 * (solvers/mechanics_solver.py:153-239 ExplicitDynamics.step; staggered_solver.py:1115-1146):
 *   u1 = u + dt v + dt^2/2 a, u1[bc] = bc_vals
 *   a1 = Minv (f_ext - f_int(u1, d)), v1 = v + dt/2 (a + a1), a1[bc] = v1[bc] = 0
 *   psi = psi+(u1), H_new = max(H_old, psi)
 * Minv is (N,) = 1/(rho M + 1e-30).  f_ext, bc_mask (uint8 (N,2)), bc_vals, psi may be null.
 * f_damp (optional, (N,2)) = beta f_int(v, d), the Kelvin-Voigt damping force of
 * phast_q4_damping_force, subtracted as f_net = (f_ext - f_int) - f_damp
 * (mechanics_solver.py:165-201, damping_ratio_max > 0).
 * Outputs must not alias inputs (neighbouring tiles read the old state). */
int phast_q4_explicit_step(const phast_q4_plan *plan, const double *u, const double *v, const double *a,
                           const double *d, const double *f_ext, const double *f_damp, const double *Minv,
                           const uint8_t *bc_mask, const double *bc_vals,
                           const double *H_old, double *u1, double *v1, double *a1, double *H_new,
                           double *psi, hipStream_t stream);

/* Damage operator pieces (solvers/damage_solver.py:2270-2301 _prepare_cg_q4, 1797-1838 _Ax_q4).
 * Gc_cell (E,) is an optional per-cell Gc field (null: the descriptor's scalar Gc). */
int phast_q4_damage_rhs_diag(const phast_q4_plan *plan, const double *H, const double *Gc_cell,
                             double *b, double *diag, hipStream_t stream);
int phast_q4_damage_apply(const phast_q4_plan *plan, const double *H, const double *Gc_cell,
                          const double *x, double *y, hipStream_t stream);

/* Box-constrained damage solve on [d_prev, 1] for every instance of the plan.
 * method 0: projected PCG with step limiting and active set (damage_solver.py:3527-3775);
 * method 1: unconstrained PCG then clamp (damage_solver.py:3377-3521).
 * d_init (optional) is the initial iterate (defaults to d_prev); fixed (uint8, optional)
 * pins nodes to fixed_vals (pf_dirichlet).  Outputs: d_out (N,), active_out (uint8 (N,),
 * optional, projected method), and host arrays iters[n_inst], status[n_inst]
 * (PHAST_CG_*).  Synchronises `stream` every `check_every` iterations. */
int phast_q4_damage_solve(phast_q4_plan *plan, int32_t method, const double *H, const double *Gc_cell,
                          const double *d_prev, const double *d_init, const uint8_t *fixed,
                          const double *fixed_vals, double tol, int32_t max_iter,
                          double *d_out, uint8_t *active_out, int32_t *iters, int32_t *status,
                          hipStream_t stream);

/* No counterpart in the reference.  phast_q4_damage_solve for callers that cannot allocate
 * device memory: code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  Same arguments, kernels and results (bit for bit); the CG scratch (vectors,
 * strip partials, active mask, CG states) is a file-scope static device array of 36 MB that
 * each GPU the library runs on holds until the process ends, so no call allocates.  Limits:
 * plans of at most 524288 nodes, 65536 strips and 64 instances (PHAST_ERR_ARG beyond).
 * Every call on one GPU shares that GPU's scratch: calls that overlap on one GPU race. */
int phast_q4_damage_solve_noalloc(phast_q4_plan *plan, int32_t method, const double *H, const double *Gc_cell,
                                  const double *d_prev, const double *d_init, const uint8_t *fixed,
                                  const double *fixed_vals, double tol, int32_t max_iter, double *d_out,
                                  uint8_t *active_out, int32_t *iters, int32_t *status, hipStream_t stream);

/* Persistent batched explicit step (no single-call counterpart in the reference: it fuses
 * the loop body of solvers/staggered_solver.py:1115-1165 with damage_every = 1, i.e.
 * ExplicitDynamics.step (mechanics_solver.py:204-239), the history max
 * (staggered_solver.py:804-845) and PhaseFieldDamageSolver.solve (damage_solver.py:
 * 3377-3775), for every instance of a plan in two launches: the fused explicit-step
 * kernel, then ONE persistent kernel running every instance's whole damage CG).
 * phast_q4_batch_prepare sizes the co-resident grid (occupancy query; blocks shared among
 * instances in proportion to their strips) and must be called once, outside graph
 * capture; it returns the grid of each damage method in grid_out[2] (optional).
 * phast_q4_batch_step launches nothing but those two kernels: it allocates nothing, never
 * synchronises and may be captured into a hipGraph.  Inputs u, v, a, d (= d_prev), H_old;
 * outputs u1, v1, a1, H_new, psi (optional), d_out; iters/status are DEVICE int32 arrays
 * [n_inst].  Outputs must not alias inputs.  Results are bitwise those of
 * phast_q4_explicit_step followed by phast_q4_damage_solve(H_new, d_prev = d).
 * phast_q4_batch_error reads (and optionally clears) the plan's barrier-timeout flag
 * (non-zero: a wait among the blocks of an instance ran out and results are invalid). */
int phast_q4_batch_prepare(phast_q4_plan *plan, int32_t *grid_out);
int phast_q4_batch_step(phast_q4_plan *plan, int32_t method, const double *u, const double *v, const double *a,
                        const double *d, const double *f_ext, const double *M_inv, const uint8_t *bc_mask,
                        const double *bc_vals, const double *H_old, const double *Gc_cell, const uint8_t *fixed,
                        const double *fixed_vals, double tol, int32_t max_iter, double *u1, double *v1, double *a1,
                        double *H_new, double *psi, double *d_out, int32_t *iters, int32_t *status,
                        hipStream_t stream);
int phast_q4_batch_error(phast_q4_plan *plan, int32_t *err, int32_t clear);

/* n_steps persistent batched steps in ONE launch (the same loop body as phast_q4_batch_step,
 * staggered_solver.py:1115-1165 with damage_every = 1, repeated).  Each instance advances on
 * its own blocks, joined only by its own barriers, so instances whose damage solves finish
 * early run ahead instead of waiting for the slowest one at every step.  The state is kept
 * in two buffer sets that the steps alternate between: step s reads (u, v, a, d, H) and
 * writes (u1, v1, a1, d1, H1) when s is even, the other way round when s is odd, so after an
 * even n_steps the result is in (u, v, a, d, H) and after an odd one in (u1, ...).  psi
 * holds the last step's psi+.  iters/status are DEVICE int32 arrays [n_steps][n_inst].
 * Results are bitwise those of n_steps calls of phast_q4_batch_step.  Needs
 * phast_q4_batch_prepare; allocates nothing, never synchronises, may be captured.  Refuses
 * an n_steps x max_iter that would overflow the per-instance barrier counters. */
int phast_q4_batch_run(phast_q4_plan *plan, int32_t method, int32_t n_steps, double *u, double *v, double *a,
                       double *d, const double *f_ext, const double *M_inv, const uint8_t *bc_mask,
                       const double *bc_vals, double *H, const double *Gc_cell, const uint8_t *fixed,
                       const double *fixed_vals, double tol, int32_t max_iter, double *u1, double *v1, double *a1,
                       double *H1, double *psi, double *d1, int32_t *iters, int32_t *status, hipStream_t stream);

/* =====================================================================================
 * 2D linear triangles (T3) on an unstructured mesh
 * ===================================================================================== */

typedef struct phast_t3_plan phast_t3_plan;

/* Mesh upload and host precompute (core/mesh.py:226-300 _precompute_t3): nodes host (N,2),
 * elements host (E,3) int32.  Builds the per-node incidence list in e*3+a order, the order in
 * which scatter_add over elements.flatten() visits a node (fem_operators.py:661-679).
 * Material constants as physics/material.py:379-388 (lam, mu, kappa) and :529 (g(d));
 * split = PHAST_SPLIT_* of psi+ and the history; stress_split that of the stress
 * (PHAST_SPLIT_ISOTROPIC under stress_degradation 'full', material.py:367-376).  pf_model
 * PHAST_PF_AT2 or PHAST_PF_AT1 (AT1: 3/4 Gc l0 on the Laplacian, no Gc/l0 in the reaction,
 * the source at1_source in the right-hand side 2H - S_H; damage_solver.py:1321-1336). */
int phast_t3_plan_create(int64_t n_nodes, int64_t n_elems, const double *nodes, const int32_t *elements,
                         double E, double nu, double Gc, double l0, double eta_residual, int32_t plane_stress,
                         int32_t split, int32_t stress_split, int32_t pf_model, double at1_source,
                         phast_t3_plan **out);
int phast_t3_plan_destroy(phast_t3_plan *plan);
/* Host copies of M_scalar (N,) = scatter(area/3) (core/mesh.py:287-289), h_min
 * (mesh.py:276-283) and the total area.  Any pointer may be null. */
int phast_t3_mesh_info(const phast_t3_plan *plan, double *M_scalar, double *h_min, double *total_area);

/* Organisation of the T3 passes: path 0 = element patches (each triangle evaluated once per
 * patch in LDS; the default when the mesh's patch lists fit the LDS budget), 1 = every
 * incidence re-evaluates its triangle.  Both give the same bits.  active_path / max_list
 * (optional) return the organisation in use and the longest patch triangle list. */
int phast_t3_set_path(phast_t3_plan *plan, int32_t path, int32_t *active_path, int32_t *max_list);
/* f (N,2) = internal force of u (N,2) at nodal damage d (N,): centroid strain
 * (fem_operators.py:223-242), element-mean degradation (:626-630), the split
 * (:277-290 isotropic, :301-333 amor, :335-390 spectral) and the scatter (:661-679). */
int phast_t3_internal_force(const phast_t3_plan *plan, const double *u, const double *d, double *f,
                            hipStream_t stream);
/* internal force plus psi+ (E,) of u (fem_operators.py:951-1030) and H_new = max(H_old, psi+)
 * (staggered_solver.py:804-845).  psi may be null. */
int phast_t3_force_history(const phast_t3_plan *plan, const double *u, const double *d, const double *H_old,
                           double *f, double *H_new, double *psi, hipStream_t stream);
/* One central-difference step with lumped mass (solvers/mechanics_solver.py:153-239), with
 * psi+(u1) and the history max, in two launches (predictor; force + corrector):
 *   u1 = u + dt v + dt^2/2 a, u1[bc] = bc_vals;  a1 = Minv (f_ext - f_int(u1, d)),
 *   v1 = v + dt/2 (a + a1), a1[bc] = v1[bc] = 0;  H_new = max(H_old, psi+(u1)).
 * Minv (N,) = 1/(rho M + 1e-30); f_ext, bc_mask (uint8 (N,2)) + bc_vals, psi may be null;
 * f_damp (optional) = phast_t3_damping_force(v, d, beta), subtracted as
 * f_net = (f_ext - f_int) - f_damp (mechanics_solver.py:165-201).  Outputs must not alias inputs. */
int phast_t3_damping_force(const phast_t3_plan *plan, const double *v, const double *d, double beta, double *f,
                           hipStream_t stream);
int phast_t3_explicit_step(const phast_t3_plan *plan, double dt, const double *u, const double *v, const double *a,
                           const double *d, const double *f_ext, const double *f_damp, const double *Minv,
                           const uint8_t *bc_mask,
                           const double *bc_vals, const double *H_old, double *u1, double *v1, double *a1,
                           double *H_new, double *psi, hipStream_t stream);
/* Element fields of u (N,2) at damage d (N,), one thread per triangle, each output optional:
 * eps (3,E) = centroid strain exx, eyy, gxy (fem_operators.py:223-242); sig (3,E) = the split's
 * degraded stress sxx, syy, sxy (:614-653; needs d); psi (E,) = psi+ (:951-1030). */
int phast_t3_element_fields(const phast_t3_plan *plan, const double *u, const double *d, double *eps, double *sig,
                            double *psi, hipStream_t stream);
/* Area-weighted projection of an element field f (E,) to the nodes (core/mesh.py:407-433):
 * out_n = sum_e f_e A_e / (3 M_n + 1e-30), M = M_scalar (device (N,)). */
int phast_t3_elem_to_node(const phast_t3_plan *plan, const double *field_e, const double *M, double *out,
                          hipStream_t stream);
/* AT2 damage operator of element history H (E,) (solvers/damage_solver.py:1715-1755 _Ax):
 *   y = Gc l0 sum_e A (grad phi_a . grad d_e) + sum_e rc_e (d_a + sum_b d_b),
 *   rc = (2H + Gc/l0) A / 12  (damage_solver.py:2168-2177);  x and y must differ.
 * Gc_e (E,) optional per-element Gc (the _Gc_l0_e / _Gc_over_l0_e branch,
 * damage_solver.py:1412-1433): then Gc_e l0 weights each element's Laplacian term. */
int phast_t3_damage_apply(const phast_t3_plan *plan, const double *H, const double *Gc_e, const double *x, double *y,
                          hipStream_t stream);
/* b = sum_e 2H A/3 (damage_solver.py:2249-2258) and the Jacobi diagonal
 * sum_e (Gc l0 A |grad phi_a|^2 + 2 rc) (damage_solver.py:2112-2119); either may be null. */
int phast_t3_damage_rhs_diag(const phast_t3_plan *plan, const double *H, const double *Gc_e, double *b,
                             double *diag, hipStream_t stream);
/* AT2 damage solve with the Jacobi-preconditioned CG of PhaseFieldDamageSolver.solve
 * (damage_solver.py:2493-2684): method 0 = projected CG with the bounds d_prev <= d <= 1
 * (:3527-3775), 1 = CG then clamp (:3377-3521); same arguments, state machine and
 * scalar steps as phast_q4_damage_solve.  fixed (uint8 (N,)) + fixed_vals pin nodes
 * (pf_dirichlet); d_init optional initial guess; active_out (uint8 (N,)) optional final
 * active set of method 0; iters/status are host int32.  Allocates its workspace on the
 * plan's first solve. */
int phast_t3_damage_solve(phast_t3_plan *plan, int32_t method, const double *H, const double *Gc_e,
                          const double *d_prev, const double *d_init, const uint8_t *fixed, const double *fixed_vals,
                          double tol, int32_t max_iter, double *d_out, uint8_t *active_out, int32_t *iters,
                          int32_t *status, hipStream_t stream);
/* Replace the plan's fracture constants (Gc, l0, and the AT1 source S_H) for the launches
 * enqueued after the call: the temporary material override that _AdjointDamageSolveScalar
 * installs around its forward and adjoint solves (damage_solver.py:574-585, 686-696). */
int phast_t3_set_fracture(phast_t3_plan *plan, double Gc, double l0, double at1_source);
/* Jacobi-PCG of the adjoint recomputation, _pcg_solve_adjoint_system
 * (damage_solver.py:423-478): solves A x = rhs with the forward operator at (H, Gc_e) and
 * M_inv = 1/(diag + 1e-30); rhs null = the forward right-hand side b (the unconstrained
 * damage d_unc).  fixed (uint8 (N,)) pins are eliminated and held at fixed_vals (0 when
 * fixed_vals is null); x_init optional (else 0).  Stops when r.r < rel_tol max(rhs.rhs,
 * 1e-30), on p.Ap <= 0, or after max_iter iterations.  iters/status are host int32. */
int phast_t3_adjoint_solve(phast_t3_plan *plan, const double *H, const double *Gc_e, const double *rhs,
                           const uint8_t *fixed, const double *fixed_vals, const double *x_init, double rel_tol,
                           int32_t max_iter, double *x_out, int32_t *iters, int32_t *status, hipStream_t stream);
/* Per-triangle adjoint terms (E,), each output optional: the history gradient
 * gH = sum_a lam_a 2A (1/3 - (dp_a + sum dp)/12) with dp = d_pull (damage_solver.py:800-810,
 * 1140-1149), and the quadratic forms of _adjoint_quadratic_forms_perelem (:284-331)
 * lKd = A grad lam . grad d_unc, lMd = A/12 (sum lam sum d_unc + lam . d_unc), lM1 = A/3 sum lam. */
int phast_t3_adjoint_terms(const phast_t3_plan *plan, const double *lam, const double *d_pull, const double *d_unc,
                           double *gH, double *lKd, double *lMd, double *lM1, hipStream_t stream);

/* ---- general quadrilateral meshes on the same plan type ------------------------------
 * The reference's Q4 branch on any quadrilateral mesh (core/mesh.py:305-352 _precompute_q4,
 * core/quad_elements.py:315-349 q4_quadrature_geometry), where the structured-plate plans
 * above need congruent axis-aligned cells.  phast_q4u_plan_create: nodes host (N,2),
 * elements host (E,4) int32 in counter-clockwise corner order; every Gauss-point Jacobian
 * must be positive (else PHAST_ERR_ARG, as the reference raises); split / stress_split
 * PHAST_SPLIT_AMOR or _ISOTROPIC; AT2.  The incidence list is in e*4+a order (the order of
 * scatter_add over elements.flatten()).  The plan then works with phast_t3_plan_destroy,
 * _mesh_info (M_scalar = scatter of sum_q N_a wdetJ, h_min = the shortest cell edge, the
 * total area), _internal_force, _damping_force, _force_history, _explicit_step,
 * _element_fields (eps (3,E,4) and psi (E,4) only), _elem_to_node, _damage_apply,
 * _damage_rhs_diag and _damage_solve, with Gauss-point arrays: H, H_old, H_new, psi and the
 * projected field are (E,4) (q = xi fastest: (-a,-a), (a,-a), (-a,a), (a,a)), g is
 * evaluated at the Gauss points from N d (fem_operators.py:626-630), the damage operator is
 * _Ax_q4 / _prepare_cg_q4 (damage_solver.py:1797-1838, 2270-2301).  Each pass is an element
 * pass (one thread per quadrilateral; geometry recomputed from the corner coordinates) that
 * writes two values per corner to plan-owned memory, then a node pass that sums them; passes
 * on one plan must therefore not overlap.  No adjoint, no patches. */
int phast_q4u_plan_create(int64_t n_nodes, int64_t n_elems, const double *nodes, const int32_t *elements, double E,
                          double nu, double Gc, double l0, double eta_residual, int32_t plane_stress, int32_t split,
                          int32_t stress_split, phast_t3_plan **out);

/* =====================================================================================
 * 3D periodic hex8 voxel cells
 * ===================================================================================== */

typedef struct phast_hex8_plan phast_hex8_plan;

/* Triply periodic nx*ny*nz voxel cell of edge lengths (hx,hy,hz)
 * (solid3d/periodic.py:61-108, 193-221).  bank: host (n_bank,6,6) engineering-Voigt
 * stiffnesses; bank_index: host (E,) int32 voxel -> bank entry; Gc, lam, mu: host (E,);
 * driver: host (E,) int8 (0 none, 1 full, 2 spectral; solid3d/operator.py:1415-1444);
 * kappa: residual stiffness; ell: length scale (solid3d/operator.py:229-251).
 * nx, ny, nz >= 2; nz <= 65535; 24 nx ny < 2^32 (a node layer is addressed with 32-bit
 * offsets).  When every bank entry is isotropic and (lam, mu) take at most 256 distinct
 * pairs (bitwise), the plan also keeps a 1-byte phase per voxel and the pairs, which the
 * voxel-parallel apply reads instead of lam and mu (same doubles, same results). */
int phast_hex8_plan_create(int32_t nx, int32_t ny, int32_t nz, double hx, double hy, double hz,
                           const double *bank, int32_t n_bank, const int32_t *bank_index,
                           const double *Gc, const double *lam, const double *mu, const int8_t *driver,
                           double kappa, double ell, phast_hex8_plan **out);
int phast_hex8_plan_destroy(phast_hex8_plan *plan);
/* Launch shape of the element apply.  (16,16), (32,8), (16,14), (32,14), (32,6), (64,6):
 * voxel-parallel kernel, one thread per voxel of a (tile_x+1) x (tile_y+1) voxel tile
 * (blocks of whole waves: 320, 320, 256, 512, 256, 512 threads), walking kz node layers.
 * The tiles of more than 256 threads run the isotropic apply without energy sums or crack
 * closure; the other variants run on the 16x14 tile with the same kz.
 * (16,8) and (8,8): node-column marching kernel
 * walking kz node layers.  Default (16,8,8). */
int phast_hex8_set_tiles(phast_hex8_plan *plan, int32_t tile_x, int32_t tile_y, int32_t kz);

/* g (E,8) = (1-kappa)(1-d_q)^2 + kappa, d_q = N d_e (solid3d/operator.py:939-946). */
int phast_hex8_degradation(const phast_hex8_plan *plan, const double *d, double *g, hipStream_t stream);

/* y (3n+6) = A(g) x for x = [u~ (3n), E (6)] (solid3d/operator.py:1108-1194 _stress_pass /
 * mech_action).  g (E,8) may be null (undamaged).  The macro-strain rows r_E use a
 * deterministic two-pass reduction. */
int phast_hex8_mech_action(const phast_hex8_plan *plan, const double *x, const double *g, double *y,
                           hipStream_t stream);
/* Same action with g formed in-kernel from nodal damage d (n,) (fused degradation). */
int phast_hex8_mech_action_d(const phast_hex8_plan *plan, const double *x, const double *d, double *y,
                             hipStream_t stream);
/* Jacobi diagonal (3n+6) of A(g) (solid3d/operator.py:1196-1254). */
int phast_hex8_mech_diagonal(const phast_hex8_plan *plan, const double *g, double *diag,
                             hipStream_t stream);
/* r_E (6) and elastic energy (1) into out7 (device, 7 doubles) (operator.py:1303-1340). */
int phast_hex8_stress_sums(const phast_hex8_plan *plan, const double *x, const double *g, double *out7,
                           hipStream_t stream);
/* Y (E,8) undamaged driving energy by driver code (operator.py:1415-1444; splits.py:68-161,
 * closed-form eigenvalues of splits.py:93-125). */
int phast_hex8_driving_force(const phast_hex8_plan *plan, const double *x, double *Y, hipStream_t stream);

/* AT2 damage operator, multiplied by `scale` (operator.py:1450-1509). H may be null. */
int phast_hex8_damage_action(const phast_hex8_plan *plan, const double *d, const double *H, double scale,
                             double *y, hipStream_t stream);
int phast_hex8_damage_rhs(const phast_hex8_plan *plan, const double *H, double scale, double *b,
                          hipStream_t stream);
int phast_hex8_damage_diagonal(const phast_hex8_plan *plan, const double *H, double scale, double *diag,
                               hipStream_t stream);

/* Device-resident solves ---------------------------------------------------------------
 * solve_constrained(mech_action(., g), diag, fixed_mask, fixed_values, rhs, x0, rtol, atol,
 * maxiter, check_every) of solid3d/linalg.py:223-287 with the Jacobi PCG of linalg.py:99-220
 * (same flag semantics: alpha/beta masked on device, flags read every check_every).
 * g (E,8) or d (n,) (fused degradation) or neither (undamaged); fixed uint8 (n_mech,);
 * rhs and x0 may be null.  info[5] = iterations, converged, rhs_norm, residual_norm,
 * relative_residual.  Breakdowns return PHAST_ERR_NONPOSITIVE / PHAST_ERR_NONFINITE. */
int phast_hex8_mech_solve(const phast_hex8_plan *plan, const double *g, const double *d, const uint8_t *fixed,
                          const double *fixed_values, const double *rhs, const double *x0, const double *diag,
                          double rtol, double atol, int32_t maxiter, int32_t check_every, int64_t n_mech,
                          double *x_out, double *info, hipStream_t stream);

/* solve_h8_scalar_dirichlet_pcg (solvers/box_kkt_h8.py:350-432) of the damage system
 * scale * A(H) (operator.py:1496-1509); b, diag are the scaled rhs and diagonal; fixed
 * uint8 (n,), values (n,), initial (n,) or null.  info[5] = iterations, converged,
 * initial_residual_norm, residual_norm, relative_residual. */
int phast_hex8_damage_dirichlet_pcg(const phast_hex8_plan *plan, const double *H, double scale, const double *b,
                                    const double *diag, const uint8_t *fixed, const double *values,
                                    const double *initial, double tol, int32_t max_iter, int64_t n, double *out,
                                    double *info, hipStream_t stream);

/* One primal active-set bookkeeping pass of solve_h8_scalar_box_kkt (box_kkt_h8.py:672-733):
 * grad = scale A(H) cand - b; counts[8] = added lower/upper, released lower/upper, free,
 * active lower, active upper (of the masks as passed), free after the update; with
 * apply != 0 the masks are updated in place. */
int phast_hex8_box_update(const phast_hex8_plan *plan, const double *H, double scale, const double *b,
                          const double *cand, const double *lower, const double *upper, uint8_t *lower_active,
                          uint8_t *upper_active, const uint8_t *prescribed, double kkt_tol, int32_t apply, int64_t n,
                          double *grad, int64_t *counts, hipStream_t stream);
/* The KKT audit of a box-constrained solution (box_kkt_h8.py:459-558) in one pass: out8
 * (host) = max |grad| on free entries, max |projected gradient|, max complementarity
 * |mult (bound gap)|, counts of lower / upper sign violations, of entries below lower and
 * above upper, and of bounds with lower > upper.  grad = A sol - b (device). */
int phast_hex8_kkt_stats(const phast_hex8_plan *plan, const double *grad, const double *sol, const double *lower,
                         const double *upper, const uint8_t *lower_active, const uint8_t *upper_active,
                         const uint8_t *prescribed, double kkt_tol, int64_t n, double *out8, hipStream_t stream);

/* Geometric multigrid on the periodic voxel hierarchy (no reference counterpart; the
 * reference's 3D solves are Jacobi-PCG only, solid3d/model.py:782).  Coarse levels halve
 * every dimension while it stays even and >= min_coarse (at most max_levels levels);
 * coarse moduli are voxel averages of g*lam, g*mu; damped-Jacobi smoothing (nu sweeps
 * pre and post, weight omega), coarse_sweeps sweeps on the coarsest level; trilinear P,
 * R = P^T; every level pins the node above `pinned` and the macro components flagged in
 * macro_fixed6 (host, 6 bytes).  The fine plan must be isotropic. */
typedef struct phast_hex8_mg phast_hex8_mg;
int phast_hex8_mg_create(const phast_hex8_plan *fine, int64_t pinned, const uint8_t *macro_fixed6,
                         int32_t max_levels, int32_t min_coarse, int32_t nu, int32_t coarse_sweeps, double omega,
                         phast_hex8_mg **out);
int phast_hex8_mg_destroy(phast_hex8_mg *mg);
/* coarse moduli and Jacobi diagonals for the fine degradation g (E,8) or nodal d (n,) */
int phast_hex8_mg_update(phast_hex8_mg *mg, const double *g, const double *d, hipStream_t stream);
/* z = V(r) on the fine level (3n+6 vectors, zero on the fixed set) */
int phast_hex8_mg_apply(phast_hex8_mg *mg, const double *r, double *z, hipStream_t stream);
int phast_hex8_mg_levels(const phast_hex8_mg *mg, int32_t *levels, int32_t *dims);
/* per-level lambda_max(D^-1 A) estimates (12 power iterations at update) and the Jacobi
 * weights omega_l = omega / (1.05 lambda_max) in use */
int phast_hex8_mg_weights(const phast_hex8_mg *mg, double *lam_max, double *omega);
/* phast_hex8_mech_solve preconditioned by the V-cycle (same flag-masked PCG). */
int phast_hex8_mech_solve_mg(const phast_hex8_plan *plan, phast_hex8_mg *mg, const double *g, const double *d,
                             const uint8_t *fixed, const double *fixed_values, const double *rhs, const double *x0,
                             const double *diag, double rtol, double atol, int32_t maxiter, int32_t check_every,
                             int64_t n_mech, double *x_out, double *info, hipStream_t stream);

/* Solver scratch of a plan (3n+6 doubles per vector: the Krylov vectors, the constrained-solve
 * vectors and the flag state).  The first solve on a plan makes it unless reserve was called;
 * it is freed with the plan.  Solves on different plans share nothing, so they may run on
 * different streams at once; two solves on one plan must not overlap (as in the reference,
 * where an operator owns its tensors). */
int phast_hex8_solver_reserve(const phast_hex8_plan *plan);
int phast_hex8_solver_release(const phast_hex8_plan *plan);

/* linalg.pcg (solid3d/linalg.py:99-220) for an operator evaluated by the caller: the Jacobi
 * PCG with the flag semantics above, whose apply writes the vector into in_buf (n doubles,
 * device), calls apply(ctx) -- which must enqueue out_buf = A(in_buf) on `stream` and return
 * 0 -- and reads out_buf; entries flagged in `fixed` (uint8 (n,), or null) are zeroed after
 * each apply.  b, diag (n,), x0 (n,) or null; x (n,) out; info[5] as phast_hex8_mech_solve.
 * Scratch is allocated for the call and freed before it returns. */
typedef int (*phast_apply_callback)(void *ctx);
int phast_pcg_callback(phast_apply_callback apply, void *ctx, int64_t n, double *in_buf, const double *out_buf,
                       const uint8_t *fixed, const double *b, const double *diag, const double *x0, double *x,
                       double rtol, double atol, int32_t maxiter, int32_t check_every, double *info,
                       hipStream_t stream);


/* ---- adjoint pieces of the converged staggered voxel state ----------------------------
 * No counterpart in the reference (it has no voxel adjoint); each is the transposed
 * derivative of a reference residual, used by a reverse sweep over load steps.
 * Gauss-point arrays are (E,8) (xi fastest), stress-like arrays (E,8,6) engineering Voigt.
 * phast_hex8_set_gc: replace the per-voxel Gc (device (E,)) used by the damage operator.
 * adj_coupling:  out_qp = w g'(d_q) eps(l)_q . C eps(x)_q   ((dF/dd)^T l before the
 *                node sum; F = A(g(d)) x, solid3d/operator.py:939-946, 1108-1194)
 * adj_driving:   S_qp6 = t_q dY/deps_q  (Y of operator.py:1415-1444 / splits.py:68-161)
 * qp_to_node:    out_n = scale sum_e sum_q N_a(q) f_q
 * qp_stress_apply: out_mech = [sum_q B_q^T S_q (3n), sum_q S_q (6)]
 * adj_damage_history: out_qp = scale 2(1-kappa) w (N_q.d - 1)(N_q.l)   ((dG/dH)^T l)
 * adj_damage_gc: out_e = scale l_e.(M/ell + ell L) d_e                   ((dG/dGc)^T l) */
int phast_hex8_set_gc(phast_hex8_plan *plan, const double *Gc, hipStream_t stream);
int phast_hex8_adj_coupling(const phast_hex8_plan *plan, const double *x, const double *l, const double *d,
                            double *out_qp, hipStream_t stream);
int phast_hex8_adj_driving(const phast_hex8_plan *plan, const double *x, const double *t_qp, double *S_qp6,
                           hipStream_t stream);
int phast_hex8_qp_to_node(const phast_hex8_plan *plan, const double *f_qp, double scale, double *out_n,
                          hipStream_t stream);
int phast_hex8_qp_stress_apply(const phast_hex8_plan *plan, const double *S_qp6, double *out_mech,
                               hipStream_t stream);
/* Element fields of PeriodicVoxelPhaseFieldOperator.field_snapshot (solid3d/operator.py:1561-1608,
 * without interface drivers or crack closure): per voxel the Gauss-point means of the total
 * engineering strain (E,6), of the degraded stress g C eps (E,6) and of 0.5 g eps . C eps (E,).
 * g_qp (E,8) may be null (g = 1). */
int phast_hex8_field_snapshot(const phast_hex8_plan *plan, const double *x, const double *g_qp, double *strain,
                              double *stress, double *psi, hipStream_t stream);
int phast_hex8_adj_damage_history(const phast_hex8_plan *plan, const double *d, const double *l, double scale,
                                  double *out_qp, hipStream_t stream);
int phast_hex8_adj_damage_gc(const phast_hex8_plan *plan, const double *d, const double *l, double scale,
                             double *out_e, hipStream_t stream);

/* ---- PF-CZM damage law and crack closure of the voxel cell -----------------------------
 * The default law of the reference's public solid3d API (solid3d/api.py:178-183).
 *
 * pfczm_setup: Wu's rational degradation g = (1-kappa) A/(A+Q) + kappa, A = (1-d)^p,
 *   Q = a1 d (1 + a2 d + a2 a3 d^2), a1 per voxel (host (E,)), p >= 2
 *   (physics/pfczm.py:17-130; operator.py:300-383 _build_pfczm).
 * pfczm_degradation: g, g', g'' (E,8) at the Gauss points of nodal d (operator.py:939-946,
 *   436-439 _pfczm_derivatives); any of the outputs may be null.
 * pfczm_energy: out3 (device) = [sum Gc (2 m.d - d.M d)/(pi ell), sum Gc ell/pi d.L d,
 *   sum w H g(d_q)]; the damage energy is their sum, the fracture energy the first two
 *   (operator.py:461-488 pfczm_damage_energy, 611-631 _pfczm_fracture_energy); H may be null.
 * pfczm_gradient: scale x the gradient of that energy (operator.py:490-512).
 * pfczm_tangent_weights: W (E,8) = max(H g'' - 2 Gc/(pi ell), 0.25 * 2 Gc/(pi ell))
 *   (operator.py:514-538); the floor everywhere when H is null.
 * pfczm_tangent_action / _diagonal: scale (Gc (2 ell/pi) L + sum_q w W_q N_q N_q^T) and its
 *   Jacobi diagonal (operator.py:540-564).
 * pfczm_dirichlet_pcg / pfczm_box_update: phast_hex8_damage_dirichlet_pcg and
 *   phast_hex8_box_update on that tangent, for the box-KKT subproblems of the projected Newton
 *   solve (operator.py:566-609 pfczm_newton_system; solid3d/staggered.py:843-1096).
 *
 * closure_setup: crack closure 'axes' of the isotropic phases (operator.py:631-690): cl_on
 *   (host uint8 (E,)) is 1 where a voxel adds (1 - g) sigma_c back (isotropic, with a damage
 *   driver); the plan must be isotropic.  Patterns are uint32 (E,): bit 4 q + k is
 *   [v_k . eps_q < 0] for Gauss point q and rows v = (trace, e_x, e_y, e_z) (splits.py:417-420).
 * closure_setup_rows: the closure of cells with a general stiffness bank, including the
 *   smeared-laminate interface band (operator.py:631-690, splits.py:424-742): cl_group (host
 *   uint8 (E,)) is 0 where a voxel adds nothing back, 1 for the isotropic group, 2 for the
 *   band group; rows (host, 28 doubles per bank entry) hold the entry's four closure rows
 *   v_k (4 x 6, engineering-strain contractions) and moduli k_k (4): (trace, e_x, e_y, e_z)
 *   with (lam, 2 mu, 2 mu, 2 mu) for an isotropic phase, (C m1, q2 + q3, q2, q3) with
 *   (1/a11, lam_S, 2 mu_S, 2 mu_S) for a band voxel (splits.py:603-670).
 * closure_pattern: newp = the pattern of state x (operator.py:792-806); with old (and g),
 *   out2 (device) = [max k |v . eps| over every bit, max (1 - g) k |v . eps| over the bits that
 *   differ from old, or -1 when none] for the isotropic group, followed by the same two
 *   values for the band group when the plan has one (out2 then holds 4 doubles) -- the test of
 *   operator.py:808-841 is out2[1] > tol * out2[0] or out2[3] > tol * out2[2].
 * mech_action_cl / stress_sums_cl / mech_diagonal_cl: phast_hex8_mech_action,
 *   phast_hex8_stress_sums and phast_hex8_mech_diagonal with the closure (operator.py:1039-1105,
 *   1196-1300); mode 1: each point's own pattern (the diagonal: every bit), 2: `pat`; g (E,8)
 *   required.
 * mech_solve_cl: phast_hex8_mech_solve with the closure at the fixed pattern `pat`
 *   (solid3d/staggered.py:550-587 _linear_mechanics with a ClosurePattern). */
int phast_hex8_pfczm_setup(phast_hex8_plan *plan, const double *a1, double a2, double a3, double p);
int phast_hex8_pfczm_degradation(const phast_hex8_plan *plan, const double *d, double *g, double *gp, double *gpp,
                                 hipStream_t stream);
int phast_hex8_pfczm_energy(const phast_hex8_plan *plan, const double *d, const double *H, double *out3,
                            hipStream_t stream);
int phast_hex8_pfczm_gradient(const phast_hex8_plan *plan, const double *d, const double *H, double scale,
                              double *out, hipStream_t stream);
int phast_hex8_pfczm_tangent_weights(const phast_hex8_plan *plan, const double *d, const double *H, double *W,
                                     hipStream_t stream);
int phast_hex8_pfczm_tangent_action(const phast_hex8_plan *plan, const double *f, const double *W, double scale,
                                    double *y, hipStream_t stream);
int phast_hex8_pfczm_tangent_diagonal(const phast_hex8_plan *plan, const double *W, double scale, double *diag,
                                      hipStream_t stream);
int phast_hex8_pfczm_dirichlet_pcg(const phast_hex8_plan *plan, const double *W, double scale, const double *b,
                                   const double *diag, const uint8_t *fixed, const double *values,
                                   const double *initial, double tol, int32_t max_iter, int64_t n, double *out,
                                   double *info, hipStream_t stream);
int phast_hex8_pfczm_box_update(const phast_hex8_plan *plan, const double *W, double scale, const double *b,
                                const double *cand, const double *lower, const double *upper,
                                uint8_t *lower_active, uint8_t *upper_active, const uint8_t *prescribed,
                                double kkt_tol, int32_t apply, int64_t n, double *grad, int64_t *counts,
                                hipStream_t stream);
int phast_hex8_closure_setup(phast_hex8_plan *plan, const uint8_t *cl_on);
int phast_hex8_closure_setup_rows(phast_hex8_plan *plan, const uint8_t *cl_group, const double *rows);
int phast_hex8_closure_pattern(const phast_hex8_plan *plan, const double *x, const double *g, const uint32_t *old,
                               uint32_t *newp, double *out2, hipStream_t stream);
int phast_hex8_mech_action_cl(const phast_hex8_plan *plan, const double *x, const double *g, const uint32_t *pat,
                              int32_t mode, double *y, hipStream_t stream);
int phast_hex8_stress_sums_cl(const phast_hex8_plan *plan, const double *x, const double *g, const uint32_t *pat,
                              int32_t mode, double *out7, hipStream_t stream);
int phast_hex8_mech_diagonal_cl(const phast_hex8_plan *plan, const double *g, const uint32_t *pat, int32_t mode,
                                double *diag, hipStream_t stream);
int phast_hex8_mech_solve_cl(const phast_hex8_plan *plan, const double *g, const uint32_t *pat,
                             const uint8_t *fixed, const double *fixed_values, const double *rhs, const double *x0,
                             const double *diag, double rtol, double atol, int32_t maxiter, int32_t check_every,
                             int64_t n_mech, double *x_out, double *info, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* PHAST_MI_H */
