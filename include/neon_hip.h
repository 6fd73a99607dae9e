/*
 * neon_hip.h — C ABI of the MI355X (gfx950) implicit Newton step for hex8 solid
 * mechanics.  Plain pointers and sizes only; every entry point returns an int status.
 *
 * The entry points replace these interfaces of dbeurle/neon (paths relative to the
 * reference repository root):
 *
 *   neon_krylov_solve_csr      conjugate_gradient_cuda::solve
 *                              (src/solver/linear/conjugate_gradient_cuda.cpp:72-192) and
 *                              biconjugate_gradient_stabilised_cuda::solve
 *                              (src/solver/linear/biconjugate_gradient_stabilised_cuda.cpp:31-200),
 *                              behind linear_solver::solve (src/solver/linear/linear_solver.hpp:18)
 *   neon_mesh_create           mechanics::solid::mesh / submesh construction
 *                              (src/mesh/mechanics/solid/mesh.cpp:25-46, submesh.cpp:26-52) with
 *                              the constitutive model from make_constitutive_model
 *                              (src/constitutive/constitutive_model_factory.cpp:26-164)
 *   neon_model_traits,         constitutive_model::is_finite_deformation / is_symmetric / variable_names
 *   neon_model_variable_names  (src/constitutive/constitutive_model.hpp:36-48); neon_mesh_model the
 *   neon_mesh_model            model behind mesh::is_symmetric (src/mesh/mechanics/solid/mesh.cpp:47-53)
 *   neon_mesh_pattern          compute_sparsity_pattern (src/assembler/sparsity_pattern.hpp:18-48)
 *   neon_mesh_set_displacement material_coordinates::update_current_configuration
 *                              (src/mesh/material_coordinates.cpp:23-27)
 *   neon_mesh_update_internal_variables
 *                              submesh::update_internal_variables (submesh.cpp:194-269) ->
 *                              constitutive_model::update_internal_variables
 *                              (src/constitutive/constitutive_model.hpp:36)
 *   neon_mesh_save_internal_variables
 *                              mesh::save_internal_variables (mesh.cpp:73-77)
 *   neon_mesh_set_ageing       ageing_micromechanical_elastomer (micromechanical_elastomer.cpp:34-53)
 *   neon_mesh_ageing_state     the ageing model's internal variables
 *                              (gaussian_ageing_affine_microsphere.cpp:17-50)
 *   neon_mesh_set_damage,      isotropic_elastic_plastic_damage (isotropic_elastic_plastic_damage.cpp:9-45)
 *   neon_mesh_damage_state     and small_strain_J2_plasticity_damage's internal variables
 *                              (small_strain_J2_plasticity_damage.cpp:17-30)
 *   neon_mesh_assemble         static_matrix::assemble_stiffness + compute_internal_force
 *                              (src/assembler/mechanics/static_matrix.hpp:189-203, 243-277) over
 *                              submesh::tangent_stiffness / internal_force (submesh.cpp:66-160)
 *   neon_mesh_enforce_dirichlet static_matrix::enforce_dirichlet_conditions (static_matrix.hpp:279-320)
 *   neon_mesh_solve            linear_solver::solve on the device-resident tangent (static_matrix.hpp:432)
 *   neon_newton_solve          static_matrix::solve (static_matrix.hpp:145-187, 402-467)
 *   neon_latin_solve           latin_matrix::solve (src/assembler/mechanics/latin_matrix.hpp:150-244)
 *                              with latin_submesh::incremental_latin_internal_force
 *                              (src/mesh/mechanics/solid/latin_submesh.hpp:62-111)
 *   neon_natural_frequency     natural_frequency_matrix::solve (natural_frequency_matrix.hpp:54-68)
 *   neon_mesh_mass             natural_frequency_matrix::assemble_mass (natural_frequency_matrix.hpp:106-137)
 *   neon_adjoint_gradient      no counterpart in the reference (new: adjoint of the converged state)
 *   neon_krylov_solve_csr_noalloc
 *                              no counterpart in the reference: neon_krylov_solve_csr on device
 *                              arrays without allocating, for callers that cannot allocate
 *
 * Threading: one host thread per context; a context owns one HIP stream and every
 * device buffer it allocates.  Host pointers are borrowed for the duration of a call.
 */
#ifndef NEON_HIP_H
#define NEON_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: > 0 recoverable (neon's computational_error), < 0 fatal ---- */
#define NEON_OK 0
#define NEON_NONCONVERGED 1      /* Krylov or Newton iteration limit */
#define NEON_NAN_OR_INF 2        /* non-finite stress or Krylov scalar (FE_INVALID) */
#define NEON_NEGATIVE_JACOBIAN 3 /* det F < 0 at a quadrature point */
#define NEON_RETURN_MAP_FAILED 4 /* J2 radial return did not converge in 50 iterations */
#define NEON_BREAKDOWN 5         /* BiCGSTAB omega == 0 */
#define NEON_FATAL -1            /* neon's std::domain_error (e.g. too many cutbacks) */
#define NEON_BAD_ARGUMENT -2
#define NEON_HIP_ERROR -3
#define NEON_OOM -4
#define NEON_NO_DEVICE -5

/* ---- enumerations ---- */
/* Constitutive models of make_constitutive_model (constitutive_model_factory.cpp:37-160).
 * The microsphere models take their unit-sphere rule (unit_sphere_quadrature.cpp:10-90) as a
 * flag: BO21 (default, 0) or NEON_SPHERE_BO33. */
#define NEON_MODEL_NEOHOOKE 0             /* compressible_neohooke */
#define NEON_MODEL_J2 1                   /* small_strain_J2_plasticity */
#define NEON_MODEL_LINEAR_ELASTIC 2       /* isotropic_linear_elasticity */
#define NEON_MODEL_MICROSPHERE_GAUSSIAN 3 /* gaussian_affine_microsphere */
#define NEON_MODEL_MICROSPHERE_LANGEVIN 4 /* affine_microsphere (Pade inverse Langevin chains) */
#define NEON_MODEL_MICROSPHERE_NONAFFINE 5 /* nonaffine_microsphere */
#define NEON_MODEL_MICROSPHERE_AGEING 6   /* gaussian_ageing_affine_microsphere (BAND network) */
#define NEON_MODEL_J2_DAMAGE 7            /* small_strain_J2_plasticity_damage (unsymmetric tangent) */
#define NEON_SPHERE_BO33 0x100            /* 33-direction rule instead of the 21-direction one */

#define NEON_TANGENT_REFERENCE 0  /* neon's tangents (neo-Hooke D without 1/J) */
#define NEON_TANGENT_CONSISTENT 1 /* exact Jacobian of f_int (D/J for neo-Hooke; the microsphere
                                     D with the term its deviatoric projections drop, over J;
                                     small-strain models integrated on X) */

#define NEON_SPECTRUM_LOWER 0 /* eigen_solver::eigen_spectrum::lower ("SM") */
#define NEON_SPECTRUM_UPPER 1 /* eigen_solver::eigen_spectrum::upper ("LM") */

#define NEON_SOLVER_CG 0
#define NEON_SOLVER_BICGSTAB 1

#define NEON_STOP_RELATIVE 0               /* CPU rule: ||r|| < tol ||b|| (Eigen) */
#define NEON_STOP_ABSOLUTE_PRECONDITIONED 1 /* CUDA rule: sqrt(r.z) <= tol */

#define NEON_PRECISION_FP64 0  /* FP64 values, vectors and arithmetic */
#define NEON_PRECISION_MIXED 1 /* FP32 tangent + inner PCG, FP64 refinement/dots/residuals */

typedef struct neon_context neon_context;

/* ---- context ---- */
int neon_device_count(int* count);
int neon_context_create(int device, neon_context** out);
int neon_context_destroy(neon_context* ctx);
/* Text of the last error on this context (valid until the next call). */
const char* neon_context_last_error(const neon_context* ctx);
/* Wait for all work on the context's stream. */
int neon_context_synchronize(neon_context* ctx);

/* ---- kernel timing (HIP events on the context's stream; no counterpart in the reference,
 * whose timing is std::chrono prints) ----
 * While enabled, every SpMV launch of the Krylov solvers and every assembly call is
 * bracketed by events; counts and summed milliseconds accumulate per kernel class until
 * read with reset.  Classes: */
#define NEON_KERNEL_SPMV_FP64 0     /* node-block SpMV, FP64 values/vectors */
#define NEON_KERNEL_SPMV_FP32 1     /* node-block SpMV, FP32 values/vectors (mixed inner) */
#define NEON_KERNEL_RESIDUAL_FP64 2 /* b - K x in FP64 (mixed reliable updates) */
#define NEON_KERNEL_ASSEMBLY 3      /* one full tangent + f_int assembly (all colours) */
#define NEON_KERNEL_CLASSES 4
int neon_context_profile(neon_context* ctx, int enable);
int neon_context_profile_read(neon_context* ctx, int kernel_class, int reset, int64_t* count, double* total_ms);

/* ---- linear solver on a host CSR matrix (drop-in for the CUDA solvers) ----
 * A is square, row-major CSR, int32 indices, FP64 values.  x is overwritten.  With
 * precision NEON_PRECISION_MIXED the inner iterations run in FP32 with FP64 reliable
 * updates of b - A x (only NEON_SOLVER_CG).  iterations / error receive the solver's
 * iteration count and final residual measure in the stopping rule's own form. */
int neon_krylov_solve_csr(neon_context* ctx, int solver, int criterion, int precision, int32_t n, int64_t nnz,
                          const int32_t* row_ptr, const int32_t* col, const double* val, const double* b,
                          double* x, double tolerance, int32_t max_iterations, int32_t* iterations,
                          double* error);
/* The same FP64 solve (NEON_SOLVER_CG or NEON_SOLVER_BICGSTAB, both stopping rules) with every
 * array in device memory and no allocation: row_ptr, col, val, b and x are device pointers; the
 * Jacobi inverse and the solver's work vectors live in a static per-GPU scratch array of the
 * library (9 vectors of 262,144 doubles, 18 MiB, held by each GPU the library runs on until the
 * process ends), so n is at most 262,144 (NEON_BAD_ARGUMENT beyond).  Malformed CSR arrays
 * (row_ptr not from 0 to nnz or not monotone, a column outside [0, n)) are found by a device
 * check and refused with NEON_BAD_ARGUMENT.  Calls that overlap on one GPU share its scratch
 * and race.  No counterpart in the reference: for callers that cannot allocate device memory
 * (code under HIP graph capture of its own buffers, tight serving loops, harnesses that forbid
 * hipMalloc). */
int neon_krylov_solve_csr_noalloc(neon_context* ctx, int solver, int criterion, int32_t n, int64_t nnz,
                                  const int32_t* row_ptr, const int32_t* col, const double* val, const double* b,
                                  double* x, double tolerance, int32_t max_iterations, int32_t* iterations,
                                  double* error);

/* ---- device-resident mesh, constitutive state and tangent ----
 * X: 3 per node, node-interleaved.  conn: 8 per element in hex8 order
 * (src/interpolations/hexahedron.cpp:21-28).  params: 4 per element: (bulk modulus, shear
 * modulus, yield stress, hardening modulus) for J2, of which neo-Hooke and linear elasticity
 * use the first two; (bulk modulus, shear modulus, segments per chain, non-affine stretch
 * parameter) for the microsphere models (micromechanical_elastomer.cpp:12-20,
 * nonaffine_microsphere.cpp:14-24).  One mesh per context. */
int neon_mesh_create(neon_context* ctx, int32_t n_nodes, const double* X, int32_t n_elem, const int32_t* conn,
                     int32_t model, const double* params, int32_t tangent_mode);
/* ---- constitutive model traits (constitutive_model::is_finite_deformation, is_symmetric and
 * variable_names, src/constitutive/constitutive_model.hpp:36-48; host only, no device use) ----
 * finite_deformation: the tangent carries the geometric term and is integrated on the current
 * configuration (submesh.cpp:75-78); symmetric: the tangent is symmetric, so neon solves with
 * conjugate gradient, otherwise with BiCGSTAB (mesh.cpp:47-53, linear_solver_factory.cpp:150-161).
 * NEON_BAD_ARGUMENT for an unknown model. */
int neon_model_traits(int32_t model, int32_t* finite_deformation, int32_t* symmetric);
/* The internal variable names the model adds to its `names` set, sorted and ';'-separated (as
 * neon's std::set iterates them); "" when it adds none; NULL for an unknown model. */
const char* neon_model_variable_names(int32_t model);
/* The model the mesh was created with (NEON_MODEL_* with its flags). */
int neon_mesh_model(neon_context* ctx, int32_t* model);
/* Replace the per-element parameters (4 per element) without rebuilding the pattern. */
int neon_mesh_set_parameters(neon_context* ctx, const double* params);
/* Sparsity pattern: nnz always; row_ptr (3N+1) and col (nnz) when non-NULL. */
int neon_mesh_pattern(neon_context* ctx, int64_t* nnz, int32_t* row_ptr, int32_t* col);
/* u: 3N displacements (host). */
int neon_mesh_set_displacement(neon_context* ctx, const double* u);
/* The displacement held by the mesh (3N, host). */
int neon_mesh_displacement(neon_context* ctx, double* u);
/* time_step is the rate models' step (the ageing microsphere integrates its network over it
 * when non-zero); neon passes the load increment at a step's first update and 0 otherwise
 * (static_matrix.hpp:151, 407, 436).  The other models ignore it. */
int neon_mesh_update_internal_variables(neon_context* ctx, double time_step);
int neon_mesh_save_internal_variables(neon_context* ctx, int32_t converged);
/* Ageing microsphere: the scission and recombination probabilities of the material
 * (micromechanical_elastomer.cpp:34-53, both >= 0; 0 before the first call).  Its per-element
 * params are (K, G, N, (1 - cure_time / 100) G): bulk modulus, shear modulus, segments per
 * chain, initial inactive shear modulus (gaussian_ageing_affine_microsphere.cpp:17-50). */
int neon_mesh_set_ageing(neon_context* ctx, double scission, double recombination);
/* Ageing microsphere: network state per QP (index e*8+l, 11 values each): active and inactive
 * shear modulus, reduction factor, active and inactive segments, secondary network's
 * intermediate stress (11, 22, 33, 23, 13, 12). */
int neon_mesh_ageing_state(neon_context* ctx, double* state);
/* J2 with damage (NEON_MODEL_J2_DAMAGE; per-element params (K, G, yield stress, unused)): the
 * material constants of isotropic_elastic_plastic_damage (isotropic_elastic_plastic_damage.cpp:
 * 9-45), in the order kinematic_hardening_modulus, softening_multiplier,
 * plasticity_viscous_denominator, plasticity_viscous_exponent, damage_viscous_denominator,
 * damage_viscous_exponent (all > 0; 1 before the first call).  The model's rate form integrates
 * over the time_step of neon_mesh_update_internal_variables (real time: neon passes the load
 * increment at a step's first update and 0 at its Newton iterates, where nothing flows). */
int neon_mesh_set_damage(neon_context* ctx, const double* constants);
/* J2 with damage: per QP (index e*8+l) the back stress (9, row-major), damage and energy release
 * rate; any pointer may be NULL.  The kinematic hardening variable is the back stress divided by
 * the kinematic hardening modulus (small_strain_J2_plasticity_damage.cpp:245). */
int neon_mesh_damage_state(neon_context* ctx, double* back_stress, double* damage, double* energy);
/* Per-QP state, index e*8+l.  Any pointer may be NULL.  sigma/eps_p 9 per QP (row-major
 * 3x3), D 36 per QP (6x6 Voigt, engineering shear in strain), detF/ebar/vm 1 per QP. */
int neon_mesh_qp_state(neon_context* ctx, double* sigma, double* D, double* detF, double* eps_p, double* ebar,
                       double* vm);
/* Assemble the tangent (FP64 or FP32 values per precision) and FP64 f_int on the device. */
int neon_mesh_assemble(neon_context* ctx, int32_t precision);
/* Download the assembled tangent values (CSR order, widened to FP64) and/or f_int. */
int neon_mesh_download(neon_context* ctx, double* val, double* f_int);
/* Zero the rows and columns of the fixed dofs, keeping the diagonal; if b is non-NULL
 * (host, length 3N) its fixed entries are zeroed too. */
int neon_mesh_enforce_dirichlet(neon_context* ctx, int32_t n_fixed, const int32_t* fixed, double* b);
/* Solve (device tangent) x = b with the given solver; b and x are host arrays of 3N. */
int neon_mesh_solve(neon_context* ctx, int solver, int criterion, double tolerance, int32_t max_iterations,
                    const double* b, double* x, int32_t* iterations, double* error);

/* Shape of the SpMV that the solvers on this mesh's tangent use from now on: with
 * NEON_SPMV_SLICED the sliced copy of the values (64 nodes per wavefront, one per lane), else
 * the CSR-order values with `lanes` (16, 32) lanes per block row; workgroups for FP64 and for
 * FP32 values (0: by size).  Default: NEON_SPMV_SLICED | NEON_SPMV_NONTEMPORAL |
 * NEON_SPMV_SYMMETRIC, workgroups by size (one slice per wavefront).  No counterpart in the
 * reference: a tuning knob; every shape gives the same products up to rounding (the order of
 * the row sums, and with NEON_SPMV_SYMMETRIC each lower block taken as the transpose of the
 * stored upper one). */
#define NEON_SPMV_NONTEMPORAL 1 /* stream the values with nontemporal loads */
#define NEON_SPMV_SLICED 256    /* the sliced copy of the values */
#define NEON_SPMV_INDEX32 1024  /* with NEON_SPMV_SLICED: 32-bit per-lane column offsets even where
                                   16 bits fit (the default uses 16 bits when every offset fits) */
#define NEON_SPMV_SYMMETRIC 2048 /* with NEON_SPMV_SLICED: where the model's tangent is symmetric and
                                    each 64-node slice's blocks have at most 40 distinct offsets
                                    (and the copy is at most 3/4 of the full one), a copy of the
                                    diagonal and upper blocks only, the lower ones read as their
                                    transposes */
int neon_mesh_spmv_config(neon_context* ctx, int32_t lanes, int32_t flags, int32_t blocks_fp64, int32_t blocks_fp32);
/* The sliced copy's shape: slices (64 nodes each), steps (block columns per lane, padding
 * included), steps whose 64 lanes share one column offset (stored once), and whether the other
 * steps' per-lane offsets are 16-bit (1) or 32-bit (0).  No counterpart in the reference. */
int neon_mesh_sliced_info(neon_context* ctx, int32_t* slices, int64_t* steps, int64_t* shared_steps, int32_t* index16);
/* The symmetric sliced copy's shape: whether the solvers on this mesh's tangent can use it
 * (a symmetric model and a pattern for which it pays: at most 40 distinct block offsets per
 * 64-node slice, a copy of at most 3/4 of the full one), the largest number of offsets of a
 * slice (0 when the pattern does not qualify) and the wave slots of its schedule.  No
 * counterpart in the reference. */
int neon_mesh_symmetric_info(neon_context* ctx, int32_t* usable, int32_t* offsets, int32_t* slots);

/* ---- Newton driver (static_matrix::solve) ---- */
typedef struct {
    int32_t n_dofs;      /* number of constrained dofs */
    const int32_t* dofs; /* global dof = 3 node + component */
    int32_t n_points;    /* piecewise-linear time history */
    const double* times;
    const double* values;
} neon_dirichlet;

/* Called by the Newton driver where neon's static_matrix::solve calls mesh.write
 * (static_matrix.hpp:150-154, 459-465): at the start of every solve attempt (the initial state,
 * and again after each cut-back, neon's solve() re-entering itself) and after every converged
 * load step, with the converged increments so far and the last converged time.  The context
 * then holds that state: the callback may read it through neon_mesh_displacement,
 * neon_mesh_qp_state and neon_mesh_download (f_int of the last assembly, neon's reaction
 * force is its negative; zero before the first), and must not change it. */
typedef void (*neon_step_callback)(void* user, int32_t step, double time);

typedef struct {
    double period, initial, minimum, maximum; /* "time" block */
    int32_t adaptive;
    double residual_tolerance, displacement_tolerance; /* "nonlinear_options" */
    int32_t max_newton_iterations;                     /* "linear_iterations" (default 10) */
    int32_t absolute_norm;                             /* "absolute_tolerance" given */
    int32_t norm_mode;                                 /* 0 neon's norms, 1 always relative */
    double solver_tolerance;                           /* "linear_solver" block */
    int32_t solver_max_iterations;
    int32_t solver_criterion; /* NEON_STOP_* */
    int32_t solver;           /* NEON_SOLVER_* */
    int32_t precision;        /* NEON_PRECISION_* */
    int32_t max_load_steps;   /* stop after this many converged increments (0: all) */
    int32_t max_total_iterations; /* stop after this many Newton iterations in all (0: no limit) */
    int32_t record_trajectory;    /* keep u and the committed history of every converged step on the
                                     device, for neon_adjoint_gradient on J2 models (0: off) */
    neon_step_callback on_step;   /* NULL: none */
    void* callback_user;
} neon_newton_options;

typedef struct {
    int32_t step;              /* converged increments before this iteration */
    int32_t iteration;         /* Newton iteration in this attempt */
    int32_t solver_iterations; /* Krylov iterations of this Newton iteration */
    int32_t status;            /* NEON_OK or the recoverable code that ended the attempt */
    double step_time;
    double displacement_norm;
    double force_norm;
} neon_newton_record;

typedef struct {
    double assemble_ms, dirichlet_ms, solve_ms, update_ms, total_ms;
    int64_t spmv_count;        /* SpMV launches in the Krylov phase */
    int64_t refinement_count;  /* FP64 residual recomputations (mixed precision) */
} neon_newton_timing;

/* u_out: 3N final displacement (host).  records may be NULL.  No nonfollower loads (f_ext = 0). */
int neon_newton_solve(neon_context* ctx, const neon_newton_options* options, int32_t n_bc,
                      const neon_dirichlet* bcs, double* u_out, neon_newton_record* records,
                      int32_t record_capacity, int32_t* n_records, neon_newton_timing* timing);

/* Nonfollower (Neumann) load of the Newton driver (static_matrix.hpp:205-241
 * compute_external_force, nonfollower_load.cpp:12-122): at a step time t it adds
 * load(t) * values[i] to f_ext[dofs[i]], where `values` is the load vector of a unit load
 * (neon_surface_load_vector / neon_volume_load_vector for traction, pressure and body force;
 * 1 at each dof for a nodal force) and load(t) interpolates (times, loads) piecewise
 * linearly (boundary.cpp:50-85).  fixed_time >= 0 evaluates the history at that time in
 * every step instead: neon applies nodal forces at value_view()'s default load factor 1
 * (static_matrix.hpp:230-236, nodal_value.hpp:29); fixed_time < 0 uses the step time. */
typedef struct {
    int32_t n_entries;
    const int32_t* dofs;   /* global dof = 3 node + component */
    const double* values;  /* unit-load vector entries */
    int32_t n_points;
    const double* times;
    const double* loads;
    double fixed_time;
} neon_neumann;

/* neon_newton_solve with nonfollower loads: the residual is f_ext - f_int and the relative
 * force norm divides by max(|r_0|, |f_ext|, |f_int|) (static_matrix.hpp:385-400); the load
 * histories' times join the step-time history (mesh.cpp:146-172).  A step time outside a
 * load's history is an error (neon interpolates past the end of its table there). */
int neon_newton_solve_loaded(neon_context* ctx, const neon_newton_options* options, int32_t n_bc,
                             const neon_dirichlet* bcs, int32_t n_loads, const neon_neumann* loads,
                             double* u_out, neon_newton_record* records, int32_t record_capacity,
                             int32_t* n_records, neon_newton_timing* timing);
/* latin_matrix::solve (latin_matrix.hpp:150-244): the iterations of neon_newton_solve_loaded,
 * except that after each load step's first iteration the right-hand side of K delta = r is the
 * LATIN force f = sum_e int B^T (sigma_old + s D (eps - eps_old) - sigma) dV on the current
 * configuration (latin_submesh.hpp:62-111; sigma_old and eps_old the committed stress and
 * strain, D the QP tangent, s = search_direction, neon's "latin_search_direction"), with the
 * constrained entries zeroed; convergence is judged on f_ext - f_int as for Newton.  Needs
 * NEON_TANGENT_REFERENCE. */
int neon_latin_solve(neon_context* ctx, const neon_newton_options* options, double search_direction, int32_t n_bc,
                     const neon_dirichlet* bcs, int32_t n_loads, const neon_neumann* loads, double* u_out,
                     neon_newton_record* records, int32_t record_capacity, int32_t* n_records,
                     neon_newton_timing* timing);

/* ---- adjoint (no counterpart in the reference) ----
 * Gradient of an objective J(u) of the final displacement with respect to the per-element
 * parameters (grad: 4 per element, in the order of params), through the adjoint of the
 * consistent tangent (needs NEON_TANGENT_CONSISTENT; the reference tangent is not dR/du).
 * dJdu (3N, host) is dJ/du at the displacement held by the context; the constrained dofs
 * `fixed` get the forward solve's Dirichlet rows and columns; every adjoint solve is the
 * Jacobi-PCG to the relative `tolerance` (K symmetric).  iterations: PCG iterations in all.
 *  - neo-Hooke (path independent): one solve K lambda = dJ/du at the held state, then
 *    dJ/dp_e = -lambda^T d f_e / d p_e.
 *  - J2 (path dependent): the load history of the last neon_newton_solve, which must have
 *    run with record_trajectory (u_n and the committed history h_{n-1} of every converged
 *    step are checkpointed on the device; the radial returns are recomputed from them).
 *    Backwards over the steps: K_n lambda_n = dJ/du [last step] + B^T (dh_n/d eps)^T mu_n,
 *    mu_{n-1} = (dh_n/dh_{n-1})^T mu_n - (d sigma_n/dh_{n-1})^T (w detJ B lambda_n), and
 *    dJ/dp += (dh_n/dp)^T mu_n - (d sigma_n/dp)^T (w detJ B lambda_n) for p = bulk, shear,
 *    yield stress, hardening.  The context's final state is restored afterwards. */
/* ---- natural frequencies (reuse of the device tangent and SpMV) ----
 * natural_frequency_matrix::solve (src/assembler/mechanics/natural_frequency_matrix.hpp:54-68)
 * with the arpack eigen_solver (src/solver/eigen/arpack.cpp:40-62): K at the mesh's held state
 * and the consistent mass (submesh::consistent_mass, submesh.cpp:162-183; density per element),
 * both with the fixed dofs' rows and columns zeroed and diagonals kept
 * (homogeneous_dirichlet.hpp:84-123), then the n_values eigenvalues of K x = lambda M x at the
 * lower or upper end of the spectrum (NEON_SPECTRUM_*), in increasing order.  Block Krylov with
 * full M-orthogonalisation and Rayleigh-Ritz over the device PCG (lower end: shift-invert,
 * OP = K^-1 M; upper end: OP = M^-1 K), stopped when every wanted Ritz pair's residual
 * |OP y - theta y|_M <= tolerance |theta|.  max_basis (0: max(4 n_values, n_values + 40)) caps
 * the basis, block (0: 4) is the block width, which must reach the largest eigenvalue
 * multiplicity wanted.  eigenvectors (3N x n_values, column-major, M-orthonormal) may be NULL;
 * solves (may be NULL) receives the number of inner PCG solves.  NEON_NONCONVERGED: "Eigenvalues
 * did not converge" within the basis cap. */
int neon_natural_frequency(neon_context* ctx, const double* density, int32_t n_fixed, const int32_t* fixed,
                           int32_t n_values, int32_t spectrum, double tolerance, int32_t max_basis, int32_t block,
                           double* eigenvalues, double* eigenvectors, int32_t* solves);
/* The consistent mass (density per element) in the pattern's CSR order (nnz values), FP64.
 * submesh::consistent_mass assembled as natural_frequency_matrix::assemble_mass (:106-137). */
int neon_mesh_mass(neon_context* ctx, const double* density, double* val);

int neon_adjoint_gradient(neon_context* ctx, const double* dJdu, int32_t n_fixed, const int32_t* fixed,
                          double tolerance, int32_t max_iterations, double* grad, int32_t* iterations);

/* ---- synthetic hex8 meshes (host only, no device use) ----
 * Structured nx*ny*nz block of size Lx*Ly*Lz; node n = i + (nx+1)(j + (ny+1)k).  Nodes are
 * jittered by U(-jitter h, +jitter h) per component with a counter-based SplitMix64 stream
 * keyed by seed, except normal to a boundary face.  Elements whose centroid lies within
 * hole_radius of the z-parallel axis through the block centre are removed (hole_radius
 * <= 0: none) and unused nodes dropped (node numbering stays increasing). */
int neon_mesh_box_size(int32_t nx, int32_t ny, int32_t nz, double Lx, double Ly, double Lz, double hole_radius,
                       int32_t* n_nodes, int32_t* n_elem);
int neon_mesh_box(int32_t nx, int32_t ny, int32_t nz, double Lx, double Ly, double Lz, double jitter,
                  uint64_t seed, double hole_radius, double* X, int32_t* conn, int32_t* node_ijk);
/* ---- unit nonfollower load vectors (host only, no device use) ----
 * Load vector of a unit load on the initial configuration X (3N), accumulated into `out`
 * (3N, dense):
 *  - neon_surface_load_vector: quadrilateral faces (4 nodes each in neon's quadrilateral4
 *    order, quadrilateral.cpp:17-46) with 2x2 Gauss points (the centre when reduced,
 *    quadrilateral_quadrature.cpp:10-24).  component 0..2: traction,
 *    out[3a+c] += sum_l w N_a j with j = |x_xi x x_eta| (neumann.hpp:104-118);
 *    component -1: pressure, out[3a+i] -= sum_l w N_a n_i j, n = (x_xi x x_eta) / j
 *    (pressure.cpp:8-31).
 *  - neon_volume_load_vector: hex8 elements (hexahedron.cpp:21-28 order), 2x2x2 Gauss points
 *    (the centre when reduced); body force, out[3a+c] += sum_l w N_a det J
 *    (neumann.hpp:159-173).
 * Triangular faces are not supported: neon's one-point triangle rule weighs 1 instead of
 * 1/2 (triangle_quadrature.cpp:12-17), doubling such loads. */
int neon_surface_load_vector(int32_t n_nodes, const double* X, int32_t n_faces, const int32_t* faces,
                             int32_t component, int32_t reduced, double* out);
int neon_volume_load_vector(int32_t n_nodes, const double* X, int32_t n_elem, const int32_t* conn,
                            int32_t component, int32_t reduced, double* out);

/* Uniform(lo, hi) per index from the counter-based SplitMix64 stream (n values). */
int neon_random_uniform(uint64_t seed, int64_t n, double lo, double hi, double* out);
/* Standard normal per index (Box-Muller on two stream values). */
int neon_random_normal(uint64_t seed, int64_t n, double* out);

#ifdef __cplusplus
}
#endif
#endif
