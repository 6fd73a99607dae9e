/*
 * CPU FP64 oracle for the implicit Newton step of hex8 solid mechanics.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against; it is never the thing measured or shipped.  Only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark may load it.
 *
 * It restates, in plain C++17 without Eigen/TBB, the algorithms of dbeurle/neon
 * (file:line citations are relative to the reference repository root):
 *   hex8 shape functions      src/interpolations/hexahedron.cpp:16-67
 *   8-point Gauss rule        src/quadrature/hexahedron/hexahedron_quadrature.cpp:27-38
 *   deformation measures      src/mesh/mechanics/solid/submesh.cpp:210-269
 *   compressible neo-Hooke    src/constitutive/mechanics/solid/compressible_neohooke.cpp:17-90
 *   small-strain J2           src/constitutive/mechanics/solid/small_strain_J2_plasticity.cpp:30-163
 *   element K and f_int       src/mesh/mechanics/solid/submesh.cpp:66-160
 *   sparsity pattern          src/assembler/sparsity_pattern.hpp:18-48
 *   assembly (atomic add_to)  src/assembler/mechanics/static_matrix.hpp:243-277,
 *                             src/numeric/thread_safe_plugin.hpp:7-23
 *   Dirichlet handling        src/assembler/mechanics/static_matrix.hpp:279-351
 *   Newton driver + norms     src/assembler/mechanics/static_matrix.hpp:145-467
 *   load stepping / cutback   src/solver/adaptive_time_step.cpp:12-165
 *   load interpolation        src/mesh/boundary/boundary.cpp:37-85
 *   CPU CG / BiCGSTAB         src/solver/linear/linear_solver.cpp:39-138 (Eigen 3.3
 *                             ConjugateGradient / BiCGSTAB, restated below)
 *   CUDA CG / BiCGSTAB        src/solver/linear/conjugate_gradient_cuda.cpp:72-192,
 *                             src/solver/linear/biconjugate_gradient_stabilised_cuda.cpp:31-200
 *   RCM reordering            src/graph/cuthill_mckee.hpp:45-97
 *   linear elasticity         src/constitutive/mechanics/solid/isotropic_linear_elasticity.cpp:27-67
 *   microsphere models        src/constitutive/mechanics/solid/{gaussian_,}affine_microsphere.cpp,
 *                             nonaffine_microsphere.cpp:26-178, unit_sphere_quadrature.cpp:10-90
 *   ageing microsphere        gaussian_ageing_affine_microsphere.cpp:17-134,
 *                             src/material/micromechanical_elastomer.cpp:22-122
 *   consistent mass           src/mesh/mechanics/solid/submesh.cpp:162-183
 *   LATIN force and driver    src/mesh/mechanics/solid/latin_submesh.hpp:62-111,
 *                             src/assembler/mechanics/latin_matrix.hpp:150-244
 *
 * What pins each restatement (closed forms, the reference's own checks and vectors, invariants)
 * is stated in the test module that checks it (tests/test_oracle.py, test_microsphere.py,
 * test_natural_frequency.py, test_latin.py, test_adjoint_oracle.py, test_ingest.py).
 *
 * Every entry point returns an int status (ORC_* codes below).
 */
#ifndef NEON_ORACLE_H
#define NEON_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_OK = 0,
    ORC_NONCONVERGED = 1,     /* Krylov or Newton iteration limit (computational_error) */
    ORC_NAN_OR_INF = 2,       /* FE_INVALID after an update (computational_error) */
    ORC_NEGATIVE_JACOBIAN = 3,/* det F sign bit set (computational_error) */
    ORC_RETURN_MAP_FAILED = 4,/* radial return did not converge (computational_error) */
    ORC_BREAKDOWN = 5,        /* BiCGSTAB omega ~ 0 (computational_error) */
    ORC_FATAL = -1,           /* std::domain_error: too many cutbacks etc. */
    ORC_BAD_ARGUMENT = -2
};

/* Constitutive models (constitutive_model_factory.cpp:37-160).  The microsphere models take
 * their unit-sphere rule as a flag: BO21 (0) or ORC_SPHERE_BO33. */
enum {
    ORC_MODEL_NEOHOOKE = 0,             /* compressible_neohooke */
    ORC_MODEL_J2 = 1,                   /* small_strain_J2_plasticity */
    ORC_MODEL_LINEAR_ELASTIC = 2,       /* isotropic_linear_elasticity */
    ORC_MODEL_MICROSPHERE_GAUSSIAN = 3, /* gaussian_affine_microsphere */
    ORC_MODEL_MICROSPHERE_LANGEVIN = 4, /* affine_microsphere */
    ORC_MODEL_MICROSPHERE_NONAFFINE = 5,/* nonaffine_microsphere */
    ORC_MODEL_MICROSPHERE_AGEING = 6,   /* gaussian_ageing_affine_microsphere (BAND) */
    ORC_MODEL_J2_DAMAGE = 7,            /* small_strain_J2_plasticity_damage */
    ORC_SPHERE_BO33 = 0x100
};

/* Krylov stopping rule: Eigen's relative ||r|| < tol ||b|| (CPU solver) or the
 * CUDA solver's absolute sqrt(r.z) <= tol. */
enum { ORC_STOP_RELATIVE = 0, ORC_STOP_ABSOLUTE_PRECONDITIONED = 1 };

/* Tangent: "reference" reproduces neon's tangents (neo-Hooke D without 1/J);
 * "consistent" divides the neo-Hooke D by J so that K = d f_int / d u. */
enum { ORC_TANGENT_REFERENCE = 0, ORC_TANGENT_CONSISTENT = 1 };

/* ---- shape functions and single quadrature-point constitutive updates ---- */
int orc_hex8_tables(double* N /*[8 qp][8 node]*/, double* dN /*[8 qp][8 node][3]*/,
                    double* w /*[8]*/);

/* neo-Hooke at one QP: F (row-major 3x3), J = det F. Outputs sigma (3x3) and D (6x6). */
int orc_neohooke_qp(const double* F, double bulk, double shear, double* sigma, double* D);

/* Microsphere model (model id with its rule flag) at one QP: F (3x3), params (bulk K, shear G,
 * segments per chain N, non-affine stretch parameter p).  consistent != 0 returns the exact
 * Cauchy-based tangent instead of neon's.  Outputs sigma (3x3) and D (6x6). */
int orc_microsphere_qp(int32_t model, const double* F, const double* params, int32_t consistent,
                       double* sigma, double* D);
/* The unit-sphere rule of a microsphere model id: count, weights w[n], directions r[n][3]. */
int orc_sphere_rule(int32_t model, double* w /*33*/, double* r /*99*/);

/* J2 at one QP: H (3x3 displacement gradient); eps_p (3x3) and ebar are in/out history. */
int orc_j2_qp(const double* H, double bulk, double shear, double yield, double hardening,
              double* eps_p, double* ebar, double* sigma, double* D, double* von_mises);

/* ---- mesh-level model ---- */
typedef struct orc_model orc_model;

/* params: 4 per element, (bulk K, shear G, yield stress, hardening modulus) for J2 and
 * neo-Hooke / linear elasticity, (K, G, segments per chain, non-affine stretch parameter) for
 * the microsphere models. */
int orc_model_create(int32_t n_nodes, const double* X, int32_t n_elem, const int32_t* conn,
                     int32_t model, const double* params, int32_t tangent_mode,
                     orc_model** out);
int orc_model_destroy(orc_model* m);
int orc_model_n_dofs(const orc_model* m);

/* submesh::update_internal_variables with x = X + u.  Returns ORC_NEGATIVE_JACOBIAN,
 * ORC_NAN_OR_INF or ORC_RETURN_MAP_FAILED like the reference's computational_error. */
int orc_update_internal_variables(orc_model* m, const double* u, double dt);
int orc_save_internal_variables(orc_model* m, int32_t converged);

/* Ageing microsphere: scission and recombination probabilities (the mesh's material); the
 * per-element params are (K, G, segments per chain, initial inactive shear modulus
 * (1 - cure_time / 100) G).  Network state per QP (11 values): active and inactive shear
 * moduli, reduction factor, active and inactive segments, and the secondary network's
 * intermediate Kirchhoff stress (11, 22, 33, 23, 13, 12). */
int orc_model_set_ageing(orc_model* m, double scission, double recombination);

/* J2 with damage: the material constants of isotropic_elastic_plastic_damage (kinematic hardening
 * modulus, softening multiplier, plasticity viscous denominator and exponent, damage viscous
 * denominator and exponent); the per-element params are (K, G, yield stress, unused).  State per
 * QP: back stress (9), damage, energy release rate; the kinematic hardening variable is the back
 * stress over the modulus. */
int orc_model_set_damage(orc_model* m, const double* constants /*6*/);
int orc_get_damage_state(const orc_model* m, double* back_stress /*9/qp or NULL*/, double* damage /*1/qp or NULL*/,
                         double* energy /*1/qp or NULL*/);
/* The return map's residual (16) and Jacobian (16x16 row-major) at unknowns y (plastic and damage
 * multipliers, kinetic stress (6), kinetic back stress (6), damage, energy release rate), for the
 * trial elastic strain (3x3) and the back stress (3x3) and damage at the start of the update. */
int orc_j2_damage_system(const double* eps_e_t, const double* back_stress, double damage, double bulk, double shear,
                         double yield, const double* constants, double dt, const double* y, double* residual,
                         double* J);
/* One QP of small_strain_J2_plasticity_damage::update_internal_variables (in place on eps_p, ebar,
 * back_stress, damage, energy); kinematic hardening (9) written; iterations of the return map. */
int orc_j2_damage_qp(const double* H, double bulk, double shear, double yield, const double* constants, double dt,
                     double* eps_p, double* ebar, double* back_stress, double* damage, double* energy,
                     double* kinematic_hardening, double* sigma, double* D, double* von_mises, int32_t* iterations);
int orc_get_ageing_state(const orc_model* m, double* state /*11/qp*/);

/* Per-QP state downloads (QP index e*8 + l). */
int orc_get_qp_state(const orc_model* m, double* sigma /*9/qp or NULL*/, double* D /*36/qp or NULL*/,
                     double* detF /*1/qp or NULL*/, double* eps_p /*9/qp or NULL*/,
                     double* ebar /*1/qp or NULL*/, double* vm /*1/qp or NULL*/);

/* Element matrices at the current configuration (tangent_stiffness / internal_force). */
int orc_element_matrices(const orc_model* m, int32_t element, double* ke /*576*/, double* fe /*24*/);
/* Adjoint: grad (4 per element) = -lambda_e . d f_e / d params_e at the current state
 * (neo-Hooke only). */
int orc_parameter_gradient(const orc_model* m, const double* lambda, double* grad);

/* compute_sparsity_pattern: pass row_ptr=NULL to get nnz only. */
int64_t orc_sparsity_nnz(const orc_model* m);
int orc_sparsity_fill(const orc_model* m, int32_t* row_ptr, int32_t* col);

/* assemble_stiffness (parallel, binary search + atomic add) + compute_internal_force (serial). */
int orc_assemble(const orc_model* m, const int32_t* row_ptr, const int32_t* col, double* val,
                 double* f_int, int32_t threads);

/* Consistent mass (submesh::consistent_mass, submesh.cpp:162-183) of one element (576 values,
 * row-major 24x24) and assembled into the CSR pattern (density per element). */
int orc_element_mass(const orc_model* m, int32_t element, double density, double* me);
int orc_assemble_mass(const orc_model* m, const int32_t* row_ptr, const int32_t* col, const double* density,
                      double* val);

/* enforce_dirichlet_conditions on a CSR matrix and rhs. */
int orc_enforce_dirichlet(int32_t n, const int32_t* row_ptr, const int32_t* col, double* val,
                          double* b, int32_t n_fixed, const int32_t* fixed);

/* ---- Krylov solvers ---- */
int orc_spmv(int32_t n, const int32_t* row_ptr, const int32_t* col, const double* val,
             const double* x, double* y, int32_t threads);
/* Eigen ConjugateGradient (Jacobi) with neon's relative criterion; iterations as Eigen reports. */
int orc_cg_eigen(int32_t n, const int32_t* row_ptr, const int32_t* col, const double* val,
                 const double* b, double* x, double tol, int32_t max_it, int32_t threads,
                 int32_t* iterations, double* error);
/* conjugate_gradient_cuda::solve algorithm (absolute sqrt(r.z) criterion). */
int orc_cg_cuda(int32_t n, const int32_t* row_ptr, const int32_t* col, const double* val,
                const double* b, double* x, double tol, int32_t max_it, int32_t threads,
                int32_t* iterations, double* error);
int orc_bicgstab_eigen(int32_t n, const int32_t* row_ptr, const int32_t* col, const double* val,
                       const double* b, double* x, double tol, int32_t max_it, int32_t threads,
                       int32_t* iterations, double* error);
int orc_bicgstab_cuda(int32_t n, const int32_t* row_ptr, const int32_t* col, const double* val,
                      const double* b, double* x, double tol, int32_t max_it, int32_t threads,
                      int32_t* iterations, double* error);
/* Reverse Cuthill-McKee permutation of the CSR graph (perm[new] = old). */
int orc_rcm(int32_t n, const int32_t* row_ptr, const int32_t* col, int32_t* perm);

/* ---- load stepping ---- */
typedef struct {
    double period;
    double initial, minimum, maximum;
    int32_t adaptive;
} orc_time_options;

typedef struct orc_time orc_time;
int orc_time_create(const orc_time_options* opt, int32_t n_hist, const double* history, orc_time** out);
int orc_time_destroy(orc_time* t);
/* returns ORC_FATAL when neon would throw std::domain_error */
int orc_time_update(orc_time* t, int32_t converged);
int orc_time_query(const orc_time* t, double* step_time, double* last_step_time, int32_t* fully_applied,
                   int32_t* successful);

/* boundary::interpolate_prescribed_load for a time-value history (NaN outside it). */
double orc_boundary_value(int32_t n_points, const double* times, const double* values, double t);

/* ---- Newton driver (static_matrix::solve) ---- */
typedef struct {
    int32_t n_dofs;        /* number of constrained dofs */
    const int32_t* dofs;   /* global dof = 3*node + component */
    int32_t n_points;      /* time-value history */
    const double* times;
    const double* values;
} orc_dirichlet;

typedef struct {
    orc_time_options time;
    double residual_tolerance;
    double displacement_tolerance;
    int32_t max_newton_iterations; /* "linear_iterations", default 10 */
    int32_t absolute_norm;         /* "absolute_tolerance" present */
    int32_t norm_mode;             /* 0: neon (is_approx fallback), 1: always relative */
    double solver_tolerance;
    int32_t solver_max_iterations;
    int32_t solver_criterion;      /* ORC_STOP_* */
    int32_t solver;                /* 0 CG, 1 BiCGSTAB */
    int32_t threads;
    int32_t rcm;                   /* apply RCM + P^T K P per solve like neon's CPU CG */
    int32_t record_trajectory;     /* keep u and the committed J2 history of every converged step */
    int32_t latin;                 /* 1: latin_matrix's iterations (latin_matrix.hpp:172-244) */
    double latin_search_direction; /* "latin_search_direction" */
} orc_newton_options;

typedef struct {
    int32_t step;           /* successful increments before this iteration */
    int32_t iteration;      /* Newton iteration within the attempt */
    int32_t solver_iterations;
    int32_t status;         /* 0, or the recoverable code that ended the attempt */
    double step_time;
    double displacement_norm;
    double force_norm;
} orc_newton_record;

/* latin_submesh::incremental_latin_internal_force assembled (latin_matrix.hpp:110-126) at the
 * model's state: f = sum_e int B^T (sigma_old + s D (eps - eps_old) - sigma) dV on the current
 * configuration, eps and eps_old the symmetric parts of the current and committed displacement
 * gradients, sigma_old the committed stress.  f has 3N entries. */
int orc_latin_force(const orc_model* m, double search_direction, double* f);

int orc_newton_solve(orc_model* m, const orc_newton_options* opt, int32_t n_bc, const orc_dirichlet* bcs,
                     double* u /* in: ignored (zero start); out: final displacement */,
                     orc_newton_record* records, int32_t record_capacity, int32_t* n_records);

/* Nonfollower load: at step time t, f_ext[dofs[i]] += load(t) values[i], load(t) the
 * piecewise-linear history (boundary.cpp:50-85) at t, or at fixed_time when >= 0 (neon's
 * nodal forces use value_view()'s default load factor 1, static_matrix.hpp:230-236). */
typedef struct {
    int32_t n_entries;
    const int32_t* dofs;
    const double* values;
    int32_t n_points;
    const double* times;
    const double* loads;
    double fixed_time;
} orc_neumann;

/* orc_newton_solve with nonfollower loads (compute_external_force, static_matrix.hpp:205-241;
 * the load times join the step-time history, mesh.cpp:146-172). */
int orc_newton_solve_loaded(orc_model* m, const orc_newton_options* opt, int32_t n_bc, const orc_dirichlet* bcs,
                            int32_t n_loads, const orc_neumann* loads, double* u, orc_newton_record* records,
                            int32_t record_capacity, int32_t* n_records);

/* Adjoint of the whole load history of a J2 model in consistent mode (no counterpart in the
 * reference), after orc_newton_solve with record_trajectory: objective J(u_N) with gradient
 * dJdu (3N) at the final converged displacement, constrained dofs `fixed`; grad (4 per
 * element) = dJ/d(bulk, shear, yield, hardening).  Local derivatives of the radial return
 * come from complex-step differentiation of the oracle's own return map; the linear solves
 * are Jacobi-PCG to `tolerance` (relative).  The model's final state is restored. */
int orc_adjoint_j2(orc_model* m, const double* dJdu, int32_t n_fixed, const int32_t* fixed, double tolerance,
                   double* grad);
/* number of recorded steps of the last trajectory */
int orc_trajectory_steps(const orc_model* m);

#ifdef __cplusplus
}
#endif
#endif
