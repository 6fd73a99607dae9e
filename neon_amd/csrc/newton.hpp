// Host-side load stepping and Newton driver declarations.
#pragma once

#include "engine.hpp"
#include "neon_hip.h"

#include <functional>
#include <queue>
#include <vector>

namespace neon_amd
{
// is_approx (src/numeric/float_compare.hpp:20-32) with boost's relative_difference, whose
// value against zero is the largest double.
bool is_approx(double x, double y, double abs_tol = 0.0, double rel_tol = 1e-10);

// adaptive_time_step (src/solver/adaptive_time_step.cpp:12-165): queue of mandatory times,
// doubling or last-increment growth after convergence, halving on failure, five
// consecutive failures fatal.
class adaptive_time_step
{
public:
    adaptive_time_step(double period, double initial, double minimum, double maximum, bool adaptive,
                       std::vector<double> history);
    // false where neon throws std::domain_error (adaptive_time_step.cpp:68-91); failure()
    // then holds neon's message
    bool update(bool converged);
    const char* failure() const { return failure_; }
    bool fully_applied() const { return applied_; }
    double step_time() const { return current_; }
    double last_step_time() const { return last_converged_; }
    double increment() const { return current_ - last_converged_; }
    int32_t successful() const { return successful_; }

private:
    std::priority_queue<double, std::vector<double>, std::greater<double>> queue_;
    int32_t increment_limit_ = 5, successful_ = 0, consecutive_converged_ = 0, consecutive_unconverged_ = 0;
    double final_ = 1, current_ = 1, last_converged_ = 0, last_size_ = 0, minimum_ = 0, maximum_ = 0;
    bool applied_ = false;
    const char* failure_ = "";
};

bool boundary_active(const neon_dirichlet& bc, double t);
double boundary_value(const neon_dirichlet& bc, double t);
// a nonfollower load's history value at step time t (fixed_time >= 0: at that time)
double load_value(const neon_neumann& load, double t);

solve_result device_solve(neon_context* ctx, int solver, int criterion, int precision, double tol, int32_t max_it,
                          const double* b_dev, double* x_dev, double* minv_dev);

// static_matrix::solve; with latin set, latin_matrix::solve (latin_matrix.hpp:150-244) with
// the search direction latin_s
int newton_solve(neon_context* ctx, const neon_newton_options* o, int32_t n_bc, const neon_dirichlet* bcs,
                 int32_t n_loads, const neon_neumann* loads, double* u_out, neon_newton_record* records,
                 int32_t cap, int32_t* n_records, neon_newton_timing* timing, bool latin = false,
                 double latin_s = 0.0);

// Adjoint of the recorded J2 load history (device buffers): grad (4 per element) of an
// objective of the final displacement with gradient dJdu; returns a status, PCG iterations
// summed into iterations.  The mesh's final converged state is restored.
int j2_adjoint(neon_context* ctx, const double* dJdu_dev, int32_t n_fixed, const int32_t* fixed_dev, double tol,
               int32_t max_it, double* grad_dev, int32_t* iterations);
// Natural frequencies (eigen.hip): the nev eigenvalues of K x = lambda M x at the lower or upper
// end of the spectrum, ascending, with M-orthonormal eigenvectors (n x nev, column-major, when
// eigenvectors is not null); solves receives the number of inner PCG solves.
int natural_frequency(neon_context* ctx, const double* density, int32_t n_fixed, const int32_t* fixed, int32_t nev,
                      int32_t spectrum, double tol, int32_t max_basis, int32_t block, double* eigenvalues,
                      double* eigenvectors, int32_t* solves);
} // namespace neon_amd
