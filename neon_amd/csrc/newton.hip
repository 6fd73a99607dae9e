// Newton driver of the device-resident path: static_matrix::solve with its load stepping,
// Dirichlet handling, norms and cutback, on top of device_mesh and krylov_engine.
//
// Reference: src/assembler/mechanics/static_matrix.hpp:145-467 (driver, norms, Dirichlet),
// src/solver/adaptive_time_step.cpp:12-165 (load stepping), src/mesh/boundary/boundary.cpp:37-85
// (prescribed values), src/numeric/float_compare.hpp:20-32 (is_approx).
#include "engine.hpp"
#include "newton.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <queue>
#include <vector>

namespace neon_amd
{
bool is_approx(double x, double y, double abs_tol, double rel_tol)
{
    double rel;
    if (x == 0.0 && y == 0.0)
        rel = 0.0;
    else if (x == 0.0 || y == 0.0)
        rel = std::numeric_limits<double>::max(); // boost::math::relative_difference to zero
    else
        rel = std::abs(x - y) / std::min(std::abs(x), std::abs(y));
    return rel <= rel_tol || std::abs(x - y) <= abs_tol;
}

adaptive_time_step::adaptive_time_step(double period, double initial, double minimum, double maximum, bool adaptive,
                                       std::vector<double> history)
{
    if (history.empty()) throw status_error(NEON_BAD_ARGUMENT, "no boundary time history");
    for (double t : history) queue_.push(t);
    if (is_approx(queue_.top(), 0.0)) queue_.pop();
    if (queue_.empty()) throw status_error(NEON_BAD_ARGUMENT, "boundary time history has only t = 0");
    minimum_ = adaptive ? minimum : initial;
    maximum_ = adaptive ? maximum : initial;
    final_ = period;
    current_ = std::min(queue_.top(), initial);
}

bool adaptive_time_step::update(bool converged)
{
    if (converged)
    {
        if (current_ >= queue_.top()) queue_.pop();
        last_size_ = current_ - last_converged_;
        last_converged_ = current_;
        applied_ = is_approx(current_, final_) || queue_.empty();
        consecutive_unconverged_ = 0;
        consecutive_converged_++;
        successful_++;
        if (applied_) return true;
        bool const nonlinear = consecutive_unconverged_ > 0 || consecutive_converged_ < 4;
        double const new_time = std::min(nonlinear ? last_size_ + current_ : 2.0 * current_, current_ + maximum_);
        current_ = std::min(queue_.top(), std::min(new_time, final_));
        return true;
    }
    double const dt = std::max(minimum_, (current_ - last_converged_) * 0.5);
    current_ = last_converged_ + dt;
    if (current_ > final_)
    {
        failure_ = "minimum increment is not small enough to resolve the time step";
        return false;
    }
    if (current_ < 0.0)
    {
        failure_ = "Step has reduced below the final time for the last load case";
        return false;
    }
    consecutive_unconverged_++;
    consecutive_converged_ = 0;
    if (consecutive_unconverged_ == increment_limit_)
    {
        failure_ = "Too many reductions.  Convergence not judged likely";
        return false;
    }
    return true;
}

bool boundary_active(const neon_dirichlet& bc, double t) { return t >= bc.times[0] && t <= bc.times[bc.n_points - 1]; }

double boundary_value(const neon_dirichlet& bc, double t)
{
    for (int32_t i = 0; i < bc.n_points; ++i)
        if (is_approx(bc.times[i], t)) return bc.values[i];
    for (int32_t i = 0; i + 1 < bc.n_points; ++i)
        if (bc.times[i] < t && t < bc.times[i + 1])
        {
            double const f = (t - bc.times[i]) / (bc.times[i + 1] - bc.times[i]);
            return (1.0 - f) * bc.values[i] + f * bc.values[i + 1];
        }
    throw status_error(NEON_FATAL, "step time outside the boundary time history");
}

double load_value(const neon_neumann& load, double t)
{
    // the same interpolation as the Dirichlet histories (boundary.cpp:50-85)
    neon_dirichlet const h{0, nullptr, load.n_points, load.times, load.loads};
    double const at = load.fixed_time >= 0.0 ? load.fixed_time : t;
    for (int32_t i = 0; i < h.n_points; ++i)
        if (is_approx(h.times[i], at)) return h.values[i];
    for (int32_t i = 0; i + 1 < h.n_points; ++i)
        if (h.times[i] < at && at < h.times[i + 1])
        {
            double const f = (at - h.times[i]) / (h.times[i + 1] - h.times[i]);
            return (1.0 - f) * h.values[i] + f * h.values[i + 1];
        }
    throw status_error(NEON_FATAL, "step time outside a load's time history");
}

namespace
{
__global__ void sub_kernel(int64_t n, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ y)
{
    int64_t const i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) y[i] = a[i] - b[i];
}

__global__ void scale_copy_kernel(int64_t n, double a, const double* __restrict__ x, double* __restrict__ y)
{
    int64_t const i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) y[i] = a * x[i];
}

__global__ void add_kernel(int64_t n, const double* __restrict__ x, double* __restrict__ y)
{
    int64_t const i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) y[i] += x[i];
}

int blocks(int64_t n) { return int(std::max<int64_t>(1, (n + kBlock - 1) / kBlock)); }

double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}
} // namespace

solve_result device_solve(neon_context* ctx, int solver, int criterion, int precision, double tol, int32_t max_it,
                          const double* b_dev, double* x_dev, double* minv_dev)
{
    device_mesh& m = *ctx->mesh;
    krylov_engine& k = *ctx->engine;
    if (precision == NEON_PRECISION_FP64)
    {
        linear_operator const A = m.op(precision);
        m.jacobi(minv_dev, criterion == NEON_STOP_RELATIVE ? 1 : 0, precision);
        if (solver == NEON_SOLVER_CG) return k.cg64(A, b_dev, x_dev, minv_dev, criterion, tol, max_it);
        return k.bicgstab64(A, b_dev, x_dev, minv_dev, criterion, tol, max_it);
    }
    if (solver != NEON_SOLVER_CG) throw status_error(NEON_BAD_ARGUMENT, "mixed precision supports CG only");
    if (criterion != NEON_STOP_RELATIVE)
        throw status_error(NEON_BAD_ARGUMENT, "mixed precision uses the relative stopping rule");
    m.jacobi(minv_dev, 1, precision);
    // on the symmetric copy the FP32 recurrence runs on the Jacobi-scaled tangent (no
    // preconditioner vector in its updates); a diagonal that is not positive keeps the
    // unscaled Jacobi-PCG
    linear_operator const A = m.op(precision);
    if (m.symmetric_selected() && k.all_positive(A.n, minv_dev))
    {
        linear_operator const scaled = m.op_scaled(minv_dev);
        return k.cg_mixed_nb(A, b_dev, x_dev, minv_dev, tol, max_it, &scaled);
    }
    return k.cg_mixed_nb(A, b_dev, x_dev, minv_dev, tol, max_it);
}

__global__ __launch_bounds__(kBlock) void zero_entries_kernel(int32_t count, const int32_t* __restrict__ idx,
                                                              double* __restrict__ v)
{
    int32_t const i = blockIdx.x * kBlock + threadIdx.x;
    if (i < count) v[idx[i]] = 0.0;
}

int newton_solve(neon_context* ctx, const neon_newton_options* o, int32_t n_bc, const neon_dirichlet* bcs,
                 int32_t n_loads, const neon_neumann* loads, double* u_out, neon_newton_record* records, int32_t cap,
                 int32_t* n_records, neon_newton_timing* timing, bool latin, double latin_s)
{
    device_mesh& m = *ctx->mesh;
    krylov_engine& k = *ctx->engine;
    hipStream_t const s = ctx->stream;
    int64_t const n = m.dofs();
    auto const t_start = std::chrono::steady_clock::now();
    neon_newton_timing tm{};

    std::vector<double> hist;
    for (int32_t b = 0; b < n_bc; ++b)
    {
        if (bcs[b].n_points < 1 || !bcs[b].times || !bcs[b].values || (bcs[b].n_dofs > 0 && !bcs[b].dofs))
            throw status_error(NEON_BAD_ARGUMENT, "malformed boundary");
        for (int32_t i = 0; i < bcs[b].n_dofs; ++i)
            if (bcs[b].dofs[i] < 0 || bcs[b].dofs[i] >= n) throw status_error(NEON_BAD_ARGUMENT, "boundary dof out of range");
        hist.insert(hist.end(), bcs[b].times, bcs[b].times + bcs[b].n_points);
    }
    for (int32_t b = 0; b < n_loads; ++b)
    {
        const neon_neumann& L = loads[b];
        if (L.n_points < 1 || !L.times || !L.loads || L.n_entries < 0 || (L.n_entries > 0 && (!L.dofs || !L.values)))
            throw status_error(NEON_BAD_ARGUMENT, "malformed nonfollower load");
        for (int32_t i = 0; i < L.n_entries; ++i)
            if (L.dofs[i] < 0 || L.dofs[i] >= n) throw status_error(NEON_BAD_ARGUMENT, "load dof out of range");
        hist.insert(hist.end(), L.times, L.times + L.n_points);
    }
    // mesh::time_history collects the boundary and load times into a std::set (mesh.cpp:146-172)
    std::sort(hist.begin(), hist.end());
    hist.erase(std::unique(hist.begin(), hist.end()), hist.end());
    adaptive_time_step ts(o->period, o->initial, o->minimum, o->maximum, o->adaptive != 0, hist);

    device_array<double> u_old, delta, mres, inc, minv, fext;
    device_array<int32_t> fixed_dev;
    u_old.resize(n);
    delta.resize(n);
    mres.resize(n);
    inc.resize(n);
    minv.resize(n);
    NEON_CHECK_HIP(hipMemsetAsync(u_old.data(), 0, n * sizeof(double), s));
    NEON_CHECK_HIP(hipMemsetAsync(m.u(), 0, n * sizeof(double), s));
    // static_matrix's f_int starts at zero (the reaction force of the initial write)
    NEON_CHECK_HIP(hipMemsetAsync(m.f_int(), 0, n * sizeof(double), s));
    if (o->record_trajectory) m.clear_trajectory();
    NEON_CHECK_HIP(hipMemsetAsync(delta.data(), 0, n * sizeof(double), s));
    device_array<double> lres;
    if (latin)
    {
        // the committed stress starts at its initial zero (internal_variables), as u_old does
        lres.resize(n);
        m.latin_reset();
    }
    double* u = m.u();
    int const max_it = o->max_newton_iterations > 0 ? o->max_newton_iterations : 10;
    int const precision = o->precision;
    double norm_initial_residual = 1.0;
    int32_t nrec = 0;
    std::vector<double> inc_h(n);
    std::vector<int32_t> fixed_h;
    std::vector<double> fext_h;
    if (n_loads > 0)
    {
        fext.resize(n);
        fext_h.resize(n);
    }

    auto record = [&](int it, int sits, int status, double dn, double fn) {
        if (records && nrec < cap)
            records[nrec] = neon_newton_record{ts.successful(), it, sits, status, ts.step_time(), dn, fn};
        ++nrec;
    };

    int result = NEON_OK;
    int steps_done = 0;
    int64_t total_iterations = 0;
    bool capped = false;
    for (;;) // solve() and its retry after a recoverable error (static_matrix.hpp:168-182)
    {
        auto t0 = std::chrono::steady_clock::now();
        int err = m.update_internal_variables(u);
        tm.update_ms += ms_since(t0);
        if (err == NEON_OK && o->on_step)
        {
            NEON_CHECK_HIP(hipStreamSynchronize(s));
            o->on_step(o->callback_user, ts.successful(), ts.last_step_time());
        }
        while (err == NEON_OK && !ts.fully_applied())
        {
            if (o->max_load_steps > 0 && steps_done >= o->max_load_steps) break;
            // compute_external_force then perform_equilibrium_iterations (static_matrix.hpp:145-187,
            // 205-241, 402-467)
            NEON_CHECK_HIP(hipMemcpyAsync(u, u_old.data(), n * sizeof(double), hipMemcpyDeviceToDevice, s));
            double const step_time = ts.step_time(), last_time = ts.last_step_time();
            double fe = 0.0; // |f_ext|
            if (n_loads > 0)
            {
                std::fill(fext_h.begin(), fext_h.end(), 0.0);
                for (int32_t b = 0; b < n_loads; ++b)
                {
                    double const v = load_value(loads[b], step_time);
                    for (int32_t i = 0; i < loads[b].n_entries; ++i) fext_h[loads[b].dofs[i]] += v * loads[b].values[i];
                }
                double ss = 0.0;
                for (double const x : fext_h) ss += x * x;
                fe = std::sqrt(ss);
                NEON_CHECK_HIP(hipMemcpyAsync(fext.data(), fext_h.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
            }
            t0 = std::chrono::steady_clock::now();
            // the step's first update carries its increment to the rate models (:407)
            err = m.update_internal_variables(u, step_time - last_time);
            tm.update_ms += ms_since(t0);
            if (err != NEON_OK) break;

            // constrained dofs and prescribed increments of the active boundaries
            fixed_h.clear();
            std::fill(inc_h.begin(), inc_h.end(), 0.0);
            for (int32_t b = 0; b < n_bc; ++b)
            {
                if (!boundary_active(bcs[b], step_time)) continue;
                double const du = boundary_value(bcs[b], step_time) - boundary_value(bcs[b], last_time);
                for (int32_t i = 0; i < bcs[b].n_dofs; ++i) inc_h[bcs[b].dofs[i]] = du;
                fixed_h.insert(fixed_h.end(), bcs[b].dofs, bcs[b].dofs + bcs[b].n_dofs);
            }
            fixed_dev.resize(std::max<size_t>(1, fixed_h.size()));
            if (!fixed_h.empty())
                NEON_CHECK_HIP(hipMemcpyAsync(fixed_dev.data(), fixed_h.data(), fixed_h.size() * sizeof(int32_t),
                                              hipMemcpyHostToDevice, s));
            NEON_CHECK_HIP(hipMemcpyAsync(inc.data(), inc_h.data(), n * sizeof(double), hipMemcpyHostToDevice, s));

            int it = 0;
            while (it < max_it)
            {
                t0 = std::chrono::steady_clock::now();
                m.assemble(u, precision);
                tm.assemble_ms += ms_since(t0);
                t0 = std::chrono::steady_clock::now();
                // minus_residual = f_ext - f_int
                if (n_loads > 0)
                    sub_kernel<<<blocks(n), kBlock, 0, s>>>(n, fext.data(), m.f_int(), mres.data());
                else
                    scale_copy_kernel<<<blocks(n), kBlock, 0, s>>>(n, -1.0, m.f_int(), mres.data());
                if (it == 0)
                {
                    // apply_displacement_boundaries: r -= K du_p; u += du_p (static_matrix.hpp:322-351)
                    linear_operator const A = m.op(precision);
                    A.residual64(s, mres.data(), inc.data(), mres.data());
                    add_kernel<<<blocks(n), kBlock, 0, s>>>(n, inc.data(), u);
                    NEON_CHECK_HIP(hipGetLastError());
                    norm_initial_residual = std::sqrt(k.dot(n, mres.data(), mres.data()));
                }
                m.enforce_dirichlet(int32_t(fixed_h.size()), fixed_dev.data(), mres.data(), precision);
                // latin_matrix::perform_equilibrium_iterations (latin_matrix.hpp:209-226): after
                // the first iteration the right-hand side is the LATIN force at the current
                // state, its fixed entries zeroed with the residual's; the norms still use
                // f_ext - f_int
                const double* rhs = mres.data();
                if (latin && it > 0)
                {
                    m.latin_force(u, u_old.data(), latin_s, lres.data());
                    if (!fixed_h.empty())
                        zero_entries_kernel<<<blocks(int64_t(fixed_h.size())), kBlock, 0, s>>>(
                            int32_t(fixed_h.size()), fixed_dev.data(), lres.data());
                    NEON_CHECK_HIP(hipGetLastError());
                    rhs = lres.data();
                }
                NEON_CHECK_HIP(hipStreamSynchronize(s));
                tm.dirichlet_ms += ms_since(t0);

                t0 = std::chrono::steady_clock::now();
                solve_result const sr = device_solve(ctx, o->solver, o->solver_criterion, precision, o->solver_tolerance,
                                                     o->solver_max_iterations, rhs, delta.data(), minv.data());
                tm.solve_ms += ms_since(t0);
                tm.spmv_count += sr.spmv;
                tm.refinement_count += sr.refinements;
                if (sr.status != NEON_OK)
                {
                    record(it, sr.iterations, sr.status, 0, 0);
                    err = sr.status;
                    break;
                }
                t0 = std::chrono::steady_clock::now();
                add_kernel<<<blocks(n), kBlock, 0, s>>>(n, delta.data(), u);
                NEON_CHECK_HIP(hipGetLastError());
                err = m.update_internal_variables(u);
                tm.update_ms += ms_since(t0);
                if (err != NEON_OK)
                {
                    record(it, sr.iterations, err, 0, 0);
                    break;
                }
                // update_relative_norms (static_matrix.hpp:385-400)
                double const un = std::sqrt(k.dot(n, u, u));
                double const fi = std::sqrt(k.dot(n, m.f_int(), m.f_int()));
                double const dn = std::sqrt(k.dot(n, delta.data(), delta.data()));
                double const rn = std::sqrt(k.dot(n, mres.data(), mres.data()));
                bool const rel = o->norm_mode == 1 ? true
                                                   : (!o->absolute_norm && !is_approx(un, 0.0, 1.0)
                                                      && !is_approx(std::max(fe, fi), 0.0, 1.0));
                double dnorm, fnorm;
                if (rel)
                {
                    dnorm = dn / un;
                    fnorm = rn / std::max(norm_initial_residual, std::max(fe, fi));
                }
                else
                {
                    dnorm = dn;
                    fnorm = rn;
                }
                record(it, sr.iterations, NEON_OK, dnorm, fnorm);
                ++total_iterations;
                if (o->max_total_iterations > 0 && total_iterations >= o->max_total_iterations)
                {
                    capped = true;
                    break;
                }
                if (dnorm <= o->displacement_tolerance && fnorm <= o->residual_tolerance) break;
                ++it;
            }
            if (err != NEON_OK || capped) break;
            if (it == max_it)
            {
                err = NEON_NONCONVERGED;
                break;
            }
            NEON_CHECK_HIP(hipMemcpyAsync(u_old.data(), u, n * sizeof(double), hipMemcpyDeviceToDevice, s));
            if (o->record_trajectory) m.record_step();
            ts.update(true);
            m.save_internal_variables(true);
            if (latin) m.latin_commit();
            ++steps_done;
            if (o->on_step)
            {
                NEON_CHECK_HIP(hipStreamSynchronize(s));
                o->on_step(o->callback_user, ts.successful(), ts.last_step_time());
            }
        }
        if (err == NEON_OK) break;
        // computational_error: cutback, revert the history, reset u (static_matrix.hpp:168-182)
        if (!ts.update(false))
        {
            result = NEON_FATAL;
            ctx->error = ts.failure();
            break;
        }
        m.save_internal_variables(false);
        NEON_CHECK_HIP(hipMemcpyAsync(u, u_old.data(), n * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    NEON_CHECK_HIP(hipMemcpyAsync(u_out, u, n * sizeof(double), hipMemcpyDeviceToHost, s));
    NEON_CHECK_HIP(hipStreamSynchronize(s));
    if (n_records) *n_records = nrec;
    tm.total_ms = ms_since(t_start);
    if (timing) *timing = tm;
    return result;
}
int j2_adjoint(neon_context* ctx, const double* dJdu_dev, int32_t n_fixed, const int32_t* fixed_dev, double tol,
               int32_t max_it, double* grad_dev, int32_t* iterations)
{
    device_mesh& m = *ctx->mesh;
    hipStream_t const s = ctx->stream;
    int32_t const steps = m.trajectory_steps();
    if (steps == 0)
        throw status_error(NEON_BAD_ARGUMENT, "no recorded load history: run neon_newton_solve with record_trajectory");
    int64_t const n = m.dofs();
    device_array<double> rhs, lam, minv;
    rhs.resize(n);
    lam.resize(n);
    minv.resize(n);
    NEON_CHECK_HIP(hipMemsetAsync(grad_dev, 0, 4ull * m.elements() * sizeof(double), s));
    m.j2_adjoint_begin();
    int32_t total = 0;
    int status = NEON_OK;
    for (int32_t step = steps - 1; step >= 0 && status == NEON_OK; --step)
    {
        // state of step n: u_n with the return map from the committed h_{n-1}
        m.restore_step(step);
        status = m.update_internal_variables(m.u());
        if (status != NEON_OK) break;
        m.assemble(m.u(), NEON_PRECISION_FP64);
        if (step == steps - 1)
            NEON_CHECK_HIP(hipMemcpyAsync(rhs.data(), dJdu_dev, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        else
        {
            NEON_CHECK_HIP(hipMemsetAsync(rhs.data(), 0, n * sizeof(double), s));
            m.j2_adjoint_rhs(rhs.data());
        }
        m.enforce_dirichlet(n_fixed, fixed_dev, rhs.data(), NEON_PRECISION_FP64);
        solve_result const r = device_solve(ctx, NEON_SOLVER_CG, NEON_STOP_RELATIVE, NEON_PRECISION_FP64, tol, max_it,
                                            rhs.data(), lam.data(), minv.data());
        total += r.iterations;
        status = r.status;
        if (status != NEON_OK) break;
        m.j2_adjoint_history(lam.data(), grad_dev);
    }
    // back to the converged final state: u_N, history returned from h_{N-1} and committed
    m.restore_step(steps - 1);
    int const st = m.update_internal_variables(m.u());
    m.save_internal_variables(true);
    NEON_CHECK_HIP(hipStreamSynchronize(s));
    if (iterations) *iterations = total;
    return status != NEON_OK ? status : st;
}

} // namespace neon_amd
