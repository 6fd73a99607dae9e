// Host engine of the gfx950 implicit Newton step: Krylov drivers and the device mesh.
#include "engine.hpp"
#include "damage_kernels.hpp"
#include "element_kernels.hpp"
#include "krylov_kernels.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iterator>
#include <map>
#include <numeric>
#include <tuple>
#include <type_traits>

#include <omp.h>

namespace neon_amd
{
namespace
{
inline int grid_for(int64_t n, int per_block = kBlock)
{
    int64_t const g = (n + per_block - 1) / per_block;
    return int(std::max<int64_t>(1, g));
}

// Round a block count to a multiple of 8 (one share per XCD) within [1, cap].
inline int xcd_blocks(int64_t want, int cap)
{
    int64_t b = std::min<int64_t>(std::max<int64_t>(want, 1), cap);
    if (b >= 8) b = (b / 8) * 8;
    return int(b);
}
} // namespace

// ================================================================ operator
int linear_operator::spmv_blocks() const
{
    if (node_block && sym)
    {
        // one wavefront per wave slot in one pass, or fewer workgroups (a multiple of 8, so that
        // each slot keeps its XCD) striding over the slots
        int64_t const needed = int64_t(sym_slots) / 4;
        int64_t want = blocks_override > 0 ? std::min<int64_t>(blocks_override, needed) : needed;
        want = std::min<int64_t>(want, kMaxPartials);
        return int(std::max<int64_t>(8, want / 8 * 8));
    }
    if (node_block && sell)
    {
        // one wavefront per slice in one pass (4 per workgroup), or fewer workgroups striding
        // over the slices
        int64_t const needed = (int64_t(slices) + 3) / 4;
        int64_t const want = blocks_override > 0 ? std::min<int64_t>(blocks_override, needed) : needed;
        return int(std::max<int64_t>(1, std::min<int64_t>(want, kMaxPartials)));
    }
    int64_t const rows = node_block ? nodes : n;
    int64_t const rows_per_pass = node_block ? kBlock / nb_lanes : kBlock / lanes_per_row;
    if (blocks_override > 0)
    {
        // at most one pass of lane groups over the rows
        int64_t const needed = ((rows + rows_per_pass - 1) / rows_per_pass + 7) / 8 * 8;
        return xcd_blocks(std::min<int64_t>(blocks_override, needed), kMaxPartials);
    }
    // about 4 passes per block, at most kMaxPartials blocks
    return xcd_blocks((rows + 4 * rows_per_pass - 1) / (4 * rows_per_pass), kMaxPartials);
}

namespace
{
template <typename VT, typename XT, typename YT, typename AT, bool DOT, int MODE>
void launch_spmv(const linear_operator& A, hipStream_t s, const VT* val, const VT* sval, const XT* x, YT* y,
                 const double* b, double* part, const krylov_state* st, int gate = 0)
{
    int const blocks = A.spmv_blocks();
    if (A.node_block && A.sym)
    {
        if (!sval) throw status_error(NEON_BAD_ARGUMENT, "symmetric operator without its values");
        sym_matrix const m{A.nodes,     A.slices,    A.sym_slots,
                           A.sym_records, A.sym_sched, A.sym_rdiag,
                           A.sym_eptr,  reinterpret_cast<const int4v*>(A.sym_ent),
                           A.sym_rslice, A.sym_roff};
        spmv_sym_kernel<VT, XT, YT, AT, DOT, MODE><<<blocks, 256, 0, s>>>(m, sval, x, y, b, part, st, gate);
        NEON_CHECK_HIP(hipGetLastError());
        return;
    }
    if (A.node_block && A.sell)
    {
        if (!sval) throw status_error(NEON_BAD_ARGUMENT, "sliced operator without sliced values");
        sell_matrix m{A.nodes, A.slices, A.sptr, A.snode, A.slane, A.scode, A.sexc, A.soff16, A.soff32};
#define NEON_SELL(I16, NT) \
    spmv_sell_kernel<I16, NT, VT, XT, YT, AT, DOT, MODE><<<blocks, 256, 0, s>>>(m, sval, x, y, b, part, st, gate)
        if (A.nb_nt)
        {
            if (A.soff16)
                NEON_SELL(true, true);
            else
                NEON_SELL(false, true);
        }
        else
        {
            if (A.soff16)
                NEON_SELL(true, false);
            else
                NEON_SELL(false, false);
        }
#undef NEON_SELL
        NEON_CHECK_HIP(hipGetLastError());
        return;
    }
    if (A.node_block)
    {
        nb_matrix m{A.nodes, A.bptr, A.bcol};
#define NEON_NB(L, NT) \
    spmv_nb_kernel<L, 1, NT, VT, XT, YT, AT, DOT, MODE><<<blocks, kBlock, 0, s>>>(m, val, x, y, b, part, st, gate)
        int const key = (A.nb_lanes == 16 ? 0 : 2) + (A.nb_nt ? 1 : 0);
        switch (key)
        {
        case 0: NEON_NB(16, false); break;
        case 1: NEON_NB(16, true); break;
        case 2: NEON_NB(32, false); break;
        default: NEON_NB(32, true); break;
        }
#undef NEON_NB
    }
    else
    {
        csr_matrix m{int32_t(A.n), A.rp, A.ci};
        switch (A.lanes_per_row)
        {
        case 1: spmv_csr_kernel<1, VT, XT, YT, AT, DOT, MODE><<<blocks, kBlock, 0, s>>>(m, val, x, y, b, part, st, gate); break;
        case 2: spmv_csr_kernel<2, VT, XT, YT, AT, DOT, MODE><<<blocks, kBlock, 0, s>>>(m, val, x, y, b, part, st, gate); break;
        case 4: spmv_csr_kernel<4, VT, XT, YT, AT, DOT, MODE><<<blocks, kBlock, 0, s>>>(m, val, x, y, b, part, st, gate); break;
        case 8: spmv_csr_kernel<8, VT, XT, YT, AT, DOT, MODE><<<blocks, kBlock, 0, s>>>(m, val, x, y, b, part, st, gate); break;
        case 16: spmv_csr_kernel<16, VT, XT, YT, AT, DOT, MODE><<<blocks, kBlock, 0, s>>>(m, val, x, y, b, part, st, gate); break;
        case 32: spmv_csr_kernel<32, VT, XT, YT, AT, DOT, MODE><<<blocks, kBlock, 0, s>>>(m, val, x, y, b, part, st, gate); break;
        default: spmv_csr_kernel<64, VT, XT, YT, AT, DOT, MODE><<<blocks, kBlock, 0, s>>>(m, val, x, y, b, part, st, gate); break;
        }
    }
    NEON_CHECK_HIP(hipGetLastError());
}
} // namespace

void linear_operator::apply64(hipStream_t s, const double* x, double* y, double* part, const krylov_state* st,
                              int gate) const
{
    if (!v64) throw status_error(NEON_BAD_ARGUMENT, "operator has no FP64 values");
    if (part)
        launch_spmv<double, double, double, double, true, 0>(*this, s, v64, s64, x, y, nullptr, part, st, gate);
    else
        launch_spmv<double, double, double, double, false, 0>(*this, s, v64, s64, x, y, nullptr, nullptr, st, gate);
}

void linear_operator::apply32(hipStream_t s, const float* x, float* y, double* part, const krylov_state* st) const
{
    if (!v32) throw status_error(NEON_BAD_ARGUMENT, "operator has no FP32 values");
    if (part)
        launch_spmv<float, float, float, float, true, 0>(*this, s, v32, s32, x, y, nullptr, part, st);
    else
        launch_spmv<float, float, float, float, false, 0>(*this, s, v32, s32, x, y, nullptr, nullptr, st);
}

void linear_operator::residual64(hipStream_t s, const double* b, const double* x, double* r, const krylov_state* st,
                                 int gate) const
{
    if (v64)
        launch_spmv<double, double, double, double, false, 1>(*this, s, v64, s64, x, r, b, nullptr, st, gate);
    else
        launch_spmv<float, double, double, double, false, 1>(*this, s, v32, s32, x, r, b, nullptr, st, gate);
}

// ================================================================ profiler
kernel_profiler::~kernel_profiler()
{
    for (hipEvent_t e : events_) (void)hipEventDestroy(e);
}

int kernel_profiler::begin(hipStream_t s, int cls)
{
    if (!enabled || pairs_ >= kMaxPairs) return -1;
    while (int(events_.size()) < 2 * (pairs_ + 1))
    {
        hipEvent_t e;
        NEON_CHECK_HIP(hipEventCreate(&e));
        events_.push_back(e);
    }
    if (int(cls_.size()) <= pairs_) cls_.resize(pairs_ + 1);
    cls_[pairs_] = cls;
    NEON_CHECK_HIP(hipEventRecord(events_[2 * pairs_], s));
    return pairs_++;
}

void kernel_profiler::end(hipStream_t s, int token)
{
    if (token < 0) return;
    NEON_CHECK_HIP(hipEventRecord(events_[2 * token + 1], s));
}

void kernel_profiler::collect(int64_t valid)
{
    for (int i = 0; i < pairs_; ++i)
    {
        if (valid >= 0 && i >= valid) break;
        float t = 0.0f;
        NEON_CHECK_HIP(hipEventSynchronize(events_[2 * i + 1]));
        NEON_CHECK_HIP(hipEventElapsedTime(&t, events_[2 * i], events_[2 * i + 1]));
        ms_[cls_[i]] += t;
        count_[cls_[i]] += 1;
    }
    pairs_ = 0;
}

void kernel_profiler::read(int cls, bool reset, int64_t* count, double* ms)
{
    if (cls < 0 || cls >= NEON_KERNEL_CLASSES) throw status_error(NEON_BAD_ARGUMENT, "unknown kernel class");
    if (count) *count = count_[cls];
    if (ms) *ms = ms_[cls];
    if (reset)
    {
        count_[cls] = 0;
        ms_[cls] = 0.0;
    }
}

// ================================================================== engine
krylov_engine::krylov_engine(hipStream_t s) : s_(s)
{
    // partials, then their total (totals_kernel)
    for (auto& p : part_) p.resize(kMaxPartials + 1);
    state_.resize(2);
    scalar_.resize(1);
    NEON_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_state_), sizeof(krylov_state)));
    NEON_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_ring_), 2 * sizeof(krylov_state)));
    for (auto& e : ring_events_) NEON_CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    NEON_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_scalar_), sizeof(double)));
}

krylov_engine::~krylov_engine()
{
    if (host_state_) (void)hipHostFree(host_state_);
    if (host_ring_) (void)hipHostFree(host_ring_);
    for (auto& e : ring_events_)
        if (e) (void)hipEventDestroy(e);
    if (host_scalar_) (void)hipHostFree(host_scalar_);
}

#ifndef NEON_VEC_BLOCKS_MAX
#define NEON_VEC_BLOCKS_MAX 1024
#endif
int krylov_engine::vec_blocks(int64_t n) const
{
    return xcd_blocks((n + 4 * kBlock - 1) / (4 * kBlock), NEON_VEC_BLOCKS_MAX);
}

void krylov_engine::ensure(int64_t n)
{
    if (n == n_) return;
    r_.resize(n);
    p_.resize(n);
    Ap_.resize(n);
    n_ = n;
}

double krylov_engine::read_scalar(const double* d)
{
    NEON_CHECK_HIP(hipMemcpyAsync(host_scalar_, d, sizeof(double), hipMemcpyDeviceToHost, s_));
    NEON_CHECK_HIP(hipStreamSynchronize(s_));
    return *host_scalar_;
}

double krylov_engine::dot(int64_t n, const double* a, const double* b)
{
    int const nb = vec_blocks(n);
    dot_kernel<<<nb, kBlock, 0, s_>>>(n, a, b, part_[0].data());
    reduce_to_kernel<<<1, kBlock, 0, s_>>>(part_[0].data(), nb, scalar_.data());
    NEON_CHECK_HIP(hipGetLastError());
    NEON_CHECK_HIP(hipMemcpyAsync(host_scalar_, scalar_.data(), sizeof(double), hipMemcpyDeviceToHost, s_));
    NEON_CHECK_HIP(hipStreamSynchronize(s_));
    return *host_scalar_;
}

void krylov_engine::jacobi64(const linear_operator& A, double* minv, int eigen_rule)
{
    csr_matrix m{int32_t(A.n), A.rp, A.ci};
    jacobi_kernel<double><<<grid_for(A.n), kBlock, 0, s_>>>(m, A.v64, minv, eigen_rule);
    NEON_CHECK_HIP(hipGetLastError());
}

namespace
{
solve_result finish(const krylov_state& s, int criterion)
{
    solve_result r;
    r.iterations = s.it;
    r.error = criterion == NEON_STOP_RELATIVE ? (s.rhs2 > 0 ? std::sqrt(s.rr / s.rhs2) : 0.0) : std::sqrt(s.rz);
    switch (s.done)
    {
    case KRYLOV_CONVERGED: r.status = NEON_OK; break;
    case KRYLOV_LIMIT: r.status = NEON_NONCONVERGED; break;
    case KRYLOV_NONFINITE: r.status = NEON_NAN_OR_INF; break;
    case KRYLOV_BREAKDOWN: r.status = NEON_BREAKDOWN; break;
    default: r.status = NEON_NONCONVERGED; break;
    }
    return r;
}
} // namespace

// Drives a device-resident Krylov loop: `batch` launches poll_every iterations (advancing
// `launched`; the state after `launched` iterations is st[launched & 1]).  The host keeps one
// batch queued ahead of the one whose final state it is waiting for, so the device never idles
// on a poll; iterations launched after convergence are no-ops.  On return the stream is idle
// and host_state_ holds the final state.
template <typename Batch>
void krylov_engine::run_batches(krylov_state* st, int64_t& launched, Batch&& batch)
{
    auto snapshot = [&](int slot) {
        NEON_CHECK_HIP(hipMemcpyAsync(host_ring_ + slot, st + (launched & 1), sizeof(krylov_state),
                                      hipMemcpyDeviceToHost, s_));
        NEON_CHECK_HIP(hipEventRecord(ring_events_[slot], s_));
    };
    snapshot(0);
    NEON_CHECK_HIP(hipEventSynchronize(ring_events_[0]));
    if (!host_ring_[0].done)
    {
        batch();
        NEON_CHECK_HIP(hipGetLastError());
        snapshot(1);
        for (int wait = 1;; wait ^= 1)
        {
            batch();
            NEON_CHECK_HIP(hipGetLastError());
            snapshot(wait ^ 1);
            NEON_CHECK_HIP(hipEventSynchronize(ring_events_[wait]));
            if (host_ring_[wait].done) break;
        }
    }
    NEON_CHECK_HIP(hipMemcpyAsync(host_state_, st + (launched & 1), sizeof(krylov_state), hipMemcpyDeviceToHost, s_));
    NEON_CHECK_HIP(hipStreamSynchronize(s_));
}

// Jacobi-PCG in FP64.  Per iteration: SpMV fused with the block partials of p.Ap, their total
// (one block: the SpMV has one partial per slice quadruple, which every update block would
// otherwise reduce), the r update fused with the partials of r.r and r.z, and the x and p
// updates with the convergence decision (each of its blocks reduces the r update's partials;
// x += alpha p sits there because that kernel reads p anyway: 10 vector passes, not 11);
// scalars stay on the device and the host polls the state every `poll_every` iterations (a
// batch queued ahead of the one it waits on).
// Iterations past convergence are no-ops, so the reported count is exact.
solve_result krylov_engine::cg64(const linear_operator& A, const double* b, double* x, const double* minv,
                                 int criterion, double tol, int32_t max_it, const krylov_work* work)
{
    int64_t const n = A.n;
    double *r = nullptr, *p = nullptr, *Ap = nullptr;
    if (work)
    {
        r = work->r, p = work->p, Ap = work->Ap;
    }
    else
    {
        ensure(n);
        r = r_.data(), p = p_.data(), Ap = Ap_.data();
    }
    int const nb = vec_blocks(n);
    int const ns = A.spmv_blocks();
    double* part_pAp = part_[0].data();
    double* part_rr = part_[1].data();
    double* part_rz = part_[2].data();
    krylov_state* st = state_.data();

    cg_init_kernel<<<nb, kBlock, 0, s_>>>(n, x, r, p, b, minv, part_rr, part_rz);
    cg_init_state_kernel<<<1, kBlock, 0, s_>>>(part_rr, part_rz, nb, st, criterion, tol, max_it);
    NEON_CHECK_HIP(hipGetLastError());

    solve_result res;
    int64_t launched = 0;
    run_batches(st, launched, [&] {
        for (int k = 0; k < poll_every; ++k, ++launched)
        {
            krylov_state* cur = st + (launched & 1);
            krylov_state* nxt = st + ((launched + 1) & 1);
            int const tok = prof ? prof->begin(s_, NEON_KERNEL_SPMV_FP64) : -1;
            A.apply64(s_, p, Ap, part_pAp, cur);
            if (prof) prof->end(s_, tok);
            totals_kernel<<<1, 1024, 0, s_>>>(part_pAp, nullptr, ns, cur);
            cg_update_xr_kernel<<<nb, kBlock, 0, s_>>>(n, r, Ap, minv, part_pAp, -1, cur, part_rr, part_rz);
            cg_update_p_kernel<<<nb, kBlock, 0, s_>>>(n, p, x, r, minv, part_pAp, part_rr, part_rz, nb, cur, nxt,
                                                       criterion, tol, max_it);
        }
    });
    res = finish(*host_state_, criterion);
    // SpMV launches that did work: Eigen's rule counts the converging iteration's SpMV
    // outside `iterations`; the CUDA rule counts every iteration.
    res.spmv = std::min<int64_t>(
        launched, res.iterations + (criterion == NEON_STOP_RELATIVE && host_state_->done == KRYLOV_CONVERGED ? 1 : 0));
    if (prof) prof->collect(res.spmv);
    return res;
}

// Jacobi-BiCGSTAB in FP64 (the unsymmetric selection) with every scalar on the device and the
// host polling the state like the CG loop (bicg_* kernels).  criterion
// NEON_STOP_ABSOLUTE_PRECONDITIONED follows biconjugate_gradient_stabilised_cuda::solve
// (biconjugate_gradient_stabilised_cuda.cpp:31-200); NEON_STOP_RELATIVE follows Eigen's BiCGSTAB
// with its restart rule (linear_solver.cpp:111-138).
solve_result krylov_engine::bicgstab64(const linear_operator& A, const double* b, double* x, const double* minv,
                                       int criterion, double tol, int32_t max_it, const krylov_work* work)
{
    int64_t const n = A.n;
    if (!work)
    {
        ensure(n);
        r0_.resize(n);
        v_.resize(n);
        s_vec_.resize(n);
        t_.resize(n);
        ph_.resize(n);
        sh_.resize(n);
    }
    double* r = work ? work->r : r_.data();
    double* p = work ? work->p : p_.data();
    double* r0 = work ? work->r0 : r0_.data();
    double* v = work ? work->v : v_.data();
    double* sv = work ? work->s : s_vec_.data();
    double* t = work ? work->t : t_.data();
    double* p_hat = work ? work->p_hat : ph_.data();
    double* s_hat = work ? work->s_hat : sh_.data();
    int const nb = vec_blocks(n);
    bool const cuda_rule = criterion == NEON_STOP_ABSOLUTE_PRECONDITIONED;
    double* part_r0v = part_[0].data();
    double* part_ss = part_[1].data();
    double* part_ts = part_[2].data();
    double* part_tt = part_[3].data();
    double* part_r0r = part_[4].data();
    double* part_rr = part_[5].data();
    krylov_state* st = state_.data();
    size_t const bytes = n * sizeof(double);
    // x = 0, r = r0 = b, p = v = 0 (x0 = 0 as both callers pass)
    NEON_CHECK_HIP(hipMemsetAsync(x, 0, bytes, s_));
    NEON_CHECK_HIP(hipMemcpyAsync(r, b, bytes, hipMemcpyDeviceToDevice, s_));
    NEON_CHECK_HIP(hipMemcpyAsync(r0, b, bytes, hipMemcpyDeviceToDevice, s_));
    NEON_CHECK_HIP(hipMemsetAsync(v, 0, bytes, s_));
    NEON_CHECK_HIP(hipMemsetAsync(p, 0, bytes, s_));
    double const bb = dot(n, b, b);
    krylov_state s0{};
    s0.rz = bb; // r0 . r = b . b
    s0.rr = bb;
    s0.rhs2 = bb;
    s0.aux0 = 1.0; // Eigen's rho_old, alpha, omega start at 1
    s0.alpha = 1.0;
    s0.omega = 1.0;
    s0.aux1 = bb; // Eigen's |r0|^2
    if (cuda_rule)
    {
        s0.aux0 = 0.0;
        s0.alpha = s0.omega = 0.0;
        if (!(std::sqrt(bb) > tol)) s0.done = KRYLOV_CONVERGED;
        else if (max_it <= 0) s0.done = KRYLOV_LIMIT;
    }
    else
    {
        s0.threshold = tol * tol * bb;
        if (bb == 0.0 || !(bb > s0.threshold)) s0.done = KRYLOV_CONVERGED;
        else if (max_it <= 0) s0.done = KRYLOV_LIMIT;
    }
    if (!std::isfinite(bb)) s0.done = KRYLOV_NONFINITE;
    *host_state_ = s0;
    NEON_CHECK_HIP(hipMemcpyAsync(st, host_state_, sizeof(krylov_state), hipMemcpyHostToDevice, s_));
    // the state is a single slot (decision kernels write it between the grid kernels that read
    // it), so run_batches reads st[launched & 1] = st[0] with launched kept even
    solve_result res;
    int64_t launched = 0;
    run_batches(st, launched, [&] {
        for (int k = 0; k < poll_every; ++k)
        {
            if (!cuda_rule)
            {
                bicg_restart_test_kernel<<<1, kBlock, 0, s_>>>(st);
                A.residual64(s_, b, x, r, st, 1);
                bicg_restart_kernel<<<nb, kBlock, 0, s_>>>(n, r, r0, part_ss, st);
                bicg_restart_finish_kernel<<<1, kBlock, 0, s_>>>(part_ss, nb, st);
            }
            bicg_p_kernel<<<nb, kBlock, 0, s_>>>(n, p, r, v, minv, p_hat, st, cuda_rule ? 1 : 0);
            int const tok = prof ? prof->begin(s_, NEON_KERNEL_SPMV_FP64) : -1;
            A.apply64(s_, p_hat, v, nullptr, st);
            if (prof) prof->end(s_, tok);
            bicg_dot_kernel<<<nb, kBlock, 0, s_>>>(n, r0, v, nullptr, nullptr, part_r0v, nullptr, st, BICG_GATE_NONE);
            bicg_s_kernel<<<nb, kBlock, 0, s_>>>(n, r, v, minv, sv, s_hat, part_r0v, nb, part_ss, st);
            bicg_alpha_kernel<<<1, kBlock, 0, s_>>>(part_r0v, part_ss, nb, st, cuda_rule ? 1 : 0, tol);
            A.apply64(s_, s_hat, t, nullptr, st, BICG_GATE_UNLESS_FLAG);
            bicg_dot_kernel<<<nb, kBlock, 0, s_>>>(n, t, sv, t, t, part_ts, part_tt, st, BICG_GATE_UNLESS_FLAG);
            bicg_update_kernel<<<nb, kBlock, 0, s_>>>(n, x, r, r0, p_hat, s_hat, sv, t, part_ts, part_tt, nb, part_r0r,
                                                      part_rr, st, cuda_rule ? 1 : 0);
            bicg_finish_kernel<<<1, kBlock, 0, s_>>>(part_ts, part_tt, part_r0r, part_rr, nb, st, cuda_rule ? 1 : 0,
                                                     tol, max_it);
        }
    });
    krylov_state const& f = *host_state_;
    res.iterations = f.it;
    res.spmv = 2ll * f.it + f.restarts;
    if (prof) prof->collect(f.it);
    if (cuda_rule)
    {
        res.error = std::sqrt(f.rz);
        switch (f.done)
        {
        case KRYLOV_CONVERGED: res.status = NEON_OK; break;
        case KRYLOV_BREAKDOWN: res.status = NEON_BREAKDOWN; break;
        case KRYLOV_NONFINITE: res.status = NEON_NAN_OR_INF; break;
        default: res.status = NEON_NONCONVERGED; break;
        }
    }
    else
    {
        res.error = f.rhs2 > 0.0 ? std::sqrt(f.rr / f.rhs2) : 0.0;
        res.status = f.done == KRYLOV_NONFINITE ? NEON_NAN_OR_INF
                                                : (f.done == KRYLOV_LIMIT ? NEON_NONCONVERGED : NEON_OK);
    }
    return res;
}

void krylov_engine::ensure_mixed(int64_t n)
{
    if (n == n_mixed_) return;
    r32_.resize(n);
    p32_.resize(n);
    Ap32_.resize(n);
    x32_.resize(n);
    minv32_.resize(n);
    r64_.resize(n);
    n_mixed_ = n;
}

// Mixed-precision Jacobi-PCG: FP32 recurrence and tangent values, FP64 dot products and
// FP64 reliable updates of r = b - K x (see mixed_* kernels).  Relative stopping rule on the
// true FP64 residual; iterations counted as Eigen counts them.
solve_result krylov_engine::cg_mixed(const linear_operator& A, const double* b, double* x, int criterion, double tol,
                                     int32_t max_it)
{
    if (criterion != NEON_STOP_RELATIVE)
        throw status_error(NEON_BAD_ARGUMENT, "mixed precision uses the relative stopping rule");
    if (!A.v32) throw status_error(NEON_BAD_ARGUMENT, "mixed precision needs FP32 tangent values");
    int64_t const n = A.n;
    ensure_mixed(n);
    int const nb = vec_blocks(n);
    int const ns = A.spmv_blocks();
    double* part_pAp = part_[0].data();
    double* part_rr = part_[1].data();
    double* part_rz = part_[2].data();
    double* part_rr_t = part_[3].data();
    double* part_rz_t = part_[4].data();
    krylov_state* st = state_.data();
    float* minv = minv32_.data();

    // Jacobi inverse (FP32) from the FP32 diagonal
    if (A.node_block)
        throw status_error(NEON_BAD_ARGUMENT, "cg_mixed: node-block operators take their Jacobi inverse from the mesh");
    csr_matrix m{int32_t(n), A.rp, A.ci};
    jacobi32_kernel<float><<<grid_for(n), kBlock, 0, s_>>>(m, A.v32, minv, 1);
    return cg_mixed_run(A, b, x, tol, max_it, nb, ns, part_pAp, part_rr, part_rz, part_rr_t, part_rz_t, st, minv,
                        nullptr, nullptr);
}

solve_result krylov_engine::cg_mixed_nb(const linear_operator& A, const double* b, double* x, const double* minv_in,
                                        double tol, int32_t max_it, const linear_operator* scaled)
{
    int64_t const n = A.n;
    ensure_mixed(n);
    convert_kernel<double, float><<<grid_for(n), kBlock, 0, s_>>>(n, minv_in, minv32_.data());
    if (scaled && (scaled->n != n || scaled->spmv_blocks() != A.spmv_blocks()))
        throw status_error(NEON_BAD_ARGUMENT, "cg_mixed_nb: the scaled operator differs in shape");
    return cg_mixed_run(A, b, x, tol, max_it, vec_blocks(n), A.spmv_blocks(), part_[0].data(), part_[1].data(),
                        part_[2].data(), part_[3].data(), part_[4].data(), state_.data(), minv32_.data(), scaled,
                        scaled ? minv_in : nullptr);
}

bool krylov_engine::all_positive(int64_t n, const double* v)
{
    int const nb = vec_blocks(n);
    nonpositive_count_kernel<<<nb, kBlock, 0, s_>>>(n, v, part_[0].data());
    reduce_to_kernel<<<1, kBlock, 0, s_>>>(part_[0].data(), nb, scalar_.data());
    NEON_CHECK_HIP(hipGetLastError());
    return read_scalar(scalar_.data()) == 0.0;
}

solve_result krylov_engine::cg_mixed_run(const linear_operator& A, const double* b, double* x, double tol,
                                         int32_t max_it, int nb, int ns, double* part_pAp, double* part_rr,
                                         double* part_rz, double* part_rr_t, double* part_rz_t, krylov_state* st,
                                         float* minv, const linear_operator* scaled_op,
                                         const double* minv64_scaled)
{
    int64_t const n = A.n;
    float* r32 = r32_.data();
    float* p32 = p32_.data();
    float* Ap32 = Ap32_.data();
    float* x32 = x32_.data();
    double* r64 = r64_.data();
    // scaled: the FP32 recurrence runs on K~ x~ = b~ (K~ = D^-1/2 K D^-1/2, the scaled copy) with
    // the identity preconditioner, the same iterates as Jacobi-PCG on K x = b in exact
    // arithmetic; x64 stays unscaled (the folds scale the steps back) and the FP64 residuals
    // r = b - K x64 use A, the unscaled FP32 values, so convergence is judged on K as without it
    bool const scaled = minv64_scaled != nullptr;
    linear_operator const& Arec = scaled ? *scaled_op : A;
    if (scaled)
        mixed_init_scaled_kernel<<<nb, kBlock, 0, s_>>>(n, b, x, r32, x32, p32, minv64_scaled, part_rr, part_rz);
    else
        mixed_init_kernel<<<nb, kBlock, 0, s_>>>(n, b, x, r32, x32, p32, minv, part_rr, part_rz);
    mixed_init_state_kernel<<<1, kBlock, 0, s_>>>(part_rr, part_rz, nb, st, tol, max_it);
    NEON_CHECK_HIP(hipGetLastError());
    // the gated true-residual product is a no-op in all but a few of its launches: a grid of at
    // most 2,048 workgroups striding over the slices keeps those launches short
    linear_operator Ar = A;
    if (Ar.node_block && (Ar.sell || Ar.sym)) Ar.blocks_override = 2048;
    int64_t launched = 0;
    run_batches(st, launched, [&] {
        for (int k = 0; k < poll_every; ++k, ++launched)
        {
            krylov_state* cur = st + (launched & 1);
            krylov_state* nxt = st + ((launched + 1) & 1);
            int const tok = prof ? prof->begin(s_, NEON_KERNEL_SPMV_FP32) : -1;
            Arec.apply32(s_, p32, Ap32, part_pAp, cur);
            if (prof) prof->end(s_, tok);
            totals_kernel<<<1, 1024, 0, s_>>>(part_pAp, nullptr, ns, cur);
            // the reliable-update / true-residual convergence test runs every refine_every
            // iterations
            int const decided = (launched + 1) % refine_every == 0 ? 1 : 0;
            if (!scaled)
                mixed_update_xr_kernel<false, true>
                    <<<nb, kBlock, 0, s_>>>(n, r32, Ap32, minv, part_pAp, -1, cur, part_rr, part_rz);
            else if (decided)
                mixed_update_xr_kernel<true, true>
                    <<<nb, kBlock, 0, s_>>>(n, r32, Ap32, minv, part_pAp, -1, cur, part_rr, part_rz);
            else
                mixed_update_xr_kernel<true, false>
                    <<<nb, kBlock, 0, s_>>>(n, r32, Ap32, nullptr, part_pAp, -1, cur, nullptr, part_rz);
            if (decided)
            {
                mixed_decide_kernel<<<1, kBlock, 0, s_>>>(part_rr, part_rz, nb, cur, reliable_delta);
                mixed_fold_kernel<<<nb, kBlock, 0, s_>>>(n, x, x32, p32, part_pAp, cur, 0, minv64_scaled);
                Ar.residual64(s_, b, x, r64, cur, 1);
                if (scaled)
                    mixed_refresh_kernel<true>
                        <<<nb, kBlock, 0, s_>>>(n, r64, r32, minv, minv64_scaled, cur, part_rr_t, part_rz_t);
                else
                    mixed_refresh_kernel<false>
                        <<<nb, kBlock, 0, s_>>>(n, r64, r32, minv, nullptr, cur, part_rr_t, part_rz_t);
            }
            if (scaled)
                mixed_update_p_kernel<true><<<nb, kBlock, 0, s_>>>(n, p32, x32, r32, nullptr, part_pAp, part_rr_t,
                                                                   part_rz_t, part_rr, part_rz, nb, nb, decided, cur,
                                                                   nxt, max_it);
            else
                mixed_update_p_kernel<false><<<nb, kBlock, 0, s_>>>(n, p32, x32, r32, minv, part_pAp, part_rr_t,
                                                                    part_rz_t, part_rr, part_rz, nb, nb, decided, cur,
                                                                    nxt, max_it);
        }
    });
    // fold whatever the recurrence holds since the last reliable update
    mixed_fold_kernel<<<nb, kBlock, 0, s_>>>(n, x, x32, p32, part_pAp, st + (launched & 1), 1, minv64_scaled);
    NEON_CHECK_HIP(hipGetLastError());
    NEON_CHECK_HIP(hipStreamSynchronize(s_));
    solve_result res = finish(*host_state_, NEON_STOP_RELATIVE);
    res.refinements = host_state_->restarts;
    int64_t const inner = res.iterations + (host_state_->done == KRYLOV_CONVERGED ? 1 : 0);
    res.spmv = inner + res.refinements;
    if (prof) prof->collect(inner);
    return res;
}

namespace
{
template <typename T>
__global__ __launch_bounds__(kBlock) void fill_kernel(int64_t n, T v, T* __restrict__ x)
{
    int64_t const i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) x[i] = v;
}
} // namespace

void fill_ones(hipStream_t s, int64_t n, double* x)
{
    fill_kernel<double><<<grid_for(n), kBlock, 0, s>>>(n, 1.0, x);
    NEON_CHECK_HIP(hipGetLastError());
}

void fill_ones(hipStream_t s, int64_t n, float* x)
{
    fill_kernel<float><<<grid_for(n), kBlock, 0, s>>>(n, 1.0f, x);
    NEON_CHECK_HIP(hipGetLastError());
}

void convert_to_float(hipStream_t s, int64_t n, const double* x, float* y)
{
    if (n <= 0) return;
    convert_kernel<double, float><<<grid_for(n), kBlock, 0, s>>>(n, x, y);
    NEON_CHECK_HIP(hipGetLastError());
}

// ==================================================================== mesh
void upload_hex8_tables()
{
    // __constant__ memory is per device: upload once for each device used
    static uint64_t done = 0;
    int dev = 0;
    NEON_CHECK_HIP(hipGetDevice(&dev));
    uint64_t const bit = dev < 64 ? uint64_t(1) << dev : 0;
    if (done & bit) return;
    hex8_tables t{};
    static const double node[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                      {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    double const qp = 1.0 / std::sqrt(3.0);
    for (int l = 0; l < 8; ++l)
    {
        double const xi = node[l][0] * qp, eta = node[l][1] * qp, zeta = node[l][2] * qp;
        t.w[l] = 1.0;
        for (int a = 0; a < 8; ++a)
        {
            double const xa = node[a][0], ya = node[a][1], za = node[a][2];
            t.dN[l][a][0] = 1.0 / 8.0 * xa * (1.0 + ya * eta) * (1.0 + za * zeta);
            t.dN[l][a][1] = 1.0 / 8.0 * (1.0 + xa * xi) * ya * (1.0 + za * zeta);
            t.dN[l][a][2] = 1.0 / 8.0 * (1.0 + xa * xi) * (1.0 + ya * eta) * za;
            t.N[l][a] = 1.0 / 8.0 * (1.0 + xa * xi) * (1.0 + ya * eta) * (1.0 + za * zeta);
        }
    }
    NEON_CHECK_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_hex8), &t, sizeof(t)));
    done |= bit;
}

// The Bazant-Oh (1986) unit-sphere rules of unit_sphere_quadrature.cpp:10-90 with their published
// 12-digit direction cosines and weights, generated from the symmetry orbits in the reference's
// order: axes, (v, +-v, 0) pairs, (a, b, 0) and (a, a, b) families.
static void sphere_rules(sphere_tables& t)
{
    auto build = [&](int k, auto&& orbits) {
        int& n = t.n[k];
        n = 0;
        auto add = [&](double w, double x, double y, double z) {
            t.w[k][n] = w;
            t.r[k][n][0] = x;
            t.r[k][n][1] = y;
            t.r[k][n][2] = z;
            ++n;
        };
        orbits(add);
    };
    static const int pair[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    auto axes = [](auto& add, double w) {
        add(w, 1.0, 0.0, 0.0);
        add(w, 0.0, 1.0, 0.0);
        add(w, 0.0, 0.0, 1.0);
    };
    auto pairs = [](auto& add, double w, double v) {
        for (auto const& p : pair)
            for (double sg : {1.0, -1.0})
            {
                double r[3] = {0, 0, 0};
                r[p[0]] = v;
                r[p[1]] = sg * v;
                add(w, r[0], r[1], r[2]);
            }
    };
    auto ab0 = [](auto& add, double w, double a, double b) {
        for (auto const& p : pair)
            for (int k = 0; k < 4; ++k)
            {
                double r[3] = {0, 0, 0};
                r[p[0]] = k < 2 ? a : b;
                r[p[1]] = (k < 2 ? b : a) * (k % 2 ? -1.0 : 1.0);
                add(w, r[0], r[1], r[2]);
            }
    };
    auto aab = [](auto& add, double w, double a, double b) {
        for (int pos : {2, 1, 0})
            for (int k = 0; k < 4; ++k)
            {
                double r[3] = {a, a, a};
                r[pos] = b;
                if (k >= 2) r[1] = -r[1];
                if (k % 2) r[2] = -r[2];
                add(w, r[0], r[1], r[2]);
            }
    };
    build(0, [&](auto& add) { // BO21
        axes(add, 0.0265214244093 * 2.0);
        pairs(add, 0.0199301476312 * 2.0, 0.707106781187);
        aab(add, 0.0250712367487 * 2.0, 0.387907304067, 0.836095596749);
    });
    build(1, [&](auto& add) { // BO33
        axes(add, 2.0 * 0.00985353993433);
        pairs(add, 2.0 * 0.0162969685886, 0.707106781187);
        ab0(add, 2.0 * 0.0134788844008, 0.933898956394, 0.357537045978);
        aab(add, 2.0 * 0.0175759129880, 0.437263676092, 0.785875915868);
    });
}

static void upload_sphere_tables()
{
    static uint64_t done = 0;
    int dev = 0;
    NEON_CHECK_HIP(hipGetDevice(&dev));
    uint64_t const bit = dev < 64 ? uint64_t(1) << dev : 0;
    if (done & bit) return;
    sphere_tables t{};
    sphere_rules(t);
    NEON_CHECK_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_sphere), &t, sizeof(t)));
    done |= bit;
}

device_mesh::device_mesh(hipStream_t s, int32_t n_nodes, const double* X, int32_t n_elem, const int32_t* conn,
                         int32_t model, const double* params, int32_t tangent_mode)
    : s_(s), N_(n_nodes), E_(n_elem), model_(model), tangent_mode_(tangent_mode)
{
    if (n_nodes <= 0 || n_elem <= 0 || !X || !conn || !params)
        throw status_error(NEON_BAD_ARGUMENT, "empty mesh or null pointer");
    if ((model & ~(0xff | NEON_SPHERE_BO33)) != 0 || (model & 0xff) > NEON_MODEL_J2_DAMAGE)
        throw status_error(NEON_BAD_ARGUMENT, "unknown model");
    if ((model & 0xff) == NEON_MODEL_MICROSPHERE_AGEING && tangent_mode == NEON_TANGENT_CONSISTENT)
        throw status_error(NEON_BAD_ARGUMENT, "the ageing microsphere model has neon's tangent only");
    if (model == NEON_MODEL_J2_DAMAGE && tangent_mode == NEON_TANGENT_CONSISTENT)
        throw status_error(NEON_BAD_ARGUMENT, "the J2 damage model has neon's tangent only");
    if ((model & NEON_SPHERE_BO33) && !is_microsphere(model))
        throw status_error(NEON_BAD_ARGUMENT, "a unit-sphere rule is only for the microsphere models");
    if (tangent_mode != NEON_TANGENT_REFERENCE && tangent_mode != NEON_TANGENT_CONSISTENT)
        throw status_error(NEON_BAD_ARGUMENT, "unknown tangent mode");
    for (int64_t i = 0; i < 8ll * n_elem; ++i)
        if (conn[i] < 0 || conn[i] >= n_nodes) throw status_error(NEON_BAD_ARGUMENT, "connectivity out of range");
    upload_hex8_tables();
    if (is_microsphere(model)) upload_sphere_tables();

    // node -> element incidence
    std::vector<int32_t> inc_ptr(size_t(N_) + 1, 0), inc;
    for (int64_t i = 0; i < 8ll * E_; ++i) inc_ptr[conn[i] + 1]++;
    for (int32_t nd = 0; nd < N_; ++nd) inc_ptr[nd + 1] += inc_ptr[nd];
    inc.resize(8ull * E_);
    {
        std::vector<int32_t> fill(inc_ptr.begin(), inc_ptr.end() - 1);
        for (int64_t e = 0; e < E_; ++e)
            for (int a = 0; a < 8; ++a) inc[fill[conn[8 * e + a]]++] = int32_t(e);
    }
    // node adjacency (sorted unique) -> block pattern, as compute_sparsity_pattern
    // (src/assembler/sparsity_pattern.hpp:18-48) yields for 3 dofs per node
    std::vector<int32_t> deg(N_, 0);
#pragma omp parallel
    {
        std::vector<int32_t> buf;
#pragma omp for schedule(dynamic, 4096)
        for (int32_t nd = 0; nd < N_; ++nd)
        {
            buf.clear();
            for (int32_t k = inc_ptr[nd]; k < inc_ptr[nd + 1]; ++k)
                for (int a = 0; a < 8; ++a) buf.push_back(conn[8ll * inc[k] + a]);
            std::sort(buf.begin(), buf.end());
            deg[nd] = int32_t(std::unique(buf.begin(), buf.end()) - buf.begin());
        }
    }
    bptr_h_.assign(size_t(N_) + 1, 0);
    for (int32_t nd = 0; nd < N_; ++nd)
    {
        bptr_h_[nd + 1] = bptr_h_[nd] + deg[nd];
        max_degree_ = std::max<int32_t>(max_degree_, deg[nd]);
    }
    int64_t const nnzb = bptr_h_[N_];
    nnz_ = 9 * nnzb;
    if (nnz_ > int64_t(INT32_MAX)) throw status_error(NEON_BAD_ARGUMENT, "nnz exceeds the int32 CSR index range");
    bcol_h_.resize(nnzb);
    std::vector<int32_t> self(N_);
#pragma omp parallel
    {
        std::vector<int32_t> buf;
#pragma omp for schedule(dynamic, 4096)
        for (int32_t nd = 0; nd < N_; ++nd)
        {
            buf.clear();
            for (int32_t k = inc_ptr[nd]; k < inc_ptr[nd + 1]; ++k)
                for (int a = 0; a < 8; ++a) buf.push_back(conn[8ll * inc[k] + a]);
            std::sort(buf.begin(), buf.end());
            buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
            std::copy(buf.begin(), buf.end(), bcol_h_.begin() + bptr_h_[nd]);
            self[nd] = int32_t(std::lower_bound(buf.begin(), buf.end(), nd) - buf.begin());
        }
    }
    // element slot table: slot of node b in node a's block row
    std::vector<uint8_t> slots(64ull * E_);
    bool overflow = false;
#pragma omp parallel for schedule(static) reduction(|| : overflow)
    for (int64_t e = 0; e < E_; ++e)
        for (int a = 0; a < 8; ++a)
        {
            int32_t const na = conn[8 * e + a];
            const int32_t* beg = bcol_h_.data() + bptr_h_[na];
            const int32_t* end = bcol_h_.data() + bptr_h_[na + 1];
            for (int b = 0; b < 8; ++b)
            {
                int64_t const s = std::lower_bound(beg, end, conn[8 * e + b]) - beg;
                if (s > 255) overflow = true;
                slots[64 * e + 8 * a + b] = uint8_t(s);
            }
        }
    if (overflow) throw status_error(NEON_BAD_ARGUMENT, "node degree above 255");
    // greedy element colouring: elements of one colour share no node
    std::vector<int8_t> colour(E_, -1);
    int ncol = 0;
    for (int64_t e = 0; e < E_; ++e)
    {
        uint64_t used = 0;
        for (int a = 0; a < 8; ++a)
        {
            int32_t const nd = conn[8 * e + a];
            for (int32_t k = inc_ptr[nd]; k < inc_ptr[nd + 1]; ++k)
            {
                int8_t const c = colour[inc[k]];
                if (c >= 0) used |= 1ull << c;
            }
        }
        int c = 0;
        while (c < 64 && (used >> c & 1ull)) ++c;
        if (c == 64) throw status_error(NEON_BAD_ARGUMENT, "element colouring needs more than 64 colours");
        colour[e] = int8_t(c);
        ncol = std::max(ncol, c + 1);
    }
    colour_offsets_.assign(ncol + 1, 0);
    for (int64_t e = 0; e < E_; ++e) colour_offsets_[colour[e] + 1]++;
    for (int c = 0; c < ncol; ++c) colour_offsets_[c + 1] += colour_offsets_[c];
    std::vector<int32_t> elems(E_);
    {
        std::vector<int32_t> fill(colour_offsets_.begin(), colour_offsets_.end() - 1);
        for (int64_t e = 0; e < E_; ++e) elems[fill[colour[e]]++] = int32_t(e);
    }

    X_.resize(3ull * N_);
    u_.resize(3ull * N_);
    fint_.resize(3ull * N_);
    params_.resize(4ull * E_);
    conn_.resize(8ull * E_);
    elems_.resize(E_);
    bptr_.resize(size_t(N_) + 1);
    bcol_.resize(nnzb);
    self_slot_.resize(N_);
    slots_.resize(64ull * E_);
    NEON_CHECK_HIP(hipMemcpyAsync(X_.data(), X, 3ull * N_ * sizeof(double), hipMemcpyHostToDevice, s_));
    NEON_CHECK_HIP(hipMemsetAsync(u_.data(), 0, 3ull * N_ * sizeof(double), s_));
    NEON_CHECK_HIP(hipMemcpyAsync(params_.data(), params, 4ull * E_ * sizeof(double), hipMemcpyHostToDevice, s_));
    NEON_CHECK_HIP(hipMemcpyAsync(conn_.data(), conn, 8ull * E_ * sizeof(int32_t), hipMemcpyHostToDevice, s_));
    NEON_CHECK_HIP(hipMemcpyAsync(elems_.data(), elems.data(), size_t(E_) * sizeof(int32_t), hipMemcpyHostToDevice, s_));
    NEON_CHECK_HIP(hipMemcpyAsync(bptr_.data(), bptr_h_.data(), (size_t(N_) + 1) * sizeof(int32_t), hipMemcpyHostToDevice, s_));
    NEON_CHECK_HIP(hipMemcpyAsync(bcol_.data(), bcol_h_.data(), size_t(nnzb) * sizeof(int32_t), hipMemcpyHostToDevice, s_));
    build_sliced_pattern();
    build_symmetric_pattern();
    NEON_CHECK_HIP(hipMemcpyAsync(self_slot_.data(), self.data(), size_t(N_) * sizeof(int32_t), hipMemcpyHostToDevice, s_));
    NEON_CHECK_HIP(hipMemcpyAsync(slots_.data(), slots.data(), slots.size(), hipMemcpyHostToDevice, s_));

    int64_t const Q = 8ll * E_;
    int64_t const ncoef = model_ == NEON_MODEL_J2_DAMAGE ? kGeneralCoefs : is_microsphere(model_) ? kFullCoefs : 9;
    sigma_.resize(6 * Q);
    coef_.resize(ncoef * Q);
    detF_.resize(Q);
    NEON_CHECK_HIP(hipMemsetAsync(sigma_.data(), 0, 6 * Q * sizeof(double), s_));
    NEON_CHECK_HIP(hipMemsetAsync(coef_.data(), 0, ncoef * Q * sizeof(double), s_));
    bool const j2_family = model_ == NEON_MODEL_J2 || model_ == NEON_MODEL_J2_DAMAGE;
    if (j2_family || model_ == NEON_MODEL_LINEAR_ELASTIC)
    {
        vm_.resize(Q);
        NEON_CHECK_HIP(hipMemsetAsync(vm_.data(), 0, Q * sizeof(double), s_));
    }
    if (model_ == NEON_MODEL_J2_DAMAGE)
    {
        // back stress 6, damage 1, energy release rate 1 per QP and the committed back stress
        // and damage; the tangent starts at C_e (isotropic_linear_elasticity.cpp:22-23)
        dmg_state_.resize(8 * Q);
        dmg_state_old_.resize(7 * Q);
        NEON_CHECK_HIP(hipMemsetAsync(dmg_state_.data(), 0, 8 * Q * sizeof(double), s_));
        NEON_CHECK_HIP(hipMemsetAsync(dmg_state_old_.data(), 0, 7 * Q * sizeof(double), s_));
        std::vector<double> c(36 * size_t(Q));
        for (int64_t q = 0; q < Q; ++q)
        {
            double const* p = params + 4 * (q / 8);
            double const lam = p[0] - 2.0 * p[1] / 3.0;
            for (int i = 0; i < 6; ++i)
                for (int j = 0; j < 6; ++j)
                    c[size_t(i * 6 + j) * Q + q] = i < 3 && j < 3 ? lam + (i == j ? 2.0 * p[1] : 0.0)
                                                                  : (i == j ? p[1] : 0.0);
        }
        NEON_CHECK_HIP(hipMemcpyAsync(coef_.data(), c.data(), c.size() * sizeof(double), hipMemcpyHostToDevice, s_));
        NEON_CHECK_HIP(hipStreamSynchronize(s_));
    }
    if (j2_family)
    {
        eps_p_.resize(6 * Q);
        ebar_.resize(Q);
        eps_p_old_.resize(6 * Q);
        ebar_old_.resize(Q);
        NEON_CHECK_HIP(hipMemsetAsync(eps_p_.data(), 0, 6 * Q * sizeof(double), s_));
        NEON_CHECK_HIP(hipMemsetAsync(ebar_.data(), 0, Q * sizeof(double), s_));
        NEON_CHECK_HIP(hipMemsetAsync(eps_p_old_.data(), 0, 6 * Q * sizeof(double), s_));
        NEON_CHECK_HIP(hipMemsetAsync(ebar_old_.data(), 0, Q * sizeof(double), s_));
    }
    if ((model_ & 0xff) == NEON_MODEL_MICROSPHERE_AGEING)
    {
        // gaussian_ageing_affine_microsphere.cpp:17-50: active shear modulus G, inactive
        // modulus (params[3]), reduction factor 1, active segments N, inactive segments 0,
        // secondary stress 0
        std::vector<double> a(11 * size_t(Q), 0.0);
        for (int64_t q = 0; q < Q; ++q)
        {
            double const* p = params + 4 * (q / 8);
            a[q] = p[1];
            a[Q + q] = p[3];
            a[2 * Q + q] = 1.0;
            a[3 * Q + q] = p[2];
        }
        age_.resize(11 * Q);
        age_old_.resize(11 * Q);
        NEON_CHECK_HIP(hipMemcpyAsync(age_.data(), a.data(), a.size() * sizeof(double), hipMemcpyHostToDevice, s_));
        NEON_CHECK_HIP(hipMemcpyAsync(age_old_.data(), a.data(), a.size() * sizeof(double), hipMemcpyHostToDevice, s_));
        NEON_CHECK_HIP(hipStreamSynchronize(s_));
    }
    status_.resize(1);
    NEON_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&status_host_), sizeof(uint32_t)));
    NEON_CHECK_HIP(hipStreamSynchronize(s_));
}

void device_mesh::set_parameters(const double* params)
{
    NEON_CHECK_HIP(hipMemcpyAsync(params_.data(), params, 4ull * E_ * sizeof(double), hipMemcpyHostToDevice, s_));
    NEON_CHECK_HIP(hipStreamSynchronize(s_));
}

void device_mesh::set_displacement_host(const double* u)
{
    NEON_CHECK_HIP(hipMemcpyAsync(u_.data(), u, 3ull * N_ * sizeof(double), hipMemcpyHostToDevice, s_));
    NEON_CHECK_HIP(hipStreamSynchronize(s_));
}

void device_mesh::latin_reset()
{
    int64_t const Q = 8ll * E_;
    sigma_old_.resize(6 * Q);
    NEON_CHECK_HIP(hipMemsetAsync(sigma_old_.data(), 0, 6 * Q * sizeof(double), s_));
}

void device_mesh::latin_commit()
{
    int64_t const Q = 8ll * E_;
    history_copy_kernel<<<grid_for(6 * Q), kBlock, 0, s_>>>(6 * Q, sigma_.data(), sigma_old_.data());
    NEON_CHECK_HIP(hipGetLastError());
}

void device_mesh::latin_force(const double* u_dev, const double* u_old_dev, double s, double* f_dev)
{
    int64_t const Q = 8ll * E_;
    if (sigma_old_.size() != size_t(6 * Q)) throw status_error(NEON_BAD_ARGUMENT, "latin_reset was not called");
    NEON_CHECK_HIP(hipMemsetAsync(f_dev, 0, 3ull * N_ * sizeof(double), s_));
    if (model_ == NEON_MODEL_J2_DAMAGE)
        throw status_error(NEON_BAD_ARGUMENT, "LATIN iterations are not available for the J2 damage model");
    bool const full = is_microsphere(model_);
    for (int c = 0; c < colours(); ++c)
    {
        int32_t const off = colour_offsets_[c];
        int32_t const cnt = colour_offsets_[c + 1] - off;
        if (cnt == 0) continue;
        int const blocks = int(std::min<int64_t>((cnt + 63) / 64, 4096));
        const int32_t* el = elems_.data() + off;
        if (full)
            latin_force_kernel<true><<<blocks, 64, 0, s_>>>(el, cnt, conn_.data(), X_.data(), u_dev, u_old_dev,
                                                            sigma_.data(), sigma_old_.data(), coef_.data(), Q, s, f_dev);
        else
            latin_force_kernel<false><<<blocks, 64, 0, s_>>>(el, cnt, conn_.data(), X_.data(), u_dev, u_old_dev,
                                                             sigma_.data(), sigma_old_.data(), coef_.data(), Q, s, f_dev);
        NEON_CHECK_HIP(hipGetLastError());
    }
}

void device_mesh::set_ageing(double scission, double recombination)
{
    if ((model_ & 0xff) != NEON_MODEL_MICROSPHERE_AGEING)
        throw status_error(NEON_BAD_ARGUMENT, "the mesh's model is not the ageing microsphere");
    if (!(scission >= 0.0) || !(recombination >= 0.0))
        throw status_error(NEON_BAD_ARGUMENT, "scission and recombination probabilities must be non-negative");
    scission_ = scission;
    recombination_ = recombination;
}

void device_mesh::download_ageing_state(double* state)
{
    if ((model_ & 0xff) != NEON_MODEL_MICROSPHERE_AGEING)
        throw status_error(NEON_BAD_ARGUMENT, "the mesh's model is not the ageing microsphere");
    int64_t const Q = 8ll * E_;
    std::vector<double> a(11 * size_t(Q));
    NEON_CHECK_HIP(hipMemcpyAsync(a.data(), age_.data(), a.size() * sizeof(double), hipMemcpyDeviceToHost, s_));
    NEON_CHECK_HIP(hipStreamSynchronize(s_));
    for (int64_t q = 0; q < Q; ++q)
        for (int k = 0; k < 11; ++k) state[11 * q + k] = a[k * Q + q];
}

void device_mesh::set_damage(const double* c)
{
    if (model_ != NEON_MODEL_J2_DAMAGE) throw status_error(NEON_BAD_ARGUMENT, "the mesh's model is not J2 with damage");
    for (int k = 0; k < 6; ++k)
        if (!(c[k] > 0.0)) throw status_error(NEON_BAD_ARGUMENT, "damage constants must be positive");
    std::copy(c, c + 6, damage_);
}

void device_mesh::download_damage_state(double* back_stress, double* damage, double* energy)
{
    if (model_ != NEON_MODEL_J2_DAMAGE) throw status_error(NEON_BAD_ARGUMENT, "the mesh's model is not J2 with damage");
    int64_t const Q = 8ll * E_;
    std::vector<double> a(8 * size_t(Q));
    NEON_CHECK_HIP(hipMemcpyAsync(a.data(), dmg_state_.data(), a.size() * sizeof(double), hipMemcpyDeviceToHost, s_));
    NEON_CHECK_HIP(hipStreamSynchronize(s_));
    for (int64_t q = 0; q < Q; ++q)
    {
        if (back_stress)
        {
            // tensor components (11, 22, 33, 23, 13, 12) -> row-major 3 x 3
            double const* v = a.data();
            double* o = back_stress + 9 * q;
            o[0] = v[q];
            o[4] = v[Q + q];
            o[8] = v[2 * Q + q];
            o[5] = o[7] = v[3 * Q + q];
            o[2] = o[6] = v[4 * Q + q];
            o[1] = o[3] = v[5 * Q + q];
        }
        if (damage) damage[q] = a[6 * Q + q];
        if (energy) energy[q] = a[7 * Q + q];
    }
}

int device_mesh::update_internal_variables(const double* u_dev, double dt)
{
    int64_t const Q = 8ll * E_;
    qp_state st{sigma_.data(), coef_.data(), detF_.data(), eps_p_.data(), ebar_.data(), vm_.data(), Q,
                eps_p_old_.data(), ebar_old_.data(), age_.data(), age_old_.data()};
    NEON_CHECK_HIP(hipMemsetAsync(status_.data(), 0, sizeof(uint32_t), s_));
    element_params prm{params_.data()};
    int const rule = (model_ & NEON_SPHERE_BO33) ? 1 : 0;
    // the ageing model keeps the last non-zero step for its secondary network
    // (gaussian_ageing_affine_microsphere.cpp:84-103)
    if (dt != 0.0) last_dt_ = dt;
    ageing_args const ag{scission_, recombination_, dt, last_dt_};
#define NEON_UPDATE(M)                                                                                             \
    update_state_kernel<M><<<grid_for(Q), kBlock, 0, s_>>>(E_, conn_.data(), X_.data(), u_dev, prm, st, tangent_mode_, \
                                                          rule, ag, status_.data())
    switch (model_ & 0xff)
    {
    case NEON_MODEL_J2_DAMAGE:
        damage_update_kernel<<<grid_for(Q), kBlock, 0, s_>>>(
            E_, conn_.data(), X_.data(), u_dev, prm, st,
            damage_state{dmg_state_.data(), dmg_state_.data() + 6 * Q, dmg_state_.data() + 7 * Q},
            damage_constants{damage_[0], damage_[1], damage_[2], damage_[3], damage_[4], damage_[5]}, dt,
            status_.data());
        break;
    case NEON_MODEL_NEOHOOKE: NEON_UPDATE(NEON_MODEL_NEOHOOKE); break;
    case NEON_MODEL_J2: NEON_UPDATE(NEON_MODEL_J2); break;
    case NEON_MODEL_LINEAR_ELASTIC: NEON_UPDATE(NEON_MODEL_LINEAR_ELASTIC); break;
    case NEON_MODEL_MICROSPHERE_GAUSSIAN: NEON_UPDATE(NEON_MODEL_MICROSPHERE_GAUSSIAN); break;
    case NEON_MODEL_MICROSPHERE_LANGEVIN: NEON_UPDATE(NEON_MODEL_MICROSPHERE_LANGEVIN); break;
    case NEON_MODEL_MICROSPHERE_NONAFFINE: NEON_UPDATE(NEON_MODEL_MICROSPHERE_NONAFFINE); break;
    default: NEON_UPDATE(NEON_MODEL_MICROSPHERE_AGEING); break;
    }
#undef NEON_UPDATE
    NEON_CHECK_HIP(hipGetLastError());
    NEON_CHECK_HIP(hipMemcpyAsync(status_host_, status_.data(), sizeof(uint32_t), hipMemcpyDeviceToHost, s_));
    NEON_CHECK_HIP(hipStreamSynchronize(s_));
    uint32_t const f = *status_host_;
    // submesh.cpp:239-269 checks det F first, then the constitutive update, then FE_INVALID
    if (f & STATUS_NEGATIVE_JACOBIAN) return NEON_NEGATIVE_JACOBIAN;
    if (f & STATUS_RETURN_MAP) return NEON_RETURN_MAP_FAILED;
    if (f & STATUS_NAN) return NEON_NAN_OR_INF;
    return NEON_OK;
}

// internal_variables::commit / revert (src/constitutive/internal_variables.hpp:203-214).  Only
// the J2 history and the ageing network are path dependent; every other variable is recomputed
// from u before use.
void device_mesh::save_internal_variables(bool converged)
{
    int64_t const Q = 8ll * E_;
    if ((model_ & 0xff) == NEON_MODEL_MICROSPHERE_AGEING)
    {
        if (converged)
            history_copy_kernel<<<grid_for(11 * Q), kBlock, 0, s_>>>(11 * Q, age_.data(), age_old_.data());
        else
            history_copy_kernel<<<grid_for(11 * Q), kBlock, 0, s_>>>(11 * Q, age_old_.data(), age_.data());
        NEON_CHECK_HIP(hipGetLastError());
        return;
    }
    if (model_ != NEON_MODEL_J2 && model_ != NEON_MODEL_J2_DAMAGE) return;
    if (converged)
    {
        history_copy_kernel<<<grid_for(6 * Q), kBlock, 0, s_>>>(6 * Q, eps_p_.data(), eps_p_old_.data());
        history_copy_kernel<<<grid_for(Q), kBlock, 0, s_>>>(Q, ebar_.data(), ebar_old_.data());
    }
    else
    {
        history_copy_kernel<<<grid_for(6 * Q), kBlock, 0, s_>>>(6 * Q, eps_p_old_.data(), eps_p_.data());
        history_copy_kernel<<<grid_for(Q), kBlock, 0, s_>>>(Q, ebar_old_.data(), ebar_.data());
    }
    if (model_ == NEON_MODEL_J2_DAMAGE)
    {
        // back stress and damage (the energy release rate is recomputed before every use)
        if (converged)
            history_copy_kernel<<<grid_for(7 * Q), kBlock, 0, s_>>>(7 * Q, dmg_state_.data(), dmg_state_old_.data());
        else
            history_copy_kernel<<<grid_for(7 * Q), kBlock, 0, s_>>>(7 * Q, dmg_state_old_.data(), dmg_state_.data());
    }
    NEON_CHECK_HIP(hipGetLastError());
}

void device_mesh::assemble(const double* u_dev, int precision)
{
    int64_t const Q = 8ll * E_;
    qp_state st{sigma_.data(), coef_.data(), detF_.data(), eps_p_.data(), ebar_.data(), vm_.data(), Q,
                eps_p_old_.data(), ebar_old_.data(), age_.data(), age_old_.data()};
    bool const finite = is_finite_model(model_); // is_finite_deformation()
    bool const general = model_ == NEON_MODEL_J2_DAMAGE;
    bool const full = is_microsphere(model_);
    // neon integrates the small-strain models on the current configuration (submesh.cpp:68,84);
    // the consistent mode integrates them on X
    if (!finite && tangent_mode_ == NEON_TANGENT_CONSISTENT) u_dev = nullptr;
    hipEvent_t e0, e1;
    NEON_CHECK_HIP(hipEventCreate(&e0));
    NEON_CHECK_HIP(hipEventCreate(&e1));
    NEON_CHECK_HIP(hipEventRecord(e0, s_));
    NEON_CHECK_HIP(hipMemsetAsync(fint_.data(), 0, 3ull * N_ * sizeof(double), s_));
    if (precision == NEON_PRECISION_FP64)
    {
        k64_.resize(nnz_ + kValuePad);
        NEON_CHECK_HIP(hipMemsetAsync(k64_.data(), 0, (nnz_ + kValuePad) * sizeof(double), s_));
    }
    else
    {
        k32_.resize(nnz_ + kValuePad);
        NEON_CHECK_HIP(hipMemsetAsync(k32_.data(), 0, (nnz_ + kValuePad) * sizeof(float), s_));
    }
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    for (int c = 0; c < colours(); ++c)
    {
        int32_t const off = colour_offsets_[c];
        int32_t const cnt = colour_offsets_[c + 1] - off;
        if (cnt == 0) continue;
        int const blocks = int(std::min<int64_t>(cnt, int64_t(cus) * assembly_blocks_per_cu));
        const int32_t* el = elems_.data() + off;
#define NEON_ASSEMBLE(KT, FIN, FULL, K)                                                                       \
    assemble_kernel<KT, FIN, FULL><<<blocks, 64, 0, s_>>>(el, cnt, conn_.data(), X_.data(), u_dev, st, slots_.data(), \
                                                          bptr_.data(), K, fint_.data())
        if (general)
        {
            if (precision == NEON_PRECISION_FP64)
                assemble_kernel<double, false, true, true><<<blocks, 64, 0, s_>>>(
                    el, cnt, conn_.data(), X_.data(), u_dev, st, slots_.data(), bptr_.data(), k64_.data(), fint_.data());
            else
                assemble_kernel<float, false, true, true><<<blocks, 64, 0, s_>>>(
                    el, cnt, conn_.data(), X_.data(), u_dev, st, slots_.data(), bptr_.data(), k32_.data(), fint_.data());
        }
        else if (precision == NEON_PRECISION_FP64)
        {
            if (full)
                NEON_ASSEMBLE(double, true, true, k64_.data());
            else if (finite)
                NEON_ASSEMBLE(double, true, false, k64_.data());
            else
                NEON_ASSEMBLE(double, false, false, k64_.data());
        }
        else
        {
            if (full)
                NEON_ASSEMBLE(float, true, true, k32_.data());
            else if (finite)
                NEON_ASSEMBLE(float, true, false, k32_.data());
            else
                NEON_ASSEMBLE(float, false, false, k32_.data());
        }
#undef NEON_ASSEMBLE
    }
    NEON_CHECK_HIP(hipGetLastError());
    NEON_CHECK_HIP(hipEventRecord(e1, s_));
    NEON_CHECK_HIP(hipEventSynchronize(e1));
    NEON_CHECK_HIP(hipEventElapsedTime(&last_assemble_ms, e0, e1));
    if (prof) prof->add(NEON_KERNEL_ASSEMBLY, last_assemble_ms);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    k64_valid_ = precision == NEON_PRECISION_FP64;
    k32_valid_ = precision != NEON_PRECISION_FP64;
    s64_fresh_ = s32_fresh_ = y64_fresh_ = y32_fresh_ = y32s_fresh_ = false;
}

void device_mesh::assemble_mass(const double* density)
{
    device_array<double> rho;
    rho.resize(E_);
    NEON_CHECK_HIP(hipMemcpyAsync(rho.data(), density, size_t(E_) * sizeof(double), hipMemcpyHostToDevice, s_));
    m64_.resize(nnz_ + kValuePad);
    NEON_CHECK_HIP(hipMemsetAsync(m64_.data(), 0, (nnz_ + kValuePad) * sizeof(double), s_));
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    for (int c = 0; c < colours(); ++c)
    {
        int32_t const off = colour_offsets_[c];
        int32_t const cnt = colour_offsets_[c + 1] - off;
        if (cnt == 0) continue;
        int const blocks = int(std::min<int64_t>(cnt, int64_t(cus) * assembly_blocks_per_cu));
        assemble_mass_kernel<<<blocks, 64, 0, s_>>>(elems_.data() + off, cnt, conn_.data(), X_.data(), rho.data(),
                                                    slots_.data(), bptr_.data(), m64_.data());
    }
    NEON_CHECK_HIP(hipGetLastError());
    NEON_CHECK_HIP(hipStreamSynchronize(s_));
    m64_valid_ = true;
}

void device_mesh::enforce_dirichlet_mass(int32_t n_fixed, const int32_t* fixed_dev)
{
    if (n_fixed <= 0) return;
    if (!m64_valid_) throw status_error(NEON_BAD_ARGUMENT, "no mass matrix assembled");
    dirichlet_kernel<double><<<grid_for(n_fixed), kBlock, 0, s_>>>(n_fixed, fixed_dev, bptr_.data(), bcol_.data(),
                                                                     m64_.data(), nullptr);
    NEON_CHECK_HIP(hipGetLastError());
}

linear_operator device_mesh::op_mass() const
{
    if (!m64_valid_) throw status_error(NEON_BAD_ARGUMENT, "no mass matrix assembled");
    linear_operator A;
    A.n = 3ll * N_;
    A.node_block = true;
    A.nodes = N_;
    A.bptr = bptr_.data();
    A.bcol = bcol_.data();
    A.v64 = m64_.data();
    A.nb_lanes = 32;
    return A;
}

void device_mesh::download_mass(double* val)
{
    if (!m64_valid_) throw status_error(NEON_BAD_ARGUMENT, "no mass matrix assembled");
    NEON_CHECK_HIP(hipMemcpyAsync(val, m64_.data(), nnz_ * sizeof(double), hipMemcpyDeviceToHost, s_));
    NEON_CHECK_HIP(hipStreamSynchronize(s_));
}

void device_mesh::enforce_dirichlet(int32_t n_fixed, const int32_t* fixed_dev, double* b_dev, int precision)
{
    if (n_fixed <= 0) return;
    s64_fresh_ = s32_fresh_ = y64_fresh_ = y32_fresh_ = y32s_fresh_ = false;
    if (precision == NEON_PRECISION_FP64)
        dirichlet_kernel<double><<<grid_for(n_fixed), kBlock, 0, s_>>>(n_fixed, fixed_dev, bptr_.data(), bcol_.data(),
                                                                         k64_.data(), b_dev);
    else
        dirichlet_kernel<float><<<grid_for(n_fixed), kBlock, 0, s_>>>(n_fixed, fixed_dev, bptr_.data(), bcol_.data(),
                                                                        k32_.data(), b_dev);
    NEON_CHECK_HIP(hipGetLastError());
}

void device_mesh::pattern(int32_t* row_ptr, int32_t* col) const
{
    if (row_ptr)
    {
        row_ptr[0] = 0;
        for (int32_t nd = 0; nd < N_; ++nd)
        {
            int32_t const d = bptr_h_[nd + 1] - bptr_h_[nd];
            for (int i = 0; i < 3; ++i) row_ptr[3 * nd + i + 1] = 9 * bptr_h_[nd] + (i + 1) * 3 * d;
        }
    }
    if (col)
    {
#pragma omp parallel for schedule(static)
        for (int32_t nd = 0; nd < N_; ++nd)
        {
            int32_t const b0 = bptr_h_[nd], d = bptr_h_[nd + 1] - b0;
            int64_t pos = 9ll * b0;
            for (int i = 0; i < 3; ++i)
                for (int32_t s = 0; s < d; ++s)
                    for (int j = 0; j < 3; ++j) col[pos++] = 3 * bcol_h_[b0 + s] + j;
        }
    }
}

void device_mesh::download(double* val, double* f_int, int)
{
    if (val)
    {
        // the device keeps the values in CSR order (FP64, or FP32 widened here)
        if (!k64_valid_ && !k32_valid_) throw status_error(NEON_BAD_ARGUMENT, "no tangent has been assembled");
        if (k64_valid_)
            NEON_CHECK_HIP(hipMemcpyAsync(val, k64_.data(), nnz_ * sizeof(double), hipMemcpyDeviceToHost, s_));
        else
        {
            device_array<double> tmp;
            tmp.resize(nnz_);
            convert_kernel<float, double><<<grid_for(nnz_), kBlock, 0, s_>>>(nnz_, k32_.data(), tmp.data());
            NEON_CHECK_HIP(hipGetLastError());
            NEON_CHECK_HIP(hipMemcpyAsync(val, tmp.data(), nnz_ * sizeof(double), hipMemcpyDeviceToHost, s_));
        }
        NEON_CHECK_HIP(hipStreamSynchronize(s_));
    }
    if (f_int) NEON_CHECK_HIP(hipMemcpyAsync(f_int, fint_.data(), 3ull * N_ * sizeof(double), hipMemcpyDeviceToHost, s_));
    NEON_CHECK_HIP(hipStreamSynchronize(s_));
}

void device_mesh::download_qp_state(double* sigma, double* D, double* detF, double* eps_p, double* ebar, double* vm)
{
    int64_t const Q = 8ll * E_;
    if (sigma || D)
    {
        device_array<double> d36, s9;
        if (D) d36.resize(36 * Q);
        if (sigma) s9.resize(9 * Q);
        expand_tangent_kernel<<<grid_for(Q), kBlock, 0, s_>>>(Q, coef_.data(), sigma_.data(), D ? d36.data() : nullptr,
                                                               sigma ? s9.data() : nullptr,
                                                               model_ == NEON_MODEL_J2_DAMAGE ? 2 : is_microsphere(model_) ? 1 : 0);
        NEON_CHECK_HIP(hipGetLastError());
        if (D) NEON_CHECK_HIP(hipMemcpyAsync(D, d36.data(), 36 * Q * sizeof(double), hipMemcpyDeviceToHost, s_));
        if (sigma) NEON_CHECK_HIP(hipMemcpyAsync(sigma, s9.data(), 9 * Q * sizeof(double), hipMemcpyDeviceToHost, s_));
        NEON_CHECK_HIP(hipStreamSynchronize(s_));
    }
    if (detF) NEON_CHECK_HIP(hipMemcpyAsync(detF, detF_.data(), Q * sizeof(double), hipMemcpyDeviceToHost, s_));
    if (model_ == NEON_MODEL_J2 || model_ == NEON_MODEL_J2_DAMAGE)
    {
        if (eps_p)
        {
            std::vector<double> e6(6 * Q);
            NEON_CHECK_HIP(hipMemcpyAsync(e6.data(), eps_p_.data(), 6 * Q * sizeof(double), hipMemcpyDeviceToHost, s_));
            NEON_CHECK_HIP(hipStreamSynchronize(s_));
            for (int64_t q = 0; q < Q; ++q)
            {
                double* o = eps_p + 9 * q;
                o[0] = e6[0 * Q + q];
                o[4] = e6[1 * Q + q];
                o[8] = e6[2 * Q + q];
                o[5] = o[7] = e6[3 * Q + q];
                o[2] = o[6] = e6[4 * Q + q];
                o[1] = o[3] = e6[5 * Q + q];
            }
        }
        if (ebar) NEON_CHECK_HIP(hipMemcpyAsync(ebar, ebar_.data(), Q * sizeof(double), hipMemcpyDeviceToHost, s_));
        if (vm) NEON_CHECK_HIP(hipMemcpyAsync(vm, vm_.data(), Q * sizeof(double), hipMemcpyDeviceToHost, s_));
    }
    else
    {
        if (eps_p) std::fill(eps_p, eps_p + 9 * Q, 0.0);
        if (ebar) std::fill(ebar, ebar + Q, 0.0);
        if (vm)
        {
            if (model_ == NEON_MODEL_LINEAR_ELASTIC)
                NEON_CHECK_HIP(hipMemcpyAsync(vm, vm_.data(), Q * sizeof(double), hipMemcpyDeviceToHost, s_));
            else
                std::fill(vm, vm + Q, 0.0);
        }
    }
    NEON_CHECK_HIP(hipStreamSynchronize(s_));
}

// Sliced layout of the pattern (sell_matrix).  Slice s holds the 64 consecutive nodes
// 64 s ... 64 s + 63 (node order keeps each step's gathers of x on few cache lines), the last
// slice padded with empty lanes; a slice has as many steps as its largest block count, and step
// kk of a lane holds the node's block kk (CSR order).  A node's signature is its block count
// and the offsets (column - node) of its blocks.  When all but at most two of a slice's nodes
// share one signature (on a structured mesh: all slices but those that straddle two kinds of
// boundary row), every step is a shared step: its code holds the offset of the common
// signature (shared_code; 0 past its block count) and its exception word the 16-bit offsets of
// the (at most two) other nodes, whose lanes the slice's exception word names.  The other
// slices' steps hold an index into the per-lane offsets (soff16 / soff32).  A lane past its
// node's block count takes offset 0 (its values are zero); an empty lane (snode -1 - m)
// gathers like node m, a node of the common signature.
void device_mesh::build_sliced_pattern()
{
    std::vector<uint64_t> sig(N_);
#pragma omp parallel for schedule(static)
    for (int64_t nd = 0; nd < N_; ++nd)
    {
        int32_t const b0 = bptr_h_[nd], d = bptr_h_[nd + 1] - b0;
        uint64_t h = 0x9e3779b97f4a7c15ull ^ uint64_t(d);
        for (int32_t k = 0; k < d; ++k)
        {
            h ^= uint64_t(int64_t(bcol_h_[size_t(b0) + k]) - nd) + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
            h *= 0xbf58476d1ce4e5b9ull;
        }
        sig[nd] = h;
    }
    int64_t const nslices = (int64_t(N_) + 63) / 64;
    if (nslices > INT32_MAX / 64) throw status_error(NEON_BAD_ARGUMENT, "sliced tangent exceeds the index range");
    slices_ = int32_t(nslices);
    auto degree = [&](int32_t nd) { return bptr_h_[nd + 1] - bptr_h_[nd]; };
    // offset of node nd's block kk (0 past its block count)
    auto offset = [&](int32_t nd, int32_t kk) -> int64_t {
        return kk < degree(nd) ? int64_t(bcol_h_[size_t(bptr_h_[nd]) + kk]) - nd : int64_t(0);
    };
    std::vector<int32_t> sptr(size_t(slices_) + 1, 0), snode(size_t(slices_) * 64), slane(slices_);
    // per slice: the lane of a node of the common signature (-1: per-lane offsets)
    std::vector<int32_t> major(slices_);
#pragma omp parallel for schedule(static)
    for (int32_t sl = 0; sl < slices_; ++sl)
    {
        int64_t const first = 64ll * sl;
        int const valid = int(std::min<int64_t>(64, N_ - first));
        // a signature that all but two lanes share is the signature of one of the first three
        int best = 0, best_count = 0;
        for (int c = 0; c < std::min(valid, 3); ++c)
        {
            int count = 0;
            for (int q = 0; q < valid; ++q) count += sig[first + q] == sig[first + c];
            if (count > best_count) best = c, best_count = count;
        }
        int exc[2] = {-1, -1}, ne = 0;
        bool ok = best_count >= valid - 2;
        int32_t const ref = int32_t(first + best);
        for (int q = 0; q < valid && ok; ++q)
        {
            int32_t const nd = int32_t(first + q);
            bool same = sig[nd] == sig[ref] && degree(nd) == degree(ref);
            // verify the signature (a hash match is not proof)
            for (int32_t kk = 0; same && kk < degree(ref); ++kk) same = offset(nd, kk) == offset(ref, kk);
            if (same) continue;
            if (ne == 2) ok = false;
            else
            {
                exc[ne++] = q;
                for (int32_t kk = 0; kk < degree(nd) && ok; ++kk)
                {
                    int64_t const o = offset(nd, kk);
                    ok = o >= INT16_MIN && o <= INT16_MAX;
                }
            }
        }
        for (int32_t kk = 0; kk < degree(ref) && ok; ++kk)
            ok = offset(ref, kk) > -kSharedBias && offset(ref, kk) < kSharedBias;
        major[sl] = ok ? best : -1;
        slane[sl] = ok ? ((exc[0] & 0xff) | ((exc[1] & 0xff) << 8)) : -1;
        for (int q = 0; q < 64; ++q)
            snode[size_t(sl) * 64 + q] = q < valid ? int32_t(first + q) : -1 - (ok ? ref : int32_t(first));
    }
    for (int32_t sl = 0; sl < slices_; ++sl)
    {
        int32_t md = 0;
        for (int q = 0; q < 64; ++q)
            if (snode[size_t(sl) * 64 + q] >= 0) md = std::max<int32_t>(md, degree(snode[size_t(sl) * 64 + q]));
        if (int64_t(sptr[sl]) + md > int64_t(INT32_MAX) / (9 * 64))
            throw status_error(NEON_BAD_ARGUMENT, "sliced tangent exceeds the index range");
        sptr[sl + 1] = sptr[sl] + md;
    }
    sell_steps_ = sptr[slices_];
    std::vector<int64_t> lane_steps(size_t(slices_) + 1, 0);
    for (int32_t sl = 0; sl < slices_; ++sl)
        lane_steps[size_t(sl) + 1] = lane_steps[sl] + (major[sl] >= 0 ? 0 : sptr[sl + 1] - sptr[sl]);
    int64_t const own_steps = lane_steps[slices_];
    if (own_steps > INT32_MAX / 64) throw status_error(NEON_BAD_ARGUMENT, "sliced tangent exceeds the index range");
    shared_steps_ = sell_steps_ - own_steps;
    std::vector<int32_t> scode(static_cast<size_t>(std::max<int64_t>(sell_steps_, 1)), 0);
    std::vector<int32_t> sexc(static_cast<size_t>(std::max<int64_t>(sell_steps_, 1)), 0);
    std::vector<int32_t> soff32(size_t(std::max<int64_t>(own_steps, 1)) * 64, 0);
    bool fits = true;
#pragma omp parallel for schedule(static) reduction(&& : fits)
    for (int32_t sl = 0; sl < slices_; ++sl)
    {
        const int32_t* g = snode.data() + size_t(sl) * 64;
        int32_t const ks = sptr[sl + 1] - sptr[sl];
        if (major[sl] >= 0)
        {
            int32_t const ref = g[major[sl]];
            int const e0 = slane[sl] & 0xff, e1 = (slane[sl] >> 8) & 0xff;
            for (int32_t kk = 0; kk < ks; ++kk)
            {
                size_t const at = size_t(sptr[sl]) + kk;
                scode[at] = shared_code(offset(ref, kk));
                int32_t const o0 = e0 != 0xff ? int32_t(offset(g[e0], kk)) : 0;
                int32_t const o1 = e1 != 0xff ? int32_t(offset(g[e1], kk)) : 0;
                sexc[at] = int32_t((uint32_t(uint16_t(int16_t(o0)))) | (uint32_t(uint16_t(int16_t(o1))) << 16));
            }
            continue;
        }
        int64_t c = lane_steps[sl];
        for (int32_t kk = 0; kk < ks; ++kk, ++c)
        {
            scode[size_t(sptr[sl]) + kk] = int32_t(c);
            for (int q = 0; q < 64; ++q)
            {
                int64_t const off = g[q] >= 0 ? offset(g[q], kk) : 0;
                soff32[size_t(c) * 64 + q] = int32_t(off);
                if (off < INT16_MIN || off > INT16_MAX) fits = false;
            }
        }
    }
    // slices with per-lane offsets have no exception lanes
    for (int32_t sl = 0; sl < slices_; ++sl)
        if (major[sl] < 0) slane[sl] = 0xffff;
    sptr_.resize(sptr.size());
    snode_.resize(snode.size());
    slane_.resize(slane.size());
    scode_.resize(scode.size());
    sexc_.resize(sexc.size());
    soff32_.resize(soff32.size());
    NEON_CHECK_HIP(hipMemcpyAsync(sptr_.data(), sptr.data(), sptr.size() * sizeof(int32_t), hipMemcpyHostToDevice, s_));
    NEON_CHECK_HIP(hipMemcpyAsync(snode_.data(), snode.data(), snode.size() * sizeof(int32_t), hipMemcpyHostToDevice, s_));
    NEON_CHECK_HIP(hipMemcpyAsync(slane_.data(), slane.data(), slane.size() * sizeof(int32_t), hipMemcpyHostToDevice, s_));
    NEON_CHECK_HIP(hipMemcpyAsync(scode_.data(), scode.data(), scode.size() * sizeof(int32_t), hipMemcpyHostToDevice, s_));
    NEON_CHECK_HIP(hipMemcpyAsync(sexc_.data(), sexc.data(), sexc.size() * sizeof(int32_t), hipMemcpyHostToDevice, s_));
    NEON_CHECK_HIP(hipMemcpyAsync(soff32_.data(), soff32.data(), soff32.size() * sizeof(int32_t), hipMemcpyHostToDevice, s_));
    soff16_ok_ = fits;
    std::vector<int16_t> soff16;
    if (fits)
    {
        soff16.assign(soff32.begin(), soff32.end());
        soff16_.resize(soff16.size());
        NEON_CHECK_HIP(hipMemcpyAsync(soff16_.data(), soff16.data(), soff16.size() * sizeof(int16_t), hipMemcpyHostToDevice, s_));
    }
    // the host vectors go out of scope: wait for the copies
    NEON_CHECK_HIP(hipStreamSynchronize(s_));
}

bool device_mesh::symmetric_selected() const
{
    return (spmv.flags & NEON_SPMV_SLICED) && (spmv.flags & NEON_SPMV_SYMMETRIC) && symmetric_usable();
}

linear_operator device_mesh::op_scaled(const double* minv) const
{
    if (!symmetric_selected()) throw status_error(NEON_BAD_ARGUMENT, "the Jacobi-scaled tangent needs the symmetric copy");
    linear_operator A = op_base(NEON_PRECISION_MIXED);
    attach_symmetric(A, NEON_PRECISION_MIXED, minv);
    return A;
}

linear_operator device_mesh::op(int precision) const
{
    linear_operator A = op_base(precision);
    // nodes of more than 170 neighbours keep the CSR-order values (the sliced copy is packed
    // through LDS rows of 3 * max degree values)
    if (symmetric_selected())
        attach_symmetric(A, precision);
    else if ((spmv.flags & NEON_SPMV_SLICED) && sliced_fits())
        attach_sliced(A, precision, (spmv.flags & NEON_SPMV_INDEX32) != 0);
    return A;
}

linear_operator device_mesh::op_base(int precision) const
{
    linear_operator A;
    A.n = 3ll * N_;
    A.node_block = true;
    A.nodes = N_;
    A.bptr = bptr_.data();
    A.bcol = bcol_.data();
    if (precision == NEON_PRECISION_FP64)
    {
        if (!k64_valid_) throw status_error(NEON_BAD_ARGUMENT, "no FP64 tangent assembled");
        A.v64 = k64_.data();
    }
    else
    {
        if (!k32_valid_) throw status_error(NEON_BAD_ARGUMENT, "no FP32 tangent assembled");
        A.v32 = k32_.data();
    }
    A.nb_lanes = spmv.lanes;
    A.nb_nt = (spmv.flags & NEON_SPMV_NONTEMPORAL) != 0;
    A.blocks_override = precision == NEON_PRECISION_FP64 ? spmv.blocks64 : spmv.blocks32;
    return A;
}

namespace
{
// sliced copy of CSR-order values (a slice's 64 rows of 3 * max degree values staged in LDS)
template <typename VT>
void pack_sliced(hipStream_t s, const sell_matrix& m, const int32_t* bptr, const VT* val, VT* sval, int32_t max_degree)
{
    int32_t const row_len = 3 * std::max<int32_t>(max_degree, 1);
    size_t const lds = size_t(64) * row_len * sizeof(VT);
    if (lds > 65536) throw status_error(NEON_BAD_ARGUMENT, "node degree too large for the sliced tangent copy");
    sell_pack_lds_kernel<VT><<<m.slices, 256, lds, s>>>(m, bptr, val, sval, row_len);
    NEON_CHECK_HIP(hipGetLastError());
}
} // namespace

bool device_mesh::sliced_fits() const
{
    return size_t(64) * 3 * std::max<int32_t>(max_degree_, 1) * sizeof(double) <= 65536;
}

void device_mesh::attach_sliced(linear_operator& A, int precision, bool index32) const
{
    if (!sliced_fits()) throw status_error(NEON_BAD_ARGUMENT, "node degree too large for the sliced tangent copy");
    sell_matrix const m{N_, slices_, sptr_.data(), snode_.data(), slane_.data(), scode_.data(), sexc_.data(), nullptr,
                        soff32_.data()};
    size_t const vals = size_t(sell_steps_) * 9 * 64;
    if (precision == NEON_PRECISION_FP64)
    {
        if (!k64_valid_) throw status_error(NEON_BAD_ARGUMENT, "no FP64 tangent assembled");
        if (!s64_fresh_)
        {
            s64_.resize(vals);
            pack_sliced<double>(s_, m, bptr_.data(), k64_.data(), s64_.data(), max_degree_);
            s64_fresh_ = true;
        }
        A.s64 = s64_.data();
    }
    else
    {
        if (!k32_valid_) throw status_error(NEON_BAD_ARGUMENT, "no FP32 tangent assembled");
        if (!s32_fresh_)
        {
            s32_.resize(vals);
            pack_sliced<float>(s_, m, bptr_.data(), k32_.data(), s32_.data(), max_degree_);
            s32_fresh_ = true;
        }
        A.s32 = s32_.data();
    }
    A.sell = true;
    A.sym = false;
    A.slices = slices_;
    A.sptr = sptr_.data();
    A.snode = snode_.data();
    A.slane = slane_.data();
    A.scode = scode_.data();
    A.sexc = sexc_.data();
    A.soff32 = soff32_.data();
    A.soff16 = index32 || !soff16_ok_ ? nullptr : soff16_.data();
}

// The symmetric sliced layout (sym_matrix).  Per slice of 64 consecutive nodes: the direct list
// (the upper offsets of its nodes' blocks, one record each after the diagonal record) and the
// lower list (the offsets of its nodes' lower neighbours), merged by offset into entries that
// carry the direct record and the records holding the transposed blocks (in slice
// floor((64 s - o) / 64) and the next one; a lower neighbour j of node i has i = j + o in its
// own direct list, so the record exists).  The copy is used only where it pays: every slice needs
// at most kMaxSymEntries entries and the records are at most 3/4 of the full copy's steps.
// Schedule: the slices sorted by (band of the in-plane index, plane), with the plane stride P and
// the row length L estimated from the 13 offsets found in most slices (for a structured block
// ascending 1, L - 1, L, L + 1, ..., P + L + 1: L = smallest above 1, plus 1; P = largest - L -
// 1) and bands of 16 rows; any estimate gives a valid order, a good one makes the transposed
// reads hit the L2.  Each XCD takes one contiguous eighth of the order.
void device_mesh::build_symmetric_pattern()
{
    sym_ok_ = false;
    sym_slots_ = sym_max_entries_ = 0;
    sym_records_ = 0;
    if (N_ <= 0 || slices_ <= 0) return;
    std::vector<std::vector<int32_t>> up(slices_), lo(slices_);
    bool too_many = false;
#pragma omp parallel for schedule(static) reduction(|| : too_many)
    for (int32_t sl = 0; sl < slices_; ++sl)
    {
        std::vector<int32_t>& u = up[sl];
        std::vector<int32_t>& l = lo[sl];
        int64_t const first = 64ll * sl, last = std::min<int64_t>(first + 64, N_);
        for (int64_t nd = first; nd < last; ++nd)
            for (int32_t b = bptr_h_[nd]; b < bptr_h_[nd + 1]; ++b)
            {
                int64_t const o = int64_t(bcol_h_[b]) - nd;
                if (o > 0) u.push_back(int32_t(o));
                if (o < 0) l.push_back(int32_t(-o));
            }
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        std::sort(l.begin(), l.end());
        l.erase(std::unique(l.begin(), l.end()), l.end());
        std::vector<int32_t> both;
        std::set_union(u.begin(), u.end(), l.begin(), l.end(), std::back_inserter(both));
        if (int(both.size()) > kMaxSymEntries) too_many = true;
    }
    if (too_many) return;
    std::vector<int32_t> base(size_t(slices_) + 1, 0);
    for (int32_t sl = 0; sl < slices_; ++sl) base[sl + 1] = base[sl] + 1 + int32_t(up[sl].size());
    int64_t const records = base[slices_];
    if (records > int64_t(3) * sell_steps_ / 4 || records > int64_t(INT32_MAX) / 64) return;
    // entries
    std::vector<int32_t> eptr(size_t(slices_) + 1, 0);
    std::vector<std::vector<int32_t>> ents(slices_);
    int32_t max_entries = 0;
#pragma omp parallel for schedule(static) reduction(max : max_entries)
    for (int32_t sl = 0; sl < slices_; ++sl)
    {
        std::vector<int32_t> both;
        std::set_union(up[sl].begin(), up[sl].end(), lo[sl].begin(), lo[sl].end(), std::back_inserter(both));
        max_entries = std::max<int32_t>(max_entries, int32_t(both.size()));
        auto record_of = [&](int64_t t, int32_t o) -> int32_t {
            // record of slice t's direct block for offset o, or -1
            if (t < 0 || t >= slices_) return -1;
            auto const& d = up[t];
            auto it = std::lower_bound(d.begin(), d.end(), o);
            return it != d.end() && *it == o ? base[t] + 1 + int32_t(it - d.begin()) : -1;
        };
        std::vector<int32_t>& e = ents[sl];
        for (int32_t o : both)
        {
            bool const lower = std::binary_search(lo[sl].begin(), lo[sl].end(), o);
            int64_t const sa = (64ll * sl - o) >> 6; // floor division by 64
            e.push_back(o);
            e.push_back(record_of(sl, o));
            e.push_back(lower ? record_of(sa, o) : -1);
            e.push_back(lower ? record_of(sa + 1, o) : -1);
        }
    }
    for (int32_t sl = 0; sl < slices_; ++sl) eptr[sl + 1] = eptr[sl] + int32_t(ents[sl].size() / 4);
    std::vector<int32_t> ent(static_cast<size_t>(eptr[slices_]) * 4);
    std::vector<int32_t> rslice(static_cast<size_t>(records)), roff(static_cast<size_t>(records), 0);
    std::vector<int32_t> rdiag(static_cast<size_t>(slices_));
#pragma omp parallel for schedule(static)
    for (int32_t sl = 0; sl < slices_; ++sl)
    {
        std::copy(ents[sl].begin(), ents[sl].end(), ent.begin() + size_t(eptr[sl]) * 4);
        rdiag[sl] = base[sl];
        rslice[size_t(base[sl])] = sl;
        for (size_t k = 0; k < up[sl].size(); ++k)
        {
            rslice[size_t(base[sl]) + 1 + k] = sl;
            roff[size_t(base[sl]) + 1 + k] = up[sl][k];
        }
    }
    // schedule: plane stride and row length from the offsets present in most slices
    std::map<int32_t, int64_t> count;
    for (int32_t sl = 0; sl < slices_; ++sl)
        for (int32_t o : up[sl]) ++count[o];
    std::vector<std::pair<int64_t, int32_t>> byc;
    for (auto const& kv : count) byc.emplace_back(-kv.second, kv.first);
    std::sort(byc.begin(), byc.end());
    std::vector<int32_t> main_offs;
    for (size_t k = 0; k < byc.size() && k < 13; ++k) main_offs.push_back(byc[k].second);
    std::sort(main_offs.begin(), main_offs.end());
    int64_t L = 64, P = int64_t(N_);
    for (int32_t o : main_offs)
        if (o > 1)
        {
            L = int64_t(o) + 1;
            break;
        }
    if (!main_offs.empty() && int64_t(main_offs.back()) - L - 1 > 0) P = int64_t(main_offs.back()) - L - 1;
    int64_t const band = std::max<int64_t>(1, 16 * L);
    std::vector<int32_t> order(slices_);
    for (int32_t sl = 0; sl < slices_; ++sl) order[sl] = sl;
    auto key = [&](int32_t sl) {
        int64_t const i0 = 64ll * sl;
        return std::make_tuple((i0 % P) / band, i0 / P, sl);
    };
    std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return key(a) < key(b); });
    int64_t per = 0;
    for (int x = 0; x < 8; ++x)
    {
        int64_t const b0 = int64_t(slices_) * x / 8, b1 = int64_t(slices_) * (x + 1) / 8;
        per = std::max<int64_t>(per, (b1 - b0 + 3) / 4);
    }
    per = std::max<int64_t>(per, 1);
    if (per * 32 > INT32_MAX) return;
    std::vector<int32_t> sched(size_t(per) * 32, -1);
    for (int x = 0; x < 8; ++x)
    {
        int64_t const b0 = int64_t(slices_) * x / 8, b1 = int64_t(slices_) * (x + 1) / 8;
        for (int64_t e = 0; e < b1 - b0; ++e) sched[size_t(((e / 4) * 8 + x) * 4 + e % 4)] = order[size_t(b0 + e)];
    }
    auto upload = [&](device_array<int32_t>& d, std::vector<int32_t> const& h) {
        d.resize(std::max<size_t>(h.size(), 4));
        if (!h.empty())
            NEON_CHECK_HIP(hipMemcpyAsync(d.data(), h.data(), h.size() * sizeof(int32_t), hipMemcpyHostToDevice, s_));
    };
    upload(sym_sched_, sched);
    upload(sym_rdiag_, rdiag);
    upload(sym_eptr_, eptr);
    upload(sym_ent_, ent);
    upload(sym_rslice_, rslice);
    upload(sym_roff_, roff);
    // the host vectors go out of scope: wait for the copies
    NEON_CHECK_HIP(hipStreamSynchronize(s_));
    sym_slots_ = int32_t(sched.size());
    sym_records_ = records;
    sym_max_entries_ = max_entries;
    sym_ok_ = true;
}

bool device_mesh::symmetric_usable() const { return sym_ok_ && model_is_symmetric(model_); }

sym_matrix device_mesh::sym_layout() const
{
    return sym_matrix{N_,
                      slices_,
                      sym_slots_,
                      sym_records_,
                      sym_sched_.data(),
                      sym_rdiag_.data(),
                      sym_eptr_.data(),
                      reinterpret_cast<const int4v*>(sym_ent_.data()),
                      sym_rslice_.data(),
                      sym_roff_.data()};
}

namespace
{
template <typename VT>
void pack_symmetric(hipStream_t s, const sym_matrix& m, const int32_t* bptr, const int32_t* bcol, const VT* val,
                    VT* sval, const double* scale_minv = nullptr)
{
    sym_pack_kernel<VT><<<int((m.records + 3) / 4), 256, 0, s>>>(m, bptr, bcol, val, sval, scale_minv);
    NEON_CHECK_HIP(hipGetLastError());
}
} // namespace

void device_mesh::attach_symmetric(linear_operator& A, int precision, const double* scale_minv) const
{
    if (!symmetric_usable()) throw status_error(NEON_BAD_ARGUMENT, "the tangent has no symmetric sliced copy");
    if (scale_minv && precision == NEON_PRECISION_FP64)
        throw status_error(NEON_BAD_ARGUMENT, "the Jacobi-scaled copy holds FP32 values only");
    sym_matrix const m = sym_layout();
    size_t const vals = size_t(sym_records_) * 576;
    if (precision == NEON_PRECISION_FP64)
    {
        if (!k64_valid_) throw status_error(NEON_BAD_ARGUMENT, "no FP64 tangent assembled");
        if (!y64_fresh_)
        {
            y64_.resize(vals);
            pack_symmetric<double>(s_, m, bptr_.data(), bcol_.data(), k64_.data(), y64_.data());
            y64_fresh_ = true;
        }
        A.s64 = y64_.data();
    }
    else
    {
        if (!k32_valid_) throw status_error(NEON_BAD_ARGUMENT, "no FP32 tangent assembled");
        if (scale_minv)
        {
            // the mixed solver's recurrence copy, D^-1/2 K D^-1/2, beside the unscaled one
            if (!y32s_fresh_)
            {
                y32s_.resize(vals);
                pack_symmetric<float>(s_, m, bptr_.data(), bcol_.data(), k32_.data(), y32s_.data(), scale_minv);
                y32s_fresh_ = true;
            }
            A.s32 = y32s_.data();
        }
        else
        {
            if (!y32_fresh_)
            {
                y32_.resize(vals);
                pack_symmetric<float>(s_, m, bptr_.data(), bcol_.data(), k32_.data(), y32_.data());
                y32_fresh_ = true;
            }
            A.s32 = y32_.data();
        }
    }
    A.sell = false;
    A.sym = true;
    A.slices = slices_;
    A.sym_slots = sym_slots_;
    A.sym_records = sym_records_;
    A.sym_sched = sym_sched_.data();
    A.sym_rdiag = sym_rdiag_.data();
    A.sym_eptr = sym_eptr_.data();
    A.sym_ent = sym_ent_.data();
    A.sym_rslice = sym_rslice_.data();
    A.sym_roff = sym_roff_.data();
}

namespace
{
template <typename VT, typename OT>
__global__ __launch_bounds__(kBlock) void nb_jacobi_kernel(int32_t N, const int32_t* __restrict__ bptr,
                                                           const int32_t* __restrict__ self,
                                                           const VT* __restrict__ val, OT* __restrict__ minv,
                                                           int eigen_rule)
{
    int64_t const r = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (r >= 3ll * N) return;
    int64_t const nd = r / 3;
    int const i = int(r % 3);
    int32_t const b0 = bptr[nd];
    int64_t const d = bptr[nd + 1] - b0;
    double const v = double(val[9ll * b0 + 3ll * i * d + 3ll * self[nd] + i]);
    double const inv = eigen_rule ? (v != 0.0 ? 1.0 / v : 1.0) : 1.0 / v;
    minv[r] = OT(inv);
}
} // namespace

void device_mesh::jacobi(double* minv, int eigen_rule, int precision)
{
    if (precision == NEON_PRECISION_FP64)
        nb_jacobi_kernel<double, double><<<grid_for(3ll * N_), kBlock, 0, s_>>>(N_, bptr_.data(), self_slot_.data(),
                                                                                  k64_.data(), minv, eigen_rule);
    else
        nb_jacobi_kernel<float, double><<<grid_for(3ll * N_), kBlock, 0, s_>>>(N_, bptr_.data(), self_slot_.data(),
                                                                                 k32_.data(), minv, eigen_rule);
    NEON_CHECK_HIP(hipGetLastError());
}

void device_mesh::jacobi_mass(double* minv)
{
    if (!m64_valid_) throw status_error(NEON_BAD_ARGUMENT, "no mass matrix assembled");
    nb_jacobi_kernel<double, double><<<grid_for(3ll * N_), kBlock, 0, s_>>>(N_, bptr_.data(), self_slot_.data(),
                                                                              m64_.data(), minv, 1);
    NEON_CHECK_HIP(hipGetLastError());
}

void device_mesh::clear_trajectory()
{
    traj_u_.clear();
    traj_ep_.clear();
    traj_eb_.clear();
}

void device_mesh::record_step()
{
    int64_t const Q = 8ll * E_;
    auto copy = [&](const double* src, int64_t count) {
        auto a = std::make_unique<device_array<double>>();
        a->resize(count);
        NEON_CHECK_HIP(hipMemcpyAsync(a->data(), src, count * sizeof(double), hipMemcpyDeviceToDevice, s_));
        return a;
    };
    traj_u_.push_back(copy(u_.data(), 3ll * N_));
    if (model_ == NEON_MODEL_J2)
    {
        traj_ep_.push_back(copy(eps_p_old_.data(), 6 * Q));
        traj_eb_.push_back(copy(ebar_old_.data(), Q));
    }
}

void device_mesh::restore_step(int32_t step)
{
    if (step < 0 || step >= trajectory_steps()) throw status_error(NEON_BAD_ARGUMENT, "no such recorded step");
    int64_t const Q = 8ll * E_;
    NEON_CHECK_HIP(hipMemcpyAsync(u_.data(), traj_u_[step]->data(), 3ll * N_ * sizeof(double),
                                  hipMemcpyDeviceToDevice, s_));
    if (model_ == NEON_MODEL_J2)
    {
        NEON_CHECK_HIP(hipMemcpyAsync(eps_p_old_.data(), traj_ep_[step]->data(), 6 * Q * sizeof(double),
                                      hipMemcpyDeviceToDevice, s_));
        NEON_CHECK_HIP(hipMemcpyAsync(ebar_old_.data(), traj_eb_[step]->data(), Q * sizeof(double),
                                      hipMemcpyDeviceToDevice, s_));
    }
}

void device_mesh::j2_adjoint_begin()
{
    int64_t const Q = 8ll * E_;
    mu_ep_.resize(6 * Q);
    mu_eb_.resize(Q);
    NEON_CHECK_HIP(hipMemsetAsync(mu_ep_.data(), 0, 6 * Q * sizeof(double), s_));
    NEON_CHECK_HIP(hipMemsetAsync(mu_eb_.data(), 0, Q * sizeof(double), s_));
}

void device_mesh::j2_adjoint_rhs(double* rhs)
{
    int64_t const Q = 8ll * E_;
    j2_history h{eps_p_old_.data(), ebar_old_.data(), mu_ep_.data(), mu_eb_.data(), Q};
    element_params prm{params_.data()};
    for (int c = 0; c + 1 < int(colour_offsets_.size()); ++c)
    {
        int32_t const count = colour_offsets_[c + 1] - colour_offsets_[c];
        if (count <= 0) continue;
        j2_adjoint_rhs_kernel<<<grid_for(count), kBlock, 0, s_>>>(elems_.data() + colour_offsets_[c], count,
                                                                  conn_.data(), X_.data(), u_.data(), prm, h, rhs);
        NEON_CHECK_HIP(hipGetLastError());
    }
}

void device_mesh::j2_adjoint_history(const double* lambda, double* grad)
{
    int64_t const Q = 8ll * E_;
    j2_history h{eps_p_old_.data(), ebar_old_.data(), mu_ep_.data(), mu_eb_.data(), Q};
    element_params prm{params_.data()};
    j2_adjoint_history_kernel<<<grid_for(E_), kBlock, 0, s_>>>(E_, conn_.data(), X_.data(), u_.data(), lambda, prm, h,
                                                               grad);
    NEON_CHECK_HIP(hipGetLastError());
}

void device_mesh::parameter_gradient(const double* u_dev, const double* lambda_dev, double* grad_dev)
{
    if (model_ != NEON_MODEL_NEOHOOKE)
        throw status_error(NEON_BAD_ARGUMENT, "the adjoint parameter gradient supports the neo-Hooke model");
    neohooke_parameter_gradient_kernel<<<grid_for(8ll * E_), kBlock, 0, s_>>>(E_, conn_.data(), X_.data(), u_dev,
                                                                             lambda_dev, grad_dev);
    NEON_CHECK_HIP(hipGetLastError());
}
} // namespace neon_amd
