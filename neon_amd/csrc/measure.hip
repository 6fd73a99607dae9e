// Measurement library (libneon_hip_measure.so, C ABI of neon_hip_measure.h): timing probes of
// the SpMV and assembly kernels of a context's mesh and the streaming kernels that calibrate the
// achievable HBM bandwidth and the rocprofv3 traffic counters.  It links against libneon_hip.so
// and works on that library's contexts; the product library itself carries no probe.
#include "engine.hpp"
#include "krylov_kernels.hpp"
#include "neon_hip_measure.h"

#include <algorithm>
#include <string>

using namespace neon_amd;

namespace
{
template <typename F>
int measured(neon_context* ctx, F&& f)
{
    if (!ctx) return NEON_BAD_ARGUMENT;
    try
    {
        NEON_CHECK_HIP(hipSetDevice(ctx->device));
        int const r = f();
        if (r == NEON_OK) ctx->error.clear();
        return r;
    }
    catch (status_error const& e)
    {
        ctx->error = e.what();
        return e.code;
    }
    catch (std::exception const& e)
    {
        ctx->error = e.what();
        return NEON_FATAL;
    }
}

int grid_for(int64_t n) { return int(std::max<int64_t>(1, (n + kBlock - 1) / kBlock)); }

// streaming read of n doubles in 16-byte nontemporal loads
__global__ __launch_bounds__(kBlock) void stream_sum_kernel(int64_t n, const double* __restrict__ v,
                                                            double* __restrict__ out)
{
    int64_t beg, end;
    xcd_chunk(n / 2, beg, end);
    typedef double d2 __attribute__((ext_vector_type(2)));
    const d2* v2 = reinterpret_cast<const d2*>(v);
    double s = 0.0;
    for (int64_t i = beg + threadIdx.x; i < end; i += kBlock)
    {
        d2 const t = __builtin_nontemporal_load(v2 + i);
        s += t.x + t.y;
    }
    s = block_sum<kBlock>(s);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// copy of n values: the reference kernel of known traffic for the FETCH_SIZE / WRITE_SIZE counters
template <typename VT>
__global__ __launch_bounds__(kBlock) void stream_copy_kernel(int64_t n, const VT* __restrict__ a, VT* __restrict__ b)
{
    int64_t const i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ __launch_bounds__(kBlock) void widen_kernel(int64_t n, const float* __restrict__ a, double* __restrict__ b)
{
    int64_t const i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) b[i] = double(a[i]);
}

double stream_ms(hipStream_t s, int64_t bytes, int reps, bool copy)
{
    int64_t const n = bytes / 8;
    device_array<double> a, b, part;
    a.resize(n);
    part.resize(4096);
    if (copy) b.resize(n);
    fill_ones(s, n, a.data());
    hipEvent_t e0, e1;
    NEON_CHECK_HIP(hipEventCreate(&e0));
    NEON_CHECK_HIP(hipEventCreate(&e1));
    auto launch = [&]() {
        if (copy)
            stream_copy_kernel<double><<<grid_for(n), kBlock, 0, s>>>(n, a.data(), b.data());
        else
            stream_sum_kernel<<<4096, kBlock, 0, s>>>(n, a.data(), part.data());
    };
    launch();
    NEON_CHECK_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) launch();
    NEON_CHECK_HIP(hipEventRecord(e1, s));
    NEON_CHECK_HIP(hipEventSynchronize(e1));
    NEON_CHECK_HIP(hipGetLastError());
    float ms = 0.0f;
    NEON_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return double(ms) / reps;
}

// average ms of `reps` launches of f on s, after one untimed launch
template <typename F>
double time_launches(hipStream_t s, int reps, F&& f)
{
    hipEvent_t e0, e1;
    NEON_CHECK_HIP(hipEventCreate(&e0));
    NEON_CHECK_HIP(hipEventCreate(&e1));
    f();
    NEON_CHECK_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) f();
    NEON_CHECK_HIP(hipEventRecord(e1, s));
    NEON_CHECK_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    NEON_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return double(ms) / reps;
}
} // namespace

extern "C" {

int neon_mesh_spmv_bench(neon_context* ctx, int32_t precision, int32_t lanes, int32_t flags, int32_t blocks,
                         int32_t reps, double* avg_ms)
{
    return measured(ctx, [&] {
        if (!ctx->mesh) throw status_error(NEON_BAD_ARGUMENT, "no mesh on this context");
        device_mesh& m = *ctx->mesh;
        if (lanes != 16 && lanes != 32) throw status_error(NEON_BAD_ARGUMENT, "lanes must be 16 or 32");
        if (reps <= 0 || !avg_ms || blocks < 0) throw status_error(NEON_BAD_ARGUMENT, "bad repetitions");
        if (flags & ~(NEON_SPMV_NONTEMPORAL | NEON_SPMV_SLICED | NEON_SPMV_INDEX32 | NEON_SPMV_SYMMETRIC | NEON_MEASURE_DOT | NEON_MEASURE_DIRTY))
            throw status_error(NEON_BAD_ARGUMENT, "bad SpMV flags");
        linear_operator A = m.op(precision);
        A.nb_lanes = lanes;
        A.nb_nt = (flags & NEON_SPMV_NONTEMPORAL) != 0;
        A.blocks_override = blocks;
        A.sell = A.sym = false;
        if ((flags & NEON_SPMV_SLICED) && (flags & NEON_SPMV_SYMMETRIC) && m.symmetric_usable())
            m.attach_symmetric(A, precision);
        else if (flags & NEON_SPMV_SLICED)
            m.attach_sliced(A, precision, (flags & NEON_SPMV_INDEX32) != 0);
        int64_t const n = m.dofs();
        hipStream_t const s = ctx->stream;
        // NEON_MEASURE_DOT: the solvers' form with partials of x.y; NEON_MEASURE_DIRTY: each
        // product follows a write of 4 vectors of n entries, as the Krylov updates do (the
        // writes are timed alone and subtracted)
        device_array<double> part, dirty;
        bool const dot = (flags & NEON_MEASURE_DOT) != 0;
        bool const dirt = (flags & NEON_MEASURE_DIRTY) != 0;
        if (dot) part.resize(kMaxPartials);
        if (dirt) dirty.resize(4 * n);
        double total = 0.0;
        if (precision == NEON_PRECISION_FP64)
        {
            device_array<double> x, y;
            x.resize(n);
            y.resize(n);
            fill_ones(s, n, x.data());
            total = time_launches(s, reps, [&] {
                if (dirt) fill_ones(s, 4 * n, dirty.data());
                A.apply64(s, x.data(), y.data(), part.data(), nullptr);
            });
        }
        else
        {
            device_array<float> x, y;
            x.resize(n);
            y.resize(n);
            fill_ones(s, n, x.data());
            total = time_launches(s, reps, [&] {
                if (dirt) fill_ones(s, 4 * n, dirty.data());
                A.apply32(s, x.data(), y.data(), part.data(), nullptr);
            });
        }
        if (dirt) total -= time_launches(s, reps, [&] { fill_ones(s, 4 * n, dirty.data()); });
        *avg_ms = total;
        return NEON_OK;
    });
}

int neon_mesh_spmv_apply(neon_context* ctx, int32_t precision, int32_t lanes, int32_t flags, int32_t blocks,
                         const double* x, double* y)
{
    return measured(ctx, [&] {
        if (!ctx->mesh) throw status_error(NEON_BAD_ARGUMENT, "no mesh on this context");
        device_mesh& m = *ctx->mesh;
        if (lanes != 16 && lanes != 32) throw status_error(NEON_BAD_ARGUMENT, "lanes must be 16 or 32");
        if (!x || !y || blocks < 0) throw status_error(NEON_BAD_ARGUMENT, "bad SpMV arguments");
        if (flags & ~(NEON_SPMV_NONTEMPORAL | NEON_SPMV_SLICED | NEON_SPMV_INDEX32 | NEON_SPMV_SYMMETRIC | NEON_MEASURE_DOT))
            throw status_error(NEON_BAD_ARGUMENT, "bad SpMV flags");
        linear_operator A = m.op(precision);
        A.nb_lanes = lanes;
        A.nb_nt = (flags & NEON_SPMV_NONTEMPORAL) != 0;
        A.blocks_override = blocks;
        A.sell = A.sym = false;
        if ((flags & NEON_SPMV_SLICED) && (flags & NEON_SPMV_SYMMETRIC) && m.symmetric_usable())
            m.attach_symmetric(A, precision);
        else if (flags & NEON_SPMV_SLICED)
            m.attach_sliced(A, precision, (flags & NEON_SPMV_INDEX32) != 0);
        int64_t const n = m.dofs();
        hipStream_t const s = ctx->stream;
        device_array<double> xd, yd, part;
        xd.resize(n);
        yd.resize(n);
        if (flags & NEON_MEASURE_DOT) part.resize(kMaxPartials);
        NEON_CHECK_HIP(hipMemcpyAsync(xd.data(), x, n * sizeof(double), hipMemcpyHostToDevice, s));
        if (precision == NEON_PRECISION_FP64)
            A.apply64(s, xd.data(), yd.data(), part.data(), nullptr);
        else
        {
            device_array<float> x32, y32;
            x32.resize(n);
            y32.resize(n);
            convert_to_float(s, n, xd.data(), x32.data());
            A.apply32(s, x32.data(), y32.data(), part.data(), nullptr);
            widen_kernel<<<grid_for(n), kBlock, 0, s>>>(n, y32.data(), yd.data());
            NEON_CHECK_HIP(hipGetLastError());
            NEON_CHECK_HIP(hipStreamSynchronize(s));
        }
        NEON_CHECK_HIP(hipMemcpyAsync(y, yd.data(), n * sizeof(double), hipMemcpyDeviceToHost, s));
        NEON_CHECK_HIP(hipStreamSynchronize(s));
        return NEON_OK;
    });
}

int neon_mesh_assembly_bench(neon_context* ctx, int32_t precision, int32_t blocks_per_cu, int32_t reps,
                             double* avg_ms)
{
    return measured(ctx, [&] {
        if (!ctx->mesh) throw status_error(NEON_BAD_ARGUMENT, "no mesh on this context");
        if (blocks_per_cu <= 0 || blocks_per_cu > 1024 || reps <= 0 || !avg_ms)
            throw status_error(NEON_BAD_ARGUMENT, "bad assembly probe arguments");
        device_mesh& m = *ctx->mesh;
        int const keep = m.assembly_blocks_per_cu;
        m.assembly_blocks_per_cu = blocks_per_cu;
        double total = 0.0;
        try
        {
            m.assemble(m.u(), precision); // warm
            for (int r = 0; r < reps; ++r)
            {
                m.assemble(m.u(), precision);
                total += m.last_assemble_ms;
            }
        }
        catch (...)
        {
            m.assembly_blocks_per_cu = keep;
            throw;
        }
        m.assembly_blocks_per_cu = keep;
        *avg_ms = total / reps;
        return NEON_OK;
    });
}

int neon_stream_probe(neon_context* ctx, int64_t bytes, int32_t reps, double* read_ms, double* copy_ms)
{
    return measured(ctx, [&] {
        if (bytes < 16 || reps <= 0) throw status_error(NEON_BAD_ARGUMENT, "bad probe arguments");
        if (read_ms) *read_ms = stream_ms(ctx->stream, bytes, reps, false);
        if (copy_ms) *copy_ms = stream_ms(ctx->stream, bytes, reps, true);
        return NEON_OK;
    });
}

} // extern "C"
