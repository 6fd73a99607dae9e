// C ABI of neon_hip.h: argument checks, exception -> status translation, host<->device
// staging.  The product path has no CPU fallback: without a device every compute entry
// point returns NEON_NO_DEVICE.
#include "engine.hpp"
#include "newton.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <new>
#include <string>
#include <vector>

using namespace neon_amd;

namespace
{
template <typename F>
int guarded(neon_context* ctx, F&& f)
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
    catch (std::bad_alloc const&)
    {
        ctx->error = "host allocation failed";
        return NEON_OOM;
    }
    catch (std::exception const& e)
    {
        ctx->error = e.what();
        return NEON_FATAL;
    }
}

int need_mesh(neon_context* ctx)
{
    if (!ctx->mesh) throw status_error(NEON_BAD_ARGUMENT, "no mesh on this context");
    return NEON_OK;
}

int lanes_for(double avg)
{
    int l = 1;
    while (l < 64 && l < avg) l <<= 1;
    return l;
}

// ---- the allocation-free CSR solve (neon_krylov_solve_csr_noalloc)
// Its device memory beyond the caller's inputs and outputs lives in `scratch`, one copy per
// GPU, held until the process ends: a validation flag, then the Jacobi inverse and the
// solver's work vectors (9 vectors of at most kNoAllocRows entries: minv and BiCGSTAB's eight;
// CG uses minv and three).  Every call on a GPU shares that GPU's copy, so calls that overlap
// on one GPU race; each call rewrites what it reads (the flag is cleared by a kernel, the
// vectors are initialised by the solvers).
constexpr int64_t kNoAllocRows = int64_t(1) << 18;
constexpr int64_t kScratchHeader = 32; // doubles before the vectors (256-byte aligned)
constexpr int64_t kScratchDoubles = kScratchHeader + 9 * kNoAllocRows;
constexpr int kMaxDevices = 64;
__device__ __align__(256) double scratch[kScratchDoubles];

std::mutex scratch_mutex;
double* scratch_ptr[kMaxDevices] = {};

// The device address of this GPU's `scratch` (the current device's), looked up once per device
double* scratch_address(int device)
{
    if (device < 0 || device >= kMaxDevices) throw status_error(NEON_BAD_ARGUMENT, "device index out of range");
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[device])
    {
        void* p = nullptr;
        hipError_t const e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw status_error(NEON_HIP_ERROR, std::string("hipGetSymbolAddress of the solver scratch failed: ")
                                                   + hipGetErrorString(e));
        scratch_ptr[device] = static_cast<double*>(p);
    }
    return scratch_ptr[device];
}

__global__ void clear_flag_kernel(double* flag)
{
    if (threadIdx.x == 0) flag[0] = 0.0;
}

// flag <- 1 if the CSR arrays are malformed: row_ptr[0] != 0, row_ptr[n] != nnz, a decreasing
// row pointer or a column outside [0, n).  Reads row_ptr[0..n] and col[0..nnz) only.
__global__ void check_csr_kernel(int32_t n, int64_t nnz, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                 double* flag)
{
    int64_t const stride = int64_t(gridDim.x) * blockDim.x;
    int64_t const t0 = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    bool bad = false;
    if (t0 == 0) bad = rp[0] != 0 || int64_t(rp[n]) != nnz;
    for (int64_t r = t0; r < n; r += stride) bad = bad || rp[r + 1] < rp[r];
    for (int64_t q = t0; q < nnz; q += stride) bad = bad || ci[q] < 0 || ci[q] >= n;
    if (bad) flag[0] = 1.0;
}
} // namespace

extern "C" {

int neon_device_count(int* count)
{
    if (!count) return NEON_BAD_ARGUMENT;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return NEON_OK;
}

int neon_context_create(int device, neon_context** out)
{
    if (!out) return NEON_BAD_ARGUMENT;
    *out = nullptr;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess || c == 0) return NEON_NO_DEVICE;
    if (device < 0 || device >= c) return NEON_BAD_ARGUMENT;
    auto* ctx = new (std::nothrow) neon_context;
    if (!ctx) return NEON_OOM;
    ctx->device = device;
    int const r = guarded(ctx, [&] {
        NEON_CHECK_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        // this device's copy of the allocation-free solve's scratch (the device is current)
        (void)scratch_address(device);
        ctx->engine = new krylov_engine(ctx->stream);
        ctx->engine->prof = &ctx->profiler;
        return NEON_OK;
    });
    if (r != NEON_OK)
    {
        neon_context_destroy(ctx);
        return r;
    }
    *out = ctx;
    return NEON_OK;
}

int neon_context_destroy(neon_context* ctx)
{
    if (!ctx) return NEON_BAD_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    delete ctx->mesh;
    delete ctx->engine;
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return NEON_OK;
}

const char* neon_context_last_error(const neon_context* ctx) { return ctx ? ctx->error.c_str() : "null context"; }

int neon_context_synchronize(neon_context* ctx)
{
    return guarded(ctx, [&] {
        NEON_CHECK_HIP(hipStreamSynchronize(ctx->stream));
        return NEON_OK;
    });
}

int neon_context_profile(neon_context* ctx, int enable)
{
    return guarded(ctx, [&] {
        ctx->profiler.discard();
        ctx->profiler.enabled = enable != 0;
        return NEON_OK;
    });
}

int neon_context_profile_read(neon_context* ctx, int kernel_class, int reset, int64_t* count, double* total_ms)
{
    return guarded(ctx, [&] {
        ctx->profiler.read(kernel_class, reset != 0, count, total_ms);
        return NEON_OK;
    });
}

int neon_krylov_solve_csr(neon_context* ctx, int solver, int criterion, int precision, int32_t n, int64_t nnz,
                          const int32_t* row_ptr, const int32_t* col, const double* val, const double* b, double* x,
                          double tolerance, int32_t max_iterations, int32_t* iterations, double* error)
{
    return guarded(ctx, [&] {
        if (n <= 0 || nnz < 0 || !row_ptr || !col || !val || !b || !x) throw status_error(NEON_BAD_ARGUMENT, "bad CSR arguments");
        if (solver != NEON_SOLVER_CG && solver != NEON_SOLVER_BICGSTAB) throw status_error(NEON_BAD_ARGUMENT, "unknown solver");
        if (criterion != NEON_STOP_RELATIVE && criterion != NEON_STOP_ABSOLUTE_PRECONDITIONED)
            throw status_error(NEON_BAD_ARGUMENT, "unknown stopping rule");
        if (row_ptr[0] != 0 || row_ptr[n] != nnz) throw status_error(NEON_BAD_ARGUMENT, "row_ptr does not span nnz");
        for (int32_t r = 0; r < n; ++r)
            if (row_ptr[r + 1] < row_ptr[r]) throw status_error(NEON_BAD_ARGUMENT, "row_ptr not monotone");
        for (int64_t p = 0; p < nnz; ++p)
            if (col[p] < 0 || col[p] >= n) throw status_error(NEON_BAD_ARGUMENT, "column index out of range");
        hipStream_t const s = ctx->stream;
        device_array<int32_t> rp, ci;
        device_array<double> v, bd, xd, minv;
        device_array<float> v32;
        rp.resize(size_t(n) + 1);
        ci.resize(std::max<int64_t>(nnz, 1));
        v.resize(std::max<int64_t>(nnz, 1));
        bd.resize(n);
        xd.resize(n);
        minv.resize(n);
        NEON_CHECK_HIP(hipMemcpyAsync(rp.data(), row_ptr, (size_t(n) + 1) * sizeof(int32_t), hipMemcpyHostToDevice, s));
        NEON_CHECK_HIP(hipMemcpyAsync(ci.data(), col, nnz * sizeof(int32_t), hipMemcpyHostToDevice, s));
        NEON_CHECK_HIP(hipMemcpyAsync(v.data(), val, nnz * sizeof(double), hipMemcpyHostToDevice, s));
        NEON_CHECK_HIP(hipMemcpyAsync(bd.data(), b, n * sizeof(double), hipMemcpyHostToDevice, s));
        linear_operator A;
        A.n = n;
        A.node_block = false;
        A.rp = rp.data();
        A.ci = ci.data();
        A.lanes_per_row = lanes_for(double(nnz) / n);
        A.v64 = v.data();
        solve_result r;
        if (precision == NEON_PRECISION_FP64)
        {
            ctx->engine->jacobi64(A, minv.data(), criterion == NEON_STOP_RELATIVE ? 1 : 0);
            r = solver == NEON_SOLVER_CG
                    ? ctx->engine->cg64(A, bd.data(), xd.data(), minv.data(), criterion, tolerance, max_iterations)
                    : ctx->engine->bicgstab64(A, bd.data(), xd.data(), minv.data(), criterion, tolerance, max_iterations);
        }
        else
        {
            if (solver != NEON_SOLVER_CG) throw status_error(NEON_BAD_ARGUMENT, "mixed precision supports CG only");
            v32.resize(std::max<int64_t>(nnz, 1));
            convert_to_float(s, nnz, v.data(), v32.data());
            A.v64 = nullptr;
            A.v32 = v32.data();
            r = ctx->engine->cg_mixed(A, bd.data(), xd.data(), criterion, tolerance, max_iterations);
        }
        NEON_CHECK_HIP(hipMemcpyAsync(x, xd.data(), n * sizeof(double), hipMemcpyDeviceToHost, s));
        NEON_CHECK_HIP(hipStreamSynchronize(s));
        if (iterations) *iterations = r.iterations;
        if (error) *error = r.error;
        return r.status;
    });
}

int neon_krylov_solve_csr_noalloc(neon_context* ctx, int solver, int criterion, int32_t n, int64_t nnz,
                                  const int32_t* row_ptr, const int32_t* col, const double* val, const double* b,
                                  double* x, double tolerance, int32_t max_iterations, int32_t* iterations,
                                  double* error)
{
    return guarded(ctx, [&] {
        if (n <= 0 || nnz < 0 || !row_ptr || !col || !val || !b || !x) throw status_error(NEON_BAD_ARGUMENT, "bad CSR arguments");
        if (n > kNoAllocRows)
            throw status_error(NEON_BAD_ARGUMENT, "the allocation-free solve takes at most " + std::to_string(kNoAllocRows)
                                                      + " rows (" + std::to_string(n) + " given)");
        if (nnz > INT32_MAX) throw status_error(NEON_BAD_ARGUMENT, "nnz exceeds 32-bit row pointers");
        if (solver != NEON_SOLVER_CG && solver != NEON_SOLVER_BICGSTAB) throw status_error(NEON_BAD_ARGUMENT, "unknown solver");
        if (criterion != NEON_STOP_RELATIVE && criterion != NEON_STOP_ABSOLUTE_PRECONDITIONED)
            throw status_error(NEON_BAD_ARGUMENT, "unknown stopping rule");
        hipStream_t const s = ctx->stream;
        double* const base = scratch_address(ctx->device);
        double* const flag = base;
        double* const vec = base + kScratchHeader;
        clear_flag_kernel<<<1, 64, 0, s>>>(flag);
        int const grid = int(std::min<int64_t>(1024, std::max<int64_t>(1, (std::max<int64_t>(nnz, n) + 255) / 256)));
        check_csr_kernel<<<grid, 256, 0, s>>>(n, nnz, row_ptr, col, flag);
        NEON_CHECK_HIP(hipGetLastError());
        if (ctx->engine->read_scalar(flag) != 0.0)
            throw status_error(NEON_BAD_ARGUMENT, "malformed CSR arrays (row pointers or column indices)");
        linear_operator A;
        A.n = n;
        A.node_block = false;
        A.rp = row_ptr;
        A.ci = col;
        A.lanes_per_row = lanes_for(double(nnz) / n);
        A.v64 = val;
        double* minv = vec;
        krylov_work w;
        double** slots[8] = {&w.r, &w.p, &w.Ap, &w.r0, &w.v, &w.s, &w.t, &w.p_hat};
        for (int k = 0; k < 8; ++k) *slots[k] = vec + (k + 1) * kNoAllocRows;
        // BiCGSTAB's eighth vector: CG's Ap slot is free there
        w.s_hat = w.Ap;
        ctx->engine->jacobi64(A, minv, criterion == NEON_STOP_RELATIVE ? 1 : 0);
        solve_result const r =
            solver == NEON_SOLVER_CG
                ? ctx->engine->cg64(A, b, x, minv, criterion, tolerance, max_iterations, &w)
                : ctx->engine->bicgstab64(A, b, x, minv, criterion, tolerance, max_iterations, &w);
        NEON_CHECK_HIP(hipStreamSynchronize(s));
        if (iterations) *iterations = r.iterations;
        if (error) *error = r.error;
        return r.status;
    });
}

int neon_mesh_create(neon_context* ctx, int32_t n_nodes, const double* X, int32_t n_elem, const int32_t* conn,
                     int32_t model, const double* params, int32_t tangent_mode)
{
    return guarded(ctx, [&] {
        delete ctx->mesh;
        ctx->mesh = nullptr;
        ctx->mesh = new device_mesh(ctx->stream, n_nodes, X, n_elem, conn, model, params, tangent_mode);
        ctx->mesh->prof = &ctx->profiler;
        return NEON_OK;
    });
}

namespace
{
// constitutive_model::is_finite_deformation / is_symmetric and the `names` each model adds
// (src/constitutive/mechanics/solid/*.cpp constructors), by model id
struct model_traits
{
    int32_t finite_deformation, symmetric;
    const char* names;
};
bool traits_of(int32_t model, model_traits& t)
{
    switch (model & 0xff)
    {
    case NEON_MODEL_NEOHOOKE: t = {1, 1, ""}; return true; // compressible_neohooke.hpp:39
    case NEON_MODEL_J2:                                       // small_strain_J2_plasticity.hpp:32
        t = {0, 1, "effective_plastic_strain;linearised_plastic_strain;linearised_strain;von_mises_stress"};
        return true;
    case NEON_MODEL_LINEAR_ELASTIC: t = {0, 1, "linearised_strain;von_mises_stress"}; return true;
    case NEON_MODEL_MICROSPHERE_GAUSSIAN:
    case NEON_MODEL_MICROSPHERE_LANGEVIN:
    case NEON_MODEL_MICROSPHERE_NONAFFINE: t = {1, 1, ""}; return true;
    case NEON_MODEL_MICROSPHERE_AGEING:
        t = {1, 1, "active_segments;active_shear_modulus;inactive_segments;inactive_shear_modulus;reduction_factor"};
        return true;
    case NEON_MODEL_J2_DAMAGE: // small_strain_J2_plasticity_damage.hpp:42-44, .cpp:21-29
        t = {0, 0,
             "back_stress;damage;effective_plastic_strain;energy_release_rate;kinematic_hardening;"
             "linearised_plastic_strain;linearised_strain;von_mises_stress"};
        return true;
    default: return false;
    }
}
} // namespace

int neon_model_traits(int32_t model, int32_t* finite_deformation, int32_t* symmetric)
{
    model_traits t{};
    if (!traits_of(model, t)) return NEON_BAD_ARGUMENT;
    if (finite_deformation) *finite_deformation = t.finite_deformation;
    if (symmetric) *symmetric = t.symmetric;
    return NEON_OK;
}

const char* neon_model_variable_names(int32_t model)
{
    model_traits t{};
    return traits_of(model, t) ? t.names : nullptr;
}

int neon_mesh_model(neon_context* ctx, int32_t* model)
{
    return guarded(ctx, [&] {
        need_mesh(ctx);
        if (model) *model = ctx->mesh->model();
        return NEON_OK;
    });
}

int neon_mesh_set_parameters(neon_context* ctx, const double* params)
{
    return guarded(ctx, [&] {
        need_mesh(ctx);
        if (!params) throw status_error(NEON_BAD_ARGUMENT, "null parameters");
        ctx->mesh->set_parameters(params);
        return NEON_OK;
    });
}

int neon_mesh_pattern(neon_context* ctx, int64_t* nnz, int32_t* row_ptr, int32_t* col)
{
    return guarded(ctx, [&] {
        need_mesh(ctx);
        if (nnz) *nnz = ctx->mesh->nnz();
        ctx->mesh->pattern(row_ptr, col);
        return NEON_OK;
    });
}

int neon_mesh_set_displacement(neon_context* ctx, const double* u)
{
    return guarded(ctx, [&] {
        need_mesh(ctx);
        if (!u) throw status_error(NEON_BAD_ARGUMENT, "null displacement");
        ctx->mesh->set_displacement_host(u);
        return NEON_OK;
    });
}

int neon_mesh_displacement(neon_context* ctx, double* u)
{
    return guarded(ctx, [&] {
        need_mesh(ctx);
        if (!u) throw status_error(NEON_BAD_ARGUMENT, "null displacement");
        device_mesh& m = *ctx->mesh;
        NEON_CHECK_HIP(hipMemcpyAsync(u, m.u(), m.dofs() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        NEON_CHECK_HIP(hipStreamSynchronize(ctx->stream));
        return NEON_OK;
    });
}

int neon_mesh_update_internal_variables(neon_context* ctx, double time_step)
{
    return guarded(ctx, [&] {
        need_mesh(ctx);
        return ctx->mesh->update_internal_variables(ctx->mesh->u(), time_step);
    });
}

int neon_mesh_set_ageing(neon_context* ctx, double scission, double recombination)
{
    return guarded(ctx, [&] {
        need_mesh(ctx);
        ctx->mesh->set_ageing(scission, recombination);
        return NEON_OK;
    });
}

int neon_mesh_set_damage(neon_context* ctx, const double* constants)
{
    return guarded(ctx, [&] {
        need_mesh(ctx);
        if (!constants) throw status_error(NEON_BAD_ARGUMENT, "null constants");
        ctx->mesh->set_damage(constants);
        return NEON_OK;
    });
}

int neon_mesh_damage_state(neon_context* ctx, double* back_stress, double* damage, double* energy)
{
    return guarded(ctx, [&] {
        need_mesh(ctx);
        ctx->mesh->download_damage_state(back_stress, damage, energy);
        return NEON_OK;
    });
}

int neon_mesh_ageing_state(neon_context* ctx, double* state)
{
    return guarded(ctx, [&] {
        need_mesh(ctx);
        if (!state) throw status_error(NEON_BAD_ARGUMENT, "null pointer");
        ctx->mesh->download_ageing_state(state);
        return NEON_OK;
    });
}

int neon_mesh_save_internal_variables(neon_context* ctx, int32_t converged)
{
    return guarded(ctx, [&] {
        need_mesh(ctx);
        ctx->mesh->save_internal_variables(converged != 0);
        NEON_CHECK_HIP(hipStreamSynchronize(ctx->stream));
        return NEON_OK;
    });
}

int neon_mesh_qp_state(neon_context* ctx, double* sigma, double* D, double* detF, double* eps_p, double* ebar, double* vm)
{
    return guarded(ctx, [&] {
        need_mesh(ctx);
        ctx->mesh->download_qp_state(sigma, D, detF, eps_p, ebar, vm);
        return NEON_OK;
    });
}

int neon_mesh_assemble(neon_context* ctx, int32_t precision)
{
    return guarded(ctx, [&] {
        need_mesh(ctx);
        if (precision != NEON_PRECISION_FP64 && precision != NEON_PRECISION_MIXED)
            throw status_error(NEON_BAD_ARGUMENT, "unknown precision");
        ctx->mesh->assemble(ctx->mesh->u(), precision);
        return NEON_OK;
    });
}

int neon_mesh_download(neon_context* ctx, double* val, double* f_int)
{
    return guarded(ctx, [&] {
        need_mesh(ctx);
        ctx->mesh->download(val, f_int, 0);
        return NEON_OK;
    });
}

int neon_mesh_enforce_dirichlet(neon_context* ctx, int32_t n_fixed, const int32_t* fixed, double* b)
{
    return guarded(ctx, [&] {
        need_mesh(ctx);
        device_mesh& m = *ctx->mesh;
        if (n_fixed < 0 || (n_fixed > 0 && !fixed)) throw status_error(NEON_BAD_ARGUMENT, "bad fixed dof list");
        for (int32_t i = 0; i < n_fixed; ++i)
            if (fixed[i] < 0 || fixed[i] >= m.dofs()) throw status_error(NEON_BAD_ARGUMENT, "fixed dof out of range");
        if (!m.has_k64() && !m.has_k32()) throw status_error(NEON_BAD_ARGUMENT, "no tangent assembled");
        device_array<int32_t> f;
        f.resize(std::max(1, n_fixed));
        if (n_fixed > 0)
            NEON_CHECK_HIP(hipMemcpyAsync(f.data(), fixed, n_fixed * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
        m.enforce_dirichlet(n_fixed, f.data(), nullptr, m.has_k64() ? NEON_PRECISION_FP64 : NEON_PRECISION_MIXED);
        NEON_CHECK_HIP(hipStreamSynchronize(ctx->stream));
        if (b)
            for (int32_t i = 0; i < n_fixed; ++i) b[fixed[i]] = 0.0;
        return NEON_OK;
    });
}

int neon_mesh_solve(neon_context* ctx, int solver, int criterion, double tolerance, int32_t max_iterations,
                    const double* b, double* x, int32_t* iterations, double* error)
{
    return guarded(ctx, [&] {
        need_mesh(ctx);
        device_mesh& m = *ctx->mesh;
        if (!b || !x) throw status_error(NEON_BAD_ARGUMENT, "null vector");
        int64_t const n = m.dofs();
        device_array<double> bd, xd, minv;
        bd.resize(n);
        xd.resize(n);
        minv.resize(n);
        NEON_CHECK_HIP(hipMemcpyAsync(bd.data(), b, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        int const precision = m.has_k64() ? NEON_PRECISION_FP64 : NEON_PRECISION_MIXED;
        solve_result const r = device_solve(ctx, solver, criterion, precision, tolerance, max_iterations, bd.data(),
                                            xd.data(), minv.data());
        NEON_CHECK_HIP(hipMemcpyAsync(x, xd.data(), n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        NEON_CHECK_HIP(hipStreamSynchronize(ctx->stream));
        if (iterations) *iterations = r.iterations;
        if (error) *error = r.error;
        return r.status;
    });
}

int neon_mesh_spmv_config(neon_context* ctx, int32_t lanes, int32_t flags, int32_t blocks_fp64, int32_t blocks_fp32)
{
    return guarded(ctx, [&] {
        need_mesh(ctx);
        if ((lanes != 16 && lanes != 32)
            || (flags & ~(NEON_SPMV_NONTEMPORAL | NEON_SPMV_SLICED | NEON_SPMV_INDEX32 | NEON_SPMV_SYMMETRIC))
            || blocks_fp64 < 0 || blocks_fp32 < 0)
            throw status_error(NEON_BAD_ARGUMENT, "bad SpMV configuration");
        spmv_config& c = ctx->mesh->spmv;
        c.lanes = lanes;
        c.flags = flags;
        c.blocks64 = blocks_fp64;
        c.blocks32 = blocks_fp32;
        return NEON_OK;
    });
}

int neon_mesh_symmetric_info(neon_context* ctx, int32_t* usable, int32_t* offsets, int32_t* slots)
{
    return guarded(ctx, [&] {
        need_mesh(ctx);
        if (usable) *usable = ctx->mesh->symmetric_usable() ? 1 : 0;
        if (offsets) *offsets = ctx->mesh->symmetric_offsets();
        if (slots) *slots = ctx->mesh->symmetric_slots();
        return NEON_OK;
    });
}

int neon_mesh_sliced_info(neon_context* ctx, int32_t* slices, int64_t* steps, int64_t* shared_steps, int32_t* index16)
{
    return guarded(ctx, [&] {
        need_mesh(ctx);
        if (slices) *slices = ctx->mesh->sliced_slices();
        if (steps) *steps = ctx->mesh->sliced_steps();
        if (shared_steps) *shared_steps = ctx->mesh->sliced_shared_steps();
        if (index16) *index16 = ctx->mesh->sliced_index16() ? 1 : 0;
        return NEON_OK;
    });
}

int neon_newton_solve(neon_context* ctx, const neon_newton_options* options, int32_t n_bc, const neon_dirichlet* bcs,
                      double* u_out, neon_newton_record* records, int32_t record_capacity, int32_t* n_records,
                      neon_newton_timing* timing)
{
    return guarded(ctx, [&] {
        need_mesh(ctx);
        if (!options || !u_out || n_bc < 0 || (n_bc > 0 && !bcs)) throw status_error(NEON_BAD_ARGUMENT, "bad arguments");
        if (options->precision != NEON_PRECISION_FP64 && options->precision != NEON_PRECISION_MIXED)
            throw status_error(NEON_BAD_ARGUMENT, "unknown precision");
        return newton_solve(ctx, options, n_bc, bcs, 0, nullptr, u_out, records, record_capacity, n_records, timing);
    });
}

int neon_newton_solve_loaded(neon_context* ctx, const neon_newton_options* options, int32_t n_bc,
                             const neon_dirichlet* bcs, int32_t n_loads, const neon_neumann* loads, double* u_out,
                             neon_newton_record* records, int32_t record_capacity, int32_t* n_records,
                             neon_newton_timing* timing)
{
    return guarded(ctx, [&] {
        need_mesh(ctx);
        if (!options || !u_out || n_bc < 0 || (n_bc > 0 && !bcs) || n_loads < 0 || (n_loads > 0 && !loads))
            throw status_error(NEON_BAD_ARGUMENT, "bad arguments");
        if (options->precision != NEON_PRECISION_FP64 && options->precision != NEON_PRECISION_MIXED)
            throw status_error(NEON_BAD_ARGUMENT, "unknown precision");
        return newton_solve(ctx, options, n_bc, bcs, n_loads, loads, u_out, records, record_capacity, n_records,
                            timing);
    });
}

int neon_latin_solve(neon_context* ctx, const neon_newton_options* options, double search_direction, int32_t n_bc,
                     const neon_dirichlet* bcs, int32_t n_loads, const neon_neumann* loads, double* u_out,
                     neon_newton_record* records, int32_t record_capacity, int32_t* n_records,
                     neon_newton_timing* timing)
{
    return guarded(ctx, [&] {
        need_mesh(ctx);
        if (!options || !u_out || n_bc < 0 || (n_bc > 0 && !bcs) || n_loads < 0 || (n_loads > 0 && !loads))
            throw status_error(NEON_BAD_ARGUMENT, "bad arguments");
        if (options->precision != NEON_PRECISION_FP64 && options->precision != NEON_PRECISION_MIXED)
            throw status_error(NEON_BAD_ARGUMENT, "unknown precision");
        if (!std::isfinite(search_direction)) throw status_error(NEON_BAD_ARGUMENT, "search direction not finite");
        if (ctx->mesh->tangent_mode() != NEON_TANGENT_REFERENCE)
            throw status_error(NEON_BAD_ARGUMENT, "the LATIN iterations use neon's tangent (NEON_TANGENT_REFERENCE)");
        return newton_solve(ctx, options, n_bc, bcs, n_loads, loads, u_out, records, record_capacity, n_records,
                            timing, true, search_direction);
    });
}

int neon_adjoint_gradient(neon_context* ctx, const double* dJdu, int32_t n_fixed, const int32_t* fixed,
                          double tolerance, int32_t max_iterations, double* grad, int32_t* iterations)
{
    return guarded(ctx, [&] {
        need_mesh(ctx);
        device_mesh& m = *ctx->mesh;
        if (!dJdu || !grad || n_fixed < 0 || (n_fixed > 0 && !fixed) || max_iterations <= 0 || !(tolerance > 0.0))
            throw status_error(NEON_BAD_ARGUMENT, "bad adjoint arguments");
        if (m.tangent_mode() != NEON_TANGENT_CONSISTENT)
            throw status_error(NEON_BAD_ARGUMENT, "the adjoint needs the consistent tangent (exact dR/du)");
        int64_t const n = m.dofs();
        if (m.model() != NEON_MODEL_NEOHOOKE && m.model() != NEON_MODEL_J2)
            throw status_error(NEON_BAD_ARGUMENT, "the adjoint gradient supports the neo-Hooke and J2 models");
        for (int32_t i = 0; i < n_fixed; ++i)
            if (fixed[i] < 0 || fixed[i] >= n) throw status_error(NEON_BAD_ARGUMENT, "fixed dof out of range");
        hipStream_t const s = ctx->stream;
        device_array<double> b, lambda, minv, g;
        device_array<int32_t> fx;
        b.resize(n);
        lambda.resize(n);
        minv.resize(n);
        g.resize(4ull * m.elements());
        NEON_CHECK_HIP(hipMemcpyAsync(b.data(), dJdu, n * sizeof(double), hipMemcpyHostToDevice, s));
        if (n_fixed > 0)
        {
            fx.resize(n_fixed);
            NEON_CHECK_HIP(hipMemcpyAsync(fx.data(), fixed, n_fixed * sizeof(int32_t), hipMemcpyHostToDevice, s));
        }
        if (m.model() == NEON_MODEL_J2)
        {
            int const r = j2_adjoint(ctx, b.data(), n_fixed, fx.data(), tolerance, max_iterations, g.data(), iterations);
            if (r != NEON_OK) return r;
            NEON_CHECK_HIP(hipMemcpyAsync(grad, g.data(), 4ull * m.elements() * sizeof(double), hipMemcpyDeviceToHost, s));
            NEON_CHECK_HIP(hipStreamSynchronize(s));
            return NEON_OK;
        }
        // K(u) at the held state, with the constrained rows and columns of the forward solve
        int const st = m.update_internal_variables(m.u());
        if (st != NEON_OK) return st;
        m.assemble(m.u(), NEON_PRECISION_FP64);
        m.enforce_dirichlet(n_fixed, fx.data(), b.data(), NEON_PRECISION_FP64);
        // K^T lambda = dJ/du (K symmetric): the forward Jacobi-PCG, relative stopping rule
        solve_result const r = device_solve(ctx, NEON_SOLVER_CG, NEON_STOP_RELATIVE, NEON_PRECISION_FP64, tolerance,
                                            max_iterations, b.data(), lambda.data(), minv.data());
        if (iterations) *iterations = r.iterations;
        if (r.status != NEON_OK) return r.status;
        m.parameter_gradient(m.u(), lambda.data(), g.data());
        NEON_CHECK_HIP(hipMemcpyAsync(grad, g.data(), 4ull * m.elements() * sizeof(double), hipMemcpyDeviceToHost, s));
        NEON_CHECK_HIP(hipStreamSynchronize(s));
        return NEON_OK;
    });
}

int neon_natural_frequency(neon_context* ctx, const double* density, int32_t n_fixed, const int32_t* fixed,
                           int32_t n_values, int32_t spectrum, double tolerance, int32_t max_basis, int32_t block,
                           double* eigenvalues, double* eigenvectors, int32_t* solves)
{
    return guarded(ctx, [&] {
        need_mesh(ctx);
        device_mesh& m = *ctx->mesh;
        if (!density || !eigenvalues || (n_fixed > 0 && !fixed) || n_fixed < 0 || max_basis < 0 || block < 0)
            throw status_error(NEON_BAD_ARGUMENT, "null pointer or negative count");
        for (int64_t e = 0; e < m.elements(); ++e)
            if (!(density[e] > 0.0)) throw status_error(NEON_BAD_ARGUMENT, "density must be positive");
        for (int32_t i = 0; i < n_fixed; ++i)
            if (fixed[i] < 0 || fixed[i] >= m.dofs()) throw status_error(NEON_BAD_ARGUMENT, "fixed dof out of range");
        return natural_frequency(ctx, density, n_fixed, fixed, n_values, spectrum, tolerance, max_basis, block,
                                 eigenvalues, eigenvectors, solves);
    });
}

int neon_mesh_mass(neon_context* ctx, const double* density, double* val)
{
    return guarded(ctx, [&] {
        need_mesh(ctx);
        if (!density || !val) throw status_error(NEON_BAD_ARGUMENT, "null pointer");
        ctx->mesh->assemble_mass(density);
        ctx->mesh->download_mass(val);
        return NEON_OK;
    });
}

} // extern "C"
