// Host-side engine of the MI355X implicit Newton step: device buffers, the
// device-resident mesh/constitutive state/tangent, and the Krylov drivers.  Everything
// here is called by the C ABI in capi.cpp and by the Newton driver in newton.cpp.
#pragma once

#include "common.hpp"
#include "neon_hip.h"

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace neon_amd
{
struct krylov_state;

template <typename T>
class device_array
{
public:
    device_array() = default;
    device_array(device_array const&) = delete;
    device_array& operator=(device_array const&) = delete;
    ~device_array() { release(); }

    void resize(size_t n)
    {
        if (n == count) return;
        release();
        if (n)
        {
            hipError_t const e = hipMalloc(reinterpret_cast<void**>(&ptr), n * sizeof(T));
            if (e != hipSuccess)
            {
                ptr = nullptr;
                throw status_error(e == hipErrorOutOfMemory ? NEON_OOM : NEON_HIP_ERROR,
                                   "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed");
            }
        }
        count = n;
    }
    void release()
    {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        count = 0;
    }
    T* data() const { return ptr; }
    size_t size() const { return count; }

private:
    T* ptr = nullptr;
    size_t count = 0;
};

// Event pairs around selected launches on one stream; elapsed times are summed per kernel
// class once the stream has been synchronised.
class kernel_profiler
{
public:
    ~kernel_profiler();
    bool enabled = false;
    // returns a token (< 0 when not recording)
    int begin(hipStream_t s, int cls);
    void end(hipStream_t s, int token);
    // accumulate pairs [0, pairs) whose launch ordinal < valid_launches (all if < 0), then clear
    void collect(int64_t valid_launches = -1);
    void read(int cls, bool reset, int64_t* count, double* ms);
    void discard() { pairs_ = 0; }
    void add(int cls, double ms)
    {
        if (!enabled) return;
        count_[cls] += 1;
        ms_[cls] += ms;
    }

private:
    static constexpr int kMaxPairs = 16384;
    std::vector<hipEvent_t> events_;
    std::vector<int> cls_;
    int pairs_ = 0;
    int64_t count_[NEON_KERNEL_CLASSES] = {};
    double ms_[NEON_KERNEL_CLASSES] = {};
};

struct solve_result
{
    int status = NEON_OK;
    int32_t iterations = 0;
    double error = 0.0;
    int64_t spmv = 0;
    int64_t refinements = 0;
};

struct sym_matrix; // krylov_kernels.hpp

// y = A x for a square matrix held on the device in one of two index layouts.
struct linear_operator
{
    int64_t n = 0;
    bool node_block = false;
    // node-block layout
    int32_t nodes = 0;
    const int32_t* bptr = nullptr;
    const int32_t* bcol = nullptr;
    // generic CSR layout
    const int32_t* rp = nullptr;
    const int32_t* ci = nullptr;
    int lanes_per_row = 8;
    // node-block SpMV shape: lanes per block row of the CSR-order kernel (16 or 32),
    // nontemporal value loads, optional workgroup count
    int nb_lanes = 32;
    bool nb_nt = false;
    int blocks_override = 0;
    // values
    const double* v64 = nullptr;
    const float* v32 = nullptr;
    // sliced copy of the node-block values (sell_matrix layout); the SpMV uses it when set
    bool sell = false;
    int32_t slices = 0;
    const int32_t* sptr = nullptr;
    const int32_t* snode = nullptr;
    const int32_t* slane = nullptr;
    const int32_t* scode = nullptr;
    const int32_t* sexc = nullptr;
    const int16_t* soff16 = nullptr; // 16-bit per-lane offsets (used when set)
    const int32_t* soff32 = nullptr;
    const double* s64 = nullptr;
    const float* s32 = nullptr;
    // symmetric sliced copy (sym_matrix layout: diagonal and upper blocks); the SpMV uses it
    // when set, with its values at s64 / s32 and `slices` slices
    bool sym = false;
    int32_t sym_slots = 0;
    int64_t sym_records = 0;
    const int32_t* sym_sched = nullptr;
    const int32_t* sym_rdiag = nullptr;
    const int32_t* sym_eptr = nullptr;
    const int32_t* sym_ent = nullptr; // 4 int32 per entry
    const int32_t* sym_rslice = nullptr;
    const int32_t* sym_roff = nullptr;

    int spmv_blocks() const;
    // FP64: y = A x (values v64); with part, also block partials of x.y (or w.y if w).
    // gate (with st): 1 runs only while st->refine is set, 2 only while it is clear
    void apply64(hipStream_t s, const double* x, double* y, double* part, const krylov_state* st, int gate = 0) const;
    // FP32 values/vectors, FP32 accumulation, FP64 partials of x.y
    void apply32(hipStream_t s, const float* x, float* y, double* part, const krylov_state* st) const;
    // r = b - A x in FP64 arithmetic from the stored values (v64 if set, else v32)
    // (skipped when st is given and the state is done, or, with gate, no reliable update is due)
    void residual64(hipStream_t s, const double* b, const double* x, double* r, const krylov_state* st = nullptr,
                    int gate = 0) const;
};

// Device work vectors of n entries each for the FP64 solvers, supplied by the caller instead of
// the engine's own (cg64 uses r, p, Ap; bicgstab64 r, p, r0, v, s, t, p_hat, s_hat).
struct krylov_work
{
    double *r = nullptr, *p = nullptr, *Ap = nullptr;
    double *r0 = nullptr, *v = nullptr, *s = nullptr, *t = nullptr, *p_hat = nullptr, *s_hat = nullptr;
};

class krylov_engine
{
public:
    explicit krylov_engine(hipStream_t s);
    ~krylov_engine();

    // Jacobi inverse of the operator's diagonal (FP64 values).
    void jacobi64(const linear_operator& A, double* minv, int eigen_rule);

    // work: the caller's work vectors (nothing allocated), else the engine's own (grown to n)
    solve_result cg64(const linear_operator& A, const double* b, double* x, const double* minv, int criterion,
                      double tol, int32_t max_it, const krylov_work* work = nullptr);
    solve_result bicgstab64(const linear_operator& A, const double* b, double* x, const double* minv,
                            int criterion, double tol, int32_t max_it, const krylov_work* work = nullptr);
    // copy one device double to the host through the engine's pinned word (synchronises)
    double read_scalar(const double* d);
    // FP32 inner CG with FP64 reliable updates r = b - A x (relative rule only).
    solve_result cg_mixed(const linear_operator& A, const double* b, double* x, int criterion, double tol,
                          int32_t max_it);
    // same, for a node-block operator whose Jacobi inverse (FP64, from the FP32 diagonal) is given;
    // scaled: the Jacobi-scaled tangent D^-1/2 K D^-1/2 (minv positive, same layout as A), on
    // which the FP32 recurrence then runs without a preconditioner vector (A still gives the
    // FP64 residuals)
    solve_result cg_mixed_nb(const linear_operator& A, const double* b, double* x, const double* minv, double tol,
                             int32_t max_it, const linear_operator* scaled = nullptr);
    // every entry positive and finite (synchronises)
    bool all_positive(int64_t n, const double* v);

    double dot(int64_t n, const double* a, const double* b);

    hipStream_t stream() const { return s_; }
    int vec_blocks(int64_t n) const;
    kernel_profiler* prof = nullptr;
    // iterations between convergence polls from the host
    int poll_every = 16;
    // reliable-update trigger of the mixed-precision CG: ||r32|| < delta ||r at last update||
#ifndef NEON_RELIABLE_DELTA
#define NEON_RELIABLE_DELTA 0.1
#endif
#ifndef NEON_REFINE_EVERY
#define NEON_REFINE_EVERY 8
#endif
    double reliable_delta = NEON_RELIABLE_DELTA;
    // the mixed CG tests the trigger (and, through the true residual, convergence) every this
    // many iterations
    int refine_every = NEON_REFINE_EVERY;

private:
    void ensure(int64_t n);
    void ensure_mixed(int64_t n);
    solve_result cg_mixed_run(const linear_operator& A, const double* b, double* x, double tol, int32_t max_it, int nb,
                              int ns, double* part_pAp, double* part_rr, double* part_rz, double* part_rr_t,
                              double* part_rz_t, krylov_state* st, float* minv, const linear_operator* scaled_op,
                              const double* minv64_scaled);
    hipStream_t s_;
    device_array<double> part_[6];
    device_array<krylov_state> state_;
    krylov_state* host_state_ = nullptr;
    // poll-ahead ring of run_batches: two pinned state snapshots and their events
    krylov_state* host_ring_ = nullptr;
    hipEvent_t ring_events_[2] = {nullptr, nullptr};
    template <typename Batch>
    void run_batches(krylov_state* st, int64_t& launched, Batch&& batch);
    double* host_scalar_ = nullptr;
    device_array<double> scalar_;
    int64_t n_ = 0, n_mixed_ = 0;
    device_array<double> r_, p_, Ap_;
    device_array<double> r0_, v_, s_vec_, t_, ph_, sh_;
    device_array<float> r32_, p32_, Ap32_, x32_, minv32_;
    device_array<double> x64_, r64_;
};

// Tangent value arrays extend this many entries past nnz (zeroed): SpMV kernels that stream a
// segment in 16-byte vectors may read up to one vector past its end.
constexpr int64_t kValuePad = 16;

// Shape of the node-block SpMV used by the Krylov solvers on a mesh tangent (the defaults
// are the fastest measured at 6.4M dofs on MI355X): lanes per block row, NEON_SPMV_* flags,
// workgroups for FP64 and FP32 values.
// constitutive_model::is_symmetric of each model id (only the J2 damage model's tangent is
// unsymmetric: small_strain_J2_plasticity_damage.hpp:44)
inline bool model_is_symmetric(int32_t model) { return (model & 0xff) != NEON_MODEL_J2_DAMAGE; }

struct spmv_config
{
    int lanes = 32;
    int flags = NEON_SPMV_SLICED | NEON_SPMV_NONTEMPORAL | NEON_SPMV_SYMMETRIC;
    int blocks64 = 0;
    int blocks32 = 0;
};

// Device-resident mesh: geometry, constitutive state, pattern, tangent and f_int.
class device_mesh
{
public:
    device_mesh(hipStream_t s, int32_t n_nodes, const double* X, int32_t n_elem, const int32_t* conn, int32_t model,
                const double* params, int32_t tangent_mode);

    int32_t nodes() const { return N_; }
    int32_t elements() const { return E_; }
    int64_t dofs() const { return 3ll * N_; }
    int64_t nnz() const { return nnz_; }
    int32_t model() const { return model_; }
    int32_t colours() const { return int32_t(colour_offsets_.size()) - 1; }

    void set_parameters(const double* params);
    void set_displacement_host(const double* u);
    // u (device) -> state; returns NEON_OK or the recoverable code.  dt is the time step of
    // the path-dependent models' rate equations: the step size at a load step's first update,
    // 0 at its Newton iterates
    int update_internal_variables(const double* u_dev, double dt = 0.0);
    // LATIN (latin_matrix.hpp): committed stress <- 0 (latin_reset) or <- current stress
    // (latin_commit); f <- the LATIN force at the current state with the committed
    // displacement u_old (device) and search direction s
    void latin_reset();
    void latin_commit();
    void latin_force(const double* u_dev, const double* u_old_dev, double s, double* f_dev);
    // ageing microsphere: the material's scission and recombination probabilities
    void set_ageing(double scission, double recombination);
    // ageing microsphere: network state, 11 per QP in QP order (host)
    void download_ageing_state(double* state);
    // J2 damage: the material constants; back stress (9 per QP, row-major), damage and energy
    // release rate (host, any may be null)
    void set_damage(const double* constants);
    void download_damage_state(double* back_stress, double* damage, double* energy);
    void save_internal_variables(bool converged);
    void assemble(const double* u_dev, int precision);
    void enforce_dirichlet(int32_t n_fixed, const int32_t* fixed_dev, double* b_dev, int precision);
    void pattern(int32_t* row_ptr, int32_t* col) const;
    void download(double* val, double* f_int, int precision_hint);
    void download_qp_state(double* sigma, double* D, double* detF, double* eps_p, double* ebar, double* vm);
    // operator over the assembled tangent (FP64 values, or FP32 values for mixed)
    linear_operator op(int precision) const;
    linear_operator op_base(int precision) const;
    // points A's SpMV at the sliced copy of the tangent (packed here when the tangent changed
    // since the last copy); index32: 32-bit per-lane offsets even where 16 bits fit
    void attach_sliced(linear_operator& A, int precision, bool index32 = false) const;
    // whether every per-lane offset of the sliced copy fits 16 bits
    bool sliced_index16() const { return soff16_ok_; }
    int32_t sliced_slices() const { return slices_; }
    int64_t sliced_steps() const { return sell_steps_; }
    // steps whose 64 lanes share one column offset
    int64_t sliced_shared_steps() const { return shared_steps_; }
    // the sliced copy is packed through LDS: nodes of at most 170 neighbours
    bool sliced_fits() const;
    // points A's SpMV at the symmetric sliced copy (diagonal and upper blocks; packed here when
    // the tangent changed since the last copy); only where symmetric_usable()
    // scale_minv (FP32 values only): attach D^-1/2 K D^-1/2 with D^-1 = scale_minv (positive)
    void attach_symmetric(linear_operator& A, int precision, const double* scale_minv = nullptr) const;
    // the solvers' operator is the symmetric copy (flags ask for it and the mesh allows it)
    bool symmetric_selected() const;
    // the symmetric FP32 copy of the Jacobi-scaled tangent (minv: the positive Jacobi inverse of
    // the FP32 diagonal), for the mixed solver
    linear_operator op_scaled(const double* minv) const;
    // a symmetric model and a pattern for which the symmetric copy pays (every slice needs at
    // most kMaxSymEntries offsets and the copy holds at most 3/4 of the full copy's steps)
    bool symmetric_usable() const;
    // the largest number of offsets (entries) of a slice (0 when the pattern does not qualify),
    // the symmetric copy's wave slots and records
    int32_t symmetric_offsets() const { return sym_ok_ ? sym_max_entries_ : 0; }
    int32_t symmetric_slots() const { return sym_slots_; }
    int64_t symmetric_records() const { return sym_records_; }
    // Jacobi inverse from the node-block diagonal
    void jacobi(double* minv, int eigen_rule, int precision);
    // consistent mass (density per element, host) in the pattern of K, FP64; Dirichlet rows and
    // columns zeroed with the diagonal kept; operator, Jacobi inverse and download
    void assemble_mass(const double* density);
    void enforce_dirichlet_mass(int32_t n_fixed, const int32_t* fixed_dev);
    linear_operator op_mass() const;
    void jacobi_mass(double* minv);
    void download_mass(double* val);
    // adjoint: grad (4 per element) = -lambda^T d f_int / d params at displacement u
    void parameter_gradient(const double* u_dev, const double* lambda_dev, double* grad_dev);

    // load-history checkpoints (J2 adjoint): u_n and the committed history h_{n-1} of every
    // converged step, kept on the device
    void clear_trajectory();
    void record_step();
    int32_t trajectory_steps() const { return int32_t(traj_u_.size()); }
    // u <- u_n, committed history <- h_{n-1}
    void restore_step(int32_t step);
    // J2 adjoint pieces at the state of restore_step + update_internal_variables:
    // rhs += B^T (dh/d eps)^T mu;  mu <- history adjoint of the step, grad += parameter adjoint
    void j2_adjoint_rhs(double* rhs);
    void j2_adjoint_history(const double* lambda, double* grad);
    void j2_adjoint_begin();
    int32_t tangent_mode() const { return tangent_mode_; }

    double* u() { return u_.data(); }
    double* f_int() { return fint_.data(); }
    hipStream_t stream() const { return s_; }
    bool has_k64() const { return k64_valid_; }
    bool has_k32() const { return k32_valid_; }

    // last assemble time split (ms) measured with events
    float last_assemble_ms = 0.0f;
    spmv_config spmv;
    // assembly grid: workgroups (one wavefront each) per compute unit and colour
    int assembly_blocks_per_cu = 128;
    int32_t max_degree() const { return max_degree_; }
    kernel_profiler* prof = nullptr;

private:
    void build_sliced_pattern();
    void build_symmetric_pattern();
    sym_matrix sym_layout() const;
    hipStream_t s_;
    int32_t N_, E_, model_, tangent_mode_;
    int64_t nnz_ = 0;
    int32_t max_degree_ = 0;
    std::vector<int32_t> bptr_h_, bcol_h_;
    std::vector<int32_t> colour_offsets_;
    device_array<double> X_, u_, params_, fint_;
    device_array<int32_t> conn_, elems_, bptr_, bcol_, self_slot_;
    device_array<uint8_t> slots_;
    device_array<double> k64_, m64_;
    device_array<float> k32_;
    bool m64_valid_ = false;
    bool k64_valid_ = false, k32_valid_ = false;
    // sliced layout of the pattern (sell_matrix) and the sliced copies of the tangent
    int32_t slices_ = 0;
    int64_t sell_steps_ = 0, shared_steps_ = 0;
    device_array<int32_t> sptr_, snode_, slane_, scode_, sexc_, soff32_;
    device_array<int16_t> soff16_;
    bool soff16_ok_ = false;
    mutable device_array<double> s64_;
    mutable device_array<float> s32_;
    mutable bool s64_fresh_ = false, s32_fresh_ = false;
    // symmetric sliced layout: per-slice entries, records, the wave-slot schedule and the copies
    bool sym_ok_ = false;
    int32_t sym_slots_ = 0, sym_max_entries_ = 0;
    int64_t sym_records_ = 0;
    device_array<int32_t> sym_sched_, sym_rdiag_, sym_eptr_, sym_ent_, sym_rslice_, sym_roff_;
    mutable device_array<double> y64_;
    mutable device_array<float> y32_, y32s_;
    mutable bool y64_fresh_ = false, y32_fresh_ = false, y32s_fresh_ = false;
    device_array<double> sigma_, coef_, detF_, eps_p_, ebar_, vm_, eps_p_old_, ebar_old_;
    device_array<double> age_, age_old_; // ageing microsphere: 11 x Q network state (SoA)
    device_array<double> sigma_old_;     // LATIN: committed stress (6 x Q)
    double scission_ = 0.0, recombination_ = 0.0, last_dt_ = 0.0;
    // J2 damage: back stress 6, damage 1, energy release rate 1 per QP (SoA) and the committed
    // back stress and damage; the material's constants (C, gamma, s_p, n_p, s_d, n_d)
    device_array<double> dmg_state_, dmg_state_old_;
    double damage_[6] = {1.0, 1.0, 1.0, 1.0, 1.0, 1.0};
    device_array<uint32_t> status_;
    std::vector<std::unique_ptr<device_array<double>>> traj_u_, traj_ep_, traj_eb_;
    device_array<double> mu_ep_, mu_eb_;
    uint32_t* status_host_ = nullptr;
};

void upload_hex8_tables();
void convert_to_float(hipStream_t s, int64_t n, const double* x, float* y);
void fill_ones(hipStream_t s, int64_t n, double* x);
void fill_ones(hipStream_t s, int64_t n, float* x);
} // namespace neon_amd

struct neon_context
{
    int device = 0;
    hipStream_t stream = nullptr;
    std::string error;
    neon_amd::krylov_engine* engine = nullptr;
    neon_amd::device_mesh* mesh = nullptr;
    neon_amd::kernel_profiler profiler;
};
