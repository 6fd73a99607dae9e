// Natural frequencies on the device-resident tangent: the generalised eigenproblem
// K x = lambda M x of natural_frequency_matrix::solve (src/assembler/mechanics/
// natural_frequency_matrix.hpp:54-68) with the consistent mass, both matrices with neon's
// homogeneous Dirichlet treatment (rows and columns zeroed, diagonal kept;
// src/assembler/homogeneous_dirichlet.hpp:84-123), for the eigenvalues at the lower or upper end
// of the spectrum (arpack.cpp:40-62 with "SM" / "LM").
//
// The solver is a block Krylov method with full M-orthogonalisation and Rayleigh-Ritz, built on
// the device SpMV and Jacobi-PCG:
//   lower end: OP = K^-1 M (shift-invert at 0; each application one PCG solve with K), Ritz
//              values theta = 1 / lambda;
//   upper end: OP = M^-1 K (one SpMV with K and one PCG solve with M), theta = lambda.
// OP is self-adjoint in the M inner product.  The basis Q (and P = M Q) is M-orthonormal; a
// block of `block` random vectors starts it, and processing basis vector j appends the
// M-orthogonalised w_j = OP q_j (classical Gram-Schmidt, twice).  The coefficients
// C(i, j) = p_i . w_j give the projected matrix H = Q^T M OP Q on the processed vectors and, for a
// Ritz pair (theta, s) of H, the exact M-norm of its residual OP y - theta y in basis
// coordinates, C s - theta [s; 0].  Blocks find multiple eigenvalues (symmetric meshes) that a
// single Krylov vector cannot.
#include "engine.hpp"
#include "newton.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

namespace neon_amd
{
namespace
{
inline int blocks_for(int64_t n) { return int(std::min<int64_t>((n + kBlock - 1) / kBlock, 8192)); }

// deterministic start vectors: a hash of (index, column) mapped to [-1, 1)
__global__ __launch_bounds__(kBlock) void random_fill_kernel(int64_t n, uint64_t column, double* __restrict__ v)
{
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBlock)
    {
        uint64_t h = uint64_t(i) * 0x9E3779B97F4A7C15ull + (column + 1) * 0xD1B54A32D192ED03ull;
        h ^= h >> 31;
        h *= 0x94D049BB133111EBull;
        h ^= h >> 29;
        h *= 0xBF58476D1CE4E5B9ull;
        h ^= h >> 32;
        v[i] = double(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
}

// part[c * gridDim.x + b] = partial dot of V_c (column c, leading dimension ld) with w
__global__ __launch_bounds__(kBlock) void multi_dot_kernel(int64_t n, const double* __restrict__ V, int64_t ld,
                                                           const double* __restrict__ w, double* __restrict__ part)
{
    const double* v = V + int64_t(blockIdx.y) * ld;
    double s = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBlock)
        s += v[i] * w[i];
    s = block_sum<kBlock>(s);
    if (threadIdx.x == 0) part[int64_t(blockIdx.y) * gridDim.x + blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void multi_reduce_kernel(const double* __restrict__ part, int nb,
                                                              double* __restrict__ out)
{
    double const s = reduce_partials<kBlock>(part + int64_t(blockIdx.x) * nb, nb);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// w = beta w + sum_c coef[c] V_c over nvec columns
__global__ __launch_bounds__(kBlock) void combine_kernel(int64_t n, const double* __restrict__ V, int64_t ld, int nvec,
                                                         const double* __restrict__ coef, double beta,
                                                         double* __restrict__ w)
{
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBlock)
    {
        double acc = beta == 0.0 ? 0.0 : beta * w[i];
        for (int c = 0; c < nvec; ++c) acc += coef[c] * V[int64_t(c) * ld + i];
        w[i] = acc;
    }
}

__global__ __launch_bounds__(kBlock) void scale_copy_kernel(int64_t n, double a, const double* __restrict__ x,
                                                            double* __restrict__ y)
{
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBlock)
        y[i] = a * x[i];
}

// Cyclic Jacobi eigen-decomposition of a symmetric k x k matrix (row-major A, destroyed):
// eigenvalues in d, eigenvectors in the columns of V.
void jacobi_eigen(int k, std::vector<double>& A, std::vector<double>& d, std::vector<double>& V)
{
    V.assign(size_t(k) * k, 0.0);
    for (int i = 0; i < k; ++i) V[size_t(i) * k + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep)
    {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < k; ++i)
        {
            diag += A[size_t(i) * k + i] * A[size_t(i) * k + i];
            for (int j = i + 1; j < k; ++j) off += A[size_t(i) * k + j] * A[size_t(i) * k + j];
        }
        if (off <= 1e-32 * diag || off == 0.0) break;
        for (int p = 0; p < k; ++p)
            for (int q = p + 1; q < k; ++q)
            {
                double const apq = A[size_t(p) * k + q];
                if (apq == 0.0) continue;
                double const app = A[size_t(p) * k + p], aqq = A[size_t(q) * k + q];
                double const theta = (aqq - app) / (2.0 * apq);
                double const t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double const c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int r = 0; r < k; ++r) // columns p, q
                {
                    double const arp = A[size_t(r) * k + p], arq = A[size_t(r) * k + q];
                    A[size_t(r) * k + p] = c * arp - s * arq;
                    A[size_t(r) * k + q] = s * arp + c * arq;
                }
                for (int r = 0; r < k; ++r) // rows p, q
                {
                    double const apr = A[size_t(p) * k + r], aqr = A[size_t(q) * k + r];
                    A[size_t(p) * k + r] = c * apr - s * aqr;
                    A[size_t(q) * k + r] = s * apr + c * aqr;
                }
                for (int r = 0; r < k; ++r)
                {
                    double const vrp = V[size_t(r) * k + p], vrq = V[size_t(r) * k + q];
                    V[size_t(r) * k + p] = c * vrp - s * vrq;
                    V[size_t(r) * k + q] = s * vrp + c * vrq;
                }
            }
    }
    d.resize(k);
    for (int i = 0; i < k; ++i) d[i] = A[size_t(i) * k + i];
}

class basis
{
public:
    basis(hipStream_t s, int64_t n, int cap) : s_(s), n_(n), cap_(cap)
    {
        Q_.resize(size_t(cap) * n);
        P_.resize(size_t(cap) * n);
        nb_ = std::max(1, std::min(blocks_for(n), 1024));
        part_.resize(size_t(cap) * nb_);
        coef_.resize(cap);
        host_coef_.resize(cap);
    }
    double* q(int i) const { return Q_.data() + int64_t(i) * n_; }
    double* p(int i) const { return P_.data() + int64_t(i) * n_; }
    int size() const { return size_; }
    int capacity() const { return cap_; }

    // c = P[0:k]^T w (device coef_, host copy returned)
    const std::vector<double>& project(const double* w, int k)
    {
        if (k == 0) return host_coef_;
        multi_dot_kernel<<<dim3(nb_, k), kBlock, 0, s_>>>(n_, P_.data(), n_, w, part_.data());
        multi_reduce_kernel<<<k, kBlock, 0, s_>>>(part_.data(), nb_, coef_.data());
        NEON_CHECK_HIP(hipGetLastError());
        NEON_CHECK_HIP(hipMemcpyAsync(host_coef_.data(), coef_.data(), k * sizeof(double), hipMemcpyDeviceToHost, s_));
        NEON_CHECK_HIP(hipStreamSynchronize(s_));
        return host_coef_;
    }
    // w -= Q[0:k] c, c the last projection (on the device)
    void subtract(double* w, int k)
    {
        if (k == 0) return;
        scale_copy_kernel<<<1, kBlock, 0, s_>>>(k, -1.0, coef_.data(), coef_.data());
        combine_kernel<<<blocks_for(n_), kBlock, 0, s_>>>(n_, Q_.data(), n_, k, coef_.data(), 1.0, w);
        NEON_CHECK_HIP(hipGetLastError());
    }
    // M-orthogonalise w against the basis (twice) and append it when its remaining M-norm is
    // above rel times its M-norm before; coefficients (first + second pass, then the remainder's
    // norm) in c.  Returns the remainder's M-norm.
    double append(const linear_operator& M, double* w, double* z, std::vector<double>& c, double rel,
                  krylov_engine& k)
    {
        int const m = size_;
        c.assign(size_t(m) + 1, 0.0);
        for (int pass = 0; pass < 2; ++pass)
        {
            std::vector<double> const& h = project(w, m);
            for (int i = 0; i < m; ++i) c[i] += h[i];
            subtract(w, m);
        }
        M.apply64(s_, w, z, nullptr, nullptr);
        double const nrm = std::sqrt(std::max(0.0, k.dot(n_, w, z)));
        c[m] = nrm;
        double total = 0.0;
        for (double v : c) total += v * v;
        if (nrm > rel * std::sqrt(total) && m < cap_)
        {
            scale_copy_kernel<<<blocks_for(n_), kBlock, 0, s_>>>(n_, 1.0 / nrm, w, q(m));
            scale_copy_kernel<<<blocks_for(n_), kBlock, 0, s_>>>(n_, 1.0 / nrm, z, p(m));
            NEON_CHECK_HIP(hipGetLastError());
            ++size_;
        }
        return nrm;
    }
    // y = Q[0:k] s (host s)
    void combine(const double* s_host, int k, double* y)
    {
        NEON_CHECK_HIP(hipMemcpyAsync(coef_.data(), s_host, k * sizeof(double), hipMemcpyHostToDevice, s_));
        combine_kernel<<<blocks_for(n_), kBlock, 0, s_>>>(n_, Q_.data(), n_, k, coef_.data(), 0.0, y);
        NEON_CHECK_HIP(hipGetLastError());
        NEON_CHECK_HIP(hipStreamSynchronize(s_));
    }

private:
    hipStream_t s_;
    int64_t n_;
    int cap_, size_ = 0, nb_ = 1;
    device_array<double> Q_, P_, part_, coef_;
    std::vector<double> host_coef_;
};
} // namespace

int natural_frequency(neon_context* ctx, const double* density, int32_t n_fixed, const int32_t* fixed, int32_t nev,
                      int32_t spectrum, double tol, int32_t max_basis, int32_t block, double* eigenvalues,
                      double* eigenvectors, int32_t* solves)
{
    device_mesh& m = *ctx->mesh;
    krylov_engine& k = *ctx->engine;
    hipStream_t const s = ctx->stream;
    int64_t const n = m.dofs();
    if (nev <= 0 || nev > n) throw status_error(NEON_BAD_ARGUMENT, "The number of requested eigenvalues must be positive");
    if (spectrum != NEON_SPECTRUM_LOWER && spectrum != NEON_SPECTRUM_UPPER)
        throw status_error(NEON_BAD_ARGUMENT, "only the lower and upper spectrum are supported");
    if (!(tol > 0.0)) throw status_error(NEON_BAD_ARGUMENT, "the eigenvalue tolerance must be positive");
    if (block <= 0) block = 4;
    if (max_basis <= 0) max_basis = std::max(4 * nev, nev + 40);
    int const cap = int(std::min<int64_t>(n, std::max<int64_t>(max_basis, int64_t(nev) + block)));
    block = std::min(block, cap);

    // K at the held state, M; neon's Dirichlet treatment on both
    int const st = m.update_internal_variables(m.u());
    if (st != NEON_OK) return st;
    m.assemble(m.u(), NEON_PRECISION_FP64);
    m.assemble_mass(density);
    device_array<int32_t> fx;
    if (n_fixed > 0)
    {
        fx.resize(n_fixed);
        NEON_CHECK_HIP(hipMemcpyAsync(fx.data(), fixed, n_fixed * sizeof(int32_t), hipMemcpyHostToDevice, s));
    }
    m.enforce_dirichlet(n_fixed, fx.data(), nullptr, NEON_PRECISION_FP64);
    m.enforce_dirichlet_mass(n_fixed, fx.data());
    linear_operator const K = m.op(NEON_PRECISION_FP64);
    linear_operator const M = m.op_mass();
    device_array<double> minv_k, minv_m, w, z, t;
    minv_k.resize(n);
    minv_m.resize(n);
    w.resize(n);
    z.resize(n);
    t.resize(n);
    m.jacobi(minv_k.data(), 1, NEON_PRECISION_FP64);
    m.jacobi_mass(minv_m.data());
    bool const lower = spectrum == NEON_SPECTRUM_LOWER;
    // the inner solves' error enters the Ritz residuals at about cond(K) times their tolerance,
    // so they are held well below the eigen tolerance (eigenvalues converge with the square of
    // the residual)
    double const inner_tol = std::clamp(1e-6 * tol, 1e-14, 1e-10);
    int32_t const inner_max = int32_t(std::min<int64_t>(std::max<int64_t>(10 * n, 1000), 1000000));
    int32_t n_solves = 0;

    basis Bq(s, n, cap);
    std::vector<double> c;
    // coefficient columns of the processed vectors over the basis
    std::vector<std::vector<double>> C;
    uint64_t next_random = 0;
    auto add_random = [&]() {
        for (int attempt = 0; attempt < 8; ++attempt)
        {
            random_fill_kernel<<<blocks_for(n), kBlock, 0, s>>>(n, next_random++, w.data());
            NEON_CHECK_HIP(hipGetLastError());
            int const before = Bq.size();
            Bq.append(M, w.data(), z.data(), c, 1e-8, k);
            if (Bq.size() > before) return true;
        }
        return false;
    };
    for (int b = 0; b < block; ++b)
        if (!add_random()) break;

    std::vector<double> theta, S, A;
    std::vector<int> order;
    bool converged = false;
    // Rayleigh-Ritz costs O(k^3) on the host: analyse at k = nev and then at growing intervals
    int next_check = nev;
    int j = 0;
    for (; j < Bq.size(); ++j)
    {
        // w = OP q_j
        if (lower)
        {
            solve_result const r = k.cg64(K, Bq.p(j), w.data(), minv_k.data(), NEON_STOP_RELATIVE, inner_tol, inner_max);
            ++n_solves;
            if (r.status != NEON_OK) return r.status;
        }
        else
        {
            K.apply64(s, Bq.q(j), t.data(), nullptr, nullptr);
            solve_result const r = k.cg64(M, t.data(), w.data(), minv_m.data(), NEON_STOP_RELATIVE, inner_tol, inner_max);
            ++n_solves;
            if (r.status != NEON_OK) return r.status;
        }
        int const before = Bq.size();
        Bq.append(M, w.data(), z.data(), c, 1e-10, k);
        C.push_back(c);
        // a vanishing remainder: the space is invariant for this direction; continue from a
        // fresh random vector so that the block keeps its width
        if (Bq.size() == before && Bq.size() < Bq.capacity()) add_random();

        int const kp = j + 1;
        bool const check = kp >= next_check || Bq.size() == kp || kp == Bq.capacity();
        if (!check) continue;
        next_check = kp + std::max(block, kp / 8);
        // Rayleigh-Ritz on the processed vectors
        A.assign(size_t(kp) * kp, 0.0);
        for (int a = 0; a < kp; ++a)
            for (int b = 0; b < kp; ++b)
            {
                // H(a, b) = p_a . w_b; a vector appended after b was processed has no
                // coefficient in w_b, so the symmetric entry stands for it
                bool const has_ab = a < int(C[b].size()), has_ba = b < int(C[a].size());
                double const hab = has_ab ? C[b][a] : 0.0, hba = has_ba ? C[a][b] : 0.0;
                A[size_t(a) * kp + b] = has_ab && has_ba ? 0.5 * (hab + hba) : hab + hba;
            }
        jacobi_eigen(kp, A, theta, S);
        order.resize(kp);
        std::iota(order.begin(), order.end(), 0);
        std::sort(order.begin(), order.end(), [&](int x, int y) { return theta[x] > theta[y]; });
        converged = true;
        for (int e = 0; e < nev && converged; ++e)
        {
            int const col = order[e];
            size_t const rows = C.back().size();
            double r2 = 0.0;
            for (size_t i = 0; i < rows; ++i)
            {
                double v = 0.0;
                for (int b = 0; b < kp; ++b)
                    if (i < C[b].size()) v += C[b][i] * S[size_t(b) * kp + col];
                if (int(i) < kp) v -= theta[col] * S[i * kp + col];
                r2 += v * v;
            }
            converged = std::sqrt(r2) <= tol * std::fabs(theta[col]);
        }
        if (converged) break;
    }
    if (solves) *solves = n_solves;
    if (!converged) throw status_error(NEON_NONCONVERGED, "Eigenvalues did not converge");

    int const kp = j + 1;
    // ascending eigenvalues: lower end lambda = 1 / theta (largest theta first), upper end the
    // nev largest theta in increasing order
    std::vector<int> pick(order.begin(), order.begin() + nev);
    if (!lower) std::reverse(pick.begin(), pick.end());
    std::vector<double> sv(kp);
    for (int e = 0; e < nev; ++e)
    {
        int const col = pick[e];
        eigenvalues[e] = lower ? 1.0 / theta[col] : theta[col];
        if (eigenvectors)
        {
            for (int b = 0; b < kp; ++b) sv[b] = S[size_t(b) * kp + col];
            Bq.combine(sv.data(), kp, w.data());
            NEON_CHECK_HIP(hipMemcpyAsync(eigenvectors + int64_t(e) * n, w.data(), n * sizeof(double),
                                          hipMemcpyDeviceToHost, s));
            NEON_CHECK_HIP(hipStreamSynchronize(s));
        }
    }
    return NEON_OK;
}
} // namespace neon_amd
