// Device kernels of the hex8 element path: per-quadrature-point deformation measures and
// constitutive update, element tangent + internal force with scatter into the node-block
// CSR, and Dirichlet row/column zeroing.
//
// Reference behaviour (paths relative to the neon repository root):
//   deformation measures   src/mesh/mechanics/solid/submesh.cpp:210-269
//   neo-Hooke              src/constitutive/mechanics/solid/compressible_neohooke.cpp:17-90
//   J2 radial return       src/constitutive/mechanics/solid/small_strain_J2_plasticity.cpp:30-163
//   linear elasticity      src/constitutive/mechanics/solid/isotropic_linear_elasticity.cpp:27-67
//   microsphere models     src/constitutive/mechanics/solid/affine_microsphere.cpp:23-111,
//                          gaussian_affine_microsphere.cpp:23-83, nonaffine_microsphere.cpp:26-178,
//                          src/constitutive/mechanics/detail/microsphere.hpp,
//                          src/quadrature/sphere/unit_sphere_quadrature.cpp:10-90
//   element K, f_int       src/mesh/mechanics/solid/submesh.cpp:66-160
//   Dirichlet zeroing      src/assembler/mechanics/static_matrix.hpp:279-320
#pragma once

#include "common.hpp"
#include "neon_hip.h"

namespace neon_amd
{
// Per-QP stored data, structure-of-arrays with stride Q = 8 E (index k*Q + qp):
//   sigma : 6  Cauchy stress, kinetic Voigt order (11, 22, 33, 23, 13, 12)
//   coef  : 9  tangent coefficients (A, Bt, C, n11, n22, n33, n23, n13, n12) such that
//              D = A m m^T + Bt diag(2,2,2,1,1,1) + C n n^T  (m = (1,1,1,0,0,0),
//              n in kinetic Voigt order); neo-Hooke: A = lambda, Bt = mu - lambda ln J, C = 0;
//              J2: A = lambda + c1/3, Bt = G - c1/2, C = c2 with c1, c2 of the algorithmic
//              tangent (small_strain_J2_plasticity.cpp:95-106).
//              Microsphere models (general symmetric D): 21 coefficients, the upper triangle
//              of D row by row, D(i,j) at kUpper[i][j].
//   detF  : 1
//   J2 history: eps_p 6 (tensor components 11,22,33,23,13,12), ebar 1, vm 1; committed
//   copies eps_p_old, ebar_old.  Linear elasticity keeps vm.
enum
{
    COEF_A = 0,
    COEF_B = 1,
    COEF_C = 2,
    COEF_N = 3
};

constexpr int kFullCoefs = 21;
__host__ __device__ constexpr int upper_index(int i, int j)
{
    return i <= j ? i * 6 - i * (i - 1) / 2 + (j - i) : j * 6 - j * (j - 1) / 2 + (i - j);
}

__host__ __device__ constexpr bool is_microsphere(int model)
{
    return (model & 0xff) >= NEON_MODEL_MICROSPHERE_GAUSSIAN && (model & 0xff) <= NEON_MODEL_MICROSPHERE_AGEING;
}
// constitutive_model::is_finite_deformation
__host__ __device__ constexpr bool is_finite_model(int model)
{
    return model == NEON_MODEL_NEOHOOKE || is_microsphere(model);
}

// Unit-sphere rules: [0] 21 directions (BO21), [1] 33 directions (BO33); w and r (x, y, z)
struct sphere_tables
{
    int n[2];
    double w[2][33];
    double r[2][33][3];
};

__constant__ sphere_tables c_sphere;

struct hex8_tables
{
    double dN[8][8][3]; // [qp][node][xi]
    double w[8];
    double N[8][8]; // [qp][node]
};

__constant__ hex8_tables c_hex8;

struct qp_state
{
    double* sigma;
    double* coef;
    double* detF;
    double* eps_p;
    double* ebar;
    double* vm;
    int64_t Q;
    const double* eps_p_old; // committed history (consistent J2 returns from it)
    const double* ebar_old;
    double* age;             // ageing microsphere: network state, 11 per QP (index k Q + qp)
    const double* age_old;   // its committed copy
};

// Ageing microsphere: the material's scission and recombination probabilities, this update's
// time step (0 inside a load step) and the last non-zero one (a model-wide value in neon).
struct ageing_args
{
    double scission, recombination, dt, last_dt;
};

struct element_params
{
    const double* p; // 4 per element: bulk, shear, yield, hardening
};

__device__ inline double det3(const double* a)
{
    return a[0] * (a[4] * a[8] - a[7] * a[5]) - a[3] * (a[1] * a[8] - a[7] * a[2])
           + a[6] * (a[1] * a[5] - a[4] * a[2]);
}

__device__ inline void inv3(const double* a, double* r)
{
    double const c00 = a[4] * a[8] - a[5] * a[7];
    double const c01 = a[5] * a[6] - a[3] * a[8];
    double const c02 = a[3] * a[7] - a[4] * a[6];
    double const det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    double const id = 1.0 / det;
    r[0] = c00 * id;
    r[1] = (a[2] * a[7] - a[1] * a[8]) * id;
    r[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    r[3] = c01 * id;
    r[4] = (a[0] * a[8] - a[2] * a[6]) * id;
    r[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    r[6] = c02 * id;
    r[7] = (a[1] * a[6] - a[0] * a[7]) * id;
    r[8] = (a[0] * a[4] - a[1] * a[3]) * id;
}

// ------------------------------------------------------------ microsphere update
// Kinetic Voigt components (11, 22, 33, 23, 13, 12) of a (x) b + symmetric
__device__ inline void sym_dyad(const double* a, double* v)
{
    v[0] = a[0] * a[0];
    v[1] = a[1] * a[1];
    v[2] = a[2] * a[2];
    v[3] = a[1] * a[2];
    v[4] = a[0] * a[2];
    v[5] = a[0] * a[1];
}

// c += s v v^T on the 21 stored entries
__device__ inline void add_outer21(double* c, double s, const double* v)
{
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = i; j < 6; ++j) c[upper_index(i, j)] += s * v[i] * v[j];
}

// Per-QP microsphere update (all three models share the volumetric part and the
// projections of affine_microsphere.cpp:50-71):
//   macro stress tau_bar and macro moduli c_bar from sphere sums over the rule's directions r
//   (t = F_bar r, l = |t|; non-affine also n = F_bar^-T r, nu = |n|):
//     Gaussian   tau_bar = 3G sum w t t,                 c_bar = 0
//     Langevin   tau_bar = G sum w pf(l) t t,            c_bar = G sum w (ps(l) - pf(l)) / l^2 tttt
//     non-affine chain stretch L = (sum w l^p)^(1/p), h = sum w l^(p-2) t t,
//                H = (p - 2) sum w l^(p-4) tttt, k = sum w nu^-1 n n,
//                T = sum w (-nu^-3 nnnn + 2 nu^-1 o(n))  (tube exponent q = 1, U = 9):
//                tau_bar = G pf(L) L^(2-p) h - G N U k,
//                c_bar = [G ps(L) L^(2-2p) - (p-1) G pf(L) L^(2-2p)] h h + G pf(L) L^(2-p) H + G N U T
//   with pf = (3N - l^2)/(N - l^2), ps = (l^4 + 3N^2)/(N - l^2)^2 (microsphere.hpp:37-65);
//   p = J U'(J), kappa = J^2 U''(J) for U = K/4 (J^2 - 1) - K/2 ln J;
//   sigma = (p I + dev tau_bar) / J;
//   D = (kappa + p) m m^T - 2p I + 2/3 tr(tau_bar) (I - m m^T / 3) + P c_bar P,
//   I = diag(1,1,1,1/2,1/2,1/2), m = (1,1,1,0,0,0), P the kinetic deviatoric projector: this is
//   neon's P [c_bar + 2/3 tr(tau_bar) I - 2/3 (tau_bar m^T + m tau_bar^T)] P, whose last term the
//   projections annihilate.  The consistent mode adds the exact tangent's
//   -2/3 (dev tau_bar m^T + m dev tau_bar^T) and divides by J.
__device__ inline void microsphere_tangent(const double* tau, const double* cb, double J, double K, bool consistent,
                                           double* sig, double* D);

template <int BASE>
__device__ void microsphere_update(const double* F, double J, const double* prm, int rule, bool consistent,
                                   double* sig, double* D)
{
    double const K = prm[0], G = prm[1], N = prm[2], p = prm[3];
    double const sc = pow(J, -1.0 / 3.0);
    double Fb[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) Fb[k] = sc * F[k];
    double tau[6] = {0, 0, 0, 0, 0, 0};
    double cb[kFullCoefs];
#pragma unroll
    for (int k = 0; k < kFullCoefs; ++k) cb[k] = 0.0;
    int const nd = c_sphere.n[rule];
    if constexpr (BASE == NEON_MODEL_MICROSPHERE_NONAFFINE)
    {
        double Fi[9];
        inv3(Fb, Fi);
        double sum_p = 0.0, h[6] = {0, 0, 0, 0, 0, 0}, kk[6] = {0, 0, 0, 0, 0, 0}, H[kFullCoefs];
#pragma unroll
        for (int k = 0; k < kFullCoefs; ++k) H[k] = 0.0;
        for (int i = 0; i < nd; ++i)
        {
            double const w = c_sphere.w[rule][i];
            double const r0 = c_sphere.r[rule][i][0], r1 = c_sphere.r[rule][i][1], r2 = c_sphere.r[rule][i][2];
            double t[3], n[3];
#pragma unroll
            for (int a = 0; a < 3; ++a)
            {
                t[a] = Fb[a * 3] * r0 + Fb[a * 3 + 1] * r1 + Fb[a * 3 + 2] * r2;
                n[a] = Fi[a] * r0 + Fi[3 + a] * r1 + Fi[6 + a] * r2; // F_bar^-T r
            }
            double const l2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
            double const nu2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
            double const nu = sqrt(nu2);
            double const lp4 = pow(l2, 0.5 * (p - 4.0)); // l^(p-4)
            double vt[6], vn[6];
            sym_dyad(t, vt);
            sym_dyad(n, vn);
            sum_p += lp4 * l2 * l2 * w;
#pragma unroll
            for (int k = 0; k < 6; ++k)
            {
                h[k] += lp4 * l2 * vt[k] * w;
                kk[k] += vn[k] * (w / nu);
            }
            add_outer21(H, lp4 * w, vt);
            // tube moduli q (q - 2) nu^(q-4) nnnn + 2 q nu^(q-2) o(n), q = 1
            add_outer21(cb, -w / (nu2 * nu), vn);
            double const o = 2.0 * w / nu;
            double const n00 = n[0] * n[0], n11 = n[1] * n[1], n22 = n[2] * n[2];
            double const n01 = n[0] * n[1], n02 = n[0] * n[2], n12 = n[1] * n[2];
            cb[upper_index(0, 0)] += o * 2.0 * n00;
            cb[upper_index(0, 4)] += o * n02;
            cb[upper_index(0, 5)] += o * n01;
            cb[upper_index(1, 1)] += o * 2.0 * n11;
            cb[upper_index(1, 3)] += o * n12;
            cb[upper_index(1, 5)] += o * n01;
            cb[upper_index(2, 2)] += o * 2.0 * n22;
            cb[upper_index(2, 3)] += o * n12;
            cb[upper_index(2, 4)] += o * n02;
            cb[upper_index(3, 3)] += o * 0.5 * (n11 + n22);
            cb[upper_index(3, 4)] += o * 0.5 * n01;
            cb[upper_index(3, 5)] += o * 0.5 * n02;
            cb[upper_index(4, 4)] += o * 0.5 * (n00 + n22);
            cb[upper_index(4, 5)] += o * 0.5 * n12;
            cb[upper_index(5, 5)] += o * 0.5 * (n00 + n11);
        }
        double const L = pow(sum_p, 1.0 / p);
        double const L2 = L * L;
        double const pf = (3.0 * N - L2) / (N - L2);
        double const ps = (L2 * L2 + 3.0 * N * N) / ((N - L2) * (N - L2));
        double const kirchhoff = G * pf * L; // micro Kirchhoff force
        double const c_f = kirchhoff * pow(L, 1.0 - p);
        double const c_hh = G * ps * pow(L, 2.0 - 2.0 * p) - (p - 1.0) * kirchhoff * pow(L, 1.0 - 2.0 * p);
        double const tube = G * N * 9.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) tau[k] = c_f * h[k] - tube * kk[k];
#pragma unroll
        for (int k = 0; k < kFullCoefs; ++k) cb[k] = tube * cb[k] + c_f * (p - 2.0) * H[k];
        add_outer21(cb, c_hh, h);
    }
    else
    {
        for (int i = 0; i < nd; ++i)
        {
            double const w = c_sphere.w[rule][i];
            double const r0 = c_sphere.r[rule][i][0], r1 = c_sphere.r[rule][i][1], r2 = c_sphere.r[rule][i][2];
            double t[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) t[a] = Fb[a * 3] * r0 + Fb[a * 3 + 1] * r1 + Fb[a * 3 + 2] * r2;
            double vt[6];
            sym_dyad(t, vt);
            if constexpr (BASE == NEON_MODEL_MICROSPHERE_GAUSSIAN)
            {
#pragma unroll
                for (int k = 0; k < 6; ++k) tau[k] += vt[k] * w;
            }
            else
            {
                double const l2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
                double const pf = (3.0 * N - l2) / (N - l2);
                double const ps = (l2 * l2 + 3.0 * N * N) / ((N - l2) * (N - l2));
#pragma unroll
                for (int k = 0; k < 6; ++k) tau[k] += pf * vt[k] * w;
                add_outer21(cb, (ps - pf) / l2 * w, vt);
            }
        }
        double const s = BASE == NEON_MODEL_MICROSPHERE_GAUSSIAN ? 3.0 * G : G;
#pragma unroll
        for (int k = 0; k < 6; ++k) tau[k] *= s;
#pragma unroll
        for (int k = 0; k < kFullCoefs; ++k) cb[k] *= G;
    }
    microsphere_tangent(tau, cb, J, K, consistent, sig, D);
}

// sigma and D of the microsphere models from the macro stress tau_bar (kinetic Voigt) and the
// macro moduli c_bar (21 upper-triangle entries), as described above microsphere_update.
__device__ inline void microsphere_tangent(const double* tau, const double* cb, double J, double K, bool consistent,
                                           double* sig, double* D)
{
    double const pr = J * (K / 2.0 * (J - 1.0 / J));
    double const kappa = J * J * (K / 2.0 * (1.0 + 1.0 / (J * J)));
    double const tr = tau[0] + tau[1] + tau[2];
    double dev[6] = {tau[0] - tr / 3.0, tau[1] - tr / 3.0, tau[2] - tr / 3.0, tau[3], tau[4], tau[5]};
#pragma unroll
    for (int k = 0; k < 6; ++k) sig[k] = ((k < 3 ? pr : 0.0) + dev[k]) / J;
    // P c_bar P: P = blockdiag(I - 1 1^T / 3, I); subtract the normal block's row / column means
    double c[6][6];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) c[i][j] = cb[upper_index(i, j)];
#pragma unroll
    for (int j = 0; j < 6; ++j)
    {
        double const mean = (c[0][j] + c[1][j] + c[2][j]) / 3.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) c[i][j] -= mean;
    }
#pragma unroll
    for (int i = 0; i < 6; ++i)
    {
        double const mean = (c[i][0] + c[i][1] + c[i][2]) / 3.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) c[i][j] -= mean;
    }
    double const a2 = 2.0 / 3.0 * tr;
    double const inv_j = consistent ? 1.0 / J : 1.0;
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = i; j < 6; ++j)
        {
            bool const nn = i < 3 && j < 3;
            double const Ik = i == j ? (i < 3 ? 1.0 : 0.5) : 0.0;
            double v = c[i][j] + (nn ? kappa + pr - a2 / 3.0 : 0.0) + (a2 - 2.0 * pr) * Ik;
            if (consistent) v -= 2.0 / 3.0 * ((j < 3 ? dev[i] : 0.0) + (i < 3 ? dev[j] : 0.0));
            D[upper_index(i, j)] = v * inv_j;
        }
}

// Ageing microsphere (gaussian_ageing_affine_microsphere.cpp:52-134 with the BAND network of
// micromechanical_elastomer.cpp:55-122): z = (active shear modulus G_a, inactive G_i, reduction
// r, active segments N_a, inactive N_i) advances by one fourth-order Runge-Kutta step over dt when
// dt != 0, with
//   alpha = 1 - (1 - p_r)^(N_i + 1), beta = 1 - (1 - p_s)^N_a, eta = 1 - (1 - p_r)^(N_a + 1),
//   nu = 1 - (1 - p_s)^N_i,
//   G_a' = alpha G_i + G_a (2 eta - beta), G_i' = 2 beta G_a + G_i (nu - 2 alpha),
//   r' = -r (beta + 2 eta), N_a' = (2 alpha G_i N_i - beta N_a G_a - N_a G_a') / G_a,
//   N_i' = (beta N_a G_a - 2 alpha N_i G_i - N_i G_i') / G_i (0 while G_i <= 1e-8);
// the secondary network's intermediate stress S = last_dt F_bar^-1 (c / r) F_bar^-T + S_old with
// the creation rate c = alpha G_i + 4 eta G_a; macro stress r (3 G sum w t t + F_bar S F_bar^T);
// then the Gaussian model's stress and tangent (zero macro moduli).
__device__ inline void band_rates(const double* y, double ps, double pr, double* f)
{
    double const Ga = y[0], Gi = y[1], r = y[2], Na = y[3], Ni = y[4];
    double const alpha = 1.0 - pow(1.0 - pr, Ni + 1.0);
    double const beta = 1.0 - pow(1.0 - ps, Na);
    double const eta = 1.0 - pow(1.0 - pr, Na + 1.0);
    double const nu = 1.0 - pow(1.0 - ps, Ni);
    f[0] = alpha * Gi + Ga * (2.0 * eta - beta);
    f[1] = 2.0 * beta * Ga + Gi * (nu - 2.0 * alpha);
    f[2] = -r * (beta + 2.0 * eta);
    f[3] = (2.0 * alpha * Gi * Ni - beta * Na * Ga - Na * f[0]) / Ga;
    f[4] = Gi > 1.0e-8 ? (beta * Na * Ga - 2.0 * alpha * Ni * Gi - Ni * f[1]) / Gi : 0.0;
}

__device__ void ageing_update(const double* F, double J, const double* prm, int rule, ageing_args ag, double* z,
                              const double* S_old, double* S, double* sig, double* D)
{
    double const K = prm[0], G = prm[1];
    if (ag.dt != 0.0)
    {
        double k1[5], k2[5], k3[5], k4[5], y[5];
        band_rates(z, ag.scission, ag.recombination, k1);
#pragma unroll
        for (int i = 0; i < 5; ++i)
        {
            k1[i] *= ag.dt;
            y[i] = z[i] + k1[i] / 2.0;
        }
        band_rates(y, ag.scission, ag.recombination, k2);
#pragma unroll
        for (int i = 0; i < 5; ++i)
        {
            k2[i] *= ag.dt;
            y[i] = z[i] + k2[i] / 2.0;
        }
        band_rates(y, ag.scission, ag.recombination, k3);
#pragma unroll
        for (int i = 0; i < 5; ++i)
        {
            k3[i] *= ag.dt;
            y[i] = z[i] + k3[i];
        }
        band_rates(y, ag.scission, ag.recombination, k4);
#pragma unroll
        for (int i = 0; i < 5; ++i)
        {
            k4[i] *= ag.dt;
            z[i] = z[i] + (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]) / 6.0;
        }
    }
    double const pr = ag.recombination;
    double const alpha = 1.0 - pow(1.0 - pr, z[4] + 1.0);
    double const eta = 1.0 - pow(1.0 - pr, z[3] + 1.0);
    double const creation = alpha * z[1] + 4.0 * eta * z[0];
    double const sc = pow(J, -1.0 / 3.0);
    double Fb[9], Fi[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) Fb[k] = sc * F[k];
    inv3(Fb, Fi);
    // primary network: 3 G sum w t t
    double tau[6] = {0, 0, 0, 0, 0, 0};
    int const nd = c_sphere.n[rule];
    for (int i = 0; i < nd; ++i)
    {
        double const w = c_sphere.w[rule][i];
        double const r0 = c_sphere.r[rule][i][0], r1 = c_sphere.r[rule][i][1], r2 = c_sphere.r[rule][i][2];
        double t[3], vt[6];
#pragma unroll
        for (int a = 0; a < 3; ++a) t[a] = Fb[a * 3] * r0 + Fb[a * 3 + 1] * r1 + Fb[a * 3 + 2] * r2;
        sym_dyad(t, vt);
#pragma unroll
        for (int k = 0; k < 6; ++k) tau[k] += vt[k] * w;
    }
    // S = a F_bar^-1 F_bar^-T + S_old, a = last_dt c / r; components (11, 22, 33, 23, 13, 12)
    double const a = ag.last_dt * creation / z[2];
    constexpr int pi[6] = {0, 1, 2, 1, 0, 0}, pj[6] = {0, 1, 2, 2, 2, 1};
#pragma unroll
    for (int k = 0; k < 6; ++k)
        S[k] = a * (Fi[pi[k] * 3] * Fi[pj[k] * 3] + Fi[pi[k] * 3 + 1] * Fi[pj[k] * 3 + 1]
                    + Fi[pi[k] * 3 + 2] * Fi[pj[k] * 3 + 2])
               + S_old[k];
    // macro = r (3 G primary + F_bar S F_bar^T)
    double const Sm[9] = {S[0], S[5], S[4], S[5], S[1], S[3], S[4], S[3], S[2]};
    double FS[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) FS[i * 3 + j] = Fb[i * 3] * Sm[j] + Fb[i * 3 + 1] * Sm[3 + j] + Fb[i * 3 + 2] * Sm[6 + j];
    double macro[6];
#pragma unroll
    for (int k = 0; k < 6; ++k)
    {
        int const i = pi[k], j = pj[k];
        double const push = FS[i * 3] * Fb[j * 3] + FS[i * 3 + 1] * Fb[j * 3 + 1] + FS[i * 3 + 2] * Fb[j * 3 + 2];
        macro[k] = z[2] * (3.0 * G * tau[k] + push);
    }
    double cb[kFullCoefs];
#pragma unroll
    for (int k = 0; k < kFullCoefs; ++k) cb[k] = 0.0;
    microsphere_tangent(macro, cb, J, K, false, sig, D);
}

// ---------------------------------------------------- constitutive update per QP
// Deformation measures of quadrature point l of element e (submesh.cpp:210-237): with
// F_0 = X dN_l, H = (x - X) dN_l F_0^-1 (displacement gradient) and F = (x dN_l) F_0^-1, x = X + u.
__device__ inline void qp_kinematics(const int32_t* __restrict__ conn, const double* __restrict__ X,
                                     const double* __restrict__ u, int64_t e, int l, double* H, double* F)
{
    double F0[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Fx[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double dXe[8][3];
    for (int a = 0; a < 8; ++a)
    {
        int64_t const nd = conn[8 * e + a];
        for (int i = 0; i < 3; ++i)
        {
            double const Xi = X[3 * nd + i];
            double const xi = Xi + u[3 * nd + i]; // x = X + u (material_coordinates.cpp:23-27)
            dXe[a][i] = xi - Xi;
            for (int j = 0; j < 3; ++j)
            {
                F0[i * 3 + j] += Xi * c_hex8.dN[l][a][j];
                Fx[i * 3 + j] += xi * c_hex8.dN[l][a][j];
            }
        }
    }
    double F0i[9];
    inv3(F0, F0i);
    for (int k = 0; k < 9; ++k) H[k] = 0.0;
    for (int a = 0; a < 8; ++a)
    {
        double B0[3];
        for (int j = 0; j < 3; ++j)
            B0[j] = c_hex8.dN[l][a][0] * F0i[0 * 3 + j] + c_hex8.dN[l][a][1] * F0i[1 * 3 + j]
                    + c_hex8.dN[l][a][2] * F0i[2 * 3 + j];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) H[i * 3 + j] += dXe[a][i] * B0[j];
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            F[i * 3 + j] = Fx[i * 3 + 0] * F0i[0 * 3 + j] + Fx[i * 3 + 1] * F0i[1 * 3 + j] + Fx[i * 3 + 2] * F0i[2 * 3 + j];
}

// One thread per quadrature point (qp = 8 e + l).
template <int MODEL>
__global__ __launch_bounds__(kBlock) void update_state_kernel(int32_t E, const int32_t* __restrict__ conn,
                                                              const double* __restrict__ X,
                                                              const double* __restrict__ u, element_params prm,
                                                              qp_state st, int tangent_mode, int rule,
                                                              ageing_args ag, uint32_t* __restrict__ status)
{
    int64_t const qp = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (qp >= 8ll * E) return;
    int64_t const e = qp >> 3;
    int const l = int(qp & 7);
    int64_t const Q = st.Q;

    double H[9], F[9];
    qp_kinematics(conn, X, u, e, l, H, F);
    double const J = det3(F);
    st.detF[qp] = J;
    if (signbit(J)) atomicOr(status, STATUS_NEGATIVE_JACOBIAN);

    if constexpr (is_microsphere(MODEL))
    {
        double const pe[4] = {prm.p[4 * e], prm.p[4 * e + 1], prm.p[4 * e + 2], prm.p[4 * e + 3]};
        double sig[6], D[kFullCoefs];
        if constexpr (MODEL == NEON_MODEL_MICROSPHERE_AGEING)
        {
            // network state advanced in place (neon's internal variables), the secondary
            // network's stress from its committed value
            double z[5], S_old[6], S[6];
            for (int k = 0; k < 5; ++k) z[k] = st.age[k * Q + qp];
            for (int k = 0; k < 6; ++k) S_old[k] = st.age_old[(5 + k) * Q + qp];
            ageing_update(F, J, pe, rule, ag, z, S_old, S, sig, D);
            for (int k = 0; k < 5; ++k) st.age[k * Q + qp] = z[k];
            for (int k = 0; k < 6; ++k) st.age[(5 + k) * Q + qp] = S[k];
        }
        else
            microsphere_update<MODEL>(F, J, pe, rule, tangent_mode == NEON_TANGENT_CONSISTENT, sig, D);
        bool finite = true;
        for (int k = 0; k < 6; ++k)
        {
            st.sigma[k * Q + qp] = sig[k];
            finite = finite && isfinite(sig[k]);
        }
        for (int k = 0; k < kFullCoefs; ++k)
        {
            st.coef[k * Q + qp] = D[k];
            finite = finite && isfinite(D[k]);
        }
        if (!finite) atomicOr(status, STATUS_NAN);
        return;
    }
    double const bulk = prm.p[4 * e], shear = prm.p[4 * e + 1];
    double const lambda = bulk - 2.0 * shear / 3.0;
    double sig[6];
    double cA, cB, cC = 0.0, nv[6] = {0, 0, 0, 0, 0, 0};
    auto vonmises = [](const double* s) {
        double const t = (s[0] + s[1] + s[2]) / 3.0;
        double const d0 = s[0] - t, d1 = s[1] - t, d2 = s[2] - t;
        double const nrm = d0 * d0 + d1 * d1 + d2 * d2 + 2.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
        return sqrt(3.0 / 2.0) * sqrt(nrm);
    };
    if constexpr (MODEL == NEON_MODEL_LINEAR_ELASTIC)
    {
        // sigma = lambda tr(eps) I + 2 G eps, eps = sym(H); constant moduli
        double const eps[6] = {H[0], H[4], H[8], 0.5 * (H[5] + H[7]), 0.5 * (H[2] + H[6]), 0.5 * (H[1] + H[3])};
        double const tr = eps[0] + eps[1] + eps[2];
        for (int k = 0; k < 6; ++k) sig[k] = 2.0 * shear * eps[k];
        for (int k = 0; k < 3; ++k) sig[k] += lambda * tr;
        st.vm[qp] = vonmises(sig);
        cA = lambda;
        cB = shear;
    }
    else if constexpr (MODEL == NEON_MODEL_NEOHOOKE)
    {
        // sigma = (lambda ln J I + mu (F F^T - I)) / J ; D: lambda, mu' = mu - lambda ln J
        double const lnJ = log(J);
        double b[6];
        b[0] = F[0] * F[0] + F[1] * F[1] + F[2] * F[2];
        b[1] = F[3] * F[3] + F[4] * F[4] + F[5] * F[5];
        b[2] = F[6] * F[6] + F[7] * F[7] + F[8] * F[8];
        b[3] = F[3] * F[6] + F[4] * F[7] + F[5] * F[8];
        b[4] = F[0] * F[6] + F[1] * F[7] + F[2] * F[8];
        b[5] = F[0] * F[3] + F[1] * F[4] + F[2] * F[5];
        for (int k = 0; k < 3; ++k) sig[k] = (lambda * lnJ + shear * (b[k] - 1.0)) / J;
        for (int k = 3; k < 6; ++k) sig[k] = (shear * b[k]) / J;
        cA = lambda;
        cB = shear - lambda * lnJ;
        if (tangent_mode == NEON_TANGENT_CONSISTENT)
        {
            cA /= J;
            cB /= J;
        }
    }
    else
    {
        // small-strain J2 with linear isotropic hardening, history updated in place
        double const G = shear, sy = prm.p[4 * e + 2], Hh = prm.p[4 * e + 3];
        // neon accumulates the history in place across Newton iterates; the consistent mode
        // returns from the last converged history
        bool const from_old = tangent_mode == NEON_TANGENT_CONSISTENT;
        const double* ep_src = from_old ? st.eps_p_old : st.eps_p;
        double ep[6];
        for (int k = 0; k < 6; ++k) ep[k] = ep_src[k * Q + qp];
        double eb = from_old ? st.ebar_old[qp] : st.ebar[qp];
        if (from_old)
        {
            for (int k = 0; k < 6; ++k) st.eps_p[k * Q + qp] = ep[k];
            st.ebar[qp] = eb;
        }
        // strain = sym(H) in tensor components (11,22,33,23,13,12)
        double eps[6] = {H[0], H[4], H[8], 0.5 * (H[5] + H[7]), 0.5 * (H[2] + H[6]), 0.5 * (H[1] + H[3])};
        double ee[6];
        for (int k = 0; k < 6; ++k) ee[k] = eps[k] - ep[k];
        double const tr = ee[0] + ee[1] + ee[2];
        for (int k = 0; k < 6; ++k) sig[k] = 2.0 * G * ee[k];
        for (int k = 0; k < 3; ++k) sig[k] += lambda * tr;
        double vm = vonmises(sig);
        cA = lambda;
        cB = G;
        double const f0 = (vm - 3.0 * G * 0.0) - (sy + (eb + 0.0) * Hh);
        if (f0 > 0.0)
        {
            double const vm_tr = vm;
            double const t = (sig[0] + sig[1] + sig[2]) / 3.0;
            double dev[6] = {sig[0] - t, sig[1] - t, sig[2] - t, sig[3], sig[4], sig[5]};
            double const nd =
                sqrt(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2] + 2.0 * (dev[3] * dev[3] + dev[4] * dev[4] + dev[5] * dev[5]));
            for (int k = 0; k < 6; ++k) nv[k] = dev[k] / nd;
            double dg = 0.0, f = f0;
            int it = 0;
            while (f > 1.0e-6 && it < 50)
            {
                dg += f / (3.0 * G + Hh);
                f = (vm_tr - 3.0 * G * dg) - (sy + (eb + dg) * Hh);
                ++it;
            }
            if (it == 50) atomicOr(status, STATUS_RETURN_MAP);
            double const s32 = sqrt(3.0 / 2.0);
            for (int k = 0; k < 6; ++k) ep[k] += dg * s32 * nv[k];
            for (int k = 0; k < 6; ++k) sig[k] -= 2.0 * G * dg * s32 * nv[k];
            vm = vonmises(sig);
            eb += dg;
            double const c1 = dg * 6.0 * G * G / vm_tr;
            double const c2 = 6.0 * G * G * (dg / vm_tr - 1.0 / (3.0 * G + Hh));
            cA = lambda + c1 / 3.0;
            cB = G - 0.5 * c1;
            cC = c2;
            for (int k = 0; k < 6; ++k) st.eps_p[k * Q + qp] = ep[k];
            st.ebar[qp] = eb;
        }
        st.vm[qp] = vm;
    }
    bool finite = isfinite(cA) && isfinite(cB);
    for (int k = 0; k < 6; ++k)
    {
        st.sigma[k * Q + qp] = sig[k];
        finite = finite && isfinite(sig[k]);
    }
    if (!finite) atomicOr(status, STATUS_NAN);
    st.coef[COEF_A * Q + qp] = cA;
    st.coef[COEF_B * Q + qp] = cB;
    st.coef[COEF_C * Q + qp] = cC;
    for (int k = 0; k < 6; ++k) st.coef[(COEF_N + k) * Q + qp] = nv[k];
}

// LATIN force (latin_submesh.hpp:62-111): one thread per element of a colour,
//   f_a += int B_a^T (sigma_old + s D (eps - eps_old) - sigma) dV
// on the current configuration x = X + u, B_a = dN_a J_x^-1; eps - eps_old is the symmetric
// part of the displacement gradient's change, sum_a (u_a - u_old_a) (x) dN_a J_X^-1, as a
// kinematic Voigt vector (engineering shears); D from the QP's tangent coefficients (the
// structured form, or the 21 stored entries when FULL); sigma_old the committed stress.
// Elements of one colour share no node, so the scatter needs no atomics.
template <bool FULL>
__global__ __launch_bounds__(64) void latin_force_kernel(const int32_t* __restrict__ elems, int32_t count,
                                                         const int32_t* __restrict__ conn,
                                                         const double* __restrict__ X, const double* __restrict__ u,
                                                         const double* __restrict__ u_old,
                                                         const double* __restrict__ sigma,
                                                         const double* __restrict__ sigma_old,
                                                         const double* __restrict__ coef, int64_t Q, double s,
                                                         double* __restrict__ f)
{
    for (int32_t t = blockIdx.x * 64 + threadIdx.x; t < count; t += gridDim.x * 64)
    {
        int64_t const e = elems[t];
        int32_t nd[8];
        double Xe[8][3], xe[8][3], du[8][3];
        for (int a = 0; a < 8; ++a)
        {
            nd[a] = conn[8 * e + a];
            for (int i = 0; i < 3; ++i)
            {
                int64_t const g = 3ll * nd[a] + i;
                Xe[a][i] = X[g];
                xe[a][i] = X[g] + u[g];
                du[a][i] = u[g] - u_old[g];
            }
        }
        double fe[8][3];
        for (int a = 0; a < 8; ++a)
            for (int i = 0; i < 3; ++i) fe[a][i] = 0.0;
        for (int l = 0; l < 8; ++l)
        {
            double Jx[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, J0[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int a = 0; a < 8; ++a)
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j)
                    {
                        Jx[i * 3 + j] += xe[a][i] * c_hex8.dN[l][a][j];
                        J0[i * 3 + j] += Xe[a][i] * c_hex8.dN[l][a][j];
                    }
            double Jxi[9], J0i[9];
            inv3(Jx, Jxi);
            inv3(J0, J0i);
            double const detJ = det3(Jx);
            double Hd[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int a = 0; a < 8; ++a)
            {
                double B0[3];
                for (int j = 0; j < 3; ++j)
                    B0[j] = c_hex8.dN[l][a][0] * J0i[j] + c_hex8.dN[l][a][1] * J0i[3 + j] + c_hex8.dN[l][a][2] * J0i[6 + j];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) Hd[i * 3 + j] += du[a][i] * B0[j];
            }
            double const ek[6] = {Hd[0], Hd[4], Hd[8], Hd[5] + Hd[7], Hd[2] + Hd[6], Hd[1] + Hd[3]};
            int64_t const qp = 8 * e + l;
            double De[6];
            if constexpr (FULL)
            {
                for (int i = 0; i < 6; ++i)
                {
                    double v = 0.0;
                    for (int j = 0; j < 6; ++j) v += coef[int64_t(upper_index(i, j)) * Q + qp] * ek[j];
                    De[i] = v;
                }
            }
            else
            {
                double const A = coef[COEF_A * Q + qp], Bt = coef[COEF_B * Q + qp], C = coef[COEF_C * Q + qp];
                double n[6], ne = 0.0;
                for (int k = 0; k < 6; ++k)
                {
                    n[k] = coef[(COEF_N + k) * Q + qp];
                    ne += n[k] * ek[k];
                }
                double const tr = ek[0] + ek[1] + ek[2];
                for (int k = 0; k < 6; ++k) De[k] = (k < 3 ? A * tr + 2.0 * Bt * ek[k] : Bt * ek[k]) + C * ne * n[k];
            }
            double d[6];
            for (int k = 0; k < 6; ++k) d[k] = sigma_old[k * Q + qp] + s * De[k] - sigma[k * Q + qp];
            double const S[9] = {d[0], d[5], d[4], d[5], d[1], d[3], d[4], d[3], d[2]};
            double const wd = detJ * c_hex8.w[l];
            for (int a = 0; a < 8; ++a)
            {
                double B[3];
                for (int j = 0; j < 3; ++j)
                    B[j] = c_hex8.dN[l][a][0] * Jxi[j] + c_hex8.dN[l][a][1] * Jxi[3 + j] + c_hex8.dN[l][a][2] * Jxi[6 + j];
                for (int i = 0; i < 3; ++i) fe[a][i] += (B[0] * S[i] + B[1] * S[3 + i] + B[2] * S[6 + i]) * wd;
            }
        }
        for (int a = 0; a < 8; ++a)
            for (int i = 0; i < 3; ++i) f[3ll * nd[a] + i] += fe[a][i];
    }
}

// commit (old <- current) or revert (current <- old) of the J2 history
__global__ __launch_bounds__(kBlock) void history_copy_kernel(int64_t n, const double* __restrict__ src,
                                                              double* __restrict__ dst)
{
    int64_t const i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// Expand the stored coefficients to the 6x6 Voigt tangent (test download).
__global__ __launch_bounds__(kBlock) void expand_tangent_kernel(int64_t Q, const double* __restrict__ coef,
                                                                const double* __restrict__ sigma,
                                                                double* __restrict__ D, double* __restrict__ s9,
                                                                int kind)
{
    // kind 0: structured coefficients; 1: 21 symmetric entries; 2: 36 entries row by row
    int64_t const qp = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (qp >= Q) return;
    bool const full = kind != 0;
    if (D && full)
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j)
                D[36 * qp + 6 * i + j] = coef[int64_t(kind == 2 ? 6 * i + j : upper_index(i, j)) * Q + qp];
    double const A = coef[COEF_A * Q + qp], B = coef[COEF_B * Q + qp], C = coef[COEF_C * Q + qp];
    double n[6];
    for (int k = 0; k < 6; ++k) n[k] = coef[(COEF_N + k) * Q + qp];
    if (D && !full)
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j)
            {
                double v = C * n[i] * n[j];
                if (i < 3 && j < 3) v += A;
                if (i == j) v += (i < 3 ? 2.0 : 1.0) * B;
                D[36 * qp + 6 * i + j] = v;
            }
    if (s9)
    {
        double const s[6] = {sigma[0 * Q + qp], sigma[1 * Q + qp], sigma[2 * Q + qp],
                             sigma[3 * Q + qp], sigma[4 * Q + qp], sigma[5 * Q + qp]};
        double* o = s9 + 9 * qp;
        o[0] = s[0];
        o[4] = s[1];
        o[8] = s[2];
        o[5] = o[7] = s[3];
        o[2] = o[6] = s[4];
        o[1] = o[3] = s[5];
    }
}

// ------------------------------------------------------------------ assembly
// One wavefront (= one 64-thread block) per element at a time; blocks stride over the
// elements of one colour, so no two elements in flight share a node and the K / f_int
// read-modify-writes need no atomics.
//
// Per element:
//   phase A, lane = (l, a): J_l = x dN_l (current configuration), det, inverse,
//     g_{l,a} = J_l^{-T} dN_{l,a} (row a of dN J^-1) -> LDS; lanes a = 0..5 of QP l stage
//     the QP coefficients, pre-multiplied by w_l det J_l.
//   phase B, lane = (a, b): 3x3 block
//     K_ab = sum_l [A g_a g_b^T + Bt (g_b g_a^T + (g_a.g_b) I) + C (N g_a)(N g_b)^T
//                   + (g_a^T sigma g_b) I (finite deformation only)]
//     which equals B_a^T D B_b (+ the geometric term) for the D above
//     (submesh.cpp:136-160, 110-134); the 9 values are added into CSR through the
//     element's slot table.
//   phase C, lanes (a, i<3): f_a,i = sum_l (sigma g_a)_i w det J (submesh.cpp:82-108).
// FULL: a general symmetric D (21 stored coefficients, microsphere models); phase B then forms
//   K_ab = sum_l B_a^T D B_b (+ the geometric term) from the engineering-shear columns of B:
//   column j of B_b has g_b,j in row j and g_b in two shear rows (kShearRow / kShearComp).
// GENERAL (with FULL): a general D, 36 stored coefficients row by row (the unsymmetric J2
//   damage tangent); K_ab = B_a^T D B_b is then not K_ba^T.
template <typename KT, bool FINITE, bool FULL = false, bool GENERAL = false>
__global__ __launch_bounds__(64) void assemble_kernel(const int32_t* __restrict__ elems, int32_t count,
                                                      const int32_t* __restrict__ conn,
                                                      const double* __restrict__ X,
                                                      const double* __restrict__ u, qp_state st,
                                                      const uint8_t* __restrict__ slots,
                                                      const int32_t* __restrict__ bptr, KT* __restrict__ K,
                                                      double* __restrict__ f_int)
{
    // coefficients per QP: A, Bt, C, n(6) (n unscaled) or the 21 of a general D; then sigma(6),
    // all scaled by w detJ
    constexpr int NCOEF = GENERAL ? 36 : FULL ? kFullCoefs : 9;
    constexpr int SIG = NCOEF;
    constexpr int NC = NCOEF + 6;
    __shared__ double s_x[8][3];
    __shared__ double s_g[8][8][3]; // [l][a][k]
    __shared__ double s_c[8][NC + 1];
    __shared__ int32_t s_node[8];

    int const lane = threadIdx.x;
    int const hi = lane >> 3, lo = lane & 7;
    int64_t const Q = st.Q;

    for (int32_t idx = blockIdx.x; idx < count; idx += gridDim.x)
    {
        int64_t const e = elems[idx];
        if (lane < 8)
        {
            int64_t const nd = conn[8 * e + lane];
            s_node[lane] = int32_t(nd);
            // current configuration x = X + u; u == nullptr integrates on X (small strain)
            for (int i = 0; i < 3; ++i) s_x[lane][i] = u ? X[3 * nd + i] + u[3 * nd + i] : X[3 * nd + i];
        }
        __syncthreads();
        // phase A: lane = (l = hi, a = lo)
        {
            int const l = hi, a = lo;
            double J[9];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                {
                    double s = 0.0;
                    for (int c = 0; c < 8; ++c) s += s_x[c][i] * c_hex8.dN[l][c][j];
                    J[i * 3 + j] = s;
                }
            double Ji[9];
            inv3(J, Ji);
            double const wdet = det3(J) * c_hex8.w[l];
            for (int k = 0; k < 3; ++k)
                s_g[l][a][k] = c_hex8.dN[l][a][0] * Ji[0 * 3 + k] + c_hex8.dN[l][a][1] * Ji[1 * 3 + k]
                               + c_hex8.dN[l][a][2] * Ji[2 * 3 + k];
            int64_t const qp = 8 * e + l;
            // coefficient slots c = a, a + 8, ...
            for (int c = a; c < NC; c += 8)
            {
                double v;
                if (c < NCOEF)
                    v = st.coef[int64_t(c) * Q + qp];
                else
                    v = st.sigma[int64_t(c - SIG) * Q + qp];
                if (FULL || c < 3 || c >= SIG) v *= wdet;
                s_c[l][c] = v;
            }
        }
        __syncthreads();
        // phase B: lane = (a = hi, b = lo)
        {
            int const a = hi, b = lo;
            double k[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            // not unrolled: unrolled, the 8 quadrature points' LDS operands are all held in
            // registers at once and occupancy falls to one wavefront per SIMD
#pragma unroll 1
            for (int l = 0; l < 8; ++l)
            {
                double const ga0 = s_g[l][a][0], ga1 = s_g[l][a][1], ga2 = s_g[l][a][2];
                double const gb0 = s_g[l][b][0], gb1 = s_g[l][b][1], gb2 = s_g[l][b][2];
                double geo = 0.0;
                if constexpr (FINITE)
                {
                    // g_a^T sigma g_b, sigma (11,22,33,23,13,12) pre-scaled
                    double const s0 = s_c[l][SIG], s1 = s_c[l][SIG + 1], s2 = s_c[l][SIG + 2],
                                 s3 = s_c[l][SIG + 3], s4 = s_c[l][SIG + 4], s5 = s_c[l][SIG + 5];
                    double const t0 = s0 * gb0 + s5 * gb1 + s4 * gb2;
                    double const t1 = s5 * gb0 + s1 * gb1 + s3 * gb2;
                    double const t2 = s4 * gb0 + s3 * gb1 + s2 * gb2;
                    geo = ga0 * t0 + ga1 * t1 + ga2 * t2;
                }
                if constexpr (FULL)
                {
                    // rows of B's column j: j (normal) and the two shear rows kRow[j][0..1], whose
                    // entries are g components kComp[j][0..1]
                    constexpr int kRow[3][2] = {{4, 5}, {3, 5}, {3, 4}};
                    constexpr int kComp[3][2] = {{2, 1}, {2, 0}, {1, 0}};
                    double const ga[3] = {ga0, ga1, ga2}, gb[3] = {gb0, gb1, gb2};
                    double Dm[6][6];
#pragma unroll
                    for (int i = 0; i < 6; ++i)
#pragma unroll
                        for (int j = 0; j < 6; ++j) Dm[i][j] = s_c[l][GENERAL ? i * 6 + j : upper_index(i, j)];
#pragma unroll
                    for (int j = 0; j < 3; ++j)
                    {
                        double db[6]; // D B_b[:, j]
#pragma unroll
                        for (int I = 0; I < 6; ++I)
                            db[I] = Dm[I][j] * gb[j] + Dm[I][kRow[j][0]] * gb[kComp[j][0]]
                                    + Dm[I][kRow[j][1]] * gb[kComp[j][1]];
#pragma unroll
                        for (int i = 0; i < 3; ++i)
                        {
                            double v = ga[i] * db[i] + ga[kComp[i][0]] * db[kRow[i][0]] + ga[kComp[i][1]] * db[kRow[i][1]];
                            if (i == j) v += geo;
                            k[i * 3 + j] += v;
                        }
                    }
                    continue;
                }
                double const A = s_c[l][0], Bt = s_c[l][1], C = s_c[l][2];
                double const gab = ga0 * gb0 + ga1 * gb1 + ga2 * gb2;
                double const diag = Bt * gab + geo;
                double const ga[3] = {ga0, ga1, ga2}, gb[3] = {gb0, gb1, gb2};
                double na[3] = {0, 0, 0}, nb[3] = {0, 0, 0};
                if (C != 0.0)
                {
                    // N g with N = [[n11,n12,n13],[n12,n22,n23],[n13,n23,n33]]
                    double const n11 = s_c[l][3], n22 = s_c[l][4], n33 = s_c[l][5], n23 = s_c[l][6],
                                 n13 = s_c[l][7], n12 = s_c[l][8];
                    na[0] = n11 * ga0 + n12 * ga1 + n13 * ga2;
                    na[1] = n12 * ga0 + n22 * ga1 + n23 * ga2;
                    na[2] = n13 * ga0 + n23 * ga1 + n33 * ga2;
                    nb[0] = n11 * gb0 + n12 * gb1 + n13 * gb2;
                    nb[1] = n12 * gb0 + n22 * gb1 + n23 * gb2;
                    nb[2] = n13 * gb0 + n23 * gb1 + n33 * gb2;
                }
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j)
                    {
                        double v = A * ga[i] * gb[j] + Bt * gb[i] * ga[j] + C * na[i] * nb[j];
                        if (i == j) v += diag;
                        k[i * 3 + j] += v;
                    }
            }
            // CSR order: row 3 n_a + i holds 3 deg(n_a) values, block slot s at 3 s
            int32_t const na_ = s_node[a];
            int32_t const b0 = bptr[na_];
            int64_t const deg = bptr[na_ + 1] - b0;
            int64_t const base = 9ll * b0 + 3ll * slots[64 * e + 8 * a + b];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j)
                {
                    int64_t const pos = base + i * 3 * deg + j;
                    K[pos] = KT(double(K[pos]) + k[i * 3 + j]);
                }
        }
        // phase C: lane = (a = hi, i = lo < 3)
        if (lo < 3)
        {
            int const a = hi, i = lo;
            double f = 0.0;
#pragma unroll 1
            for (int l = 0; l < 8; ++l)
            {
                double const s0 = s_c[l][SIG], s1 = s_c[l][SIG + 1], s2 = s_c[l][SIG + 2], s3 = s_c[l][SIG + 3],
                             s4 = s_c[l][SIG + 4], s5 = s_c[l][SIG + 5];
                double const g0 = s_g[l][a][0], g1 = s_g[l][a][1], g2 = s_g[l][a][2];
                double v;
                if (i == 0)
                    v = g0 * s0 + g1 * s5 + g2 * s4;
                else if (i == 1)
                    v = g0 * s5 + g1 * s1 + g2 * s3;
                else
                    v = g0 * s4 + g1 * s3 + g2 * s2;
                f += v;
            }
            int64_t const pos = 3ll * s_node[a] + i;
            f_int[pos] += f;
        }
        __syncthreads();
    }
}

// -------------------------------------------------------------- consistent mass
// M_ab = rho_e sum_l w_l N_a(l) N_b(l) det(X dN(l)) I_3 on the initial configuration
// (submesh::consistent_mass, submesh.cpp:162-183), added into the diagonal entries of the node
// blocks (a, b) through the slot table; one wavefront per element of one colour as in
// assemble_kernel (lanes 0..7 compute the 8 Jacobian determinants, lane (a, b) one block).
__global__ __launch_bounds__(64) void assemble_mass_kernel(const int32_t* __restrict__ elems, int32_t count,
                                                           const int32_t* __restrict__ conn,
                                                           const double* __restrict__ X,
                                                           const double* __restrict__ density,
                                                           const uint8_t* __restrict__ slots,
                                                           const int32_t* __restrict__ bptr, double* __restrict__ M)
{
    __shared__ double s_x[8][3];
    __shared__ double s_w[8]; // w_l det J_l
    __shared__ int32_t s_node[8];
    int const lane = threadIdx.x;
    int const a = lane >> 3, b = lane & 7;
    for (int32_t idx = blockIdx.x; idx < count; idx += gridDim.x)
    {
        int64_t const e = elems[idx];
        if (lane < 8)
        {
            int64_t const nd = conn[8 * e + lane];
            s_node[lane] = int32_t(nd);
            for (int i = 0; i < 3; ++i) s_x[lane][i] = X[3 * nd + i];
        }
        __syncthreads();
        if (lane < 8)
        {
            int const l = lane;
            double J[9];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                {
                    double acc = 0.0;
                    for (int c = 0; c < 8; ++c) acc += s_x[c][i] * c_hex8.dN[l][c][j];
                    J[i * 3 + j] = acc;
                }
            s_w[l] = det3(J) * c_hex8.w[l];
        }
        __syncthreads();
        double mab = 0.0;
        for (int l = 0; l < 8; ++l) mab += c_hex8.N[l][a] * c_hex8.N[l][b] * s_w[l];
        mab *= density[e];
        int32_t const na = s_node[a];
        int32_t const b0 = bptr[na];
        int64_t const deg = bptr[na + 1] - b0;
        int64_t const base = 9ll * b0 + 3ll * slots[64 * e + 8 * a + b];
        for (int i = 0; i < 3; ++i) M[base + i * 3 * deg + i] += mab;
        __syncthreads();
    }
}

// ----------------------------------------------------------------- Dirichlet
// For each fixed dof d: zero row d and column d except the diagonal.  All writes are
// zeros and the diagonal is never written, so concurrent dofs need no ordering; the
// result equals the sequential loop of enforce_dirichlet_conditions.
template <typename KT>
__global__ __launch_bounds__(kBlock) void dirichlet_kernel(int32_t n_fixed, const int32_t* __restrict__ fixed,
                                                           const int32_t* __restrict__ bptr,
                                                           const int32_t* __restrict__ bcol, KT* __restrict__ K,
                                                           double* __restrict__ b)
{
    int32_t const t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n_fixed) return;
    int32_t const d = fixed[t];
    int32_t const nd = d / 3, i = d % 3;
    int32_t const b0 = bptr[nd];
    int32_t const deg = bptr[nd + 1] - b0;
    if (b) b[d] = 0.0;
    for (int32_t s = 0; s < deg; ++s)
    {
        int32_t const m = bcol[b0 + s];
        // row d: entries (i, j) of block (nd, m)
        for (int j = 0; j < 3; ++j)
        {
            if (m == nd && j == i) continue;
            K[9ll * b0 + 3ll * i * deg + 3ll * s + j] = KT(0);
        }
        // column d in the rows of node m: find the slot of nd in m's block row
        int32_t const c0 = bptr[m];
        int32_t const cdeg = bptr[m + 1] - c0;
        int32_t slot = -1;
        for (int32_t q = 0; q < cdeg; ++q)
            if (bcol[c0 + q] == nd)
            {
                slot = q;
                break;
            }
        if (slot < 0) continue;
        // column d: entries (j, i) of block (m, nd)
        for (int j = 0; j < 3; ++j)
        {
            if (m == nd && j == i) continue;
            K[9ll * c0 + 3ll * j * cdeg + 3ll * slot + i] = KT(0);
        }
    }
}

// ----------------------------------------------------------------- adjoint
// dJ/dp_e = -lambda_e^T d f_e / d p_e for the compressible neo-Hooke element force
// f_e,a = sum_l w_l detJx_l sigma_l L_a (L_a = dN_a Jx^{-1}, current configuration,
// submesh.cpp:82-108) with sigma = (lambda ln J I + mu (b - I)) / J, lambda = K - 2 mu / 3
// (compressible_neohooke.cpp:37-40).  sigma is linear in (K, mu):
//   d sigma / dK  = ln J / J I,   d sigma / dmu = (b - I - 2/3 ln J I) / J,
// so with G = sum_a lambda_a (x) L_a, lambda_e^T d f_e/dp = sum_l w_l detJx_l (d sigma/dp : G).
// One thread per quadrature point; the 8 points of an element reduce by lane shuffles and
// the first writes the 4 gradients of the element (bulk, shear, yield, hardening; the last
// two are zero for this model).
__global__ __launch_bounds__(kBlock) void neohooke_parameter_gradient_kernel(
    int32_t E, const int32_t* __restrict__ conn, const double* __restrict__ X, const double* __restrict__ u,
    const double* __restrict__ lambda, double* __restrict__ grad)
{
    int64_t const qp = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    bool const on = qp < 8ll * E;
    int64_t const e = on ? qp >> 3 : 0;
    int const l = int(qp & 7);
    double gK = 0.0, gmu = 0.0;
    if (on)
    {
        double F0[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Fx[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        double le[8][3];
        for (int a = 0; a < 8; ++a)
        {
            int64_t const nd = conn[8 * e + a];
            for (int i = 0; i < 3; ++i)
            {
                double const Xi = X[3 * nd + i];
                double const xi = Xi + u[3 * nd + i];
                le[a][i] = lambda[3 * nd + i];
                for (int j = 0; j < 3; ++j)
                {
                    F0[i * 3 + j] += Xi * c_hex8.dN[l][a][j];
                    Fx[i * 3 + j] += xi * c_hex8.dN[l][a][j];
                }
            }
        }
        double F0i[9], Fxi[9];
        inv3(F0, F0i);
        inv3(Fx, Fxi);
        double F[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                F[i * 3 + j] = Fx[i * 3 + 0] * F0i[0 * 3 + j] + Fx[i * 3 + 1] * F0i[1 * 3 + j] + Fx[i * 3 + 2] * F0i[2 * 3 + j];
        double const J = det3(F);
        double const lnJ = log(J);
        double G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int a = 0; a < 8; ++a)
        {
            double L[3];
            for (int j = 0; j < 3; ++j)
                L[j] = c_hex8.dN[l][a][0] * Fxi[0 * 3 + j] + c_hex8.dN[l][a][1] * Fxi[1 * 3 + j]
                       + c_hex8.dN[l][a][2] * Fxi[2 * 3 + j];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) G[i * 3 + j] += le[a][i] * L[j];
        }
        double const trG = G[0] + G[4] + G[8];
        double bG = 0.0; // (F F^T) : G
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                bG += (F[i * 3 + 0] * F[j * 3 + 0] + F[i * 3 + 1] * F[j * 3 + 1] + F[i * 3 + 2] * F[j * 3 + 2]) * G[i * 3 + j];
        double const wdet = c_hex8.w[l] * det3(Fx);
        gK = -wdet * lnJ / J * trG;
        gmu = -wdet / J * (bG - trG - 2.0 / 3.0 * lnJ * trG);
    }
    for (int off = 4; off > 0; off >>= 1)
    {
        gK += __shfl_xor(gK, off, 8);
        gmu += __shfl_xor(gmu, off, 8);
    }
    if (on && l == 0)
    {
        grad[4 * e] = gK;
        grad[4 * e + 1] = gmu;
        grad[4 * e + 2] = 0.0;
        grad[4 * e + 3] = 0.0;
    }
}
// ------------------------------------------------------- J2 load-history adjoint
// Symmetric tensors in 6 components (11, 22, 33, 23, 13, 12); a : b is the full contraction.
__device__ inline double ddot6(const double* a, const double* b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + 2.0 * (a[3] * b[3] + a[4] * b[4] + a[5] * b[5]);
}

// Vector-Jacobian product of the consistent-mode radial return (small_strain_J2_plasticity.cpp
// :45-92 started from the committed history): (sigma, eps_p, ebar) = F(eps, eps_p_old,
// ebar_old; K, G, sigma_y, H).  With e = eps - eps_p_old, d = dev e, N = d / |d|,
// q = sqrt(3/2) 2G |d|, f = q - (sigma_y + H ebar_old) and, when f > 0, dgamma = f / (3G + H)
// (the exact result of the reference's iteration for linear hardening):
//   sigma = C : e - 2G sqrt(3/2) dgamma N,  eps_p = eps_p_old + sqrt(3/2) dgamma N,
//   ebar = ebar_old + dgamma,  d N = (dev de - N (N : de)) / |d|.
// Seeds sbar (sigma), epbar (eps_p), ebbar (ebar) give e_bar (= eps_bar), the history
// adjoints and pbar = (K, G, sigma_y, H) adjoints.
__device__ inline void j2_return_vjp(const double* eps, const double* ep_old, double eb_old, double K, double G,
                                     double sy, double Hh, const double* sbar, const double* epbar, double ebbar,
                                     double* e_bar, double* ep_old_bar, double* eb_old_bar, double* pbar)
{
    double const lambda = K - 2.0 * G / 3.0;
    double e[6];
    for (int k = 0; k < 6; ++k) e[k] = eps[k] - ep_old[k];
    double const tr = e[0] + e[1] + e[2];
    double const d[6] = {e[0] - tr / 3.0, e[1] - tr / 3.0, e[2] - tr / 3.0, e[3], e[4], e[5]};
    double const dn = sqrt(ddot6(d, d));
    double const a = sqrt(1.5);
    double const f0 = a * 2.0 * G * dn - (sy + Hh * eb_old);
    double const trs = sbar[0] + sbar[1] + sbar[2];
    double Cs[6];
    for (int k = 0; k < 6; ++k) Cs[k] = 2.0 * G * sbar[k] + (k < 3 ? lambda * trs : 0.0);
    double const Kbar = trs * tr;
    double const Gbar_el = 2.0 * ddot6(sbar, d);
    if (!(f0 > 0.0))
    {
        for (int k = 0; k < 6; ++k)
        {
            e_bar[k] = Cs[k];
            ep_old_bar[k] = epbar[k] - Cs[k];
        }
        *eb_old_bar = ebbar;
        pbar[0] = Kbar;
        pbar[1] = Gbar_el;
        pbar[2] = 0.0;
        pbar[3] = 0.0;
        return;
    }
    double const h3 = 3.0 * G + Hh;
    double const dg = f0 / h3;
    double N[6];
    for (int k = 0; k < 6; ++k) N[k] = d[k] / dn;
    double const sN = ddot6(sbar, N), eN = ddot6(epbar, N);
    double const gdg = -2.0 * G * a * sN + a * eN + ebbar; // adjoint of dgamma
    double W[6];
    for (int k = 0; k < 6; ++k) W[k] = dg * a * (epbar[k] - 2.0 * G * sbar[k]); // adjoint of N
    double const WN = ddot6(W, N);
    double const trW = W[0] + W[1] + W[2];
    for (int k = 0; k < 6; ++k)
    {
        double const devW = W[k] - (k < 3 ? trW / 3.0 : 0.0);
        e_bar[k] = Cs[k] + gdg * a * 2.0 * G / h3 * N[k] + (devW - WN * N[k]) / dn;
        ep_old_bar[k] = epbar[k] - e_bar[k];
    }
    *eb_old_bar = ebbar - gdg * Hh / h3;
    pbar[0] = Kbar;
    pbar[1] = Gbar_el - 2.0 * a * dg * sN + gdg * (a * 2.0 * dn - 3.0 * dg) / h3;
    pbar[2] = -gdg / h3;
    pbar[3] = -gdg * (eb_old + dg) / h3;
}

// gradient operator g_a = dN_a J_X^{-1} (reference configuration) at QP l and w detJ_X
__device__ inline double reference_gradients(const double (*Xe)[3], int l, double (*g)[3])
{
    double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int a = 0; a < 8; ++a)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) J[i * 3 + j] += Xe[a][i] * c_hex8.dN[l][a][j];
    double Ji[9];
    inv3(J, Ji);
    for (int a = 0; a < 8; ++a)
        for (int k = 0; k < 3; ++k)
            g[a][k] = c_hex8.dN[l][a][0] * Ji[0 * 3 + k] + c_hex8.dN[l][a][1] * Ji[1 * 3 + k] + c_hex8.dN[l][a][2] * Ji[2 * 3 + k];
    return c_hex8.w[l] * det3(J);
}

// symmetric part of sum_a v_a (x) g_a in 6 components
__device__ inline void sym_gradient(const double (*ve)[3], const double (*g)[3], double* s)
{
    double G9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int a = 0; a < 8; ++a)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) G9[i * 3 + j] += ve[a][i] * g[a][j];
    s[0] = G9[0];
    s[1] = G9[4];
    s[2] = G9[8];
    s[3] = 0.5 * (G9[5] + G9[7]);
    s[4] = 0.5 * (G9[2] + G9[6]);
    s[5] = 0.5 * (G9[1] + G9[3]);
}

struct j2_history
{
    const double* ep_old; // 6 per QP, SoA k * Q + q: committed history h_{n-1}
    const double* eb_old; // 1 per QP
    double* mu_ep;        // 6 per QP: multiplier of eps_p_n (in), of eps_p_{n-1} (out)
    double* mu_eb;        // 1 per QP
    int64_t Q;
};

// rhs += sum_QP B^T (dh_n / d eps)^T mu_n, one thread per element of one colour (elements of
// a colour share no node, so the read-modify-writes need no atomics)
__global__ __launch_bounds__(kBlock) void j2_adjoint_rhs_kernel(const int32_t* __restrict__ elems, int32_t count,
                                                                const int32_t* __restrict__ conn,
                                                                const double* __restrict__ X,
                                                                const double* __restrict__ u, element_params prm,
                                                                j2_history h, double* __restrict__ rhs)
{
    int32_t const t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= count) return;
    int64_t const e = elems[t];
    double Xe[8][3], ue[8][3], r[8][3];
    for (int a = 0; a < 8; ++a)
    {
        int64_t const nd = conn[8 * e + a];
        for (int i = 0; i < 3; ++i)
        {
            Xe[a][i] = X[3 * nd + i];
            ue[a][i] = u[3 * nd + i];
            r[a][i] = 0.0;
        }
    }
    double const K = prm.p[4 * e], G = prm.p[4 * e + 1], sy = prm.p[4 * e + 2], Hh = prm.p[4 * e + 3];
    double const zero[6] = {0, 0, 0, 0, 0, 0};
    for (int l = 0; l < 8; ++l)
    {
        int64_t const q = 8 * e + l;
        double g[8][3];
        reference_gradients(Xe, l, g);
        double eps[6], ep_old[6], mu[6];
        sym_gradient(ue, g, eps);
        for (int k = 0; k < 6; ++k)
        {
            ep_old[k] = h.ep_old[k * h.Q + q];
            mu[k] = h.mu_ep[k * h.Q + q];
        }
        double eb_[6], epb[6], ebb, pb[4];
        j2_return_vjp(eps, ep_old, h.eb_old[q], K, G, sy, Hh, zero, mu, h.mu_eb[q], eb_, epb, &ebb, pb);
        double const E9[9] = {eb_[0], eb_[5], eb_[4], eb_[5], eb_[1], eb_[3], eb_[4], eb_[3], eb_[2]};
        for (int a = 0; a < 8; ++a)
            for (int i = 0; i < 3; ++i)
                r[a][i] += E9[i * 3 + 0] * g[a][0] + E9[i * 3 + 1] * g[a][1] + E9[i * 3 + 2] * g[a][2];
    }
    for (int a = 0; a < 8; ++a)
    {
        int64_t const nd = conn[8 * e + a];
        for (int i = 0; i < 3; ++i) rhs[3 * nd + i] += r[a][i];
    }
}

// mu_{n-1} = (dh_n/dh_{n-1})^T mu_n - (d sigma_n/dh_{n-1})^T (w detJ B lambda_n) in place, and
// grad_e += (dh_n/dp)^T mu_n - (d sigma_n/dp)^T (w detJ B lambda_n); one thread per element
__global__ __launch_bounds__(kBlock) void j2_adjoint_history_kernel(int32_t E, const int32_t* __restrict__ conn,
                                                                    const double* __restrict__ X,
                                                                    const double* __restrict__ u,
                                                                    const double* __restrict__ lambda,
                                                                    element_params prm, j2_history h,
                                                                    double* __restrict__ grad)
{
    int64_t const e = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (e >= E) return;
    double Xe[8][3], ue[8][3], le[8][3];
    for (int a = 0; a < 8; ++a)
    {
        int64_t const nd = conn[8 * e + a];
        for (int i = 0; i < 3; ++i)
        {
            Xe[a][i] = X[3 * nd + i];
            ue[a][i] = u[3 * nd + i];
            le[a][i] = lambda[3 * nd + i];
        }
    }
    double const K = prm.p[4 * e], G = prm.p[4 * e + 1], sy = prm.p[4 * e + 2], Hh = prm.p[4 * e + 3];
    double acc[4] = {0, 0, 0, 0};
    for (int l = 0; l < 8; ++l)
    {
        int64_t const q = 8 * e + l;
        double g[8][3];
        double const wdet = reference_gradients(Xe, l, g);
        double eps[6], gl[6], sbar[6], ep_old[6], mu[6];
        sym_gradient(ue, g, eps);
        sym_gradient(le, g, gl);
        for (int k = 0; k < 6; ++k)
        {
            sbar[k] = -wdet * gl[k];
            ep_old[k] = h.ep_old[k * h.Q + q];
            mu[k] = h.mu_ep[k * h.Q + q];
        }
        double eb_[6], epb[6], ebb, pb[4];
        j2_return_vjp(eps, ep_old, h.eb_old[q], K, G, sy, Hh, sbar, mu, h.mu_eb[q], eb_, epb, &ebb, pb);
        for (int k = 0; k < 6; ++k) h.mu_ep[k * h.Q + q] = epb[k];
        h.mu_eb[q] = ebb;
        for (int k = 0; k < 4; ++k) acc[k] += pb[k];
    }
    for (int k = 0; k < 4; ++k) grad[4 * e + k] += acc[k];
}

} // namespace neon_amd

