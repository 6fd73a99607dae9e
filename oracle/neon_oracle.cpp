// CPU FP64 oracle for the hex8 implicit Newton step.  TEST INFRASTRUCTURE ONLY:
// the checker for the HIP path and the CPU baseline of the benchmark, never the
// product path.  See neon_oracle.h for the reference files each part restates.
#include "neon_oracle.h"

#include <algorithm>
#include <cfenv>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>
#include <numeric>
#include <queue>
#include <vector>

#include <omp.h>

namespace
{
using std::vector;

// ---------------------------------------------------------------- 3x3 helpers
// Row-major 3x3.
inline double det3(const double* a)
{
    return a[0] * (a[4] * a[8] - a[7] * a[5]) - a[3] * (a[1] * a[8] - a[7] * a[2])
           + a[6] * (a[1] * a[5] - a[4] * a[2]);
}

// Inverse by cofactors (the method Eigen uses for fixed 3x3 matrices).
inline void inv3(const double* a, double* r)
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

inline void matmul3(const double* a, const double* b, double* c)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
        {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += a[i * 3 + k] * b[k * 3 + j];
            c[i * 3 + j] = s;
        }
}

// --------------------------------------------------------------- hex8 tables
// hexahedron8::precompute_shape_functions (src/interpolations/hexahedron.cpp:16-67)
// at the 8-point rule of src/quadrature/hexahedron/hexahedron_quadrature.cpp:27-38.
struct hex8
{
    double N[8][8];
    double dN[8][8][3];
    double w[8];
    hex8()
    {
        static const double node[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                          {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
        double const qp = 1.0 / std::sqrt(3.0);
        for (int l = 0; l < 8; ++l)
        {
            double const xi = node[l][0] * qp, eta = node[l][1] * qp, zeta = node[l][2] * qp;
            w[l] = 1.0;
            for (int a = 0; a < 8; ++a)
            {
                double const xa = node[a][0], ya = node[a][1], za = node[a][2];
                N[l][a] = 1.0 / 8.0 * (1.0 + xa * xi) * (1.0 + ya * eta) * (1.0 + za * zeta);
                dN[l][a][0] = 1.0 / 8.0 * xa * (1.0 + ya * eta) * (1.0 + za * zeta);
                dN[l][a][1] = 1.0 / 8.0 * (1.0 + xa * xi) * ya * (1.0 + za * zeta);
                dN[l][a][2] = 1.0 / 8.0 * (1.0 + xa * xi) * (1.0 + ya * eta) * za;
            }
        }
    }
};
const hex8& shape()
{
    static hex8 h;
    return h;
}

// configuration (3x8, column a = node a) * dN (8x3): local_deformation_gradient
// (src/numeric/deformation_gradient.hpp:15-18)
inline void local_F(const double* xe /*node-major 8x3*/, const double (*dN)[3], double* J)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
        {
            double s = 0.0;
            for (int a = 0; a < 8; ++a) s += xe[a * 3 + i] * dN[a][j];
            J[i * 3 + j] = s;
        }
}

// ------------------------------------------------------ constitutive updates
// compressible_neohooke::update_internal_variables
// (src/constitutive/mechanics/solid/compressible_neohooke.cpp:27-89)
void neohooke(const double* F, double J, double bulk, double shear, double* sigma, double* D)
{
    double const lambda = bulk - 2.0 * shear / 3.0; // isotropic_elastic_property.hpp:28
    double const lnJ = std::log(J);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
        {
            double b = 0.0;
            for (int k = 0; k < 3; ++k) b += F[i * 3 + k] * F[j * 3 + k];
            double const I = i == j ? 1.0 : 0.0;
            sigma[i * 3 + j] = (lambda * lnJ * I + shear * (b - I)) / J;
        }
    double const mu = shear - lambda * lnJ;
    std::fill(D, D + 36, 0.0);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) D[i * 6 + j] = lambda + (i == j ? 2.0 * mu : 0.0);
    for (int i = 3; i < 6; ++i) D[i * 6 + i] = mu;
}

// isotropic_linear_elasticity::elastic_moduli
// (src/constitutive/mechanics/solid/isotropic_linear_elasticity.cpp:55-67)
void elastic_moduli(double bulk, double shear, double* C)
{
    double const lambda = bulk - 2.0 * shear / 3.0;
    std::fill(C, C + 36, 0.0);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C[i * 6 + j] = lambda + (i == j ? 2.0 * shear : 0.0);
    for (int i = 3; i < 6; ++i) C[i * 6 + i] = shear;
}

inline void deviatoric3(const double* a, double* d)
{
    double const t = (a[0] + a[4] + a[8]) / 3.0;
    for (int i = 0; i < 9; ++i) d[i] = a[i];
    d[0] -= t;
    d[4] -= t;
    d[8] -= t;
}
inline double norm9(const double* a)
{
    double s = 0.0;
    for (int i = 0; i < 9; ++i) s += a[i] * a[i];
    return std::sqrt(s);
}
// von_mises_stress (src/numeric/stress_routines.hpp:19-22)
inline double von_mises(const double* s)
{
    double d[9];
    deviatoric3(s, d);
    return std::sqrt(3.0 / 2.0) * norm9(d);
}

// small_strain_J2_plasticity::update_internal_variables body for one QP
// (src/constitutive/mechanics/solid/small_strain_J2_plasticity.cpp:45-92), with
// perform_radial_return (:108-155), evaluate_yield_function (:157-163),
// algorithmic_tangent (:95-106), isotropic_elastic_plastic::yield_stress (:31-36).
int j2(const double* H, double bulk, double shear, double sy, double Hh, double* eps_p,
       double* ebar, double* sigma, double* D, double* vm)
{
    double const G = shear;
    double const lambda = bulk - 2.0 * shear / 3.0;
    double strain[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) strain[i * 3 + j] = 0.5 * (H[i * 3 + j] + H[j * 3 + i]);
    double ee[9];
    for (int i = 0; i < 9; ++i) ee[i] = strain[i] - eps_p[i];
    double const tr = ee[0] + ee[4] + ee[8];
    for (int i = 0; i < 9; ++i) sigma[i] = 2.0 * G * ee[i];
    sigma[0] += lambda * tr;
    sigma[4] += lambda * tr;
    sigma[8] += lambda * tr;
    // compute_cauchy_stress = lambda tr(e) I + 2 G e (stress_routines.hpp:31-36); the
    // order above adds the same two terms.
    *vm = von_mises(sigma);
    auto yield = [&](double vm_, double e_, double dg) { return (vm_ - 3.0 * G * dg) - (sy + (e_ + dg) * Hh); };
    double C[36];
    elastic_moduli(bulk, shear, C);
    if (yield(*vm, *ebar, 0.0) <= 0.0)
    {
        std::memcpy(D, C, sizeof(C));
        return ORC_OK;
    }
    double const vm_trial = *vm;
    double dev[9];
    deviatoric3(sigma, dev);
    double const ndev = norm9(dev);
    double normal[9];
    for (int i = 0; i < 9; ++i) normal[i] = dev[i] / ndev;

    double dg = 0.0;
    double f = yield(vm_trial, *ebar, 0.0);
    int it = 0;
    while (f > 1.0e-6 && it < 50)
    {
        dg += f / (3.0 * G + Hh);
        f = yield(vm_trial, *ebar, dg);
        ++it;
    }
    if (it == 50) return ORC_RETURN_MAP_FAILED;

    double const s32 = std::sqrt(3.0 / 2.0);
    for (int i = 0; i < 9; ++i) eps_p[i] += dg * s32 * normal[i];
    for (int i = 0; i < 9; ++i) sigma[i] -= 2.0 * G * dg * s32 * normal[i];
    *vm = von_mises(sigma);
    *ebar += dg;

    // C_e - dg 6 G^2 / q_tr I_dev + 6 G^2 (dg/q_tr - 1/(3G+H)) n (x) n
    static const double Idev[36] = {2.0 / 3, -1.0 / 3, -1.0 / 3, 0, 0, 0, -1.0 / 3, 2.0 / 3, -1.0 / 3, 0, 0, 0,
                                    -1.0 / 3, -1.0 / 3, 2.0 / 3, 0, 0, 0, 0, 0, 0, 0.5, 0, 0,
                                    0, 0, 0, 0, 0.5, 0, 0, 0, 0, 0, 0, 0.5};
    double const nv[6] = {normal[0], normal[4], normal[8], normal[5], normal[2], normal[1]};
    double const c1 = dg * 6.0 * G * G / vm_trial;
    double const c2 = 6.0 * G * G * (dg / vm_trial - 1.0 / (3.0 * G + Hh));
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) D[i * 6 + j] = C[i * 6 + j] - c1 * Idev[i * 6 + j] + c2 * nv[i] * nv[j];
    return ORC_OK;
}
// isotropic_linear_elasticity::update_internal_variables
// (src/constitutive/mechanics/solid/isotropic_linear_elasticity.cpp:27-53): sigma = C : sym(H)
// with the constant moduli of elastic_moduli (:55-67), and the von Mises stress.
void linear_elastic(const double* H, double bulk, double shear, double* sigma, double* D, double* vm)
{
    double const lambda = bulk - 2.0 * shear / 3.0;
    double e[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) e[i * 3 + j] = 0.5 * (H[i * 3 + j] + H[j * 3 + i]);
    double const tr = e[0] + e[4] + e[8];
    for (int i = 0; i < 9; ++i) sigma[i] = 2.0 * shear * e[i];
    sigma[0] += lambda * tr;
    sigma[4] += lambda * tr;
    sigma[8] += lambda * tr;
    *vm = von_mises(sigma);
    elastic_moduli(bulk, shear, D);
}

// ------------------------------------------------------------ microsphere models
// unit_sphere_quadrature (src/quadrature/sphere/unit_sphere_quadrature.cpp:10-90): the
// Bazant-Oh (1986) rules with 21 and 33 directions on the half sphere (antipodal symmetry),
// weights summing to one.  The published direction cosines and weights (12 digits) are
// used as published, generated here from their symmetry orbits in the same order.
struct sphere_rule
{
    int n = 0;
    double w[33];
    double r[33][3];
    void add(double wt, double a, double b, double c)
    {
        w[n] = wt;
        r[n][0] = a;
        r[n][1] = b;
        r[n][2] = c;
        ++n;
    }
    void axes(double wt)
    {
        add(wt, 1.0, 0.0, 0.0);
        add(wt, 0.0, 1.0, 0.0);
        add(wt, 0.0, 0.0, 1.0);
    }
    // (v, +-v, 0) over the coordinate pairs (x,y), (x,z), (y,z)
    void pairs(double wt, double v)
    {
        static const int ij[3][2] = {{0, 1}, {0, 2}, {1, 2}};
        for (auto const& p : ij)
            for (double s : {1.0, -1.0})
            {
                double t[3] = {0, 0, 0};
                t[p[0]] = v;
                t[p[1]] = s * v;
                add(wt, t[0], t[1], t[2]);
            }
    }
    // (a, b, 0) type: for the zero in z, y, x: (a, b), (a, -b), (b, a), (b, -a)
    void ab0(double wt, double a, double b)
    {
        static const int ij[3][2] = {{0, 1}, {0, 2}, {1, 2}};
        for (auto const& p : ij)
            for (int k = 0; k < 4; ++k)
            {
                double t[3] = {0, 0, 0};
                t[p[0]] = k < 2 ? a : b;
                t[p[1]] = (k < 2 ? b : a) * (k % 2 ? -1.0 : 1.0);
                add(wt, t[0], t[1], t[2]);
            }
    }
    // (a, a, b) type: b in the z, y, x slot; signs (+,+), (+,-), (-,+), (-,-) on y and z
    void aab(double wt, double a, double b)
    {
        for (int pos : {2, 1, 0})
            for (int k = 0; k < 4; ++k)
            {
                double t[3] = {a, a, a};
                t[pos] = b;
                if (k >= 2) t[1] = -t[1];
                if (k % 2) t[2] = -t[2];
                add(wt, t[0], t[1], t[2]);
            }
    }
};

const sphere_rule& sphere(int which)
{
    static const sphere_rule bo21 = [] {
        sphere_rule q;
        q.axes(0.0265214244093 * 2.0);
        q.pairs(0.0199301476312 * 2.0, 0.707106781187);
        q.aab(0.0250712367487 * 2.0, 0.387907304067, 0.836095596749);
        return q;
    }();
    static const sphere_rule bo33 = [] {
        sphere_rule q;
        q.axes(2.0 * 0.00985353993433);
        q.pairs(2.0 * 0.0162969685886, 0.707106781187);
        q.ab0(2.0 * 0.0134788844008, 0.933898956394, 0.357537045978);
        q.aab(2.0 * 0.0175759129880, 0.437263676092, 0.785875915868);
        return q;
    }();
    return which == ORC_SPHERE_BO33 ? bo33 : bo21;
}

// kinetic Voigt vector (11, 22, 33, 23, 13, 12) of a 3x3 tensor (tensor_operations.hpp:373-376)
inline void kinetic6(const double* a, double* v)
{
    v[0] = a[0];
    v[1] = a[4];
    v[2] = a[8];
    v[3] = a[5];
    v[4] = a[2];
    v[5] = a[1];
}
// outer_product(matrix3 a, matrix3 b) = kinetic(a) kinetic(b)^T (tensor_operations.hpp:480-483)
inline void outer66(const double* a, const double* b, double* M)
{
    double va[6], vb[6];
    kinetic6(a, va);
    kinetic6(b, vb);
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) M[i * 6 + j] = va[i] * vb[j];
}
inline void matmul6(const double* A, const double* B, double* C)
{
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j)
        {
            double s = 0.0;
            for (int k = 0; k < 6; ++k) s += A[i * 6 + k] * B[k * 6 + j];
            C[i * 6 + j] = s;
        }
}
inline void outer3(const double* t, const double* u, double* M)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M[i * 3 + j] = t[i] * u[j];
}

// ---- small_strain_J2_plasticity_damage (src/constitutive/mechanics/solid/
// small_strain_J2_plasticity_damage.cpp:34-290): viscoplastic J2 with nonlinear kinematic
// hardening and Lemaitre-Chaboche scalar damage.  Tensors are 3x3 row-major; "kinetic" 6-vectors
// are (11, 22, 33, 23, 13, 12), "kinematic" ones double the shear entries
// (tensor_operations.hpp:207-214, 373-376).

// Solves A x = b (n x n row-major, destroyed) by Gaussian elimination with complete pivoting, the
// algorithm behind Eigen's FullPivLU::solve: at step k the entry of largest magnitude in the
// trailing block is swapped to (k, k) by a row and a column exchange.  Returns false when a pivot
// is exactly zero.
bool fullpiv_solve(int n, double* A, const double* b, double* x)
{
    int colp[16];
    double r[16];
    for (int i = 0; i < n; ++i)
    {
        colp[i] = i;
        r[i] = b[i];
    }
    for (int k = 0; k < n; ++k)
    {
        int pr = k, pc = k;
        double big = -1.0;
        for (int i = k; i < n; ++i)
            for (int j = k; j < n; ++j)
                if (std::abs(A[i * n + j]) > big)
                {
                    big = std::abs(A[i * n + j]);
                    pr = i;
                    pc = j;
                }
        if (big == 0.0) return false;
        if (pr != k)
        {
            for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[pr * n + j]);
            std::swap(r[k], r[pr]);
        }
        if (pc != k)
        {
            for (int i = 0; i < n; ++i) std::swap(A[i * n + k], A[i * n + pc]);
            std::swap(colp[k], colp[pc]);
        }
        for (int i = k + 1; i < n; ++i)
        {
            double const f = A[i * n + k] / A[k * n + k];
            if (f == 0.0) continue;
            for (int j = k; j < n; ++j) A[i * n + j] -= f * A[k * n + j];
            r[i] -= f * r[k];
        }
    }
    double y[16];
    for (int k = n - 1; k >= 0; --k)
    {
        double s = r[k];
        for (int j = k + 1; j < n; ++j) s -= A[k * n + j] * y[j];
        y[k] = s / A[k * n + k];
    }
    for (int k = 0; k < n; ++k) x[colp[k]] = y[k];
    return true;
}

inline void kinetic_from6(const double* v, double* a)
{
    a[0] = v[0], a[4] = v[1], a[8] = v[2];
    a[5] = a[7] = v[3];
    a[2] = a[6] = v[4];
    a[1] = a[3] = v[5];
}
inline void kinematic6(const double* a, double* v)
{
    v[0] = a[0], v[1] = a[4], v[2] = a[8];
    v[3] = 2.0 * a[5], v[4] = 2.0 * a[2], v[5] = 2.0 * a[1];
}
// compute_stress_like_vector_voigt (:280-285): C * kinematic(e)
inline void stress_like6(const double* C, const double* e, double* s)
{
    double v[6];
    kinematic6(e, v);
    for (int i = 0; i < 6; ++i)
    {
        double t = 0.0;
        for (int j = 0; j < 6; ++j) t += C[i * 6 + j] * v[j];
        s[i] = t;
    }
}
inline double ddot9(const double* a, const double* b)
{
    double s = 0.0;
    for (int i = 0; i < 9; ++i) s += a[i] * b[i];
    return s;
}
// mandel_notation (tensor_operations.hpp:509-515): shear blocks by sqrt 2 and 2
inline void mandel6(const double* A, double* M)
{
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j)
        {
            double f = 1.0;
            if ((i < 3) != (j < 3)) f = std::sqrt(2.0);
            if (i >= 3 && j >= 3) f = 2.0;
            M[i * 6 + j] = A[i * 6 + j] * f;
        }
}

// Material constants of isotropic_elastic_plastic_damage (isotropic_elastic_plastic_damage.cpp:
// 9-45): kinematic hardening modulus C, softening multiplier gamma, plasticity viscous
// denominator s_p and exponent n_p, damage viscous denominator s_d and exponent n_d.
struct damage_constants
{
    double C, gamma, sp, np, sd, nd;
};

// The residual (:180-189) and Jacobian (:191-236) of the return map's 16 equations at the
// unknowns y = (plastic multiplier, damage multiplier, kinetic stress (6), kinetic back stress
// (6), damage, energy release rate), for the trial elastic strain eps_e_t and the point's back
// stress and damage at the start of the update.  Jm holds the entries that the reference's
// loop does not overwrite (its initial identity entries) on entry.
void damage_system(double G, double lam, double sy, double E, const double* Ce, const damage_constants& dc, double dt,
                   const double* eps_e_t, const double* beta_t, double dmg_t, const double* y, double* res, double* Jm)
{
    double const kp = std::pow(dc.sp, -dc.np), kd = std::pow(dc.sd, -dc.nd);
    double const s32 = std::sqrt(3.0 / 2.0);
    static const double Pdev[36] = {2.0 / 3, -1.0 / 3, -1.0 / 3, 0, 0, 0, -1.0 / 3, 2.0 / 3, -1.0 / 3, 0, 0, 0,
                                    -1.0 / 3, -1.0 / 3, 2.0 / 3, 0, 0, 0, 0, 0, 0, 1, 0, 0,
                                    0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 1}; // voigt::kinetic::deviatoric
    double const dlp = y[0], dld = y[1];
    double sig[9], bet[9];
    kinetic_from6(y + 2, sig);
    kinetic_from6(y + 8, bet);
    double const D = y[14], omD = 1.0 - D, Y = y[15];
    double rel[9], dev[9];
    for (int i = 0; i < 9; ++i) rel[i] = sig[i] - bet[i];
    deviatoric3(rel, dev);
    double const nd = norm9(dev);
    double const vm = s32 * nd;
    double N[9];
    for (int i = 0; i < 9; ++i) N[i] = s32 * dev[i] / (nd * omD);
    double ee[9];
    for (int i = 0; i < 9; ++i) ee[i] = eps_e_t[i] - dlp * N[i];
    // evaluate_yield_function (:252-260), evaluate_damage_yield_function (:262-265)
    double const fp_raw = vm / omD + 0.5 * dc.gamma / dc.C * ddot9(bet, bet) - sy;
    double const fd_raw = Y - sy * sy / (2.0 * E);
    double const fp = std::max(0.0, fp_raw), fd = std::max(0.0, fd_raw);
    double dfdb[9];
    for (int i = 0; i < 9; ++i) dfdb[i] = dc.gamma / dc.C * bet[i] - N[i];
    double const dfdD = vm / (omD * omD);
    double dNds[36], od[36];
    outer66(dev, dev, od);
    for (int k = 0; k < 36; ++k) dNds[k] = s32 / omD * (Pdev[k] / nd - od[k] / (nd * nd * nd));
    double dNdD[9];
    for (int i = 0; i < 9; ++i) dNdD[i] = N[i] / omD;
    // residual (:180-189)
    double se[9]; // (1 - D) (lambda tr(ee) I + 2 G ee)
    double const tr = ee[0] + ee[4] + ee[8];
    for (int i = 0; i < 9; ++i) se[i] = omD * (2.0 * G * ee[i] + (i % 4 == 0 ? lam * tr : 0.0));
    double Cee[6], Ceem[9];
    stress_like6(Ce, ee, Cee);
    kinetic_from6(Cee, Ceem);
    res[0] = dlp - kp * std::pow(fp, dc.np) * dt;
    res[1] = dld - kd * std::pow(fd, dc.nd) * dt;
    {
        double t[9], v[6];
        for (int i = 0; i < 9; ++i) t[i] = sig[i] - se[i];
        kinetic6(t, v);
        for (int i = 0; i < 6; ++i) res[2 + i] = v[i];
        for (int i = 0; i < 9; ++i) t[i] = bet[i] - beta_t[i] - dlp * (dc.C * N[i] - dc.gamma * bet[i]);
        kinetic6(t, v);
        for (int i = 0; i < 6; ++i) res[8 + i] = v[i];
    }
    res[14] = D - dmg_t - dld;
    res[15] = Y - 0.5 * ddot9(ee, Ceem);
    // Jacobian (:191-236)
    double const cfp = -dc.np * kp * std::pow(fp, dc.np - 1.0) * dt;
    {
        double v[6], t[9];
        for (int i = 0; i < 9; ++i) t[i] = cfp * N[i];
        kinetic6(t, v);
        for (int j = 0; j < 6; ++j) Jm[0 * 16 + 2 + j] = v[j];
        for (int i = 0; i < 9; ++i) t[i] = cfp * dfdb[i];
        kinetic6(t, v);
        for (int j = 0; j < 6; ++j) Jm[0 * 16 + 8 + j] = v[j];
    }
    Jm[0 * 16 + 14] = cfp * dfdD;
    Jm[1 * 16 + 15] = -dc.nd * kd * std::pow(fd, dc.nd - 1.0) * dt;
    double omC[36], MomC[36], Mt[36], prod[36], dl_dNds[36];
    for (int k = 0; k < 36; ++k) omC[k] = omD * Ce[k];
    {
        double v[6];
        stress_like6(omC, N, v);
        for (int i = 0; i < 6; ++i) Jm[(2 + i) * 16 + 0] = v[i];
    }
    mandel6(omC, MomC);
    for (int k = 0; k < 36; ++k) dl_dNds[k] = dlp * dNds[k];
    mandel6(dl_dNds, Mt);
    matmul6(MomC, Mt, prod);
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) Jm[(2 + i) * 16 + 2 + j] = (i == j ? 1.0 : 0.0) + prod[i * 6 + j];
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) Jm[(2 + i) * 16 + 8 + j] = -prod[i * 6 + j]; // dN/dbeta = -dN/dsigma
    {
        double t[9], v[6];
        for (int i = 0; i < 9; ++i) t[i] = omD * dlp * dNdD[i] + ee[i];
        stress_like6(Ce, t, v);
        for (int i = 0; i < 6; ++i) Jm[(2 + i) * 16 + 14] = v[i];
        for (int i = 0; i < 9; ++i) t[i] = dc.gamma * bet[i] - dc.C * N[i];
        kinetic6(t, v);
        for (int i = 0; i < 6; ++i) Jm[(8 + i) * 16 + 0] = v[i];
        for (int i = 0; i < 9; ++i) t[i] = -dc.C * dlp * dNdD[i];
        kinetic6(t, v);
        for (int i = 0; i < 6; ++i) Jm[(8 + i) * 16 + 14] = v[i];
    }
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j)
        {
            Jm[(8 + i) * 16 + 2 + j] = -dc.C * dlp * dNds[i * 6 + j];
            Jm[(8 + i) * 16 + 8 + j] = (i == j ? 1.0 + dc.gamma * dlp : 0.0) + dc.C * dlp * dNds[i * 6 + j];
        }
    Jm[14 * 16 + 1] = -1.0;
    Jm[15 * 16 + 0] = ddot9(N, Ceem);
    for (int j = 0; j < 6; ++j)
    {
        double t = 0.0;
        for (int k = 0; k < 6; ++k) t += dNds[j * 6 + k] * Cee[k];
        Jm[15 * 16 + 2 + j] = dlp * t;
        Jm[15 * 16 + 8 + j] = -dlp * t;
    }
    Jm[15 * 16 + 14] = dlp * ddot9(dNdD, Ceem);
}

// perform_radial_return (:128-250) from the trial stress; the in/out state is the QP's back
// stress, damage, energy release rate (and kinematic hardening = back stress / C); returns the
// plastic increment times dt (as :249 does) or a negative status on failure.
int damage_return(double G, double lam, double sy, double E, const double* Ce, const damage_constants& dc, double dt,
                  const double* eps_e_t, double* sigma_t, double* beta_t, double* dmg_t, double* energy_t,
                  double* kin_t, double* tangent, double* plastic_increment, int* iterations)
{
    double y[16];
    y[0] = y[1] = 0.0;
    kinetic6(sigma_t, y + 2);
    kinetic6(beta_t, y + 8);
    y[14] = *dmg_t;
    y[15] = *energy_t;
    double res[16];
    for (double& v : res) v = 1.0;
    double Jm[256] = {};
    Jm[0 * 16 + 0] = Jm[1 * 16 + 1] = Jm[14 * 16 + 14] = Jm[15 * 16 + 15] = 1.0;
    for (int i = 0; i < 6; ++i) Jm[(2 + i) * 16 + 2 + i] = Jm[(8 + i) * 16 + 8 + i] = 1.0;
    double MCe[36];
    int it = 0;
    auto rnorm = [&] {
        double t = 0.0;
        for (double v : res) t += v * v;
        return std::sqrt(t);
    };
    while (rnorm() > 1.0e-6 && it < 50)
    {
        damage_system(G, lam, sy, E, Ce, dc, dt, eps_e_t, beta_t, *dmg_t, y, res, Jm);
        double A[256], dx[16];
        std::memcpy(A, Jm, sizeof(A));
        if (!fullpiv_solve(16, A, res, dx)) return ORC_RETURN_MAP_FAILED;
        for (int i = 0; i < 16; ++i) y[i] -= dx[i];
        ++it;
    }
    if (iterations) *iterations = it;
    if (y[14] > 1.0) return ORC_RETURN_MAP_FAILED; // "Damage is greater than one" (:238-243)
    if (it == 50) return ORC_RETURN_MAP_FAILED;
    double const pinc = y[0] > 0.0 ? y[0] : 0.0;
    kinetic_from6(y + 2, sigma_t);
    kinetic_from6(y + 8, beta_t);
    *dmg_t = y[14] > *dmg_t ? y[14] : *dmg_t;
    *energy_t = y[15];
    for (int i = 0; i < 9; ++i) kin_t[i] = beta_t[i] / dc.C;
    // (1 - D) mandel(C_e) mandel(inverse(J)[2:8, 2:8]) with the Jacobian of the last iteration (:245-247)
    double inv22[36];
    for (int c = 0; c < 6; ++c)
    {
        double A[256], e[16] = {}, col[16];
        std::memcpy(A, Jm, sizeof(A));
        e[2 + c] = 1.0;
        if (!fullpiv_solve(16, A, e, col)) return ORC_RETURN_MAP_FAILED;
        for (int r = 0; r < 6; ++r) inv22[r * 6 + c] = col[2 + r];
    }
    mandel6(Ce, MCe);
    double Mi[36], prod[36];
    mandel6(inv22, Mi);
    matmul6(MCe, Mi, prod);
    for (int k = 0; k < 36; ++k) tangent[k] = (1.0 - *dmg_t) * prod[k];
    *plastic_increment = pinc * dt;
    return ORC_OK;
}

// small_strain_J2_plasticity_damage::update_internal_variables (:34-120) for one QP.  The
// plastic strain, effective plastic strain, back stress and damage are updated in place, as
// neon does at every call.
int j2_damage(const double* H, double bulk, double shear, double sy, const damage_constants& dc, double dt,
              double* eps_p, double* ebar, double* beta, double* dmg, double* energy, double* kin, double* sigma,
              double* D, double* vm, int* iterations)
{
    double const G = shear, lam = bulk - 2.0 * shear / 3.0;
    double const E = 9.0 * bulk * shear / (3.0 * bulk + shear); // isotropic_elastic_property.hpp:16
    double Ce[36];
    elastic_moduli(bulk, shear, Ce);
    double strain[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) strain[i * 3 + j] = 0.5 * (H[i * 3 + j] + H[j * 3 + i]);
    double ee[9];
    for (int i = 0; i < 9; ++i) ee[i] = strain[i] - eps_p[i];
    double const tr = ee[0] + ee[4] + ee[8];
    // compute_cauchy_stress_scalar_damage (stress_routines.hpp:38-44)
    for (int i = 0; i < 9; ++i) sigma[i] = (1.0 - *dmg) * (lam * (i % 4 == 0 ? tr : 0.0) + 2.0 * G * ee[i]);
    double rel[9];
    for (int i = 0; i < 9; ++i) rel[i] = sigma[i] - beta[i];
    *vm = von_mises(rel);
    {
        double Cee[6], Ceem[9];
        stress_like6(Ce, ee, Cee);
        kinetic_from6(Cee, Ceem);
        *energy = 0.5 * ddot9(ee, Ceem); // compute_energy_release_rate (:267-272)
    }
    if (iterations) *iterations = 0;
    double const fp = *vm / (1.0 - *dmg) + 0.5 * dc.gamma / dc.C * ddot9(beta, beta) - sy;
    double const fd = *energy - sy * sy / (2.0 * E);
    if (fp <= 0.0 && fd <= 0.0)
    {
        std::memcpy(D, Ce, sizeof(Ce));
        return ORC_OK;
    }
    double pinc = 0.0;
    int const s = damage_return(G, lam, sy, E, Ce, dc, dt, ee, sigma, beta, dmg, energy, kin, D, &pinc, iterations);
    if (s != ORC_OK) return s;
    for (int i = 0; i < 9; ++i) rel[i] = sigma[i] - beta[i];
    *vm = von_mises(rel);
    double dev[9];
    deviatoric3(rel, dev);
    double const nd = norm9(dev);
    for (int i = 0; i < 9; ++i) eps_p[i] += pinc * std::sqrt(3.0 / 2.0) * dev[i] / (nd * (1.0 - *dmg));
    *ebar += pinc / (1.0 - *dmg);
    return ORC_OK;
}

// microsphere.hpp:37-65 (Pade approximations of the inverse Langevin function)
inline double pade_first(double l, double N) { return (3.0 * N - l * l) / (N - l * l); }
inline double pade_second(double l, double N) { return (std::pow(l, 4) + 3.0 * N * N) / std::pow(N - l * l, 2); }

// Macro stress tau_bar (3x3) and macro moduli c_bar (6x6) of the affine and non-affine
// microsphere models from the unimodular F_bar:
//   gaussian_affine_microsphere::compute_macro_stress (gaussian_affine_microsphere.cpp:71-83),
//     zero macro moduli (:41)
//   affine_microsphere::compute_macro_stress / compute_macro_moduli
//     (affine_microsphere.cpp:73-111)
//   nonaffine_microsphere::update_internal_variables (nonaffine_microsphere.cpp:26-88) with
//     compute_nonaffine_stretch, compute_h/H/k/K/G_tensor (:90-178), compute_o_dot_product
//     (nonaffine_microsphere.hpp:111-121); tube parameters U = 9, q = 1 (hpp:104-106)
// Each sphere sum adds f(r_i) * w_i in rule order (numerical_quadrature.hpp:47-55).
void microsphere_macro(int base, const sphere_rule& q, const double* Fb, double G, double N, double p, double* tau,
                       double* cb)
{
    std::fill(tau, tau + 9, 0.0);
    std::fill(cb, cb + 36, 0.0);
    auto tangent = [&](int i, double* t) { // deformed_tangent = F_bar r (microsphere.hpp:17-21)
        for (int a = 0; a < 3; ++a) t[a] = Fb[a * 3] * q.r[i][0] + Fb[a * 3 + 1] * q.r[i][1] + Fb[a * 3 + 2] * q.r[i][2];
    };
    auto norm = [](const double* t) { return std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]); };
    if (base == ORC_MODEL_MICROSPHERE_GAUSSIAN)
    {
        for (int i = 0; i < q.n; ++i)
        {
            double t[3], tt[9];
            tangent(i, t);
            outer3(t, t, tt);
            for (int k = 0; k < 9; ++k) tau[k] += tt[k] * q.w[i];
        }
        for (int k = 0; k < 9; ++k) tau[k] *= 3.0 * G;
        return;
    }
    if (base == ORC_MODEL_MICROSPHERE_LANGEVIN)
    {
        for (int i = 0; i < q.n; ++i)
        {
            double t[3], tt[9], tttt[36];
            tangent(i, t);
            double const l = norm(t);
            outer3(t, t, tt);
            outer66(tt, tt, tttt);
            double const pf = pade_first(l, N);
            double const a = std::pow(l, -2) * (pade_second(l, N) - pf);
            for (int k = 0; k < 9; ++k) tau[k] += pf * tt[k] * q.w[i];
            for (int k = 0; k < 36; ++k) cb[k] += a * tttt[k] * q.w[i];
        }
        for (int k = 0; k < 9; ++k) tau[k] *= G;
        for (int k = 0; k < 36; ++k) cb[k] *= G;
        return;
    }
    // non-affine: chain stretch lambda = [sum lambda_i^p w_i]^(1/p), tube area stretches nu_i
    double Fbi[9], FbiT[9];
    inv3(Fb, Fbi);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) FbiT[a * 3 + b] = Fbi[b * 3 + a];
    double sum_p = 0.0, h[9] = {0}, H[36] = {0}, k3[9] = {0}, K6[36] = {0}, G6[36] = {0};
    double const qt = 1.0, U = 9.0;
    for (int i = 0; i < q.n; ++i)
    {
        double t[3], n[3], tt[9], nn[9], tttt[36], nnnn[36], o[36];
        tangent(i, t);
        for (int a = 0; a < 3; ++a)
            n[a] = FbiT[a * 3] * q.r[i][0] + FbiT[a * 3 + 1] * q.r[i][1] + FbiT[a * 3 + 2] * q.r[i][2];
        double const l = norm(t), nu = norm(n);
        outer3(t, t, tt);
        outer3(n, n, nn);
        outer66(tt, tt, tttt);
        outer66(nn, nn, nnnn);
        sum_p += std::pow(l, p) * q.w[i];
        for (int k = 0; k < 9; ++k) h[k] += std::pow(l, p - 2.0) * tt[k] * q.w[i];
        for (int k = 0; k < 36; ++k) H[k] += std::pow(l, p - 4.0) * tttt[k] * q.w[i];
        for (int k = 0; k < 9; ++k) k3[k] += std::pow(nu, qt - 2.0) * nn[k] * q.w[i];
        for (int k = 0; k < 36; ++k) K6[k] += std::pow(nu, qt - 4.0) * nnnn[k] * q.w[i];
        double const o_[36] = {2 * n[0] * n[0], 0, 0, 0, n[0] * n[2], n[0] * n[1],
                               0, 2 * n[1] * n[1], 0, n[1] * n[2], 0, n[0] * n[1],
                               0, 0, 2 * n[2] * n[2], n[1] * n[2], n[0] * n[2], 0,
                               0, n[1] * n[2], n[1] * n[2], 0.5 * (n[1] * n[1] + n[2] * n[2]), 0.5 * n[0] * n[1], 0.5 * n[0] * n[2],
                               n[0] * n[2], 0, n[0] * n[2], 0.5 * n[0] * n[1], 0.5 * (n[0] * n[0] + n[2] * n[2]), 0.5 * n[2] * n[1],
                               n[0] * n[1], n[0] * n[1], 0, 0.5 * n[0] * n[2], 0.5 * n[2] * n[1], 0.5 * (n[0] * n[0] + n[1] * n[1])};
        std::memcpy(o, o_, sizeof(o));
        for (int k = 0; k < 36; ++k) G6[k] += std::pow(nu, qt - 2.0) * o[k] * q.w[i];
    }
    double const lna = std::pow(sum_p, 1.0 / p);
    for (int k = 0; k < 36; ++k) H[k] *= p - 2.0;
    for (int k = 0; k < 9; ++k) k3[k] *= qt;
    for (int k = 0; k < 36; ++k) K6[k] *= qt * (qt - 2.0);
    for (int k = 0; k < 36; ++k) G6[k] *= 2.0 * qt;
    double const micro_kirchhoff = G * pade_first(lna, N) * lna;
    double const micro_moduli = G * pade_second(lna, N);
    double hh[36];
    outer66(h, h, hh);
    double const c_hh = micro_moduli * std::pow(lna, 2.0 - 2.0 * p) - (p - 1.0) * micro_kirchhoff * std::pow(lna, 1.0 - 2.0 * p);
    double const c_f = micro_kirchhoff * std::pow(lna, 1.0 - p);
    for (int k = 0; k < 9; ++k) tau[k] = c_f * h[k] - G * N * U * k3[k];
    for (int k = 0; k < 36; ++k) cb[k] = (c_hh * hh[k] + c_f * H[k]) + G * N * U * (K6[k] + G6[k]);
}

// affine_microsphere::update_internal_variables / compute_kirchhoff_stress /
// compute_material_tangent (affine_microsphere.cpp:23-71), shared by all three models:
//   p = J U'(J), kappa = J^2 U''(J) with U = K/4 (J^2 - 1) - K/2 ln J
//     (volumetric_free_energy.hpp:17-33);
//   sigma = (p I + dev tau_bar) / J;
//   D = (kappa + p) 1(x)1 - 2 p I + P [c_bar + 2/3 tr(tau_bar) I - 2/3 (tau_bar(x)1 + 1(x)tau_bar)] P
//   with I = diag(1,1,1,1/2,1/2,1/2) and P the kinetic deviatoric projector.
// The projections annihilate the last bracket term, so D lacks the -2/3 (dev tau_bar (x) 1 +
// 1 (x) dev tau_bar) of the exact Kirchhoff tangent; the consistent mode adds it and divides
// by J, which makes K = d f_int / d u.
// The volumetric part and the projections shared by the microsphere models, from the macro
// stress tau_bar (3x3) and macro moduli c_bar (6x6, NULL for zero).
void microsphere_finish(const double* tau, const double* cb, double J, double K, int consistent, double* sigma,
                        double* D)
{
    double const pressure = J * (K / 2.0 * (J - 1.0 / J));
    double const kappa = std::pow(J, 2) * (K / 2.0 * (1.0 + 1.0 / std::pow(J, 2)));
    double dtau[9];
    deviatoric3(tau, dtau);
    for (int k = 0; k < 9; ++k) sigma[k] = ((k % 4 == 0 ? pressure : 0.0) + dtau[k]) / J;
    static const double Ikin[6] = {1, 1, 1, 0.5, 0.5, 0.5};
    static const double Pdev[36] = {2.0 / 3.0, -1.0 / 3.0, -1.0 / 3.0, 0, 0, 0,
                                    -1.0 / 3.0, 2.0 / 3.0, -1.0 / 3.0, 0, 0, 0,
                                    -1.0 / 3.0, -1.0 / 3.0, 2.0 / 3.0, 0, 0, 0,
                                    0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 1};
    static const double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    double tI[36], It[36], Dm[36], PD[36], PDP[36];
    outer66(tau, I3, tI);
    outer66(I3, tau, It);
    double const trt = tau[0] + tau[4] + tau[8];
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j)
            Dm[i * 6 + j] = (cb ? cb[i * 6 + j] : 0.0) + (i == j ? 2.0 / 3.0 * trt * Ikin[i] : 0.0)
                            - 2.0 / 3.0 * (tI[i * 6 + j] + It[i * 6 + j]);
    matmul6(Pdev, Dm, PD);
    matmul6(PD, Pdev, PDP);
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j)
            D[i * 6 + j] = (kappa + pressure) * (i < 3 && j < 3 ? 1.0 : 0.0) - 2.0 * pressure * (i == j ? Ikin[i] : 0.0)
                           + PDP[i * 6 + j];
    if (consistent)
    {
        double dI[36], Id[36];
        outer66(dtau, I3, dI);
        outer66(I3, dtau, Id);
        for (int k = 0; k < 36; ++k) D[k] = (D[k] - 2.0 / 3.0 * (dI[k] + Id[k])) / J;
    }
}

void microsphere(int model, const double* F, double J, const double* prm, int consistent, double* sigma, double* D)
{
    int const base = model & 0xff;
    sphere_rule const& q = sphere(model & ORC_SPHERE_BO33);
    double const K = prm[0], G = prm[1], N = prm[2], p = prm[3];
    double Fb[9];
    double const s = std::pow(det3(F), -1.0 / 3.0); // unimodular (tensor_operations.hpp:70-73)
    for (int k = 0; k < 9; ++k) Fb[k] = s * F[k];
    double tau[9], cb[36];
    microsphere_macro(base, q, Fb, G, N, p, tau, cb);
    microsphere_finish(tau, cb, J, K, consistent, sigma, D);
}

inline bool is_microsphere(int model)
{
    int const b = model & 0xff;
    return b == ORC_MODEL_MICROSPHERE_GAUSSIAN || b == ORC_MODEL_MICROSPHERE_LANGEVIN
           || b == ORC_MODEL_MICROSPHERE_NONAFFINE || b == ORC_MODEL_MICROSPHERE_AGEING;
}

// ageing_micromechanical_elastomer::integrate (micromechanical_elastomer.cpp:70-122): one
// fourth-order Runge-Kutta step (runge_kutta_integration.hpp:23-32) of the network evolution
// z = (active shear modulus, inactive shear modulus, reduction factor, active segments,
// inactive segments) over dt.
void band_rates(const double* y, double ps, double pr, double* f)
{
    double const Ga = y[0], Gi = y[1], r = y[2], Na = y[3], Ni = y[4];
    double const alpha = 1.0 - std::pow(1.0 - pr, Ni + 1.0); // inactive set recombination
    double const beta = 1.0 - std::pow(1.0 - ps, Na);        // active set scission
    double const eta = 1.0 - std::pow(1.0 - pr, Na + 1.0);   // active set generation
    double const nu = 1.0 - std::pow(1.0 - ps, Ni);          // inactive set generation
    f[0] = alpha * Gi + Ga * (2.0 * eta - beta);
    f[1] = 2.0 * beta * Ga + Gi * (nu - 2.0 * alpha);
    f[2] = -r * (beta + 2.0 * eta);
    f[3] = (2.0 * alpha * Gi * Ni - beta * Na * Ga - Na * f[0]) / Ga;
    f[4] = Gi > 1.0e-8 ? (beta * Na * Ga - 2.0 * alpha * Ni * Gi - Ni * f[1]) / Gi : 0.0;
}

void band_integrate(double* z, double dt, double ps, double pr)
{
    double k1[5], k2[5], k3[5], k4[5], y[5];
    band_rates(z, ps, pr, k1);
    for (int i = 0; i < 5; ++i) k1[i] *= dt;
    for (int i = 0; i < 5; ++i) y[i] = z[i] + k1[i] / 2.0;
    band_rates(y, ps, pr, k2);
    for (int i = 0; i < 5; ++i) k2[i] *= dt;
    for (int i = 0; i < 5; ++i) y[i] = z[i] + k2[i] / 2.0;
    band_rates(y, ps, pr, k3);
    for (int i = 0; i < 5; ++i) k3[i] *= dt;
    for (int i = 0; i < 5; ++i) y[i] = z[i] + k3[i];
    band_rates(y, ps, pr, k4);
    for (int i = 0; i < 5; ++i) k4[i] *= dt;
    for (int i = 0; i < 5; ++i) z[i] = z[i] + (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]) / 6.0;
}

// gaussian_ageing_affine_microsphere::update_internal_variables for one QP
// (gaussian_ageing_affine_microsphere.cpp:52-134): the network ODE over dt when dt != 0, the
// creation rate (micromechanical_elastomer.cpp:55-67), the primary network's Gaussian macro
// stress 3 G sum w t t, the secondary network's intermediate stress by an implicit Euler step
// S = last_dt F_bar^-1 (c / r I) F_bar^-T + S_old, macro stress r (primary + F_bar S F_bar^T),
// and the Gaussian model's stress and tangent (zero macro moduli).
void ageing(int model, const double* F, double J, const double* prm, double ps, double pr, double dt,
            double last_dt, double* z, const double* S_old, double* S, double* sigma, double* D)
{
    sphere_rule const& q = sphere(model & ORC_SPHERE_BO33);
    double const K = prm[0], G = prm[1];
    if (dt != 0.0) band_integrate(z, dt, ps, pr);
    double const alpha = 1.0 - std::pow(1.0 - pr, z[4] + 1.0);
    double const eta = 1.0 - std::pow(1.0 - pr, z[3] + 1.0);
    double const creation = alpha * z[1] + 4.0 * eta * z[0];
    double Fb[9], Fbi[9];
    double const sc = std::pow(det3(F), -1.0 / 3.0);
    for (int k = 0; k < 9; ++k) Fb[k] = sc * F[k];
    inv3(Fb, Fbi);
    double primary[9], cb[36];
    microsphere_macro(ORC_MODEL_MICROSPHERE_GAUSSIAN, q, Fb, G, 0.0, 0.0, primary, cb);
    // S (tensor order 11, 22, 33, 23, 13, 12) = last_dt F_bar^-1 (c / r) F_bar^-T + S_old
    double const a = last_dt * creation / z[2];
    double St[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
        {
            double v = 0.0;
            for (int k = 0; k < 3; ++k) v += Fbi[i * 3 + k] * Fbi[j * 3 + k];
            St[i * 3 + j] = a * v;
        }
    static const int pair[6][2] = {{0, 0}, {1, 1}, {2, 2}, {1, 2}, {0, 2}, {0, 1}};
    for (int k = 0; k < 6; ++k)
    {
        S[k] = St[pair[k][0] * 3 + pair[k][1]] + S_old[k];
        St[pair[k][0] * 3 + pair[k][1]] = St[pair[k][1] * 3 + pair[k][0]] = S[k];
    }
    double FS[9], macro[9];
    matmul3(Fb, St, FS);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
        {
            double v = 0.0;
            for (int k = 0; k < 3; ++k) v += FS[i * 3 + k] * Fb[j * 3 + k];
            macro[i * 3 + j] = z[2] * (primary[i * 3 + j] + v);
        }
    microsphere_finish(macro, nullptr, J, K, 0, sigma, D);
}
// constitutive_model::is_finite_deformation
inline bool is_finite(int model) { return model == ORC_MODEL_NEOHOOKE || is_microsphere(model); }

// The return map of j2() in consistent mode as a function of its inputs, templated on the
// scalar so that complex-step differentiation gives its exact partial derivatives (adjoint
// only).  Full 3x3 tensors; comparisons use real parts, so the complex evaluation takes the
// same branch and iteration count as the real one.
template <typename T>
double re(T v)
{
    return std::real(v);
}

template <typename T>
void j2_map(const T* strain, const T* ep_old, T eb_old, T bulk, T shear, T sy, T Hh, T* sigma, T* ep, T* eb)
{
    T const G = shear;
    T const lambda = bulk - T(2.0) * shear / T(3.0);
    T ee[9];
    for (int i = 0; i < 9; ++i) ee[i] = strain[i] - ep_old[i];
    T const tr = ee[0] + ee[4] + ee[8];
    for (int i = 0; i < 9; ++i) sigma[i] = T(2.0) * G * ee[i];
    sigma[0] += lambda * tr;
    sigma[4] += lambda * tr;
    sigma[8] += lambda * tr;
    T const t = (sigma[0] + sigma[4] + sigma[8]) / T(3.0);
    T dev[9];
    for (int i = 0; i < 9; ++i) dev[i] = sigma[i];
    dev[0] -= t;
    dev[4] -= t;
    dev[8] -= t;
    T nn = T(0.0);
    for (int i = 0; i < 9; ++i) nn += dev[i] * dev[i];
    T const ndev = std::sqrt(nn);
    T const vm_trial = std::sqrt(T(1.5)) * ndev;
    for (int i = 0; i < 9; ++i) ep[i] = ep_old[i];
    *eb = eb_old;
    auto yield = [&](T dg) { return (vm_trial - T(3.0) * G * dg) - (sy + (eb_old + dg) * Hh); };
    if (re(yield(T(0.0))) <= 0.0) return;
    T dg = T(0.0);
    T f = yield(dg);
    for (int it = 0; re(f) > 1.0e-6 && it < 50; ++it)
    {
        dg += f / (T(3.0) * G + Hh);
        f = yield(dg);
    }
    T const s32 = std::sqrt(T(1.5));
    for (int i = 0; i < 9; ++i)
    {
        T const nrm = dev[i] / ndev;
        ep[i] += dg * s32 * nrm;
        sigma[i] -= T(2.0) * G * dg * s32 * nrm;
    }
    *eb = eb_old + dg;
}

// Vector-Jacobian product of j2_map at one point: given seeds (sig_bar, ep_bar: 3x3 tensors
// paired by full contraction; eb_bar) returns the adjoints of strain (symmetric tensor E
// with E : d strain), of ep_old (tensor), eb_old and the 4 parameters.  Each input
// direction is a unit symmetric tensor ((e_i e_j + e_j e_i) / 2 off the diagonal), whose
// adjoint equals the tensor component.
void j2_map_vjp(const double* strain, const double* ep_old, double eb_old, const double* prm, const double* sig_bar,
                const double* ep_bar, double eb_bar, double* strain_bar, double* ep_old_bar, double* eb_old_bar,
                double* prm_bar)
{
    typedef std::complex<double> cplx;
    double const h = 1e-30;
    static const int pair[6][2] = {{0, 0}, {1, 1}, {2, 2}, {1, 2}, {0, 2}, {0, 1}};
    auto evaluate = [&](int group, int k) {
        cplx s[9], e0[9], ps[4], eb0 = eb_old;
        for (int i = 0; i < 9; ++i)
        {
            s[i] = strain[i];
            e0[i] = ep_old[i];
        }
        for (int i = 0; i < 4; ++i) ps[i] = prm[i];
        cplx const d(0.0, h);
        if (group < 2)
        {
            cplx* t = group == 0 ? s : e0;
            int const a = pair[k][0], b = pair[k][1];
            if (a == b)
                t[3 * a + a] += d;
            else
            {
                t[3 * a + b] += 0.5 * d;
                t[3 * b + a] += 0.5 * d;
            }
        }
        else if (group == 2)
            eb0 += d;
        else
            ps[k] += d;
        cplx sig[9], ep[9], eb;
        j2_map<cplx>(s, e0, eb0, ps[0], ps[1], ps[2], ps[3], sig, ep, &eb);
        double acc = eb_bar * eb.imag() / h;
        for (int i = 0; i < 9; ++i) acc += (sig_bar[i] * sig[i].imag() + ep_bar[i] * ep[i].imag()) / h;
        return acc;
    };
    for (int k = 0; k < 6; ++k)
    {
        int const a = pair[k][0], b = pair[k][1];
        double const gs = evaluate(0, k), ge = evaluate(1, k);
        strain_bar[3 * a + b] = strain_bar[3 * b + a] = gs;
        ep_old_bar[3 * a + b] = ep_old_bar[3 * b + a] = ge;
    }
    *eb_old_bar = evaluate(2, 0);
    for (int k = 0; k < 4; ++k) prm_bar[k] = evaluate(3, k);
}

// gradient operator of element e at QP l on the reference configuration: Bt = dN J_X^{-1}
void reference_gradients(const orc_model& m, int64_t e, int l, double (*Bt)[3], double& wdetJ);

} // namespace

// ======================================================================= model
struct orc_model
{
    int32_t N = 0, E = 0;
    vector<double> X;
    vector<int32_t> conn;
    int32_t model = 0;
    int32_t tangent_mode = 0;
    vector<double> params; // 4 per element
    vector<double> x;      // current configuration
    // per-QP variables (internal_variables); "old" copies for commit/revert of
    // scalars and second order tensors (src/constitutive/internal_variables.hpp:203-214)
    vector<double> Hg, F, sigma, detF, eps_p, ebar, vm;
    vector<double> Hg_o, F_o, sigma_o, detF_o, eps_p_o, ebar_o, vm_o;
    vector<double> D; // fourth-order tangent, not saved
    // ageing microsphere: network state (11 per QP) and its committed copy, the material's
    // probabilities and the model's last non-zero time step (a model-wide member in neon)
    vector<double> age, age_o;
    double scission = 0.0, recombination = 0.0, last_dt = 0.0;
    // J2 with damage: back stress (9 per QP), damage, energy release rate and their committed
    // copies, kinematic hardening (back stress / C), the material's constants
    vector<double> beta, dmg, energy, beta_o, dmg_o, energy_o, kin;
    damage_constants dconst{1.0, 1.0, 1.0, 1.0, 1.0, 1.0};
    // recorded load history: u_n and the committed history h_{n-1} of each converged step
    vector<vector<double>> traj_u, traj_ep, traj_eb;
};

extern "C" {

int orc_hex8_tables(double* N, double* dN, double* w)
{
    hex8 const& h = shape();
    if (N) std::memcpy(N, h.N, sizeof(h.N));
    if (dN) std::memcpy(dN, h.dN, sizeof(h.dN));
    if (w) std::memcpy(w, h.w, sizeof(h.w));
    return ORC_OK;
}

int orc_neohooke_qp(const double* F, double bulk, double shear, double* sigma, double* D)
{
    neohooke(F, det3(F), bulk, shear, sigma, D);
    return ORC_OK;
}

int orc_microsphere_qp(int32_t model, const double* F, const double* params, int32_t consistent, double* sigma,
                       double* D)
{
    if (!F || !params || !sigma || !D || !is_microsphere(model) || (model & 0xff) == ORC_MODEL_MICROSPHERE_AGEING)
        return ORC_BAD_ARGUMENT;
    microsphere(model, F, det3(F), params, consistent, sigma, D);
    return ORC_OK;
}

int orc_sphere_rule(int32_t model, double* w, double* r)
{
    sphere_rule const& q = sphere(model & ORC_SPHERE_BO33);
    for (int i = 0; i < q.n; ++i)
    {
        if (w) w[i] = q.w[i];
        if (r)
            for (int a = 0; a < 3; ++a) r[3 * i + a] = q.r[i][a];
    }
    return q.n;
}

int orc_j2_qp(const double* H, double bulk, double shear, double yield, double hardening, double* eps_p,
              double* ebar, double* sigma, double* D, double* von_mises_out)
{
    return j2(H, bulk, shear, yield, hardening, eps_p, ebar, sigma, D, von_mises_out);
}

int orc_model_create(int32_t n_nodes, const double* X, int32_t n_elem, const int32_t* conn, int32_t model,
                     const double* params, int32_t tangent_mode, orc_model** out)
{
    if (!out || n_nodes <= 0 || n_elem <= 0 || !X || !conn || !params) return ORC_BAD_ARGUMENT;
    int const base = model & 0xff;
    if ((model & ~(0xff | ORC_SPHERE_BO33)) != 0 || base > ORC_MODEL_J2_DAMAGE) return ORC_BAD_ARGUMENT;
    if ((base == ORC_MODEL_MICROSPHERE_AGEING || base == ORC_MODEL_J2_DAMAGE) && tangent_mode == ORC_TANGENT_CONSISTENT)
        return ORC_BAD_ARGUMENT;
    if ((model & ORC_SPHERE_BO33) && !is_microsphere(model)) return ORC_BAD_ARGUMENT;
    for (int64_t i = 0; i < int64_t(n_elem) * 8; ++i)
        if (conn[i] < 0 || conn[i] >= n_nodes) return ORC_BAD_ARGUMENT;
    auto* m = new orc_model;
    m->N = n_nodes;
    m->E = n_elem;
    m->X.assign(X, X + 3 * size_t(n_nodes));
    m->x = m->X;
    m->conn.assign(conn, conn + 8 * size_t(n_elem));
    m->model = model;
    m->tangent_mode = tangent_mode;
    m->params.assign(params, params + 4 * size_t(n_elem));
    size_t const Q = size_t(n_elem) * 8;
    m->Hg.assign(9 * Q, 0.0);
    m->F.assign(9 * Q, 0.0);
    for (size_t q = 0; q < Q; ++q) m->F[9 * q] = m->F[9 * q + 4] = m->F[9 * q + 8] = 1.0; // submesh.cpp:45-47
    m->sigma.assign(9 * Q, 0.0);
    m->detF.assign(Q, 1.0);
    m->eps_p.assign(9 * Q, 0.0);
    m->ebar.assign(Q, 0.0);
    m->vm.assign(Q, 0.0);
    m->D.assign(36 * Q, 0.0);
    if (model == ORC_MODEL_J2 || model == ORC_MODEL_LINEAR_ELASTIC
        || model == ORC_MODEL_J2_DAMAGE) // isotropic_linear_elasticity.cpp:22-23
        for (size_t q = 0; q < Q; ++q)
            elastic_moduli(m->params[4 * (q / 8)], m->params[4 * (q / 8) + 1], &m->D[36 * q]);
    if (model == ORC_MODEL_J2_DAMAGE)
    {
        m->beta.assign(9 * Q, 0.0);
        m->kin.assign(9 * Q, 0.0);
        m->dmg.assign(Q, 0.0);
        m->energy.assign(Q, 0.0);
        m->beta_o = m->beta;
        m->dmg_o = m->dmg;
        m->energy_o = m->energy;
    }
    if (base == ORC_MODEL_MICROSPHERE_AGEING)
    {
        // gaussian_ageing_affine_microsphere.cpp:17-50: active modulus G, inactive modulus
        // (1 - cure_time / 100) G, reduction 1, active segments N, inactive segments 0
        m->age.assign(11 * Q, 0.0);
        for (size_t q = 0; q < Q; ++q)
        {
            double const* p = &m->params[4 * (q / 8)];
            double* a = &m->age[11 * q];
            a[0] = p[1];
            a[1] = p[3];
            a[2] = 1.0;
            a[3] = p[2];
        }
        m->age_o = m->age;
    }
    m->Hg_o = m->Hg;
    m->F_o = m->F;
    m->sigma_o = m->sigma;
    m->detF_o = m->detF;
    m->eps_p_o = m->eps_p;
    m->ebar_o = m->ebar;
    m->vm_o = m->vm;
    *out = m;
    return ORC_OK;
}

int orc_model_destroy(orc_model* m)
{
    delete m;
    return ORC_OK;
}

int orc_model_n_dofs(const orc_model* m) { return 3 * m->N; }

// mesh::update_internal_variables (src/mesh/mechanics/solid/mesh.cpp:57-71) ->
// material_coordinates::update_current_configuration (src/mesh/material_coordinates.cpp:23-27)
// -> submesh::update_internal_variables (submesh.cpp:194-208).
int orc_update_internal_variables(orc_model* m, const double* u, double dt)
{
    std::feclearexcept(FE_ALL_EXCEPT);
    for (int64_t i = 0; i < 3 * int64_t(m->N); ++i) m->x[i] = m->X[i] + u[i];
    hex8 const& h = shape();

    // update_deformation_measures (submesh.cpp:210-237)
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m->E; ++e)
    {
        double Xe[24], xe[24];
        for (int a = 0; a < 8; ++a)
            for (int i = 0; i < 3; ++i)
            {
                Xe[a * 3 + i] = m->X[3 * size_t(m->conn[8 * e + a]) + i];
                xe[a * 3 + i] = m->x[3 * size_t(m->conn[8 * e + a]) + i];
            }
        for (int l = 0; l < 8; ++l)
        {
            double F0[9], Fx[9], F0i[9];
            local_F(Xe, h.dN[l], F0);
            local_F(xe, h.dN[l], Fx);
            inv3(F0, F0i);
            double B0t[8][3]; // rhea * F_0^{-1}
            for (int a = 0; a < 8; ++a)
                for (int j = 0; j < 3; ++j)
                {
                    double s = 0.0;
                    for (int k = 0; k < 3; ++k) s += h.dN[l][a][k] * F0i[k * 3 + j];
                    B0t[a][j] = s;
                }
            double* H = &m->Hg[9 * (8 * e + l)];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                {
                    double s = 0.0;
                    for (int a = 0; a < 8; ++a) s += (xe[a * 3 + i] - Xe[a * 3 + i]) * B0t[a][j];
                    H[i * 3 + j] = s;
                }
            matmul3(Fx, F0i, &m->F[9 * (8 * e + l)]);
        }
    }
    // update_Jacobian_determinants (submesh.cpp:239-269)
    size_t const Q = size_t(m->E) * 8;
    for (size_t q = 0; q < Q; ++q) m->detF[q] = det3(&m->F[9 * q]);
    for (size_t q = 0; q < Q; ++q)
        if (std::signbit(m->detF[q])) return ORC_NEGATIVE_JACOBIAN;

    int status = ORC_OK;
    if (m->model == ORC_MODEL_NEOHOOKE)
    {
        for (size_t q = 0; q < Q; ++q)
        {
            double const* p = &m->params[4 * (q / 8)];
            neohooke(&m->F[9 * q], m->detF[q], p[0], p[1], &m->sigma[9 * q], &m->D[36 * q]);
            if (m->tangent_mode == ORC_TANGENT_CONSISTENT)
                for (int k = 0; k < 36; ++k) m->D[36 * q + k] /= m->detF[q];
        }
    }
    else if (m->model == ORC_MODEL_LINEAR_ELASTIC)
    {
        for (size_t q = 0; q < Q; ++q)
        {
            double const* p = &m->params[4 * (q / 8)];
            linear_elastic(&m->Hg[9 * q], p[0], p[1], &m->sigma[9 * q], &m->D[36 * q], &m->vm[q]);
        }
    }
    else if ((m->model & 0xff) == ORC_MODEL_MICROSPHERE_AGEING)
    {
        if (dt != 0.0) m->last_dt = dt; // gaussian_ageing_affine_microsphere.cpp:84-103
#pragma omp parallel for schedule(static)
        for (int64_t q = 0; q < int64_t(Q); ++q)
            ageing(m->model, &m->F[9 * q], m->detF[q], &m->params[4 * (q / 8)], m->scission, m->recombination, dt,
                   m->last_dt, &m->age[11 * q], &m->age_o[11 * q + 5], &m->age[11 * q + 5], &m->sigma[9 * q],
                   &m->D[36 * q]);
    }
    else if (m->model == ORC_MODEL_J2_DAMAGE)
    {
        int failed = 0;
#pragma omp parallel for schedule(static) reduction(| : failed)
        for (int64_t q = 0; q < int64_t(Q); ++q)
        {
            double const* p = &m->params[4 * (q / 8)];
            int const s = j2_damage(&m->Hg[9 * q], p[0], p[1], p[2], m->dconst, dt, &m->eps_p[9 * q], &m->ebar[q],
                                    &m->beta[9 * q], &m->dmg[q], &m->energy[q], &m->kin[9 * q], &m->sigma[9 * q],
                                    &m->D[36 * q], &m->vm[q], nullptr);
            if (s != ORC_OK) failed |= 1;
        }
        if (failed) status = ORC_RETURN_MAP_FAILED;
    }
    else if (is_microsphere(m->model))
    {
        int const consistent = m->tangent_mode == ORC_TANGENT_CONSISTENT;
#pragma omp parallel for schedule(static)
        for (int64_t q = 0; q < int64_t(Q); ++q)
            microsphere(m->model, &m->F[9 * q], m->detF[q], &m->params[4 * (q / 8)], consistent, &m->sigma[9 * q],
                        &m->D[36 * q]);
    }
    else
    {
        int failed = 0;
#pragma omp parallel for schedule(static) reduction(| : failed)
        for (int64_t q = 0; q < int64_t(Q); ++q)
        {
            double const* p = &m->params[4 * (q / 8)];
            if (m->tangent_mode == ORC_TANGENT_CONSISTENT)
            {
                // consistent mode: return map from the last converged history (backward Euler
                // from t_n), instead of neon's in-place accumulation across Newton iterates
                std::copy(&m->eps_p_o[9 * q], &m->eps_p_o[9 * q] + 9, &m->eps_p[9 * q]);
                m->ebar[q] = m->ebar_o[q];
            }
            int const s = j2(&m->Hg[9 * q], p[0], p[1], p[2], p[3], &m->eps_p[9 * q], &m->ebar[q],
                             &m->sigma[9 * q], &m->D[36 * q], &m->vm[q]);
            if (s != ORC_OK) failed |= 1;
        }
        if (failed) status = ORC_RETURN_MAP_FAILED;
    }
    if (status != ORC_OK) return status;
    // FE_INVALID test of submesh.cpp:204-207.  The per-QP loops run on worker threads whose
    // floating-point flags the calling thread cannot see, so non-finite results are scanned.
    if (std::fetestexcept(FE_INVALID)) return ORC_NAN_OR_INF;
    for (size_t k = 0; k < 9 * Q; ++k)
        if (!std::isfinite(m->sigma[k])) return ORC_NAN_OR_INF;
    return ORC_OK;
}

// submesh::save_internal_variables (submesh.cpp:54-64)
int orc_save_internal_variables(orc_model* m, int32_t converged)
{
    if (converged)
    {
        m->Hg_o = m->Hg;
        m->F_o = m->F;
        m->sigma_o = m->sigma;
        m->detF_o = m->detF;
        m->eps_p_o = m->eps_p;
        m->ebar_o = m->ebar;
        m->vm_o = m->vm;
        m->age_o = m->age;
        m->beta_o = m->beta;
        m->dmg_o = m->dmg;
        m->energy_o = m->energy;
    }
    else
    {
        m->Hg = m->Hg_o;
        m->F = m->F_o;
        m->sigma = m->sigma_o;
        m->detF = m->detF_o;
        m->eps_p = m->eps_p_o;
        m->ebar = m->ebar_o;
        m->vm = m->vm_o;
        m->age = m->age_o;
        m->beta = m->beta_o;
        m->dmg = m->dmg_o;
        m->energy = m->energy_o;
        // the kinematic hardening variable is the back stress over the modulus at every point
        for (size_t k = 0; k < m->kin.size(); ++k) m->kin[k] = m->beta[k] / m->dconst.C;
    }
    return ORC_OK;
}

int orc_model_set_damage(orc_model* m, const double* c)
{
    if (!m || !c || m->model != ORC_MODEL_J2_DAMAGE) return ORC_BAD_ARGUMENT;
    for (int k = 0; k < 6; ++k)
        if (!(c[k] > 0.0)) return ORC_BAD_ARGUMENT;
    m->dconst = damage_constants{c[0], c[1], c[2], c[3], c[4], c[5]};
    return ORC_OK;
}

int orc_get_damage_state(const orc_model* m, double* back_stress, double* damage, double* energy)
{
    if (!m || m->model != ORC_MODEL_J2_DAMAGE) return ORC_BAD_ARGUMENT;
    if (back_stress) std::copy(m->beta.begin(), m->beta.end(), back_stress);
    if (damage) std::copy(m->dmg.begin(), m->dmg.end(), damage);
    if (energy) std::copy(m->energy.begin(), m->energy.end(), energy);
    return ORC_OK;
}

int orc_j2_damage_system(const double* eps_e_t, const double* back_stress, double damage, double bulk, double shear,
                         double yield, const double* c, double dt, const double* y, double* residual, double* J)
{
    damage_constants const dc{c[0], c[1], c[2], c[3], c[4], c[5]};
    double Ce[36];
    elastic_moduli(bulk, shear, Ce);
    double const E = 9.0 * bulk * shear / (3.0 * bulk + shear);
    for (int k = 0; k < 256; ++k) J[k] = 0.0;
    J[0 * 16 + 0] = J[1 * 16 + 1] = J[14 * 16 + 14] = J[15 * 16 + 15] = 1.0;
    for (int i = 0; i < 6; ++i) J[(2 + i) * 16 + 2 + i] = J[(8 + i) * 16 + 8 + i] = 1.0;
    damage_system(shear, bulk - 2.0 * shear / 3.0, yield, E, Ce, dc, dt, eps_e_t, back_stress, damage, y, residual, J);
    return ORC_OK;
}

int orc_j2_damage_qp(const double* H, double bulk, double shear, double yield, const double* c, double dt,
                     double* eps_p, double* ebar, double* back_stress, double* damage, double* energy,
                     double* kinematic_hardening, double* sigma, double* D, double* von_mises, int32_t* iterations)
{
    damage_constants const dc{c[0], c[1], c[2], c[3], c[4], c[5]};
    int its = 0;
    int const s = j2_damage(H, bulk, shear, yield, dc, dt, eps_p, ebar, back_stress, damage, energy,
                            kinematic_hardening, sigma, D, von_mises, &its);
    if (iterations) *iterations = its;
    return s;
}

int orc_model_set_ageing(orc_model* m, double scission, double recombination)
{
    if (!m || scission < 0.0 || recombination < 0.0) return ORC_BAD_ARGUMENT;
    m->scission = scission;
    m->recombination = recombination;
    return ORC_OK;
}

int orc_get_ageing_state(const orc_model* m, double* state)
{
    if (!m || !state || m->age.empty()) return ORC_BAD_ARGUMENT;
    std::memcpy(state, m->age.data(), m->age.size() * sizeof(double));
    return ORC_OK;
}

int orc_get_qp_state(const orc_model* m, double* sigma, double* D, double* detF, double* eps_p, double* ebar,
                     double* vm)
{
    size_t const Q = size_t(m->E) * 8;
    if (sigma) std::memcpy(sigma, m->sigma.data(), 9 * Q * sizeof(double));
    if (D) std::memcpy(D, m->D.data(), 36 * Q * sizeof(double));
    if (detF) std::memcpy(detF, m->detF.data(), Q * sizeof(double));
    if (eps_p) std::memcpy(eps_p, m->eps_p.data(), 9 * Q * sizeof(double));
    if (ebar) std::memcpy(ebar, m->ebar.data(), Q * sizeof(double));
    if (vm) std::memcpy(vm, m->vm.data(), Q * sizeof(double));
    return ORC_OK;
}

} // extern "C"

namespace
{
// submesh::tangent_stiffness = material_tangent_stiffness + geometric_tangent_stiffness
// (submesh.cpp:66-80, 110-160) and submesh::internal_force (submesh.cpp:82-108), using the
// dense B (6x24) of symmetric_gradient<3> (src/numeric/gradient_operator.hpp:38-57).
// sigma_e: optional 8 x 9 Cauchy stresses replacing the element's stored ones (f_e only)
void element_matrices(const orc_model& m, int64_t e, double* ke, double* fe, const double* sigma_e = nullptr)
{
    hex8 const& h = shape();
    double xe[24];
    // neon integrates on the current configuration x (submesh.cpp:68,84), also for the
    // small-strain J2 model; consistent mode integrates small strain on X.
    bool const material_frame = (m.model == ORC_MODEL_J2 || m.model == ORC_MODEL_LINEAR_ELASTIC)
                                && m.tangent_mode == ORC_TANGENT_CONSISTENT;
    const std::vector<double>& cfg = material_frame ? m.X : m.x;
    for (int a = 0; a < 8; ++a)
        for (int i = 0; i < 3; ++i) xe[a * 3 + i] = cfg[3 * size_t(m.conn[8 * e + a]) + i];
    bool const want_k = ke != nullptr; // internal_force alone skips the stiffness
    if (want_k) std::fill(ke, ke + 576, 0.0);
    std::fill(fe, fe + 24, 0.0);
    double kgeo[64];
    std::fill(kgeo, kgeo + 64, 0.0);
    bool const finite = is_finite(m.model); // is_finite_deformation()
    for (int l = 0; l < 8; ++l)
    {
        double J[9], Ji[9];
        local_F(xe, h.dN[l], J);
        double const detJ = det3(J);
        inv3(J, Ji);
        double Bt[8][3]; // dN * J^{-1}
        for (int a = 0; a < 8; ++a)
            for (int k = 0; k < 3; ++k)
            {
                double s = 0.0;
                for (int j = 0; j < 3; ++j) s += h.dN[l][a][j] * Ji[j * 3 + k];
                Bt[a][k] = s;
            }
        double const wl = h.w[l];
        double const* sg = sigma_e ? sigma_e + 9 * l : &m.sigma[9 * (8 * e + l)];
        for (int a = 0; a < 8; ++a)
            for (int i = 0; i < 3; ++i)
            {
                double s = 0.0;
                for (int j = 0; j < 3; ++j) s += Bt[a][j] * sg[j * 3 + i];
                fe[a * 3 + i] += s * detJ * wl;
            }
        if (!want_k) continue;
        double B[6][24];
        std::memset(B, 0, sizeof(B));
        for (int a = 0; a < 8; ++a)
        {
            int const b = 3 * a;
            B[0][b] = Bt[a][0];
            B[4][b] = Bt[a][2];
            B[5][b] = Bt[a][1];
            B[1][b + 1] = Bt[a][1];
            B[3][b + 1] = Bt[a][2];
            B[5][b + 1] = Bt[a][0];
            B[2][b + 2] = Bt[a][2];
            B[3][b + 2] = Bt[a][1];
            B[4][b + 2] = Bt[a][0];
        }
        double const* D = &m.D[36 * (8 * e + l)];
        double DB[6][24];
        for (int i = 0; i < 6; ++i)
            for (int c = 0; c < 24; ++c)
            {
                double s = 0.0;
                for (int k = 0; k < 6; ++k) s += D[i * 6 + k] * B[k][c];
                DB[i][c] = s;
            }
        for (int r = 0; r < 24; ++r)
            for (int c = 0; c < 24; ++c)
            {
                double s = 0.0;
                for (int k = 0; k < 6; ++k) s += B[k][r] * DB[k][c];
                ke[r * 24 + c] += s * detJ * wl;
            }
        if (finite)
            for (int a = 0; a < 8; ++a)
                for (int b = 0; b < 8; ++b)
                {
                    double s = 0.0;
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j) s += Bt[a][i] * sg[i * 3 + j] * Bt[b][j];
                    kgeo[a * 8 + b] += s * detJ * wl;
                }
    }
    if (finite && want_k) // identity_expansion_inplace<3> (src/numeric/tensor_operations.hpp:122-135)
        for (int a = 0; a < 8; ++a)
            for (int b = 0; b < 8; ++b)
                for (int k = 0; k < 3; ++k) ke[(a * 3 + k) * 24 + b * 3 + k] += kgeo[a * 8 + b];
}

// node adjacency of the element graph: sorted unique neighbour lists
void node_adjacency(const orc_model& m, vector<vector<int32_t>>& adj)
{
    adj.assign(m.N, {});
    for (int64_t e = 0; e < m.E; ++e)
        for (int a = 0; a < 8; ++a)
            for (int b = 0; b < 8; ++b) adj[m.conn[8 * e + a]].push_back(m.conn[8 * e + b]);
    for (auto& l : adj)
    {
        std::sort(l.begin(), l.end());
        l.erase(std::unique(l.begin(), l.end()), l.end());
    }
}
} // namespace

extern "C" {

int orc_element_matrices(const orc_model* m, int32_t element, double* ke, double* fe)
{
    if (element < 0 || element >= m->E) return ORC_BAD_ARGUMENT;
    element_matrices(*m, element, ke, fe);
    return ORC_OK;
}

// Adjoint parameter gradient (no counterpart in the reference): grad_e,k = -lambda_e .
// d f_e / d p_k for the neo-Hooke element force at the current state.  The Cauchy stress
// (compressible_neohooke.cpp:37-40) is linear in (bulk, shear), so d f_e / d bulk and
// d f_e / d shear are exactly the element forces of the stresses evaluated with unit
// parameters (1, 0) and (0, 1) at the stored F; yield and hardening get zero.
int orc_parameter_gradient(const orc_model* m, const double* lambda, double* grad)
{
    if (!m || !lambda || !grad || m->model != ORC_MODEL_NEOHOOKE) return ORC_BAD_ARGUMENT;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m->E; ++e)
    {
        double sk[72], sm[72], D[36], fk[24], fm[24];
        for (int l = 0; l < 8; ++l)
        {
            size_t const q = 8 * size_t(e) + l;
            neohooke(&m->F[9 * q], m->detF[q], 1.0, 0.0, sk + 9 * l, D);
            neohooke(&m->F[9 * q], m->detF[q], 0.0, 1.0, sm + 9 * l, D);
        }
        element_matrices(*m, e, nullptr, fk, sk);
        element_matrices(*m, e, nullptr, fm, sm);
        double gk = 0.0, gm = 0.0;
        for (int a = 0; a < 8; ++a)
            for (int i = 0; i < 3; ++i)
            {
                double const la = lambda[3 * size_t(m->conn[8 * e + a]) + i];
                gk += la * fk[3 * a + i];
                gm += la * fm[3 * a + i];
            }
        grad[4 * e] = -gk;
        grad[4 * e + 1] = -gm;
        grad[4 * e + 2] = 0.0;
        grad[4 * e + 3] = 0.0;
    }
    return ORC_OK;
}

// compute_sparsity_pattern (src/assembler/sparsity_pattern.hpp:18-48): the CSR pattern
// of all element dof pairs with sorted, unique column indices per row; dof = 3 node + i
// (src/mesh/dof_allocator.hpp:41-47).
int64_t orc_sparsity_nnz(const orc_model* m)
{
    vector<vector<int32_t>> adj;
    node_adjacency(*m, adj);
    int64_t nnz = 0;
    for (auto const& l : adj) nnz += 9 * int64_t(l.size());
    return nnz;
}

int orc_sparsity_fill(const orc_model* m, int32_t* row_ptr, int32_t* col)
{
    vector<vector<int32_t>> adj;
    node_adjacency(*m, adj);
    int64_t pos = 0;
    row_ptr[0] = 0;
    for (int32_t n = 0; n < m->N; ++n)
        for (int i = 0; i < 3; ++i)
        {
            for (int32_t b : adj[n])
                for (int j = 0; j < 3; ++j) col[pos++] = 3 * b + j;
            row_ptr[3 * n + i + 1] = int32_t(pos);
        }
    return ORC_OK;
}

// assemble_stiffness (static_matrix.hpp:243-277) with add_to's binary search and atomic
// add (thread_safe_plugin.hpp:7-23); compute_internal_force serial loop (static_matrix.hpp:189-203).
// submesh::consistent_mass (submesh.cpp:162-183): rho sum_l w_l N(l) N(l)^T det(X dN(l)) on the
// initial configuration, expanded to 3x3 identity blocks (identity_expansion_inplace<3>).
int orc_element_mass(const orc_model* m, int32_t element, double density, double* me)
{
    if (!m || !me || element < 0 || element >= m->E) return ORC_BAD_ARGUMENT;
    hex8 const& h = shape();
    double Xe[24];
    for (int a = 0; a < 8; ++a)
        for (int i = 0; i < 3; ++i) Xe[a * 3 + i] = m->X[3 * size_t(m->conn[8 * int64_t(element) + a]) + i];
    double mm[64] = {0};
    for (int l = 0; l < 8; ++l)
    {
        double J[9];
        local_F(Xe, h.dN[l], J);
        double const d = det3(J);
        for (int a = 0; a < 8; ++a)
            for (int b = 0; b < 8; ++b) mm[a * 8 + b] += density * h.N[l][a] * h.N[l][b] * d * h.w[l];
    }
    std::fill(me, me + 576, 0.0);
    for (int a = 0; a < 8; ++a)
        for (int b = 0; b < 8; ++b)
            for (int i = 0; i < 3; ++i) me[(a * 3 + i) * 24 + b * 3 + i] = mm[a * 8 + b];
    return ORC_OK;
}

// natural_frequency_matrix::assemble_mass (natural_frequency_matrix.hpp:106-137)
int orc_assemble_mass(const orc_model* m, const int32_t* row_ptr, const int32_t* col, const double* density,
                      double* val)
{
    if (!m || !row_ptr || !col || !density || !val) return ORC_BAD_ARGUMENT;
    int64_t const n = 3 * int64_t(m->N);
    std::fill(val, val + row_ptr[n], 0.0);
    double me[576];
    for (int64_t e = 0; e < m->E; ++e)
    {
        orc_element_mass(m, int32_t(e), density[e], me);
        int32_t dofs[24];
        for (int a = 0; a < 8; ++a)
            for (int i = 0; i < 3; ++i) dofs[a * 3 + i] = 3 * m->conn[8 * e + a] + i;
        for (int a = 0; a < 24; ++a)
            for (int b = 0; b < 24; ++b)
            {
                int32_t const r = dofs[a], c = dofs[b];
                const int32_t* lo = std::lower_bound(col + row_ptr[r], col + row_ptr[r + 1], c);
                val[lo - col] += me[a * 24 + b];
            }
    }
    return ORC_OK;
}

int orc_assemble(const orc_model* m, const int32_t* row_ptr, const int32_t* col, double* val, double* f_int,
                 int32_t threads)
{
    int64_t const n = 3 * int64_t(m->N);
    if (val)
    {
        std::fill(val, val + row_ptr[n], 0.0);
        if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
        {
            double ke[576], fe[24];
#pragma omp for schedule(dynamic, 64)
            for (int64_t e = 0; e < m->E; ++e)
            {
                element_matrices(*m, e, ke, fe);
                int32_t dofs[24];
                for (int a = 0; a < 8; ++a)
                    for (int i = 0; i < 3; ++i) dofs[a * 3 + i] = 3 * m->conn[8 * e + a] + i;
                for (int b = 0; b < 24; ++b)
                    for (int a = 0; a < 24; ++a)
                    {
                        int32_t const r = dofs[a], c = dofs[b];
                        const int32_t* lo = std::lower_bound(col + row_ptr[r], col + row_ptr[r + 1], c);
                        int64_t const p = lo - col;
#pragma omp atomic
                        val[p] += ke[a * 24 + b];
                    }
            }
        }
    }
    if (f_int)
    {
        std::fill(f_int, f_int + n, 0.0);
        double fe[24];
        for (int64_t e = 0; e < m->E; ++e)
        {
            element_matrices(*m, e, nullptr, fe);
            for (int a = 0; a < 8; ++a)
                for (int i = 0; i < 3; ++i) f_int[3 * int64_t(m->conn[8 * e + a]) + i] += fe[a * 3 + i];
        }
    }
    return ORC_OK;
}

// latin_submesh::incremental_latin_internal_force (latin_submesh.hpp:62-111) per element,
// summed as latin_matrix::compute_incremental_latin_internal_force (latin_matrix.hpp:110-126):
// at each QP sigma_global = sigma_old + s D (eps - eps_old) with D applied to the kinematic
// Voigt strain and read back as a kinetic Voigt stress, then int B^T (sigma_global - sigma)
// on the current configuration.
int orc_latin_force(const orc_model* m, double search_direction, double* f)
{
    if (!m || !f) return ORC_BAD_ARGUMENT;
    std::fill(f, f + 3 * int64_t(m->N), 0.0);
    static const int pair[6][2] = {{0, 0}, {1, 1}, {2, 2}, {1, 2}, {0, 2}, {0, 1}};
    for (int64_t e = 0; e < m->E; ++e)
    {
        double diff[72], fe[24];
        for (int l = 0; l < 8; ++l)
        {
            size_t const q = 8 * size_t(e) + l;
            const double* H = &m->Hg[9 * q];
            const double* Ho = &m->Hg_o[9 * q];
            double eps[6]; // kinematic Voigt: engineering shears
            for (int k = 0; k < 6; ++k)
            {
                int const i = pair[k][0], j = pair[k][1];
                double const cur = 0.5 * (H[i * 3 + j] + H[j * 3 + i]);
                double const old = 0.5 * (Ho[i * 3 + j] + Ho[j * 3 + i]);
                eps[k] = (k < 3 ? 1.0 : 2.0) * (cur - old);
            }
            double sg[6];
            for (int r = 0; r < 6; ++r)
            {
                double v = 0.0;
                for (int c = 0; c < 6; ++c) v += m->D[36 * q + r * 6 + c] * eps[c];
                sg[r] = v;
            }
            double* d = diff + 9 * l;
            for (int k = 0; k < 6; ++k)
            {
                int const i = pair[k][0], j = pair[k][1];
                double const v = m->sigma_o[9 * q + i * 3 + j] + search_direction * sg[k] - m->sigma[9 * q + i * 3 + j];
                d[i * 3 + j] = d[j * 3 + i] = v;
            }
        }
        element_matrices(*m, e, nullptr, fe, diff);
        for (int a = 0; a < 8; ++a)
            for (int i = 0; i < 3; ++i) f[3 * int64_t(m->conn[8 * e + a]) + i] += fe[a * 3 + i];
    }
    return ORC_OK;
}

// enforce_dirichlet_conditions (static_matrix.hpp:279-320)
int orc_enforce_dirichlet(int32_t n, const int32_t* row_ptr, const int32_t* col, double* val, double* b,
                          int32_t n_fixed, const int32_t* fixed)
{
    auto coeff_ref = [&](int32_t r, int32_t c) -> double* {
        const int32_t* lo = std::lower_bound(col + row_ptr[r], col + row_ptr[r + 1], c);
        if (lo == col + row_ptr[r + 1] || *lo != c) return nullptr;
        return &val[lo - col];
    };
    for (int32_t k = 0; k < n_fixed; ++k)
    {
        int32_t const d = fixed[k];
        if (d < 0 || d >= n) return ORC_BAD_ARGUMENT;
        double* dp = coeff_ref(d, d);
        double const diag = dp ? *dp : 0.0;
        if (b) b[d] = 0.0;
        for (int32_t p = row_ptr[d]; p < row_ptr[d + 1]; ++p)
        {
            val[p] = 0.0;
            if (double* q = coeff_ref(col[p], d)) *q = 0.0;
        }
        if (dp) *dp = diag;
    }
    return ORC_OK;
}

int orc_spmv(int32_t n, const int32_t* row_ptr, const int32_t* col, const double* val, const double* x,
             double* y, int32_t threads)
{
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(static)
    for (int32_t r = 0; r < n; ++r)
    {
        double s = 0.0;
        for (int32_t p = row_ptr[r]; p < row_ptr[r + 1]; ++p) s += val[p] * x[col[p]];
        y[r] = s;
    }
    return ORC_OK;
}

} // extern "C"

namespace
{
struct csr
{
    int32_t n;
    const int32_t* rp;
    const int32_t* ci;
    const double* v;
};

double dot(const vector<double>& a, const vector<double>& b)
{
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}

void spmv(const csr& A, const double* x, double* y)
{
#pragma omp parallel for schedule(static)
    for (int32_t r = 0; r < A.n; ++r)
    {
        double s = 0.0;
        for (int32_t p = A.rp[r]; p < A.rp[r + 1]; ++p) s += A.v[p] * x[A.ci[p]];
        y[r] = s;
    }
}

// Eigen DiagonalPreconditioner: 1/diag, or 1 where the diagonal is missing or zero.
vector<double> jacobi_inverse(const csr& A)
{
    vector<double> d(A.n, 1.0);
    for (int32_t r = 0; r < A.n; ++r)
        for (int32_t p = A.rp[r]; p < A.rp[r + 1]; ++p)
            if (A.ci[p] == r && A.v[p] != 0.0) d[r] = 1.0 / A.v[p];
    return d;
}
} // namespace

extern "C" {

// Eigen 3.3 internal::conjugate_gradient (IterativeLinearSolvers/ConjugateGradient.h), the
// algorithm behind conjugate_gradient::solve (src/solver/linear/linear_solver.cpp:67-109):
// x0 = 0, threshold tol^2 ||b||^2 on the unpreconditioned residual, iterations counted as
// Eigen counts them.
int orc_cg_eigen(int32_t n, const int32_t* row_ptr, const int32_t* col, const double* val, const double* b,
                 double* x, double tol, int32_t max_it, int32_t threads, int32_t* iterations, double* error)
{
    if (threads > 0) omp_set_num_threads(threads);
    std::feclearexcept(FE_ALL_EXCEPT);
    csr A{n, row_ptr, col, val};
    vector<double> invd = jacobi_inverse(A);
    vector<double> r(b, b + n), p(n), z(n), tmp(n), xv(n, 0.0);
    double const rhs2 = dot(r, r);
    int32_t it = 0;
    double err = 0.0;
    if (rhs2 == 0.0)
    {
        std::fill(x, x + n, 0.0);
        *iterations = 0;
        *error = 0.0;
        return ORC_OK;
    }
    double const threshold = std::max(tol * tol * rhs2, std::numeric_limits<double>::min());
    double rn2 = rhs2;
    if (rn2 < threshold)
    {
        err = std::sqrt(rn2 / rhs2);
    }
    else
    {
        for (int32_t i = 0; i < n; ++i) p[i] = invd[i] * r[i];
        double absNew = dot(r, p);
        while (it < max_it)
        {
            spmv(A, p.data(), tmp.data());
            double const alpha = absNew / dot(p, tmp);
#pragma omp parallel for schedule(static)
            for (int32_t i = 0; i < n; ++i)
            {
                xv[i] += alpha * p[i];
                r[i] -= alpha * tmp[i];
            }
            rn2 = dot(r, r);
            if (rn2 < threshold) break;
            for (int32_t i = 0; i < n; ++i) z[i] = invd[i] * r[i];
            double const absOld = absNew;
            absNew = dot(r, z);
            double const beta = absNew / absOld;
#pragma omp parallel for schedule(static)
            for (int32_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
            ++it;
        }
        err = std::sqrt(rn2 / rhs2);
    }
    std::copy(xv.begin(), xv.end(), x);
    *iterations = it;
    *error = err;
    if (std::fetestexcept(FE_INVALID)) return ORC_NAN_OR_INF; // linear_solver.cpp:99-102
    if (it >= max_it) return ORC_NONCONVERGED;                 // linear_solver.cpp:104-108
    return ORC_OK;
}

// conjugate_gradient_cuda::solve (src/solver/linear/conjugate_gradient_cuda.cpp:95-186)
int orc_cg_cuda(int32_t n, const int32_t* row_ptr, const int32_t* col, const double* val, const double* b,
                double* x, double tol, int32_t max_it, int32_t threads, int32_t* iterations, double* error)
{
    if (threads > 0) omp_set_num_threads(threads);
    csr A{n, row_ptr, col, val};
    vector<double> minv(n);
    for (int32_t r = 0; r < n; ++r)
    {
        double d = 0.0;
        for (int32_t p = row_ptr[r]; p < row_ptr[r + 1]; ++p)
            if (col[p] == r) d = val[p];
        minv[r] = 1.0 / d; // A.diagonal().cwiseInverse()
    }
    vector<double> xv(n, 0.0), r(b, b + n), z(n), p(n), Ap(n);
    for (int32_t i = 0; i < n; ++i) z[i] = minv[i] * r[i];
    double residual = dot(r, z), residual_old = 0.0;
    int32_t k = 0;
    while (std::sqrt(residual) > tol && k < max_it)
    {
        ++k;
        if (k == 1)
            p = z;
        else
        {
            double const beta = residual / residual_old;
            for (int32_t i = 0; i < n; ++i) p[i] = beta * p[i] + z[i];
        }
        spmv(A, p.data(), Ap.data());
        double const alpha = residual / dot(p, Ap);
        for (int32_t i = 0; i < n; ++i)
        {
            xv[i] += alpha * p[i];
            r[i] += -alpha * Ap[i];
        }
        residual_old = residual;
        for (int32_t i = 0; i < n; ++i) z[i] = minv[i] * r[i];
        residual = dot(z, r);
    }
    std::copy(xv.begin(), xv.end(), x);
    *iterations = k;
    *error = std::sqrt(residual);
    if (k >= max_it) return ORC_NONCONVERGED;
    return ORC_OK;
}

// Eigen 3.3 internal::bicgstab (IterativeLinearSolvers/BiCGSTAB.h) with its restart rule,
// behind biconjugate_gradient_stabilised::solve (linear_solver.cpp:111-138).
int orc_bicgstab_eigen(int32_t n, const int32_t* row_ptr, const int32_t* col, const double* val,
                       const double* b, double* x, double tol, int32_t max_it, int32_t threads,
                       int32_t* iterations, double* error)
{
    if (threads > 0) omp_set_num_threads(threads);
    std::feclearexcept(FE_ALL_EXCEPT);
    csr A{n, row_ptr, col, val};
    vector<double> invd = jacobi_inverse(A);
    vector<double> xv(n, 0.0), r(b, b + n), r0 = r, v(n, 0.0), p(n, 0.0), y(n), z(n), s(n), t(n);
    double r0_sqnorm = dot(r0, r0);
    double const rhs_sqnorm = dot(r, r);
    if (rhs_sqnorm == 0.0)
    {
        std::fill(x, x + n, 0.0);
        *iterations = 0;
        *error = 0.0;
        return ORC_OK;
    }
    double rho = 1, alpha = 1, w = 1;
    double const tol2 = tol * tol * rhs_sqnorm;
    double const eps2 = std::numeric_limits<double>::epsilon() * std::numeric_limits<double>::epsilon();
    int32_t i = 0, restarts = 0;
    while (dot(r, r) > tol2 && i < max_it)
    {
        double const rho_old = rho;
        rho = dot(r0, r);
        if (std::abs(rho) < eps2 * r0_sqnorm)
        {
            spmv(A, xv.data(), t.data());
            for (int32_t k = 0; k < n; ++k) r[k] = b[k] - t[k];
            r0 = r;
            rho = r0_sqnorm = dot(r, r);
            if (restarts++ == 0) i = 0;
        }
        double const beta = (rho / rho_old) * (alpha / w);
        for (int32_t k = 0; k < n; ++k) p[k] = r[k] + beta * (p[k] - w * v[k]);
        for (int32_t k = 0; k < n; ++k) y[k] = invd[k] * p[k];
        spmv(A, y.data(), v.data());
        alpha = rho / dot(r0, v);
        for (int32_t k = 0; k < n; ++k) s[k] = r[k] - alpha * v[k];
        for (int32_t k = 0; k < n; ++k) z[k] = invd[k] * s[k];
        spmv(A, z.data(), t.data());
        double const tmp = dot(t, t);
        w = tmp > 0.0 ? dot(t, s) / tmp : 0.0;
        for (int32_t k = 0; k < n; ++k)
        {
            xv[k] += alpha * y[k] + w * z[k];
            r[k] = s[k] - w * t[k];
        }
        ++i;
    }
    std::copy(xv.begin(), xv.end(), x);
    *iterations = i;
    *error = std::sqrt(dot(r, r) / rhs_sqnorm);
    if (std::fetestexcept(FE_INVALID)) return ORC_NAN_OR_INF;
    if (i >= max_it) return ORC_NONCONVERGED;
    return ORC_OK;
}

// biconjugate_gradient_stabilised_cuda::solve (biconjugate_gradient_stabilised_cuda.cpp:31-200)
int orc_bicgstab_cuda(int32_t n, const int32_t* row_ptr, const int32_t* col, const double* val,
                      const double* b, double* x, double tol, int32_t max_it, int32_t threads,
                      int32_t* iterations, double* error)
{
    if (threads > 0) omp_set_num_threads(threads);
    csr A{n, row_ptr, col, val};
    vector<double> minv(n);
    for (int32_t r = 0; r < n; ++r)
    {
        double d = 0.0;
        for (int32_t p = row_ptr[r]; p < row_ptr[r + 1]; ++p)
            if (col[p] == r) d = val[p];
        minv[r] = 1.0 / d;
    }
    vector<double> xv(n, 0.0), r(b, b + n), r0 = r, p(n, 0.0), ph(n), v(n, 0.0), s(n), sh(n), t(n);
    int32_t k = 0;
    double residual_old = 0.0, residual = dot(r, r0);
    double alpha = 0.0, omega = 0.0;
    int status = ORC_OK;
    while (std::sqrt(residual) > tol && k < max_it)
    {
        ++k;
        if (k == 1)
            p = r;
        else
        {
            double const beta = residual / residual_old * alpha / omega;
            for (int32_t i = 0; i < n; ++i)
            {
                v[i] = omega * v[i];
                p[i] = p[i] - v[i];
                p[i] = beta * p[i];
                p[i] = p[i] + r[i];
            }
        }
        for (int32_t i = 0; i < n; ++i) ph[i] = minv[i] * p[i];
        spmv(A, ph.data(), v.data());
        alpha = residual / dot(r0, v);
        for (int32_t i = 0; i < n; ++i) s[i] = r[i] + (-alpha) * v[i];
        double const norm_s = std::sqrt(dot(s, s));
        if (norm_s < tol)
        {
            for (int32_t i = 0; i < n; ++i) xv[i] += alpha * ph[i];
            residual = norm_s * norm_s;
            break;
        }
        for (int32_t i = 0; i < n; ++i) sh[i] = minv[i] * s[i];
        spmv(A, sh.data(), t.data());
        double const t_dot_s = dot(t, s), t_dot_t = dot(t, t);
        omega = t_dot_s / t_dot_t;
        for (int32_t i = 0; i < n; ++i)
        {
            xv[i] += alpha * ph[i];
            xv[i] += omega * sh[i];
            r[i] = s[i] + (-omega) * t[i];
        }
        residual_old = residual;
        residual = dot(r0, r);
        if (omega == 0.0) // is_approx(omega, 0.0): relative difference to 0 is never <= 1e-10
        {
            status = ORC_BREAKDOWN;
            break;
        }
    }
    std::copy(xv.begin(), xv.end(), x);
    *iterations = k;
    *error = std::sqrt(residual);
    if (status != ORC_OK) return status;
    if (k >= max_it) return ORC_NONCONVERGED;
    return ORC_OK;
}

// reverse_cuthill_mcgee::compute (src/graph/cuthill_mckee.hpp:45-97) over the adjacency
// graph of src/graph/adjacency_graph.hpp (diagonal removed, children sorted by degree).
int orc_rcm(int32_t n, const int32_t* row_ptr, const int32_t* col, int32_t* perm)
{
    vector<vector<int32_t>> adj(n);
    for (int32_t r = 0; r < n; ++r)
        for (int32_t p = row_ptr[r]; p < row_ptr[r + 1]; ++p)
            if (col[p] != r) adj[r].push_back(col[p]);
    for (auto& l : adj)
        std::stable_sort(l.begin(), l.end(),
                         [&](int32_t a, int32_t b) { return adj[a].size() < adj[b].size(); });
    vector<char> done(n, 0);
    int32_t count = 0;
    std::queue<int32_t> q;
    while (count != n)
    {
        int32_t best = n;
        size_t deg = size_t(n);
        for (int32_t r = 0; r < n; ++r)
            if (adj[r].size() < deg && !done[r])
            {
                deg = adj[r].size();
                best = r;
            }
        perm[count++] = best;
        done[best] = 1;
        for (int32_t c : adj[best])
            if (!done[c]) q.push(c);
        while (!q.empty())
        {
            int32_t const f = q.front();
            if (!done[f])
            {
                done[f] = 1;
                perm[count++] = f;
                for (int32_t c : adj[f])
                    if (!done[c]) q.push(c);
            }
            q.pop();
        }
    }
    std::reverse(perm, perm + n);
    return ORC_OK;
}

} // extern "C"

// ============================================================ load stepping
namespace
{
// is_approx (src/numeric/float_compare.hpp:20-32) with boost's relative_difference:
// relative difference to zero is the largest double, so only the absolute test can pass.
bool is_approx(double x, double y, double abs_tol = 0.0, double rel_tol = 1e-10)
{
    double rel;
    if (x == 0.0 && y == 0.0)
        rel = 0.0;
    else if (x == 0.0 || y == 0.0)
        rel = std::numeric_limits<double>::max();
    else
        rel = std::abs(x - y) / std::min(std::abs(x), std::abs(y));
    return rel <= rel_tol || std::abs(x - y) <= abs_tol;
}
} // namespace

// adaptive_time_step (src/solver/adaptive_time_step.cpp:12-165, adaptive_time_step.hpp:72-137)
struct orc_time
{
    std::priority_queue<double, vector<double>, std::greater<double>> queue;
    int32_t increment_limit = 5, successful = 0, consecutive_converged = 0, consecutive_unconverged = 0;
    double initial = 1, final_time = 1, current = 1, total = 0, last_converged = 0, last_size = 0;
    double minimum = 0, maximum = 0;
    bool applied = false;
    bool highly_nonlinear() const { return consecutive_unconverged > 0 || consecutive_converged < 4; }
};

extern "C" {

int orc_time_create(const orc_time_options* opt, int32_t n_hist, const double* history, orc_time** out)
{
    if (!opt || n_hist <= 0 || !out) return ORC_BAD_ARGUMENT;
    auto* t = new orc_time;
    for (int32_t i = 0; i < n_hist; ++i) t->queue.push(history[i]);
    if (is_approx(t->queue.top(), 0.0)) t->queue.pop();
    t->initial = opt->initial;
    t->minimum = opt->adaptive ? opt->minimum : opt->initial;
    t->maximum = opt->adaptive ? opt->maximum : opt->initial;
    t->final_time = opt->period;
    t->current = std::min(t->queue.top(), t->initial);
    *out = t;
    return ORC_OK;
}

int orc_time_destroy(orc_time* t)
{
    delete t;
    return ORC_OK;
}

int orc_time_update(orc_time* t, int32_t converged)
{
    if (converged)
    {
        if (t->current >= t->queue.top()) t->queue.pop();
        t->last_size = t->current - t->last_converged;
        t->last_converged = t->current;
        t->applied = is_approx(t->current, t->final_time) || t->queue.empty();
        t->consecutive_unconverged = 0;
        t->consecutive_converged++;
        t->successful++;
        if (t->applied) return ORC_OK;
        double const new_time = std::min(t->highly_nonlinear() ? t->last_size + t->current : 2.0 * t->current,
                                          t->current + t->maximum);
        t->current = std::min(t->queue.top(), std::min(new_time, t->final_time));
        return ORC_OK;
    }
    double const dt = std::max(t->minimum, (t->current - t->last_converged) * 0.5);
    t->current = t->last_converged + dt;
    if (t->current > t->final_time) return ORC_FATAL;
    if (t->current < 0.0) return ORC_FATAL;
    t->consecutive_unconverged++;
    t->consecutive_converged = 0;
    if (t->consecutive_unconverged == t->increment_limit) return ORC_FATAL;
    return ORC_OK;
}

int orc_time_query(const orc_time* t, double* step_time, double* last_step_time, int32_t* fully_applied,
                   int32_t* successful)
{
    if (step_time) *step_time = t->current;
    if (last_step_time) *last_step_time = t->last_converged;
    if (fully_applied) *fully_applied = t->applied;
    if (successful) *successful = t->successful;
    return ORC_OK;
}

} // extern "C"

// ============================================================= Newton driver
namespace
{
// boundary::is_active / interpolate_prescribed_load (src/mesh/boundary/boundary.cpp:37-85)
bool bc_active(const orc_dirichlet& bc, double t) { return t >= bc.times[0] && t <= bc.times[bc.n_points - 1]; }
double bc_value(const orc_dirichlet& bc, double t)
{
    for (int32_t i = 0; i < bc.n_points; ++i)
        if (is_approx(bc.times[i], t)) return bc.values[i];
    for (int32_t i = 0; i + 1 < bc.n_points; ++i)
        if (bc.times[i] < t && t < bc.times[i + 1])
        {
            double const f = (t - bc.times[i]) / (bc.times[i + 1] - bc.times[i]);
            return (1.0 - f) * bc.values[i] + f * bc.values[i + 1];
        }
    return std::numeric_limits<double>::quiet_NaN();
}

double norm(const vector<double>& v) { return std::sqrt(dot(v, v)); }

// P^T K P and P^T b with perm[new] = old (apply_permutation, linear_solver.cpp:39-44)
void permute(int32_t n, const int32_t* rp, const int32_t* ci, const double* v, const vector<int32_t>& perm,
             const vector<int32_t>& inv, vector<int32_t>& prp, vector<int32_t>& pci, vector<double>& pv)
{
    prp.assign(n + 1, 0);
    for (int32_t r = 0; r < n; ++r) prp[r + 1] = prp[r] + (rp[perm[r] + 1] - rp[perm[r]]);
    pci.resize(rp[n]);
    pv.resize(rp[n]);
#pragma omp parallel for schedule(static)
    for (int32_t r = 0; r < n; ++r)
    {
        int32_t const o = perm[r];
        int32_t const len = rp[o + 1] - rp[o];
        vector<std::pair<int32_t, double>> row(len);
        for (int32_t k = 0; k < len; ++k) row[k] = {inv[ci[rp[o] + k]], v[rp[o] + k]};
        std::sort(row.begin(), row.end(), [](auto& a, auto& b) { return a.first < b.first; });
        for (int32_t k = 0; k < len; ++k)
        {
            pci[prp[r] + k] = row[k].first;
            pv[prp[r] + k] = row[k].second;
        }
    }
}
} // namespace

extern "C" {

double orc_boundary_value(int32_t n_points, const double* times, const double* values, double t)
{
    orc_dirichlet bc{0, nullptr, n_points, times, values};
    return bc_value(bc, t);
}

int orc_newton_solve(orc_model* m, const orc_newton_options* opt, int32_t n_bc, const orc_dirichlet* bcs,
                     double* u_out, orc_newton_record* records, int32_t cap, int32_t* n_records)
{
    return orc_newton_solve_loaded(m, opt, n_bc, bcs, 0, nullptr, u_out, records, cap, n_records);
}

// static_matrix::solve / perform_equilibrium_iterations (static_matrix.hpp:145-187, 402-467)
int orc_newton_solve_loaded(orc_model* m, const orc_newton_options* opt, int32_t n_bc, const orc_dirichlet* bcs,
                            int32_t n_loads, const orc_neumann* loads, double* u_out, orc_newton_record* records,
                            int32_t cap, int32_t* n_records)
{
    if (!m || !opt || (n_bc > 0 && !bcs) || !u_out || n_loads < 0 || (n_loads > 0 && !loads)) return ORC_BAD_ARGUMENT;
    for (int32_t k = 0; k < n_loads; ++k)
    {
        if (loads[k].n_points < 1 || !loads[k].times || !loads[k].loads || loads[k].n_entries < 0
            || (loads[k].n_entries > 0 && (!loads[k].dofs || !loads[k].values)))
            return ORC_BAD_ARGUMENT;
        for (int32_t i = 0; i < loads[k].n_entries; ++i)
            if (loads[k].dofs[i] < 0 || loads[k].dofs[i] >= 3 * m->N) return ORC_BAD_ARGUMENT;
    }
    int32_t const threads = opt->threads;
    if (threads > 0) omp_set_num_threads(threads);
    int32_t const n = 3 * m->N;
    int32_t nrec = 0;

    vector<double> hist;
    for (int32_t k = 0; k < n_bc; ++k)
        for (int32_t i = 0; i < bcs[k].n_points; ++i) hist.push_back(bcs[k].times[i]);
    for (int32_t k = 0; k < n_loads; ++k)
        for (int32_t i = 0; i < loads[k].n_points; ++i) hist.push_back(loads[k].times[i]);
    // mesh::time_history collects the boundary and load times into a std::set (mesh.cpp:146-172)
    std::sort(hist.begin(), hist.end());
    hist.erase(std::unique(hist.begin(), hist.end()), hist.end());
    orc_time* ts = nullptr;
    if (orc_time_create(&opt->time, int32_t(hist.size()), hist.data(), &ts) != ORC_OK) return ORC_BAD_ARGUMENT;

    int64_t const nnz = orc_sparsity_nnz(m);
    vector<int32_t> rp(n + 1), ci(nnz);
    orc_sparsity_fill(m, rp.data(), ci.data());
    vector<double> K(nnz);
    vector<double> f_int(n, 0.0), f_ext(n, 0.0), u(n, 0.0), u_old(n, 0.0), delta(n, 0.0), mres(n, 0.0), lres;
    if (opt->record_trajectory)
    {
        m->traj_u.clear();
        m->traj_ep.clear();
        m->traj_eb.clear();
    }
    vector<int32_t> perm, inv;
    vector<int32_t> prp, pci;
    vector<double> pv, pb, px;

    bool const relative = !opt->absolute_norm;
    int const max_it = opt->max_newton_iterations > 0 ? opt->max_newton_iterations : 10;
    double norm_initial_residual = 1.0;
    int result = ORC_OK;

    auto record = [&](int it, int solver_its, int status, double dn, double fn) {
        if (records && nrec < cap)
        {
            double st;
            int32_t succ;
            orc_time_query(ts, &st, nullptr, nullptr, &succ);
            records[nrec] = orc_newton_record{succ, it, solver_its, status, st, dn, fn};
        }
        ++nrec;
    };

    for (;;) // solve() and its recursive retry after a computational_error
    {
        int err = orc_update_internal_variables(m, u.data(), 0.0);
        bool recovered = false;
        while (err == ORC_OK)
        {
            int32_t applied;
            orc_time_query(ts, nullptr, nullptr, &applied, nullptr);
            if (applied) break;
            // perform_equilibrium_iterations
            u = u_old;
            double step_time, last_time;
            orc_time_query(ts, &step_time, &last_time, nullptr, nullptr);
            // compute_external_force (static_matrix.hpp:205-241) at the step time
            std::fill(f_ext.begin(), f_ext.end(), 0.0);
            bool load_outside = false;
            for (int32_t k = 0; k < n_loads; ++k)
            {
                double const at = loads[k].fixed_time >= 0.0 ? loads[k].fixed_time : step_time;
                orc_dirichlet const h{0, nullptr, loads[k].n_points, loads[k].times, loads[k].loads};
                if (!(at >= h.times[0] && at <= h.times[h.n_points - 1]))
                {
                    load_outside = true;
                    break;
                }
                double const v = bc_value(h, at);
                for (int32_t i = 0; i < loads[k].n_entries; ++i) f_ext[loads[k].dofs[i]] += v * loads[k].values[i];
            }
            if (load_outside)
            {
                err = ORC_FATAL;
                break;
            }
            err = orc_update_internal_variables(m, u.data(), step_time - last_time);
            if (err != ORC_OK) break;
            int it = 0;
            double dnorm = 0, fnorm = 0;
            while (it < max_it)
            {
                orc_assemble(m, rp.data(), ci.data(), K.data(), f_int.data(), threads);
                for (int32_t i = 0; i < n; ++i) mres[i] = f_ext[i] - f_int[i];
                vector<int32_t> fixed;
                if (it == 0) // apply_displacement_boundaries (static_matrix.hpp:322-351)
                {
                    vector<double> inc(n, 0.0), Kinc(n);
                    vector<char> touched(n, 0);
                    for (int32_t k = 0; k < n_bc; ++k)
                    {
                        if (!bc_active(bcs[k], step_time)) continue;
                        double const du = bc_value(bcs[k], step_time) - bc_value(bcs[k], last_time);
                        for (int32_t i = 0; i < bcs[k].n_dofs; ++i)
                        {
                            inc[bcs[k].dofs[i]] = du;
                            touched[bcs[k].dofs[i]] = 1;
                        }
                    }
                    spmv(csr{n, rp.data(), ci.data(), K.data()}, inc.data(), Kinc.data());
                    for (int32_t i = 0; i < n; ++i)
                    {
                        mres[i] -= Kinc[i];
                        u[i] += inc[i];
                    }
                    norm_initial_residual = norm(mres);
                }
                for (int32_t k = 0; k < n_bc; ++k)
                    if (bc_active(bcs[k], step_time)) fixed.insert(fixed.end(), bcs[k].dofs, bcs[k].dofs + bcs[k].n_dofs);
                orc_enforce_dirichlet(n, rp.data(), ci.data(), K.data(), mres.data(), int32_t(fixed.size()),
                                      fixed.data());
                // latin_matrix::perform_equilibrium_iterations (latin_matrix.hpp:209-226): after
                // the first iteration the right-hand side is the LATIN force, its fixed entries
                // zeroed with the residual's
                const double* rhs = mres.data();
                if (opt->latin && it > 0)
                {
                    lres.resize(n);
                    orc_latin_force(m, opt->latin_search_direction, lres.data());
                    for (int32_t d : fixed) lres[d] = 0.0;
                    rhs = lres.data();
                }

                int32_t its = 0;
                double e = 0;
                int s;
                if (opt->rcm)
                {
                    if (perm.empty())
                    {
                        perm.resize(n);
                        inv.resize(n);
                        orc_rcm(n, rp.data(), ci.data(), perm.data());
                        for (int32_t i = 0; i < n; ++i) inv[perm[i]] = i;
                    }
                    permute(n, rp.data(), ci.data(), K.data(), perm, inv, prp, pci, pv);
                    pb.resize(n);
                    px.resize(n);
                    for (int32_t i = 0; i < n; ++i) pb[i] = rhs[perm[i]];
                }
                const int32_t* srp = opt->rcm ? prp.data() : rp.data();
                const int32_t* sci = opt->rcm ? pci.data() : ci.data();
                const double* sv = opt->rcm ? pv.data() : K.data();
                const double* sb = opt->rcm ? pb.data() : rhs;
                double* sx = opt->rcm ? px.data() : delta.data();
                if (opt->solver == 0)
                    s = opt->solver_criterion == ORC_STOP_RELATIVE
                            ? orc_cg_eigen(n, srp, sci, sv, sb, sx, opt->solver_tolerance, opt->solver_max_iterations,
                                           threads, &its, &e)
                            : orc_cg_cuda(n, srp, sci, sv, sb, sx, opt->solver_tolerance, opt->solver_max_iterations,
                                          threads, &its, &e);
                else
                    s = opt->solver_criterion == ORC_STOP_RELATIVE
                            ? orc_bicgstab_eigen(n, srp, sci, sv, sb, sx, opt->solver_tolerance,
                                                 opt->solver_max_iterations, threads, &its, &e)
                            : orc_bicgstab_cuda(n, srp, sci, sv, sb, sx, opt->solver_tolerance,
                                                opt->solver_max_iterations, threads, &its, &e);
                if (opt->rcm)
                    for (int32_t i = 0; i < n; ++i) delta[perm[i]] = px[i];
                if (s != ORC_OK)
                {
                    record(it, its, s, 0, 0);
                    err = s;
                    break;
                }
                for (int32_t i = 0; i < n; ++i) u[i] += delta[i];
                err = orc_update_internal_variables(m, u.data(), 0.0);
                if (err != ORC_OK)
                {
                    record(it, its, err, 0, 0);
                    break;
                }
                // update_relative_norms (static_matrix.hpp:385-400)
                double const un = norm(u), fe = norm(f_ext), fi = norm(f_int);
                bool const use_rel = opt->norm_mode == 1
                                         ? true
                                         : (relative && !is_approx(un, 0.0, 1.0) && !is_approx(std::max(fe, fi), 0.0, 1.0));
                if (use_rel)
                {
                    dnorm = norm(delta) / un;
                    fnorm = norm(mres) / std::max(norm_initial_residual, std::max(fe, fi));
                }
                else
                {
                    dnorm = norm(delta);
                    fnorm = norm(mres);
                }
                record(it, its, 0, dnorm, fnorm);
                if (dnorm <= opt->displacement_tolerance && fnorm <= opt->residual_tolerance) break;
                ++it;
            }
            if (err != ORC_OK) break;
            if (it == max_it)
            {
                err = ORC_NONCONVERGED;
                break;
            }
            u_old = u;
            if (opt->record_trajectory)
            {
                m->traj_u.push_back(u);
                m->traj_ep.push_back(m->eps_p_o);
                m->traj_eb.push_back(m->ebar_o);
            }
            orc_time_update(ts, 1);
            orc_save_internal_variables(m, 1);
        }
        if (err == ORC_OK) break;
        if (err < 0)
        {
            result = err;
            break;
        }
        // catch (computational_error&): cutback, revert, reset u, retry (static_matrix.hpp:168-182)
        if (orc_time_update(ts, 0) != ORC_OK)
        {
            result = ORC_FATAL;
            break;
        }
        orc_save_internal_variables(m, 0);
        u = u_old;
        recovered = true;
        (void)recovered;
    }
    std::copy(u.begin(), u.end(), u_out);
    if (n_records) *n_records = nrec;
    orc_time_destroy(ts);
    return result;
}

} // extern "C"

namespace
{
void reference_gradients(const orc_model& m, int64_t e, int l, double (*Bt)[3], double& wdetJ)
{
    hex8 const& h = shape();
    double Xe[24];
    for (int a = 0; a < 8; ++a)
        for (int i = 0; i < 3; ++i) Xe[a * 3 + i] = m.X[3 * size_t(m.conn[8 * e + a]) + i];
    double J[9], Ji[9];
    local_F(Xe, h.dN[l], J);
    inv3(J, Ji);
    wdetJ = h.w[l] * det3(J);
    for (int a = 0; a < 8; ++a)
        for (int k = 0; k < 3; ++k)
        {
            double s = 0.0;
            for (int j = 0; j < 3; ++j) s += h.dN[l][a][j] * Ji[j * 3 + k];
            Bt[a][k] = s;
        }
}
} // namespace

extern "C" {

int orc_trajectory_steps(const orc_model* m) { return m ? int(m->traj_u.size()) : 0; }

// Discrete adjoint of the load history (consistent mode: the return map of step n starts
// from the committed history h_{n-1} and strains are small on X).  With
//   R_n(u_n, h_{n-1}) = f_int - f_ext = 0,  h_n = H(eps(u_n), h_{n-1}),
// and multipliers lambda_n (dofs) and mu_n (per QP history), going backwards from n = N:
//   K_n lambda_n = dJ/du [n = N] + sum_QP B^T (dH_n/d eps)^T mu_n,
//   mu_{n-1} = (dH_n/dh_{n-1})^T mu_n - (d sigma_n/dh_{n-1})^T (w detJ B lambda_n),   mu_N = 0,
//   dJ/dp += (dH_n/dp)^T mu_n - (d sigma_n/dp)^T (w detJ B lambda_n),
// every partial derivative from j2_map_vjp.
int orc_adjoint_j2(orc_model* m, const double* dJdu, int32_t n_fixed, const int32_t* fixed, double tolerance,
                   double* grad)
{
    if (!m || !dJdu || !grad || (n_fixed > 0 && !fixed)) return ORC_BAD_ARGUMENT;
    if (m->model != ORC_MODEL_J2 || m->tangent_mode != ORC_TANGENT_CONSISTENT) return ORC_BAD_ARGUMENT;
    int const steps = int(m->traj_u.size());
    if (steps == 0) return ORC_BAD_ARGUMENT;
    int32_t const n = 3 * m->N;
    size_t const Q = 8 * size_t(m->E);
    // final state, restored at the end
    vector<double> const ep_final = m->eps_p_o, eb_final = m->ebar_o, u_final = m->traj_u.back();

    int64_t const nnz = orc_sparsity_nnz(m);
    vector<int32_t> rp(n + 1), ci(nnz);
    orc_sparsity_fill(m, rp.data(), ci.data());
    vector<double> K(nnz), f(n), rhs(n), lam(n);
    vector<double> mu_ep(9 * Q, 0.0), mu_eb(Q, 0.0);
    std::fill(grad, grad + 4 * size_t(m->E), 0.0);
    for (int step = steps - 1; step >= 0; --step)
    {
        m->eps_p_o = m->traj_ep[step];
        m->ebar_o = m->traj_eb[step];
        int const st = orc_update_internal_variables(m, m->traj_u[step].data(), 0.0);
        if (st != ORC_OK) return st;
        orc_assemble(m, rp.data(), ci.data(), K.data(), f.data(), 0);
        // right-hand side: dJ/du at the last step, plus the history multipliers' pull-back
        for (int32_t i = 0; i < n; ++i) rhs[i] = step == steps - 1 ? dJdu[i] : 0.0;
        if (step < steps - 1)
            for (int64_t e = 0; e < m->E; ++e)
                for (int l = 0; l < 8; ++l)
                {
                    size_t const q = 8 * size_t(e) + l;
                    double Bt[8][3], wdetJ;
                    reference_gradients(*m, e, l, Bt, wdetJ);
                    double strain[9], zero[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                    const double* H = &m->Hg[9 * q];
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j) strain[3 * i + j] = 0.5 * (H[3 * i + j] + H[3 * j + i]);
                    double E[9], epb[9], ebb, pb[4];
                    j2_map_vjp(strain, &m->eps_p_o[9 * q], m->ebar_o[q], &m->params[4 * e], zero, &mu_ep[9 * q],
                               mu_eb[q], E, epb, &ebb, pb);
                    for (int a = 0; a < 8; ++a)
                        for (int i = 0; i < 3; ++i)
                        {
                            double s = 0.0;
                            for (int j = 0; j < 3; ++j) s += E[3 * i + j] * Bt[a][j];
                            rhs[3 * size_t(m->conn[8 * e + a]) + i] += s;
                        }
                }
        orc_enforce_dirichlet(n, rp.data(), ci.data(), K.data(), rhs.data(), n_fixed, fixed);
        int32_t its = 0;
        double err = 0.0;
        int const cg = orc_cg_eigen(n, rp.data(), ci.data(), K.data(), rhs.data(), lam.data(), tolerance, 1000000, 0,
                                    &its, &err);
        if (cg != ORC_OK) return cg;
        // history multipliers of the previous step and the parameter gradient
        for (int64_t e = 0; e < m->E; ++e)
            for (int l = 0; l < 8; ++l)
            {
                size_t const q = 8 * size_t(e) + l;
                double Bt[8][3], wdetJ;
                reference_gradients(*m, e, l, Bt, wdetJ);
                double gl[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                for (int a = 0; a < 8; ++a)
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j) gl[3 * i + j] += lam[3 * size_t(m->conn[8 * e + a]) + i] * Bt[a][j];
                double sb[9], strain[9];
                const double* H = &m->Hg[9 * q];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j)
                    {
                        sb[3 * i + j] = -wdetJ * 0.5 * (gl[3 * i + j] + gl[3 * j + i]);
                        strain[3 * i + j] = 0.5 * (H[3 * i + j] + H[3 * j + i]);
                    }
                double E[9], epb[9], ebb, pb[4];
                j2_map_vjp(strain, &m->eps_p_o[9 * q], m->ebar_o[q], &m->params[4 * e], sb, &mu_ep[9 * q], mu_eb[q],
                           E, epb, &ebb, pb);
                for (int i = 0; i < 9; ++i) mu_ep[9 * q + i] = epb[i];
                mu_eb[q] = ebb;
                for (int k = 0; k < 4; ++k) grad[4 * e + k] += pb[k];
            }
    }
    m->eps_p_o = ep_final;
    m->ebar_o = eb_final;
    return orc_update_internal_variables(m, u_final.data(), 0.0);
}

} // extern "C"
