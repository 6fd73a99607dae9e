// Device kernel of small_strain_J2_plasticity_damage (viscoplastic J2 with nonlinear kinematic
// hardening and Lemaitre-Chaboche scalar damage), one thread per quadrature point.
//
// Reference behaviour (paths relative to the neon repository root):
//   update_internal_variables  src/constitutive/mechanics/solid/small_strain_J2_plasticity_damage.cpp:34-120
//   perform_radial_return      :122-250 (16 unknowns: plastic and damage multipliers, kinetic
//                              stress (6), kinetic back stress (6), damage, energy release rate)
//   yield functions            :252-265; energy release rate :267-285
//   material constants         src/material/isotropic_elastic_plastic_damage.cpp:9-45
//   Voigt / Mandel conventions src/numeric/tensor_operations.hpp:207-214, 373-376, 425-439, 509-515
//
// The return map's 16 x 16 Newton systems are solved by Gaussian elimination with complete
// pivoting (the algorithm of Eigen's FullPivLU the reference calls), and the tangent takes the
// 6 x 6 stress block of the last iteration's inverse from one factorisation and six solves.  The
// system lives in per-thread private memory (about 2.3 KB); only yielding points run it.
// History (plastic strain, effective plastic strain, back stress, damage) is updated in place
// at every call, as neon does; the tangent is general (36 coefficients, row-major 6 x 6).
#pragma once

#include "element_kernels.hpp"

namespace neon_amd
{
constexpr int kGeneralCoefs = 36;

// isotropic_elastic_plastic_damage constants: kinematic hardening modulus C, softening multiplier
// gamma, plasticity viscous denominator s_p and exponent n_p, damage viscous denominator s_d and
// exponent n_d
struct damage_constants
{
    double C, gamma, sp, np, sd, nd;
};

// J2 damage state beyond the J2 arrays of qp_state (stride Q): back stress 6 (tensor components
// 11, 22, 33, 23, 13, 12), damage 1, energy release rate 1
struct damage_state
{
    double* beta;
    double* dmg;
    double* energy;
};

namespace damage_detail
{
// kinetic Voigt (11, 22, 33, 23, 13, 12) <-> 3 x 3 row-major
__device__ inline void kin(const double* a, double* v)
{
    v[0] = a[0], v[1] = a[4], v[2] = a[8], v[3] = a[5], v[4] = a[2], v[5] = a[1];
}
__device__ inline void kin_from(const double* v, double* a)
{
    a[0] = v[0], a[4] = v[1], a[8] = v[2];
    a[5] = a[7] = v[3];
    a[2] = a[6] = v[4];
    a[1] = a[3] = v[5];
}
// C * kinematic(e) for the isotropic C_e (lambda, G) scaled by s: the stress-like vector of
// compute_stress_like_vector_voigt (:280-285)
__device__ inline void ce_times(double lam, double G, double s, const double* e, double* v)
{
    double const tr = e[0] + e[4] + e[8];
    v[0] = s * (lam * tr + 2.0 * G * e[0]);
    v[1] = s * (lam * tr + 2.0 * G * e[4]);
    v[2] = s * (lam * tr + 2.0 * G * e[8]);
    v[3] = s * G * 2.0 * e[5];
    v[4] = s * G * 2.0 * e[2];
    v[5] = s * G * 2.0 * e[1];
}
__device__ inline double ddot9(const double* a, const double* b)
{
    double s = 0.0;
    for (int i = 0; i < 9; ++i) s += a[i] * b[i];
    return s;
}
__device__ inline double mandel_factor(int i, int j)
{
    if (i >= 3 && j >= 3) return 2.0;
    if ((i < 3) != (j < 3)) return 1.4142135623730951; // sqrt(2)
    return 1.0;
}
// the isotropic elastic moduli (isotropic_linear_elasticity.cpp:55-67) entry (i, j)
__device__ inline double ce_entry(double lam, double G, int i, int j)
{
    if (i < 3 && j < 3) return lam + (i == j ? 2.0 * G : 0.0);
    return i == j ? G : 0.0;
}

// residual and Jacobian (16 x 16 row-major, fully written) of the return map at y
__device__ void system(double G, double lam, double sy, double E, const damage_constants& dc, double dt,
                       const double* eps_e_t, const double* beta_t, double dmg_t, const double* y, double* res,
                       double* J)
{
    double const kp = pow(dc.sp, -dc.np), kd = pow(dc.sd, -dc.nd);
    double const s32 = sqrt(1.5);
    double const dlp = y[0], dld = y[1];
    double sig[9], bet[9];
    kin_from(y + 2, sig);
    kin_from(y + 8, bet);
    double const D = y[14], omD = 1.0 - D, Y = y[15];
    double dev[9];
    for (int i = 0; i < 9; ++i) dev[i] = sig[i] - bet[i];
    double const m = (dev[0] + dev[4] + dev[8]) / 3.0;
    dev[0] -= m, dev[4] -= m, dev[8] -= m;
    double const nd = sqrt(ddot9(dev, dev));
    double const vm = s32 * nd;
    double N[9], ee[9], dNdD[9], dfdb[9];
    for (int i = 0; i < 9; ++i) N[i] = s32 * dev[i] / (nd * omD);
    for (int i = 0; i < 9; ++i) ee[i] = eps_e_t[i] - dlp * N[i];
    double const fp = fmax(0.0, vm / omD + 0.5 * dc.gamma / dc.C * ddot9(bet, bet) - sy);
    double const fd = fmax(0.0, Y - sy * sy / (2.0 * E));
    for (int i = 0; i < 9; ++i) dfdb[i] = dc.gamma / dc.C * bet[i] - N[i];
    for (int i = 0; i < 9; ++i) dNdD[i] = N[i] / omD;
    double const dfdD = vm / (omD * omD);
    // dN/dsigma in kinetic Voigt: sqrt(3/2) / (1 - D) (P / |dev| - dev (x) dev / |dev|^3)
    double dv[6];
    kin(dev, dv);
    double const a1 = s32 / omD / nd, a2 = s32 / omD / (nd * nd * nd);
    auto dNds = [&](int i, int j) {
        double p = 0.0;
        if (i < 3 && j < 3)
            p = i == j ? 2.0 / 3.0 : -1.0 / 3.0;
        else if (i == j)
            p = 1.0;
        return a1 * p - a2 * dv[i] * dv[j];
    };
    for (int k = 0; k < 256; ++k) J[k] = 0.0;
    J[0 * 16 + 0] = J[1 * 16 + 1] = J[14 * 16 + 14] = J[15 * 16 + 15] = 1.0;
    double Cee[6], Ceem[9];
    ce_times(lam, G, 1.0, ee, Cee);
    kin_from(Cee, Ceem);
    res[0] = dlp - kp * pow(fp, dc.np) * dt;
    res[1] = dld - kd * pow(fd, dc.nd) * dt;
    {
        double const tr = ee[0] + ee[4] + ee[8];
        double t[9], v[6];
        for (int i = 0; i < 9; ++i) t[i] = sig[i] - omD * ((i % 4 == 0 ? lam * tr : 0.0) + 2.0 * G * ee[i]);
        kin(t, v);
        for (int i = 0; i < 6; ++i) res[2 + i] = v[i];
        for (int i = 0; i < 9; ++i) t[i] = bet[i] - beta_t[i] - dlp * (dc.C * N[i] - dc.gamma * bet[i]);
        kin(t, v);
        for (int i = 0; i < 6; ++i) res[8 + i] = v[i];
    }
    res[14] = D - dmg_t - dld;
    res[15] = Y - 0.5 * ddot9(ee, Ceem);
    double const cfp = -dc.np * kp * pow(fp, dc.np - 1.0) * dt;
    {
        double v[6], w[6];
        kin(N, v);
        kin(dfdb, w);
        for (int j = 0; j < 6; ++j)
        {
            J[0 * 16 + 2 + j] = cfp * v[j];
            J[0 * 16 + 8 + j] = cfp * w[j];
        }
    }
    J[0 * 16 + 14] = cfp * dfdD;
    J[1 * 16 + 15] = -dc.nd * kd * pow(fd, dc.nd - 1.0) * dt;
    {
        double v[6];
        ce_times(lam, G, omD, N, v);
        for (int i = 0; i < 6; ++i) J[(2 + i) * 16 + 0] = v[i];
    }
    // mandel((1 - D) C_e) mandel(dlp dN/dsigma); the back-stress block is its negative
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j)
        {
            double p = 0.0;
            for (int k = 0; k < 6; ++k)
                p += omD * ce_entry(lam, G, i, k) * mandel_factor(i, k) * dlp * dNds(k, j) * mandel_factor(k, j);
            J[(2 + i) * 16 + 2 + j] = (i == j ? 1.0 : 0.0) + p;
            J[(2 + i) * 16 + 8 + j] = -p;
        }
    {
        double t[9], v[6];
        for (int i = 0; i < 9; ++i) t[i] = omD * dlp * dNdD[i] + ee[i];
        ce_times(lam, G, 1.0, t, v);
        for (int i = 0; i < 6; ++i) J[(2 + i) * 16 + 14] = v[i];
        for (int i = 0; i < 9; ++i) t[i] = dc.gamma * bet[i] - dc.C * N[i];
        kin(t, v);
        for (int i = 0; i < 6; ++i) J[(8 + i) * 16 + 0] = v[i];
        for (int i = 0; i < 9; ++i) t[i] = -dc.C * dlp * dNdD[i];
        kin(t, v);
        for (int i = 0; i < 6; ++i) J[(8 + i) * 16 + 14] = v[i];
    }
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j)
        {
            double const d = dNds(i, j);
            J[(8 + i) * 16 + 2 + j] = -dc.C * dlp * d;
            J[(8 + i) * 16 + 8 + j] = (i == j ? 1.0 + dc.gamma * dlp : 0.0) + dc.C * dlp * d;
        }
    J[14 * 16 + 1] = -1.0;
    J[15 * 16 + 0] = ddot9(N, Ceem);
    for (int j = 0; j < 6; ++j)
    {
        double t = 0.0;
        for (int k = 0; k < 6; ++k) t += dNds(j, k) * Cee[k];
        J[15 * 16 + 2 + j] = dlp * t;
        J[15 * 16 + 8 + j] = -dlp * t;
    }
    J[15 * 16 + 14] = dlp * ddot9(dNdD, Ceem);
}

// In-place LU with complete pivoting of the 16 x 16 row-major A: rowp/colp receive the
// exchanges; multipliers below the diagonal.  false when a pivot is exactly zero.
__device__ bool lu_full(double* A, int* rowp, int* colp)
{
    for (int i = 0; i < 16; ++i) colp[i] = i;
    for (int k = 0; k < 16; ++k)
    {
        int pr = k, pc = k;
        double big = -1.0;
        for (int i = k; i < 16; ++i)
            for (int j = k; j < 16; ++j)
            {
                double const v = fabs(A[i * 16 + j]);
                if (v > big)
                {
                    big = v;
                    pr = i;
                    pc = j;
                }
            }
        if (big == 0.0) return false;
        rowp[k] = pr;
        // the trailing columns only: each stored multiplier stays with the position it was
        // computed at, so lu_solve can apply exchanges and eliminations step by step
        if (pr != k)
            for (int j = k; j < 16; ++j)
            {
                double const t = A[k * 16 + j];
                A[k * 16 + j] = A[pr * 16 + j];
                A[pr * 16 + j] = t;
            }
        if (pc != k)
        {
            for (int i = 0; i < 16; ++i)
            {
                double const t = A[i * 16 + k];
                A[i * 16 + k] = A[i * 16 + pc];
                A[i * 16 + pc] = t;
            }
            int const t = colp[k];
            colp[k] = colp[pc];
            colp[pc] = t;
        }
        for (int i = k + 1; i < 16; ++i)
        {
            double const f = A[i * 16 + k] / A[k * 16 + k];
            A[i * 16 + k] = f;
            if (f == 0.0) continue;
            for (int j = k + 1; j < 16; ++j) A[i * 16 + j] -= f * A[k * 16 + j];
        }
    }
    return true;
}

// x = A^-1 b from lu_full's factors (b is not modified)
__device__ void lu_solve(const double* A, const int* rowp, const int* colp, const double* b, double* x)
{
    double r[16];
    for (int i = 0; i < 16; ++i) r[i] = b[i];
    for (int k = 0; k < 16; ++k)
    {
        if (rowp[k] != k)
        {
            double const t = r[k];
            r[k] = r[rowp[k]];
            r[rowp[k]] = t;
        }
        for (int i = k + 1; i < 16; ++i)
        {
            double const f = A[i * 16 + k];
            if (f != 0.0) r[i] -= f * r[k];
        }
    }
    double yv[16];
    for (int k = 15; k >= 0; --k)
    {
        double s = r[k];
        for (int j = k + 1; j < 16; ++j) s -= A[k * 16 + j] * yv[j];
        yv[k] = s / A[k * 16 + k];
    }
    for (int k = 0; k < 16; ++k) x[colp[k]] = yv[k];
}
} // namespace damage_detail

// One thread per quadrature point (qp = 8 e + l); the per-element parameters are (K, G, yield
// stress, unused).
__global__ __launch_bounds__(kBlock) void damage_update_kernel(int32_t E, const int32_t* __restrict__ conn,
                                                               const double* __restrict__ X,
                                                               const double* __restrict__ u, element_params prm,
                                                               qp_state st, damage_state ds, damage_constants dc,
                                                               double dt, uint32_t* __restrict__ status)
{
    using namespace damage_detail;
    int64_t const qp = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (qp >= 8ll * E) return;
    int64_t const e = qp >> 3;
    int const l = int(qp & 7);
    int64_t const Q = st.Q;
    double H[9], F[9];
    qp_kinematics(conn, X, u, e, l, H, F);
    double const Jdet = det3(F);
    st.detF[qp] = Jdet;
    if (signbit(Jdet)) atomicOr(status, STATUS_NEGATIVE_JACOBIAN);

    double const bulk = prm.p[4 * e], G = prm.p[4 * e + 1], sy = prm.p[4 * e + 2];
    double const lam = bulk - 2.0 * G / 3.0;
    double const Emod = 9.0 * bulk * G / (3.0 * bulk + G); // isotropic_elastic_property.hpp:16
    double ep[9], beta[9], ev[6];
    for (int k = 0; k < 6; ++k) ev[k] = st.eps_p[k * Q + qp];
    kin_from(ev, ep);
    for (int k = 0; k < 6; ++k) ev[k] = ds.beta[k * Q + qp];
    kin_from(ev, beta);
    double dmg = ds.dmg[qp];
    double eb = st.ebar[qp];
    double ee[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) ee[i * 3 + j] = 0.5 * (H[i * 3 + j] + H[j * 3 + i]) - ep[i * 3 + j];
    // elastic predictor with damage (stress_routines.hpp:38-44), relative stress, energy
    double sig[9];
    {
        double const tr = ee[0] + ee[4] + ee[8];
        for (int i = 0; i < 9; ++i) sig[i] = (1.0 - dmg) * ((i % 4 == 0 ? lam * tr : 0.0) + 2.0 * G * ee[i]);
    }
    auto von_mises_rel = [&](const double* s, const double* b) {
        double d[9];
        for (int i = 0; i < 9; ++i) d[i] = s[i] - b[i];
        double const m = (d[0] + d[4] + d[8]) / 3.0;
        d[0] -= m, d[4] -= m, d[8] -= m;
        return sqrt(1.5) * sqrt(ddot9(d, d));
    };
    double vm = von_mises_rel(sig, beta);
    double energy;
    {
        double v[6], m[9];
        ce_times(lam, G, 1.0, ee, v);
        kin_from(v, m);
        energy = 0.5 * ddot9(ee, m);
    }
    double const fp0 = vm / (1.0 - dmg) + 0.5 * dc.gamma / dc.C * ddot9(beta, beta) - sy;
    double const fd0 = energy - sy * sy / (2.0 * Emod);
    double* coef = st.coef;
    bool ok = true;
    if (fp0 <= 0.0 && fd0 <= 0.0)
    {
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) coef[int64_t(i * 6 + j) * Q + qp] = ce_entry(lam, G, i, j);
    }
    else
    {
        // perform_radial_return
        double y[16], yprev[16], res[16], Jm[256], dx[16];
        int rowp[16], colp[16];
        y[0] = y[1] = 0.0;
        kin(sig, y + 2);
        kin(beta, y + 8);
        y[14] = dmg;
        y[15] = energy;
        double rn = 2.0; // the reference starts from a residual of ones (norm 4)
        int it = 0;
        while (rn > 1.0e-6 && it < 50)
        {
            system(G, lam, sy, Emod, dc, dt, ee, beta, dmg, y, res, Jm);
            double s = 0.0;
            for (int i = 0; i < 16; ++i) s += res[i] * res[i];
            rn = sqrt(s);
            if (!lu_full(Jm, rowp, colp))
            {
                ok = false;
                break;
            }
            lu_solve(Jm, rowp, colp, res, dx);
            for (int i = 0; i < 16; ++i)
            {
                yprev[i] = y[i];
                y[i] -= dx[i];
            }
            ++it;
            // the reference tests the residual of the iterate it just solved from (its loop
            // condition reads the residual the last system computed)
        }
        if (!ok || it == 50 || y[14] > 1.0) atomicOr(status, STATUS_RETURN_MAP);
        double const pinc = y[0] > 0.0 ? y[0] : 0.0;
        kin_from(y + 2, sig);
        kin_from(y + 8, beta);
        dmg = y[14] > dmg ? y[14] : dmg;
        energy = y[15];
        // tangent (1 - D) mandel(C_e) mandel(inverse(J)[2:8, 2:8]) with the Jacobian of the
        // last iteration: rebuild it at the iterate it was formed at and factor once
        double inv22[36];
        if (it > 0 && ok)
        {
            double const dmg0 = ds.dmg[qp];
            double b0[9], v0[6];
            for (int k = 0; k < 6; ++k) v0[k] = ds.beta[k * Q + qp];
            kin_from(v0, b0);
            system(G, lam, sy, Emod, dc, dt, ee, b0, dmg0, yprev, res, Jm);
            if (!lu_full(Jm, rowp, colp))
                atomicOr(status, STATUS_RETURN_MAP);
            else
                for (int c = 0; c < 6; ++c)
                {
                    double ecol[16], col[16];
                    for (int i = 0; i < 16; ++i) ecol[i] = i == 2 + c ? 1.0 : 0.0;
                    lu_solve(Jm, rowp, colp, ecol, col);
                    for (int r = 0; r < 6; ++r) inv22[r * 6 + c] = col[2 + r];
                }
        }
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j)
            {
                double p = 0.0;
                for (int k = 0; k < 6; ++k)
                    p += ce_entry(lam, G, i, k) * mandel_factor(i, k) * inv22[k * 6 + j] * mandel_factor(k, j);
                coef[int64_t(i * 6 + j) * Q + qp] = (1.0 - dmg) * p;
            }
        // update_internal_variables after the return map (:104-119)
        vm = von_mises_rel(sig, beta);
        double d[9];
        for (int i = 0; i < 9; ++i) d[i] = sig[i] - beta[i];
        double const m = (d[0] + d[4] + d[8]) / 3.0;
        d[0] -= m, d[4] -= m, d[8] -= m;
        double const nd = sqrt(ddot9(d, d));
        double const inc = pinc * dt;
        for (int i = 0; i < 9; ++i) ep[i] += inc * sqrt(1.5) * d[i] / (nd * (1.0 - dmg));
        eb += inc / (1.0 - dmg);
        kin(ep, ev);
        for (int k = 0; k < 6; ++k) st.eps_p[k * Q + qp] = ev[k];
        st.ebar[qp] = eb;
        kin(beta, ev);
        for (int k = 0; k < 6; ++k) ds.beta[k * Q + qp] = ev[k];
        ds.dmg[qp] = dmg;
    }
    ds.energy[qp] = energy;
    st.vm[qp] = vm;
    double sv[6];
    kin(sig, sv);
    bool finite = true;
    for (int k = 0; k < 6; ++k)
    {
        st.sigma[k * Q + qp] = sv[k];
        finite = finite && isfinite(sv[k]);
    }
    if (!finite) atomicOr(status, STATUS_NAN);
}
} // namespace neon_amd
