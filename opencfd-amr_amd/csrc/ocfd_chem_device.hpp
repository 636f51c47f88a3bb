// Li-Dryer H2/air chemistry on the device: NASA7 thermodynamics, Newton
// temperature inversion, 19-reaction production rates with their mass-action
// Jacobian, the 9x9 pivoted Gauss solve and the linearised-implicit step of
// one cell.  FP64 throughout, in the reference's operation order and without
// FMA contraction, so results track the reference arithmetic to libm rounding.
//
// Species order: H2 O2 H2O H O OH HO2 H2O2 N2 (Source/Constants.cpp:13-22).
#pragma once

#include <hip/hip_runtime.h>

namespace ocfd {
namespace chem {

constexpr int NS = 9;
constexpr int NR = 19;
constexpr double Ru = 8.31446261815324;
constexpr double Patm = 101325.0;

// kg/mol (Source/Constants.cpp:13-22)
__device__ constexpr double mw[NS] = {2.016 * 1e-3, 31.998 * 1e-3, 18.015 * 1e-3, 1.008 * 1e-3, 15.999 * 1e-3,
                                      17.007 * 1e-3, 33.006 * 1e-3, 34.014 * 1e-3, 28.014 * 1e-3};

// NASA7 low (T < 1000 K) / high branches (Source/Constants.cpp:23-73)
__device__ constexpr double lo[NS][7] = {
    {3.29812, 8.24944e-4, -81.43015e-8, -9.47543e-11, 413.48718e-15, -0.01013e+5, -3.29409},
    {3.21294, 11.27486e-4, -57.56150e-8, 131.38770e-11, -876.8554e-15, -0.01005e+5, 6.03474},
    {3.38684, 34.74982e-4, -635.4696e-8, 696.8581e-11, -2506.588e-15, -0.30208e+5, 2.59023},
    {2.5, 0.0, 0.0, 0.0, 0.0, 0.25472e+5, -0.46012},
    {2.94643, -16.3817e-4, 242.10320e-8, -160.284e-11, 389.06961e-15, 0.29148e+5, 2.96399},
    {3.63727, 1.85091e-4, -167.6165e-8, 238.72031e-11, -843.1442e-15, 0.03607e+5, 1.35886},
    {2.97996, 49.96697e-4, -379.0997e-8, 235.41919e-11, -808.9024e-15, 0.00176e+5, 9.22272},
    {3.38875, 65.69226e-4, -14.85013e-8, -462.58060e-11, 2471.515e-15, -0.17663e+5, 6.78536},
    {3.29868, 14.08240e-4, -396.3222e-8, 564.15149e-11, -2444.855e-15, -0.01021e+5, 3.95037}};
__device__ constexpr double hi[NS][7] = {
    {2.99142, 7.00064e-4, -5.63383e-8, -0.92316e-11, 1.58275e-15, -0.00835e+5, -1.35511},
    {3.69758, 6.13520e-4, -12.58842e-8, 1.77528e-11, -1.13644e-15, -0.01234e+5, 3.18917},
    {2.67215, 30.56293e-4, -87.30260e-8, 12.00996e-11, -6.39162e-15, -0.29899e+5, 6.86282},
    {2.5, 0.0, 0.0, 0.0, 0.0, 0.25472e+5, -0.46012},
    {2.54206, -0.27551e-4, -0.31028e-8, 0.45511e-11, -0.43681e-15, 0.29231e+5, 4.92031},
    {2.88273, 10.13974e-4, -22.76877e-8, 2.17468e-11, -0.51263e-15, 0.03887e+5, 5.59571},
    {4.07219, 21.31296e-4, -53.08145e-8, 6.11227e-11, -2.84116e-15, -0.00158e+5, 3.47603},
    {4.57317, 43.36136e-4, -147.4689e-8, 23.48904e-11, -14.31654e-15, -0.18007e+5, 0.50114},
    {2.92664, 14.87977e-4, -56.84761e-8, 10.09704e-11, -6.75335e-15, -0.00923e+5, 5.98053}};

#pragma clang fp contract(off)

// sensible energy density (J/m^3) and cv density (J/(m^3 K)) at T, CKUBMS /
// CKCVBS of Source/CHEM/LiDryer.H:697-816
__device__ __forceinline__ void e_cv(double T, const double* rhoi, double& e, double& cv, bool want_cv)
{
    const double tc1 = T, tc2 = T * T, tc3 = T * T * T, tc4 = T * T * T * T;
    const double invT = 1.0 / T;
    const bool low = T < 1000;
    double re = 0, rc = 0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const double* a = low ? lo[s] : hi[s];
        double u = a[0] - 1.0 + a[1] / 2.0 * tc1 + a[2] / 3.0 * tc2 + a[3] / 4.0 * tc3 + a[4] / 5.0 * tc4;
        if (!low) u = u + (hi[s][5] - lo[s][5]) * invT;
        re += rhoi[s] * u / mw[s];
        if (want_cv) {
            double c = a[0] + a[1] * tc1 + a[2] * tc2 + a[3] * tc3 + a[4] * tc4;
            rc += (c - 1.0) * rhoi[s] / mw[s];
        }
    }
    e = re * (Ru * T);
    cv = rc * Ru;
}

__device__ __forceinline__ double ckubms(double T, const double* rhoi)
{
    double e, cv;
    e_cv(T, rhoi, e, cv, false);
    return e;
}

__device__ __forceinline__ double ckcvbs(double T, const double* rhoi)
{
    double e, cv;
    e_cv(T, rhoi, e, cv, true);
    return cv;
}

// GET_T_GIVEN_EY, Source/CHEM/LiDryer.H:836-888
__device__ __forceinline__ double T_given_ey(double ein, const double* rhoi)
{
    const double tmin = 250, tmax = 3500, tol = 1e-7;
    const double emin = ckubms(tmin, rhoi), emax = ckubms(tmax, rhoi);
    if (ein < emin) return tmin - (emin - ein) / ckcvbs(tmin, rhoi);
    if (ein > emax) return tmax - (emax - ein) / ckcvbs(tmax, rhoi);
    double As = 0;
#pragma unroll
    for (int n = 0; n < NS; ++n) As += (lo[n][0] - 1.0) * Ru / mw[n] * rhoi[n];
    double t1 = ein / As;
    if (t1 < tmin || t1 > tmax) t1 = tmin + (tmax - tmin) / (emax - emin) * (ein - emin);
    for (int i = 0; i < 200; ++i) {
        double e1, cv;
        e_cv(t1, rhoi, e1, cv, true);
        double dt = (ein - e1) / cv;
        if (dt > 100.0) dt = 100.0;
        else if (dt < -100.0) dt = -100.0;
        else if (fabs(dt) < tol) break;
        else if (t1 + dt == t1) break;
        t1 += dt;
    }
    return t1;
}

// forward rate constants (CGS), Source/CHEM/LiDryer.H:171-191
__device__ __forceinline__ void rate_constants(double lnT, double invT, double* k)
{
    k[0] = 3.55e+15 * exp(-0.41 * lnT - 8353.444256079512 * invT);
    k[1] = 50800.0 * exp(2.67 * lnT - 3165.2508657072367 * invT);
    k[2] = 2.16e+08 * exp(1.51 * lnT - 1726.0429999007667 * invT);
    k[3] = 2.97e+06 * exp(2.02 * lnT - 6743.141748883461 * invT);
    k[4] = 4.58e+19 * exp(-1.4 * lnT - 52526.05490660117 * invT);
    k[5] = 6.16e+15 * exp(-0.5 * lnT);
    k[6] = 4.71e+18 * exp(-1.0 * lnT);
    k[7] = 3.8e+22 * exp(-2.0 * lnT);
    k[8] = 1.48e+12 * exp(0.6 * lnT);
    k[9] = 1.66e+13 * exp(-412.64001746898793 * invT);
    k[10] = 7.08e+13 * exp(-150.96586004962973 * invT);
    k[11] = 3.25e+13;
    k[12] = 2.89e+13 * exp(+251.60976674938289 * invT);
    k[13] = 4.2e+14 * exp(-6028.570011315214 * invT);
    k[14] = 1.3e+11 * exp(+820.2478396029883 * invT);
    k[15] = 2.95e+14 * exp(-24355.825421340265 * invT);
    k[16] = 2.41e+13 * exp(-1997.7815479901 * invT);
    k[17] = 4.82e+13 * exp(-4000.595291315188 * invT);
    k[18] = 9.55e+06 * exp(2.0 * lnT - 1997.7815479901 * invT);
    k[19] = 1.0e+12;
    k[20] = 5.8e+14 * exp(-4810.778740248201 * invT);
}

// Troe fall-off factor k/k_inf, Source/CHEM/LiDryer.H:337-357,431-449
__device__ __forceinline__ double troe(double redP, double Fc)
{
    const double logPred = log10(redP);
    const double logFcent = log10(Fc);
    const double troe_c = -0.4 - 0.67 * logFcent;
    const double troe_n = 0.75 - 1.27 * logFcent;
    const double t = (troe_c + logPred) / (troe_n - 0.14 * (troe_c + logPred));
    const double F = pow(10.0, logFcent / (1.0 + t * t));
    return (redP / (1.0 + redP)) * F;
}

// Reaction table: forward/reverse species (-1 = absent), stoichiometric
// coefficients of the participating species, rate index, duplicate, third body
struct Rx {
    int f0, f1, r0, r1, kf, dup, tb;
    int np;                 // participating species
    int sp[4];              // species index
    int nuf[4], nur[4];     // stoichiometric coefficients of those species
};

__device__ constexpr Rx rx[NR] = {
    {3, 1, 4, 5, 0, -1, 0, 4, {1, 3, 4, 5}, {1, 1, 0, 0}, {0, 0, 1, 1}},
    {4, 0, 3, 5, 1, -1, 0, 4, {0, 3, 4, 5}, {1, 0, 1, 0}, {0, 1, 0, 1}},
    {0, 5, 2, 3, 2, -1, 0, 4, {0, 2, 3, 5}, {1, 0, 0, 1}, {0, 1, 1, 0}},
    {4, 2, 5, 5, 3, -1, 0, 3, {2, 4, 5, -1}, {1, 1, 0, 0}, {0, 0, 2, 0}},
    {0, -1, 3, 3, 4, -1, 1, 2, {0, 3, -1, -1}, {1, 0, 0, 0}, {0, 2, 0, 0}},
    {4, 4, 1, -1, 5, -1, 1, 2, {1, 4, -1, -1}, {0, 2, 0, 0}, {1, 0, 0, 0}},
    {4, 3, 5, -1, 6, -1, 1, 3, {3, 4, 5, -1}, {1, 1, 0, 0}, {0, 0, 1, 0}},
    {3, 5, 2, -1, 7, -1, 1, 3, {2, 3, 5, -1}, {0, 1, 1, 0}, {1, 0, 0, 0}},
    {3, 1, 6, -1, 8, -1, 2, 3, {1, 3, 6, -1}, {1, 1, 0, 0}, {0, 0, 1, 0}},
    {6, 3, 0, 1, 9, -1, 0, 4, {0, 1, 3, 6}, {0, 0, 1, 1}, {1, 1, 0, 0}},
    {6, 3, 5, 5, 10, -1, 0, 3, {3, 5, 6, -1}, {1, 0, 1, 0}, {0, 2, 0, 0}},
    {6, 4, 1, 5, 11, -1, 0, 4, {1, 4, 5, 6}, {0, 1, 0, 1}, {1, 0, 1, 0}},
    {6, 5, 2, 1, 12, -1, 0, 4, {1, 2, 5, 6}, {0, 0, 1, 1}, {1, 1, 0, 0}},
    {6, 6, 7, 1, 13, 14, 0, 3, {1, 6, 7, -1}, {0, 2, 0, 0}, {1, 0, 1, 0}},
    {7, -1, 5, 5, 15, -1, 3, 2, {5, 7, -1, -1}, {0, 1, 0, 0}, {2, 0, 0, 0}},
    {7, 3, 2, 5, 16, -1, 0, 4, {2, 3, 5, 7}, {0, 1, 0, 1}, {1, 0, 1, 0}},
    {7, 3, 6, 0, 17, -1, 0, 4, {0, 3, 6, 7}, {0, 1, 0, 1}, {1, 0, 1, 0}},
    {7, 4, 5, 6, 18, -1, 0, 4, {4, 5, 6, 7}, {1, 0, 0, 1}, {0, 1, 1, 0}},
    {7, 5, 6, 2, 19, 20, 0, 4, {2, 5, 6, 7}, {0, 1, 0, 1}, {1, 0, 1, 0}}};

// equilibrium-constant species groups: exp((g[a0]+g[a1]) - (g[b0]+g[b1])),
// times 1, 1/RsT or RsT (Source/CHEM/LiDryer.H:197-219)
struct KcDef { int a0, a1, b0, b1, fac; };
__device__ constexpr KcDef kcd[21] = {
    {3, 1, 4, 5, 0}, {4, 0, 3, 5, 0}, {0, 5, 2, 3, 0}, {4, 2, 5, 5, 0}, {0, -1, 3, 3, 1}, {4, 4, 1, -1, 2},
    {4, 3, 5, -1, 2}, {3, 5, 2, -1, 2}, {3, 1, 6, -1, 2}, {6, 3, 0, 1, 0}, {6, 3, 5, 5, 0}, {6, 4, 1, 5, 0},
    {6, 5, 2, 1, 0}, {6, 6, 7, 1, 0}, {6, 6, 7, 1, 0}, {7, -1, 5, 5, 1}, {7, 3, 2, 5, 0}, {7, 3, 6, 0, 0},
    {7, 4, 5, 6, 0}, {7, 5, 6, 2, 0}, {7, 5, 6, 2, 0}};

__device__ __forceinline__ double sc2(const double* sc, int a, int b) { return b < 0 ? sc[a] : sc[a] * sc[b]; }

// Production rates wdot (mol/cm^3/s) and the mass-action Jacobian A[l][n] =
// d wdot_l / d c_n at fixed T, Source/CHEM/LiDryer.H:156-529.  Species that do
// not take part in a reaction contribute exact zeros in the reference's dense
// loops; skipping them leaves every sum unchanged.
template <bool JAC>
__device__ __forceinline__ void production_rate_t(const double* sc, double T, double* wdot, double (*A)[NS])
{
    const double eps = 1e-20;
    const double lnT = log(T);
    const double invT = 1.0 / T;
    double kfs[21];
    rate_constants(lnT, invT, kfs);
    // g/(RT) (gibbs, Source/CHEM/LiDryer.H:67-154)
    double g[NS];
    {
        const double tc1 = T, tc2 = T * (T), tc3 = tc2 * tc1, tc4 = tc2 * tc2;
        const double ginv = 1 / T;
        const bool low = T < 1000;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const double* a = low ? lo[s] : hi[s];
            g[s] = a[0] * (1.0 - lnT) - a[1] / 2.0 * tc1 - a[2] / 6.0 * tc2 - a[3] / 12.0 * tc3 -
                   a[4] / 20.0 * tc4 + a[5] * ginv - a[6];
        }
    }
    const double RsT = Ru / Patm * 1e6 * T;
    double mixture = 0.0;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
        mixture += sc[n];
        wdot[n] = 0.0;
        if constexpr (JAC) {
#pragma unroll
            for (int l = 0; l < NS; ++l) A[n][l] = 0.0;
        }
    }
#pragma unroll
    for (int m = 0; m < NR; ++m) {
        const Rx& R = rx[m];
        const double phi_f = sc2(sc, R.f0, R.f1);
        const double phi_r = sc2(sc, R.r0, R.r1);
        double k_f = kfs[R.kf];
        if (R.tb == 1) {
            const double alpha = mixture + 1.5 * sc[0] + 11 * sc[2];
            k_f = alpha * k_f;
        } else if (R.tb == 2) {
            const double alpha = mixture + sc[0] + 10 * sc[2] + -0.22 * sc[1];
            const double redP = alpha / k_f * 6.37e+20 * exp(-1.72 * lnT - 261.6741574193582 * invT);
            k_f *= troe(redP, 0.8);
        } else if (R.tb == 3) {
            const double alpha = mixture + 1.5 * sc[0] + 11 * sc[2];
            const double redP = alpha / k_f * 1.20e+17 * exp(-22846.166820843966 * invT);
            k_f *= troe(redP, 0.5);
        }
        double Kc;
        {
            const KcDef& d = kcd[R.kf];
            const double sa = d.a1 < 0 ? g[d.a0] : g[d.a0] + g[d.a1];
            const double sb = d.b1 < 0 ? g[d.b0] : g[d.b0] + g[d.b1];
            const double e = exp((sa) - (sb));
            Kc = d.fac == 0 ? e : (d.fac == 1 ? 1.0 / RsT * e : RsT * e);
        }
        double qf = phi_f * k_f;
        double qr = phi_r * (k_f / Kc);
        if (R.dup >= 0) {
            const KcDef& d = kcd[R.dup];
            const double sa = d.a1 < 0 ? g[d.a0] : g[d.a0] + g[d.a1];
            const double sb = d.b1 < 0 ? g[d.b0] : g[d.b0] + g[d.b1];
            const double Kc2 = exp((sa) - (sb));
            const double k2 = kfs[R.dup];
            qf += phi_f * k2;
            qr += phi_r * (k2 / Kc2);
        }
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            if (a >= R.np) break;
            const int n = R.sp[a];
            wdot[n] += (qf - qr) * (double)(R.nur[a] - R.nuf[a]);
        }
        if constexpr (JAC) {
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                if (a >= R.np) break;
                const int n = R.sp[a];
                const double Awf = R.nuf[a] * qf / (sc[n] + eps);
                const double Awr = R.nur[a] * qr / (sc[n] + eps);
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    if (b >= R.np) break;
                    const int l = R.sp[b];
                    A[l][n] += (Awf - Awr) * (double)(R.nur[b] - R.nuf[b]);
                }
            }
        }
    }
}

__device__ __forceinline__ void production_rate(const double* sc, double T, double* wdot, double (*A)[NS])
{
    production_rate_t<true>(sc, T, wdot, A);
}

// Gaussian elimination with partial pivoting (first maximum), in place on the
// augmented system U[9][10]; Source/CHEM/LiDryer.H:890-942
__device__ __forceinline__ void gauss_solve(double (*U)[NS + 1], double* x)
{
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        int p = k;
        double maxPivot = 0.0;
#pragma unroll
        for (int i = k; i < NS; ++i) {
            const double v = fabs(U[i][k]);
            if (v > maxPivot) { maxPivot = v; p = i; }
        }
        if (p != k) {
#pragma unroll
            for (int i = k + 1; i < NS; ++i) {
                const bool s = (p == i);
#pragma unroll
                for (int j = k; j < NS + 1; ++j) {
                    const double a = U[i][j], b = U[k][j];
                    U[i][j] = s ? b : a;
                    U[k][j] = s ? a : b;
                }
            }
        }
#pragma unroll
        for (int i = k + 1; i < NS; ++i) {
            const double m = U[i][k] / U[k][k];
#pragma unroll
            for (int j = k; j < NS + 1; ++j) U[i][j] -= m * U[k][j];
        }
    }
#pragma unroll
    for (int k = NS - 1; k >= 0; --k) {
        double sum = U[k][NS];
#pragma unroll
        for (int j = k + 1; j < NS; ++j) sum -= U[k][j] * x[j];
        x[k] = sum / U[k][k];
    }
}

// Linearised-implicit chemistry step of one cell at rest-frame energy ei,
// Source/CHEM/ChemAdvance.cpp:48-81: rhoi updated (clipped at 0) in place,
// returns the change of the sensible energy density.
__device__ __forceinline__ double cell_step(double ei, double* rhoi, double dt)
{
    const double T = T_given_ey(ei, rhoi);
    double c[NS], wdot[NS], A[NS][NS];
#pragma unroll
    for (int n = 0; n < NS; ++n) c[n] = rhoi[n] / mw[n] * 1e-6;
    production_rate(c, T, wdot, A);
    double U[NS][NS + 1];
#pragma unroll
    for (int i = 0; i < NS; ++i) {
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            double v = (i == j ? 1.0 : 0.0);
            v -= A[i][j] * mw[i] / mw[j] * dt;
            U[i][j] = v;
        }
        U[i][NS] = wdot[i] * mw[i] * 1e6 * dt;
    }
    double drho[NS];
    gauss_solve(U, drho);
    double dei = 0;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
        rhoi[n] += drho[n];
        if (rhoi[n] < 0) rhoi[n] = 0.0;
        dei += -lo[n][5] * Ru / mw[n] * drho[n];
    }
    return dei;
}

// ---------------------------------------------------------------------------
// Rosenbrock (RODAS3) integration of one constant-volume reactor: no
// counterpart in the reference, whose chemistry takes fixed linearised-
// implicit steps.  The ODE d rho_k/dt = W_k wdot_k(c, T), T from the sensible
// energy E(rho) = E0 - sum_k hf_k (rho_k - rho_k0); RODAS3 (Sandu et al.,
// Atmos. Environ. 31 (1997) 3459-3472: 4 stages, order 3, embedded order 2,
// L-stable, stiffly accurate) in the transformed form
//   (I/(h gamma) - J) K_i = f(y + sum_j a_ij K_j) + sum_j (c_ij/h) K_j,
// J the full Jacobian: the complete fixed-T derivative plus (df/dT)(dT/drho).
// ---------------------------------------------------------------------------

// Troe factor P(Pr) = Pr/(1+Pr) F(Pr) (troe above) and its slope dP/dPr
__device__ __forceinline__ void troe_and_slope(double redP, double Fc, double& P, double& dP)
{
    const double logPred = log10(redP);
    const double logFcent = log10(Fc);
    const double troe_c = -0.4 - 0.67 * logFcent;
    const double troe_n = 0.75 - 1.27 * logFcent;
    const double uu = troe_c + logPred;
    const double den = troe_n - 0.14 * uu;
    const double t = uu / den;
    const double F = pow(10.0, logFcent / (1.0 + t * t));
    P = (redP / (1.0 + redP)) * F;
    const double q = 1.0 + t * t;
    dP = F / ((1.0 + redP) * (1.0 + redP)) + F / (1.0 + redP) * (-2.0 * t * logFcent * troe_n) / (q * q * den * den);
}

// A 9x9 matrix of one lane in LDS: element (i, j) at base[(9 i + j) * 64]
// (a block's 64 lanes interleaved, so a wave's accesses never conflict)
struct LaneMat {
    double* base;
    __device__ __forceinline__ double& operator()(int i, int j) const { return base[(i * NS + j) * 64]; }
};

// the complete d wdot/d c (CGS) at fixed T: mass action by the product rule
// (exact where a species is absent), third-body efficiencies and the Troe
// fall-off, with the rates of production_rate
__device__ __forceinline__ void jacobian_fixed_T(const double* sc, double T, LaneMat J)
{
    const double lnT = log(T);
    const double invT = 1.0 / T;
    double kfs[21];
    rate_constants(lnT, invT, kfs);
    double g[NS];
    {
        const double tc1 = T, tc2 = T * (T), tc3 = tc2 * tc1, tc4 = tc2 * tc2;
        const double ginv = 1 / T;
        const bool low = T < 1000;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const double* a = low ? lo[s] : hi[s];
            g[s] = a[0] * (1.0 - lnT) - a[1] / 2.0 * tc1 - a[2] / 6.0 * tc2 - a[3] / 12.0 * tc3 -
                   a[4] / 20.0 * tc4 + a[5] * ginv - a[6];
        }
    }
    const double RsT = Ru / Patm * 1e6 * T;
    double mixture = 0.0;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
        mixture += sc[n];
#pragma unroll
        for (int l = 0; l < NS; ++l) J(n, l) = 0.0;
    }
#pragma unroll
    for (int m = 0; m < NR; ++m) {
        const Rx& R = rx[m];
        const double phi_f = sc2(sc, R.f0, R.f1), phi_r = sc2(sc, R.r0, R.r1);
        double dpf[NS], dpr[NS];
#pragma unroll
        for (int n = 0; n < NS; ++n) { dpf[n] = 0.0; dpr[n] = 0.0; }
        if (R.f1 < 0) dpf[R.f0] += 1.0;
        else { dpf[R.f0] += sc[R.f1]; dpf[R.f1] += sc[R.f0]; }
        if (R.r1 < 0) dpr[R.r0] += 1.0;
        else { dpr[R.r0] += sc[R.r1]; dpr[R.r1] += sc[R.r0]; }
        const double k0 = kfs[R.kf];
        double keff = k0, dk[NS];
#pragma unroll
        for (int n = 0; n < NS; ++n) dk[n] = 0.0;
        if (R.tb != 0) {
            double eff[NS];
#pragma unroll
            for (int n = 0; n < NS; ++n) eff[n] = 1.0;
            double alpha;
            if (R.tb == 2) {
                eff[0] += 1.0; eff[2] += 10.0; eff[1] += -0.22;
                alpha = mixture + sc[0] + 10 * sc[2] + -0.22 * sc[1];
            } else {
                eff[0] += 1.5; eff[2] += 11.0;
                alpha = mixture + 1.5 * sc[0] + 11 * sc[2];
            }
            if (R.tb == 1) {
                keff = alpha * k0;
#pragma unroll
                for (int n = 0; n < NS; ++n) dk[n] = eff[n] * k0;
            } else {
                const double kl = R.tb == 2 ? 6.37e+20 * exp(-1.72 * lnT - 261.6741574193582 * invT)
                                            : 1.20e+17 * exp(-22846.166820843966 * invT);
                const double redP = alpha / k0 * kl;
                double P, dP;
                troe_and_slope(redP, R.tb == 2 ? 0.8 : 0.5, P, dP);
                keff = k0 * P;
#pragma unroll
                for (int n = 0; n < NS; ++n) dk[n] = dP * kl * eff[n];
            }
        }
        double Kc;
        {
            const KcDef& d = kcd[R.kf];
            const double sa = d.a1 < 0 ? g[d.a0] : g[d.a0] + g[d.a1];
            const double sb = d.b1 < 0 ? g[d.b0] : g[d.b0] + g[d.b1];
            const double e = exp((sa) - (sb));
            Kc = d.fac == 0 ? e : (d.fac == 1 ? 1.0 / RsT * e : RsT * e);
        }
        const double diff = phi_f - phi_r / Kc;
        double dq[NS];
#pragma unroll
        for (int n = 0; n < NS; ++n) dq[n] = keff * (dpf[n] - dpr[n] / Kc) + dk[n] * diff;
        if (R.dup >= 0) {
            const KcDef& d = kcd[R.dup];
            const double sa = d.a1 < 0 ? g[d.a0] : g[d.a0] + g[d.a1];
            const double sb = d.b1 < 0 ? g[d.b0] : g[d.b0] + g[d.b1];
            const double Kc2 = exp((sa) - (sb));
            const double k2 = kfs[R.dup];
#pragma unroll
            for (int n = 0; n < NS; ++n) dq[n] += k2 * (dpf[n] - dpr[n] / Kc2);
        }
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            if (b >= R.np) break;
            const int l = R.sp[b];
            const int nu = R.nur[b] - R.nuf[b];
            if (nu == 0) continue;
#pragma unroll
            for (int n = 0; n < NS; ++n) J(l, n) += nu * dq[n];
        }
    }
}

// Newton inversion of the sensible energy from a temperature guess (the
// iteration of T_given_ey); T_given_ey itself when the guess or an iterate
// leaves [250, 3500] K
__device__ __forceinline__ double T_newton(double ein, const double* rhoi, double Tg)
{
    const double tol = 1e-7;
    if (!(Tg >= 250.0 && Tg <= 3500.0)) return T_given_ey(ein, rhoi);
    double t1 = Tg;
    for (int i = 0; i < 50; ++i) {
        double e1, cv;
        e_cv(t1, rhoi, e1, cv, true);
        double dt = (ein - e1) / cv;
        if (dt > 100.0) dt = 100.0;
        else if (dt < -100.0) dt = -100.0;
        else if (fabs(dt) < tol) return t1;
        else if (t1 + dt == t1) return t1;
        t1 += dt;
        if (t1 < 250.0 || t1 > 3500.0) return T_given_ey(ein, rhoi);
    }
    return t1;
}

// f = d rho/dt (kg/m^3/s) at T
__device__ __forceinline__ void reactor_f(const double* y, double T, double* f)
{
    double c[NS], wdot[NS];
#pragma unroll
    for (int n = 0; n < NS; ++n) c[n] = y[n] / mw[n] * 1e-6;
    production_rate_t<false>(c, T, wdot, nullptr);
#pragma unroll
    for (int i = 0; i < NS; ++i) f[i] = wdot[i] * mw[i] * 1e6;
}

// LU with partial pivoting (first maximum) in place on the lane's LDS matrix;
// a row exchange at step k moves columns k.. only
__device__ __forceinline__ void lu_factor(LaneMat M, int* piv)
{
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        int p = k;
        double mx = 0.0;
#pragma unroll
        for (int i = k; i < NS; ++i) {
            const double v = fabs(M(i, k));
            if (v > mx) { mx = v; p = i; }
        }
        piv[k] = p;
        if (p != k) {
#pragma unroll
            for (int j = k; j < NS; ++j) {
                const double a = M(p, j);
                M(p, j) = M(k, j);
                M(k, j) = a;
            }
        }
        const double pv = M(k, k);
#pragma unroll
        for (int i = k + 1; i < NS; ++i) {
            const double m = M(i, k) / pv;
            M(i, k) = m;
#pragma unroll
            for (int j = k + 1; j < NS; ++j) M(i, j) -= m * M(k, j);
        }
    }
}

// b <- the solution of (the factored M) x = b; b in registers (exchanges by
// selects: no dynamic register indexing)
__device__ __forceinline__ void lu_solve(LaneMat M, const int* piv, double* b)
{
#pragma unroll
    for (int k = 0; k < NS; ++k) {
#pragma unroll
        for (int i = k + 1; i < NS; ++i) {
            const bool sw = (piv[k] == i);
            const double a = b[i], c = b[k];
            b[i] = sw ? c : a;
            b[k] = sw ? a : c;
        }
#pragma unroll
        for (int i = k + 1; i < NS; ++i) b[i] -= M(i, k) * b[k];
    }
#pragma unroll
    for (int k = NS - 1; k >= 0; --k) {
        double sum = b[k];
#pragma unroll
        for (int j = k + 1; j < NS; ++j) sum -= M(k, j) * b[j];
        b[k] = sum / M(k, k);
    }
}

// energy change of a change dy of the partial densities (formation enthalpies)
__device__ __forceinline__ double energy_of(const double* dy)
{
    double de = 0.0;
#pragma unroll
    for (int n = 0; n < NS; ++n) de += -lo[n][5] * Ru / mw[n] * dy[n];
    return de;
}

// One RODAS3 step of h from (y, E) with temperature guess Tg: the new state in
// ynew/Enew, the last stage's temperature in Tlast; returns the error
// max_k |K4_k| / (atol rho + rtol max(|y_k|, |ynew_k|)).  The four rate
// evaluations (y at T0, y at T0 + dT, the stages at y + 2 K1 and y + 2 K1 +
// K3) run as one loop, so the kernel holds one copy of the rate code.
__device__ __forceinline__ double rodas3_step(const double* y, double E, double Tg, double h, double rho, double rtol,
                                              double atol, double* ynew, double& Enew, double& Tlast, LaneMat M)
{
    double f0[NS], K1[NS], K2[NS], K3[NS], K4[NS], dy[NS];
    int piv[NS];
    const double hinv = 1.0 / h;
    double T0 = 0.0, Tprev = Tg, T4 = 0.0;
#pragma unroll 1
    for (int it = 0; it < 4; ++it) {
        double ys[NS], fs[NS], Ecur = E;
#pragma unroll
        for (int n = 0; n < NS; ++n) {
            dy[n] = it == 2 ? 2.0 * K1[n] : (it == 3 ? 2.0 * K1[n] + K3[n] : 0.0);
            ys[n] = it >= 2 ? y[n] + dy[n] : y[n];
        }
        if (it >= 2) Ecur = E + energy_of(dy);
        double Tc;
        if (it == 1) Tc = T0 + 1e-7 * T0;
        else Tc = T_newton(Ecur, ys, Tprev);
        reactor_f(ys, Tc, fs);
        if (it == 0) {
            T0 = Tc;
#pragma unroll
            for (int n = 0; n < NS; ++n) f0[n] = fs[n];
        } else if (it == 1) {
            // M = I/(h gamma) - J, J = the fixed-T Jacobian + (df/dT)(dT/drho)
            const double dT = 1e-7 * T0;
            double c[NS];
#pragma unroll
            for (int n = 0; n < NS; ++n) c[n] = y[n] / mw[n] * 1e-6;
            jacobian_fixed_T(c, T0, M);
            double dTdy[NS];
            {
                const double tc1 = T0, tc2 = T0 * T0, tc3 = T0 * T0 * T0, tc4 = T0 * T0 * T0 * T0;
                const double invT = 1.0 / T0;
                const bool low = T0 < 1000;
                const double cv = ckcvbs(T0, y);
#pragma unroll
                for (int s2 = 0; s2 < NS; ++s2) {
                    const double* a = low ? lo[s2] : hi[s2];
                    double u = a[0] - 1.0 + a[1] / 2.0 * tc1 + a[2] / 3.0 * tc2 + a[3] / 4.0 * tc3 + a[4] / 5.0 * tc4;
                    if (!low) u = u + (hi[s2][5] - lo[s2][5]) * invT;
                    dTdy[s2] = (-lo[s2][5] * Ru / mw[s2] - u * Ru * T0 / mw[s2]) / cv;
                }
            }
            const double gh = 1.0 / (h * 0.5);
#pragma unroll
            for (int i = 0; i < NS; ++i) {
                const double dfdT = (fs[i] - f0[i]) / dT;
#pragma unroll
                for (int j = 0; j < NS; ++j)
                    M(i, j) = -(M(i, j) * mw[i] / mw[j] + dfdT * dTdy[j]) + (i == j ? gh : 0.0);
            }
            lu_factor(M, piv);
#pragma unroll
            for (int n = 0; n < NS; ++n) K1[n] = f0[n];
            lu_solve(M, piv, K1);
#pragma unroll
            for (int n = 0; n < NS; ++n) K2[n] = f0[n] + (4.0 * hinv) * K1[n];
            lu_solve(M, piv, K2);
            Tprev = T0;
        } else if (it == 2) {
#pragma unroll
            for (int n = 0; n < NS; ++n) K3[n] = fs[n] + hinv * K1[n] - hinv * K2[n];
            lu_solve(M, piv, K3);
            Tprev = Tc;
        } else {
#pragma unroll
            for (int n = 0; n < NS; ++n) K4[n] = fs[n] + hinv * K1[n] - hinv * K2[n] - (8.0 / 3.0 * hinv) * K3[n];
            lu_solve(M, piv, K4);
            T4 = Tc;
        }
    }
    double err = 0.0;
    bool finite = true;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
        dy[n] = 2.0 * K1[n] + K3[n] + K4[n];
        ynew[n] = y[n] + dy[n];
        const double e = fabs(K4[n]) / (atol * rho + rtol * fmax(fabs(y[n]), fabs(ynew[n])));
        finite = finite && __builtin_isfinite(e) && __builtin_isfinite(ynew[n]);
        err = fmax(err, e);
    }
    Enew = E + energy_of(dy);
    Tlast = T4;
    // a stage that left the thermodynamic tables (a non-physical temperature,
    // non-finite rates) fails the step: it is retried with a smaller h
    return finite ? err : __builtin_huge_val();
}

#pragma clang fp contract(on)

} // namespace chem
} // namespace ocfd
