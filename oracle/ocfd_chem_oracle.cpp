// Li-Dryer H2/air chemistry -- CPU ORACLE (test infrastructure only).
//
// FP64 restatement of the reference's 9-species / 19-reaction kinetics, NASA7
// thermodynamics, temperature inversion, 9x9 Gauss solve, the per-cell
// linearised-implicit chemistry step and the species renormalisation, in the
// reference's operation order (built with -ffp-contract=off).  Pinned by the
// reference's own AMReX-free copy of these functions (Test/chemical_test.cpp,
// compiled into oracle/_ref by oracle/ref.mk), the reference's known answers
// (Exec/7_JET/prob.cpp:18-37), element conservation and a library ODE solve.
//
// Species order (Source/Constants.cpp:13-22): H2 O2 H2O H O OH HO2 H2O2 N2.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>

namespace {

constexpr int NS = 9;
constexpr int NR = 19;
constexpr double Ru = 8.31446261815324;
constexpr double Patm = 101325.0;

// kg/mol, Source/Constants.cpp:13-22
const double mw[NS] = {2.016 * 1e-3, 31.998 * 1e-3, 18.015 * 1e-3, 1.008 * 1e-3, 15.999 * 1e-3,
                       17.007 * 1e-3, 33.006 * 1e-3, 34.014 * 1e-3, 28.014 * 1e-3};

// NASA7 coefficients (low branch T < 1000 K, high branch), Source/Constants.cpp:23-73
const double nasa_lo[NS][7] = {
    {3.29812, 8.24944e-4, -81.43015e-8, -9.47543e-11, 413.48718e-15, -0.01013e+5, -3.29409},
    {3.21294, 11.27486e-4, -57.56150e-8, 131.38770e-11, -876.8554e-15, -0.01005e+5, 6.03474},
    {3.38684, 34.74982e-4, -635.4696e-8, 696.8581e-11, -2506.588e-15, -0.30208e+5, 2.59023},
    {2.5, 0.0, 0.0, 0.0, 0.0, 0.25472e+5, -0.46012},
    {2.94643, -16.3817e-4, 242.10320e-8, -160.284e-11, 389.06961e-15, 0.29148e+5, 2.96399},
    {3.63727, 1.85091e-4, -167.6165e-8, 238.72031e-11, -843.1442e-15, 0.03607e+5, 1.35886},
    {2.97996, 49.96697e-4, -379.0997e-8, 235.41919e-11, -808.9024e-15, 0.00176e+5, 9.22272},
    {3.38875, 65.69226e-4, -14.85013e-8, -462.58060e-11, 2471.515e-15, -0.17663e+5, 6.78536},
    {3.29868, 14.08240e-4, -396.3222e-8, 564.15149e-11, -2444.855e-15, -0.01021e+5, 3.95037}};
const double nasa_hi[NS][7] = {
    {2.99142, 7.00064e-4, -5.63383e-8, -0.92316e-11, 1.58275e-15, -0.00835e+5, -1.35511},
    {3.69758, 6.13520e-4, -12.58842e-8, 1.77528e-11, -1.13644e-15, -0.01234e+5, 3.18917},
    {2.67215, 30.56293e-4, -87.30260e-8, 12.00996e-11, -6.39162e-15, -0.29899e+5, 6.86282},
    {2.5, 0.0, 0.0, 0.0, 0.0, 0.25472e+5, -0.46012},
    {2.54206, -0.27551e-4, -0.31028e-8, 0.45511e-11, -0.43681e-15, 0.29231e+5, 4.92031},
    {2.88273, 10.13974e-4, -22.76877e-8, 2.17468e-11, -0.51263e-15, 0.03887e+5, 5.59571},
    {4.07219, 21.31296e-4, -53.08145e-8, 6.11227e-11, -2.84116e-15, -0.00158e+5, 3.47603},
    {4.57317, 43.36136e-4, -147.4689e-8, 23.48904e-11, -14.31654e-15, -0.18007e+5, 0.50114},
    {2.92664, 14.87977e-4, -56.84761e-8, 10.09704e-11, -6.75335e-15, -0.00923e+5, 5.98053}};

inline double HP(int n) { return nasa_lo[n][5]; }          // Constants.cpp:78
inline double A0(int n) { return nasa_lo[n][0] - 1.0; }    // Constants.cpp:81

// g/(R T), Source/CHEM/LiDryer.H:67-154
void gibbs(double* g, const double* tc)
{
    double T = tc[1];
    double invT = 1 / T;
    for (int s = 0; s < NS; ++s) {
        const double* a = T < 1000 ? nasa_lo[s] : nasa_hi[s];
        g[s] = a[0] * (1.0 - tc[0]) - a[1] / 2.0 * tc[1] - a[2] / 6.0 * tc[2] - a[3] / 12.0 * tc[3] -
               a[4] / 20.0 * tc[4] + a[5] * invT - a[6];
    }
}

// sensible e/(R T), Source/CHEM/LiDryer.H:531-614
void species_internal_energy(double* u, const double* tc)
{
    double T = tc[1];
    double invT = 1.0 / T;
    for (int s = 0; s < NS; ++s) {
        if (T < 1000) {
            const double* a = nasa_lo[s];
            u[s] = a[0] - 1.0 + a[1] / 2.0 * tc[1] + a[2] / 3.0 * tc[2] + a[3] / 4.0 * tc[3] + a[4] / 5.0 * tc[4];
        } else {
            const double* a = nasa_hi[s];
            u[s] = a[0] - 1.0 + a[1] / 2.0 * tc[1] + a[2] / 3.0 * tc[2] + a[3] / 4.0 * tc[3] + a[4] / 5.0 * tc[4] +
                   (nasa_hi[s][5] - nasa_lo[s][5]) * invT;
        }
    }
}

// sensible h/R in K, Source/CHEM/LiDryer.H:616-695
void species_enthalpy(double T, double* h)
{
    double tc[] = {T, T * T, T * T * T, T * T * T * T, T * T * T * T * T};
    for (int s = 0; s < NS; ++s) {
        if (T < 1000) {
            const double* a = nasa_lo[s];
            h[s] = a[0] * tc[0] + a[1] / 2.0 * tc[1] + a[2] / 3.0 * tc[2] + a[3] / 4.0 * tc[3] + a[4] / 5.0 * tc[4];
        } else {
            const double* a = nasa_hi[s];
            h[s] = a[0] * tc[0] + a[1] / 2.0 * tc[1] + a[2] / 3.0 * tc[2] + a[3] / 4.0 * tc[3] + a[4] / 5.0 * tc[4] +
                   (nasa_hi[s][5] - nasa_lo[s][5]);
        }
    }
}

// cp/R, Source/CHEM/LiDryer.H:717-787
void cp_R(double* c, const double* tc)
{
    double T = tc[1];
    for (int s = 0; s < NS; ++s) {
        const double* a = T < 1000 ? nasa_lo[s] : nasa_hi[s];
        c[s] = a[0] + a[1] * tc[1] + a[2] * tc[2] + a[3] * tc[3] + a[4] * tc[4];
    }
}

// J/m^3, Source/CHEM/LiDryer.H:697-715
double CKUBMS(double T, const double* rhoi)
{
    double result = 0;
    double tc[] = {0, T, T * T, T * T * T, T * T * T * T};
    double ums[NS];
    double RT = Ru * T;
    species_internal_energy(ums, tc);
    for (int n = 0; n < NS; ++n) result += rhoi[n] * ums[n] / mw[n];
    return result * RT;
}

// J/(m^3 K), Source/CHEM/LiDryer.H:789-816
double CKCPBS(double T, const double* rhoi)
{
    double result = 0;
    double tc[] = {0, T, T * T, T * T * T, T * T * T * T};
    double cpor[NS];
    cp_R(cpor, tc);
    for (int n = 0; n < NS; ++n) result += cpor[n] * rhoi[n] / mw[n];
    return result * Ru;
}

double CKCVBS(double T, const double* rhoi)
{
    double result = 0;
    double tc[] = {0, T, T * T, T * T * T, T * T * T * T};
    double cpor[NS];
    cp_R(cpor, tc);
    for (int n = 0; n < NS; ++n) result += (cpor[n] - 1.0) * rhoi[n] / mw[n];
    return result * Ru;
}

// Source/CHEM/LiDryer.H:818-834
double CKGAMMA(double T, const double* rhoi)
{
    double cp = 0, cv = 0;
    double tc[] = {0, T, T * T, T * T * T, T * T * T * T};
    double cpor[NS];
    cp_R(cpor, tc);
    for (int n = 0; n < NS; ++n) {
        cp += cpor[n] * rhoi[n] / mw[n];
        cv += (cpor[n] - 1.0) * rhoi[n] / mw[n];
    }
    return cp / cv;
}

// Source/CHEM/LiDryer.H:10-66
double CKPY(const double* rhoi, double T)
{
    double YOW = 0;
    for (int i = 0; i < NS; ++i) YOW += rhoi[i] / mw[i];
    return Ru * T * YOW;
}

double CKRHOY(double P, double T, const double* y)
{
    double YOW = 0, tmp[NS];
    for (int i = 0; i < NS; ++i) tmp[i] = y[i] / mw[i];
    for (int i = 0; i < NS; ++i) YOW += tmp[i];
    return P / (Ru * T * YOW);
}

void CKYTX(const double* y, double* x)
{
    double YOW = 0, tmp[NS];
    for (int i = 0; i < NS; i++) tmp[i] = y[i] / mw[i];
    for (int i = 0; i < NS; ++i) YOW += tmp[i];
    double YOWINV = 1.0 / YOW;
    for (int i = 0; i < NS; i++) x[i] = y[i] / mw[i] * YOWINV;
}

void CKXTY(const double* x, double* y)
{
    double XW = 0.0;
    for (int i = 0; i < NS; ++i) XW += x[i] * mw[i];
    for (int i = 0; i < NS; ++i) y[i] = x[i] * mw[i] / XW;
}

// Newton inversion of the sensible energy, Source/CHEM/LiDryer.H:836-888
double T_given_ey(double e, const double* rhoi, int* iters)
{
    const int maxiter = 200;
    const double tol = 1e-7;
    double ein = e;
    double tmin = 250, tmax = 3500;
    double emin = CKUBMS(tmin, rhoi);
    double emax = CKUBMS(tmax, rhoi);
    if (iters) *iters = 0;
    if (ein < emin) return tmin - (emin - ein) / CKCVBS(tmin, rhoi);
    if (ein > emax) return tmax - (emax - ein) / CKCVBS(tmax, rhoi);
    double As = 0;
    for (int n = 0; n < NS; ++n) As += A0(n) * Ru / mw[n] * rhoi[n];
    double t1 = e / As;
    if (t1 < tmin || t1 > tmax) t1 = tmin + (tmax - tmin) / (emax - emin) * (ein - emin);
    int i;
    for (i = 0; i < maxiter; ++i) {
        double e1 = CKUBMS(t1, rhoi);
        double cv = CKCVBS(t1, rhoi);
        double dt = (ein - e1) / cv;
        if (dt > 100.0) dt = 100.0;
        else if (dt < -100.0) dt = -100.0;
        else if (std::fabs(dt) < tol) break;
        else if (t1 + dt == t1) break;
        t1 += dt;
    }
    if (iters) *iters = i;
    return t1;
}

// Arrhenius forward rate constants (CGS), Source/CHEM/LiDryer.H:171-191:
// form 0: A exp(beta lnT - TA/T); 1: A exp(beta lnT); 2: A exp(c/T); 3: A
struct Arr { double A, beta, c; int form; };
const Arr kf_tab[21] = {
    {3.55e+15, -0.41, 8353.444256079512, 0}, {50800.0, 2.67, 3165.2508657072367, 0},
    {2.16e+08, 1.51, 1726.0429999007667, 0}, {2.97e+06, 2.02, 6743.141748883461, 0},
    {4.58e+19, -1.4, 52526.05490660117, 0}, {6.16e+15, -0.5, 0, 1}, {4.71e+18, -1, 0, 1}, {3.8e+22, -2, 0, 1},
    {1.48e+12, 0.6, 0, 1}, {1.66e+13, 0, -412.64001746898793, 2}, {7.08e+13, 0, -150.96586004962973, 2},
    {3.25e+13, 0, 0, 3}, {2.89e+13, 0, +251.60976674938289, 2}, {4.2e+14, 0, -6028.570011315214, 2},
    {1.3e+11, 0, +820.2478396029883, 2}, {2.95e+14, 0, -24355.825421340265, 2}, {2.41e+13, 0, -1997.7815479901, 2},
    {4.82e+13, 0, -4000.595291315188, 2}, {9.55e+06, 2, 1997.7815479901, 0}, {1.0e+12, 0, 0, 3},
    {5.8e+14, 0, -4810.778740248201, 2}};

// Equilibrium constants, Source/CHEM/LiDryer.H:197-219: exp((g[a0]+g[a1]) - (g[b0]+g[b1]))
// times 1 (fac 0), 1/RsT (fac 1) or RsT (fac 2); -1 = absent term
struct KcDef { int a0, a1, b0, b1, fac; };
const KcDef kc_tab[21] = {
    {3, 1, 4, 5, 0}, {4, 0, 3, 5, 0}, {0, 5, 2, 3, 0}, {4, 2, 5, 5, 0}, {0, -1, 3, 3, 1}, {4, 4, 1, -1, 2},
    {4, 3, 5, -1, 2}, {3, 5, 2, -1, 2}, {3, 1, 6, -1, 2}, {6, 3, 0, 1, 0}, {6, 3, 5, 5, 0}, {6, 4, 1, 5, 0},
    {6, 5, 2, 1, 0}, {6, 6, 7, 1, 0}, {6, 6, 7, 1, 0}, {7, -1, 5, 5, 1}, {7, 3, 2, 5, 0}, {7, 3, 6, 0, 0},
    {7, 4, 5, 6, 0}, {7, 5, 6, 2, 0}, {7, 5, 6, 2, 0}};

// Reactions, Source/CHEM/LiDryer.H:232-510.  f0,f1 / r0,r1: species of the
// forward / reverse concentration products (-1 = absent); kf: rate index;
// dup: second rate index of a duplicate (-1 none); tb: third body
// (0 none, 1 M with H2 2.5 H2O 12, 2 falloff HO2 (Fcent 0.8), 3 falloff H2O2 (Fcent 0.5)).
struct Reac { int f0, f1, r0, r1, kf, dup, tb; int nuf[NS]; int nur[NS]; };
const Reac rx[NR] = {
    {3, 1, 4, 5, 0, -1, 0, {0, 1, 0, 1, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 1, 1, 0, 0, 0}},
    {4, 0, 3, 5, 1, -1, 0, {1, 0, 0, 0, 1, 0, 0, 0, 0}, {0, 0, 0, 1, 0, 1, 0, 0, 0}},
    {0, 5, 2, 3, 2, -1, 0, {1, 0, 0, 0, 0, 1, 0, 0, 0}, {0, 0, 1, 1, 0, 0, 0, 0, 0}},
    {4, 2, 5, 5, 3, -1, 0, {0, 0, 1, 0, 1, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 2, 0, 0, 0}},
    {0, -1, 3, 3, 4, -1, 1, {1, 0, 0, 0, 0, 0, 0, 0, 0}, {0, 0, 0, 2, 0, 0, 0, 0, 0}},
    {4, 4, 1, -1, 5, -1, 1, {0, 0, 0, 0, 2, 0, 0, 0, 0}, {0, 1, 0, 0, 0, 0, 0, 0, 0}},
    {4, 3, 5, -1, 6, -1, 1, {0, 0, 0, 1, 1, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 1, 0, 0, 0}},
    {3, 5, 2, -1, 7, -1, 1, {0, 0, 0, 1, 0, 1, 0, 0, 0}, {0, 0, 1, 0, 0, 0, 0, 0, 0}},
    {3, 1, 6, -1, 8, -1, 2, {0, 1, 0, 1, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0, 1, 0, 0}},
    {6, 3, 0, 1, 9, -1, 0, {0, 0, 0, 1, 0, 0, 1, 0, 0}, {1, 1, 0, 0, 0, 0, 0, 0, 0}},
    {6, 3, 5, 5, 10, -1, 0, {0, 0, 0, 1, 0, 0, 1, 0, 0}, {0, 0, 0, 0, 0, 2, 0, 0, 0}},
    {6, 4, 1, 5, 11, -1, 0, {0, 0, 0, 0, 1, 0, 1, 0, 0}, {0, 1, 0, 0, 0, 1, 0, 0, 0}},
    {6, 5, 2, 1, 12, -1, 0, {0, 0, 0, 0, 0, 1, 1, 0, 0}, {0, 1, 1, 0, 0, 0, 0, 0, 0}},
    {6, 6, 7, 1, 13, 14, 0, {0, 0, 0, 0, 0, 0, 2, 0, 0}, {0, 1, 0, 0, 0, 0, 0, 1, 0}},
    {7, -1, 5, 5, 15, -1, 3, {0, 0, 0, 0, 0, 0, 0, 1, 0}, {0, 0, 0, 0, 0, 2, 0, 0, 0}},
    {7, 3, 2, 5, 16, -1, 0, {0, 0, 0, 1, 0, 0, 0, 1, 0}, {0, 0, 1, 0, 0, 1, 0, 0, 0}},
    {7, 3, 6, 0, 17, -1, 0, {0, 0, 0, 1, 0, 0, 0, 1, 0}, {1, 0, 0, 0, 0, 0, 1, 0, 0}},
    {7, 4, 5, 6, 18, -1, 0, {0, 0, 0, 0, 1, 0, 0, 1, 0}, {0, 0, 0, 0, 0, 1, 1, 0, 0}},
    {7, 5, 6, 2, 19, 20, 0, {0, 0, 0, 0, 0, 1, 0, 1, 0}, {0, 0, 1, 0, 0, 0, 1, 0, 0}}};

inline double prod(const double* sc, int a, int b) { return b < 0 ? sc[a] : sc[a] * sc[b]; }

// Production rates (mol/cm^3/s) and their Jacobian d wdot_l / d c_n at fixed T,
// Source/CHEM/LiDryer.H:156-529.
void production_rate(double* wdot, double (*Arate)[NS], const double* sc, double T)
{
    const double eps = 1e-20;
    double tc[5];
    tc[0] = std::log(T);
    tc[1] = T;
    tc[2] = T * (T);
    tc[3] = tc[2] * tc[1];
    tc[4] = tc[2] * tc[2];
    double invT = 1.0 / T;
    double kfs[21];
    for (int r = 0; r < 21; ++r) {
        const Arr& a = kf_tab[r];
        if (a.form == 0) kfs[r] = a.A * std::exp(a.beta * tc[0] - a.c * invT);
        else if (a.form == 1) kfs[r] = a.A * std::exp(a.beta * tc[0]);
        else if (a.form == 2) kfs[r] = a.A * std::exp(a.c * invT);
        else kfs[r] = a.A;
    }
    double g[NS];
    gibbs(g, tc);
    double RsT = Ru / Patm * 1e6 * T;
    double Kc[21];
    for (int r = 0; r < 21; ++r) {
        const KcDef& d = kc_tab[r];
        double sa = d.a1 < 0 ? g[d.a0] : g[d.a0] + g[d.a1];
        double sb = d.b1 < 0 ? g[d.b0] : g[d.b0] + g[d.b1];
        double e = std::exp((sa) - (sb));
        Kc[r] = d.fac == 0 ? e : (d.fac == 1 ? 1.0 / RsT * e : RsT * e);
    }
    double mixture = 0.0;
    for (int n = 0; n < NS; n++) {
        mixture += sc[n];
        wdot[n] = 0.0;
        for (int l = 0; l < NS; l++) Arate[n][l] = 0.0;
    }
    double q_f[NR], q_r[NR];
    for (int m = 0; m < NR; ++m) {
        const Reac& R = rx[m];
        double phi_f = prod(sc, R.f0, R.f1);
        double phi_r = prod(sc, R.r0, R.r1);
        double k_f = kfs[R.kf];
        if (R.tb == 1) {
            double alpha = mixture + 1.5 * sc[0] + 11 * sc[2];
            k_f = alpha * k_f;
        } else if (R.tb >= 2) {
            double alpha, redP, Fc;
            if (R.tb == 2) {
                alpha = mixture + sc[0] + 10 * sc[2] + -0.22 * sc[1];
                redP = alpha / k_f * 6.37e+20 * std::exp(-1.72 * tc[0] - 261.6741574193582 * invT);
                Fc = 0.8;
            } else {
                alpha = mixture + 1.5 * sc[0] + 11 * sc[2];
                redP = alpha / k_f * 1.20e+17 * std::exp(-22846.166820843966 * invT);
                Fc = 0.5;
            }
            double logPred = std::log10(redP);
            double logFcent = std::log10(Fc);
            double troe_c = -0.4 - 0.67 * logFcent;
            double troe_n = 0.75 - 1.27 * logFcent;
            double troe = (troe_c + logPred) / (troe_n - 0.14 * (troe_c + logPred));
            double F_troe = std::pow(10.0, logFcent / (1.0 + troe * troe));
            k_f *= (redP / (1.0 + redP)) * F_troe;
        }
        q_f[m] = phi_f * k_f;
        q_r[m] = phi_r * (k_f / Kc[R.kf]);
        if (R.dup >= 0) {
            double k2 = kfs[R.dup];
            q_f[m] += phi_f * k2;
            q_r[m] += phi_r * (k2 / Kc[R.dup]);
        }
    }
    for (int m = 0; m < NR; ++m) {
        const Reac& R = rx[m];
        double wf1 = q_f[m], wr1 = q_r[m];
        for (int n = 0; n < NS; ++n) wdot[n] += (wf1 - wr1) * (R.nur[n] - R.nuf[n]);
        for (int n = 0; n < NS; ++n) {
            double Awf = R.nuf[n] * wf1 / (sc[n] + eps);
            double Awr = R.nur[n] * wr1 / (sc[n] + eps);
            for (int l = 0; l < NS; ++l) Arate[l][n] += (Awf - Awr) * (R.nur[l] - R.nuf[l]);
        }
    }
}

// 9x9 Gaussian elimination with partial pivoting, Source/CHEM/LiDryer.H:890-942
void gauss(double* x, const double (*A)[NS], const double* b)
{
    double U[NS][NS + 1];
    for (int i = 0; i < NS; i++) {
        U[i][NS] = b[i];
        for (int j = 0; j < NS; j++) U[i][j] = A[i][j];
    }
    for (int k = 0; k < NS; k++) {
        int p = k;
        double maxPivot = 0.0;
        for (int i = k; i < NS; i++)
            if (std::fabs(U[i][k]) > maxPivot) { maxPivot = std::fabs(U[i][k]); p = i; }
        if (p != k)
            for (int i = k; i < NS + 1; i++) std::swap(U[p][i], U[k][i]);
        for (int i = k + 1; i < NS; i++) {
            double m = U[i][k] / U[k][k];
            for (int j = k; j < NS + 1; j++) U[i][j] -= m * U[k][j];
        }
    }
    for (int k = NS - 1; k >= 0; k--) {
        double sum = U[k][NS];
        for (int j = k + 1; j < NS; j++) sum -= U[k][j] * x[j];
        x[k] = sum / U[k][k];
    }
}

// One linearised-implicit chemistry step of one cell (Source/CHEM/ChemAdvance.cpp:36-81):
// u[5] conserved state (E sensible + kinetic), rhoi[9] partial densities, in place.
void chem_step(double* u, double* rhoi, double dt, double smallr)
{
    double rho = std::max(u[0], smallr);
    double rhoinv = 1.0 / rho;
    double ux = u[1] * rhoinv, uy = u[2] * rhoinv, uz = u[3] * rhoinv;
    double kineng = 0.5 * (ux * ux + uy * uy + uz * uz);
    double ei = u[4] - rho * kineng;
    double rhoi_0[NS], c[NS], wdot[NS], Arate[NS][NS], A1[NS][NS], rhs[NS], drho[NS];
    for (int n = 0; n < NS; ++n) rhoi_0[n] = rhoi[n];
    double T = T_given_ey(ei, rhoi_0, nullptr);
    for (int n = 0; n < NS; n++) c[n] = rhoi_0[n] / mw[n] * 1e-6;
    production_rate(wdot, Arate, c, T);
    for (int i = 0; i < NS; ++i)
        for (int j = 0; j < NS; ++j) {
            A1[i][j] = (i == j ? 1.0 : 0.0);
            A1[i][j] -= Arate[i][j] * mw[i] / mw[j] * dt;
        }
    for (int n = 0; n < NS; ++n) rhs[n] = wdot[n] * mw[n] * 1e6 * dt;
    gauss(drho, A1, rhs);
    for (int n = 0; n < NS; ++n) {
        rhoi[n] += drho[n];
        if (rhoi[n] < 0) rhoi[n] = 0.0;
    }
    double dei = 0;
    for (int n = 0; n < NS; ++n) dei += -HP(n) * Ru / mw[n] * drho[n];
    u[4] += dei;
}

// Species clip + rescale to rho, Source/EBR.cpp:341-355
void renorm(const double* u, double* rhoi)
{
    double rho0 = 0;
    for (int n = 0; n < NS; ++n) {
        if (rhoi[n] < 0) rhoi[n] = 0.0;
        rho0 += rhoi[n];
    }
    double tmp = u[0] / rho0;
    for (int n = 0; n < NS; ++n) rhoi[n] *= tmp;
}

// Adaptive integration of one constant-volume reactor with step doubling on
// the linearised-implicit step: one step of h (y1) against two of h/2 (y2);
// error max_k |y2-y1| / (atol*rho + rtol*|y2|) over the partial densities; on
// acceptance the Richardson value 2*y2 - y1 (second order, densities clipped at
// 0) is kept; h <- h * clamp(0.9/sqrt(err), 0.2, 4).  Returns accepted steps,
// or -1 when max_steps attempts did not reach t_end.
int reactor_adaptive(double* u, double* rhoi, double t_end, double h0, double rtol, double atol, int max_steps,
                     int* rejects)
{
    double t = 0.0, h = h0;
    int acc = 0, rej = 0;
    const double rho = u[0];
    while (t < t_end) {
        if (acc + rej >= max_steps) { if (rejects) *rejects = rej; return -1; }
        if (t + h > t_end) h = t_end - t;
        double u1[5], r1[NS], u2[5], r2[NS];
        std::memcpy(u1, u, sizeof(u1)); std::memcpy(r1, rhoi, sizeof(r1));
        std::memcpy(u2, u, sizeof(u2)); std::memcpy(r2, rhoi, sizeof(r2));
        chem_step(u1, r1, h, 1e-19);
        chem_step(u2, r2, 0.5 * h, 1e-19);
        chem_step(u2, r2, 0.5 * h, 1e-19);
        double err = 0.0;
        for (int n = 0; n < NS; ++n) {
            double e = std::fabs(r2[n] - r1[n]) / (atol * rho + rtol * std::fabs(r2[n]));
            err = std::max(err, e);
        }
        if (err <= 1.0) {
            for (int n = 0; n < NS; ++n) rhoi[n] = std::max(2.0 * r2[n] - r1[n], 0.0);
            u[4] = 2.0 * u2[4] - u1[4];
            t += h;
            ++acc;
            double f = err > 0.0 ? 0.9 / std::sqrt(err) : 4.0;
            h *= std::min(4.0, std::max(0.2, f));
        } else {
            ++rej;
            h *= std::max(0.2, 0.9 / std::sqrt(err));
        }
    }
    if (rejects) *rejects = rej;
    return acc;
}

} // namespace

extern "C" {

void orc_chem_tables(double* mw_out, double* lo_out, double* hi_out)
{
    for (int n = 0; n < NS; ++n) {
        mw_out[n] = mw[n];
        for (int c = 0; c < 7; ++c) { lo_out[7 * n + c] = nasa_lo[n][c]; hi_out[7 * n + c] = nasa_hi[n][c]; }
    }
}

void orc_production_rate(const double* c, double T, double* wdot, double* jac)
{
    double A[NS][NS];
    production_rate(wdot, A, c, T);
    std::memcpy(jac, A, sizeof(A));
}

void orc_gauss(const double* A, const double* b, double* x)
{
    gauss(x, (const double(*)[NS])A, b);
}

double orc_T_given_ey(double e, const double* rhoi, int* iters) { return T_given_ey(e, rhoi, iters); }
double orc_CKUBMS(double T, const double* rhoi) { return CKUBMS(T, rhoi); }
double orc_CKCVBS(double T, const double* rhoi) { return CKCVBS(T, rhoi); }
double orc_CKCPBS(double T, const double* rhoi) { return CKCPBS(T, rhoi); }
double orc_CKGAMMA(double T, const double* rhoi) { return CKGAMMA(T, rhoi); }
double orc_CKPY(const double* rhoi, double T) { return CKPY(rhoi, T); }
double orc_CKRHOY(double P, double T, const double* y) { return CKRHOY(P, T, y); }
void orc_CKXTY(const double* x, double* y) { CKXTY(x, y); }
void orc_CKYTX(const double* y, double* x) { CKYTX(y, x); }
void orc_species_enthalpy(double T, double* h) { species_enthalpy(T, h); }

// n cells: u (n x 5), rhoi (n x 9), in place, one step of dt each
void orc_chem_step_many(long n, double* u, double* rhoi, double dt, double smallr)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (long m = 0; m < n; ++m) chem_step(u + 5 * m, rhoi + NS * m, dt, smallr);
}

void orc_renorm_many(long n, const double* u, double* rhoi)
{
    for (long m = 0; m < n; ++m) renorm(u + 5 * m, rhoi + NS * m);
}

// Constant-volume H2/air reactor at (T0, p0, phi): X(H2):X(O2):X(N2) = 2phi : 1 : 3.76
void orc_reactor_init(long n, const double* T0, const double* p0, const double* phi, double* u, double* rhoi)
{
    for (long m = 0; m < n; ++m) {
        double X[NS] = {0}, Y[NS];
        double tot = 2.0 * phi[m] + 1.0 + 3.76;
        X[0] = 2.0 * phi[m] / tot;
        X[1] = 1.0 / tot;
        X[8] = 3.76 / tot;
        CKXTY(X, Y);
        double rho = CKRHOY(p0[m], T0[m], Y);
        double* r = rhoi + NS * m;
        for (int k = 0; k < NS; ++k) r[k] = Y[k] * rho;
        double* uu = u + 5 * m;
        uu[0] = rho; uu[1] = uu[2] = uu[3] = 0.0;
        uu[4] = CKUBMS(T0[m], r);
    }
}

// Fixed-step integration: nsteps linearised-implicit steps of dt
void orc_reactor_fixed(long n, double* u, double* rhoi, double dt, int nsteps)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (long m = 0; m < n; ++m)
        for (int s = 0; s < nsteps; ++s) chem_step(u + 5 * m, rhoi + NS * m, dt, 1e-19);
}

// Adaptive integration to t_end; steps[m] = accepted steps (-1 on failure)
void orc_reactor_adaptive(long n, double* u, double* rhoi, double t_end, double h0, double rtol, double atol,
                          int max_steps, int* steps, int* rejects)
{
#pragma omp parallel for schedule(dynamic, 4)
    for (long m = 0; m < n; ++m)
        steps[m] = reactor_adaptive(u + 5 * m, rhoi + NS * m, t_end, h0, rtol, atol, max_steps,
                                    rejects ? rejects + m : nullptr);
}

// Temperature of each reactor (sensible energy inversion)
void orc_reactor_T(long n, const double* u, const double* rhoi, double* T)
{
    for (long m = 0; m < n; ++m) {
        const double* uu = u + 5 * m;
        double rho = uu[0];
        double ux = uu[1] / rho, uy = uu[2] / rho, uz = uu[3] / rho;
        double ei = uu[4] - rho * (0.5 * (ux * ux + uy * uy + uz * uz));
        T[m] = T_given_ey(ei, rhoi + NS * m, nullptr);
    }
}

} // extern "C"
