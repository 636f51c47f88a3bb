// OpenCFD-AMR per-box advance -- CPU ORACLE (test infrastructure only).
//
// This file is the checker, never the product: only tests/, the smoke check in
// __graft_entry__.py and the cpu_baseline leg of bench.py may load it.  It is a
// plain FP64 C++/OpenMP restatement of the reference algorithm, written to follow
// the reference's operation order so that the HIP path can be compared against
// it at 1e-10 relative.  Build: oracle/Makefile (g++ -O2 -fopenmp
// -ffp-contract=off, so no FMA contraction changes the reference's rounding).
//
// Each function cites the reference file:line it restates.  How each is pinned
// independently of this file (closed forms, invariants, textbook solutions and
// the reference's own known answers) is listed in tests/test_oracle_*.py.
//
// Layout follows amrex::Array4: component-major SoA, i fastest, indices are
// global cell indices, a FAB covers an inclusive box [lo, hi].

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdlib>
#include <algorithm>
#include <limits>
#include <vector>

extern "C" {

// Source/Constants.H:7-64 (non-CHEM members, same order)
typedef struct orc_parm {
    double eos_gamma, eos_m, Patm, cv, cp, Ru, Rg, kOverMu, Pr, C_s, T_s,
           smallr, smallp, eps, twoThirds, oneSix;
} orc_parm;

typedef struct orc_fab {
    double* p;
    int lo[3];
    int hi[3];
    int ncomp;
} orc_fab;

// Embedded-boundary fields of one box over an index range [lo, hi] (at least
// grow(bx,4)): EBCellFlag words, volume fraction, low-face area fractions.
typedef struct orc_ebfab {
    const uint32_t* flag;
    const double* vfrac;
    const double* apx;
    const double* apy;
    const double* apz;
    int lo[3];
    int hi[3];
} orc_ebfab;

// Source/Constants.H:9-29 defaults + Parm::Initialize, Source/Constants.cpp:7-11
void orc_parm_default(orc_parm* p)
{
    p->eos_gamma = 1.4;
    p->eos_m = 28.97e-3;
    p->Patm = 101325.0;
    p->Ru = 8.31446261815324;
    p->Pr = 0.72;
    p->C_s = 1.458e-6;
    p->T_s = 110.4;
    p->smallr = 1.e-19;
    p->smallp = 1.e-10;
    p->eps = 1.e-40;
    p->twoThirds = 2.0 / 3.0;
    p->oneSix = 1.0 / 6.0;
    p->Rg = p->Ru / p->eos_m;
    p->cv = p->Rg / (p->eos_gamma - 1.0);
    p->cp = p->eos_gamma * p->cv;
    p->kOverMu = p->cp / p->Pr;
}

// Parm::Initialize after the caller set eos_gamma/eos_m/Pr (EBR.cpp:522-528)
void orc_parm_initialize(orc_parm* p)
{
    p->Rg = p->Ru / p->eos_m;
    p->cv = p->Rg / (p->eos_gamma - 1.0);
    p->cp = p->eos_gamma * p->cv;
    p->kOverMu = p->cp / p->Pr;
}

} // extern "C"

namespace {

enum { URHO = 0, UMX = 1, UMY = 2, UMZ = 3, UEDEN = 4, NCONS = 5 };   // IndexDefines.H:4-9
enum { QRHO = 0, QU = 1, QV = 2, QW = 3, QPRES = 4, QT = 5, QC = 6, NPRIM = 7 }; // IndexDefines.H:11-22

struct View {
    double* p;
    int lo[3], hi[3];
    long nx, ny, nz;
    int nc;
    View() : p(nullptr), nx(0), ny(0), nz(0), nc(0) {}
    View(const orc_fab& f) : p(f.p), nc(f.ncomp) {
        for (int d = 0; d < 3; ++d) { lo[d] = f.lo[d]; hi[d] = f.hi[d]; }
        nx = hi[0] - lo[0] + 1; ny = hi[1] - lo[1] + 1; nz = hi[2] - lo[2] + 1;
    }
    View(double* ptr, const int* l, const int* h, int ncomp) : p(ptr), nc(ncomp) {
        for (int d = 0; d < 3; ++d) { lo[d] = l[d]; hi[d] = h[d]; }
        nx = hi[0] - lo[0] + 1; ny = hi[1] - lo[1] + 1; nz = hi[2] - lo[2] + 1;
    }
    inline double& operator()(int i, int j, int k, int n) const {
        return p[(((long)n * nz + (k - lo[2])) * ny + (j - lo[1])) * nx + (i - lo[0])];
    }
    long npts() const { return nx * ny * nz; }
};

struct Box {
    int lo[3], hi[3];
};

inline Box grow(const Box& b, int g) {
    Box r = b;
    for (int d = 0; d < 3; ++d) { r.lo[d] -= g; r.hi[d] += g; }
    return r;
}

inline Box surrounding_nodes(const Box& b, int d) {
    Box r = b;
    r.hi[d] += 1;
    return r;
}

// Source/kernel/Kernels.H:13-40
inline void c2prim(int i, int j, int k, const View& u, const View& q, const orc_parm& parm)
{
    double rho = std::max(u(i, j, k, URHO), parm.smallr);
    double rhoinv = 1.0 / rho;
    double ux = u(i, j, k, UMX) * rhoinv;
    double uy = u(i, j, k, UMY) * rhoinv;
    double uz = u(i, j, k, UMZ) * rhoinv;
    double kineng = 0.5 * (ux * ux + uy * uy + uz * uz);
    double ei = u(i, j, k, UEDEN) - rho * kineng;
    double p = std::max((parm.eos_gamma - 1.0) * ei, parm.smallp);
    double T = p / (rho * parm.Rg);
    q(i, j, k, QRHO) = rho;
    q(i, j, k, QU) = ux;
    q(i, j, k, QV) = uy;
    q(i, j, k, QW) = uz;
    q(i, j, k, QPRES) = p;
    q(i, j, k, QT) = T;
    q(i, j, k, QC) = std::sqrt(parm.eos_gamma * parm.Rg * T);
}

// Steger-Warming split, Source/kernel/FluxSplit.H:11-70 (x), :72-131 (y), :133-193 (z).
// The three directions differ only in which velocity is normal; the sums
// vvc1/vvc2 keep the reference's x,y,z term order with the normal term replaced.
inline void sw_point(int dir, const double* qv, double* fpv, double* fmv, double gamma)
{
    double d = qv[QRHO], u = qv[QU], v = qv[QV], w = qv[QW], c = qv[QC];
    double tmp1 = 1.0 / (2.0 * gamma);
    double tmp2 = 2.0 * (gamma - 1.0);
    double tmp3 = (3.0 - gamma) / (2.0 * (gamma - 1.0));
    double un = dir == 0 ? u : (dir == 1 ? v : w);
    double uc1 = un - c;
    double uc2 = un + c;
    double vvc1, vvc2;
    if (dir == 0) {
        vvc1 = 0.5 * (uc1 * uc1 + v * v + w * w);
        vvc2 = 0.5 * (uc2 * uc2 + v * v + w * w);
    } else if (dir == 1) {
        vvc1 = 0.5 * (u * u + uc1 * uc1 + w * w);
        vvc2 = 0.5 * (u * u + uc2 * uc2 + w * w);
    } else {
        vvc1 = 0.5 * (u * u + v * v + uc1 * uc1);
        vvc2 = 0.5 * (u * u + v * v + uc2 * uc2);
    }
    double v2 = (gamma - 1) * (u * u + v * v + w * w);
    double w2 = tmp3 * c * c;
    double E1 = un, E2 = un - c, E3 = un + c;
    double E1P = (E1 + std::fabs(E1)) * 0.5;
    double E2P = (E2 + std::fabs(E2)) * 0.5;
    double E3P = (E3 + std::fabs(E3)) * 0.5;
    double E1M = E1 - E1P, E2M = E2 - E2P, E3M = E3 - E3P;
    double tmp0 = d * tmp1;
    const double vel[3] = {u, v, w};
    for (int side = 0; side < 2; ++side) {
        double a = side == 0 ? E1P : E1M;
        double b = side == 0 ? E2P : E2M;
        double e = side == 0 ? E3P : E3M;
        double* f = side == 0 ? fpv : fmv;
        f[URHO] = tmp0 * (tmp2 * a + b + e);
        for (int m = 0; m < 3; ++m) {
            if (m == dir) f[UMX + m] = tmp0 * (tmp2 * a * un + b * uc1 + e * uc2);
            else          f[UMX + m] = tmp0 * vel[m] * (tmp2 * a + b + e);
        }
        f[UEDEN] = tmp0 * (a * v2 + b * vvc1 + e * vvc2 + w2 * (b + e));
    }
}

// WENO-Z5 upwind-biased value from f+ at cells (f-3..f+1) of face f,
// Source/kernel/Reconstruction.H:83-110.
inline double weno_plus(double V1, double V2, double V3, double V4, double V5)
{
    double eps = 1e-6;
    double s11 = (13.0/12.0)*(V1-2.0*V2+V3)*(V1-2.0*V2+V3) + 0.25*(V1-4.0*V2+3.0*V3)*(V1-4.0*V2+3.0*V3);
    double s22 = (13.0/12.0)*(V2-2.0*V3+V4)*(V2-2.0*V3+V4) + 0.25*(V2-V4)*(V2-V4);
    double s33 = (13.0/12.0)*(V3-2.0*V4+V5)*(V3-2.0*V4+V5) + 0.25*(3.0*V3-4.0*V4+V5)*(3.0*V3-4.0*V4+V5);
    double s55 = std::fabs(s11 - s33);
    s11 = 1.0 + (s55/(eps+s11))*(s55/(eps+s11));
    s22 = 1.0 + (s55/(eps+s22))*(s55/(eps+s22));
    s33 = 1.0 + (s55/(eps+s33))*(s55/(eps+s33));
    double a1 = s11;
    double a2 = 6.0*s22;
    double a3 = 3.0*s33;
    double invsum = 1.0/(a1+a2+a3);
    double v1 = 2.0*V1-7.0*V2+11.0*V3;
    double v2 = -V2+5.0*V3+2.0*V4;
    double v3 = 2.0*V3+5.0*V4-V5;
    return (1.0/6.0)*invsum*(a1*v1+a2*v2+a3*v3);
}

// WENO-Z5 value from f- at cells (f-2..f+2) of face f,
// Source/kernel/Reconstruction.H:112-137.
inline double weno_minus(double V1, double V2, double V3, double V4, double V5)
{
    double eps = 1e-6;
    double s11 = (13.0/12.0)*(V5-2.0*V4+V3)*(V5-2.0*V4+V3) + 0.25*(V5-4.0*V4+3.0*V3)*(V5-4.0*V4+3.0*V3);
    double s22 = (13.0/12.0)*(V2-2.0*V3+V4)*(V2-2.0*V3+V4) + 0.25*(V4-V2)*(V4-V2);
    double s33 = (13.0/12.0)*(V3-2.0*V2+V1)*(V3-2.0*V2+V1) + 0.25*(3.0*V3-4.0*V2+V1)*(3.0*V3-4.0*V2+V1);
    double s55 = std::fabs(s11 - s33);
    s11 = 1.0 + (s55/(eps+s11))*(s55/(eps+s11));
    s22 = 1.0 + (s55/(eps+s22))*(s55/(eps+s22));
    s33 = 1.0 + (s55/(eps+s33))*(s55/(eps+s33));
    double a1 = s11;
    double a2 = 6.0*s22;
    double a3 = 3.0*s33;
    double invsum = 1.0/(a1+a2+a3);
    double v1 = 11.0*V3-7.0*V4+2.0*V5;
    double v2 = -V4+5.0*V3+2.0*V2;
    double v3 = 2.0*V3+5.0*V2-V1;
    return (1.0/6.0)*invsum*(a1*v1+a2*v2+a3*v3);
}

inline void sutherland(double T, const orc_parm& parm, double& mu)
{
    mu = parm.C_s*T*std::sqrt(T)/(T+parm.T_s);
}

// Source/kernel/Diffusion.H:7-47
inline void visc_x(int i, int j, int k, const View& q, const View& fx, const double* dxinv, const orc_parm& parm)
{
    double Tr = q(i,j,k,QT);
    double Tl = q(i-1,j,k,QT);
    double mutl, mutr;
    sutherland(Tl, parm, mutl);
    sutherland(Tr, parm, mutr);
    double ktl = mutl * parm.kOverMu;
    double ktr = mutr * parm.kOverMu;
    double dTdx = (Tr-Tl)*dxinv[0];
    double dudx = (q(i,j,k,QU)-q(i-1,j,k,QU))*dxinv[0];
    double dvdx = (q(i,j,k,QV)-q(i-1,j,k,QV))*dxinv[0];
    double dwdx = (q(i,j,k,QW)-q(i-1,j,k,QW))*dxinv[0];
    double dudy = (q(i,j+1,k,QU)+q(i-1,j+1,k,QU)-q(i,j-1,k,QU)-q(i-1,j-1,k,QU))*(0.25*dxinv[1]);
    double dvdy = (q(i,j+1,k,QV)+q(i-1,j+1,k,QV)-q(i,j-1,k,QV)-q(i-1,j-1,k,QV))*(0.25*dxinv[1]);
    double dudz = (q(i,j,k+1,QU)+q(i-1,j,k+1,QU)-q(i,j,k-1,QU)-q(i-1,j,k-1,QU))*(0.25*dxinv[2]);
    double dwdz = (q(i,j,k+1,QW)+q(i-1,j,k+1,QW)-q(i,j,k-1,QW)-q(i-1,j,k-1,QW))*(0.25*dxinv[2]);
    double divu = dudx + dvdy + dwdz;
    double muf = 0.5*(mutl+mutr);
    double tauxx = muf*(2.0*dudx-parm.twoThirds*divu);
    double tauxy = muf*(dudy+dvdx);
    double tauxz = muf*(dudz+dwdx);
    fx(i,j,k,UMX)   += -tauxx;
    fx(i,j,k,UMY)   += -tauxy;
    fx(i,j,k,UMZ)   += -tauxz;
    fx(i,j,k,UEDEN) += -0.5*((q(i,j,k,QU)+q(i-1,j,k,QU))*tauxx
                            +(q(i,j,k,QV)+q(i-1,j,k,QV))*tauxy
                            +(q(i,j,k,QW)+q(i-1,j,k,QW))*tauxz
                            +(ktl+ktr)*dTdx);
}

// Source/kernel/Diffusion.H:49-89
inline void visc_y(int i, int j, int k, const View& q, const View& fy, const double* dxinv, const orc_parm& parm)
{
    double Tr = q(i,j,k,QT);
    double Tl = q(i,j-1,k,QT);
    double mutl, mutr;
    sutherland(Tl, parm, mutl);
    sutherland(Tr, parm, mutr);
    double ktl = mutl * parm.kOverMu;
    double ktr = mutr * parm.kOverMu;
    double dTdy = (Tr-Tl)*dxinv[1];
    double dudy = (q(i,j,k,QU)-q(i,j-1,k,QU))*dxinv[1];
    double dvdy = (q(i,j,k,QV)-q(i,j-1,k,QV))*dxinv[1];
    double dwdy = (q(i,j,k,QW)-q(i,j-1,k,QW))*dxinv[1];
    double dudx = (q(i+1,j,k,QU)+q(i+1,j-1,k,QU)-q(i-1,j,k,QU)-q(i-1,j-1,k,QU))*(0.25*dxinv[0]);
    double dvdx = (q(i+1,j,k,QV)+q(i+1,j-1,k,QV)-q(i-1,j,k,QV)-q(i-1,j-1,k,QV))*(0.25*dxinv[0]);
    double dvdz = (q(i,j,k+1,QV)+q(i,j-1,k+1,QV)-q(i,j,k-1,QV)-q(i,j-1,k-1,QV))*(0.25*dxinv[2]);
    double dwdz = (q(i,j,k+1,QW)+q(i,j-1,k+1,QW)-q(i,j,k-1,QW)-q(i,j-1,k-1,QW))*(0.25*dxinv[2]);
    double divu = dudx + dvdy + dwdz;
    double muf = 0.5*(mutl+mutr);
    double tauyy = muf*(2.0*dvdy-parm.twoThirds*divu);
    double tauxy = muf*(dudy+dvdx);
    double tauyz = muf*(dwdy+dvdz);
    fy(i,j,k,UMX)   += -tauxy;
    fy(i,j,k,UMY)   += -tauyy;
    fy(i,j,k,UMZ)   += -tauyz;
    fy(i,j,k,UEDEN) += -0.5*((q(i,j,k,QU)+q(i,j-1,k,QU))*tauxy
                            +(q(i,j,k,QV)+q(i,j-1,k,QV))*tauyy
                            +(q(i,j,k,QW)+q(i,j-1,k,QW))*tauyz
                            +(ktl+ktr)*dTdy);
}

// Source/kernel/Diffusion.H:91-131
inline void visc_z(int i, int j, int k, const View& q, const View& fz, const double* dxinv, const orc_parm& parm)
{
    double Tr = q(i,j,k,QT);
    double Tl = q(i,j,k-1,QT);
    double mutl, mutr;
    sutherland(Tl, parm, mutl);
    sutherland(Tr, parm, mutr);
    double ktl = mutl * parm.kOverMu;
    double ktr = mutr * parm.kOverMu;
    double dTdz = (Tr-Tl)*dxinv[2];
    double dudz = (q(i,j,k,QU)-q(i,j,k-1,QU))*dxinv[2];
    double dvdz = (q(i,j,k,QV)-q(i,j,k-1,QV))*dxinv[2];
    double dwdz = (q(i,j,k,QW)-q(i,j,k-1,QW))*dxinv[2];
    double dudx = (q(i+1,j,k,QU)+q(i+1,j,k-1,QU)-q(i-1,j,k,QU)-q(i-1,j,k-1,QU))*(0.25*dxinv[0]);
    double dwdx = (q(i+1,j,k,QW)+q(i+1,j,k-1,QW)-q(i-1,j,k,QW)-q(i-1,j,k-1,QW))*(0.25*dxinv[0]);
    double dvdy = (q(i,j+1,k,QV)+q(i,j+1,k-1,QV)-q(i,j-1,k,QV)-q(i,j-1,k-1,QV))*(0.25*dxinv[1]);
    double dwdy = (q(i,j+1,k,QW)+q(i,j+1,k-1,QW)-q(i,j-1,k,QW)-q(i,j-1,k-1,QW))*(0.25*dxinv[1]);
    double divu = dudx + dvdy + dwdz;
    double muf = 0.5*(mutl+mutr);
    double tauxz = muf*(dudz+dwdx);
    double tauyz = muf*(dvdz+dwdy);
    double tauzz = muf*(2.0*dwdz-parm.twoThirds*divu);
    fz(i,j,k,UMX)   += -tauxz;
    fz(i,j,k,UMY)   += -tauyz;
    fz(i,j,k,UMZ)   += -tauzz;
    fz(i,j,k,UEDEN) += -0.5*((q(i,j,k,QU)+q(i,j,k-1,QU))*tauxz
                            +(q(i,j,k,QV)+q(i,j,k-1,QV))*tauyz
                            +(q(i,j,k,QW)+q(i,j,k-1,QW))*tauzz
                            +(ktl+ktr)*dTdz);
}

template <class F>
inline void for_box(const Box& b, F&& f)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = b.lo[2]; k <= b.hi[2]; ++k)
        for (int j = b.lo[1]; j <= b.hi[1]; ++j)
            for (int i = b.lo[0]; i <= b.hi[0]; ++i)
                f(i, j, k);
}

constexpr int NUM_GROW = 4;  // EBR.H:207

// Regular-box branch of EBR::compute_dSdt, Source/Advance.cpp:219-342:
// c2prim on grow(bx,4); per direction SW split on grow(bx,4), WENO-Z5 on the
// face box, optional viscous flux; divop on bx; optional gravity.  Temporaries
// live in host memory, as the reference's async-arena FABs do on the device.
void compute_dsdt_box(const Box& bx, const View& s, const View& dsdt, const double* dxinv,
                      int do_visc, int do_gravity, const orc_parm& parm, double* face_out[3])
{
    const Box bxg = grow(bx, NUM_GROW);
    std::vector<double> qbuf((size_t)NPRIM * (bxg.hi[0]-bxg.lo[0]+1) * (bxg.hi[1]-bxg.lo[1]+1) * (bxg.hi[2]-bxg.lo[2]+1));
    std::vector<double> fpbuf((size_t)NCONS * (bxg.hi[0]-bxg.lo[0]+1) * (bxg.hi[1]-bxg.lo[1]+1) * (bxg.hi[2]-bxg.lo[2]+1));
    std::vector<double> fmbuf(fpbuf.size());
    View q(qbuf.data(), bxg.lo, bxg.hi, NPRIM);
    View fp(fpbuf.data(), bxg.lo, bxg.hi, NCONS);
    View fm(fmbuf.data(), bxg.lo, bxg.hi, NCONS);

    Box fb[3];
    std::vector<double> fbuf[3];
    View fl[3];
    for (int d = 0; d < 3; ++d) {
        fb[d] = surrounding_nodes(bx, d);
        fbuf[d].assign((size_t)NCONS * (fb[d].hi[0]-fb[d].lo[0]+1) * (fb[d].hi[1]-fb[d].lo[1]+1) * (fb[d].hi[2]-fb[d].lo[2]+1), 0.0);
        fl[d] = View(fbuf[d].data(), fb[d].lo, fb[d].hi, NCONS);
    }

    for_box(bxg, [&](int i, int j, int k) { c2prim(i, j, k, s, q, parm); });

    for (int d = 0; d < 3; ++d) {
        for_box(bxg, [&](int i, int j, int k) {
            double qv[NPRIM], a[NCONS], b[NCONS];
            for (int n = 0; n < NPRIM; ++n) qv[n] = q(i, j, k, n);
            sw_point(d, qv, a, b, parm.eos_gamma);
            for (int n = 0; n < NCONS; ++n) { fp(i, j, k, n) = a[n]; fm(i, j, k, n) = b[n]; }
        });
        const int di = d == 0, dj = d == 1, dk = d == 2;
        const View F = fl[d];
        for_box(fb[d], [&](int i, int j, int k) {
            for (int n = 0; n < NCONS; ++n) {
                double pv = weno_plus(fp(i-3*di, j-3*dj, k-3*dk, n), fp(i-2*di, j-2*dj, k-2*dk, n),
                                      fp(i-di, j-dj, k-dk, n), fp(i, j, k, n), fp(i+di, j+dj, k+dk, n));
                double mv = weno_minus(fm(i-2*di, j-2*dj, k-2*dk, n), fm(i-di, j-dj, k-dk, n),
                                       fm(i, j, k, n), fm(i+di, j+dj, k+dk, n), fm(i+2*di, j+2*dj, k+2*dk, n));
                F(i, j, k, n) = pv + mv;
            }
        });
        if (do_visc) {
            for_box(fb[d], [&](int i, int j, int k) {
                if (d == 0) visc_x(i, j, k, q, F, dxinv, parm);
                else if (d == 1) visc_y(i, j, k, q, F, dxinv, parm);
                else visc_z(i, j, k, q, F, dxinv, parm);
            });
        }
    }

    // divop, Source/kernel/Kernels.H:84-97
    for_box(bx, [&](int i, int j, int k) {
        for (int n = 0; n < NCONS; ++n) {
            dsdt(i, j, k, n) = dxinv[0] * (fl[0](i, j, k, n) - fl[0](i+1, j, k, n))
                +              dxinv[1] * (fl[1](i, j, k, n) - fl[1](i, j+1, k, n))
                +              dxinv[2] * (fl[2](i, j, k, n) - fl[2](i, j, k+1, n));
        }
    });

    // gravity, Source/Advance.cpp:331-342
    if (do_gravity) {
        const double g = -9.8;
        for_box(bx, [&](int i, int j, int k) {
            dsdt(i, j, k, UMZ) += g * s(i, j, k, URHO);
            dsdt(i, j, k, UEDEN) += g * s(i, j, k, UMZ);
        });
    }

    if (face_out) {
        for (int d = 0; d < 3; ++d)
            if (face_out[d]) std::memcpy(face_out[d], fbuf[d].data(), fbuf[d].size() * sizeof(double));
    }
}


// ---------------------------------------------------------------------------
// Embedded-boundary box path, Source/EB/AdvanceBox.cpp:15-228 with
// Source/EB/EBkernels.H and Source/EB/EBdiffusion.H
// ---------------------------------------------------------------------------
struct EBV {
    const orc_ebfab* e;
    long nx, ny;
    explicit EBV(const orc_ebfab* f) : e(f), nx(f->hi[0] - f->lo[0] + 1), ny(f->hi[1] - f->lo[1] + 1) {}
    long at(int i, int j, int k) const { return ((long)(k - e->lo[2]) * ny + (j - e->lo[1])) * nx + (i - e->lo[0]); }
    uint32_t flag(int i, int j, int k) const { return e->flag[at(i, j, k)]; }
    double vf(int i, int j, int k) const { return e->vfrac[at(i, j, k)]; }
    double ax(int i, int j, int k) const { return e->apx[at(i, j, k)]; }
    double ay(int i, int j, int k) const { return e->apy[at(i, j, k)]; }
    double az(int i, int j, int k) const { return e->apz[at(i, j, k)]; }
};

// amrex::EBCellFlag (bit layout of AMReX_EBCellFlag.H): type in bits 0-1
// (0 regular, 1 single-valued, 3 covered), neighbour (i,j,k) connected at bit
// 5 + (i+1) + 3(j+1) + 9(k+1).  The reference also asks for offsets of 2
// (EBkernels.H:20,40,60); the same bit formula is applied to them, which is
// what the flag word answers for such an offset.
inline bool eb_conn(uint32_t f, int i, int j, int k) { return (f >> (5 + (i + 1) + 3 * (j + 1) + 9 * (k + 1))) & 1u; }
inline bool eb_covered(uint32_t f) { return (f & 3u) == 3u; }
inline bool eb_regular(uint32_t f) { return (f & 3u) == 0u; }

// Source/kernel/Reconstruction.H:9-19
inline double minmod(double a, double b)
{
    if (a * b > 0) return std::fabs(a) > std::fabs(b) ? b : a;
    return 0.0;
}

// eb_recon_{x,y,z}, Source/EB/EBkernels.H:10-71: left/right face states from
// the neighbour cells, with half minmod slopes where the cells connect
inline void eb_recon(int d, int i, int j, int k, int n, const View& q, const EBV& g, double& ql, double& qr)
{
    const int di = d == 0, dj = d == 1, dk = d == 2;
    const uint32_t f = g.flag(i, j, k);
    ql = q(i - di, j - dj, k - dk, n);
    qr = q(i, j, k, n);
    if (eb_conn(f, -di, -dj, -dk)) {
        if (eb_conn(f, -2 * di, -2 * dj, -2 * dk))
            ql = q(i - di, j - dj, k - dk, n) + 0.5 * minmod(q(i, j, k, n) - q(i - di, j - dj, k - dk, n),
                                                              q(i - di, j - dj, k - dk, n) - q(i - 2 * di, j - 2 * dj, k - 2 * dk, n));
        if (eb_conn(f, di, dj, dk))
            qr = q(i, j, k, n) - 0.5 * minmod(q(i + di, j + dj, k + dk, n) - q(i, j, k, n),
                                             q(i, j, k, n) - q(i - di, j - dj, k - dk, n));
    }
}

// One side of Van Leer's flux-vector splitting (B. van Leer, Flux-vector
// splitting for the Euler equations, Lecture Notes in Physics 170 (1982)
// 507-512) along direction cdir for q = (rho, u, v, w, p, c, ...): sgn = +1
// the forward part of the left state, -1 the backward part of the right one.
// Supersonic towards the part's side: the whole Euler flux; subsonic: the
// split mass flux sgn rho c (1 + sgn M)^2 / 4 carrying the normal momentum
// ((gamma-1) u_n + 2 sgn c)/gamma, the tangential velocities and the energy
// ((gamma-1) u_n + 2 sgn c)^2 / (2 (gamma^2-1)) + |u_t|^2 / 2; else nothing.
inline void vl_part(int cdir, const double* q, double sgn, double* f, const orc_parm& parm)
{
    const double rho = q[QRHO], p = std::max(q[QPRES], parm.smallp), c = q[QC];
    const double gamma = parm.eos_gamma;
    const double vel[3] = {q[QU], q[QV], q[QW]};
    const double un = vel[cdir];
    const double M = un / c;
    const double uu = vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2];
    if (sgn * M >= 1.0) {
        const double m = rho * un;
        f[0] = m;
        for (int a = 0; a < 3; ++a) f[1 + a] = m * vel[a] + (a == cdir ? p : 0.0);
        f[4] = un * (gamma * p / (gamma - 1.0) + 0.5 * rho * uu);
    } else if (std::fabs(M) < 1.0) {
        const double w = 1.0 + sgn * M;
        const double m = rho * c * (sgn * 0.250 * w * w);
        const double a2 = (gamma - 1) * un + sgn * 2.0 * c;
        f[0] = m;
        for (int a = 0; a < 3; ++a) f[1 + a] = a == cdir ? m * ((-un + sgn * 2.0 * c) / gamma + un) : m * vel[a];
        f[4] = m * (a2 * a2 / (2 * (gamma * gamma - 1)) + (uu - un * un) / 2);
    } else {
        for (int n = 0; n < NCONS; ++n) f[n] = 0.0;
    }
}

// the cut-cell face flux of Source/EB/EBkernels.H:73-171 (non-CHEM build):
// F = F+(left) + F-(right)
inline void eb_flux_point(int cdir, const double* ql, const double* qr, double* F, const orc_parm& parm)
{
    double fp[NCONS], fm[NCONS];
    vl_part(cdir, ql, 1.0, fp, parm);
    vl_part(cdir, qr, -1.0, fm, parm);
    for (int n = 0; n < NCONS; ++n) F[n] = fp[n] + fm[n];
}

// eb_compute_visc_{x,y,z}, Source/EB/EBdiffusion.H:8-203: as the regular
// viscous flux, with tangential differences one-sided where a cell does not
// connect, weighted by {0, 1, 0.5}[hi - lo] (AdvanceBox.cpp:40-43).  Written
// for direction d with (t1, t2) the two tangential directions in x,y,z order.
inline void eb_visc(int d, int i, int j, int k, const View& q, double* F, const EBV& g, const double* dxinv,
                    const orc_parm& parm)
{
    static const double weights[3] = {0.0, 1.0, 0.5};
    int e[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    const int* ed = e[d];
    const int im = i - ed[0], jm = j - ed[1], km = k - ed[2];
    double Tl = q(i, j, k, QT);
    double Tr = q(im, jm, km, QT);
    double mutl = parm.C_s * Tl * std::sqrt(Tl) / (Tl + parm.T_s);
    double mutr = parm.C_s * Tr * std::sqrt(Tr) / (Tr + parm.T_s);
    double muf = 0.5 * (mutl + mutr);
    double ktl = mutl * parm.kOverMu;
    double ktr = mutr * parm.kOverMu;
    double dTn = (Tl - Tr) * dxinv[d];
    double dn[3];   // normal derivatives of u, v, w
    for (int c = 0; c < 3; ++c) dn[c] = (q(i, j, k, QU + c) - q(im, jm, km, QU + c)) * dxinv[d];
    // tangential derivative of velocity component c along direction t
    auto tang = [&](int t, int c1, int c2, double& o1, double& o2) {
        const int* et = e[t];
        const uint32_t fh = g.flag(i, j, k), fl = g.flag(im, jm, km);
        int hp = eb_conn(fh, et[0], et[1], et[2]) ? 1 : 0;
        int hm = eb_conn(fh, -et[0], -et[1], -et[2]) ? 1 : 0;
        int lp = eb_conn(fl, et[0], et[1], et[2]) ? 1 : 0;
        int lm = eb_conn(fl, -et[0], -et[1], -et[2]) ? 1 : 0;
        double whi = weights[hp + hm];
        double wlo = weights[lp + lm];
        auto Q = [&](int ii, int jj, int kk, int sgn, int comp) {
            return q(ii + sgn * et[0], jj + sgn * et[1], kk + sgn * et[2], comp);
        };
        o1 = (0.5 * dxinv[t]) * ((Q(i, j, k, hp, c1) - Q(i, j, k, -hm, c1)) * whi
                                 + (Q(im, jm, km, lp, c1) - Q(im, jm, km, -lm, c1)) * wlo);
        o2 = (0.5 * dxinv[t]) * ((Q(i, j, k, hp, c2) - Q(i, j, k, -hm, c2)) * whi
                                 + (Q(im, jm, km, lp, c2) - Q(im, jm, km, -lm, c2)) * wlo);
    };
    double t1, t2, t3;   // the three stress components of this face, reference order
    if (d == 0) {
        double dudy, dvdy, dudz, dwdz;
        tang(1, QU, QV, dudy, dvdy);
        tang(2, QU, QW, dudz, dwdz);
        double divu = dn[0] + dvdy + dwdz;
        t1 = muf * (2.0 * dn[0] - parm.twoThirds * divu);   // tauxx
        t2 = muf * (dudy + dn[1]);                           // tauxy
        t3 = muf * (dudz + dn[2]);                           // tauxz
    } else if (d == 1) {
        double dudx, dvdx, dvdz, dwdz;
        tang(0, QU, QV, dudx, dvdx);
        tang(2, QV, QW, dvdz, dwdz);
        double divu = dudx + dn[1] + dwdz;
        double tauyy = muf * (2.0 * dn[1] - parm.twoThirds * divu);
        double tauxy = muf * (dn[0] + dvdx);
        double tauyz = muf * (dn[2] + dvdz);
        t1 = tauxy; t2 = tauyy; t3 = tauyz;
    } else {
        double dudx, dwdx, dvdy, dwdy;
        tang(0, QU, QW, dudx, dwdx);
        tang(1, QV, QW, dvdy, dwdy);
        double divu = dudx + dvdy + dn[2];
        double tauxz = muf * (dn[0] + dwdx);
        double tauyz = muf * (dn[1] + dwdy);
        double tauzz = muf * (2.0 * dn[2] - parm.twoThirds * divu);
        t1 = tauxz; t2 = tauyz; t3 = tauzz;
    }
    F[UMX] += -t1;
    F[UMY] += -t2;
    F[UMZ] += -t3;
    F[UEDEN] += -0.5 * ((q(i, j, k, QU) + q(im, jm, km, QU)) * t1 + (q(i, j, k, QV) + q(im, jm, km, QV)) * t2
                        + (q(i, j, k, QW) + q(im, jm, km, QW)) * t3 + (ktl + ktr) * dTn);
}

// eb_wallflux, Source/EB/EBkernels.H:173-214
inline void eb_wallflux(const double* qw, double axm, double axp, double aym, double ayp, double azm, double azp,
                        double* fluxw, const orc_parm& parm)
{
    double fp[NCONS], fm[NCONS];
    double d = qw[QRHO], u = qw[QU], v = qw[QV], w = qw[QW], p = qw[QPRES], c = qw[QC];
    double apnorm = std::sqrt((axm - axp) * (axm - axp) + (aym - ayp) * (aym - ayp) + (azm - azp) * (azm - azp));
    double un = u * (axm - axp) + v * (aym - ayp) + w * (azm - azp);
    double lambda0 = std::fabs(un) + c * apnorm;
    double gamma = parm.eos_gamma;
    double E0 = p / (gamma - 1) + 0.50 * d * (u * u + v * v + w * w);
    fp[0] = 0.5 * (d * un + lambda0 * d);
    fp[1] = 0.5 * (d * u * un + (axm - axp) * p + lambda0 * d * u);
    fp[2] = 0.5 * (d * v * un + (aym - ayp) * p + lambda0 * d * v);
    fp[3] = 0.5 * (d * w * un + (azm - azp) * p + lambda0 * d * w);
    fp[4] = 0.5 * ((E0 + p) * un + lambda0 * E0);
    fm[0] = 0.5 * (d * (-un) - lambda0 * d);
    fm[1] = 0.5 * (d * u * un + (axm - axp) * p + lambda0 * d * u);
    fm[2] = 0.5 * (d * v * un + (aym - ayp) * p + lambda0 * d * v);
    fm[3] = 0.5 * (d * w * un + (azm - azp) * p + lambda0 * d * w);
    fm[4] = 0.5 * (-(E0 + p) * un - lambda0 * E0);
    for (int n = 0; n < NCONS; ++n) fluxw[n] = fp[n] + fm[n];
}

// FabType of a region: 0 regular (all regular), 2 covered (all covered), else 1
inline int eb_region_type(const EBV& g, const Box& b)
{
    bool allreg = true, allcov = true;
    for (int k = b.lo[2]; k <= b.hi[2]; ++k)
        for (int j = b.lo[1]; j <= b.hi[1]; ++j)
            for (int i = b.lo[0]; i <= b.hi[0]; ++i) {
                const uint32_t f = g.flag(i, j, k);
                allreg = allreg && eb_regular(f);
                allcov = allcov && eb_covered(f);
            }
    return allcov ? 2 : (allreg ? 0 : 1);
}

// Box kind of EBR::compute_dSdt (Source/Advance.cpp:213,231): 2 covered (bx all
// covered), 0 regular (grow(bx,4) all regular), 1 the EB path.
inline int eb_box_kind(const EBV& g, const Box& bx)
{
    if (eb_region_type(g, bx) == 2) return 2;
    return eb_region_type(g, grow(bx, NUM_GROW)) == 0 ? 0 : 1;
}

// EBR::eb_compute_dSdt_box, Source/EB/AdvanceBox.cpp:15-228
void eb_compute_dsdt_box(const Box& bx, const View& s, const View& dsdt, const double* dxinv, const EBV& g,
                         int do_visc, int do_gravity, int do_redist, const orc_parm& parm)
{
    const Box bxg = grow(bx, NUM_GROW);
    std::vector<double> qbuf((size_t)NPRIM * (bxg.hi[0]-bxg.lo[0]+1) * (bxg.hi[1]-bxg.lo[1]+1) * (bxg.hi[2]-bxg.lo[2]+1));
    View q(qbuf.data(), bxg.lo, bxg.hi, NPRIM);
    for_box(bxg, [&](int i, int j, int k) { c2prim(i, j, k, s, q, parm); });

    Box fb[3];
    std::vector<double> fbuf[3];
    View fl[3];
    for (int d = 0; d < 3; ++d) {
        fb[d] = surrounding_nodes(bx, d);
        fbuf[d].assign((size_t)NCONS * (fb[d].hi[0]-fb[d].lo[0]+1) * (fb[d].hi[1]-fb[d].lo[1]+1) * (fb[d].hi[2]-fb[d].lo[2]+1), 0.0);
        fl[d] = View(fbuf[d].data(), fb[d].lo, fb[d].hi, NCONS);
        const View F = fl[d];
        for_box(fb[d], [&](int i, int j, int k) {
            double ql[NPRIM], qr[NPRIM], f[NCONS];
            for (int n = 0; n < NPRIM; ++n) eb_recon(d, i, j, k, n, q, g, ql[n], qr[n]);
            eb_flux_point(d, ql, qr, f, parm);
            for (int n = 0; n < NCONS; ++n) F(i, j, k, n) = f[n];
        });
        if (do_visc) {
            for_box(fb[d], [&](int i, int j, int k) {
                double f[NCONS];
                for (int n = 0; n < NCONS; ++n) f[n] = F(i, j, k, n);
                eb_visc(d, i, j, k, q, f, g, dxinv, parm);
                for (int n = 0; n < NCONS; ++n) F(i, j, k, n) = f[n];
            });
        }
    }
    // eb_compute_div(_visc), EBkernels.H:216-338, then flux_redist (:341-377)
    // or a copy (AdvanceBox.cpp:192-214)
    for_box(bx, [&](int i, int j, int k) {
        const uint32_t f = g.flag(i, j, k);
        double L[NCONS];
        double tmp = dxinv[0] / g.vf(i, j, k);
        if (eb_covered(f)) {
            for (int n = 0; n < NCONS; ++n) L[n] = 0.0;
        } else if (eb_regular(f)) {
            for (int n = 0; n < NCONS; ++n)
                L[n] = dxinv[0] * (fl[0](i + 1, j, k, n) - fl[0](i, j, k, n)) +
                       dxinv[1] * (fl[1](i, j + 1, k, n) - fl[1](i, j, k, n)) +
                       dxinv[2] * (fl[2](i, j, k + 1, n) - fl[2](i, j, k, n));
        } else {
            for (int n = 0; n < NCONS; ++n)
                L[n] = tmp * (g.ax(i + 1, j, k) * fl[0](i + 1, j, k, n) - g.ax(i, j, k) * fl[0](i, j, k, n) +
                              g.ay(i, j + 1, k) * fl[1](i, j + 1, k, n) - g.ay(i, j, k) * fl[1](i, j, k, n) +
                              g.az(i, j, k + 1) * fl[2](i, j, k + 1, n) - g.az(i, j, k) * fl[2](i, j, k, n));
            double fw[NCONS];
            if (do_visc) {
                double qw[NPRIM];
                for (int n = 0; n < NPRIM; ++n) qw[n] = q(i, j, k, n);
                eb_wallflux(qw, g.ax(i, j, k), g.ax(i + 1, j, k), g.ay(i, j, k), g.ay(i, j + 1, k), g.az(i, j, k),
                            g.az(i, j, k + 1), fw, parm);
            } else {
                fw[URHO] = 0.0;
                fw[UMX] = (g.ax(i, j, k) - g.ax(i + 1, j, k)) * q(i, j, k, QPRES);
                fw[UMY] = (g.ay(i, j, k) - g.ay(i, j + 1, k)) * q(i, j, k, QPRES);
                fw[UMZ] = (g.az(i, j, k) - g.az(i, j, k + 1)) * q(i, j, k, QPRES);
                fw[UEDEN] = 0.0;
            }
            for (int n = 0; n < NCONS; ++n) L[n] += fw[n] * tmp;
        }
        for (int n = 0; n < NCONS; ++n) L[n] *= -1.0;
        const double vf = g.vf(i, j, k);
        const bool scale = do_redist && vf < 0.5 && vf > 0.0;
        for (int n = 0; n < NCONS; ++n) dsdt(i, j, k, n) = scale ? vf * L[n] : L[n];
    });
    if (do_gravity) {
        const double gr = -9.8;
        for_box(bx, [&](int i, int j, int k) {
            dsdt(i, j, k, UMZ) += gr * s(i, j, k, URHO);
            dsdt(i, j, k, UEDEN) += gr * s(i, j, k, UMZ);
        });
    }
}

// Per-box dispatch of EBR::compute_dSdt (Source/Advance.cpp:207-231)
void compute_dsdt_any(const Box& bx, const View& s, const View& dsdt, const double* dxinv, const orc_ebfab* eb,
                      int do_visc, int do_gravity, int do_redist, const orc_parm& parm)
{
    if (!eb) { compute_dsdt_box(bx, s, dsdt, dxinv, do_visc, do_gravity, parm, nullptr); return; }
    EBV g(eb);
    const int kind = eb_box_kind(g, bx);
    if (kind == 2) for_box(bx, [&](int i, int j, int k) { for (int n = 0; n < NCONS; ++n) dsdt(i, j, k, n) = 0.0; });
    else if (kind == 0) compute_dsdt_box(bx, s, dsdt, dxinv, do_visc, do_gravity, parm, nullptr);
    else eb_compute_dsdt_box(bx, s, dsdt, dxinv, g, do_visc, do_gravity, do_redist, parm);
}

// EB fields of the level the box-level drivers below use (per box, or empty)
std::vector<orc_ebfab> g_eb;
int g_eb_redist = 0;

// ebr_estdt, Source/Advance.cpp:10-52 (non-CHEM)
double estdt_box(const Box& bx, const View& state, const double* dx, const orc_parm& parm)
{
    double dt = std::numeric_limits<double>::max();
#pragma omp parallel for collapse(2) reduction(min:dt)
    for (int k = bx.lo[2]; k <= bx.hi[2]; ++k)
        for (int j = bx.lo[1]; j <= bx.hi[1]; ++j)
            for (int i = bx.lo[0]; i <= bx.hi[0]; ++i) {
                double rho = state(i,j,k,URHO);
                double mx  = state(i,j,k,UMX);
                double my  = state(i,j,k,UMY);
                double mz  = state(i,j,k,UMZ);
                double E   = state(i,j,k,UEDEN);
                double rhoinv = 1.0/std::max(rho,parm.smallr);
                double vx = mx*rhoinv;
                double vy = my*rhoinv;
                double vz = mz*rhoinv;
                double p = std::max((parm.eos_gamma-1.0)*(E-0.5*rho*(vx*vx+vy*vy+vz*vz)), parm.smallp);
                double cs = std::sqrt(parm.eos_gamma*p*rhoinv);
                double dtx = dx[0]/(std::fabs(vx)+cs);
                double dty = dx[1]/(std::fabs(vy)+cs);
                double dtz = dx[2]/(std::fabs(vz)+cs);
                dt = std::min(dt,std::min(dtx,std::min(dty,dtz)));
            }
    return dt;
}

} // namespace

extern "C" {

// ---------------------------------------------------------------------------
// point-kernel entry points used by the pinning tests
// ---------------------------------------------------------------------------

// c2prim on n cells; u: n x 5 (row-major), q: n x 7
void orc_c2prim_many(long n, const double* u, double* q, const orc_parm* parm)
{
    for (long m = 0; m < n; ++m) {
        double uu[NCONS];
        for (int c = 0; c < NCONS; ++c) uu[c] = u[m * NCONS + c];
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        // a one-cell FAB per component stride 1
        double ucell[NCONS], qcell[NPRIM];
        for (int c = 0; c < NCONS; ++c) ucell[c] = uu[c];
        View uv(ucell, lo, hi, NCONS), qv(qcell, lo, hi, NPRIM);
        c2prim(0, 0, 0, uv, qv, *parm);
        for (int c = 0; c < NPRIM; ++c) q[m * NPRIM + c] = qcell[c];
    }
}

// SW split on n prim states; q: n x 7, fp/fm: n x 5
void orc_flux_split_many(int dir, long n, const double* q, double* fp, double* fm, const orc_parm* parm)
{
    for (long m = 0; m < n; ++m) sw_point(dir, q + m * NPRIM, fp + m * NCONS, fm + m * NCONS, parm->eos_gamma);
}

// WENO-Z5 on n stencils of 5 values: side 0 = f+ stencil (cells f-3..f+1),
// side 1 = f- stencil (cells f-2..f+2)
void orc_weno_many(int side, long n, const double* v, double* out)
{
    for (long m = 0; m < n; ++m) {
        const double* s = v + 5 * m;
        out[m] = side == 0 ? weno_plus(s[0], s[1], s[2], s[3], s[4]) : weno_minus(s[0], s[1], s[2], s[3], s[4]);
    }
}

// ---------------------------------------------------------------------------
// box-level entry points
// ---------------------------------------------------------------------------

// EBR::compute_dSdt regular path for one box (Advance.cpp:219-342).
// S covers grow(bx,4) (or more); dSdt covers bx (or more).  face[d] (optional)
// receives the face fluxes on surroundingNodes(bx,d), 5 comps.
void orc_compute_dsdt_box(const int* lo, const int* hi, const orc_fab* S, const orc_fab* dSdt,
                          const double* dxinv, int do_visc, int do_gravity, const orc_parm* parm,
                          double* fx, double* fy, double* fz)
{
    Box bx;
    for (int d = 0; d < 3; ++d) { bx.lo[d] = lo[d]; bx.hi[d] = hi[d]; }
    double* faces[3] = {fx, fy, fz};
    compute_dsdt_box(bx, View(*S), View(*dSdt), dxinv, do_visc, do_gravity, *parm,
                     (fx || fy || fz) ? faces : nullptr);
}

// EBR::compute_dSdt for one box with embedded-boundary fields (Advance.cpp:207-231
// dispatch: covered / regular / EB path); eb covers at least grow(bx,4).
void orc_compute_dsdt_box_eb(const int* lo, const int* hi, const orc_fab* S, const orc_fab* dSdt,
                             const double* dxinv, const orc_ebfab* eb, int do_visc, int do_gravity, int do_redist,
                             const orc_parm* parm)
{
    Box bx;
    for (int d = 0; d < 3; ++d) { bx.lo[d] = lo[d]; bx.hi[d] = hi[d]; }
    compute_dsdt_any(bx, View(*S), View(*dSdt), dxinv, eb, do_visc, do_gravity, do_redist, *parm);
}

// Van Leer face flux on n (left, right) state pairs; ql, qr: n x 7, F: n x 5
void orc_eb_flux_many(int dir, long n, const double* ql, const double* qr, double* F, const orc_parm* parm)
{
    for (long m = 0; m < n; ++m) eb_flux_point(dir, ql + 7 * m, qr + 7 * m, F + 5 * m, *parm);
}

// eb_recon (EBkernels.H:10-71) on every face of surroundingNodes(bx,d), for
// each component of q: ql, qr laid out (ncomp, faces) like a FAB over the
// face box.  eb covers the cells the stencils reach (grow(bx,2) suffices).
void orc_eb_recon_box(int d, const int* lo, const int* hi, const orc_fab* q, const orc_ebfab* eb, double* ql,
                      double* qr)
{
    Box bx;
    for (int e = 0; e < 3; ++e) { bx.lo[e] = lo[e]; bx.hi[e] = hi[e]; }
    const Box fb = surrounding_nodes(bx, d);
    const View qv(*q), L(ql, fb.lo, fb.hi, q->ncomp), R(qr, fb.lo, fb.hi, q->ncomp);
    const EBV g(eb);
    for_box(fb, [&](int i, int j, int k) {
        for (int n = 0; n < q->ncomp; ++n) eb_recon(d, i, j, k, n, qv, g, L(i, j, k, n), R(i, j, k, n));
    });
}

// eb_compute_visc (EBdiffusion.H:8-203) alone on every face of
// surroundingNodes(bx,d): F (5 comps, zeroed first) receives the viscous and
// heat-conduction terms.  q: primitive FAB (7 comps: rho,u,v,w,p,T,c).
void orc_eb_visc_box(int d, const int* lo, const int* hi, const orc_fab* q, const orc_ebfab* eb, const double* dxinv,
                     const orc_parm* parm, double* F)
{
    Box bx;
    for (int e = 0; e < 3; ++e) { bx.lo[e] = lo[e]; bx.hi[e] = hi[e]; }
    const Box fb = surrounding_nodes(bx, d);
    const View qv(*q), Fv(F, fb.lo, fb.hi, NCONS);
    const EBV g(eb);
    for_box(fb, [&](int i, int j, int k) {
        double f[NCONS] = {0, 0, 0, 0, 0};
        eb_visc(d, i, j, k, qv, f, g, dxinv, *parm);
        for (int n = 0; n < NCONS; ++n) Fv(i, j, k, n) = f[n];
    });
}

// eb_wallflux (EBkernels.H:173-214) on n cells: qw n x 7 primitives, area
// n x 6 (apx lo, apx hi, apy lo, apy hi, apz lo, apz hi), out n x 5
void orc_eb_wallflux_many(long n, const double* qw, const double* area, double* out, const orc_parm* parm)
{
    for (long m = 0; m < n; ++m) {
        const double* a = area + 6 * m;
        eb_wallflux(qw + NPRIM * m, a[0], a[1], a[2], a[3], a[4], a[5], out + NCONS * m, *parm);
    }
}

// Box kind as EBR::compute_dSdt decides it: 0 regular, 1 EB path, 2 covered.
int orc_eb_box_kind(const int* lo, const int* hi, const orc_ebfab* eb)
{
    Box bx;
    for (int d = 0; d < 3; ++d) { bx.lo[d] = lo[d]; bx.hi[d] = hi[d]; }
    return eb_box_kind(EBV(eb), bx);
}

// Set (nbox > 0) or clear (nbox == 0) the EB fields the level drivers
// (orc_flow_advance, orc_run) use for box b: eb[b].
void orc_eb_set(int nbox, const orc_ebfab* eb, int do_redist)
{
    g_eb.assign(eb, eb + nbox);
    g_eb_redist = do_redist;
}

// EBR::estTimeStep (Advance.cpp:61-79): cfl * min over boxes of ebr_estdt.
double orc_estdt(int nbox, const int* lo, const int* hi, const orc_fab* S, const double* dx,
                 double cfl, const orc_parm* parm)
{
    double dt = std::numeric_limits<double>::max();
    for (int b = 0; b < nbox; ++b) {
        Box bx;
        for (int d = 0; d < 3; ++d) { bx.lo[d] = lo[3*b+d]; bx.hi[d] = hi[3*b+d]; }
        dt = std::min(dt, estdt_box(bx, View(S[b]), dx, *parm));
    }
    return dt * cfl;
}

} // extern "C"

// ---------------------------------------------------------------------------
// Ghost fill (FillPatch on one level), RK drivers and case initial data.
// ---------------------------------------------------------------------------
extern "C" {

// Physical BC codes of ebr.lo_bc / ebr.hi_bc (Exec/*/inputs:10-14):
// 0 Interior, 1 Inflow, 2 Outflow, 3 Symmetry, 4 SlipWall, 5 NoSlipWall.
typedef struct orc_geom {
    int dom_lo[3], dom_hi[3];
    double prob_lo[3];
    double dx[3];
    int periodic[3];
    int lo_bc[3], hi_bc[3];
    int probtype;          // 0: no user fill, 1: double Mach reflection inflow fill
    double left_state[5];  // DMR post-shock state (Exec/1_DMR/prob.cpp:28-40)
    double right_state[5]; // DMR pre-shock state
} orc_geom;

}

namespace {

enum BCType { INT_DIR = 0, EXT_DIR = 3, FOEXTRAP = 2, REFLECT_EVEN = 1, REFLECT_ODD = -1 };

// Source/Setup.cpp:21-34
const int scalar_bc[6]   = {INT_DIR, EXT_DIR, FOEXTRAP, REFLECT_EVEN, REFLECT_EVEN, REFLECT_EVEN};
const int norm_vel_bc[6] = {INT_DIR, EXT_DIR, FOEXTRAP, REFLECT_ODD,  REFLECT_ODD,  REFLECT_ODD};
const int tang_vel_bc[6] = {INT_DIR, EXT_DIR, FOEXTRAP, REFLECT_EVEN, REFLECT_EVEN, REFLECT_ODD};

// BCType of component n on side (d, hi) -- Setup.cpp:36-92 and :201-205
inline int comp_bc(const orc_geom& g, int n, int d, int hi)
{
    int phys = hi ? g.hi_bc[d] : g.lo_bc[d];
    if (n == URHO || n == UEDEN) return scalar_bc[phys];
    int veldir = n - UMX;
    return veldir == d ? norm_vel_bc[phys] : tang_vel_bc[phys];
}

struct Level {
    int nbox;
    std::vector<Box> boxes;
    std::vector<View> fabs;
    int find(int i, int j, int k, int hint) const {
        if (hint >= 0) {
            const Box& b = boxes[hint];
            if (i >= b.lo[0] && i <= b.hi[0] && j >= b.lo[1] && j <= b.hi[1] && k >= b.lo[2] && k <= b.hi[2]) return hint;
        }
        for (int m = 0; m < nbox; ++m) {
            const Box& b = boxes[m];
            if (i >= b.lo[0] && i <= b.hi[0] && j >= b.lo[1] && j <= b.hi[1] && k >= b.lo[2] && k <= b.hi[2]) return m;
        }
        return -1;
    }
};

// Fill the ghost cells of every FAB in lev from the valid cells of the level:
// periodic images, then per-direction BC maps (AMReX FilccCell semantics:
// foextrap copies the boundary cell, reflect_even/odd mirror about the face
// with sign +1/-1; a cell outside in several directions composes the maps),
// then the case's ext_dir fill (Exec/1_DMR/prob_bcfill.cpp:14-58).
// species: fill a 9-component species MultiFab; every component takes the
// scalar (density) BC, Source/Setup.cpp:241-247, and no ext_dir values exist.
// species written into ext_dir ghost cells of species MultiFabs (set with
// orc_set_inflow_species; the library's ocfd_level_set_inflow_species)
bool g_spec_in_set = false;
double g_spec_in[9];

int fill_level(const Level& lev, const orc_geom& g, double time, int species = 0)
{
    int bad = 0;
    for (int b = 0; b < lev.nbox; ++b) {
        const View& f = lev.fabs[b];
        const Box& vb = lev.boxes[b];
#pragma omp parallel for collapse(2) reduction(+:bad)
        for (int k = f.lo[2]; k <= f.hi[2]; ++k)
            for (int j = f.lo[1]; j <= f.hi[1]; ++j) {
                int hint = b;
                for (int i = f.lo[0]; i <= f.hi[0]; ++i) {
                    if (i >= vb.lo[0] && i <= vb.hi[0] && j >= vb.lo[1] && j <= vb.hi[1] && k >= vb.lo[2] && k <= vb.hi[2])
                        continue;
                    int iv[3] = {i, j, k};
                    // periodic wrap
                    for (int d = 0; d < 3; ++d) {
                        if (!g.periodic[d]) continue;
                        int len = g.dom_hi[d] - g.dom_lo[d] + 1;
                        while (iv[d] < g.dom_lo[d]) iv[d] += len;
                        while (iv[d] > g.dom_hi[d]) iv[d] -= len;
                    }
                    int outside[3];
                    bool any_out = false, ext = false;
                    for (int d = 0; d < 3; ++d) {
                        outside[d] = iv[d] < g.dom_lo[d] ? -1 : (iv[d] > g.dom_hi[d] ? 1 : 0);
                        if (outside[d]) any_out = true;
                    }
                    const int ncomp = species ? f.nc : NCONS;
                    for (int n = 0; n < ncomp; ++n) {
                        int src[3] = {iv[0], iv[1], iv[2]};
                        double sign = 1.0;
                        bool user = false;
                        for (int d = 0; d < 3; ++d) {
                            if (!outside[d]) continue;
                            int bc = comp_bc(g, species ? 0 : n, d, outside[d] > 0);
                            int edge = outside[d] < 0 ? g.dom_lo[d] : g.dom_hi[d];
                            if (bc == FOEXTRAP) src[d] = edge;
                            else if (bc == REFLECT_EVEN || bc == REFLECT_ODD) {
                                src[d] = outside[d] < 0 ? 2 * edge - src[d] - 1 : 2 * edge - src[d] + 1;
                                if (bc == REFLECT_ODD) sign = -sign;
                            } else if (bc == EXT_DIR) user = true;
                        }
                        if (user) { ext = true; continue; }
                        int m = lev.find(src[0], src[1], src[2], hint);
                        if (m < 0) { bad += 1; continue; }
                        hint = m;
                        f(i, j, k, n) = sign * lev.fabs[m](src[0], src[1], src[2], n);
                    }
                    (void)any_out;
                    if (species) {
                        // ext_dir sides of a species MultiFab: the inflow
                        // species when set, else they keep their values
                        if (ext && g_spec_in_set)
                            for (int n = 0; n < f.nc; ++n) f(i, j, k, n) = g_spec_in[n];
                    } else if (g.probtype == 1) {
                        // Exec/1_DMR/prob_bcfill.cpp:27-57; bc of component 0 decides
                        double x = g.prob_lo[0] + (i + 0.5) * g.dx[0];
                        double x0 = 1.0/6.0 + std::sqrt(1/3.0) + time*10.0/std::sin(M_PI*(1.0/3.0));
                        int bc_hi1 = comp_bc(g, 0, 1, 1), bc_lo1 = comp_bc(g, 0, 1, 0), bc_lo0 = comp_bc(g, 0, 0, 0);
                        if (bc_hi1 == EXT_DIR && j > g.dom_hi[1]) {
                            for (int n = 0; n < NCONS; ++n) f(i, j, k, n) = x < x0 ? g.left_state[n] : g.right_state[n];
                        }
                        if (bc_lo1 == REFLECT_EVEN && j < g.dom_lo[1]) {
                            if (x < 1.0/6.0)
                                for (int n = 0; n < NCONS; ++n) f(i, j, k, n) = g.left_state[n];
                        }
                        if (bc_lo0 == EXT_DIR && i < g.dom_lo[0]) {
                            for (int n = 0; n < NCONS; ++n) f(i, j, k, n) = g.left_state[n];
                        }
                    } else if (g.probtype == 2) {
                        // inflow at x-lo, Exec/3_FSTEP/prob_bcfill.cpp:8-33
                        if (comp_bc(g, 0, 0, 0) == EXT_DIR && i < g.dom_lo[0]) {
                            for (int n = 0; n < NCONS; ++n) f(i, j, k, n) = g.left_state[n];
                        } else if (ext) {
                            bad += 1;
                        }
                    } else if (ext) {
                        bad += 1;
                    }
                }
            }
    }
    return bad;
}

Level make_level(int nbox, const int* lo, const int* hi, const orc_fab* fabs)
{
    Level L;
    L.nbox = nbox;
    for (int b = 0; b < nbox; ++b) {
        Box bx;
        for (int d = 0; d < 3; ++d) { bx.lo[d] = lo[3*b+d]; bx.hi[d] = hi[3*b+d]; }
        L.boxes.push_back(bx);
        L.fabs.push_back(View(fabs[b]));
    }
    return L;
}

} // namespace

extern "C" {

int orc_fill_ghost(int nbox, const int* lo, const int* hi, const orc_fab* fabs, const orc_geom* g, double time)
{
    Level L = make_level(nbox, lo, hi, fabs);
    return fill_level(L, *g, time);
}

// One time step of EBR::flow_advance on a single level (Advance.cpp:81-150).
// scheme 0: Euler (:117-122); 1: RK2 Heun (:123-139); 2: SSP-RK3 in Shu-Osher
// form U1=Un+dtL(Un), U2=3/4Un+1/4U1+dt/4 L(U1), Un+1=1/3Un+2/3U2+2dt/3 L(U2)
// with stage times t, t+dt, t+dt/2 (AMReX RK(3,...), Advance.cpp:141-148).
// A: U^n in, U^(n+1) out (ghosts filled on return for the next step's use);
// B, C: scratch FABs of the same boxes.  All FABs hold 4 ghost cells.
int orc_flow_advance(int nbox, const int* lo, const int* hi, const orc_fab* A, const orc_fab* B, const orc_fab* C,
                     const orc_geom* g, const orc_parm* parm, double time, double dt, int scheme,
                     int do_visc, int do_gravity)
{
    Level LA = make_level(nbox, lo, hi, A), LB = make_level(nbox, lo, hi, B), LC = make_level(nbox, lo, hi, C);
    double dxinv[3] = {1.0 / g->dx[0], 1.0 / g->dx[1], 1.0 / g->dx[2]};
    int bad = 0;
    std::vector<std::vector<double>> dsdt(nbox);
    std::vector<View> dv(nbox);
    for (int b = 0; b < nbox; ++b) {
        const Box& bx = LA.boxes[b];
        dsdt[b].assign((size_t)NCONS * (bx.hi[0]-bx.lo[0]+1) * (bx.hi[1]-bx.lo[1]+1) * (bx.hi[2]-bx.lo[2]+1), 0.0);
        dv[b] = View(dsdt[b].data(), bx.lo, bx.hi, NCONS);
    }
    // stage: out = (c_cur*cur + c_old*old) + c_dt*L(cur)  (LinComb / Saxpy forms)
    auto stage = [&](const Level& cur, const Level* old, const Level& out, double c_cur, double c_old, double c_dt) {
        for (int b = 0; b < nbox; ++b)
            compute_dsdt_any(cur.boxes[b], cur.fabs[b], dv[b], dxinv, g_eb.empty() ? nullptr : &g_eb[b], do_visc,
                             do_gravity, g_eb_redist, *parm);
        for (int b = 0; b < nbox; ++b) {
            const View& u = cur.fabs[b];
            const View& o = out.fabs[b];
            const View* uo = old ? &old->fabs[b] : nullptr;
            const View& L = dv[b];
            for_box(cur.boxes[b], [&](int i, int j, int k) {
                for (int n = 0; n < NCONS; ++n) {
                    double v = uo ? c_cur * u(i, j, k, n) + c_old * (*uo)(i, j, k, n) : c_cur * u(i, j, k, n);
                    o(i, j, k, n) = v + c_dt * L(i, j, k, n);
                }
            });
        }
    };
    bad += fill_level(LA, *g, time);
    if (scheme == 0) {
        stage(LA, nullptr, LB, 1.0, 0.0, dt);
        bad += fill_level(LB, *g, time + dt);
        for (int b = 0; b < nbox; ++b) std::memcpy(LA.fabs[b].p, LB.fabs[b].p, sizeof(double) * LA.fabs[b].npts() * NCONS);
    } else if (scheme == 1) {
        stage(LA, nullptr, LB, 1.0, 0.0, dt);
        bad += fill_level(LB, *g, time + dt);
        stage(LB, &LA, LC, 0.5, 0.5, 0.5 * dt);
        bad += fill_level(LC, *g, time + dt);
        for (int b = 0; b < nbox; ++b) std::memcpy(LA.fabs[b].p, LC.fabs[b].p, sizeof(double) * LA.fabs[b].npts() * NCONS);
    } else {
        stage(LA, nullptr, LB, 1.0, 0.0, dt);
        bad += fill_level(LB, *g, time + dt);
        stage(LB, &LA, LC, 0.25, 0.75, 0.25 * dt);
        bad += fill_level(LC, *g, time + 0.5 * dt);
        stage(LC, &LA, LB, 2.0 / 3.0, 1.0 / 3.0, (2.0 / 3.0) * dt);
        bad += fill_level(LB, *g, time + dt);
        for (int b = 0; b < nbox; ++b) std::memcpy(LA.fabs[b].p, LB.fabs[b].p, sizeof(double) * LA.fabs[b].npts() * NCONS);
    }
    return bad;
}

// Single-level time loop: initial dt = estTimeStep (EBR.cpp:561-563,189-229),
// then per step computeNewDt with change_max 1.1 and the stop_time clip
// (EBR.cpp:231-306).  Runs until time reaches stop_time or nsteps steps.
// Returns the number of steps taken; *time_io and *dt_io carry the time and
// the next dt.
int orc_run(int nbox, const int* lo, const int* hi, const orc_fab* A, const orc_fab* B, const orc_fab* C,
            const orc_geom* g, const orc_parm* parm, double cfl, int scheme, int do_visc, int do_gravity,
            double stop_time, int nsteps, double* time_io, double* dt_io)
{
    Level LA = make_level(nbox, lo, hi, A);
    double time = *time_io;
    double dt = *dt_io;
    if (dt <= 0.0) {
        dt = orc_estdt(nbox, lo, hi, A, g->dx, cfl, parm);
        const double eps = 0.001 * dt;
        if (stop_time >= 0.0 && (time + dt) > (stop_time - eps)) dt = stop_time - time;
    }
    int steps = 0;
    while ((nsteps < 0 || steps < nsteps) && (stop_time < 0.0 || time < stop_time)) {
        if (orc_flow_advance(nbox, lo, hi, A, B, C, g, parm, time, dt, scheme, do_visc, do_gravity) != 0)
            return -1;   // a ghost cell had no source (domain thinner than the ghost width, or uncovered)
        time += dt;
        ++steps;
        double est = orc_estdt(nbox, lo, hi, A, g->dx, cfl, parm);
        double dt_new = std::min(est, 1.1 * dt);
        const double eps = 0.001 * dt_new;
        if (stop_time >= 0.0 && (time + dt_new) > (stop_time - eps)) dt_new = stop_time - time;
        dt = dt_new;
    }
    *time_io = time;
    *dt_io = dt;
    return steps;
}

// ---------------------------------------------------------------------------
// Case initial data (valid cells only)
// ---------------------------------------------------------------------------

// Exec/0_SOD/prob.H:11-38 with Exec/0_SOD/prob_parm.H:9-12 defaults
void orc_init_sod(const int* lo, const int* hi, const orc_fab* S, const orc_geom* g, const orc_parm* parm)
{
    View s(*S);
    Box bx;
    for (int d = 0; d < 3; ++d) { bx.lo[d] = lo[d]; bx.hi[d] = hi[d]; }
    const double p_l = 1.0, p_r = 0.1, rho_l = 1.0, rho_r = 0.125;
    for_box(bx, [&](int i, int j, int k) {
        double x = g->prob_lo[0] + (i + 0.5) * g->dx[0];
        double Pt, rhot;
        if (x < 0.5) { Pt = p_l; rhot = rho_l; } else { Pt = p_r; rhot = rho_r; }
        s(i, j, k, URHO) = rhot;
        s(i, j, k, UMX) = 0.0;
        s(i, j, k, UMY) = 0.0;
        s(i, j, k, UMZ) = 0.0;
        s(i, j, k, UEDEN) = Pt / (parm->eos_gamma - 1.0);
    });
}

// Exec/1_DMR/prob.H:11-45 with Exec/1_DMR/prob_parm.H:10-17 defaults
void orc_init_dmr(const int* lo, const int* hi, const orc_fab* S, const orc_geom* g, const orc_parm* parm)
{
    View s(*S);
    Box bx;
    for (int d = 0; d < 3; ++d) { bx.lo[d] = lo[d]; bx.hi[d] = hi[d]; }
    const double p_l = 116.5, p_r = 1.0, rho_l = 8.0, rho_r = 1.4;
    const double u_l = 8.25 * std::cos(M_PI * (1.0 / 6.0)), u_r = 0.0;
    const double v_l = -8.25 * std::sin(M_PI * (1.0 / 6.0)), v_r = 0.0;
    for_box(bx, [&](int i, int j, int k) {
        double x = g->prob_lo[0] + (i + 0.5) * g->dx[0];
        double y = g->prob_lo[1] + (j + 0.5) * g->dx[1];
        double x0 = 1.0 / 6.0;
        double Pt, rhot, uxt, vxt;
        if (x < x0 + std::sqrt(1 / 3.0) * y) { Pt = p_l; rhot = rho_l; uxt = u_l; vxt = v_l; }
        else { Pt = p_r; rhot = rho_r; uxt = u_r; vxt = v_r; }
        s(i, j, k, URHO) = rhot;
        s(i, j, k, UMX) = rhot * uxt;
        s(i, j, k, UMY) = rhot * vxt;
        s(i, j, k, UMZ) = 0.0;
        double et = Pt / (parm->eos_gamma - 1.0);
        s(i, j, k, UEDEN) = et + 0.5 * rhot * (uxt * uxt + vxt * vxt);
    });
}

// Exec/3_FSTEP/prob.H:11-40 with Exec/3_FSTEP/prob_parm.H defaults; flag holds the
// EBCellFlag words of the valid box (i fastest)
void orc_init_eb_inflow(const int* lo, const int* hi, const orc_fab* S, const uint32_t* flag, const orc_parm* parm)
{
    View s(*S);
    const double Pt = 2026.5, rhot = 0.08938, uxt = 534.49;
    const long nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1;
    for (int k = lo[2]; k <= hi[2]; ++k)
        for (int j = lo[1]; j <= hi[1]; ++j)
            for (int i = lo[0]; i <= hi[0]; ++i) {
                const bool cov = (flag[((long)(k - lo[2]) * ny + (j - lo[1])) * nx + (i - lo[0])] & 3u) == 3u;
                if (cov) {
                    s(i, j, k, URHO) = rhot;
                    s(i, j, k, UMX) = 0.0;
                    s(i, j, k, UMY) = 0.0;
                    s(i, j, k, UMZ) = 0.0;
                    double et = Pt / (parm->eos_gamma - 1.0);
                    s(i, j, k, UEDEN) = et + 0.5 * rhot * uxt * uxt;
                } else {
                    s(i, j, k, URHO) = rhot;
                    s(i, j, k, UMX) = rhot * uxt;
                    s(i, j, k, UMY) = 0.0;
                    s(i, j, k, UMZ) = 0.0;
                    double et = Pt / (parm->eos_gamma - 1.0);
                    s(i, j, k, UEDEN) = et + 0.5 * rhot * uxt * uxt;
                }
            }
}

// Exec/3_FSTEP/prob.cpp:24-29 (literal 0.4)
void orc_inflow_state(double* st)
{
    const double p = 2026.5, rho = 0.08938, u = 534.49;
    st[0] = rho;
    st[1] = u * rho;
    st[2] = 0;
    st[3] = 0;
    st[4] = p / 0.4 + 0.5 * rho * u * u;
}

// DMR inflow states, Exec/1_DMR/prob.cpp:28-40 (note the literal 0.4)
void orc_dmr_states(double* left, double* right)
{
    const double p_l = 116.5, p_r = 1.0, rho_l = 8.0, rho_r = 1.4;
    const double u_l = 8.25 * std::cos(M_PI * (1.0 / 6.0)), u_r = 0.0;
    const double v_l = -8.25 * std::sin(M_PI * (1.0 / 6.0)), v_r = 0.0;
    left[0] = rho_l; left[1] = u_l * rho_l; left[2] = v_l * rho_l; left[3] = 0;
    left[4] = p_l / 0.4 + 0.5 * rho_l * (u_l * u_l + v_l * v_l);
    right[0] = rho_r; right[1] = u_r * rho_r; right[2] = v_r * rho_r; right[3] = 0;
    right[4] = p_r / 0.4 + 0.5 * rho_r * (u_r * u_r + v_r * v_r);
}

// Counter-based uniform variate in [0,1): splitmix64 finaliser of
// (seed * golden-ratio constant + counter), top 53 bits.
double orc_rng_uniform(uint64_t seed, uint64_t counter)
{
    uint64_t z = seed * 0x9E3779B97F4A7C15ull + counter;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// 3D viscous Kelvin-Helmholtz shear layer on a periodic unit cube (the case
// of the 3D KH workload; the reference's Exec/8_KH is a 2D inviscid variant):
// air at p0 = 101325 Pa, outer stream at 300 K, density ratio 2 inside the
// band 1/4 < y < 3/4, tanh profiles of thickness delta = 1/32, streams at
// +-Mach 0.5 of the outer sound speed, seeded uniform velocity perturbation of
// amplitude 1e-2 of the stream speed.
void orc_init_kh3d(const int* lo, const int* hi, const orc_fab* S, const orc_geom* g, const orc_parm* parm,
                   uint64_t seed)
{
    View s(*S);
    Box bx;
    for (int d = 0; d < 3; ++d) { bx.lo[d] = lo[d]; bx.hi[d] = hi[d]; }
    const double gamma = parm->eos_gamma;
    const double p0 = 101325.0;
    const double rho_out = p0 / (parm->Rg * 300.0);
    const double rho_in = 2.0 * rho_out;
    const double c_out = std::sqrt(gamma * p0 / rho_out);
    const double U = 0.5 * c_out;
    const double delta = 1.0 / 32.0;
    const double amp = 1.0e-2 * U;
    const long nx = g->dom_hi[0] - g->dom_lo[0] + 1, ny = g->dom_hi[1] - g->dom_lo[1] + 1;
    for_box(bx, [&](int i, int j, int k) {
        double y = g->prob_lo[1] + (j + 0.5) * g->dx[1];
        double sh = 0.5 * (std::tanh((y - 0.25) / delta) - std::tanh((y - 0.75) / delta));
        double rho = rho_out + (rho_in - rho_out) * sh;
        uint64_t cell = ((uint64_t)(k - g->dom_lo[2]) * ny + (j - g->dom_lo[1])) * nx + (i - g->dom_lo[0]);
        double u = U * (2.0 * sh - 1.0) + amp * (2.0 * orc_rng_uniform(seed, 3 * cell + 0) - 1.0);
        double v = amp * (2.0 * orc_rng_uniform(seed, 3 * cell + 1) - 1.0);
        double w = amp * (2.0 * orc_rng_uniform(seed, 3 * cell + 2) - 1.0);
        s(i, j, k, URHO) = rho;
        s(i, j, k, UMX) = rho * u;
        s(i, j, k, UMY) = rho * v;
        s(i, j, k, UMZ) = rho * w;
        s(i, j, k, UEDEN) = p0 / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w);
    });
}

} // extern "C"

// ---------------------------------------------------------------------------
// Multispecies (CHEM) flow path: EBR::compute_dSdt_multi on regular boxes
// (Source/CHEM/FlowAdvance.cpp:92-496), EBR::flow_advance_multi (:9-90) and
// the Strang-split EBR::advance (Source/Advance.cpp:152-177).
// ---------------------------------------------------------------------------
extern "C" {
double orc_T_given_ey(double e, const double* rhoi, int* iters);
double orc_CKPY(const double* rhoi, double T);
void orc_CKYTX(const double* y, double* x);
void orc_species_enthalpy(double T, double* h);
void orc_mixture_properties(double T, const double* X, double* mu, double* lambda);
void orc_mix_diff_coeffs_mass(double T, double p, const double* X, double* D);
void orc_chem_step_many(long n, double* u, double* rhoi, double dt, double smallr);
void orc_renorm_many(long n, const double* u, double* rhoi);
}

namespace {

constexpr int NSP = 9;
constexpr int QGAMA = 7, NPRIM_M = 8;   // IndexDefines.H:19-21 (CHEM)

// c2prim_rgas, Source/kernel/Kernels.H:42-81
inline void c2prim_rgas(int i, int j, int k, const View& u, const View& spec, const View& q, const orc_parm& parm)
{
    double rho = std::max(u(i, j, k, URHO), parm.smallr);
    double rhoinv = 1.0 / rho;
    double ux = u(i, j, k, UMX) * rhoinv;
    double uy = u(i, j, k, UMY) * rhoinv;
    double uz = u(i, j, k, UMZ) * rhoinv;
    double kineng = 0.5 * (ux * ux + uy * uy + uz * uz);
    double ei = u(i, j, k, UEDEN) - rho * kineng;
    double rhoi[NSP];
    q(i, j, k, QRHO) = rho;
    q(i, j, k, QU) = ux;
    q(i, j, k, QV) = uy;
    q(i, j, k, QW) = uz;
    for (int n = 0; n < NSP; ++n) rhoi[n] = spec(i, j, k, n);
    double T = orc_T_given_ey(ei, rhoi, nullptr);
    double p = orc_CKPY(rhoi, T);
    q(i, j, k, QT) = T;
    q(i, j, k, QPRES) = p;
    double gama = p / ei + 1.0;
    q(i, j, k, QGAMA) = gama;
    q(i, j, k, QC) = std::sqrt(gama * p / rho);
}

// Steger-Warming split of the partial densities, FlowAdvance.cpp:214-242
inline void sw_species(int d, const View& q, const View& spec, int i, int j, int k, double* fp, double* fm)
{
    double un = q(i, j, k, QU + d);
    double c = q(i, j, k, QC);
    double gamma = q(i, j, k, QGAMA);
    double E1 = un, E2 = un - c, E3 = un + c;
    double E1P = (E1 + std::fabs(E1)) * 0.5;
    double E2P = (E2 + std::fabs(E2)) * 0.5;
    double E3P = (E3 + std::fabs(E3)) * 0.5;
    double E1M = E1 - E1P, E2M = E2 - E2P, E3M = E3 - E3P;
    double tmp1 = 1.0 / (2.0 * gamma);
    double tmp2 = 2.0 * (gamma - 1.0);
    double tmp2_p = tmp2 * E1P + E2P + E3P;
    double tmp2_m = tmp2 * E1M + E2M + E3M;
    for (int n = 0; n < NSP; ++n) {
        double tmp0 = spec(i, j, k, n) * tmp1;
        fp[n] = tmp0 * tmp2_p;
        fm[n] = tmp0 * tmp2_m;
    }
}

// compute_dSdt_multi on one regular box (FlowAdvance.cpp:173-403)
void compute_dsdt_multi_box(const Box& bx, const View& s, const View& spec, const View& dsdt, const View& dsdt_spec,
                            const double* dxinv, int do_visc, const orc_parm& parm, double* const* face_out = nullptr)
{
    const Box bxg = grow(bx, NUM_GROW);
    const size_t ng = (size_t)(bxg.hi[0]-bxg.lo[0]+1) * (bxg.hi[1]-bxg.lo[1]+1) * (bxg.hi[2]-bxg.lo[2]+1);
    std::vector<double> qb(NPRIM_M * ng), mub(ng), lamb(ng), Db(NSP * ng);
    std::vector<double> fpb(NCONS * ng), fmb(NCONS * ng), fpsb(NSP * ng), fmsb(NSP * ng);
    View q(qb.data(), bxg.lo, bxg.hi, NPRIM_M), mu(mub.data(), bxg.lo, bxg.hi, 1), lam(lamb.data(), bxg.lo, bxg.hi, 1);
    View D(Db.data(), bxg.lo, bxg.hi, NSP);
    View fp(fpb.data(), bxg.lo, bxg.hi, NCONS), fm(fmb.data(), bxg.lo, bxg.hi, NCONS);
    View fps(fpsb.data(), bxg.lo, bxg.hi, NSP), fms(fmsb.data(), bxg.lo, bxg.hi, NSP);

    for_box(bxg, [&](int i, int j, int k) { c2prim_rgas(i, j, k, s, spec, q, parm); });
    if (do_visc) {
        // FlowAdvance.cpp:188-202
        for_box(bxg, [&](int i, int j, int k) {
            double Yt[NSP], Xt[NSP], Dk[NSP];
            double T = q(i, j, k, QT);
            double p = q(i, j, k, QPRES);
            for (int n = 0; n < NSP; ++n) Yt[n] = spec(i, j, k, n) / q(i, j, k, QRHO);
            orc_CKYTX(Yt, Xt);
            orc_mixture_properties(T, Xt, &mu(i, j, k, 0), &lam(i, j, k, 0));
            orc_mix_diff_coeffs_mass(T, p, Xt, Dk);
            for (int n = 0; n < NSP; ++n) D(i, j, k, n) = Dk[n];
        });
    }
    Box fb[3];
    std::vector<double> fbuf[3], fsbuf[3], fdbuf[3];
    View F[3], Fs[3], Fd[3];
    for (int d = 0; d < 3; ++d) {
        fb[d] = surrounding_nodes(bx, d);
        const size_t nf = (size_t)(fb[d].hi[0]-fb[d].lo[0]+1) * (fb[d].hi[1]-fb[d].lo[1]+1) * (fb[d].hi[2]-fb[d].lo[2]+1);
        fbuf[d].assign(NCONS * nf, 0.0); fsbuf[d].assign(NSP * nf, 0.0); fdbuf[d].assign(NSP * nf, 0.0);
        F[d] = View(fbuf[d].data(), fb[d].lo, fb[d].hi, NCONS);
        Fs[d] = View(fsbuf[d].data(), fb[d].lo, fb[d].hi, NSP);
        Fd[d] = View(fdbuf[d].data(), fb[d].lo, fb[d].hi, NSP);
        const int di = d == 0, dj = d == 1, dk = d == 2;
        for_box(bxg, [&](int i, int j, int k) {
            double qv[NPRIM], a[NCONS], b[NCONS], as[NSP], bs[NSP];
            for (int n = 0; n < NPRIM; ++n) qv[n] = q(i, j, k, n);
            sw_point(d, qv, a, b, q(i, j, k, QGAMA));
            for (int n = 0; n < NCONS; ++n) { fp(i, j, k, n) = a[n]; fm(i, j, k, n) = b[n]; }
            sw_species(d, q, spec, i, j, k, as, bs);
            for (int n = 0; n < NSP; ++n) { fps(i, j, k, n) = as[n]; fms(i, j, k, n) = bs[n]; }
        });
        auto weno = [&](const View& P, const View& M, const View& out, int ncomp) {
            for_box(fb[d], [&](int i, int j, int k) {
                for (int n = 0; n < ncomp; ++n) {
                    double pv = weno_plus(P(i-3*di, j-3*dj, k-3*dk, n), P(i-2*di, j-2*dj, k-2*dk, n),
                                          P(i-di, j-dj, k-dk, n), P(i, j, k, n), P(i+di, j+dj, k+dk, n));
                    double mv = weno_minus(M(i-2*di, j-2*dj, k-2*dk, n), M(i-di, j-dj, k-dk, n),
                                           M(i, j, k, n), M(i+di, j+dj, k+dk, n), M(i+2*di, j+2*dj, k+2*dk, n));
                    out(i, j, k, n) = pv + mv;
                }
            });
        };
        weno(fp, fm, F[d], NCONS);
        weno(fps, fms, Fs[d], NSP);
        if (do_visc) {
            const View Fv = F[d], Fsv = Fs[d], Fdv = Fd[d];
            // diffusion_{x,y,z}, Source/CHEM/ChemViscous.H:108-199
            for_box(fb[d], [&](int i, int j, int k) {
                const int im = i - di, jm = j - dj, km = k - dk;
                double rhor = q(i, j, k, QRHO);
                double rhol = q(im, jm, km, QRHO);
                double rho = 0.5 * (rhol + rhor);
                double dn[NSP], Ytr[NSP], Ytl[NSP];
                for (int n = 0; n < NSP; ++n) {
                    Ytr[n] = spec(i, j, k, n) / rhor;
                    Ytl[n] = spec(im, jm, km, n) / rhol;
                    dn[n] = 0.5 * (D(i, j, k, n) + D(im, jm, km, n));
                }
                for (int n = 0; n < NSP; ++n) {
                    double tmp = rho * dn[n] * (Ytl[n] - Ytr[n]) * dxinv[d];
                    Fdv(i, j, k, n) = tmp;
                    Fsv(i, j, k, n) += tmp;
                }
            });
            // compute_visc_{x,y,z}_multi, Source/CHEM/ChemViscous.H:201-357
            for_box(fb[d], [&](int i, int j, int k) {
                const int e[3] = {di, dj, dk};
                const int im = i - di, jm = j - dj, km = k - dk;
                double Tr = q(i, j, k, QT);
                double Tl = q(im, jm, km, QT);
                double T = 0.5 * (Tl + Tr);
                double muf = 0.5 * (mu(i, j, k, 0) + mu(im, jm, km, 0));
                double lamf = 0.5 * (lam(i, j, k, 0) + lam(im, jm, km, 0));
                double hi[NSP];
                orc_species_enthalpy(T, hi);
                double tmp = 0;
                for (int n = 0; n < NSP; ++n) tmp += Fdv(i, j, k, n) * hi[n];
                double dTn = (Tr - Tl) * dxinv[d];
                double dn[3];
                for (int c = 0; c < 3; ++c) dn[c] = (q(i, j, k, QU + c) - q(im, jm, km, QU + c)) * dxinv[d];
                // along each tangential direction t: 4-point differences of u_d and u_t
                double g_d[3] = {0, 0, 0}, g_t[3] = {0, 0, 0};
                for (int t = 0; t < 3; ++t) {
                    if (t == d) continue;
                    const int ti = t == 0, tj = t == 1, tk = t == 2;
                    auto four = [&](int comp) {
                        return (q(i + ti, j + tj, k + tk, comp) + q(im + ti, jm + tj, km + tk, comp) -
                                q(i - ti, j - tj, k - tk, comp) - q(im - ti, jm - tj, km - tk, comp)) *
                               (0.25 * dxinv[t]);
                    };
                    g_d[t] = four(QU + d);
                    g_t[t] = four(QU + t);
                    (void)e;
                }
                double divu = 0;
                {
                    double c0 = d == 0 ? dn[0] : g_t[0];
                    double c1 = d == 1 ? dn[1] : g_t[1];
                    double c2 = d == 2 ? dn[2] : g_t[2];
                    divu = c0 + c1 + c2;
                }
                double tau[3];
                for (int t = 0; t < 3; ++t) {
                    if (t == d) tau[t] = muf * (2.0 * dn[d] - parm.twoThirds * divu);
                    else tau[t] = muf * (g_d[t] + dn[t]);
                }
                Fv(i, j, k, UMX) += -tau[0];
                Fv(i, j, k, UMY) += -tau[1];
                Fv(i, j, k, UMZ) += -tau[2];
                Fv(i, j, k, UEDEN) += -0.5 * ((q(i, j, k, QU) + q(im, jm, km, QU)) * tau[0] +
                                              (q(i, j, k, QV) + q(im, jm, km, QV)) * tau[1] +
                                              (q(i, j, k, QW) + q(im, jm, km, QW)) * tau[2]) - lamf * dTn - tmp;
            });
        }
    }
    // face_out (optional): F[0..2] (5 comps), Fs[0..2] (9), Fd[0..2] (9), each
    // on surroundingNodes(bx,d); null entries are skipped
    if (face_out)
        for (int d = 0; d < 3; ++d) {
            if (face_out[d]) std::memcpy(face_out[d], fbuf[d].data(), fbuf[d].size() * sizeof(double));
            if (face_out[3 + d]) std::memcpy(face_out[3 + d], fsbuf[d].data(), fsbuf[d].size() * sizeof(double));
            if (face_out[6 + d]) std::memcpy(face_out[6 + d], fdbuf[d].data(), fdbuf[d].size() * sizeof(double));
        }
    // divop for both, FlowAdvance.cpp:394-403
    for_box(bx, [&](int i, int j, int k) {
        for (int n = 0; n < NCONS; ++n)
            dsdt(i, j, k, n) = dxinv[0] * (F[0](i, j, k, n) - F[0](i+1, j, k, n))
                +              dxinv[1] * (F[1](i, j, k, n) - F[1](i, j+1, k, n))
                +              dxinv[2] * (F[2](i, j, k, n) - F[2](i, j, k+1, n));
        for (int n = 0; n < NSP; ++n)
            dsdt_spec(i, j, k, n) = dxinv[0] * (Fs[0](i, j, k, n) - Fs[0](i+1, j, k, n))
                +                   dxinv[1] * (Fs[1](i, j, k, n) - Fs[1](i, j+1, k, n))
                +                   dxinv[2] * (Fs[2](i, j, k, n) - Fs[2](i, j, k+1, n));
    });
}

// Fickian species flux of face (i,j,k) in direction d, Source/CHEM/ChemViscous.H:108-199
inline void diffusion_face(int d, int i, int j, int k, const View& q, const View& spec, const View& D,
                           const double* dxinv, double* J)
{
    const int im = i - (d == 0), jm = j - (d == 1), km = k - (d == 2);
    double rhor = q(i, j, k, QRHO);
    double rhol = q(im, jm, km, QRHO);
    double rho = 0.5 * (rhol + rhor);
    for (int n = 0; n < NSP; ++n) {
        double Ytr = spec(i, j, k, n) / rhor;
        double Ytl = spec(im, jm, km, n) / rhol;
        double dn = 0.5 * (D(i, j, k, n) + D(im, jm, km, n));
        J[n] = rho * dn * (Ytl - Ytr) * dxinv[d];
    }
}

// Van Leer mass-flux split of the partial densities at one face,
// Source/CHEM/FlowAdvanceBox.cpp:174-223: normal velocities ul/ur and sound
// speeds cL/cR of the reconstructed left/right states, partial densities
// rl/rr; F = f+(left) + f-(right) per species.
// The mass flux of Van Leer's splitting for a partial density r moving with
// normal velocity U at sound speed c: sgn = +1 the forward part, -1 the
// backward part (the whole flux r U when supersonic towards the side, the
// split r c sgn (1 + sgn M)^2 / 4 when subsonic, else nothing)
inline double vl_mass_part(double r, double U, double c, double sgn)
{
    const double M = U / c;
    if (sgn * M >= 1.0) return r * U;
    if (std::fabs(M) < 1.0) {
        const double w = 1.0 + sgn * M;
        return r * c * (sgn * 0.250 * w * w);
    }
    return 0.0;
}

inline void vl_species_flux(double ul, double ur, double cL, double cR, const double* rl, const double* rr, double* F)
{
    for (int n = 0; n < NSP; ++n) F[n] = vl_mass_part(rl[n], ul, cL, 1.0) + vl_mass_part(rr[n], ur, cR, -1.0);
}

// EBR::ebchem_compute_dSdt_box, Source/CHEM/FlowAdvanceBox.cpp:16-525: MUSCL
// left/right states of the normal velocity, sound speed and partial densities
// (eb_recon), the Van Leer mass-flux split for the species (:174-223), the
// Fickian flux (viscous), the EB divergence with the species wall term
// (:414-460).  The flow divergence is computed by the reference but never
// copied into dSdt (:501-524), so the flow dSdt of these boxes is 0.
void ebchem_dsdt_box(const Box& bx, const View& s, const View& spec, const View& dsdt, const View& dsdt_spec,
                     const double* dxinv, const EBV& g, int do_visc, const orc_parm& parm)
{
    const Box bxg = grow(bx, NUM_GROW);
    const size_t ng = (size_t)(bxg.hi[0]-bxg.lo[0]+1) * (bxg.hi[1]-bxg.lo[1]+1) * (bxg.hi[2]-bxg.lo[2]+1);
    std::vector<double> qb(NPRIM_M * ng), mub(ng), lamb(ng), Db(NSP * ng);
    View q(qb.data(), bxg.lo, bxg.hi, NPRIM_M), D(Db.data(), bxg.lo, bxg.hi, NSP);
    for_box(bxg, [&](int i, int j, int k) { c2prim_rgas(i, j, k, s, spec, q, parm); });
    if (do_visc) {
        for_box(bxg, [&](int i, int j, int k) {
            double Yt[NSP], Xt[NSP], Dk[NSP], mu, lam;
            for (int n = 0; n < NSP; ++n) Yt[n] = spec(i, j, k, n) / q(i, j, k, QRHO);
            orc_CKYTX(Yt, Xt);
            orc_mixture_properties(q(i, j, k, QT), Xt, &mu, &lam);
            orc_mix_diff_coeffs_mass(q(i, j, k, QT), q(i, j, k, QPRES), Xt, Dk);
            for (int n = 0; n < NSP; ++n) D(i, j, k, n) = Dk[n];
        });
    }
    Box fb[3];
    std::vector<double> fsbuf[3];
    View Fs[3];
    for (int d = 0; d < 3; ++d) {
        fb[d] = surrounding_nodes(bx, d);
        const size_t nf = (size_t)(fb[d].hi[0]-fb[d].lo[0]+1) * (fb[d].hi[1]-fb[d].lo[1]+1) * (fb[d].hi[2]-fb[d].lo[2]+1);
        fsbuf[d].assign(NSP * nf, 0.0);
        Fs[d] = View(fsbuf[d].data(), fb[d].lo, fb[d].hi, NSP);
        const View F = Fs[d];
        for_box(fb[d], [&](int i, int j, int k) {
            double ul, ur, cL, cR, rl[NSP], rr[NSP], fs[NSP];
            eb_recon(d, i, j, k, QU + d, q, g, ul, ur);
            eb_recon(d, i, j, k, QC, q, g, cL, cR);
            for (int n = 0; n < NSP; ++n) eb_recon(d, i, j, k, n, spec, g, rl[n], rr[n]);
            vl_species_flux(ul, ur, cL, cR, rl, rr, fs);
            for (int n = 0; n < NSP; ++n) F(i, j, k, n) = fs[n];
            if (do_visc) {
                double J[NSP];
                diffusion_face(d, i, j, k, q, spec, D, dxinv, J);
                for (int n = 0; n < NSP; ++n) F(i, j, k, n) += J[n];
            }
        });
    }
    for_box(bx, [&](int i, int j, int k) {
        for (int n = 0; n < NCONS; ++n) dsdt(i, j, k, n) = 0.0;
        const uint32_t f = g.flag(i, j, k);
        double tmp = dxinv[0] / g.vf(i, j, k);
        double L[NSP];
        if (eb_covered(f)) {
            for (int n = 0; n < NSP; ++n) L[n] = 0.0;
        } else if (eb_regular(f)) {
            for (int n = 0; n < NSP; ++n)
                L[n] = dxinv[0] * (Fs[0](i + 1, j, k, n) - Fs[0](i, j, k, n)) +
                       dxinv[1] * (Fs[1](i, j + 1, k, n) - Fs[1](i, j, k, n)) +
                       dxinv[2] * (Fs[2](i, j, k + 1, n) - Fs[2](i, j, k, n));
        } else {
            for (int n = 0; n < NSP; ++n)
                L[n] = tmp * (g.ax(i + 1, j, k) * Fs[0](i + 1, j, k, n) - g.ax(i, j, k) * Fs[0](i, j, k, n) +
                              g.ay(i, j + 1, k) * Fs[1](i, j + 1, k, n) - g.ay(i, j, k) * Fs[1](i, j, k, n) +
                              g.az(i, j, k + 1) * Fs[2](i, j, k + 1, n) - g.az(i, j, k) * Fs[2](i, j, k, n));
            if (do_visc) {
                double axm = g.ax(i, j, k), axp = g.ax(i + 1, j, k), aym = g.ay(i, j, k), ayp = g.ay(i, j + 1, k);
                double azm = g.az(i, j, k), azp = g.az(i, j, k + 1);
                double apnorm = std::sqrt((axm - axp) * (axm - axp) + (aym - ayp) * (aym - ayp) + (azm - azp) * (azm - azp));
                double un = q(i, j, k, QU) * (axm - axp) + q(i, j, k, QV) * (aym - ayp) + q(i, j, k, QW) * (azm - azp);
                double lambda0 = std::fabs(un) + q(i, j, k, QC) * apnorm;
                for (int n = 0; n < NSP; ++n) {
                    double fp = 0.5 * (spec(i, j, k, n) * un + lambda0 * spec(i, j, k, n));
                    double fm = -0.5 * (spec(i, j, k, n) * un + lambda0 * spec(i, j, k, n));
                    L[n] += (fp + fm) * tmp;
                }
            }
        }
        for (int n = 0; n < NSP; ++n) dsdt_spec(i, j, k, n) = -1.0 * L[n];
    });
}

// Per-box dispatch of EBR::compute_dSdt_multi (FlowAdvance.cpp:127-173)
void compute_dsdt_multi_any(const Box& bx, const View& s, const View& spec, const View& dsdt, const View& dsdt_spec,
                            const double* dxinv, const orc_ebfab* eb, int do_visc, const orc_parm& parm)
{
    if (!eb) { compute_dsdt_multi_box(bx, s, spec, dsdt, dsdt_spec, dxinv, do_visc, parm); return; }
    EBV g(eb);
    const int kind = eb_box_kind(g, bx);
    if (kind == 2) {
        for_box(bx, [&](int i, int j, int k) {
            for (int n = 0; n < NCONS; ++n) dsdt(i, j, k, n) = 0.0;
            for (int n = 0; n < NSP; ++n) dsdt_spec(i, j, k, n) = 0.0;
        });
    } else if (kind == 0) {
        compute_dsdt_multi_box(bx, s, spec, dsdt, dsdt_spec, dxinv, do_visc, parm);
    } else {
        ebchem_dsdt_box(bx, s, spec, dsdt, dsdt_spec, dxinv, g, do_visc, parm);
    }
}

// chemical_advance(dt) on the valid cells of every box (ChemAdvance.cpp:7-85)
void chem_level(const Level& S, const Level& Sp, double dt, const orc_parm& parm)
{
    for (int b = 0; b < S.nbox; ++b) {
        const Box& bx = S.boxes[b];
        if (!g_eb.empty() && eb_region_type(EBV(&g_eb[b]), bx) == 2) continue;   // ChemAdvance.cpp:20
        std::vector<double> u, r;
        for (int k = bx.lo[2]; k <= bx.hi[2]; ++k)
            for (int j = bx.lo[1]; j <= bx.hi[1]; ++j)
                for (int i = bx.lo[0]; i <= bx.hi[0]; ++i) {
                    for (int n = 0; n < NCONS; ++n) u.push_back(S.fabs[b](i, j, k, n));
                    for (int n = 0; n < NSP; ++n) r.push_back(Sp.fabs[b](i, j, k, n));
                }
        orc_chem_step_many((long)(u.size() / NCONS), u.data(), r.data(), dt, parm.smallr);
        size_t c = 0;
        for (int k = bx.lo[2]; k <= bx.hi[2]; ++k)
            for (int j = bx.lo[1]; j <= bx.hi[1]; ++j)
                for (int i = bx.lo[0]; i <= bx.hi[0]; ++i, ++c) {
                    S.fabs[b](i, j, k, UEDEN) = u[c * NCONS + UEDEN];
                    for (int n = 0; n < NSP; ++n) Sp.fabs[b](i, j, k, n) = r[c * NSP + n];
                }
    }
}

} // namespace

extern "C" {

// The face fluxes of compute_dSdt_multi on one regular box (test use): the
// nine outputs F_x,y,z (5 comps), species F_x,y,z (9) and Fickian J_x,y,z
// (9), each on surroundingNodes(bx,d) (null pointers skipped)
void orc_compute_dsdt_multi_box_faces(const int* lo, const int* hi, const orc_fab* S, const orc_fab* Spec,
                                      const double* dxinv, int do_visc, const orc_parm* parm, double* const* faces)
{
    Box bx;
    for (int d = 0; d < 3; ++d) { bx.lo[d] = lo[d]; bx.hi[d] = hi[d]; }
    std::vector<double> a(NCONS * (size_t)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1));
    std::vector<double> b(NSP * (size_t)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1));
    compute_dsdt_multi_box(bx, View(*S), View(*Spec), View(a.data(), lo, hi, NCONS), View(b.data(), lo, hi, NSP),
                           dxinv, do_visc, *parm, faces);
}

// Van Leer species mass-flux split on n faces; rl, rr, F: n x 9
void orc_vl_species_flux_many(long n, const double* ul, const double* ur, const double* cL, const double* cR,
                              const double* rl, const double* rr, double* F)
{
    for (long m = 0; m < n; ++m) vl_species_flux(ul[m], ur[m], cL[m], cR[m], rl + NSP * m, rr + NSP * m, F + NSP * m);
}

// inflow species of the species ghost fill (NULL clears)
void orc_set_inflow_species(const double* spec)
{
    g_spec_in_set = spec != nullptr;
    for (int n = 0; n < 9; ++n) g_spec_in[n] = spec ? spec[n] : 0.0;
}

// compute_dSdt_multi for one regular box; S, Spec cover grow(bx,4)
void orc_compute_dsdt_multi_box(const int* lo, const int* hi, const orc_fab* S, const orc_fab* Spec,
                                const orc_fab* dSdt, const orc_fab* dSdt_spec, const double* dxinv, int do_visc,
                                const orc_parm* parm)
{
    Box bx;
    for (int d = 0; d < 3; ++d) { bx.lo[d] = lo[d]; bx.hi[d] = hi[d]; }
    compute_dsdt_multi_box(bx, View(*S), View(*Spec), View(*dSdt), View(*dSdt_spec), dxinv, do_visc, *parm);
}

// compute_dSdt_multi for one box with EB fields (covered / regular / EB-CHEM)
void orc_compute_dsdt_multi_box_eb(const int* lo, const int* hi, const orc_fab* S, const orc_fab* Spec,
                                   const orc_fab* dSdt, const orc_fab* dSdt_spec, const double* dxinv,
                                   const orc_ebfab* eb, int do_visc, const orc_parm* parm)
{
    Box bx;
    for (int d = 0; d < 3; ++d) { bx.lo[d] = lo[d]; bx.hi[d] = hi[d]; }
    compute_dsdt_multi_any(bx, View(*S), View(*Spec), View(*dSdt), View(*dSdt_spec), dxinv, eb, do_visc, *parm);
}

int orc_fill_ghost_spec(int nbox, const int* lo, const int* hi, const orc_fab* fabs, const orc_geom* g, double time)
{
    Level L = make_level(nbox, lo, hi, fabs);
    return fill_level(L, *g, time, 1);
}

// EBR::flow_advance_multi (FlowAdvance.cpp:9-90): scheme 0 Euler, 1 RK2.
// A/As: state and species (U^n in, U^(n+1) out); B/Bs scratch.
int orc_flow_advance_multi(int nbox, const int* lo, const int* hi, const orc_fab* A, const orc_fab* As,
                           const orc_fab* B, const orc_fab* Bs, const orc_geom* g, const orc_parm* parm,
                           double time, double dt, int scheme, int do_visc)
{
    Level LA = make_level(nbox, lo, hi, A), LAs = make_level(nbox, lo, hi, As);
    Level LB = make_level(nbox, lo, hi, B), LBs = make_level(nbox, lo, hi, Bs);
    double dxinv[3] = {1.0 / g->dx[0], 1.0 / g->dx[1], 1.0 / g->dx[2]};
    int bad = 0;
    std::vector<std::vector<double>> d1(nbox), d2(nbox);
    std::vector<View> v1(nbox), v2(nbox);
    for (int b = 0; b < nbox; ++b) {
        const Box& bx = LA.boxes[b];
        const size_t nv = (size_t)(bx.hi[0]-bx.lo[0]+1) * (bx.hi[1]-bx.lo[1]+1) * (bx.hi[2]-bx.lo[2]+1);
        d1[b].assign(NCONS * nv, 0.0); d2[b].assign(NSP * nv, 0.0);
        v1[b] = View(d1[b].data(), bx.lo, bx.hi, NCONS);
        v2[b] = View(d2[b].data(), bx.lo, bx.hi, NSP);
    }
    auto stage = [&](const Level& cur, const Level& curs, const Level* old, const Level* olds, const Level& out,
                     const Level& outs, double c_cur, double c_old, double c_dt) {
        for (int b = 0; b < nbox; ++b)
            compute_dsdt_multi_any(cur.boxes[b], cur.fabs[b], curs.fabs[b], v1[b], v2[b], dxinv,
                                   g_eb.empty() ? nullptr : &g_eb[b], do_visc, *parm);
        for (int b = 0; b < nbox; ++b) {
            for_box(cur.boxes[b], [&](int i, int j, int k) {
                for (int n = 0; n < NCONS; ++n) {
                    double v = old ? c_cur * cur.fabs[b](i, j, k, n) + c_old * old->fabs[b](i, j, k, n)
                                   : c_cur * cur.fabs[b](i, j, k, n);
                    out.fabs[b](i, j, k, n) = v + c_dt * v1[b](i, j, k, n);
                }
                for (int n = 0; n < NSP; ++n) {
                    double v = olds ? c_cur * curs.fabs[b](i, j, k, n) + c_old * olds->fabs[b](i, j, k, n)
                                    : c_cur * curs.fabs[b](i, j, k, n);
                    outs.fabs[b](i, j, k, n) = v + c_dt * v2[b](i, j, k, n);
                }
            });
        }
    };
    bad += fill_level(LA, *g, time);
    bad += fill_level(LAs, *g, time, 1);
    stage(LA, LAs, nullptr, nullptr, LB, LBs, 1.0, 0.0, dt);
    if (scheme == 0) {
        for (int b = 0; b < nbox; ++b) {
            std::memcpy(LA.fabs[b].p, LB.fabs[b].p, sizeof(double) * LA.fabs[b].npts() * NCONS);
            std::memcpy(LAs.fabs[b].p, LBs.fabs[b].p, sizeof(double) * LAs.fabs[b].npts() * NSP);
        }
    } else {
        bad += fill_level(LB, *g, time + dt);
        bad += fill_level(LBs, *g, time + dt, 1);
        stage(LB, LBs, &LA, &LAs, LA, LAs, 0.5, 0.5, 0.5 * dt);
    }
    return bad;
}

// EBR::advance with chemistry (Source/Advance.cpp:152-177): iter = 4/2^level
// (level <= 2, else 1) chemistry steps of dt/(2 iter), the flow step, and
// iter more chemistry steps; then the species renormalisation of
// EBR::post_timestep (Source/EBR.cpp:333-357).
int orc_advance_chem(int nbox, const int* lo, const int* hi, const orc_fab* A, const orc_fab* As, const orc_fab* B,
                     const orc_fab* Bs, const orc_geom* g, const orc_parm* parm, double time, double dt, int level,
                     int scheme, int do_visc, int renorm)
{
    Level LA = make_level(nbox, lo, hi, A), LAs = make_level(nbox, lo, hi, As);
    int iter = level <= 2 ? 4 / (int)std::pow(2, level) : 1;
    double dt1 = 0.5 * dt;
    double dt2 = dt1 / iter;
    for (int n = 0; n < iter; ++n) chem_level(LA, LAs, dt2, *parm);
    int bad = orc_flow_advance_multi(nbox, lo, hi, A, As, B, Bs, g, parm, time, dt, scheme, do_visc);
    for (int n = 0; n < iter; ++n) chem_level(LA, LAs, dt2, *parm);
    if (renorm) {
        for (int b = 0; b < nbox; ++b) {
            const Box& bx = LA.boxes[b];
            for_box(bx, [&](int i, int j, int k) {
                double u[NCONS], r[NSP];
                for (int n = 0; n < NCONS; ++n) u[n] = LA.fabs[b](i, j, k, n);
                for (int n = 0; n < NSP; ++n) r[n] = LAs.fabs[b](i, j, k, n);
                orc_renorm_many(1, u, r);
                for (int n = 0; n < NSP; ++n) LAs.fabs[b](i, j, k, n) = r[n];
            });
        }
    }
    return bad;
}

// Exec/5_CHEMSOD/prob.H:11-82 with prob_parm.H defaults: H2:O2 = 0.667:0.333,
// 20 atm and 1500 K for x < 0.0004, 0.2 atm and 295 K elsewhere, at rest
void orc_init_chemsod(const int* lo, const int* hi, const orc_fab* S, const orc_fab* Spec, const orc_geom* g)
{
    extern double orc_CKRHOY(double P, double T, const double* y);
    extern void orc_CKXTY(const double* x, double* y);
    extern double orc_CKUBMS(double T, const double* rhoi);
    View s(*S), sp(*Spec);
    for (int k = lo[2]; k <= hi[2]; ++k)
        for (int j = lo[1]; j <= hi[1]; ++j)
            for (int i = lo[0]; i <= hi[0]; ++i) {
                double Pt = 0.2 * 101325.0, temp = 295;
                double Xt[NSP] = {0.0}, Yt[NSP], rhoi[NSP];
                Xt[0] = 0.667;
                Xt[1] = 0.333;
                double x = g->prob_lo[0] + (i + 0.5) * g->dx[0];
                if (x < 0.0004) { Pt = 20 * 101325.0; temp = 1500; }
                orc_CKXTY(Xt, Yt);
                double rho = orc_CKRHOY(Pt, temp, Yt);
                for (int n = 0; n < NSP; ++n) rhoi[n] = rho * Yt[n];
                double et = orc_CKUBMS(temp, rhoi);
                s(i, j, k, URHO) = rho;
                s(i, j, k, UMX) = 0.0;
                s(i, j, k, UMY) = 0.0;
                s(i, j, k, UMZ) = 0.0;
                s(i, j, k, UEDEN) = et;
                // ebr_initspec: rho of the state times the same Y
                for (int n = 0; n < NSP; ++n) sp(i, j, k, n) = s(i, j, k, URHO) * Yt[n];
            }
}

} // extern "C"
