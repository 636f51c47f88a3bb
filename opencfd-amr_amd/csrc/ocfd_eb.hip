// Embedded-boundary (cut-cell) box path: EBR::eb_compute_dSdt_box
// (Source/EB/AdvanceBox.cpp:15-228) with the kernels of Source/EB/EBkernels.H
// and Source/EB/EBdiffusion.H, for the boxes of a level that touch the body.
//
// Three kernels per RK stage over the cut boxes' work-lists:
//   k_eb_prims   c2prim (7 comps, FP64) on grow(bx,2)
//   k_eb_face<D> per face of surroundingNodes(bx,D): MUSCL/minmod left and
//                right states where the cells connect, Van Leer flux, and
//                the viscous flux with one-sided tangential differences
//   k_eb_div     per valid cell (cut and covered boxes): conservative
//                divergence with area fractions, the wall flux, the
//                small-cell scaling, gravity and the fused stage update
// All in FP64: cut boxes are a thin layer around the body.
#include "ocfd_device.hpp"
#include "ocfd_internal.hpp"

namespace ocfd {

enum { QRHO = 0, QU = 1, QV = 2, QW = 3, QPRES = 4, QT = 5, QC = 6, NQ = 7 };

// amrex::EBCellFlag bit layout: type in bits 0-1, neighbour (i,j,k)
// connected at bit 5 + (i+1) + 3(j+1) + 9(k+1).  The reference also queries
// offsets of 2 (EBkernels.H:20,40,60); the same formula answers them.
__device__ __forceinline__ bool eb_conn(uint32_t f, int i, int j, int k)
{
    return (f >> (5 + (i + 1) + 3 * (j + 1) + 9 * (k + 1))) & 1u;
}
__device__ __forceinline__ bool eb_covered(uint32_t f) { return (f & 3u) == 3u; }
__device__ __forceinline__ bool eb_regular(uint32_t f) { return (f & 3u) == 0u; }

__device__ __forceinline__ long gloc(const BoxDesc& B, int i, int j, int k)
{
    return ((long)(k - B.lo[2] + NG) * B.nyg + (j - B.lo[1] + NG)) * B.nxg + (i - B.lo[0] + NG);
}

struct EBArgs {
    const uint32_t* flag;
    const double* geo;
    double* q;
    double* F[3];
};

// Source/kernel/Reconstruction.H:9-19
__device__ __forceinline__ double minmod(double a, double b)
{
    if (a * b > 0) return fabs(a) > fabs(b) ? b : a;
    return 0.0;
}

// c2prim (Source/kernel/Kernels.H:13-40) on grow(bx,2) of a cut box
__global__ __launch_bounds__(256) void k_eb_prims(const BoxDesc* __restrict__ boxes, const int4* __restrict__ tiles,
                                                  StageArgs a, EBArgs e)
{
    if (ctl_done(a.ctl)) return;
    const int4 t = tiles[blockIdx.x];
    if ((int)threadIdx.x >= t.z) return;
    const BoxDesc B = boxes[t.x];
    const int g = t.w;
    const int nx = B.hi[0] - B.lo[0] + 1 + 2 * g, ny = B.hi[1] - B.lo[1] + 1 + 2 * g;
    const int c = t.y + threadIdx.x;
    const int i = B.lo[0] - g + c % nx, j = B.lo[1] - g + (c / nx) % ny, k = B.lo[2] - g + c / (nx * ny);
    const long s0 = sidx(B, i, j, k), st = B.nptg;
    const PhysArgs& ph = a.ph;
    const double rho = fmax(a.scur[s0], ph.smallr);
    const double rhoinv = 1.0 / rho;
    const double ux = a.scur[s0 + st] * rhoinv;
    const double uy = a.scur[s0 + 2 * st] * rhoinv;
    const double uz = a.scur[s0 + 3 * st] * rhoinv;
    const double kineng = 0.5 * (ux * ux + uy * uy + uz * uz);
    const double ei = a.scur[s0 + 4 * st] - rho * kineng;
    const double p = fmax((ph.gamma - 1.0) * ei, ph.smallp);
    const double T = p / (rho * ph.Rg);
    double* q = e.q + B.off_e * NQ + gloc(B, i, j, k);
    q[0] = rho; q[st] = ux; q[2 * st] = uy; q[3 * st] = uz; q[4 * st] = p; q[5 * st] = T;
    q[6 * st] = sqrt(ph.gamma * ph.Rg * T);
}

// One side of Van Leer's flux-vector splitting (B. van Leer, Flux-vector
// splitting for the Euler equations, Lecture Notes in Physics 170 (1982)
// 507-512) along direction D, for a state q = (rho, u, v, w, p, c): s = +1
// gives the forward part F+ of the left state, s = -1 the backward part F- of
// the right state.  Supersonic towards the part's side: the whole Euler
// flux; subsonic: the split mass flux s rho c (1 + s M)^2 / 4 carrying the
// normal momentum ((gamma-1) u_n + 2 s c)/gamma, the tangential velocities
// and the energy ((gamma-1) u_n + 2 s c)^2 / (2 (gamma^2-1)) + |u_t|^2 / 2;
// otherwise nothing.  Used as the cut-cell face flux of
// Source/EB/EBkernels.H:73-171 (non-CHEM build, gamma constant).
template <int D>
__device__ __forceinline__ void vl_part(const double* q, double sgn, double gamma, double smallp, double* f)
{
    const double rho = q[0], p = fmax(q[4], smallp), c = q[5];
    const double un = q[1 + D];
    const double M = un / c;
    const double uu = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    if (sgn * M >= 1.0) {
        const double m = rho * un;
        f[0] = m;
#pragma unroll
        for (int a = 0; a < 3; ++a) f[1 + a] = m * q[1 + a] + (a == D ? p : 0.0);
        f[4] = un * (gamma * p / (gamma - 1.0) + 0.5 * rho * uu);
    } else if (fabs(M) < 1.0) {
        const double w = 1.0 + sgn * M;
        const double m = rho * c * (sgn * 0.250 * w * w);
        const double a2 = (gamma - 1) * un + sgn * 2.0 * c;
        f[0] = m;
#pragma unroll
        for (int a = 0; a < 3; ++a) f[1 + a] = a == D ? m * ((-un + sgn * 2.0 * c) / gamma + un) : m * q[1 + a];
        f[4] = m * (a2 * a2 / (2 * (gamma * gamma - 1)) + (uu - un * un) / 2);
    } else {
#pragma unroll
        for (int n = 0; n < NC; ++n) f[n] = 0.0;
    }
}

// the face flux F = F+(left) + F-(right)
template <int D>
__device__ __forceinline__ void vl_flux(const double* l, const double* r, double gamma, double smallp, double* F)
{
    double fp[NC], fm[NC];
    vl_part<D>(l, 1.0, gamma, smallp, fp);
    vl_part<D>(r, -1.0, gamma, smallp, fm);
#pragma unroll
    for (int n = 0; n < NC; ++n) F[n] = fp[n] + fm[n];
}

// One face (the low face of cell (i,j,k) in direction D) of a cut box:
// eb_recon (EBkernels.H:10-71), eb_flux and eb_compute_visc_D
// (EBdiffusion.H:8-203).
template <int D>
__global__ __launch_bounds__(256) void k_eb_face(const BoxDesc* __restrict__ boxes, const int4* __restrict__ tiles,
                                                 StageArgs a, EBArgs e)
{
    if (ctl_done(a.ctl)) return;
    const int4 t = tiles[blockIdx.x];
    if ((int)threadIdx.x >= t.z) return;
    const BoxDesc B = boxes[t.x];
    const int di = D == 0, dj = D == 1, dk = D == 2;
    const int nx = B.hi[0] - B.lo[0] + 1 + di, ny = B.hi[1] - B.lo[1] + 1 + dj;
    const int c = t.y + threadIdx.x;
    const int i = B.lo[0] + c % nx, j = B.lo[1] + (c / nx) % ny, k = B.lo[2] + c / (nx * ny);
    const long st = B.nptg;
    const long dstep = D == 0 ? 1 : (D == 1 ? (long)B.nxg : (long)B.nxg * B.nyg);
    const long g0 = gloc(B, i, j, k);
    const double* q = e.q + B.off_e * NQ;
    const uint32_t* fl = e.flag + B.off_e;
    const uint32_t f = fl[g0];
    const bool cm = eb_conn(f, -di, -dj, -dk), cm2 = eb_conn(f, -2 * di, -2 * dj, -2 * dk), cp = eb_conn(f, di, dj, dk);
    // left/right states of rho, u, v, w, p, c (T is not used by the flux)
    double l[6], r[6];
    const int comp[6] = {QRHO, QU, QV, QW, QPRES, QC};
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        const double* qn = q + comp[m] * st + g0;
        const double q0 = qn[0], qm = qn[-dstep];
        double ql = qm, qr = q0;
        if (cm) {
            if (cm2) ql = qm + 0.5 * minmod(q0 - qm, qm - qn[-2 * dstep]);
            if (cp) qr = q0 - 0.5 * minmod(qn[dstep] - q0, q0 - qm);
        }
        l[m] = ql;
        r[m] = qr;
    }
    double F[NC];
    vl_flux<D>(l, r, a.ph.gamma, a.ph.smallp, F);

    if (a.flags & OCFD_VISC) {
        const PhysArgs& ph = a.ph;
        const double weights[3] = {0.0, 1.0, 0.5};
        const long gm = g0 - dstep;
        const double Tl = q[QT * st + g0], Tr = q[QT * st + gm];
        const double mutl = ph.C_s * Tl * sqrt(Tl) / (Tl + ph.T_s);
        const double mutr = ph.C_s * Tr * sqrt(Tr) / (Tr + ph.T_s);
        const double muf = 0.5 * (mutl + mutr);
        const double ktl = mutl * ph.kOverMu, ktr = mutr * ph.kOverMu;
        const double dxn = B.dxinv[D];
        const double dTn = (Tl - Tr) * dxn;
        double dn[3];
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) dn[cc] = (q[(QU + cc) * st + g0] - q[(QU + cc) * st + gm]) * dxn;
        const uint32_t fh = f, fo = fl[gm];
        // tangential difference of comps c1, c2 along direction T
        auto tang = [&](int T, int c1, int c2, double& o1, double& o2) {
            const int ti = T == 0, tj = T == 1, tk = T == 2;
            const long ts = T == 0 ? 1 : (T == 1 ? (long)B.nxg : (long)B.nxg * B.nyg);
            const int hp = eb_conn(fh, ti, tj, tk) ? 1 : 0, hm = eb_conn(fh, -ti, -tj, -tk) ? 1 : 0;
            const int lp = eb_conn(fo, ti, tj, tk) ? 1 : 0, lm = eb_conn(fo, -ti, -tj, -tk) ? 1 : 0;
            const double whi = weights[hp + hm], wlo = weights[lp + lm];
            const double h = 0.5 * B.dxinv[T];
            const double* a1 = q + c1 * st;
            const double* a2 = q + c2 * st;
            o1 = h * ((a1[g0 + hp * ts] - a1[g0 - hm * ts]) * whi + (a1[gm + lp * ts] - a1[gm - lm * ts]) * wlo);
            o2 = h * ((a2[g0 + hp * ts] - a2[g0 - hm * ts]) * whi + (a2[gm + lp * ts] - a2[gm - lm * ts]) * wlo);
        };
        double t1, t2, t3;
        if (D == 0) {
            double dudy, dvdy, dudz, dwdz;
            tang(1, QU, QV, dudy, dvdy);
            tang(2, QU, QW, dudz, dwdz);
            const double divu = dn[0] + dvdy + dwdz;
            t1 = muf * (2.0 * dn[0] - ph.twoThirds * divu);
            t2 = muf * (dudy + dn[1]);
            t3 = muf * (dudz + dn[2]);
        } else if (D == 1) {
            double dudx, dvdx, dvdz, dwdz;
            tang(0, QU, QV, dudx, dvdx);
            tang(2, QV, QW, dvdz, dwdz);
            const double divu = dudx + dn[1] + dwdz;
            t2 = muf * (2.0 * dn[1] - ph.twoThirds * divu);
            t1 = muf * (dn[0] + dvdx);
            t3 = muf * (dn[2] + dvdz);
        } else {
            double dudx, dwdx, dvdy, dwdy;
            tang(0, QU, QW, dudx, dwdx);
            tang(1, QV, QW, dvdy, dwdy);
            const double divu = dudx + dvdy + dn[2];
            t1 = muf * (dn[0] + dwdx);
            t2 = muf * (dn[1] + dwdy);
            t3 = muf * (2.0 * dn[2] - ph.twoThirds * divu);
        }
        F[UMX] += -t1;
        F[UMY] += -t2;
        F[UMZ] += -t3;
        F[UEDEN] += -0.5 * ((q[QU * st + g0] + q[QU * st + gm]) * t1 + (q[QV * st + g0] + q[QV * st + gm]) * t2 +
                            (q[QW * st + g0] + q[QW * st + gm]) * t3 + (ktl + ktr) * dTn);
    }
    double* Fo = e.F[D] + B.off_e * NC + g0;
#pragma unroll
    for (int n = 0; n < NC; ++n) Fo[n * st] = F[n];
}

// eb_compute_div(_visc) (EBkernels.H:216-338), flux_redist scaling
// (EBkernels.H:341-377, AdvanceBox.cpp:192-214), gravity and the stage update,
// per valid cell of a cut box; covered boxes get dSdt = 0
// (Source/Advance.cpp:213-214).
__global__ __launch_bounds__(256) void k_eb_div(const BoxDesc* __restrict__ boxes, const int4* __restrict__ tiles,
                                                StageArgs a, EBArgs e)
{
    if (ctl_done(a.ctl)) return;
    const int4 t = tiles[blockIdx.x];
    if ((int)threadIdx.x >= t.z) return;
    const BoxDesc B = boxes[t.x];
    const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1;
    const int c = t.y + threadIdx.x;
    const int i = B.lo[0] + c % nx, j = B.lo[1] + (c / nx) % ny, k = B.lo[2] + c / (nx * ny);
    double L[NC];
    if (B.type == 2) {
#pragma unroll
        for (int n = 0; n < NC; ++n) L[n] = 0.0;
        StageArgs b = a;
        b.flags &= ~OCFD_GRAVITY;
        finish_cell(b, B, i, j, k, L);
        return;
    }
    const long st = B.nptg;
    const long g0 = gloc(B, i, j, k);
    const long sy = B.nxg, sz = (long)B.nxg * B.nyg;
    const uint32_t f = e.flag[B.off_e + g0];
    const double* geo = e.geo + B.off_e * 4;
    const double vf = geo[g0];
    const double* ax = geo + st;
    const double* ay = geo + 2 * st;
    const double* az = geo + 3 * st;
    const double* Fx = e.F[0] + B.off_e * NC;
    const double* Fy = e.F[1] + B.off_e * NC;
    const double* Fz = e.F[2] + B.off_e * NC;
    const double tmp = B.dxinv[0] / vf;
    if (eb_covered(f)) {
#pragma unroll
        for (int n = 0; n < NC; ++n) L[n] = 0.0;
    } else if (eb_regular(f)) {
#pragma unroll
        for (int n = 0; n < NC; ++n)
            L[n] = B.dxinv[0] * (Fx[n * st + g0 + 1] - Fx[n * st + g0]) +
                   B.dxinv[1] * (Fy[n * st + g0 + sy] - Fy[n * st + g0]) +
                   B.dxinv[2] * (Fz[n * st + g0 + sz] - Fz[n * st + g0]);
    } else {
        const double axm = ax[g0], axp = ax[g0 + 1], aym = ay[g0], ayp = ay[g0 + sy], azm = az[g0], azp = az[g0 + sz];
#pragma unroll
        for (int n = 0; n < NC; ++n)
            L[n] = tmp * (axp * Fx[n * st + g0 + 1] - axm * Fx[n * st + g0] + ayp * Fy[n * st + g0 + sy] -
                          aym * Fy[n * st + g0] + azp * Fz[n * st + g0 + sz] - azm * Fz[n * st + g0]);
        const double* q = e.q + B.off_e * NQ + g0;
        double fw[NC];
        const double p = q[QPRES * st];
        if (a.flags & OCFD_VISC) {
            // eb_wallflux, EBkernels.H:173-214
            const double d = q[0], u = q[QU * st], v = q[QV * st], w = q[QW * st], cs = q[QC * st];
            const double apnorm = sqrt((axm - axp) * (axm - axp) + (aym - ayp) * (aym - ayp) + (azm - azp) * (azm - azp));
            const double un = u * (axm - axp) + v * (aym - ayp) + w * (azm - azp);
            const double lambda0 = fabs(un) + cs * apnorm;
            const double gamma = a.ph.gamma;
            const double E0 = p / (gamma - 1) + 0.50 * d * (u * u + v * v + w * w);
            const double fp0 = 0.5 * (d * un + lambda0 * d);
            const double fp1 = 0.5 * (d * u * un + (axm - axp) * p + lambda0 * d * u);
            const double fp2 = 0.5 * (d * v * un + (aym - ayp) * p + lambda0 * d * v);
            const double fp3 = 0.5 * (d * w * un + (azm - azp) * p + lambda0 * d * w);
            const double fp4 = 0.5 * ((E0 + p) * un + lambda0 * E0);
            const double fm0 = 0.5 * (d * (-un) - lambda0 * d);
            const double fm4 = 0.5 * (-(E0 + p) * un - lambda0 * E0);
            fw[0] = fp0 + fm0;
            fw[1] = fp1 + fp1;
            fw[2] = fp2 + fp2;
            fw[3] = fp3 + fp3;
            fw[4] = fp4 + fm4;
        } else {
            fw[URHO] = 0.0;
            fw[UMX] = (axm - axp) * p;
            fw[UMY] = (aym - ayp) * p;
            fw[UMZ] = (azm - azp) * p;
            fw[UEDEN] = 0.0;
        }
#pragma unroll
        for (int n = 0; n < NC; ++n) L[n] += fw[n] * tmp;
    }
    const bool scale = (a.flags & OCFD_EB_REDIST) && vf < 0.5 && vf > 0.0;
#pragma unroll
    for (int n = 0; n < NC; ++n) {
        L[n] *= -1.0;
        if (scale) L[n] = vf * L[n];
    }
    finish_cell(a, B, i, j, k, L);
}

void launch_eb(const LevelDev& L, const StageArgs& a, hipStream_t s)
{
    if (!L.ntile_ed) return;
    EBArgs e{L.d_ebflag, L.d_ebgeo, L.d_ebq, {L.d_ebF[0], L.d_ebF[1], L.d_ebF[2]}};
    if (L.ntile_eq) {
        hipLaunchKernelGGL(k_eb_prims, dim3(L.ntile_eq), dim3(256), 0, s, L.d_boxes, L.d_tile_eq, a, e);
        hipLaunchKernelGGL(k_eb_face<0>, dim3(L.ntile_ef[0]), dim3(256), 0, s, L.d_boxes, L.d_tile_ef[0], a, e);
        hipLaunchKernelGGL(k_eb_face<1>, dim3(L.ntile_ef[1]), dim3(256), 0, s, L.d_boxes, L.d_tile_ef[1], a, e);
        hipLaunchKernelGGL(k_eb_face<2>, dim3(L.ntile_ef[2]), dim3(256), 0, s, L.d_boxes, L.d_tile_ef[2], a, e);
    }
    hipLaunchKernelGGL(k_eb_div, dim3(L.ntile_ed), dim3(256), 0, s, L.d_boxes, L.d_tile_ed, a, e);
}

} // namespace ocfd
