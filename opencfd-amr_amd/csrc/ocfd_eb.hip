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

// Van Leer flux vector splitting, Source/EB/EBkernels.H:73-171 (non-CHEM).
// l/r: rho, u, v, w, p, c of the left and right states.
template <int D>
__device__ __forceinline__ void vl_flux(const double* l, const double* r, double gamma, double smallp, double* F)
{
    const double dl = l[0], ul = l[1], vl = l[2], wl = l[3], pl = fmax(l[4], smallp);
    const double dr = r[0], ur = r[1], vr = r[2], wr = r[3], pr = fmax(r[4], smallp);
    const double cL = l[5], cR = r[5];
    const double UL = D == 0 ? ul : (D == 1 ? vl : wl);
    const double UR = D == 0 ? ur : (D == 1 ? vr : wr);
    const double ML = UL / cL, MR = UR / cR;
    const int nx = D == 0, ny = D == 1, nz = D == 2;
    double fp[NC] = {0, 0, 0, 0, 0}, fm[NC] = {0, 0, 0, 0, 0};
    if (ML >= 1.0) {
        fp[0] = dl * UL;
        fp[1] = dl * UL * ul + nx * pl;
        fp[2] = dl * UL * vl + ny * pl;
        fp[3] = dl * UL * wl + nz * pl;
        fp[4] = UL * (gamma * pl / (gamma - 1.0) + 0.5 * dl * (ul * ul + vl * vl + wl * wl));
    } else if (fabs(ML) < 1.0) {
        const double Mp = 0.250 * (1.0 + ML) * (1.0 + ML);
        const double tmp0 = dl * cL * Mp;
        fp[0] = tmp0;
        fp[1] = tmp0 * (nx * (-UL + 2.0 * cL) / gamma + ul);
        fp[2] = tmp0 * (ny * (-UL + 2.0 * cL) / gamma + vl);
        fp[3] = tmp0 * (nz * (-UL + 2.0 * cL) / gamma + wl);
        fp[4] = tmp0 * (((gamma - 1) * UL + 2 * cL) * ((gamma - 1) * UL + 2 * cL) / (2 * (gamma * gamma - 1)) +
                        (ul * ul + vl * vl + wl * wl - UL * UL) / 2);
    }
    if (fabs(MR) < 1.0) {
        const double Mm = -0.250 * (MR - 1.0) * (MR - 1.0);
        const double tmp0 = dr * cR * Mm;
        fm[0] = tmp0;
        fm[1] = tmp0 * (nx * (-UR - 2.0 * cR) / gamma + ur);
        fm[2] = tmp0 * (ny * (-UR - 2.0 * cR) / gamma + vr);
        fm[3] = tmp0 * (nz * (-UR - 2.0 * cR) / gamma + wr);
        fm[4] = tmp0 * (((gamma - 1) * UR - 2 * cR) * ((gamma - 1) * UR - 2 * cR) / (2 * (gamma * gamma - 1)) +
                        (ur * ur + vr * vr + wr * wr - UR * UR) / 2);
    } else if (MR <= -1.0) {
        fm[0] = dr * UR;
        fm[1] = dr * UR * ur + nx * pr;
        fm[2] = dr * UR * vr + ny * pr;
        fm[3] = dr * UR * wr + nz * pr;
        fm[4] = UR * (gamma * pr / (gamma - 1.0) + 0.5 * dr * (ur * ur + vr * vr + wr * wr));
    }
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
