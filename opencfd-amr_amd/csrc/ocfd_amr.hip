// AMR level machinery between two levels of refinement ratio 2:
//   * density-gradient tagging, tag_dengrad (Source/Tagging.H:8-36);
//   * the coarse part of the two-level FillPatch: fine ghost cells that no
//     fine box owns are interpolated from the coarse level, already combined
//     in time and ghost-filled at the stage time, with the cell-conservative
//     linear interpolation of the state descriptor (Source/Setup.cpp:196);
//   * the same interpolation for the valid cells of a new fine region, and the
//     copy of an old level's cells into a new one (regrid);
//   * average-down, EBR::avgDown (Source/EBR.cpp:535-559) with unit volumes;
//   * the flux register: coarse and fine face fluxes of the coarse/fine
//     interface weighted by the RK stage dt (the CrseAdd / FineAdd calls of
//     Source/Advance.cpp:344-360) and the reflux of EBR::post_timestep
//     (Source/EBR.cpp:322-330).
//
// The interface fluxes are recomputed at the interface faces only, from the
// primitive scratch the stage's sweeps have just used, so the fused sweeps
// carry no flux-register work.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include <array>

#include "ocfd_device.hpp"
#include "ocfd_internal.hpp"
#include "ocfd_host_pool.hpp"

namespace ocfd {

// ---------------------------------------------------------------------------
// face flux of the regular path at one face
// ---------------------------------------------------------------------------
// Steger-Warming + WENO-Z5 (+ viscous) flux of the low face of the cell at
// primitive index s (grown layout of box B) in direction D: the value the
// sweeps compute for that face, with the same operations in the same order.
template <int D, class FT, bool VISC>
__device__ __forceinline__ void face_flux(const FT* __restrict__ P, long s, const BoxDesc& B, const PhysArgs& ph,
                                          double* Fout)
{
    const long st = B.nptg;
    const long ds = D == 0 ? 1 : (D == 1 ? (long)B.nxg : (long)B.nxg * B.nyg);
    const FT gam = (FT)ph.gamma;
    FT fp[6][5], fm[6][5];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        Prim<FT> q = ldprim<FT>(P, s + (c - 3) * ds, st);
        sw_split<D, FT>(q, gam, sw_t1<FT>(ph), sw_t3<FT>(ph), fp[c], fm[c]);
    }
    FT F[5];
    if constexpr (D == 0 && sizeof(FT) == 4) {
        // the FP32 x sweep's form: sub-stencil quantities by centre cell
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            const WenoCentre p1 = weno_centre(fp[0][n], fp[1][n], fp[2][n]);
            const WenoCentre p2 = weno_centre(fp[1][n], fp[2][n], fp[3][n]);
            const WenoCentre p3 = weno_centre(fp[2][n], fp[3][n], fp[4][n]);
            const WenoCentre m2 = weno_centre(fm[1][n], fm[2][n], fm[3][n]);
            const WenoCentre m3 = weno_centre(fm[2][n], fm[3][n], fm[4][n]);
            const WenoCentre m4 = weno_centre(fm[3][n], fm[4][n], fm[5][n]);
            F[n] = wenoz_centred(p1.b0, p2.b1, p3.b2, p1.p0, p2.p1, p3.p2) +
                   wenoz_centred(m4.b2, m3.b1, m2.b0, m4.m0, m3.p2, m2.p1);
        }
    } else {
#pragma unroll
        for (int n = 0; n < 5; ++n)
            F[n] = weno_p<FT>(fp[0][n], fp[1][n], fp[2][n], fp[3][n], fp[4][n]) +
                   weno_m<FT>(fm[1][n], fm[2][n], fm[3][n], fm[4][n], fm[5][n]);
    }
    if (redo_scaled(F))   // as the sweeps do
        weno_face_scaled<FT>([&](int c, int n) { return fp[c][n]; }, [&](int c, int n) { return fm[c + 1][n]; }, F);
    if (VISC) {
        const long sL = s - ds;
        const Prim<FT> L = ldprim<FT>(P, sL, st), R = ldprim<FT>(P, s, st);
        // lateral directions a (first of x,y,z other than D) and b (second)
        constexpr int A = D == 0 ? 1 : 0;
        constexpr int Bd = D == 2 ? 1 : 2;
        const long sa = A == 0 ? 1 : (long)B.nxg;
        const long sb = Bd == 1 ? (long)B.nxg : (long)B.nxg * B.nyg;
        // velocity components of the four 4-point differences:
        //   x face: (u,v) along y, (u,w) along z
        //   y face: (u,v) along x, (v,w) along z
        //   z face: (u,w) along x, (v,w) along y
        constexpr int ca = 0, cb = D == 2 ? 2 : 1, cc = D == 0 ? 0 : 1, cd = 2;
        FT v[8][3];
        const long off[8] = {s + sa, sL + sa, s - sa, sL - sa, s + sb, sL + sb, s - sb, sL - sb};
#pragma unroll
        for (int m = 0; m < 8; ++m) ldvel<FT>(P, off[m], st, v[m][0], v[m][1], v[m][2]);
        const FT ta = lat(v[0][ca], v[1][ca], v[2][ca], v[3][ca]);
        const FT tb = lat(v[0][cb], v[1][cb], v[2][cb], v[3][cb]);
        const FT tc = lat(v[4][cc], v[5][cc], v[6][cc], v[7][cc]);
        const FT td = lat(v[4][cd], v[5][cd], v[6][cd], v[7][cd]);
        visc_face<D, FT>(L, R, ta, tb, tc, td, B.dxinv, ph, F);
    }
#pragma unroll
    for (int n = 0; n < 5; ++n) Fout[n] = (double)F[n];
}

// one interface face, seen from the cell (box, i, j, k) that owns it on its
// level: dir, and up = 1 when it is the high face of that cell
struct CFFace {
    int box, i, j, k;
    int dir, up;
    int fbox;   // coarse faces: the fine box across the face
    int pad1;
};

// reg[f] += w * F(face f) for every listed face; one thread per face, each
// slot written by one thread per launch, so the sums are deterministic
// zrep: a z-uniform level whose sweeps computed the primitives of the first
// z layer only; every x or y face takes the flux of its first-layer twin.
// FINE: the fine side, 4 sub-faces per listed coarse face (thread 4f + s):
// sub-face s of face f is the low or high face (across the interface) of
// fine cell (2I, 2J, 2K) + (a1, a2) in the two tangential directions of the
// first fine layer beyond the coarse cell, a1 = s & 1, a2 = s >> 1, wrapped
// to the fine domain on periodic sides, in the fine box c.fbox.
__device__ __forceinline__ int wrap_dev(const GeomDesc& G, int d, int v)
{
    if (!G.periodic[d] || (v >= G.dom_lo[d] && v <= G.dom_hi[d])) return v;
    const int len = G.dom_hi[d] - G.dom_lo[d] + 1;
    return G.dom_lo[d] + (((v - G.dom_lo[d]) % len) + len) % len;
}

template <class FT, bool VISC, bool FINE>
__global__ __launch_bounds__(128) void k_cf_flux(const BoxDesc* __restrict__ boxes, const CFFace* __restrict__ faces,
                                                 int nface, const void* prim, PhysArgs ph, double w,
                                                 double* __restrict__ reg, int zrep,
                                                 const GeomDesc* __restrict__ fgeom)
{
    const int f = blockIdx.x * 128 + threadIdx.x;
    if (f >= nface) return;
    CFFace c = faces[FINE ? f >> 2 : f];
    if (FINE) {
        const int d = c.dir, side = c.up;
        int nb[3] = {c.i, c.j, c.k};
        nb[d] += side ? 1 : -1;
        int g[3] = {2 * c.i, 2 * c.j, 2 * c.k};
        g[d] = side ? 2 * nb[d] : 2 * nb[d] + 1;
        const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
        g[t1] += f & 1;
        g[t2] += (f >> 1) & 1;
        const GeomDesc& G = *fgeom;
        c.box = c.fbox;
        c.i = wrap_dev(G, 0, g[0]); c.j = wrap_dev(G, 1, g[1]); c.k = wrap_dev(G, 2, g[2]);
        c.up = side ? 0 : 1;   // fine level above the face: its low face
    }
    const BoxDesc B = boxes[c.box];
    const int i = c.i + (c.up && c.dir == 0), j = c.j + (c.up && c.dir == 1);
    const int k = zrep && c.dir != 2 ? B.lo[2] : c.k + (c.up && c.dir == 2);
    const long s = sidx(B, i, j, k) - B.off_s + B.off_p;
    const FT* P = (const FT*)prim;
    double F[5];
    if (c.dir == 0) face_flux<0, FT, VISC>(P, s, B, ph, F);
    else if (c.dir == 1) face_flux<1, FT, VISC>(P, s, B, ph, F);
    else face_flux<2, FT, VISC>(P, s, B, ph, F);
#pragma unroll
    for (int n = 0; n < 5; ++n) reg[(long)f * 5 + n] += w * F[n];
}

__global__ void k_zero(double* __restrict__ p, long n)
{
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) p[i] = 0.0;
}

// reflux: per coarse cell next to the fine level, over its interface faces in
// the fixed order (x lo, x hi, y lo, y hi, z lo, z hi),
//   U += -s * (mean of the 4 fine sub-face sums - coarse sum) * dxinv,
// s = +1 for the cell's high face, -1 for its low face
__global__ __launch_bounds__(128) void k_reflux(const BoxDesc* __restrict__ boxes, const int4* __restrict__ cells,
                                                const int* __restrict__ cellfaces, int ncell,
                                                const CFFace* __restrict__ cfaces, const double* __restrict__ regc,
                                                const double* __restrict__ regf, double* __restrict__ S)
{
    const int c = blockIdx.x * 128 + threadIdx.x;
    if (c >= ncell) return;
    const int4 cl = cells[c];
    const BoxDesc B = boxes[cl.x];
    const long s0 = sidx(B, cl.y, cl.z, cl.w), st = B.nptg;
    double U[5];
#pragma unroll
    for (int n = 0; n < 5; ++n) U[n] = S[s0 + n * st];
    for (int m = 0; m < 6; ++m) {
        const int f = cellfaces[c * 6 + m];
        if (f < 0) continue;
        const CFFace F = cfaces[f];
        const double sgn = F.up ? 1.0 : -1.0;
        const double dxi = B.dxinv[F.dir];
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            const double fine = (((regf[(4L * f + 0) * 5 + n] + regf[(4L * f + 1) * 5 + n]) + regf[(4L * f + 2) * 5 + n]) +
                                 regf[(4L * f + 3) * 5 + n]) * 0.25;
            U[n] += -sgn * (fine - regc[(long)f * 5 + n]) * dxi;
        }
    }
#pragma unroll
    for (int n = 0; n < 5; ++n) S[s0 + n * st] = U[n];
}

__device__ __forceinline__ void valid_cell(const BoxDesc& B, int c, int& i, int& j, int& k)
{
    const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1;
    i = B.lo[0] + c % nx;
    j = B.lo[1] + (c / nx) % ny;
    k = B.lo[2] + c / (nx * ny);
}

// average-down: every coarse cell under a fine box takes the volume-weighted
// mean of its 8 fine cells (EB_average_down with unit volume fractions):
// sum in i-fastest order, then divided by the summed weight 8.
// tiles: (fine box, first coarse cell of the coarsened box, count, 0)
__global__ __launch_bounds__(256) void k_avg_down(const BoxDesc* __restrict__ fboxes, const BoxDesc* __restrict__ cboxes,
                                                  const GeomDesc* __restrict__ cgeom, const int4* __restrict__ tiles,
                                                  const double* __restrict__ Sf, double* __restrict__ Sc, int ncomp)
{
    const int4 t = tiles[blockIdx.x];
    if ((int)threadIdx.x >= t.z) return;
    const BoxDesc Bf = fboxes[t.x];
    const int nx = (Bf.hi[0] - Bf.lo[0] + 1) >> 1, ny = (Bf.hi[1] - Bf.lo[1] + 1) >> 1;
    const int c = t.y + threadIdx.x;
    const int I = (Bf.lo[0] >> 1) + c % nx, J = (Bf.lo[1] >> 1) + (c / nx) % ny, K = (Bf.lo[2] >> 1) + c / (nx * ny);
    const int m = level_owner(*cgeom, cboxes, I, J, K);
    if (m < 0) return;
    const BoxDesc Bc = cboxes[m];
    const long dst = Bc.off_s / NC * ncomp + (sidx(Bc, I, J, K) - Bc.off_s);
    const long src = Bf.off_s / NC * ncomp + (sidx(Bf, 2 * I, 2 * J, 2 * K) - Bf.off_s);
    const long sy = Bf.nxg, sz = (long)Bf.nxg * Bf.nyg;
    for (int n = 0; n < ncomp; ++n) {
        const double* f = Sf + src + n * Bf.nptg;
        double v = 0.0;
#pragma unroll
        for (int kk = 0; kk < 2; ++kk)
#pragma unroll
            for (int jj = 0; jj < 2; ++jj)
#pragma unroll
                for (int ii = 0; ii < 2; ++ii) v += f[kk * sz + jj * sy + ii];
        Sc[dst + n * Bc.nptg] = v / 8.0;
    }
}

// Central slope of one component along one direction; at a domain edge with an
// external-Dirichlet BC, the third-order one-sided form that takes the ghost
// value as the boundary-face value (the slopes of AMReX's cell-conservative
// linear interpolation).
__device__ __forceinline__ double cc_slope(const double* u, long s, long d, bool ext_lo, bool ext_hi)
{
    double dc = 0.5 * (u[s + d] - u[s - d]);
    if (ext_lo) dc = -(16.0 / 15.0) * u[s - d] + 0.5 * u[s] + (2.0 / 3.0) * u[s + d] - 0.1 * u[s + 2 * d];
    if (ext_hi) dc = (16.0 / 15.0) * u[s + d] - 0.5 * u[s] - (2.0 / 3.0) * u[s - d] + 0.1 * u[s - 2 * d];
    return dc;
}

// MC limiting of a central slope by twice the one-sided differences
__device__ __forceinline__ double mc_limit(double dc, double um, double u0, double up)
{
    const double df = 2.0 * (up - u0), db = 2.0 * (u0 - um);
    const double s = (df * db >= 0.0) ? fmin(fabs(df), fabs(db)) : 0.0;
    return copysign(1.0, dc) * fmin(s, fabs(dc));
}

// Cell-conservative linear interpolation (refinement ratio 2) of component n
// of the coarse state C at fine cell (i,j,k): MC-limited central slopes, then
// one factor alpha <= 1 that keeps every fine value of the parent within the
// min/max of the parent's 27 coarse neighbours; fine value =
// u + xoff*sx + yoff*sy + zoff*sz with offsets +-1/4.  cmap gives, per coarse
// cell, the coarse box to read it from: its owner, or else a box whose filled
// ghost region holds the cell and its neighbours (a coarse level that does not
// cover the domain).  Returns false when there is none.
// coarse box the interpolation of fine cell (i,j,k) reads, -1 if none
__device__ __forceinline__ int interp_box(const BoxDesc* __restrict__ cboxes, const GeomDesc& CG,
                                          const int* __restrict__ cmap, int i, int j, int k)
{
    const int I = i >> 1, J = j >> 1, K = k >> 1;
    if (I < CG.dom_lo[0] || I > CG.dom_hi[0] || J < CG.dom_lo[1] || J > CG.dom_hi[1] || K < CG.dom_lo[2] ||
        K > CG.dom_hi[2])
        return -1;
    const long ntile = (long)CG.map_n[0] * CG.map_n[1] * CG.map_n[2];
    return tile_owner(cmap, cmap + ntile, CG.map_n, CG.dom_lo, CG.map_shift, cboxes, I, J, K, 3);
}

// component n of fine cell (i,j,k) interpolated from coarse box Bc
// czuni: the coarse level is z-uniform (planar): every layer holds the first
// layer's values, and its ghost fill covers the first layers only, so the
// parent and its x/y neighbours are read in the first layer of its box, the z
// slope is zero and the limiter's z neighbours repeat that layer (the values a
// full read of z-uniform data gives, bit for bit, with a third of the loads).
__device__ __forceinline__ double interp_value(const BoxDesc& Bc, const GeomDesc& CG, const double* __restrict__ C,
                                               int i, int j, int k, int n, int czuni)
{
    const int I = i >> 1, J = j >> 1, K = czuni ? Bc.lo[2] : k >> 1;
    const double* u = C + n * Bc.nptg;
    const long s = sidx(Bc, I, J, K);
    const long dy = Bc.nxg, dz = (long)Bc.nxg * Bc.nyg;
    const long dd[3] = {1, dy, dz};
    const int iv[3] = {I, J, K};
    const double u0 = u[s];
    double sl[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d == 2 && czuni) break;
        const bool elo = iv[d] == CG.dom_lo[d] && CG.bc[n][d][0] == BC_EXT && !CG.periodic[d];
        const bool ehi = iv[d] == CG.dom_hi[d] && CG.bc[n][d][1] == BC_EXT && !CG.periodic[d];
        const double dc = cc_slope(u, s, dd[d], elo, ehi);
        sl[d] = mc_limit(dc, u[s - dd[d]], u0, u[s + dd[d]]);
    }
    double alpha = 1.0;
    if (sl[0] != 0.0 || sl[1] != 0.0 || sl[2] != 0.0) {
        const double dumax = fabs(sl[0]) * 0.25 + fabs(sl[1]) * 0.25 + fabs(sl[2]) * 0.25;
        double umax = u0, umin = u0;
        const int kz = czuni ? 0 : 1;
        for (int kk = -kz; kk <= kz; ++kk)
            for (int jj = -1; jj <= 1; ++jj)
                for (int ii = -1; ii <= 1; ++ii) {
                    const double w = u[s + kk * dz + jj * dy + ii];
                    umax = fmax(umax, w);
                    umin = fmin(umin, w);
                }
        if (dumax * alpha > umax - u0) alpha = (umax - u0) / dumax;
        if (dumax * alpha > u0 - umin) alpha = (u0 - umin) / dumax;
    }
    const double xo = (i & 1) ? 0.25 : -0.25, yo = (j & 1) ? 0.25 : -0.25, zo = (k & 1) ? 0.25 : -0.25;
    return u0 + xo * (sl[0] * alpha) + yo * (sl[1] * alpha) + zo * (sl[2] * alpha);
}

// the coarse part of the two-level ghost fill: fine ghost cells whose source
// cell (after the periodic wrap and the BC map) no fine box owns take the
// interpolated coarse value at that source; runs after the level's own fill,
// which leaves exactly those cells untouched.  Cells the case's own boundary
// function sets (external BCs, the DMR wall override) are left alone too.
__device__ __forceinline__ void fill_cf_ghost(const BoxDesc* __restrict__ fboxes, const GeomDesc& G,
                                              double* __restrict__ Sf, const BoxDesc* __restrict__ cboxes,
                                              const GeomDesc* __restrict__ cgeom, const int* __restrict__ cmap,
                                              const double* __restrict__ Sc, int czuni, const BoxDesc& B, int i,
                                              int j, int k)
{
    if (G.probtype == 1 && j < G.dom_lo[1] && G.bc[0][1][0] == BC_EVEN) {
        // the DMR post-shock wall region x < 1/6 below y-lo is set by the case
        const double x = G.prob_lo[0] + (i + 0.5) * G.dx[0];
        if (x < 1.0 / 6.0) return;
    }
    const long dst = sidx(B, i, j, k);
    int iv[3], outside[3];
    ghost_wrap(G, i, j, k, iv, outside);
    // the source cell is normally the same for every component (the BC maps
    // mirror or clamp alike): its owner and coarse box are looked up once
    int last[3] = {INT_MIN, INT_MIN, INT_MIN};
    int own = -1, m = -1;
    for (int n = 0; n < NC; ++n) {
        int src[3];
        double sign;
        if (!bc_source(G, n, iv, outside, src, sign)) continue;
        if (src[0] != last[0] || src[1] != last[1] || src[2] != last[2]) {
            own = level_owner(G, fboxes, src[0], src[1], src[2]);
            m = own >= 0 ? -1 : interp_box(cboxes, *cgeom, cmap, src[0], src[1], src[2]);
            last[0] = src[0]; last[1] = src[1]; last[2] = src[2];
        }
        if (own >= 0 || m < 0) continue;   // the level fill copied it / nothing to interpolate from
        Sf[dst + n * B.nptg] = sign * interp_value(cboxes[m], *cgeom, Sc, src[0], src[1], src[2], n, czuni);
    }
}

__global__ __launch_bounds__(256) void k_fill_cf(const BoxDesc* __restrict__ fboxes, const GeomDesc* __restrict__ fgeom,
                                                 const int4* __restrict__ tiles, double* __restrict__ Sf,
                                                 const BoxDesc* __restrict__ cboxes, const GeomDesc* __restrict__ cgeom,
                                                 const int* __restrict__ cmap, const double* __restrict__ Sc, int czuni)
{
    const int4 t = tiles[blockIdx.x];
    if ((int)threadIdx.x >= t.z) return;
    const BoxDesc B = fboxes[t.x];
    int i, j, k;
    ghost_cell(B, t.y + threadIdx.x, t.w, i, j, k);
    fill_cf_ghost(fboxes, *fgeom, Sf, cboxes, cgeom, cmap, Sc, czuni, B, i, j, k);
}

// the same over the fine level's list of the ghost cells the coarse/fine fill
// may write (box, ghost index at depth gz)
__global__ __launch_bounds__(256) void k_fill_cf_list(const BoxDesc* __restrict__ fboxes,
                                                      const GeomDesc* __restrict__ fgeom, const int2* __restrict__ list,
                                                      const unsigned* __restrict__ nl, int gz, double* __restrict__ Sf,
                                                      const BoxDesc* __restrict__ cboxes,
                                                      const GeomDesc* __restrict__ cgeom, const int* __restrict__ cmap,
                                                      const double* __restrict__ Sc, int czuni)
{
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= (int)*nl) return;
    const int2 e = list[q];
    const BoxDesc B = fboxes[e.x];
    int i, j, k;
    ghost_cell(B, e.y, gz, i, j, k);
    fill_cf_ghost(fboxes, *fgeom, Sf, cboxes, cgeom, cmap, Sc, czuni, B, i, j, k);
}

// Both ghost fills of a level in one launch, over its two ghost lists: the
// first ca threads take list A (the level fill), the others list B (the
// coarse/fine fill).  The two write disjoint (cell, component) pairs -- the
// level fill a component whose mapped source a level box owns, the
// coarse/fine fill one whose source no box owns, neither an ext_dir component
// the other sets -- so their order does not matter.
__global__ __launch_bounds__(256) void k_fill_both(const BoxDesc* __restrict__ fboxes,
                                                   const GeomDesc* __restrict__ fgeom, const int2* __restrict__ la,
                                                   const unsigned* __restrict__ na, int ca,
                                                   const int2* __restrict__ lb, const unsigned* __restrict__ nb, int gz,
                                                   double* __restrict__ Sf, double time,
                                                   const BoxDesc* __restrict__ cboxes,
                                                   const GeomDesc* __restrict__ cgeom, const int* __restrict__ cmap,
                                                   const double* __restrict__ Sc, int czuni)
{
    int q = blockIdx.x * 256 + threadIdx.x;
    const bool a = q < ca;
    if (!a) q -= ca;
    if (q >= (int)(a ? *na : *nb)) return;
    const int2 e = (a ? la : lb)[q];
    const BoxDesc B = fboxes[e.x];
    int i, j, k;
    ghost_cell(B, e.y, gz, i, j, k);
    if (a) fill_ghost(fboxes, fgeom[B.geom], Sf, time, B, i, j, k);
    else fill_cf_ghost(fboxes, *fgeom, Sf, cboxes, cgeom, cmap, Sc, czuni, B, i, j, k);
}

// valid cells of the fine level from the coarse level (a new fine region);
// cells that the old fine level (when given) covers are skipped
__global__ __launch_bounds__(256) void k_interp_valid(const BoxDesc* __restrict__ fboxes, const int4* __restrict__ tiles,
                                                      double* __restrict__ Sf, const BoxDesc* __restrict__ cboxes,
                                                      const GeomDesc* __restrict__ cgeom, const int* __restrict__ cmap,
                                                      const double* __restrict__ Sc, const BoxDesc* __restrict__ oboxes,
                                                      const GeomDesc* __restrict__ ogeom, int czuni)
{
    const int4 t = tiles[blockIdx.x];
    const BoxDesc B = fboxes[t.x];
    for (int c = threadIdx.x; c < t.z; c += 256) {
        int i, j, k;
        valid_cell(B, t.y + c, i, j, k);
        if (ogeom) {
            const int m = level_owner(*ogeom, oboxes, i, j, k);
            if (m >= 0) {
                const BoxDesc& O = oboxes[m];
                if (i >= O.lo[0] && i <= O.hi[0] && j >= O.lo[1] && j <= O.hi[1] && k >= O.lo[2] && k <= O.hi[2])
                    continue;
            }
        }
        const long dst = sidx(B, i, j, k);
        const int m = interp_box(cboxes, *cgeom, cmap, i, j, k);
        if (m < 0) continue;
        for (int n = 0; n < NC; ++n) Sf[dst + n * B.nptg] = interp_value(cboxes[m], *cgeom, Sc, i, j, k, n, czuni);
    }
}

// valid cells of a new level from an old level of the same resolution where
// the old one covers them (regrid)
__global__ __launch_bounds__(256) void k_copy_level(const BoxDesc* __restrict__ nboxes, const int4* __restrict__ tiles,
                                                    double* __restrict__ Sn, const BoxDesc* __restrict__ oboxes,
                                                    const GeomDesc* __restrict__ ogeom, const double* __restrict__ So)
{
    const int4 t = tiles[blockIdx.x];
    const BoxDesc B = nboxes[t.x];
    for (int c = threadIdx.x; c < t.z; c += 256) {
        int i, j, k;
        valid_cell(B, t.y + c, i, j, k);
        const int m = level_owner(*ogeom, oboxes, i, j, k);
        if (m < 0) continue;
        const BoxDesc O = oboxes[m];
        if (i < O.lo[0] || i > O.hi[0] || j < O.lo[1] || j > O.hi[1] || k < O.lo[2] || k > O.hi[2]) continue;
        const long dst = sidx(B, i, j, k), src = sidx(O, i, j, k);
        for (int n = 0; n < NC; ++n) Sn[dst + n * B.nptg] = So[src + n * O.nptg];
    }
}

// tag_dengrad, Source/Tagging.H:8-36: the larger one-sided density jump in
// any direction >= threshold (ghost cells of S filled); covered boxes untagged
// zuni: a z-uniform (planar) level, whose ghost cells are filled in its first
// two layers only: every cell takes the tag of its column's first-layer cell
// (the same value, as the layers hold the same data)
__global__ __launch_bounds__(256) void k_tag(const BoxDesc* __restrict__ boxes, const int4* __restrict__ tiles,
                                             const double* __restrict__ S, double thr, uint8_t* __restrict__ tags,
                                             int zuni)
{
    const int4 t = tiles[blockIdx.x];
    const BoxDesc B = boxes[t.x];
    for (int c = threadIdx.x; c < t.z; c += 256) {
        int i, j, k;
        valid_cell(B, t.y + c, i, j, k);
        uint8_t tg = 0;
        if (B.type != 2) {
            const long s = sidx(B, i, j, zuni ? B.lo[2] : k);
            const long sy = B.nxg, sz = (long)B.nxg * B.nyg;
            const double r = S[s];
            double ax = fabs(S[s + 1] - r), ay = fabs(S[s + sy] - r), az = fabs(S[s + sz] - r);
            ax = fmax(ax, fabs(r - S[s - 1]));
            ay = fmax(ay, fabs(r - S[s - sy]));
            az = fmax(az, fabs(r - S[s - sz]));
            tg = fmax(ax, fmax(ay, az)) >= thr ? 1 : 0;
        }
        tags[B.off_d / NC + (didx(B, i, j, k) - B.off_d)] = tg;
    }
}

// k_tag's flag in bit 0 and, in bit 1, whether the cell lies in the level's
// proper nesting region: every cell within Chebyshev distance n_proper
// (periodic images included, cells beyond a non-periodic domain side not
// counted) is covered by the level's boxes
__global__ __launch_bounds__(256) void k_tag_nest(const BoxDesc* __restrict__ boxes, const GeomDesc* __restrict__ geom,
                                                  const int4* __restrict__ tiles, const double* __restrict__ S,
                                                  double thr, int n_proper, uint8_t* __restrict__ flags, int zuni)
{
    const int4 t = tiles[blockIdx.x];
    const BoxDesc B = boxes[t.x];
    const GeomDesc& G = *geom;
    for (int c = threadIdx.x; c < t.z; c += 256) {
        int i, j, k;
        valid_cell(B, t.y + c, i, j, k);
        uint8_t tg = 0;
        if (B.type != 2) {
            const long s = sidx(B, i, j, zuni ? B.lo[2] : k);
            const long sy = B.nxg, sz = (long)B.nxg * B.nyg;
            const double r = S[s];
            double ax = fabs(S[s + 1] - r), ay = fabs(S[s + sy] - r), az = fabs(S[s + sz] - r);
            ax = fmax(ax, fabs(r - S[s - 1]));
            ay = fmax(ay, fabs(r - S[s - sy]));
            az = fmax(az, fabs(r - S[s - sz]));
            tg = fmax(ax, fmax(ay, az)) >= thr ? 1 : 0;
        }
        // a neighbourhood inside the cell's own box is covered
        bool in = true;
        const bool own = i - n_proper >= B.lo[0] && i + n_proper <= B.hi[0] && j - n_proper >= B.lo[1] &&
                         j + n_proper <= B.hi[1] && k - n_proper >= B.lo[2] && k + n_proper <= B.hi[2];
        for (int c2 = -n_proper; c2 <= n_proper && in && !own; ++c2)
            for (int c1 = -n_proper; c1 <= n_proper && in; ++c1)
                for (int c0 = -n_proper; c0 <= n_proper && in; ++c0) {
                    int q[3] = {i + c0, j + c1, k + c2};
                    bool inside = true;
                    for (int d = 0; d < 3; ++d) {
                        const int len = G.dom_hi[d] - G.dom_lo[d] + 1;
                        if (q[d] >= G.dom_lo[d] && q[d] <= G.dom_hi[d]) continue;
                        if (G.periodic[d] && len > 1) q[d] = G.dom_lo[d] + (((q[d] - G.dom_lo[d]) % len) + len) % len;
                        else inside = false;
                    }
                    if (inside) in = level_owner(G, boxes, q[0], q[1], q[2]) >= 0;
                }
        flags[B.off_d / NC + (didx(B, i, j, k) - B.off_d)] = (uint8_t)(tg | (in ? 2 : 0));
    }
}

// out = a*x + b*y, the two-time combination of FillPatchSingleLevel
__global__ void k_lincomb(long n, double a, const double* __restrict__ x, double b, const double* __restrict__ y,
                          double* __restrict__ out)
{
#pragma clang fp contract(off)
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) out[i] = a * x[i] + b * y[i];
}

// the same over the valid cells of a level's boxes only (the ghost cells are
// filled afterwards); tiles: (box, first valid cell, count, 0)
__global__ __launch_bounds__(256) void k_lincomb_valid(const BoxDesc* __restrict__ boxes, const int4* __restrict__ tiles,
                                                       double a, const double* __restrict__ x, double b,
                                                       const double* __restrict__ y, double* __restrict__ out)
{
#pragma clang fp contract(off)
    const int4 t = tiles[blockIdx.x];
    const BoxDesc B = boxes[t.x];
    for (int c = threadIdx.x; c < t.z; c += 256) {
        int i, j, k;
        valid_cell(B, t.y + c, i, j, k);
        const long s0 = sidx(B, i, j, k);
        for (int n = 0; n < NC; ++n) {
            const long q = s0 + n * B.nptg;
            out[q] = a * x[q] + b * y[q];
        }
    }
}

} // namespace ocfd

using namespace ocfd;

// ---------------------------------------------------------------------------
// host side: the coarse/fine interface and the C ABI
// ---------------------------------------------------------------------------
struct ocfd_cf_s {
    ocfd_level coarse = nullptr, fine = nullptr;
    int nface = 0, ncell = 0;
    CFFace* d_cf = nullptr;      // coarse faces
    int4* d_cells = nullptr;     // coarse cells next to the fine level (box, i, j, k)
    int* d_cellfaces = nullptr;  // 6 face indices per cell, -1 = none
    double* d_regc = nullptr;    // 5 per coarse face
    double* d_regf = nullptr;    // 5 per fine sub-face
    int4* d_tile_avg = nullptr;  // coarsened fine boxes, 256 coarse cells per tile
    int ntile_avg = 0;
    std::vector<CFFace> h_cf;    // host copy of the coarse faces
    void* d_arena = nullptr;     // one allocation holding the lists and the map below
    int* d_cmap = nullptr;       // coarse cell -> box to interpolate from (owner, else the first box
                                 // whose grown-by-3 region holds it): a tile map of the coarse
                                 // level's tiles, its box lists following the tiles
};

namespace {

int find_box(const LevelDev& L, int i, int j, int k)
{
    const GeomDesc& G = L.geom;
    if (i < G.dom_lo[0] || i > G.dom_hi[0] || j < G.dom_lo[1] || j > G.dom_hi[1] || k < G.dom_lo[2] || k > G.dom_hi[2])
        return -1;
    return tile_owner(L.owner_map.data(), L.owner_list.data(), G.map_n, G.dom_lo, G.map_shift, L.boxes.data(), i, j,
                      k, 0);
}

int wrap(const GeomDesc& G, int d, int v)
{
    if (!G.periodic[d] || (v >= G.dom_lo[d] && v <= G.dom_hi[d])) return v;
    const int len = G.dom_hi[d] - G.dom_lo[d] + 1;
    return G.dom_lo[d] + (((v - G.dom_lo[d]) % len) + len) % len;
}

int launched(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string(what) + ": " + hipGetErrorString(e));
        return OCFD_EHIP;
    }
    return OCFD_OK;
}

void free_cf(ocfd_cf c)
{
    // the interface lists, the interpolation map and the registers live in c->d_arena
    dev_free({c->d_arena});
}

bool refined_by_2(const LevelDev& C, const LevelDev& F)
{
    for (int d = 0; d < 3; ++d)
        if (F.geom.dom_lo[d] != 2 * C.geom.dom_lo[d] || F.geom.dom_hi[d] != 2 * C.geom.dom_hi[d] + 1 ||
            F.geom.periodic[d] != C.geom.periodic[d])
            return false;
    return true;
}

bool aligned_by_2(const LevelDev& F)
{
    for (const BoxDesc& B : F.boxes)
        for (int d = 0; d < 3; ++d)
            if ((B.lo[d] & 1) || !(B.hi[d] & 1)) return false;
    return true;
}

} // namespace

extern "C" {

int ocfd_cf_create(ocfd_level coarse, ocfd_level fine, ocfd_cf* out)
{
    if (!coarse || !fine || !out) {
        set_error("ocfd_cf_create: invalid arguments");
        return OCFD_EINVAL;
    }
    const LevelDev& C = coarse->dev;
    const LevelDev& F = fine->dev;
    if (!refined_by_2(C, F) || !aligned_by_2(F)) {
        set_error("ocfd_cf_create: the fine level must refine the coarse domain by 2 with boxes aligned to 2");
        return OCFD_EINVAL;
    }
    if (C.has_eb || F.has_eb) {
        set_error("ocfd_cf_create: the flux register covers regular levels only");
        return OCFD_EUNSUPPORTED;
    }
    // coverage of the coarse domain by the fine level (a tile map of the
    // coarsened fine boxes), and the uncovered coarse cells next to it (the
    // only ones that can own an interface face)
    const GeomDesc& CG = C.geom;
    std::vector<BoxDesc> fcrs(F.boxes.size());
    for (size_t b = 0; b < F.boxes.size(); ++b)
        for (int d = 0; d < 3; ++d) {
            fcrs[b].lo[d] = F.boxes[b].lo[d] >> 1;
            fcrs[b].hi[d] = F.boxes[b].hi[d] >> 1;
        }
    // (tiles of 4^3 cells: the lookups below all lie at fine-box boundaries,
    // where coarser tiles would hold several boxes each)
    const int fshift = 2;
    std::vector<int> fmap, flist;
    int fmn[3];
    build_tile_map(CG.dom_lo, CG.dom_hi, fshift, fcrs, 0, fmn, fmap, flist);
    // the fine box covering coarse cell (I,J,K) after the periodic wrap (fine
    // boxes are aligned to 2, so it holds all 8 fine cells under it), -1 none,
    // -2 outside a non-periodic domain side
    auto covering = [&](int I, int J, int K) {
        int iv[3] = {I, J, K};
        for (int d = 0; d < 3; ++d) {
            iv[d] = wrap(CG, d, iv[d]);
            if (iv[d] < CG.dom_lo[d] || iv[d] > CG.dom_hi[d]) return -2;
        }
        return tile_owner(fmap.data(), flist.data(), fmn, CG.dom_lo, fshift, fcrs.data(), iv[0], iv[1], iv[2], 0);
    };
    // Interface cells and faces, coarse box by coarse box on the host threads.
    // A candidate is a coarse cell face-adjacent to a coarsened fine box (an
    // edge or corner neighbour of one box that owns a face is face-adjacent
    // to another box): each side of each coarsened fine box gives a one-cell
    // slab (its plane wrapped on a periodic side; none on a non-periodic
    // domain side, nor for a box spanning a periodic direction, whose image
    // is the box itself).  The slabs are binned to the coarse boxes they meet
    // (through the coarse level's tile map); each coarse box then lists the
    // uncovered slab cells it owns in (K, J, I) order, once each, with their
    // faces toward covered neighbours in (direction, side) order, and the
    // lists are joined in coarse-box order: the order a global sort of
    // (coarse box, K, J, I) gives, whatever the number of threads.
    struct Slab {
        int lo[3], hi[3];
    };
    std::vector<Slab> slabs;
    slabs.reserve(6 * fcrs.size());
    for (const BoxDesc& B : fcrs)
        for (int d = 0; d < 3; ++d) {
            const int len = CG.dom_hi[d] - CG.dom_lo[d] + 1;
            if (CG.periodic[d] && B.hi[d] - B.lo[d] + 1 >= len) continue;
            for (int side = 0; side < 2; ++side) {
                const int at = side ? B.hi[d] + 1 : B.lo[d] - 1;
                if (!CG.periodic[d] && (at < CG.dom_lo[d] || at > CG.dom_hi[d])) continue;
                Slab sl;
                for (int e = 0; e < 3; ++e) { sl.lo[e] = B.lo[e]; sl.hi[e] = B.hi[e]; }
                sl.lo[d] = sl.hi[d] = wrap(CG, d, at);
                slabs.push_back(sl);
            }
        }
    // slabs per coarse box, from the tiles of the coarse owner map each meets
    std::vector<std::vector<int>> cslab(C.nbox);
    {
        const int sh = CG.map_shift;
        std::vector<int> seen(C.nbox, -1);
        for (int q = 0; q < (int)slabs.size(); ++q) {
            const Slab& sl = slabs[q];
            int t0[3], t1[3];
            for (int e = 0; e < 3; ++e) {
                t0[e] = (sl.lo[e] - CG.dom_lo[e]) >> sh;
                t1[e] = (sl.hi[e] - CG.dom_lo[e]) >> sh;
            }
            auto take = [&](int cb) {
                if (cb < 0 || seen[cb] == q) return;
                seen[cb] = q;
                cslab[cb].push_back(q);
            };
            for (int tk = t0[2]; tk <= t1[2]; ++tk)
                for (int tj = t0[1]; tj <= t1[1]; ++tj)
                    for (int ti = t0[0]; ti <= t1[0]; ++ti) {
                        const int v = C.owner_map[((long)tk * CG.map_n[1] + tj) * CG.map_n[0] + ti];
                        if (v >= -1) { take(v); continue; }
                        const int* l = C.owner_list.data() + (-v - 2);
                        for (int m = 1; m <= l[0]; ++m) take(l[m]);
                    }
        }
    }
    struct FaceLists {
        std::vector<CFFace> cf;
        std::vector<int4> cells;
        std::vector<int> cellfaces;
    };
    std::vector<FaceLists> part(C.nbox);
    parallel_for(C.nbox, [&](int cb) {
        if (cslab[cb].empty()) return;
        const BoxDesc& A = C.boxes[cb];
        std::vector<std::array<int, 3>> cand;   // (K, J, I)
        for (int q : cslab[cb]) {
            const Slab& sl = slabs[q];
            int a[3], z[3];
            for (int e = 0; e < 3; ++e) {
                a[e] = std::max(sl.lo[e], A.lo[e]);
                z[e] = std::min(sl.hi[e], A.hi[e]);
            }
            for (int K = a[2]; K <= z[2]; ++K)
                for (int J = a[1]; J <= z[1]; ++J)
                    for (int I = a[0]; I <= z[0]; ++I)
                        if (covering(I, J, K) < 0) cand.push_back({K, J, I});
        }
        std::sort(cand.begin(), cand.end());
        cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
        FaceLists& P = part[cb];
        P.cf.reserve(2 * cand.size());
        for (const auto& q : cand) {
            const int K = q[0], J = q[1], I = q[2];
            int fidx[6] = {-1, -1, -1, -1, -1, -1};
            for (int d = 0; d < 3; ++d)
                for (int side = 0; side < 2; ++side) {
                    int nb[3] = {I, J, K};
                    nb[d] += side ? 1 : -1;
                    const int fb = covering(nb[0], nb[1], nb[2]);
                    if (fb < 0) continue;
                    CFFace f{};
                    f.box = cb; f.i = I; f.j = J; f.k = K; f.dir = d; f.up = side;
                    f.fbox = fb;   // covers the coarse neighbour; its 4 sub-faces are found on the device
                    fidx[2 * d + side] = (int)P.cf.size();   // local: offset when the parts are joined
                    P.cf.push_back(f);
                }
            bool any = false;
            for (int m = 0; m < 6; ++m) any = any || fidx[m] >= 0;
            if (any) {
                P.cells.push_back(make_int4(cb, I, J, K));
                P.cellfaces.insert(P.cellfaces.end(), fidx, fidx + 6);
            }
        }
    });
    std::vector<CFFace> cf;
    std::vector<int4> cells;
    std::vector<int> cellfaces;
    {
        size_t nf = 0, nc = 0;
        for (const FaceLists& P : part) { nf += P.cf.size(); nc += P.cells.size(); }
        cf.reserve(nf); cells.reserve(nc); cellfaces.reserve(6 * nc);
    }
    for (int cb = 0; cb < C.nbox; ++cb) {
        const int off = (int)cf.size();
        for (int& v : part[cb].cellfaces)
            if (v >= 0) v += off;
        cf.insert(cf.end(), part[cb].cf.begin(), part[cb].cf.end());
        cells.insert(cells.end(), part[cb].cells.begin(), part[cb].cells.end());
        cellfaces.insert(cellfaces.end(), part[cb].cellfaces.begin(), part[cb].cellfaces.end());
    }
    std::vector<int4> ta;
    for (int b = 0; b < F.nbox; ++b) {
        const BoxDesc& B = F.boxes[b];
        const long n = (long)((B.hi[0] - B.lo[0] + 1) / 2) * ((B.hi[1] - B.lo[1] + 1) / 2) * ((B.hi[2] - B.lo[2] + 1) / 2);
        for (long s = 0; s < n; s += 256) ta.push_back(make_int4(b, (int)s, (int)std::min<long>(256, n - s), 0));
    }
    // coarse cell -> box to interpolate from: the owner, else the first box
    // whose grown-by-3 region holds the cell (its 27 neighbours then lie in
    // that box's filled ghost region)
    std::vector<int> cmap, clist;
    int cmn[3];
    build_tile_map(CG.dom_lo, CG.dom_hi, CG.map_shift, C.boxes, 3, cmn, cmap, clist);
    cmap.insert(cmap.end(), clist.begin(), clist.end());
    // the fine level must lie in the coarse level's valid cells: the coarse
    // boxes are disjoint, so a coarsened fine box is covered when its
    // intersections with them add up to its volume
    for (const BoxDesc& B : fcrs) {
        long vol = 1, got = 0;
        for (int d = 0; d < 3; ++d) vol *= B.hi[d] - B.lo[d] + 1;
        for (const BoxDesc& A : C.boxes) {
            long v = 1;
            for (int d = 0; d < 3 && v > 0; ++d) v *= std::max(0, std::min(A.hi[d], B.hi[d]) - std::max(A.lo[d], B.lo[d]) + 1);
            got += v;
        }
        if (got != vol) {
            set_error("ocfd_cf_create: the fine level is not nested in the coarse level");
            return OCFD_EINVAL;
        }
    }
    ocfd_cf c = new ocfd_cf_s;
    c->coarse = coarse; c->fine = fine;
    c->nface = (int)cf.size(); c->ncell = (int)cells.size();
    c->h_cf = cf;
    c->ntile_avg = (int)ta.size();
    int rc;
    Staging stg;
    stg.add(&c->d_cf, cf); stg.add(&c->d_cells, cells); stg.add(&c->d_cellfaces, cellfaces);
    stg.add(&c->d_tile_avg, ta); stg.add(&c->d_cmap, cmap);
    stg.reserve(&c->d_regc, 5 * (size_t)c->nface);
    stg.reserve(&c->d_regf, 20 * (size_t)c->nface);
    if ((rc = stg.commit(&c->d_arena))) {
        set_error("ocfd_cf_create: device allocation failed");
        free_cf(c);
        delete c;
        return rc;
    }
    *out = c;
    return OCFD_OK;
}

int ocfd_cf_destroy(ocfd_cf c)
{
    if (!c) return OCFD_OK;
    free_cf(c);
    delete c;
    return OCFD_OK;
}

int ocfd_cf_nfaces(ocfd_cf c) { return c ? c->nface : -1; }

int ocfd_cf_faces(ocfd_cf c, int* out)
{
    if (!c || !out) {
        set_error("ocfd_cf_faces: invalid arguments");
        return OCFD_EINVAL;
    }
    for (int f = 0; f < c->nface; ++f) {
        const CFFace& x = c->h_cf[f];
        const int v[6] = {x.box, x.i, x.j, x.k, x.dir, x.up};
        std::copy(v, v + 6, out + 6 * f);
    }
    return OCFD_OK;
}

int ocfd_cf_reset(ocfd_cf c, void* stream)
{
    if (!c) {
        set_error("ocfd_cf_reset: null register");
        return OCFD_EINVAL;
    }
    if (c->nface) {
        hipLaunchKernelGGL(k_zero, dim3(std::min(1024, (c->nface * 5 + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                           c->d_regc, (long)c->nface * 5);
        hipLaunchKernelGGL(k_zero, dim3(std::min(1024, (c->nface * 20 + 255) / 256)), dim3(256), 0,
                           (hipStream_t)stream, c->d_regf, (long)c->nface * 20);
    }
    return launched("ocfd_cf_reset");
}

int ocfd_cf_add(ocfd_cf c, int fine_side, double w_dt, const ocfd_parm* parm, int flags, void* stream)
{
    if (!c || !parm || (fine_side != 0 && fine_side != 1)) {
        set_error("ocfd_cf_add: invalid arguments");
        return OCFD_EINVAL;
    }
    if (!c->nface) return OCFD_OK;
    const LevelDev& L = fine_side ? c->fine->dev : c->coarse->dev;
    const CFFace* faces = c->d_cf;   // the fine side derives its 4 sub-faces per face
    const GeomDesc* fg = c->fine->dev.d_geom;
    double* reg = fine_side ? c->d_regf : c->d_regc;
    const int n = fine_side ? 4 * c->nface : c->nface;
    const PhysArgs ph = phys_of(*parm);
    const bool f32 = (flags & OCFD_FLUX_FP32) != 0, visc = (flags & OCFD_VISC) != 0;
    const void* prim = f32 ? (const void*)L.d_prim32 : (const void*)L.d_prim64;
    const dim3 g((n + 127) / 128), blk(128);
    const int zrep = (flags & OCFD_PLANAR) && L.fill_gz == 1;   // as the level's planar sweeps
    hipStream_t s = (hipStream_t)stream;
#define OCFD_CF(FT, V)                                                                                   \
    do {                                                                                                 \
        if (fine_side)                                                                                   \
            hipLaunchKernelGGL((k_cf_flux<FT, V, true>), g, blk, 0, s, L.d_boxes, faces, n, prim, ph, w_dt, reg, \
                               zrep, fg);                                                                \
        else                                                                                             \
            hipLaunchKernelGGL((k_cf_flux<FT, V, false>), g, blk, 0, s, L.d_boxes, faces, n, prim, ph, w_dt, reg, \
                               zrep, fg);                                                                \
    } while (0)
    if (f32) {
        if (visc) OCFD_CF(float, true); else OCFD_CF(float, false);
    } else {
        if (visc) OCFD_CF(double, true); else OCFD_CF(double, false);
    }
#undef OCFD_CF
    return launched("ocfd_cf_add");
}

int ocfd_compute_dsdt_fr(ocfd_level lev, const double* S, double* dSdt, double dt, ocfd_cf fine, ocfd_cf current,
                         const ocfd_parm* parm, int flags, void* stream)
{
    // compute_dSdt(S, dSdt, dt, fine, current), Advance.cpp:182 and 344-360:
    // `fine` joins this level to the next finer one (this level is its coarse
    // side: CrseAdd), `current` joins it to the next coarser one (fine side:
    // FineAdd); dt is the stage's weight
    if (!lev || (fine && fine->coarse != lev) || (current && current->fine != lev)) {
        set_error("ocfd_compute_dsdt_fr: a register does not border this level on the named side");
        return OCFD_EINVAL;
    }
    int rc = ocfd_compute_dsdt(lev, S, dSdt, dt, parm, flags, stream);
    if (!rc && fine) rc = ocfd_cf_add(fine, 0, dt, parm, flags, stream);
    if (!rc && current) rc = ocfd_cf_add(current, 1, dt, parm, flags, stream);
    return rc;
}

int ocfd_cf_registers(ocfd_cf c, double* regc, double* regf)
{
    if (!c || !regc || !regf) {
        set_error("ocfd_cf_registers: invalid arguments");
        return OCFD_EINVAL;
    }
    if (c->nface && (hipMemcpy(regc, c->d_regc, sizeof(double) * 5 * c->nface, hipMemcpyDeviceToHost) != hipSuccess ||
                     hipMemcpy(regf, c->d_regf, sizeof(double) * 20 * c->nface, hipMemcpyDeviceToHost) != hipSuccess)) {
        set_error("ocfd_cf_registers: copy failed");
        return OCFD_EHIP;
    }
    return OCFD_OK;
}

int ocfd_cf_reflux(ocfd_cf c, double* S_coarse, void* stream)
{
    if (!c || !S_coarse) {
        set_error("ocfd_cf_reflux: invalid arguments");
        return OCFD_EINVAL;
    }
    tcache_clear(c->coarse->dev);
    if (c->ncell)
        hipLaunchKernelGGL(k_reflux, dim3((c->ncell + 127) / 128), dim3(128), 0, (hipStream_t)stream,
                           c->coarse->dev.d_boxes, c->d_cells, c->d_cellfaces, c->ncell, c->d_cf, c->d_regc, c->d_regf,
                           S_coarse);
    return launched("ocfd_cf_reflux");
}

int ocfd_cf_avg_down(ocfd_cf c, const double* S_fine, double* S_coarse, int ncomp, void* stream)
{
    if (!c || !S_fine || !S_coarse || ncomp <= 0) {
        set_error("ocfd_cf_avg_down: invalid arguments");
        return OCFD_EINVAL;
    }
    tcache_clear(c->coarse->dev);
    if (c->ntile_avg)
        hipLaunchKernelGGL(k_avg_down, dim3(c->ntile_avg), dim3(256), 0, (hipStream_t)stream, c->fine->dev.d_boxes,
                           c->coarse->dev.d_boxes, c->coarse->dev.d_geom, c->d_tile_avg, S_fine, S_coarse, ncomp);
    return launched("ocfd_cf_avg_down");
}

int ocfd_cf_fill_fine(ocfd_cf c, double* S_fine, const double* S_coarse, void* stream)
{
    if (!c || !S_fine || !S_coarse) {
        set_error("ocfd_cf_fill_fine: invalid arguments");
        return OCFD_EINVAL;
    }
    const LevelDev& F = c->fine->dev;
    const LevelDev& C = c->coarse->dev;
    const int czuni = C.fill_gz == 1;
    const auto* gl = ghost_list(F, 1, (hipStream_t)stream);
    if (gl) {   // the fine level's ghost cells without a same-level source
        if (gl->cap)
            hipLaunchKernelGGL(k_fill_cf_list, dim3((gl->cap + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                               F.d_boxes, F.d_geom, gl->d, gl->cnt, F.fill_gz, S_fine, C.d_boxes, C.d_geom, c->d_cmap,
                               S_coarse, czuni);
    } else {
        int ng;
        const int4* tg = fill_tiles(F, ng);
        if (ng)
            hipLaunchKernelGGL(k_fill_cf, dim3(ng), dim3(256), 0, (hipStream_t)stream, F.d_boxes, F.d_geom, tg,
                               S_fine, C.d_boxes, C.d_geom, c->d_cmap, S_coarse, czuni);
    }
    return launched("ocfd_cf_fill_fine");
}

int ocfd_cf_interp_fine(ocfd_cf c, double* S_fine, const double* S_coarse, ocfd_level old_fine, void* stream)
{
    if (!c || !S_fine || !S_coarse) {
        set_error("ocfd_cf_interp_fine: invalid arguments");
        return OCFD_EINVAL;
    }
    const LevelDev& F = c->fine->dev;
    const LevelDev& C = c->coarse->dev;
    if (old_fine) {
        for (int d = 0; d < 3; ++d)
            if (old_fine->dev.geom.dom_lo[d] != F.geom.dom_lo[d] || old_fine->dev.geom.dom_hi[d] != F.geom.dom_hi[d]) {
                set_error("ocfd_cf_interp_fine: the old fine level must share the fine domain");
                return OCFD_EINVAL;
            }
    }
    tcache_clear(c->fine->dev);
    if (F.ntile_c1)
        hipLaunchKernelGGL(k_interp_valid, dim3(F.ntile_c1), dim3(256), 0, (hipStream_t)stream, F.d_boxes, F.d_tile_c1,
                           S_fine, C.d_boxes, C.d_geom, c->d_cmap, S_coarse, old_fine ? old_fine->dev.d_boxes : nullptr,
                           old_fine ? old_fine->dev.d_geom : nullptr, (int)(C.fill_gz == 1));
    return launched("ocfd_cf_interp_fine");
}

int ocfd_copy_level(ocfd_level dst, double* S_dst, ocfd_level src, const double* S_src, void* stream)
{
    if (!dst || !src || !S_dst || !S_src) {
        set_error("ocfd_copy_level: invalid arguments");
        return OCFD_EINVAL;
    }
    const LevelDev& N = dst->dev;
    const LevelDev& O = src->dev;
    for (int d = 0; d < 3; ++d)
        if (N.geom.dom_lo[d] != O.geom.dom_lo[d] || N.geom.dom_hi[d] != O.geom.dom_hi[d]) {
            set_error("ocfd_copy_level: the levels must share the domain");
            return OCFD_EINVAL;
        }
    tcache_clear(dst->dev);
    if (N.ntile_c1)
        hipLaunchKernelGGL(k_copy_level, dim3(N.ntile_c1), dim3(256), 0, (hipStream_t)stream, N.d_boxes, N.d_tile_c1,
                           S_dst, O.d_boxes, O.d_geom, S_src);
    return launched("ocfd_copy_level");
}

int ocfd_tag_dengrad(ocfd_level lev, const double* S, double threshold, unsigned char* tags, void* stream)
{
    if (!lev || !S || !tags) {
        set_error("ocfd_tag_dengrad: invalid arguments");
        return OCFD_EINVAL;
    }
    const LevelDev& L = lev->dev;
    if (L.ntile_c1)
        hipLaunchKernelGGL(k_tag, dim3(L.ntile_c1), dim3(256), 0, (hipStream_t)stream, L.d_boxes, L.d_tile_c1, S,
                           threshold, (uint8_t*)tags, (int)(L.fill_gz == 1));
    return launched("ocfd_tag_dengrad");
}

int ocfd_tag_dengrad_nest(ocfd_level lev, const double* S, double threshold, int n_proper, unsigned char* flags,
                          void* stream)
{
    if (!lev || !S || !flags || n_proper < 0) {
        set_error("ocfd_tag_dengrad_nest: invalid arguments");
        return OCFD_EINVAL;
    }
    const LevelDev& L = lev->dev;
    if (L.ntile_c1)
        hipLaunchKernelGGL(k_tag_nest, dim3(L.ntile_c1), dim3(256), 0, (hipStream_t)stream, L.d_boxes, L.d_geom,
                           L.d_tile_c1, S, threshold, n_proper, (uint8_t*)flags, (int)(L.fill_gz == 1));
    return launched("ocfd_tag_dengrad_nest");
}

int ocfd_lincomb(long n, double a, const double* x, double b, const double* y, double* out, void* stream)
{
    if (n < 0 || (n > 0 && (!x || !y || !out))) {
        set_error("ocfd_lincomb: invalid arguments");
        return OCFD_EINVAL;
    }
    if (n)
        hipLaunchKernelGGL(k_lincomb, dim3((unsigned)std::min<long>(4096, (n + 255) / 256)), dim3(256), 0,
                           (hipStream_t)stream, n, a, x, b, y, out);
    return launched("ocfd_lincomb");
}

} // extern "C"

namespace {

// (stage time offset, register weight) in units of dt, per scheme
const double kStageT[3][3] = {{0.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 1.0, 0.5}};
const double kStageW[3][3] = {{1.0, 0.0, 0.0}, {0.5, 0.5, 0.0}, {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0}};

// stage s of the scheme: out = (c_cur*cur + c_old*old) + c_dt*dt*L(cur)
void stage_coefs(int scheme, int s, double& c_cur, double& c_old, double& c_dt)
{
    if (scheme == 0) { c_cur = 1.0; c_old = 0.0; c_dt = 1.0; return; }
    if (scheme == 1) {
        if (s == 0) { c_cur = 1.0; c_old = 0.0; c_dt = 1.0; }
        else { c_cur = 0.5; c_old = 0.5; c_dt = 0.5; }
        return;
    }
    if (s == 0) { c_cur = 1.0; c_old = 0.0; c_dt = 1.0; }
    else if (s == 1) { c_cur = 0.25; c_old = 0.75; c_dt = 0.25; }
    else { c_cur = 2.0 / 3.0; c_old = 1.0 / 3.0; c_dt = 2.0 / 3.0; }
}

} // namespace

namespace {

// time-combined states a level keeps for its finer levels' stages: the
// distinct stage times of one coarse step (up to 8 for two finer levels of
// SSP-RK3 with subcycling)
constexpr int kTimeStates = 8;

// A hierarchy during one level advance: fills stage ghosts of level l from the
// level itself and, recursively, from the coarser levels at the same time.
struct Hier {
    int nlev;
    const ocfd_level* levels;
    const ocfd_cf* cfs;
    ocfd_amr_level_state* st;
    hipStream_t s;

    // ghost cells of X (a state of level l) at time ts
    int fill(int l, double* X, double ts)
    {
        int rc;
        if (l == 0) return ocfd_fill_boundary(levels[l], X, ts, s);
        // the coarser state at ts first (it touches only coarser levels), then
        // both fills of this level: one launch over its ghost lists when it
        // keeps them, else the level fill and the coarse/fine fill
        const double* C = nullptr;
        if ((rc = at(l - 1, ts, C))) return rc;
        const LevelDev& F = levels[l]->dev;
        const LevelDev::GhostList* ga = ghost_list(F, 0, s);
        const LevelDev::GhostList* gb = ga ? ghost_list(F, 1, s) : nullptr;
        if (ga && gb) {
            if (ga->cap + gb->cap > 0)
                hipLaunchKernelGGL(k_fill_both, dim3((ga->cap + gb->cap + 255) / 256), dim3(256), 0, s, F.d_boxes,
                                   F.d_geom, ga->d, ga->cnt, ga->cap, gb->d, gb->cnt, F.fill_gz, X, ts,
                                   levels[l - 1]->dev.d_boxes, levels[l - 1]->dev.d_geom, cfs[l]->d_cmap, C,
                                   levels[l - 1]->dev.fill_gz == 1);
            return launched("ocfd_amr fill");
        }
        if ((rc = ocfd_fill_boundary(levels[l], X, ts, s))) return rc;
        return ocfd_cf_fill_fine(cfs[l], X, C, s);
    }

    // state of level m at time ts (between its t_old and t_new) with filled ghosts
    int at(int m, double ts, const double*& out)
    {
        ocfd_amr_level_state& q = st[m];
        const double span = q.t_new - q.t_old;
        int rc;
        if (std::fabs(ts - q.t_new) <= 1e-10 * span) {   // also a level not advanced yet (span 0)
            if ((rc = fill(m, q.cur, q.t_new))) return rc;
            out = q.cur;
            return OCFD_OK;
        }
        if (std::fabs(ts - q.t_old) <= 1e-10 * span) {
            out = q.old;        // ghosts filled at t_old by the level's first stage
            return OCFD_OK;
        }
        // valid cells combined in time, then the ghost cells filled at ts; the
        // result is kept for the other finer stages at the same time (the
        // stages of a level's second substep start where the first one ended,
        // and the finest level's stage times repeat for every coarser level)
        LevelDev& L = levels[m]->dev;
        for (auto& e : L.tcache)
            if (e.valid && e.ts == ts && e.t_old == q.t_old && e.t_new == q.t_new && e.old == q.old &&
                e.cur == q.cur) {
                e.use = ++L.tcache_clock;
                out = e.buf;
                return OCFD_OK;
            }
        double* dst = q.tmp;
        LevelDev::TimeState* slot = nullptr;
        // at most kTimeStates entries and 2 GiB per level
        const size_t bytes = sizeof(double) * (size_t)L.mf_state;
        const int cap = (int)std::min<size_t>(kTimeStates, ((size_t)2 << 30) / std::max<size_t>(bytes, 1));
        if ((int)L.tcache.size() < cap) {
            void* b = nullptr;
            if (dev_alloc(&b, bytes)) {
                L.tcache.emplace_back();
                L.tcache.back().buf = (double*)b;
                slot = &L.tcache.back();
            }
        }
        if (!slot && !L.tcache.empty()) {   // the least recently used entry
            slot = &L.tcache[0];
            for (auto& e : L.tcache)
                if (e.use < slot->use) slot = &e;
        }
        if (slot) {
            slot->valid = false;
            dst = slot->buf;
        }
        if (L.ntile_c1)
            hipLaunchKernelGGL(k_lincomb_valid, dim3(L.ntile_c1), dim3(256), 0, s, L.d_boxes, L.d_tile_c1,
                               (q.t_new - ts) / span, q.old, (ts - q.t_old) / span, q.cur, dst);
        if ((rc = launched("ocfd_amr_advance_level"))) return rc;
        if ((rc = fill(m, dst, ts))) return rc;
        if (slot) {
            slot->ts = ts; slot->t_old = q.t_old; slot->t_new = q.t_new; slot->old = q.old; slot->cur = q.cur;
            slot->valid = true;
            slot->use = ++L.tcache_clock;
        }
        out = dst;
        return OCFD_OK;
    }
};

} // namespace

extern "C" int ocfd_amr_advance_level(int nlev, const ocfd_level* levels, const ocfd_cf* cfs, ocfd_amr_level_state* st,
                                      int l, double t, double dt, int scheme, const ocfd_parm* parm, int flags,
                                      int do_reflux, void* stream)
{
    if (nlev < 1 || !levels || !st || l < 0 || l >= nlev || !parm || scheme < 0 || scheme > 2 || !(dt > 0.0) ||
        (nlev > 1 && !cfs)) {
        set_error("ocfd_amr_advance_level: invalid arguments");
        return OCFD_EINVAL;
    }
    for (int m = 0; m <= std::min(l + 1, nlev - 1); ++m) {
        const ocfd_amr_level_state& q = st[m];
        if (!levels[m] || !q.old || !q.cur || (m == l && (!q.s1 || (scheme == 2 && !q.s2))) ||
            (m < l && !q.tmp) || (m > 0 && (!cfs[m] || cfs[m]->coarse != levels[m - 1] || cfs[m]->fine != levels[m]))) {
            set_error("ocfd_amr_advance_level: level " + std::to_string(m) + " is incomplete");
            return OCFD_EINVAL;
        }
    }
    hipStream_t s = (hipStream_t)stream;
    Hier h{nlev, levels, cfs, st, s};
    ocfd_amr_level_state& q = st[l];
    tcache_clear(levels[l]->dev);   // the level's states change now
    std::swap(q.old, q.cur);
    q.t_old = t;
    q.t_new = t + dt;
    ocfd_cf reg_c = (do_reflux && l + 1 < nlev) ? cfs[l + 1] : nullptr;   // this level as the coarse side
    ocfd_cf reg_f = (do_reflux && l > 0) ? cfs[l] : nullptr;              // this level as the fine side
    int rc;
    if (reg_c && (rc = ocfd_cf_reset(reg_c, s))) return rc;
    const int ns = scheme + 1;
    double* in[3] = {q.old, q.s1, q.s2};
    double* dst[3] = {q.s1, q.s2, q.cur};
    dst[ns - 1] = q.cur;
    for (int k = 0; k < ns; ++k) {
        if ((rc = h.fill(l, in[k], t + kStageT[scheme][k] * dt))) return rc;
        double c_cur, c_old, c_dt;
        stage_coefs(scheme, k, c_cur, c_old, c_dt);
        if ((rc = ocfd_rk_stage(levels[l], in[k], k ? q.old : nullptr, dst[k], c_cur, c_old, c_dt * dt, parm, flags, s)))
            return rc;
        if (reg_c && (rc = ocfd_cf_add(reg_c, 0, kStageW[scheme][k] * dt, parm, flags, s))) return rc;
        if (reg_f && (rc = ocfd_cf_add(reg_f, 1, kStageW[scheme][k] * dt, parm, flags, s))) return rc;
    }
    return OCFD_OK;
}

extern "C" int ocfd_amr_fill_level(int nlev, const ocfd_level* levels, const ocfd_cf* cfs, ocfd_amr_level_state* st,
                                   int l, double* X, double t, void* stream)
{
    if (nlev < 1 || !levels || !st || l < 0 || l >= nlev || !X || (l > 0 && !cfs)) {
        set_error("ocfd_amr_fill_level: invalid arguments");
        return OCFD_EINVAL;
    }
    for (int m = 0; m < l; ++m)
        if (!levels[m] || !st[m].cur || !st[m].old || !st[m].tmp || !cfs[m + 1]) {
            set_error("ocfd_amr_fill_level: coarser level " + std::to_string(m) + " is incomplete");
            return OCFD_EINVAL;
        }
    Hier h{nlev, levels, cfs, st, (hipStream_t)stream};
    return h.fill(l, X, t);
}
