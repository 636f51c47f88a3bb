// Device-side building blocks of the per-box advance: box/geometry
// descriptors and the point kernels (primitive recovery, Steger-Warming split,
// WENO-Z5, viscous face flux, stage update), written for CDNA4.
//
// FP64 forms follow the reference's operation order term by term, so the FP64
// path agrees with the reference arithmetic to rounding (FMA contraction is
// the only difference).  The FT=float forms are the FP32-flux variant: the
// same formulas in single precision with hardware reciprocals in the WENO
// weights; primitive recovery stays FP64 so pressure does not suffer the
// E - 1/2 rho u^2 cancellation in single precision.
#pragma once

#include <hip/hip_runtime.h>
#include <climits>
#include <cmath>
#include <cstdint>
#include "ocfd.h"

namespace ocfd {

constexpr int NG = OCFD_NGROW;
constexpr int NC = OCFD_NCONS;
enum { URHO = 0, UMX = 1, UMY = 2, UMZ = 3, UEDEN = 4 };

// BC types of the ghost fill (amrex::BCType meanings)
enum : signed char { BC_INT = 0, BC_EVEN = 1, BC_FOEXT = 2, BC_EXT = 3, BC_ODD = -1 };

struct BoxDesc {
    int lo[3], hi[3];     // valid box, global cell indices
    int nxg, nyg, nzg;    // extents of the grown (4-ghost) FAB
    int geom;             // index into the geometry table
    long off_s;           // offset (doubles) of this box's grown FAB in a state MultiFab
    long off_d;           // offset (doubles) of this box's valid FAB in a dSdt MultiFab
    long nptg;            // component stride of the grown FAB
    long nptv;            // component stride of the valid FAB
    long off_p;           // offset (elements) of this box's primitive FAB (6 comps, grown layout)
    long off_e;           // cut boxes: offset (grown-FAB units) in the embedded-boundary arrays, else -1
    int type;             // 0 regular, 1 cut (embedded-boundary path), 2 covered
    int pad;
    double dxinv[3];
    double dx[3];
};

struct GeomDesc {
    int dom_lo[3], dom_hi[3];
    int periodic[3];
    signed char bc[NC][3][2];   // BC type per component, direction, side (0 lo, 1 hi)
    int probtype;
    double prob_lo[3], dx[3];
    double left[NC], right[NC];
    double spec_in[OCFD_NSPEC];  // species written on ext_dir sides (has_spec_in), e.g. an inflow mixture
    int has_spec_in;
    int map_shift;              // log2 of the owner-map tile edge
    int map_n[3];               // owner-map tiles per direction
    const int* map;             // tile -> the box wholly covering it (>= 0), -1 none, <= -2 a list:
    const int* map_list;        //   at -(v+2): count, then the boxes meeting the tile, ascending
};

// Box of a tile map holding cell (i,j,k) of the domain: the box whose valid
// region contains it, else (grow > 0) the first box whose region grown by
// `grow` contains it; -1 if none.  Host and device.
__host__ __device__ inline int tile_owner(const int* map, const int* list, const int* mn, const int* dlo, int shift,
                                          const BoxDesc* boxes, int i, int j, int k, int grow)
{
    const int v = map[((long)((k - dlo[2]) >> shift) * mn[1] + ((j - dlo[1]) >> shift)) * mn[0] +
                      ((i - dlo[0]) >> shift)];
    if (v >= -1) return v;
    const int* l = list + (-v - 2);
    for (int q = 1; q <= l[0]; ++q) {
        const BoxDesc& B = boxes[l[q]];
        if (i >= B.lo[0] && i <= B.hi[0] && j >= B.lo[1] && j <= B.hi[1] && k >= B.lo[2] && k <= B.hi[2]) return l[q];
    }
    if (grow > 0)
        for (int q = 1; q <= l[0]; ++q) {
            const BoxDesc& B = boxes[l[q]];
            if (i >= B.lo[0] - grow && i <= B.hi[0] + grow && j >= B.lo[1] - grow && j <= B.hi[1] + grow &&
                k >= B.lo[2] - grow && k <= B.hi[2] + grow)
                return l[q];
        }
    return -1;
}

// owner box of a cell inside the level's domain, -1 if not covered
__host__ __device__ inline int level_owner(const GeomDesc& G, const BoxDesc* boxes, int i, int j, int k)
{
    return tile_owner(G.map, G.map_list, G.map_n, G.dom_lo, G.map_shift, boxes, i, j, k, 0);
}

// Parm members the kernels use, by value in the launch arguments
struct PhysArgs {
    double gamma, smallr, smallp, Rg, C_s, T_s, kOverMu, twoThirds;
    // constants of the Steger-Warming split, 1/(2 gamma) and
    // (3 - gamma)/(2 (gamma - 1)), rounded as the split would form them in
    // FP64 and in FP32 (the host's IEEE division gives the same bits), and 1/Rg
    double sw_t1, sw_t3;
    float sw_t1f, sw_t3f;
    double Rinv;
};

template <class FT> __device__ __forceinline__ FT sw_t1(const PhysArgs& ph);
template <> __device__ __forceinline__ double sw_t1<double>(const PhysArgs& ph) { return ph.sw_t1; }
template <> __device__ __forceinline__ float sw_t1<float>(const PhysArgs& ph) { return ph.sw_t1f; }
template <class FT> __device__ __forceinline__ FT sw_t3(const PhysArgs& ph);
template <> __device__ __forceinline__ double sw_t3<double>(const PhysArgs& ph) { return ph.sw_t3; }
template <> __device__ __forceinline__ float sw_t3<float>(const PhysArgs& ph) { return ph.sw_t3f; }

struct StageArgs {
    const void* prim;     // primitives of scur (FT, 6 comps r,u,v,w,c,T, grown layout)
    const double* scur;   // state being differentiated (4 ghosts, filled)
    const double* sold;   // U^n for the stage combination, or null
    double* snew;         // stage output (valid cells), or null for a plain dSdt
    void* part;           // partial divergence dx-term (+dy-term), FT, valid layout, level-owned
    double* acc;          // dSdt output of a plain compute_dSdt (valid layout)
    double c_cur, c_old, c_dt;  // with ctl: c_dt multiplies ctl->dt
    int flags;
    PhysArgs ph;
    const ocfd_step_ctl* ctl;    // device time control, or null (host-driven)
    int zrep;                    // z-uniform level: sweep its first z layer only and copy the result to the others
    // CFL reduction folded into the stage that finishes a time step (null:
    // none): atomic minimum of the cells' dt bit patterns, non-finite flag
    unsigned long long* red;
    unsigned int* bad;
};

// CFL time step of one cell, min_d dx_d/(|u_d| + c) (ebr_estdt,
// Source/Advance.cpp:10-52), in the reference's operation order without FMA
// contraction, so every kernel that reduces it gives the same bits.  With
// dx = dy = dz one division serves: rounded division by a positive number is
// monotone, so min_d dx/a_d equals dx/max_d a_d exactly.
__device__ __forceinline__ double cell_dt(double rho, double mx, double my, double mz, double E, const PhysArgs& ph,
                                          const double* dx, bool dx_equal)
{
#pragma clang fp contract(off)
    const double rhoinv = 1.0 / fmax(rho, ph.smallr);
    const double vx = mx * rhoinv;
    const double vy = my * rhoinv;
    const double vz = mz * rhoinv;
    const double p = fmax((ph.gamma - 1.0) * (E - 0.5 * rho * (vx * vx + vy * vy + vz * vz)), ph.smallp);
    const double cs = sqrt(ph.gamma * p * rhoinv);
    const double ax = fabs(vx) + cs, ay = fabs(vy) + cs, az = fabs(vz) + cs;
    if (dx_equal) return dx[0] / fmax(ax, fmax(ay, az));
    return fmin(dx[0] / ax, fmin(dx[1] / ay, dx[2] / az));
}

// a cell whose conserved state holds a NaN or an inf
__device__ __forceinline__ bool cell_nonfinite(double rho, double mx, double my, double mz, double E)
{
    return !__builtin_isfinite(rho + mx + my + mz + E);
}

__device__ __forceinline__ bool box_dx_equal(const BoxDesc& B) { return B.dx[0] == B.dx[1] && B.dx[1] == B.dx[2]; }

// The control block's fields are written by single-thread kernels with plain
// stores and read with plain loads by the following kernels, like any other
// buffer; its reduction slot sits on a separate cache line and is touched only
// by device-scope atomics (a plain load of a line that an atomic updated may
// hit a stale copy in the reading XCD's L2).
__device__ __forceinline__ bool ctl_done(const ocfd_step_ctl* c) { return c && c->done; }

// Workgroups are dealt round-robin over the 8 XCDs (workgroup b runs on XCD
// b % 8), each with its own L2.  Give every XCD a contiguous range of the
// work-list instead, so tiles that share halo cells share an L2.
__device__ __forceinline__ int xcd_tile(int b, int n)
{
    const int q = n >> 3, r = n & 7, x = b & 7;
    return x * q + min(x, r) + (b >> 3);
}

__device__ __forceinline__ long sidx(const BoxDesc& B, int i, int j, int k)
{
    return B.off_s + ((long)(k - B.lo[2] + NG) * B.nyg + (j - B.lo[1] + NG)) * B.nxg + (i - B.lo[0] + NG);
}

__device__ __forceinline__ long didx(const BoxDesc& B, int i, int j, int k)
{
    const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1;
    return B.off_d + ((long)(k - B.lo[2]) * ny + (j - B.lo[1])) * nx + (i - B.lo[0]);
}

// Ghost cells of box B, numbered: z slabs gz deep (full grown x-y extent),
// then y slabs, then x slabs, over the valid layers.  gz = NG: every ghost
// cell.  gz = 1: the ghost region of a z-uniform (planar) level, whose sweeps
// read the first z layer and its two neighbours only: the z slab below, the y
// and x slabs of the first two layers, and the z slab above only when the box
// has one layer.
__host__ __device__ inline int ghost_zslabs(const BoxDesc& B, int gz)
{
    return (gz == 1 && B.hi[2] > B.lo[2]) ? 1 : 2;
}

__host__ __device__ inline int ghost_layers(const BoxDesc& B, int gz)
{
    const int nz = B.hi[2] - B.lo[2] + 1;
    return gz == 1 ? (nz < 2 ? nz : 2) : nz;
}

__host__ __device__ inline long ghost_count(const BoxDesc& B, int gz)
{
    const long ny = B.hi[1] - B.lo[1] + 1, nk = ghost_layers(B, gz);
    return (long)ghost_zslabs(B, gz) * gz * B.nxg * B.nyg + 2L * NG * B.nxg * nk + 2L * NG * ny * nk;
}

// Decode the g-th ghost cell of box B in that numbering.
__device__ __forceinline__ void ghost_cell(const BoxDesc& B, int g, int gz, int& i, int& j, int& k)
{
    const int ny = B.hi[1] - B.lo[1] + 1, nk = ghost_layers(B, gz);
    const int nxg = B.nxg, nyg = B.nyg;
    const int zs = gz * nxg * nyg;
    if (g < ghost_zslabs(B, gz) * zs) {
        const int side = g / zs, r = g - side * zs;
        const int kk = r / (nxg * nyg), rem = r - kk * nxg * nyg;
        k = side ? B.hi[2] + 1 + kk : B.lo[2] - gz + kk;
        j = B.lo[1] - NG + rem / nxg;
        i = B.lo[0] - NG + rem % nxg;
        return;
    }
    g -= ghost_zslabs(B, gz) * zs;
    const int ys = NG * nxg * nk;
    if (g < 2 * ys) {
        const int side = g / ys, r = g - side * ys;
        const int kk = r / (NG * nxg), rem = r - kk * NG * nxg;
        k = B.lo[2] + kk;
        const int jj = rem / nxg;
        j = side ? B.hi[1] + 1 + jj : B.lo[1] - NG + jj;
        i = B.lo[0] - NG + rem % nxg;
        return;
    }
    g -= 2 * ys;
    const int xs = NG * ny * nk;
    const int side = g / xs, r = g - side * xs;
    const int kk = r / (NG * ny), rem = r - kk * NG * ny;
    k = B.lo[2] + kk;
    j = B.lo[1] + rem / NG;
    const int ii = rem % NG;
    i = side ? B.hi[0] + 1 + ii : B.lo[0] - NG + ii;
}

// Periodic image of cell (i,j,k) (a domain thinner than the ghost width wraps
// more than once) and, per direction, whether it lies below (-1) or above (+1)
// the domain.
__device__ __forceinline__ void ghost_wrap(const GeomDesc& G, int i, int j, int k, int iv[3], int outside[3])
{
    iv[0] = i; iv[1] = j; iv[2] = k;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (G.periodic[d]) {
            const int len = G.dom_hi[d] - G.dom_lo[d] + 1;
            int r = (iv[d] - G.dom_lo[d]) % len;
            if (r < 0) r += len;
            iv[d] = G.dom_lo[d] + r;
        }
        outside[d] = iv[d] < G.dom_lo[d] ? -1 : (iv[d] > G.dom_hi[d] ? 1 : 0);
    }
}

// Source cell and sign of component n of a ghost cell outside the domain by
// `outside` (AMReX FilccCell: foextrap copies the boundary cell,
// reflect_even/odd mirror about the face with sign +-1, composed when the cell
// is outside in several directions).  False for ext_dir (the case's boundary
// fill sets it) and for a source outside a domain thinner than the ghost width.
__device__ __forceinline__ bool bc_source(const GeomDesc& G, int n, const int iv[3], const int outside[3], int src[3],
                                          double& sign)
{
    src[0] = iv[0]; src[1] = iv[1]; src[2] = iv[2];
    sign = 1.0;
    bool user = false;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (!outside[d]) continue;
        const int bc = G.bc[n][d][outside[d] > 0];
        const int edge = outside[d] < 0 ? G.dom_lo[d] : G.dom_hi[d];
        if (bc == BC_FOEXT) src[d] = edge;
        else if (bc == BC_EVEN || bc == BC_ODD) {
            src[d] = outside[d] < 0 ? 2 * edge - src[d] - 1 : 2 * edge - src[d] + 1;
            if (bc == BC_ODD) sign = -sign;
        } else if (bc == BC_EXT) user = true;
    }
    if (user) return false;
    return src[0] >= G.dom_lo[0] && src[0] <= G.dom_hi[0] && src[1] >= G.dom_lo[1] && src[1] <= G.dom_hi[1] &&
           src[2] >= G.dom_lo[2] && src[2] <= G.dom_hi[2];
}

// Level ghost fill of one ghost cell (i,j,k) of box B at time `time`:
// periodic images and the per-direction BC maps from the level's own boxes,
// then the case's ext_dir values (Exec/1_DMR/prob_bcfill.cpp:14-58,
// Exec/3_FSTEP/prob_bcfill.cpp:8-33).  Components whose source no box of the
// level owns are left alone (the coarse/fine fill's part).
__device__ __forceinline__ void fill_ghost(const BoxDesc* __restrict__ boxes, const GeomDesc& G,
                                           double* __restrict__ S, double time, const BoxDesc& B, int i, int j, int k)
{
    const long dst = sidx(B, i, j, k);
    int iv[3], outside[3];
    ghost_wrap(G, i, j, k, iv, outside);
    // the BC maps mirror or clamp every component alike, so the source cell and
    // its owner are normally found once for all five
    int last[3] = {INT_MIN, INT_MIN, INT_MIN};
    long so = -1, sst = 0;
    for (int n = 0; n < NC; ++n) {
        int src[3];
        double sign;
        if (!bc_source(G, n, iv, outside, src, sign)) continue;
        if (src[0] != last[0] || src[1] != last[1] || src[2] != last[2]) {
            const int m = level_owner(G, boxes, src[0], src[1], src[2]);
            so = -1;
            if (m >= 0) {   // else not covered on this level (coarse/fine boundary)
                const BoxDesc& O = boxes[m];
                so = sidx(O, src[0], src[1], src[2]);
                sst = O.nptg;
            }
            last[0] = src[0]; last[1] = src[1]; last[2] = src[2];
        }
        if (so < 0) continue;
        S[dst + n * B.nptg] = sign * S[so + n * sst];
    }
    if (G.probtype == 1) {
        const double x = G.prob_lo[0] + (i + 0.5) * G.dx[0];
        const double x0 = 1.0 / 6.0 + sqrt(1 / 3.0) + time * 10.0 / sin(M_PI * (1.0 / 3.0));
        if (G.bc[0][1][1] == BC_EXT && j > G.dom_hi[1]) {
            for (int n = 0; n < NC; ++n) S[dst + n * B.nptg] = x < x0 ? G.left[n] : G.right[n];
        }
        if (G.bc[0][1][0] == BC_EVEN && j < G.dom_lo[1]) {
            if (x < 1.0 / 6.0)
                for (int n = 0; n < NC; ++n) S[dst + n * B.nptg] = G.left[n];
        }
        if (G.bc[0][0][0] == BC_EXT && i < G.dom_lo[0]) {
            for (int n = 0; n < NC; ++n) S[dst + n * B.nptg] = G.left[n];
        }
    } else if (G.probtype == 2) {
        // inflow state at x-lo (Exec/3_FSTEP/prob_bcfill.cpp:8-33)
        if (G.bc[0][0][0] == BC_EXT && i < G.dom_lo[0])
            for (int n = 0; n < NC; ++n) S[dst + n * B.nptg] = G.left[n];
    }
}

// Primitive state used by the flux kernels
template <class FT>
struct Prim {
    FT r, u, v, w, c, T;
};

// c2prim, Source/kernel/Kernels.H:13-40 (always FP64, then narrowed).  The
// FP32-flux form, whose primitives are narrowed to FP32, takes 1/rho as the
// hardware FP64 reciprocal refined by two Newton steps (about an FP64 ulp),
// T as p (1/rho)(1/Rg), and c as the hardware FP32 square root of the FP64
// product gamma Rg T (positive, far from the denormal range).  The energy
// difference, where the cancellation is, stays in FP64; every output keeps
// its FP32 rounding but for a last-bit difference now and then.
template <class FT>
__device__ __forceinline__ Prim<FT> c2prim(const double* __restrict__ s, long idx, long stride, const PhysArgs& ph)
{
    double rho = fmax(s[idx], ph.smallr);
    if constexpr (sizeof(FT) == 4) {
        double r = __builtin_amdgcn_rcp(rho);
        r = __builtin_fma(r, __builtin_fma(-rho, r, 1.0), r);
        r = __builtin_fma(r, __builtin_fma(-rho, r, 1.0), r);
        const double ux = s[idx + stride] * r, uy = s[idx + 2 * stride] * r, uz = s[idx + 3 * stride] * r;
        const double kineng = 0.5 * (ux * ux + uy * uy + uz * uz);
        const double ei = s[idx + 4 * stride] - rho * kineng;
        const double p = fmax((ph.gamma - 1.0) * ei, ph.smallp);
        const double T = p * (r * ph.Rinv);
        Prim<FT> q;
        q.r = (FT)rho; q.u = (FT)ux; q.v = (FT)uy; q.w = (FT)uz;
        q.T = (FT)T;
        q.c = __builtin_amdgcn_sqrtf((float)(ph.gamma * ph.Rg * T));
        return q;
    }
    double rhoinv = 1.0 / rho;
    double ux = s[idx + stride] * rhoinv;
    double uy = s[idx + 2 * stride] * rhoinv;
    double uz = s[idx + 3 * stride] * rhoinv;
    double kineng = 0.5 * (ux * ux + uy * uy + uz * uz);
    double ei = s[idx + 4 * stride] - rho * kineng;
    double p = fmax((ph.gamma - 1.0) * ei, ph.smallp);
    double T = p / (rho * ph.Rg);
    Prim<FT> q;
    q.r = (FT)rho; q.u = (FT)ux; q.v = (FT)uy; q.w = (FT)uz;
    q.T = (FT)T;
    q.c = (FT)sqrt(ph.gamma * ph.Rg * T);
    return q;
}

// primitive state of one cell from the primitive MultiFab
template <class FT>
__device__ __forceinline__ Prim<FT> ldprim(const FT* __restrict__ P, long idx, long stride)
{
    Prim<FT> q;
    q.r = P[idx]; q.u = P[idx + stride]; q.v = P[idx + 2 * stride]; q.w = P[idx + 3 * stride];
    q.c = P[idx + 4 * stride]; q.T = P[idx + 5 * stride];
    return q;
}

template <class FT>
__device__ __forceinline__ void ldvel(const FT* __restrict__ P, long idx, long stride, FT& u, FT& v, FT& w)
{
    u = P[idx + stride]; v = P[idx + 2 * stride]; w = P[idx + 3 * stride];
}

// velocity only (viscous tangential stencils)
template <class FT>
__device__ __forceinline__ void c2vel(const double* __restrict__ s, long idx, long stride, const PhysArgs& ph,
                                      FT& u, FT& v, FT& w)
{
    double rho = fmax(s[idx], ph.smallr);
    double rhoinv = 1.0 / rho;
    u = (FT)(s[idx + stride] * rhoinv);
    v = (FT)(s[idx + 2 * stride] * rhoinv);
    w = (FT)(s[idx + 3 * stride] * rhoinv);
}

// Velocity of a neighbouring cell for the tangential derivatives of a
// viscous face: exact (as c2vel) in the FP64 form; in the FP32-flux form,
// where the face flux is evaluated in FP32, from the FP32-rounded momenta and
// the hardware reciprocal of the density (a few FP32 ulps).
template <class FT>
__device__ __forceinline__ void c2vel_nb(const double* __restrict__ s, long idx, long stride, const PhysArgs& ph,
                                         FT& u, FT& v, FT& w)
{
    c2vel<FT>(s, idx, stride, ph, u, v, w);
}

template <>
__device__ __forceinline__ void c2vel_nb<float>(const double* __restrict__ s, long idx, long stride,
                                                const PhysArgs& ph, float& u, float& v, float& w)
{
    const float ri = __builtin_amdgcn_rcpf((float)fmax(s[idx], ph.smallr));
    u = (float)s[idx + stride] * ri;
    v = (float)s[idx + 2 * stride] * ri;
    w = (float)s[idx + 3 * stride] * ri;
}

// Steger-Warming split in direction D, Source/kernel/FluxSplit.H:11-193.
template <int D, class FT>
__device__ __forceinline__ void sw_split(const Prim<FT>& q, FT gamma, FT t1, FT t3, FT* __restrict__ fp,
                                         FT* __restrict__ fm)
{
    const FT d = q.r, u = q.u, v = q.v, w = q.w, c = q.c;
    const FT tmp1 = t1;      // 1/(2 gamma)
    const FT tmp2 = FT(2.0) * (gamma - FT(1.0));
    const FT tmp3 = t3;      // (3 - gamma)/(2 (gamma - 1))
    const FT un = D == 0 ? u : (D == 1 ? v : w);
    const FT uc1 = un - c;
    const FT uc2 = un + c;
    FT vvc1, vvc2;
    if (D == 0) {
        vvc1 = FT(0.5) * (uc1 * uc1 + v * v + w * w);
        vvc2 = FT(0.5) * (uc2 * uc2 + v * v + w * w);
    } else if (D == 1) {
        vvc1 = FT(0.5) * (u * u + uc1 * uc1 + w * w);
        vvc2 = FT(0.5) * (u * u + uc2 * uc2 + w * w);
    } else {
        vvc1 = FT(0.5) * (u * u + v * v + uc1 * uc1);
        vvc2 = FT(0.5) * (u * u + v * v + uc2 * uc2);
    }
    const FT v2 = (gamma - FT(1)) * (u * u + v * v + w * w);
    const FT w2 = tmp3 * c * c;
    const FT E1 = un, E2 = un - c, E3 = un + c;
    const FT E1P = (E1 + fabs(E1)) * FT(0.5);
    const FT E2P = (E2 + fabs(E2)) * FT(0.5);
    const FT E3P = (E3 + fabs(E3)) * FT(0.5);
    const FT E1M = E1 - E1P, E2M = E2 - E2P, E3M = E3 - E3P;
    const FT tmp0 = d * tmp1;
    const FT sp = tmp2 * E1P + E2P + E3P;
    const FT sm = tmp2 * E1M + E2M + E3M;
    fp[URHO] = tmp0 * sp;
    fm[URHO] = tmp0 * sm;
    const FT vel[3] = {u, v, w};
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        if (m == D) {
            fp[UMX + m] = tmp0 * (tmp2 * E1P * un + E2P * uc1 + E3P * uc2);
            fm[UMX + m] = tmp0 * (tmp2 * E1M * un + E2M * uc1 + E3M * uc2);
        } else {
            fp[UMX + m] = tmp0 * vel[m] * sp;
            fm[UMX + m] = tmp0 * vel[m] * sm;
        }
    }
    fp[UEDEN] = tmp0 * (E1P * v2 + E2P * vvc1 + E3P * vvc2 + w2 * (E2P + E3P));
    fm[UEDEN] = tmp0 * (E1M * v2 + E2M * vvc1 + E3M * vvc2 + w2 * (E2M + E3M));
}

// reciprocal: IEEE division in FP64, hardware v_rcp_f32 in FP32
__device__ __forceinline__ double rcp_(double x) { return 1.0 / x; }
__device__ __forceinline__ float rcp_(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double div_(double a, double b) { return a / b; }
__device__ __forceinline__ float div_(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// WENO-Z combination of the three candidate values v_k with the weights
// alpha_k = d_k (1 + r_k^2), r_k = tau5/(eps + beta_k), d = (1, 6, 3)
// (Source/kernel/Reconstruction.H:96-110), in the reference's operation order.
__device__ __forceinline__ double wenoz_combine(double s11, double s22, double s33, double s55, double v1, double v2,
                                                double v3)
{
    const double eps = 1e-6;
    double r1 = s55 / (eps + s11), r2 = s55 / (eps + s22), r3 = s55 / (eps + s33);
    double a1 = 1.0 + r1 * r1;
    double a2 = 6.0 * (1.0 + r2 * r2);
    double a3 = 3.0 * (1.0 + r3 * r3);
    double invsum = 1.0 / (a1 + a2 + a3);
    return (1.0 / 6.0) * invsum * (a1 * v1 + a2 * v2 + a3 * v3);
}

// FP32 form: the same arithmetic with hardware reciprocals.
__device__ __forceinline__ float wenoz_combine(float s11, float s22, float s33, float s55, float v1, float v2,
                                               float v3)
{
    const float eps = 1e-6f;
    float r1 = s55 * __builtin_amdgcn_rcpf(eps + s11), r2 = s55 * __builtin_amdgcn_rcpf(eps + s22),
          r3 = s55 * __builtin_amdgcn_rcpf(eps + s33);
    float a1 = 1.0f + r1 * r1;
    float a2 = 6.0f * (1.0f + r2 * r2);
    float a3 = 3.0f * (1.0f + r3 * r3);
    float invsum = __builtin_amdgcn_rcpf(a1 + a2 + a3);
    return (1.0f / 6.0f) * invsum * (a1 * v1 + a2 * v2 + a3 * v3);
}

// The same weights in FP32 without overflow, for faces whose plain FP32 flux
// came out non-finite: squaring r_k overflows single precision once
// tau5/(eps + beta_k) exceeds about 1.8e19 -- a jump of a few 1e6 in a split
// flux next to an exactly smooth sub-stencil, as in dimensional shear layers
// and shocks (the 256^3 Kelvin-Helmholtz layer reaches it within four steps).
// Each alpha_k is divided by (1 + r_max^2), r_max = tau5/z_min with
// z_k = eps + beta_k: alpha_k' = d_k (g + u_k^2 (1 - g)), u_k = z_min/z_k in
// (0, 1], g = 1/(1 + r_max^2) in [0, 1].  Every alpha_k' is at most 10 and the
// largest u_k is 1, so nothing overflows whatever the scale of the data.
__device__ __forceinline__ float wenoz_scaled(float s11, float s22, float s33, float s55, float v1, float v2, float v3)
{
    const float eps = 1e-6f;
    const float z1 = eps + s11, z2 = eps + s22, z3 = eps + s33;
    const float zm = fminf(z1, fminf(z2, z3));
    const float u1 = zm * __builtin_amdgcn_rcpf(z1), u2 = zm * __builtin_amdgcn_rcpf(z2),
                u3 = zm * __builtin_amdgcn_rcpf(z3);
    const float rm = s55 * __builtin_amdgcn_rcpf(zm);
    const float g = __builtin_amdgcn_rcpf(1.0f + rm * rm);
    const float h = 1.0f - g;
    const float a1 = g + u1 * u1 * h, a2 = 6.0f * (g + u2 * u2 * h), a3 = 3.0f * (g + u3 * u3 * h);
    const float inv = __builtin_amdgcn_rcpf(a1 + a2 + a3);
    return (1.0f / 6.0f) * ((a1 * inv) * v1 + (a2 * inv) * v2 + (a3 * inv) * v3);
}

// WENO-Z5 by stencil centre (FP32 form of the x sweep).  Every 3-cell
// sub-stencil (a, b, c) of Source/kernel/Reconstruction.H:83-137 is centred on
// one cell and serves three faces; with D2 = a - 2b + c and D1 = c - a its
// three smoothness indicators are
//   beta1 = 13/12 D2^2 + 1/4 D1^2,
//   beta0 = 13/12 D2^2 + 1/4 (2 D2 + D1)^2 = beta1 + D2 (D2 + D1),
//   beta2 = 13/12 D2^2 + 1/4 (2 D2 - D1)^2 = beta1 + D2 (D2 - D1),
// and its candidate values (over 6) are
//   P0 = (2a - 7b + 11c)/6 = b + 13/12 D2 + 3/4 D1,
//   P1 = (-a + 5b + 2c)/6  = b + 1/12 D2 + 1/4 D1,
//   P2 = (2a + 5b - c)/6   = b + 1/12 D2 - 1/4 D1,
//   M0 = (11a - 7b + 2c)/6 = b + 13/12 D2 - 3/4 D1.
// The face between cells f-1 and f takes, for f+, (beta0, P0) of centre f-2,
// (beta1, P1) of f-1 and (beta2, P2) of f; for f-, (beta2, M0) of centre f+1,
// (beta1, P2) of f and (beta0, P1) of f-1 -- the sub-stencils, weights and
// candidates of weno_p / weno_m below.  A cell's quantities are computed
// once instead of once per face (11 + 8 operations per cell and series
// instead of 32 per face); the results differ from weno_p / weno_m by
// rounding only.
struct WenoCentre {
    float b0, b1, b2, p0, p1, p2, m0;
};

// (explicit FMAs and no contraction, so every kernel that evaluates a face
// this way -- the x sweep and the flux registers -- gets the same bits)
__device__ __forceinline__ WenoCentre weno_centre(float a, float b, float c)
{
#pragma clang fp contract(off)
    const float c13 = 13.0f / 12.0f;
    const float D2 = __builtin_fmaf(-2.0f, b, a) + c, D1 = c - a;
    const float h = 0.5f * D1;
    WenoCentre w;
    w.b1 = __builtin_fmaf(c13 * D2, D2, h * h);
    w.b0 = __builtin_fmaf(D2, D2 + D1, w.b1);
    w.b2 = __builtin_fmaf(D2, D2 - D1, w.b1);
    const float e = __builtin_fmaf(1.0f / 12.0f, D2, b), g = __builtin_fmaf(c13, D2, b);
    w.p0 = __builtin_fmaf(0.75f, D1, g);
    w.m0 = __builtin_fmaf(-0.75f, D1, g);
    w.p1 = __builtin_fmaf(0.25f, D1, e);
    w.p2 = __builtin_fmaf(-0.25f, D1, e);
    return w;
}

// The value of x held by lane - 1 of the wave (a DPP wave shift; lane 0
// gets 0).  Every lane must be active where it is called.
__device__ __forceinline__ float lane_from_below(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x138, 0xf, 0xf, false));
}

// The WENO-Z combination of three sub-stencils (indicators s11, s22, s33 with
// weights 1, 6, 3) of candidates already divided by 6.
__device__ __forceinline__ float wenoz_centred(float s11, float s22, float s33, float v1, float v2, float v3)
{
#pragma clang fp contract(off)
    const float eps = 1e-6f;
    const float s55 = fabsf(s11 - s33);
    const float r1 = s55 * __builtin_amdgcn_rcpf(eps + s11), r2 = s55 * __builtin_amdgcn_rcpf(eps + s22),
                r3 = s55 * __builtin_amdgcn_rcpf(eps + s33);
    const float a1 = __builtin_fmaf(r1, r1, 1.0f);
    const float a2 = 6.0f * __builtin_fmaf(r2, r2, 1.0f);
    const float a3 = 3.0f * __builtin_fmaf(r3, r3, 1.0f);
    const float num = __builtin_fmaf(a3, v3, __builtin_fmaf(a2, v2, a1 * v1));
    return __builtin_amdgcn_rcpf((a1 + a2) + a3) * num;
}

// WENO-Z5 from f+ at cells (f-3..f+1) of face f, Source/kernel/Reconstruction.H:83-110
template <class FT, bool SCALED = false>
__device__ __forceinline__ FT weno_p(FT V1, FT V2, FT V3, FT V4, FT V5)
{
    const FT c13 = FT(13.0 / 12.0);
    FT t1 = V1 - FT(2.0) * V2 + V3, t2 = V1 - FT(4.0) * V2 + FT(3.0) * V3;
    FT s11 = c13 * t1 * t1 + FT(0.25) * t2 * t2;
    FT t3 = V2 - FT(2.0) * V3 + V4, t4 = V2 - V4;
    FT s22 = c13 * t3 * t3 + FT(0.25) * t4 * t4;
    FT t5 = V3 - FT(2.0) * V4 + V5, t6 = FT(3.0) * V3 - FT(4.0) * V4 + V5;
    FT s33 = c13 * t5 * t5 + FT(0.25) * t6 * t6;
    FT s55 = fabs(s11 - s33);
    FT v1 = FT(2.0) * V1 - FT(7.0) * V2 + FT(11.0) * V3;
    FT v2 = -V2 + FT(5.0) * V3 + FT(2.0) * V4;
    FT v3 = FT(2.0) * V3 + FT(5.0) * V4 - V5;
    if constexpr (SCALED && sizeof(FT) == 4) return wenoz_scaled(s11, s22, s33, s55, v1, v2, v3);
    else return wenoz_combine(s11, s22, s33, s55, v1, v2, v3);
}

// WENO-Z5 from f- at cells (f-2..f+2) of face f, Source/kernel/Reconstruction.H:112-137
template <class FT, bool SCALED = false>
__device__ __forceinline__ FT weno_m(FT V1, FT V2, FT V3, FT V4, FT V5)
{
    const FT c13 = FT(13.0 / 12.0);
    FT t1 = V5 - FT(2.0) * V4 + V3, t2 = V5 - FT(4.0) * V4 + FT(3.0) * V3;
    FT s11 = c13 * t1 * t1 + FT(0.25) * t2 * t2;
    FT t3 = V2 - FT(2.0) * V3 + V4, t4 = V4 - V2;
    FT s22 = c13 * t3 * t3 + FT(0.25) * t4 * t4;
    FT t5 = V3 - FT(2.0) * V2 + V1, t6 = FT(3.0) * V3 - FT(4.0) * V2 + V1;
    FT s33 = c13 * t5 * t5 + FT(0.25) * t6 * t6;
    FT s55 = fabs(s11 - s33);
    FT v1 = FT(11.0) * V3 - FT(7.0) * V4 + FT(2.0) * V5;
    FT v2 = -V4 + FT(5.0) * V3 + FT(2.0) * V2;
    FT v3 = FT(2.0) * V3 + FT(5.0) * V2 - V1;
    if constexpr (SCALED && sizeof(FT) == 4) return wenoz_scaled(s11, s22, s33, s55, v1, v2, v3);
    else return wenoz_combine(s11, s22, s33, s55, v1, v2, v3);
}

// Whether the FP32 form must redo a face's reconstructions with the scaled
// weights: its flux (5 components) is not finite.  A sum that overflows only
// sends a finite face through the scaled path, which gives the same weights.
// Never in FP64.
template <class FT>
__device__ __forceinline__ bool redo_scaled(const FT* F)
{
    if (sizeof(FT) != 4) return false;
    return __builtin_expect(!__builtin_isfinite(F[0] + F[1] + F[2] + F[3] + F[4]), 0);
}

// Convective face flux of the five components with the scaled weights, from
// accessors of the split fluxes f+ at the face's cells -3..+1 (gp(c, n)) and
// f- at -2..+2 (gm(c, n)).  The inputs pass through an empty asm so the
// compiler recomputes the smoothness indicators here instead of keeping those
// of the plain evaluation live across the (rare) branch.
template <class FT, class GP, class GM>
__device__ __forceinline__ void weno_face_scaled(GP gp, GM gm, FT* F)
{
#pragma unroll
    for (int n = 0; n < 5; ++n) {
        FT p[5], m[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            p[c] = gp(c, n);
            m[c] = gm(c, n);
            asm volatile("" : "+v"(p[c]), "+v"(m[c]));
        }
        F[n] = weno_p<FT, true>(p[0], p[1], p[2], p[3], p[4]) + weno_m<FT, true>(m[0], m[1], m[2], m[3], m[4]);
    }
}

// Sutherland viscosity, Source/kernel/Diffusion.H:21-22
template <class FT>
__device__ __forceinline__ FT sutherland(FT T, const PhysArgs& ph)
{
    return div_(FT(ph.C_s) * T * sqrt(T), T + FT(ph.T_s));
}

// Tangential 4-point difference q(+,R)+q(+,L)-q(-,R)-q(-,L), Diffusion.H:31-34
template <class FT>
__device__ __forceinline__ FT lat(FT pR, FT pL, FT mR, FT mL) { return pR + pL - mR - mL; }

// Viscous + heat flux on the face between cell L (low) and cell R (high) in
// direction D, added to F (Source/kernel/Diffusion.H:7-131).  The tangential
// derivatives come in as the two 4-point differences the reference forms:
//   D=0: ta = lat(u,y) tb = lat(v,y) tc = lat(u,z) td = lat(w,z)
//   D=1: ta = lat(u,x) tb = lat(v,x) tc = lat(v,z) td = lat(w,z)
//   D=2: ta = lat(u,x) tb = lat(w,x) tc = lat(v,y) td = lat(w,y)
// mutl, mutr: the Sutherland viscosities of the two cells (a march carries
// the high cell's value to the next face).
template <int D, class FT>
__device__ __forceinline__ void visc_face_mu(const Prim<FT>& L, const Prim<FT>& R, FT mutl, FT mutr, FT ta, FT tb,
                                             FT tc, FT td, const double* dxinv, const PhysArgs& ph, FT* F)
{
    const FT ktl = mutl * FT(ph.kOverMu), ktr = mutr * FT(ph.kOverMu);
    const FT dn = FT(dxinv[D]);
    const FT dT = (R.T - L.T) * dn;
    const FT du = (R.u - L.u) * dn;
    const FT dv = (R.v - L.v) * dn;
    const FT dw = (R.w - L.w) * dn;
    const FT muf = FT(0.5) * (mutl + mutr);
    const FT tt = FT(ph.twoThirds);
    FT t1, t2, t3;
    if (D == 0) {
        FT dudy = ta * (FT(0.25) * FT(dxinv[1]));
        FT dvdy = tb * (FT(0.25) * FT(dxinv[1]));
        FT dudz = tc * (FT(0.25) * FT(dxinv[2]));
        FT dwdz = td * (FT(0.25) * FT(dxinv[2]));
        FT divu = du + dvdy + dwdz;
        t1 = muf * (FT(2.0) * du - tt * divu);   // tauxx
        t2 = muf * (dudy + dv);                  // tauxy
        t3 = muf * (dudz + dw);                  // tauxz
    } else if (D == 1) {
        FT dudx = ta * (FT(0.25) * FT(dxinv[0]));
        FT dvdx = tb * (FT(0.25) * FT(dxinv[0]));
        FT dvdz = tc * (FT(0.25) * FT(dxinv[2]));
        FT dwdz = td * (FT(0.25) * FT(dxinv[2]));
        FT divu = dudx + dv + dwdz;
        t2 = muf * (FT(2.0) * dv - tt * divu);   // tauyy
        t1 = muf * (du + dvdx);                  // tauxy
        t3 = muf * (dw + dvdz);                  // tauyz
    } else {
        FT dudx = ta * (FT(0.25) * FT(dxinv[0]));
        FT dwdx = tb * (FT(0.25) * FT(dxinv[0]));
        FT dvdy = tc * (FT(0.25) * FT(dxinv[1]));
        FT dwdy = td * (FT(0.25) * FT(dxinv[1]));
        FT divu = dudx + dvdy + dw;
        t1 = muf * (du + dwdx);                  // tauxz
        t2 = muf * (dv + dwdy);                  // tauyz
        t3 = muf * (FT(2.0) * dw - tt * divu);   // tauzz
    }
    F[UMX] += -t1;
    F[UMY] += -t2;
    F[UMZ] += -t3;
    F[UEDEN] += FT(-0.5) * ((R.u + L.u) * t1 + (R.v + L.v) * t2 + (R.w + L.w) * t3 + (ktl + ktr) * dT);
}

template <int D, class FT>
__device__ __forceinline__ void visc_face(const Prim<FT>& L, const Prim<FT>& R, FT ta, FT tb, FT tc, FT td,
                                          const double* dxinv, const PhysArgs& ph, FT* F)
{
    visc_face_mu<D, FT>(L, R, sutherland(L.T, ph), sutherland(R.T, ph), ta, tb, tc, td, dxinv, ph, F);
}

// Gravity source (Source/Advance.cpp:331-342) and the stage combination of
// MultiFab::LinComb / Saxpy (Source/Advance.cpp:117-148); with no output
// state the divergence itself is stored (compute_dSdt).  With an output
// state, L holds the cell's new state on return.
__device__ __forceinline__ void finish_cell(const StageArgs& a, const BoxDesc& B, int i, int j, int k, double* L)
{
    const long s0 = sidx(B, i, j, k);
    const long st = B.nptg;
    if (a.flags & OCFD_GRAVITY) {
        const double g = -9.8;
        L[UMZ] += g * a.scur[s0];
        L[UEDEN] += g * a.scur[s0 + 3 * st];
    }
    if (a.snew) {
        const double cdt = a.ctl ? a.c_dt * a.ctl->dt : a.c_dt;
#pragma unroll
        for (int n = 0; n < NC; ++n) {
            double v = a.sold ? a.c_cur * a.scur[s0 + n * st] + a.c_old * a.sold[s0 + n * st]
                              : a.c_cur * a.scur[s0 + n * st];
            L[n] = v + cdt * L[n];
            a.snew[s0 + n * st] = L[n];
        }
    } else {
        const long d0 = didx(B, i, j, k);
#pragma unroll
        for (int n = 0; n < NC; ++n) a.acc[d0 + n * B.nptv] = L[n];
    }
}

// Per-wave part of the CFL reduction folded into a finishing sweep: each lane
// keeps the minimum dt bit pattern of its cells and whether one of them was
// not finite; the lanes still running meet in the wave's LDS slot (which
// tolerates lanes that left early), and one of them updates the global slot.
struct CflFold {
    unsigned long long m = 0xffffffffffffffffull;
    bool nf = false;
    __device__ __forceinline__ void cell(const StageArgs& a, const BoxDesc& B, const double* U)
    {
        if (cell_nonfinite(U[0], U[1], U[2], U[3], U[4])) { nf = true; return; }
        const double dt = cell_dt(U[0], U[1], U[2], U[3], U[4], a.ph, B.dx, box_dx_equal(B));
        const unsigned long long b = (unsigned long long)__double_as_longlong(dt);
        m = b < m ? b : m;
    }
};

__device__ __forceinline__ void cfl_wave_init(unsigned long long* slot, int lane)
{
    if (lane == 0) *slot = 0xffffffffffffffffull;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void cfl_wave_finish(const StageArgs& a, unsigned long long* slot, const CflFold& f)
{
    atomicMin(slot, f.m);
    const bool any_nf = __any(f.nf);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const unsigned long long act = __ballot(1);
    if ((int)(threadIdx.x & 63) == __ffsll((long long)act) - 1) {
        const unsigned long long m = *slot;
        if (m != 0xffffffffffffffffull) atomicMin(a.red, m);
        if (any_nf) atomicOr(a.bad, 1u);
    }
}

// z-uniform level: the stage result of cell (i,j,k) (first z layer) copied to
// the box's other z layers
__device__ __forceinline__ void copy_to_layers(const StageArgs& a, const BoxDesc& B, int i, int j, int k)
{
    if (a.snew) {
        const long s0 = sidx(B, i, j, k), st = B.nptg, dz = (long)B.nxg * B.nyg;
        for (int n = 0; n < NC; ++n) {
            const double v = a.snew[s0 + n * st];
            for (int kk = B.lo[2] + 1; kk <= B.hi[2]; ++kk) a.snew[s0 + (kk - k) * dz + n * st] = v;
        }
    } else {
        const long d0 = didx(B, i, j, k);
        const long dz = (long)(B.hi[0] - B.lo[0] + 1) * (B.hi[1] - B.lo[1] + 1);
        for (int n = 0; n < NC; ++n) {
            const double v = a.acc[d0 + n * B.nptv];
            for (int kk = B.lo[2] + 1; kk <= B.hi[2]; ++kk) a.acc[d0 + (kk - k) * dz + n * B.nptv] = v;
        }
    }
}

} // namespace ocfd
