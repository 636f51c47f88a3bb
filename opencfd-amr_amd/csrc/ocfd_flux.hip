// Flux-divergence sweeps of the per-box advance (EBR::compute_dSdt regular
// path, Source/Advance.cpp:231-342), one kernel per direction over the box
// work-list of a whole level (or of many levels/cases at once).
//
//   sweep_x : each z plane of the primitive region is one flattened run of
//             its rows; wave = 64 consecutive cells of that run (windows
//             advance by 58 cells and may straddle rows), a block holds 4
//             consecutive windows.  Each lane recovers the primitives of its
//             cell from the state (FP64, c2prim) and stores them for the y
//             and z sweeps; on rows inside the box it also splits its cell
//             (FT), the split fluxes are exchanged through the wave's LDS
//             rows, lanes 3..61 evaluate the faces whose stencil lies in one
//             row, and the x-part of -div F is stored.
//   sweep_y : lane = one i column, marches j through a chunk of cells with
//             a register window of split fluxes (6 cells of f+, 5 of f-), and
//             adds the y-part into the accumulator.
//   sweep_z : same march along k; adds the z-part and finishes the cell
//             (gravity, RK stage combination or plain dSdt store).
//
// The divergence is summed in FP64 in the reference's order
// dx-term + dy-term + dz-term (Source/kernel/Kernels.H:94-96); the partial
// sums between sweeps are stored in the flux type FT (FP64 in the plain form,
// FP32 in the FP32-flux form, where they are differences of FP32 fluxes).
// The y and z sweeps read the primitive MultiFab the x sweep wrote.
#include "ocfd_device.hpp"
#include "ocfd_internal.hpp"

namespace ocfd {

constexpr int XSEG = 58;   // cells an x-sweep window advances (64 lanes less a 6-cell stencil halo)
constexpr int CH = 32;     // cells per march chunk (y and z sweeps); small levels use shorter ones

// Orders the LDS accesses of one wave (each wave of the x sweep owns its own
// LDS rows, so no block barrier is needed between its phases).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// x sweep over the primitive region grow(bx, (3,3,gz)) of one z plane, read
// as one flattened run of rows (nx+6 cells each, y row after y row): wave
// `win` takes cells [win*XSEG, win*XSEG+64) of the run, so a row's 6 halo
// cells cost 6 lanes rather than a partly idle wave per row, and windows
// straddle rows.  A face is taken by the window whose lanes 3..61 hold its
// cell (its 6-cell stencil then lies inside the window and inside one row).
template <class FT, bool VISC>
__global__ __launch_bounds__(256) void k_sweep_x(const BoxDesc* __restrict__ boxes, const int4* __restrict__ tiles,
                                                 StageArgs a)
{
    // FP32: WENO quantities by stencil centre, the viscous state of the
    // face's low cell from the neighbour lane by a wave shift (no LDS)
    constexpr bool CENTRED = sizeof(FT) == 4;
    __shared__ FT sf[4][10][64];
    __shared__ FT sv[VISC && !CENTRED ? 4 : 1][VISC && !CENTRED ? 13 : 1][64];
    __shared__ float sq[CENTRED ? 4 : 1][CENTRED ? 8 : 1][64];
    if (ctl_done(a.ctl)) return;
    const int4 t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
    const BoxDesc B = boxes[t.x];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1;
    // t.y = first window of the block, t.z = z row of the region, t.w = z growth
    const int gz = t.w, R = nx + 6, ncell = (ny + 6) * R;
    const int nwin = max(1, (ncell - 6 + XSEG - 1) / XSEG);
    const int k = B.lo[2] - gz + t.z;
    // a z-uniform level (zrep) needs the first layer and its two neighbours
    // only; the test is uniform over the block (one z plane per block)
    if (a.zrep && (k < B.lo[2] - 1 || k > B.lo[2] + 1)) return;
    const int win = __builtin_amdgcn_readfirstlane(t.y + w);
    const int g0 = win * XSEG, jr0 = g0 / R;
    // row and column of this lane's cell: off < R + 64, one exact correction
    const int off = g0 - jr0 * R + lane;
    int rq = (int)((float)off * __builtin_amdgcn_rcpf((float)R));
    int ir = off - rq * R;
    if (ir >= R) { ++rq; ir -= R; }
    if (ir < 0) { --rq; ir += R; }
    const int jr = jr0 + rq;
    const bool cellok = win < nwin && g0 + lane < ncell;
    const int i = B.lo[0] - 3 + ir, j = B.lo[1] - 3 + jr;
    const bool frow = cellok && jr >= 3 && jr <= ny + 2 && k >= B.lo[2] && k <= (a.zrep ? B.lo[2] : B.hi[2]);
    const FT gam = (FT)a.ph.gamma;
    FT fp[5], fm[5];
    // FP32 viscous state of the lane's cell: T, u, v, w, then the tangential
    // neighbours' (u, v) at j+1, j-1 and (u, w) at k+1, k-1, then mu
    FT vs[13];

    if (cellok) {
        const long s0 = sidx(B, i, j, k);
        const long st = B.nptg;
        const Prim<FT> q = c2prim<FT>(a.scur, s0, st, a.ph);
        // each primitive is stored once: by the first window that holds it
        // in lanes 0..XSEG-1, the last window's other lanes by that window
        if (lane < XSEG || win == nwin - 1) {
            FT* __restrict__ P = (FT*)a.prim;
            const long p0 = s0 - B.off_s + B.off_p;
            P[p0] = q.r; P[p0 + st] = q.u; P[p0 + 2 * st] = q.v; P[p0 + 3 * st] = q.w;
            P[p0 + 4 * st] = q.c; P[p0 + 5 * st] = q.T;
        }
        if (frow) {
            sw_split<0, FT>(q, gam, sw_t1<FT>(a.ph), sw_t3<FT>(a.ph), fp, fm);
#pragma unroll
            for (int n = 0; n < 5; ++n) { sf[w][n][lane] = fp[n]; sf[w][5 + n][lane] = fm[n]; }
            if (VISC && CENTRED) {
                // tangential neighbours' velocities from the state itself
                // (their primitives are being written by other waves)
                vs[0] = q.T; vs[1] = q.u; vs[2] = q.v; vs[3] = q.w;
                vs[12] = sutherland(q.T, a.ph);   // once per cell, used by both faces
                FT u, v, ww;
                c2vel_nb<FT>(a.scur, s0 + B.nxg, st, a.ph, u, v, ww);               // (i, j+1)
                vs[4] = u; vs[5] = v;
                c2vel_nb<FT>(a.scur, s0 - B.nxg, st, a.ph, u, v, ww);               // (i, j-1)
                vs[6] = u; vs[7] = v;
                c2vel_nb<FT>(a.scur, s0 + (long)B.nxg * B.nyg, st, a.ph, u, v, ww);  // (i, k+1)
                vs[8] = u; vs[9] = ww;
                c2vel_nb<FT>(a.scur, s0 - (long)B.nxg * B.nyg, st, a.ph, u, v, ww);  // (i, k-1)
                vs[10] = u; vs[11] = ww;
            } else if (VISC) {
                // tangential neighbours' velocities from the state itself
                // (their primitives are being written by other waves)
                sv[w][0][lane] = q.T; sv[w][1][lane] = q.u; sv[w][2][lane] = q.v; sv[w][3][lane] = q.w;
                sv[w][12][lane] = sutherland(q.T, a.ph);   // once per cell, read by both faces
                FT u, v, ww;
                c2vel_nb<FT>(a.scur, s0 + B.nxg, st, a.ph, u, v, ww);               // (i, j+1)
                sv[w][4][lane] = u; sv[w][5][lane] = v;
                c2vel_nb<FT>(a.scur, s0 - B.nxg, st, a.ph, u, v, ww);               // (i, j-1)
                sv[w][6][lane] = u; sv[w][7][lane] = v;
                c2vel_nb<FT>(a.scur, s0 + (long)B.nxg * B.nyg, st, a.ph, u, v, ww);  // (i, k+1)
                sv[w][8][lane] = u; sv[w][9][lane] = ww;
                c2vel_nb<FT>(a.scur, s0 - (long)B.nxg * B.nyg, st, a.ph, u, v, ww);  // (i, k-1)
                sv[w][10][lane] = u; sv[w][11][lane] = ww;
            }
        }
    }
    wave_sync();

    const bool faceok = frow && lane >= 3 && lane <= 61 && ir >= 3 && ir <= nx + 3;
    FT F[5];
    // FP32 viscous terms of the face, formed before the WENO part (each is
    // one addition to its component, so the sum is the same either way) from
    // this lane's cell and the low cell's state shifted in from lane - 1
    FT V[5] = {0, 0, 0, 0, 0};
    if constexpr (CENTRED && VISC) {
        FT vl[13];
#pragma unroll
        for (int m = 0; m < 13; ++m) vl[m] = lane_from_below(vs[m]);
        Prim<FT> L, R;
        L.T = vl[0]; L.u = vl[1]; L.v = vl[2]; L.w = vl[3];
        R.T = vs[0]; R.u = vs[1]; R.v = vs[2]; R.w = vs[3];
        const FT ta = lat(vs[4], vl[4], vs[6], vl[6]);
        const FT tb = lat(vs[5], vl[5], vs[7], vl[7]);
        const FT tc = lat(vs[8], vl[8], vs[10], vl[10]);
        const FT td = lat(vs[9], vl[9], vs[11], vl[11]);
        visc_face_mu<0, FT>(L, R, vl[12], vs[12], ta, tb, tc, td, B.dxinv, a.ph, V);
    }
    if constexpr (CENTRED) {
        // per component: every lane forms the quantities of the sub-stencils
        // centred on its cell and publishes those its neighbour faces take,
        // then every lane combines the three centres of its face.  All lanes
        // run both phases (lane indices clamped to the wave, no branches);
        // only the faces of faceok lanes, whose 6 cells lie in their row and
        // were split above, are used.
        const int lm1 = max(lane - 1, 0), lm2 = max(lane - 2, 0), lp1 = min(lane + 1, 63);
#pragma unroll
        for (int n = 0; n < 5; ++n) {
            const WenoCentre P = weno_centre(sf[w][n][lm1], fp[n], sf[w][n][lp1]);
            const WenoCentre M = weno_centre(sf[w][5 + n][lm1], fm[n], sf[w][5 + n][lp1]);
            sq[w][0][lane] = P.b0; sq[w][1][lane] = P.p0; sq[w][2][lane] = P.b1; sq[w][3][lane] = P.p1;
            sq[w][4][lane] = M.b2; sq[w][5][lane] = M.m0; sq[w][6][lane] = M.b0; sq[w][7][lane] = M.p1;
            wave_sync();
            F[n] = wenoz_centred(sq[w][0][lm2], sq[w][2][lm1], P.b2, sq[w][1][lm2], sq[w][3][lm1], P.p2)
                 + wenoz_centred(sq[w][4][lp1], M.b1, sq[w][6][lm1], sq[w][5][lp1], M.p2, sq[w][7][lm1]);
            wave_sync();
        }
    }
    if (faceok) {
        if constexpr (!CENTRED) {
#pragma unroll
            for (int n = 0; n < 5; ++n) {
                F[n] = weno_p<FT>(sf[w][n][lane - 3], sf[w][n][lane - 2], sf[w][n][lane - 1], sf[w][n][lane], sf[w][n][lane + 1])
                     + weno_m<FT>(sf[w][5 + n][lane - 2], sf[w][5 + n][lane - 1], sf[w][5 + n][lane], sf[w][5 + n][lane + 1],
                                  sf[w][5 + n][lane + 2]);
            }
        }
        auto add_visc = [&](FT* Fo) {
            if constexpr (CENTRED) {
#pragma unroll
                for (int n = 1; n < 5; ++n) Fo[n] += V[n];
                return;
            }
            Prim<FT> L, R;
            L.T = sv[w][0][lane - 1]; L.u = sv[w][1][lane - 1]; L.v = sv[w][2][lane - 1]; L.w = sv[w][3][lane - 1];
            R.T = sv[w][0][lane];     R.u = sv[w][1][lane];     R.v = sv[w][2][lane];     R.w = sv[w][3][lane];
            FT ta = lat(sv[w][4][lane], sv[w][4][lane - 1], sv[w][6][lane], sv[w][6][lane - 1]);
            FT tb = lat(sv[w][5][lane], sv[w][5][lane - 1], sv[w][7][lane], sv[w][7][lane - 1]);
            FT tc = lat(sv[w][8][lane], sv[w][8][lane - 1], sv[w][10][lane], sv[w][10][lane - 1]);
            FT td = lat(sv[w][9][lane], sv[w][9][lane - 1], sv[w][11][lane], sv[w][11][lane - 1]);
            visc_face_mu<0, FT>(L, R, sv[w][12][lane - 1], sv[w][12][lane], ta, tb, tc, td, B.dxinv, a.ph, Fo);
        };
        if (VISC) add_visc(F);
        // checked after the viscous part, so the face's LDS reads and
        // arithmetic stay in one block; the rare redo recomputes both parts
        if (redo_scaled(F)) {
            weno_face_scaled<FT>([&](int c, int n) { return sf[w][n][lane - 3 + c]; },
                                 [&](int c, int n) { return sf[w][5 + n][lane - 2 + c]; }, F);
            if (VISC) add_visc(F);
        }
    }
    wave_sync();
    if (faceok) {
#pragma unroll
        for (int n = 0; n < 5; ++n) sf[w][n][lane] = F[n];
    }
    wave_sync();
    if (frow && lane >= 3 && lane <= 60 && ir >= 3 && ir <= nx + 2) {
        const long d0 = didx(B, i, j, k);
        FT* __restrict__ part = (FT*)a.part;
#pragma unroll
        for (int n = 0; n < 5; ++n)
            part[d0 + n * B.nptv] = (FT)(B.dxinv[0] * ((double)sf[w][n][lane] - (double)sf[w][n][lane + 1]));
    }
}

// Marching sweep along direction D (1 = y, 2 = z).  FINAL: this sweep
// completes the divergence and finishes the cell.  ZREP: a z-uniform level
// whose sweep covers the first z layer and copies each finished cell to the
// other layers (a template flag so the 3D instantiations keep their registers).
// The FP32 forms of the 3D sweeps are held to 128 VGPRs: four waves per
// SIMD hide the march's load latency (three waves measured about 19% slower).
template <int D, class FT, bool VISC, bool FINAL, bool ZREP = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(FT) == 4 && (D == 2 || !FINAL) ? 4 : 1)))
void k_sweep_march(const BoxDesc* __restrict__ boxes, const int4* __restrict__ tiles,
                                                     StageArgs a)
{
    __shared__ unsigned long long cflslot[FINAL ? 4 : 1];
    if (ctl_done(a.ctl)) return;
    const int4 t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
    const BoxDesc B = boxes[t.x];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (FINAL && a.red) cfl_wave_init(&cflslot[w], lane);   // before any lane leaves
    // the "other" transverse direction: z for the y sweep, y for the z sweep
    constexpr int O = D == 1 ? 2 : 1;
    int i, o, m0;
    if (ZREP) {
        // first z layer only; sub-rows of 2^lw lanes, sub-row k on chunk k of the tile
        const int lw = t.y >> 16, sub = lane >> lw;
        i = B.lo[0] + ((t.y & 0xffff) << lw) + (lane & ((1 << lw) - 1));
        o = B.lo[O];
        m0 = B.lo[D] + (t.w & 0xfffff) + sub * (t.w >> 20);
        if (i > B.hi[0] || m0 > B.hi[D]) return;
    } else {
        i = B.lo[0] + t.y * 64 + lane;
        o = B.lo[O] + t.z * 4 + w;
        m0 = B.lo[D] + (t.w & 0xfffff);
        if (i > B.hi[0] || o > B.hi[O]) return;
    }
    const int m1 = min(m0 + (t.w >> 20) - 1, B.hi[D]);
    const FT gam = (FT)a.ph.gamma;
    const long st = B.nptg;
    const long dstep = D == 1 ? (long)B.nxg : (long)B.nxg * B.nyg;   // index step along D
    const long ostep = O == 1 ? (long)B.nxg : (long)B.nxg * B.nyg;   // index step along O
    const double dxi = B.dxinv[D];
    int jj = D == 1 ? m0 : o, kk = D == 1 ? o : m0;
    const long base = sidx(B, i, jj, kk) - B.off_s + B.off_p;   // primitive index of cell m0
    const FT* __restrict__ P = (const FT*)a.prim;
    const long dbase = didx(B, i, jj, kk);
    const long dd = D == 1 ? (long)(B.hi[0] - B.lo[0] + 1)
                           : (long)(B.hi[0] - B.lo[0] + 1) * (B.hi[1] - B.lo[1] + 1);

    FT Wp[6][5], Wm[5][5];
    // prologue: cells m0-3 .. m0+1
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        Prim<FT> q = ldprim<FT>(P, base + (c - 3) * dstep, st);
        FT fp[5], fm[5];
        sw_split<D, FT>(q, gam, sw_t1<FT>(a.ph), sw_t3<FT>(a.ph), fp, fm);
#pragma unroll
        for (int n = 0; n < 5; ++n) Wp[c][n] = fp[n];
        if (c >= 1) {
#pragma unroll
            for (int n = 0; n < 5; ++n) Wm[c - 1][n] = fm[n];
        }
    }
    // viscous state of the low cell of the first face (m0-1)
    Prim<FT> PL;
    FT muL = 0;
    FT lxpL_a = 0, lxpL_b = 0, lxmL_a = 0, lxmL_b = 0, lopL_a = 0, lopL_b = 0, lomL_a = 0, lomL_b = 0;
    if (VISC) {
        const long s = base - dstep;
        PL = ldprim<FT>(P, s, st);
        muL = sutherland(PL.T, a.ph);
        FT u, v, w2;
        ldvel<FT>(P, s + 1, st, u, v, w2);
        lxpL_a = u; lxpL_b = D == 1 ? v : w2;
        ldvel<FT>(P, s - 1, st, u, v, w2);
        lxmL_a = u; lxmL_b = D == 1 ? v : w2;
        ldvel<FT>(P, s + ostep, st, u, v, w2);
        lopL_a = D == 1 ? v : v; lopL_b = w2;
        ldvel<FT>(P, s - ostep, st, u, v, w2);
        lomL_a = v; lomL_b = w2;
    }

    FT Fprev[5];
    CflFold cfl;
    for (int f = m0; f <= m1 + 1; ++f) {
        // load cell f+2
        {
            Prim<FT> q = ldprim<FT>(P, base + (long)(f + 2 - m0) * dstep, st);
            FT fp[5], fm[5];
            sw_split<D, FT>(q, gam, sw_t1<FT>(a.ph), sw_t3<FT>(a.ph), fp, fm);
#pragma unroll
            for (int n = 0; n < 5; ++n) { Wp[5][n] = fp[n]; Wm[4][n] = fm[n]; }
        }
        FT F[5];
#pragma unroll
        for (int n = 0; n < 5; ++n)
            F[n] = weno_p<FT>(Wp[0][n], Wp[1][n], Wp[2][n], Wp[3][n], Wp[4][n])
                 + weno_m<FT>(Wm[0][n], Wm[1][n], Wm[2][n], Wm[3][n], Wm[4][n]);
        // checked before the viscous part: the window's first cell is dead
        // after this point, which keeps the march within its 128 VGPRs
        if (redo_scaled(F))
            weno_face_scaled<FT>([&](int c, int n) { return Wp[c][n]; }, [&](int c, int n) { return Wm[c][n]; }, F);
        if (VISC) {
            const long s = base + (long)(f - m0) * dstep;
            Prim<FT> PR = ldprim<FT>(P, s, st);
            FT u, v, w2;
            ldvel<FT>(P, s + 1, st, u, v, w2);
            FT xpa = u, xpb = D == 1 ? v : w2;
            ldvel<FT>(P, s - 1, st, u, v, w2);
            FT xma = u, xmb = D == 1 ? v : w2;
            ldvel<FT>(P, s + ostep, st, u, v, w2);
            FT opa = v, opb = w2;
            ldvel<FT>(P, s - ostep, st, u, v, w2);
            FT oma = v, omb = w2;
            // D=1: ta=lat(u,x) tb=lat(v,x) tc=lat(v,z) td=lat(w,z)
            // D=2: ta=lat(u,x) tb=lat(w,x) tc=lat(v,y) td=lat(w,y)
            FT ta = lat(xpa, lxpL_a, xma, lxmL_a);
            FT tb = lat(xpb, lxpL_b, xmb, lxmL_b);
            FT tc = lat(opa, lopL_a, oma, lomL_a);
            FT td = lat(opb, lopL_b, omb, lomL_b);
            const FT muR = sutherland(PR.T, a.ph);
            visc_face_mu<D, FT>(PL, PR, muL, muR, ta, tb, tc, td, B.dxinv, a.ph, F);
            PL = PR;
            muL = muR;
            lxpL_a = xpa; lxpL_b = xpb; lxmL_a = xma; lxmL_b = xmb;
            lopL_a = opa; lopL_b = opb; lomL_a = oma; lomL_b = omb;
        }
        if (f > m0) {
            // cell f-1 has both faces
            const long d0 = dbase + (long)(f - 1 - m0) * dd;
            double L[5];
#pragma unroll
            for (int n = 0; n < 5; ++n)
                L[n] = (double)((const FT*)a.part)[d0 + n * B.nptv] + dxi * ((double)Fprev[n] - (double)F[n]);
            if (FINAL) {
                const int c = f - 1;
                finish_cell(a, B, i, D == 1 ? c : o, D == 1 ? o : c, L);
                if (ZREP) copy_to_layers(a, B, i, D == 1 ? c : o, D == 1 ? o : c);
                if (a.red) cfl.cell(a, B, L);
            } else {
#pragma unroll
                for (int n = 0; n < 5; ++n) ((FT*)a.part)[d0 + n * B.nptv] = (FT)L[n];
            }
        }
#pragma unroll
        for (int n = 0; n < 5; ++n) Fprev[n] = F[n];
        // shift the window
#pragma unroll
        for (int c = 0; c < 5; ++c) {
#pragma unroll
            for (int n = 0; n < 5; ++n) Wp[c][n] = Wp[c + 1][n];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
#pragma unroll
            for (int n = 0; n < 5; ++n) Wm[c][n] = Wm[c + 1][n];
        }
    }
    if (FINAL && a.red) cfl_wave_finish(a, &cflslot[w], cfl);
}

// Primitive recovery (c2prim, Source/kernel/Kernels.H:13-40) once per cell
// of the stencil region of every box, in FP64, stored narrowed to FT.
// tiles: (box, first cell of the region, count, z growth of the region).
template <class FT>
__global__ __launch_bounds__(256) void k_prims(const BoxDesc* __restrict__ boxes, const int4* __restrict__ tiles,
                                               StageArgs a)
{
    if (ctl_done(a.ctl)) return;
    const int4 t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
    if ((int)threadIdx.x >= t.z) return;
    const BoxDesc B = boxes[t.x];
    const int gz = t.w;
    const int ex = B.hi[0] - B.lo[0] + 7, ey = B.hi[1] - B.lo[1] + 7;
    const int g = t.y + threadIdx.x;
    const int i = B.lo[0] - 3 + g % ex;
    const int j = B.lo[1] - 3 + (g / ex) % ey;
    const int k = B.lo[2] - gz + g / (ex * ey);
    const long s0 = sidx(B, i, j, k);
    const Prim<FT> q = c2prim<FT>(a.scur, s0, B.nptg, a.ph);
    FT* __restrict__ P = (FT*)a.prim;
    const long p0 = s0 - B.off_s + B.off_p;
    const long st = B.nptg;
    P[p0] = q.r; P[p0 + st] = q.u; P[p0 + 2 * st] = q.v; P[p0 + 3 * st] = q.w;
    P[p0 + 4 * st] = q.c; P[p0 + 5 * st] = q.T;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <class FT, bool VISC>
static void launch_sweeps_t(const LevelDev& L, const StageArgs& a0, bool planar, hipStream_t s)
{
    StageArgs a = a0;
    a.prim = sizeof(FT) == 4 ? (const void*)L.d_prim32 : (const void*)L.d_prim64;
    a.part = (void*)L.d_acc;
    // a planar level (one z ghost layer filled) holds z-uniform data: its
    // sweeps take the first z layer of each box and copy the result to the rest
    a.zrep = planar && L.fill_gz == 1;
    const int nx = a.zrep ? L.ntile_xz : (planar ? L.ntile_xp : L.ntile_x);
    if (nx) hipLaunchKernelGGL((k_sweep_x<FT, VISC>), dim3(nx), dim3(256), 0, s, L.d_boxes,
                               planar ? L.d_tile_xp : L.d_tile_x, a);
    if (planar) {
        // one wave per march tile when only the first z layer is swept
        if (a.zrep) {
            if (L.ntile_yp)
                hipLaunchKernelGGL((k_sweep_march<1, FT, VISC, true, true>), dim3(L.ntile_yp), dim3(64), 0, s,
                                   L.d_boxes, L.d_tile_y + L.ntile_y, a);
        } else if (L.ntile_y)
            hipLaunchKernelGGL((k_sweep_march<1, FT, VISC, true>), dim3(L.ntile_y), dim3(256), 0, s, L.d_boxes,
                               L.d_tile_y, a);
    } else {
        if (L.ntile_y) hipLaunchKernelGGL((k_sweep_march<1, FT, VISC, false>), dim3(L.ntile_y), dim3(256), 0, s,
                                          L.d_boxes, L.d_tile_y, a);
        if (L.ntile_z) hipLaunchKernelGGL((k_sweep_march<2, FT, VISC, true>), dim3(L.ntile_z), dim3(256), 0, s,
                                          L.d_boxes, L.d_tile_z, a);
    }
}

void launch_sweeps(const LevelDev& L, const StageArgs& a, hipStream_t s)
{
    const bool planar = (a.flags & OCFD_PLANAR) != 0;
    const bool visc = (a.flags & OCFD_VISC) != 0;
    if (a.flags & OCFD_FLUX_FP32) {
        if (visc) launch_sweeps_t<float, true>(L, a, planar, s);
        else launch_sweeps_t<float, false>(L, a, planar, s);
    } else {
        if (visc) launch_sweeps_t<double, true>(L, a, planar, s);
        else launch_sweeps_t<double, false>(L, a, planar, s);
    }
    launch_eb(L, a, s);   // cut and covered boxes
}

// One sweep alone (0 = x, 1 = y, 2 = z final, 4 = primitives, 5 = the first-layer y
// march of a z-uniform level with a one-layer fill), for per-kernel timing.
void launch_sweep_one(const LevelDev& L, const StageArgs& a0, int which, hipStream_t s)
{
    const bool visc = (a0.flags & OCFD_VISC) != 0;
    const bool f32 = (a0.flags & OCFD_FLUX_FP32) != 0;
    StageArgs a = a0;
    a.prim = f32 ? (const void*)L.d_prim32 : (const void*)L.d_prim64;
    a.part = (void*)L.d_acc;
#define OCFD_SW(FT, V)                                                                                           \
    do {                                                                                                         \
        if (which == 4) hipLaunchKernelGGL((k_prims<FT>), dim3(L.ntile_p), dim3(256), 0, s, L.d_boxes,           \
                                           L.d_tile_p, a);                                                       \
        else if (which == 0) hipLaunchKernelGGL((k_sweep_x<FT, V>), dim3(L.ntile_x), dim3(256), 0, s, L.d_boxes,      \
                                           L.d_tile_x, a);                                                       \
        else if (which == 1) hipLaunchKernelGGL((k_sweep_march<1, FT, V, false>), dim3(L.ntile_y), dim3(256), 0, \
                                                s, L.d_boxes, L.d_tile_y, a);                                     \
        else if (which == 5) {                                                                                   \
            if (L.ntile_yp) hipLaunchKernelGGL((k_sweep_march<1, FT, V, true, true>), dim3(L.ntile_yp), dim3(64), \
                                               0, s, L.d_boxes, L.d_tile_y + L.ntile_y, a);                     \
        }                                                                                                        \
        else hipLaunchKernelGGL((k_sweep_march<2, FT, V, true>), dim3(L.ntile_z), dim3(256), 0, s, L.d_boxes,    \
                                L.d_tile_z, a);                                                                  \
    } while (0)
    if (f32) { if (visc) OCFD_SW(float, true); else OCFD_SW(float, false); }
    else { if (visc) OCFD_SW(double, true); else OCFD_SW(double, false); }
#undef OCFD_SW
}

int tile_xseg() { return XSEG; }
int tile_chunk() { return CH; }

} // namespace ocfd
