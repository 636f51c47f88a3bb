// Level set-up, ghost fill (FillPatch on one level), CFL reduction and the
// C ABI of include/ocfd.h.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "ocfd_device.hpp"
#include "ocfd_internal.hpp"

namespace ocfd {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

#define OCFD_HIP(call)                                                                  \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            set_error(std::string(#call) + ": " + hipGetErrorString(e_));               \
            return OCFD_EHIP;                                                           \
        }                                                                               \
    } while (0)

static int check_launch(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string(what) + ": " + hipGetErrorString(e));
        return OCFD_EHIP;
    }
    return OCFD_OK;
}

// ---------------------------------------------------------------------------
// ghost fill
// ---------------------------------------------------------------------------

// Ghost cells from periodic images, per-direction BC maps (AMReX FilccCell:
// foextrap copies the boundary cell, reflect_even/odd mirror about the face
// with sign +-1, composed when a cell is outside in several directions) and
// the case's ext_dir fill (Exec/1_DMR/prob_bcfill.cpp:14-58).
// With device time control the fill time is ctl->time + toff * ctl->dt and a
// finished run fills nothing.

__global__ __launch_bounds__(256) void k_fill(const BoxDesc* __restrict__ boxes, const GeomDesc* __restrict__ geoms,
                                              const int4* __restrict__ tiles, double* __restrict__ S, double time,
                                              const ocfd_step_ctl* __restrict__ ctl, double toff)
{
    if (ctl) {
        if (ctl->done) return;
        time = ctl->time + toff * ctl->dt;
    }
    const int4 t = tiles[blockIdx.x];
    if ((int)threadIdx.x >= t.z) return;
    const BoxDesc B = boxes[t.x];
    int i, j, k;
    ghost_cell(B, t.y + threadIdx.x, t.w, i, j, k);
    fill_ghost(boxes, geoms[B.geom], S, time, B, i, j, k);
}

// the same over a list of ghost cells (box, ghost index at depth gz): the
// cells of a level the level fill writes (ocfd_level's ghost lists)
__global__ __launch_bounds__(256) void k_fill_list(const BoxDesc* __restrict__ boxes,
                                                   const GeomDesc* __restrict__ geoms, const int2* __restrict__ list,
                                                   const unsigned* __restrict__ n, int gz, double* __restrict__ S,
                                                   double time, const ocfd_step_ctl* __restrict__ ctl, double toff)
{
    if (ctl) {
        if (ctl->done) return;
        time = ctl->time + toff * ctl->dt;
    }
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= (int)*n) return;
    const int2 e = list[q];
    const BoxDesc B = boxes[e.x];
    int i, j, k;
    ghost_cell(B, e.y, gz, i, j, k);
    fill_ghost(boxes, geoms[B.geom], S, time, B, i, j, k);
}

// Ghost lists of a level at fill depth t.w: list A, the cells the level fill
// writes for some component (a box of the level owns the BC-mapped source) or
// may set (outside the domain: the case's boundary fill); list B, the cells
// the coarse/fine fill may write (for some component the mapped source lies in
// the domain and no box of the level owns it).  A cell can be in both.  Each
// wave appends its cells with one atomic add per list; the order within a
// list varies from run to run, the cells a fill writes and their values do not.
__global__ __launch_bounds__(256) void k_ghost_lists(const BoxDesc* __restrict__ boxes,
                                                     const GeomDesc* __restrict__ geoms, const int4* __restrict__ tiles,
                                                     int2* __restrict__ la, int2* __restrict__ lb,
                                                     unsigned* __restrict__ counts)
{
    const int4 t = tiles[blockIdx.x];
    const int lane = threadIdx.x & 63;
    const int g = t.y + (int)threadIdx.x;
    bool a = false, b = false;
    if ((int)threadIdx.x < t.z) {
        const BoxDesc B = boxes[t.x];
        const GeomDesc& G = geoms[B.geom];
        int i, j, k;
        ghost_cell(B, g, t.w, i, j, k);
        int iv[3], outside[3];
        ghost_wrap(G, i, j, k, iv, outside);
        a = outside[0] || outside[1] || outside[2];
        if (!a) {
            // inside the domain every component's source is the wrapped cell
            if (level_owner(G, boxes, iv[0], iv[1], iv[2]) >= 0) a = true;
            else b = true;
        } else {
            // outside: one owner lookup per distinct source cell
            int last[3] = {0, 0, 0};
            bool have = false, own = false;
            for (int n = 0; n < NC; ++n) {
                int src[3];
                double sign;
                if (!bc_source(G, n, iv, outside, src, sign)) continue;
                if (!have || src[0] != last[0] || src[1] != last[1] || src[2] != last[2]) {
                    own = level_owner(G, boxes, src[0], src[1], src[2]) >= 0;
                    last[0] = src[0]; last[1] = src[1]; last[2] = src[2];
                    have = true;
                }
                if (!own) b = true;
            }
        }
    }
    const unsigned long long ma = __ballot(a), mb = __ballot(b);
    unsigned ba = 0, bb = 0;
    if (lane == 0) {
        if (ma) ba = atomicAdd(&counts[0], (unsigned)__popcll(ma));
        if (mb) bb = atomicAdd(&counts[1], (unsigned)__popcll(mb));
    }
    ba = __shfl(ba, 0);
    bb = __shfl(bb, 0);
    const unsigned long long below = (1ull << lane) - 1ull;
    if (a) la[ba + __popcll(ma & below)] = make_int2(t.x, g);
    if (b) lb[bb + __popcll(mb & below)] = make_int2(t.x, g);
}

// ---------------------------------------------------------------------------
// CFL reduction, Source/Advance.cpp:10-79
// ---------------------------------------------------------------------------
__global__ void k_red_init(unsigned long long* r, unsigned int* bad)
{
    const double big = std::numeric_limits<double>::max();
    *r = (unsigned long long)__double_as_longlong(big);
    *bad = 0;
}

// bad: set to 1 when a valid cell holds a NaN or inf (such cells do not
// enter the minimum)
__global__ __launch_bounds__(256) void k_estdt(const BoxDesc* __restrict__ boxes, const int4* __restrict__ tiles,
                                               const double* __restrict__ S, PhysArgs ph,
                                               unsigned long long* __restrict__ out, unsigned int* __restrict__ bad,
                                               const ocfd_step_ctl* ctl)
{
    __shared__ double red[4];
    if (ctl_done(ctl)) return;
    const int4 t = tiles[blockIdx.x];
    const BoxDesc B = boxes[t.x];
    const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1;
    const bool dxe = box_dx_equal(B);
    double dt = std::numeric_limits<double>::max();
    int nf = 0;
    for (int c = threadIdx.x; c < t.z; c += 256) {
        const int g = t.y + c;
        const int i = B.lo[0] + g % nx;
        const int j = B.lo[1] + (g / nx) % ny;
        const int k = B.lo[2] + g / (nx * ny);
        const long s0 = sidx(B, i, j, k);
        const long st = B.nptg;
        const double rho = S[s0], mx = S[s0 + st], my = S[s0 + 2 * st], mz = S[s0 + 3 * st], E = S[s0 + 4 * st];
        if (cell_nonfinite(rho, mx, my, mz, E)) nf = 1;
        else dt = fmin(dt, cell_dt(rho, mx, my, mz, E, ph, B.dx, dxe));
    }
    for (int off = 32; off > 0; off >>= 1) dt = fmin(dt, __shfl_xor(dt, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = dt;
    nf = __syncthreads_or(nf);
    if (threadIdx.x == 0) {
        double m = fmin(fmin(red[0], red[1]), fmin(red[2], red[3]));
        // positive doubles order like their bit patterns
        atomicMin(out, (unsigned long long)__double_as_longlong(m));
        if (nf) atomicOr(bad, 1u);
    }
}

// ---------------------------------------------------------------------------
// device time control: the driver's dt sequence (Source/EBR.cpp:189-306)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double red_value(unsigned long long bits) { return __longlong_as_double((long long)bits); }
__device__ __forceinline__ unsigned long long red_big()
{
    return (unsigned long long)__double_as_longlong(std::numeric_limits<double>::max());
}

__global__ void k_ctl_init(ocfd_step_ctl* c, double time, double cfl, double stop_time, int max_steps)
{
    c->time = time; c->dt = 0.0; c->cfl = cfl; c->stop_time = stop_time;
    c->steps = 0; c->max_steps = max_steps; c->done = 0; c->status = OCFD_OK;
    atomicExch(&c->red, red_big());
    atomicExch(&c->bad, 0u);
}

// a non-finite cell seen by the last reduction ends the run with a status
__device__ void ctl_check_bad(ocfd_step_ctl* c)
{
    if (atomicExch(&c->bad, 0u)) {
        c->status = OCFD_ENONFINITE;
        c->done = 1;
    }
}

// dt_new from the CFL minimum, clipped so the run ends on stop_time; then the
// done flag (the driver's loop condition, Source/main.cpp:61-70)
__device__ void ctl_set_dt(ocfd_step_ctl* c, double dt_new, double time, int steps)
{
#pragma clang fp contract(off)
    const double stop = c->stop_time;
    const int max_steps = c->max_steps;
    const double eps = 0.001 * dt_new;
    if (stop >= 0.0 && time + dt_new > stop - eps) dt_new = stop - time;
    c->dt = dt_new;
    c->time = time;
    c->steps = steps;
    c->done = (int)((max_steps >= 0 && steps >= max_steps) || (stop >= 0.0 && !(time < stop)));
}

// initial dt: cfl * estdt (EBR::computeInitialDt, Source/EBR.cpp:189-229 with
// the single-level ratio 1)
__global__ void k_ctl_first(ocfd_step_ctl* c)
{
#pragma clang fp contract(off)
    const double dt = red_value(atomicExch(&c->red, red_big())) * c->cfl;
    ctl_set_dt(c, dt, c->time, c->steps);
    ctl_check_bad(c);
}

// after a step: t += dt, then dt_new = min(cfl*estdt, change_max*dt) with
// change_max 1.1 and the stop_time clip (EBR::computeNewDt, Source/EBR.cpp:231-306)
__global__ void k_ctl_update(ocfd_step_ctl* c)
{
#pragma clang fp contract(off)
    if (c->done) return;
    const double est = red_value(atomicExch(&c->red, red_big())) * c->cfl;
    const double dt = c->dt;
    const double time = c->time + dt;
    ctl_set_dt(c, fmin(est, 1.1 * dt), time, c->steps + 1);
    ctl_check_bad(c);
}

// the CFL estimate replaced by a fixed value (the CHEM build's estTimeStep,
// Source/Advance.cpp:54-57), after a reduction that only checked finiteness
__global__ void k_ctl_fixed_est(ocfd_step_ctl* c, double est)
{
    if (c->done) return;
    atomicExch(&c->red, (unsigned long long)__double_as_longlong(est));
}

// A <- B for the forward-Euler step under device control (a no-op once done)
__global__ __launch_bounds__(256) void k_copy_ctl(double* __restrict__ dst, const double* __restrict__ src, long n,
                                                  const ocfd_step_ctl* __restrict__ ctl)
{
    if (ctl_done(ctl)) return;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) dst[i] = src[i];
}


// Source/Setup.cpp:21-34
static const signed char scalar_bc[6]   = {BC_INT, BC_EXT, BC_FOEXT, BC_EVEN, BC_EVEN, BC_EVEN};
static const signed char norm_vel_bc[6] = {BC_INT, BC_EXT, BC_FOEXT, BC_ODD,  BC_ODD,  BC_ODD};
static const signed char tang_vel_bc[6] = {BC_INT, BC_EXT, BC_FOEXT, BC_EVEN, BC_EVEN, BC_ODD};

hipStream_t upload_stream()
{
    static std::mutex mu;
    static std::map<int, hipStream_t> streams;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    auto it = streams.find(dev);
    if (it != streams.end()) return it->second;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
    streams[dev] = s;
    return s;
}

template <class T>
static int upload(T** dptr, const std::vector<T>& h)
{
    if (h.empty()) { *dptr = nullptr; return OCFD_OK; }
    if (!dev_alloc((void**)dptr, sizeof(T) * h.size())) {
        set_error("out of device memory for the level tables");
        return OCFD_ENOMEM;
    }
    return upload_bytes(*dptr, h.data(), sizeof(T) * h.size());
}

// The level's read-only tables (owner map, work-lists, box descriptors, EB
// geometry) share one device allocation (Staging): a level is rebuilt at
// every regrid, and an allocation and a copy per table cost more than the
// tables themselves.
static void free_level(LevelDev& L)
{
    for (auto& e : L.tcache) dev_free({e.buf});
    for (hipEvent_t e : L.gl_event)
        if (e) (void)hipEventDestroy(e);
    dev_free({L.d_gl[0], L.d_gl[1]});
    // tables inside the arena are not freed on their own
    dev_free({L.d_arena, L.d_bufs, L.own_tile_yz ? (void*)L.d_tile_y : nullptr,
              L.own_tile_yz ? (void*)L.d_tile_z : nullptr, L.d_ebq, L.d_ebF[0], L.d_ebF[1], L.d_ebF[2], L.d_mF[0],
              L.d_mF[1], L.d_mF[2], L.d_mq});
    L = LevelDev{};
}

__global__ void k_zero_u32x2(unsigned* p) { p[threadIdx.x] = 0u; }

// The ghost lists of depth d (k_ghost_lists) of a level that does not cover
// its domain, each with room for every ghost cell of that depth, built on
// stream s: the lengths cleared by a kernel, then one classification pass.
bool build_ghost_lists(const LevelDev& L, int d, hipStream_t s)
{
    const int4* tiles = d ? L.d_tile_gp : L.d_tile_g;
    const int ntile = d ? L.ntile_gp : L.ntile_g;
    if (!ntile) return false;
    long cap = 0;
    for (const BoxDesc& B : L.boxes) cap += ghost_count(B, d ? 1 : NG);
    const size_t bytes = 256 + sizeof(int2) * 2 * (size_t)cap;
    void* mem = nullptr;
    if (!dev_alloc(&mem, bytes)) return false;
    unsigned* cnt = (unsigned*)mem;   // the two lengths on the first 256 bytes
    int2* p = (int2*)((char*)mem + 256);
    L.gl[d][0] = LevelDev::GhostList{p, cnt, (int)cap};
    L.gl[d][1] = LevelDev::GhostList{p + cap, cnt + 1, (int)cap};
    hipLaunchKernelGGL(k_zero_u32x2, dim3(1), dim3(2), 0, s, cnt);
    hipLaunchKernelGGL(k_ghost_lists, dim3(ntile), dim3(256), 0, s, L.d_boxes, L.d_geom, tiles, L.gl[d][0].d,
                       L.gl[d][1].d, cnt);
    hipEvent_t ev = nullptr;
    if (hipGetLastError() != hipSuccess || hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess ||
        hipEventRecord(ev, s) != hipSuccess) {
        if (ev) (void)hipEventDestroy(ev);
        dev_free({mem});
        return false;
    }
    L.d_gl[d] = mem;
    L.gl_stream[d] = s;
    L.gl_event[d] = ev;
    return true;
}

// whether a level needs ghost lists: its boxes (disjoint) do not cover the
// domain, so some ghost cells may have no same-level or physical source
static void set_ghost_need(LevelDev& L)
{
    long vol = 0, dom = 1;
    for (const BoxDesc& B : L.boxes)
        vol += (long)(B.hi[0] - B.lo[0] + 1) * (B.hi[1] - B.lo[1] + 1) * (B.hi[2] - B.lo[2] + 1);
    for (int d = 0; d < 3; ++d) dom *= L.geom.dom_hi[d] - L.geom.dom_lo[d] + 1;
    L.gl_need = vol < dom && L.ntile_g > 0;
}

} // namespace ocfd

using namespace ocfd;

extern "C" {

const char* ocfd_last_error(void) { return g_err.c_str(); }

int ocfd_parm_initialize(ocfd_parm* p)
{
    if (!p) { set_error("ocfd_parm_initialize: null parm"); return OCFD_EINVAL; }
    if (p->eos_gamma == 0.0) {
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
    }
    p->Rg = p->Ru / p->eos_m;
    p->cv = p->Rg / (p->eos_gamma - 1.0);
    p->cp = p->eos_gamma * p->cv;
    p->kOverMu = p->cp / p->Pr;
    return OCFD_OK;
}

// march tiles of the y and z sweeps over the regular boxes: i segment fastest,
// then chunk, then row group, so a contiguous range of the list (one XCD's
// share) holds tiles that share chunk halos and transverse neighbour rows;
// .w = first cell of the chunk (low 20 bits) | chunk length << 20
// ty holds the 3D y-march tiles, then (from index ny3) those of the planar
// first-layer y march: one wave per tile, packing 64 / w sub-rows of w lanes
// (w the power of two >= the box width, at most 64), sub-row k taking chunk
// c0 + k of the same x segment, so narrow boxes do not leave lanes idle;
// .y = x segment | log2(w) << 16
static void march_tiles(const LevelDev& L, int ch, std::vector<int4>& ty, std::vector<int4>& tz, int& ny3)
{
    ty.clear();
    tz.clear();
    for (int b = 0; b < L.nbox; ++b) {
        const BoxDesc& B = L.boxes[b];
        if (B.type != 0) continue;
        const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1, nz = B.hi[2] - B.lo[2] + 1;
        const int iseg = (nx + 63) / 64;
        for (int kg = 0; kg < (nz + 3) / 4; ++kg)
            for (int c = 0; c < (ny + ch - 1) / ch; ++c)
                for (int s = 0; s < iseg; ++s) ty.push_back(make_int4(b, s, kg, (c * ch) | (ch << 20)));
        for (int jg = 0; jg < (ny + 3) / 4; ++jg)
            for (int c = 0; c < (nz + ch - 1) / ch; ++c)
                for (int s = 0; s < iseg; ++s) tz.push_back(make_int4(b, s, jg, (c * ch) | (ch << 20)));
    }
    ny3 = (int)ty.size();
    for (int b = 0; b < L.nbox; ++b) {
        const BoxDesc& B = L.boxes[b];
        if (B.type != 0) continue;
        const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1;
        int lw = 1;
        while (lw < 6 && (1 << lw) < nx) ++lw;
        const int nsub = 64 >> lw, iseg = (nx + (1 << lw) - 1) >> lw, nc = (ny + ch - 1) / ch;
        for (int c0 = 0; c0 < nc; c0 += nsub)
            for (int s = 0; s < iseg; ++s) ty.push_back(make_int4(b, s | (lw << 16), 0, (c0 * ch) | (ch << 20)));
    }
}

int ocfd_level_create(int nbox, const int* lo, const int* hi, const ocfd_geom* g, ocfd_level* out)
{
    return ocfd_level_create_eb(nbox, lo, hi, g, nullptr, out);
}

int ocfd_level_box_types(ocfd_level lev, int* types)
{
    if (!lev || !types) { set_error("ocfd_level_box_types: invalid arguments"); return OCFD_EINVAL; }
    for (int b = 0; b < lev->dev.nbox; ++b) types[b] = lev->dev.boxes[b].type;
    return OCFD_OK;
}

int ocfd_level_create_eb(int nbox, const int* lo, const int* hi, const ocfd_geom* g, const ocfd_eb_spec* spec,
                         ocfd_level* out)
{
    if (!out || !g || nbox <= 0 || !lo || !hi) { set_error("ocfd_level_create: invalid arguments"); return OCFD_EINVAL; }
    *out = nullptr;
    ocfd_level lev = new ocfd_level_s;
    LevelDev& L = lev->dev;
    L.nbox = nbox;
    long off_s = 0, off_d = 0;
    for (int b = 0; b < nbox; ++b) {
        BoxDesc B{};
        for (int d = 0; d < 3; ++d) {
            B.lo[d] = lo[3 * b + d];
            B.hi[d] = hi[3 * b + d];
            if (B.hi[d] < B.lo[d] || B.lo[d] < g->dom_lo[d] || B.hi[d] > g->dom_hi[d]) {
                set_error("ocfd_level_create: box " + std::to_string(b) + " empty or outside the domain");
                delete lev;
                return OCFD_EINVAL;
            }
            if (B.hi[d] - B.lo[d] + 1 >= (1 << 20)) {   // march tiles hold in-box offsets in 20 bits
                set_error("ocfd_level_create: box " + std::to_string(b) + " has 2^20 cells or more along a side");
                delete lev;
                return OCFD_EINVAL;
            }
            B.dx[d] = g->dx[d];
            B.dxinv[d] = 1.0 / g->dx[d];
        }
        B.nxg = B.hi[0] - B.lo[0] + 1 + 2 * NG;
        B.nyg = B.hi[1] - B.lo[1] + 1 + 2 * NG;
        B.nzg = B.hi[2] - B.lo[2] + 1 + 2 * NG;
        B.nptg = (long)B.nxg * B.nyg * B.nzg;
        B.nptv = (long)(B.hi[0] - B.lo[0] + 1) * (B.hi[1] - B.lo[1] + 1) * (B.hi[2] - B.lo[2] + 1);
        B.off_s = off_s;
        B.off_d = off_d;
        B.off_p = off_s / NC * 6;
        B.off_e = -1;
        B.type = 0;
        B.geom = 0;
        off_s += B.nptg * NC;
        off_d += B.nptv * NC;
        L.boxes.push_back(B);
    }
    L.mf_state = off_s;
    L.mf_valid = off_d;

    // embedded boundary: per-box geometry over grow(bx,4) and the box kinds of
    // EBR::compute_dSdt (Source/Advance.cpp:213,231)
    std::vector<uint32_t> eb_flag;
    std::vector<double> eb_geo;
    if (spec && spec->kind != OCFD_EB_NONE) {
        for (int b = 0; b < nbox; ++b) {
            BoxDesc& B = L.boxes[b];
            std::vector<uint32_t> fl(B.nptg);
            std::vector<double> vf(B.nptg), ax(B.nptg), ay(B.nptg), az(B.nptg);
            int rc = ocfd_eb_geometry(spec, g, B.lo, B.hi, NG, fl.data(), vf.data(), ax.data(), ay.data(), az.data(),
                                      nullptr, nullptr, nullptr, nullptr);
            if (rc) { set_error("ocfd_level_create_eb: bad embedded-boundary spec"); delete lev; return rc; }
            bool allreg = true, allcov = true;
            for (long c = 0; c < B.nptg; ++c) allreg = allreg && (fl[c] & 3u) == 0u;
            for (int k = 0; k <= B.hi[2] - B.lo[2]; ++k)
                for (int j = 0; j <= B.hi[1] - B.lo[1]; ++j)
                    for (int i = 0; i <= B.hi[0] - B.lo[0]; ++i)
                        allcov = allcov && (fl[((long)(k + NG) * B.nyg + j + NG) * B.nxg + i + NG] & 3u) == 3u;
            B.type = allcov ? 2 : (allreg ? 0 : 1);
            L.has_eb = true;
            L.eb_spec = *spec;
            if (B.type == 1) {
                B.off_e = L.mf_eb;
                L.mf_eb += B.nptg;
                eb_flag.insert(eb_flag.end(), fl.begin(), fl.end());
                // geometry comps box by box: vfrac, apx, apy, apz FABs
                eb_geo.insert(eb_geo.end(), vf.begin(), vf.end());
                eb_geo.insert(eb_geo.end(), ax.begin(), ax.end());
                eb_geo.insert(eb_geo.end(), ay.begin(), ay.end());
                eb_geo.insert(eb_geo.end(), az.begin(), az.end());
                ++L.n_cut;
            } else if (B.type == 2) {
                ++L.n_covered;
            }
        }
    }

    GeomDesc& G = L.geom;
    for (int d = 0; d < 3; ++d) {
        G.dom_lo[d] = g->dom_lo[d]; G.dom_hi[d] = g->dom_hi[d];
        G.periodic[d] = g->periodic[d];
        G.prob_lo[d] = g->prob_lo[d]; G.dx[d] = g->dx[d];
        for (int side = 0; side < 2; ++side) {
            int phys = side ? g->hi_bc[d] : g->lo_bc[d];
            if (g->periodic[d]) phys = 0;
            if (phys < 0 || phys > 5) { set_error("ocfd_level_create: bad BC code"); delete lev; return OCFD_EINVAL; }
            for (int n = 0; n < NC; ++n) {
                signed char bc;
                if (n == URHO || n == UEDEN) bc = scalar_bc[phys];
                else bc = (n - UMX) == d ? norm_vel_bc[phys] : tang_vel_bc[phys];
                G.bc[n][d][side] = bc;
            }
        }
    }
    G.probtype = g->probtype;
    for (int n = 0; n < NC; ++n) { G.left[n] = g->left_state[n]; G.right[n] = g->right_state[n]; }
    G.has_spec_in = 0;

    // owner map: tiles of 8^3 cells, with box lists where boxes share a tile
    G.map_shift = kMapShift;
    if (!build_tile_map(G.dom_lo, G.dom_hi, kMapShift, L.boxes, 0, G.map_n, L.owner_map, L.owner_list)) {
        set_error("ocfd_level_create: overlapping boxes");
        delete lev;
        return OCFD_EINVAL;
    }

    // work-lists
    const int xseg = tile_xseg();
    // march chunk length: the full chunk unless the level has too few march
    // tiles to fill the chip (small or thin levels), where shorter chunks
    // trade a longer halo prologue per cell for more independent waves
    int ch = tile_chunk();
    auto march_count = [&](int c) {
        long ny_t = 0, nz_t = 0;
        for (const BoxDesc& B : L.boxes) {
            if (B.type != 0) continue;
            const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1, nz = B.hi[2] - B.lo[2] + 1;
            const long is = (nx + 63) / 64;
            ny_t += is * ((nz + 3) / 4) * ((ny + c - 1) / c);
            nz_t += is * ((ny + 3) / 4) * ((nz + c - 1) / c);
        }
        return std::max(ny_t, nz_t);
    };
    while (ch > 2 && march_count(ch) < 2048) ch /= 2;
    std::vector<int4> tx, txz, txp, ty, tz, tg, tgp, tc, tc1, tp, tpp, tv, teq, tef[3], ted, tmf[3], tmd;
    for (int b = 0; b < nbox; ++b) {
        const BoxDesc& B = L.boxes[b];
        const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1, nz = B.hi[2] - B.lo[2] + 1;
        const long ngh = ghost_count(B, NG);
        for (long s = 0; s < ngh; s += 256) tg.push_back(make_int4(b, (int)s, (int)std::min<long>(256, ngh - s), NG));
        // planar fill: the z ghost layer below, the y and x ghosts of the
        // first two layers (the cells the first-layer sweeps read)
        const long ngp = ghost_count(B, 1);
        for (long s = 0; s < ngp; s += 256) tgp.push_back(make_int4(b, (int)s, (int)std::min<long>(256, ngp - s), 1));
        for (long s = 0; s < B.nptv; s += 2048) tc.push_back(make_int4(b, (int)s, (int)std::min<long>(2048, B.nptv - s), 0));
        for (long s = 0; s < B.nptv; s += 256) tc1.push_back(make_int4(b, (int)s, (int)std::min<long>(256, B.nptv - s), 0));
        for (long s = 0; s < B.nptv; s += 64) tv.push_back(make_int4(b, (int)s, (int)std::min<long>(64, B.nptv - s), 0));
        if (B.type != 0) {
            if (B.type == 1) {
                const long nq = (long)(nx + 4) * (ny + 4) * (nz + 4);
                for (long s = 0; s < nq; s += 256) teq.push_back(make_int4(b, (int)s, (int)std::min<long>(256, nq - s), 2));
                for (int d = 0; d < 3; ++d) {
                    const long nf = (long)(nx + (d == 0)) * (ny + (d == 1)) * (nz + (d == 2));
                    for (long s = 0; s < nf; s += 256)
                        tef[d].push_back(make_int4(b, (int)s, (int)std::min<long>(256, nf - s), d));
                }
            }
            for (long s = 0; s < B.nptv; s += 256) ted.push_back(make_int4(b, (int)s, (int)std::min<long>(256, B.nptv - s), 0));
            continue;
        }
        for (int d = 0; d < 3; ++d) {
            const long nf = (long)(nx + (d == 0)) * (ny + (d == 1)) * (nz + (d == 2));
            for (long s = 0; s < nf; s += 256) tmf[d].push_back(make_int4(b, (int)s, (int)std::min<long>(256, nf - s), d));
        }
        for (long s = 0; s < B.nptv; s += 256) tmd.push_back(make_int4(b, (int)s, (int)std::min<long>(256, B.nptv - s), 0));
        // x sweep: each z plane of the primitive region grow(bx, (3,3,gz)) is
        // one flattened run of (ny+6) rows of nx+6 cells, cut into 64-cell
        // windows that advance by xseg cells; a block takes 4 windows.  Listed
        // in slabs of 8 z planes, window group by window group, so the rows
        // whose states a row's viscous faces read (y +- 1, z +- 1) are taken
        // up close in time, while their lines are still in L2.
        const long nrun = (long)(ny + 6) * (nx + 6);
        const int nwin = (int)std::max<long>(1, (nrun - 6 + xseg - 1) / xseg), ngw = (nwin + 3) / 4;
        for (int k0 = 0; k0 < nz + 6; k0 += 8)
            for (int g = 0; g < ngw; ++g)
                for (int kr = k0; kr < std::min(k0 + 8, nz + 6); ++kr) tx.push_back(make_int4(b, 4 * g, kr, 3));
        // planar (z growth 1): the first layer and its two neighbours, which a
        // z-uniform level sweeps, for all boxes first (txz), then the other layers
        for (int g = 0; g < ngw; ++g)
            for (int kr = 0; kr < 3; ++kr) txz.push_back(make_int4(b, 4 * g, kr, 1));
        for (int k0 = 3; k0 < nz + 2; k0 += 8)
            for (int g = 0; g < ngw; ++g)
                for (int kr = k0; kr < std::min(k0 + 8, nz + 2); ++kr) txp.push_back(make_int4(b, 4 * g, kr, 1));
        // march tiles: i segment fastest, then chunk, then row group, so a
        // contiguous range of the list (one XCD's share) holds tiles that
        // share chunk halos and transverse neighbour rows

        for (int gz : {3, 1}) {
            const long np = (long)(nx + 6) * (ny + 6) * (nz + 2 * gz);
            std::vector<int4>& tl = gz == 3 ? tp : tpp;
            for (long s = 0; s < np; s += 256) tl.push_back(make_int4(b, (int)s, (int)std::min<long>(256, np - s), gz));
        }
    }
    L.ntile_xz = (int)txz.size();
    txz.insert(txz.end(), txp.begin(), txp.end());
    txp.swap(txz);
    L.march_ch = ch;
    int ny3 = 0;
    march_tiles(L, ch, ty, tz, ny3);
    int rc = OCFD_OK;
    Staging stg;
    stg.add(&L.d_map, L.owner_map); stg.add(&L.d_map_list, L.owner_list);
    stg.add(&L.d_tile_x, tx); stg.add(&L.d_tile_xp, txp); stg.add(&L.d_tile_g, tg); stg.add(&L.d_tile_gp, tgp); stg.add(&L.d_tile_c, tc); stg.add(&L.d_tile_c1, tc1);
    stg.add(&L.d_tile_p, tp); stg.add(&L.d_tile_pp, tpp); stg.add(&L.d_tile_v, tv); stg.add(&L.d_tile_eq, teq);
    for (int d = 0; d < 3; ++d) stg.add(&L.d_tile_ef[d], tef[d]);
    stg.add(&L.d_tile_ed, ted); stg.add(&L.d_ebflag, eb_flag); stg.add(&L.d_ebgeo, eb_geo);
    for (int d = 0; d < 3; ++d) stg.add(&L.d_tile_mf[d], tmf[d]);
    stg.add(&L.d_tile_md, tmd);
    stg.add(&L.d_boxes, L.boxes);
    stg.add(&L.d_tile_y, ty); stg.add(&L.d_tile_z, tz);
    const size_t geom_off = stg.reserve(&L.d_geom, 1);
    if ((rc = stg.allocate(&L.d_arena))) { free_level(L); delete lev; return rc; }
    L.ntile_eq = (int)teq.size(); L.ntile_ed = (int)ted.size(); L.ntile_md = (int)tmd.size();
    for (int d = 0; d < 3; ++d) L.ntile_mf[d] = (int)tmf[d].size();
    for (int d = 0; d < 3; ++d) L.ntile_ef[d] = (int)tef[d].size();
    if (L.mf_eb) {
        bool ok = dev_alloc((void**)&L.d_ebq, sizeof(double) * 7 * L.mf_eb);
        for (int d = 0; d < 3 && ok; ++d) ok = dev_alloc((void**)&L.d_ebF[d], sizeof(double) * NC * L.mf_eb);
        if (!ok) { set_error("ocfd_level_create_eb: out of device memory"); free_level(L); delete lev; return OCFD_ENOMEM; }
    }
    L.ntile_x = (int)tx.size(); L.ntile_xp = (int)txp.size(); L.ntile_y = ny3; L.ntile_z = (int)tz.size();
    L.ntile_yp = (int)ty.size() - ny3;
    L.ntile_g = (int)tg.size(); L.ntile_gp = (int)tgp.size(); L.ntile_c = (int)tc.size(); L.ntile_c1 = (int)tc1.size();
    L.ntile_p = (int)tp.size(); L.ntile_pp = (int)tpp.size(); L.ntile_v = (int)tv.size();
    L.mf_prim = L.mf_state / NC * 6;
    G.map = L.d_map;
    G.map_list = L.d_map_list;
    stg.put(geom_off, G);
    if ((rc = stg.copy(L.d_arena))) { free_level(L); delete lev; return rc; }
    // stage scratch in one allocation: partial divergence, FP32 and FP64
    // primitives, the reduction slot
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t o_p32 = al(sizeof(double) * L.mf_valid), o_p64 = o_p32 + al(sizeof(float) * L.mf_prim),
                 o_red = o_p64 + al(sizeof(double) * L.mf_prim), nbufs = o_red + 256;
    if (!dev_alloc(&L.d_bufs, nbufs)) {
        L.d_bufs = nullptr;
        set_error("ocfd_level_create: out of device memory");
        free_level(L); delete lev; return OCFD_ENOMEM;
    }
    L.d_acc = (double*)L.d_bufs;
    L.d_prim32 = (float*)((char*)L.d_bufs + o_p32);
    L.d_prim64 = (double*)((char*)L.d_bufs + o_p64);
    L.d_red = (unsigned long long*)((char*)L.d_bufs + o_red);
    L.dev_bytes = (long)sizeof(int) * (long)(L.owner_map.size() + L.owner_list.size()) +
                  (long)sizeof(int4) * (L.ntile_x + L.ntile_xp + L.ntile_y + L.ntile_yp + L.ntile_z + L.ntile_g + L.ntile_gp + L.ntile_c + L.ntile_c1 + L.ntile_p +
                                        L.ntile_pp + L.ntile_v) +
                  (long)sizeof(GeomDesc) + (long)sizeof(BoxDesc) * L.nbox + (long)sizeof(double) * L.mf_valid +
                  (long)(sizeof(float) + sizeof(double)) * L.mf_prim + (long)sizeof(unsigned long long) +
                  (long)sizeof(int4) * (L.ntile_eq + L.ntile_ef[0] + L.ntile_ef[1] + L.ntile_ef[2] + L.ntile_ed) +
                  L.mf_eb * (long)(sizeof(uint32_t) + sizeof(double) * (4 + 7 + 3 * NC));
    set_ghost_need(L);
    *out = lev;
    return OCFD_OK;
}

int ocfd_level_destroy(ocfd_level lev)
{
    if (!lev) return OCFD_OK;
    if (lev->dev.live_graphs > 0) {
        set_error("ocfd_level_destroy: step graphs of this level still exist (destroy them first)");
        return OCFD_EINVAL;
    }
    free_level(lev->dev);
    delete lev;
    return OCFD_OK;
}

int ocfd_level_nbox(ocfd_level lev) { return lev ? lev->dev.nbox : 0; }

int ocfd_level_set_inflow_species(ocfd_level lev, const double* spec)
{
    if (!lev) { set_error("ocfd_level_set_inflow_species: invalid arguments"); return OCFD_EINVAL; }
    LevelDev& L = lev->dev;
    L.geom.has_spec_in = spec != nullptr;
    for (int n = 0; n < OCFD_NSPEC; ++n) L.geom.spec_in[n] = spec ? spec[n] : 0.0;
    OCFD_HIP(hipDeviceSynchronize());   // no launch may still read the old table
    OCFD_HIP(hipMemcpy(L.d_geom, &L.geom, sizeof(GeomDesc), hipMemcpyHostToDevice));
    return OCFD_OK;
}

int ocfd_level_set_planar(ocfd_level lev, int on)
{
    if (!lev) { set_error("ocfd_level_set_planar: invalid arguments"); return OCFD_EINVAL; }
    if (on && lev->dev.n_cut) {
        set_error("ocfd_level_set_planar: cut boxes read two z ghost layers; keep the full fill depth");
        return OCFD_EUNSUPPORTED;
    }
    lev->dev.fill_gz = on ? 1 : NG;
    return OCFD_OK;
}

long ocfd_level_device_bytes(ocfd_level lev) { return lev ? lev->dev.dev_bytes : -1; }

long ocfd_level_mf_size(ocfd_level lev, int ngrow, int ncomp)
{
    if (!lev) return -1;
    long n = 0;
    for (const BoxDesc& B : lev->dev.boxes) {
        long c = 1;
        for (int d = 0; d < 3; ++d) c *= (B.hi[d] - B.lo[d] + 1 + 2 * ngrow);
        n += c * ncomp;
    }
    return n;
}

long ocfd_level_fab_offset(ocfd_level lev, int box, int ngrow, int ncomp)
{
    if (!lev || box < 0 || box >= lev->dev.nbox) return -1;
    long n = 0;
    for (int b = 0; b < box; ++b) {
        const BoxDesc& B = lev->dev.boxes[b];
        long c = 1;
        for (int d = 0; d < 3; ++d) c *= (B.hi[d] - B.lo[d] + 1 + 2 * ngrow);
        n += c * ncomp;
    }
    return n;
}

// stage flags, and a planar fill depth only with planar stepping
static int valid_flags(const LevelDev& L, int flags)
{
    if (L.fill_gz != NG && !(flags & OCFD_PLANAR)) return 0;
    return (flags & ~(OCFD_VISC | OCFD_GRAVITY | OCFD_FLUX_FP32 | OCFD_PLANAR | OCFD_EB_REDIST)) == 0;
}

int ocfd_compute_dsdt(ocfd_level lev, const double* S, double* dSdt, double dt, const ocfd_parm* parm, int flags,
                      void* stream)
{
    (void)dt;
    if (!lev || !S || !dSdt || !parm || !valid_flags(lev->dev, flags)) { set_error("ocfd_compute_dsdt: invalid arguments"); return OCFD_EINVAL; }
    StageArgs a{};
    a.scur = S; a.sold = nullptr; a.snew = nullptr; a.acc = dSdt;
    a.c_cur = 1.0; a.c_old = 0.0; a.c_dt = 0.0;
    a.flags = flags; a.ph = phys_of(*parm);
    launch_sweeps(lev->dev, a, (hipStream_t)stream);
    return check_launch("ocfd_compute_dsdt");
}

int ocfd_rk_stage(ocfd_level lev, const double* S_cur, const double* S_old, double* S_new, double c_cur, double c_old,
                  double c_dt, const ocfd_parm* parm, int flags, void* stream)
{
    if (!lev || !S_cur || !S_new || !parm || !valid_flags(lev->dev, flags) || S_cur == S_new) {
        set_error("ocfd_rk_stage: invalid arguments (S_new must differ from S_cur)");
        return OCFD_EINVAL;
    }
    StageArgs a{};
    a.scur = S_cur; a.sold = S_old; a.snew = S_new; a.acc = lev->dev.d_acc;
    a.c_cur = c_cur; a.c_old = c_old; a.c_dt = c_dt;
    a.flags = flags; a.ph = phys_of(*parm);
    launch_sweeps(lev->dev, a, (hipStream_t)stream);
    return check_launch("ocfd_rk_stage");
}

int ocfd_fill_boundary(ocfd_level lev, double* S, double time, void* stream)
{
    if (!lev || !S) { set_error("ocfd_fill_boundary: invalid arguments"); return OCFD_EINVAL; }
    launch_fill_ctl(lev->dev, S, nullptr, 0.0, (hipStream_t)stream, time);
    return check_launch("ocfd_fill_boundary");
}

int ocfd_estdt(ocfd_level lev, const double* S, double cfl, const ocfd_parm* parm, double* dt_out, void* stream)
{
    if (!lev || !S || !parm || !dt_out) { set_error("ocfd_estdt: invalid arguments"); return OCFD_EINVAL; }
    const LevelDev& L = lev->dev;
    hipStream_t s = (hipStream_t)stream;
    unsigned int* bad = (unsigned int*)(L.d_red + 1);
    hipLaunchKernelGGL(k_red_init, dim3(1), dim3(1), 0, s, L.d_red, bad);
    hipLaunchKernelGGL(k_estdt, dim3(L.ntile_c), dim3(256), 0, s, L.d_boxes, L.d_tile_c, S, phys_of(*parm), L.d_red,
                       bad, (const ocfd_step_ctl*)nullptr);
    int rc = check_launch("ocfd_estdt");
    if (rc) return rc;
    unsigned long long res[2] = {0, 0};   // the slot and the flag word after it
    OCFD_HIP(hipMemcpyAsync(res, L.d_red, sizeof(res), hipMemcpyDeviceToHost, s));
    OCFD_HIP(hipStreamSynchronize(s));
    double m;
    std::memcpy(&m, &res[0], sizeof(m));
    *dt_out = m * cfl;
    if ((unsigned int)res[1]) {
        set_error("ocfd_estdt: the state holds a non-finite value");
        return OCFD_ENONFINITE;
    }
    return OCFD_OK;
}

int ocfd_amr_estdt(int nlev, const ocfd_level* levels, const double* const* states, double cfl,
                   const ocfd_parm* parm, double* dt_out, void* stream)
{
    if (nlev < 1 || !levels || !states || !parm || !dt_out) {
        set_error("ocfd_amr_estdt: invalid arguments");
        return OCFD_EINVAL;
    }
    for (int l = 0; l < nlev; ++l)
        if (!levels[l] || !states[l]) {
            set_error("ocfd_amr_estdt: level " + std::to_string(l) + " or its state is null");
            return OCFD_EINVAL;
        }
    hipStream_t s = (hipStream_t)stream;
    const PhysArgs ph = phys_of(*parm);
    for (int l = 0; l < nlev; ++l) {
        const LevelDev& L = levels[l]->dev;
        unsigned int* bad = (unsigned int*)(L.d_red + 1);
        hipLaunchKernelGGL(k_red_init, dim3(1), dim3(1), 0, s, L.d_red, bad);
        hipLaunchKernelGGL(k_estdt, dim3(L.ntile_c), dim3(256), 0, s, L.d_boxes, L.d_tile_c, states[l], ph, L.d_red,
                           bad, (const ocfd_step_ctl*)nullptr);
    }
    int rc = check_launch("ocfd_amr_estdt");
    if (rc) return rc;
    // every level's slot and flag word into one pinned buffer, one wait
    static std::mutex mu;
    static unsigned long long* pin = nullptr;
    static int cap = 0;
    std::lock_guard<std::mutex> lk(mu);
    if (cap < nlev) {
        if (pin) (void)hipHostFree(pin);
        pin = nullptr;
        cap = 0;
        OCFD_HIP(hipHostMalloc((void**)&pin, sizeof(unsigned long long) * 2 * nlev, hipHostMallocDefault));
        cap = nlev;
    }
    for (int l = 0; l < nlev; ++l)
        OCFD_HIP(hipMemcpyAsync(pin + 2 * l, levels[l]->dev.d_red, 2 * sizeof(unsigned long long),
                                hipMemcpyDeviceToHost, s));
    OCFD_HIP(hipStreamSynchronize(s));
    bool bad = false;
    for (int l = 0; l < nlev; ++l) {
        double m;
        std::memcpy(&m, &pin[2 * l], sizeof(m));
        dt_out[l] = m * cfl;
        bad = bad || (unsigned int)pin[2 * l + 1] != 0u;
    }
    if (bad) {
        set_error("ocfd_amr_estdt: a level's state holds a non-finite value");
        return OCFD_ENONFINITE;
    }
    return OCFD_OK;
}

int ocfd_advance(ocfd_level lev, double* A, double* B, double* C, double time, double dt, int scheme,
                 const ocfd_parm* parm, int flags, void* stream)
{
    if (!lev || !A || !B || !parm || (scheme == 2 && !C) || scheme < 0 || scheme > 2) {
        set_error("ocfd_advance: invalid arguments");
        return OCFD_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    int rc;
    if ((rc = ocfd_fill_boundary(lev, A, time, s))) return rc;
    if (scheme == 0) {
        // Advance.cpp:117-122
        if ((rc = ocfd_rk_stage(lev, A, nullptr, B, 1.0, 0.0, dt, parm, flags, s))) return rc;
        OCFD_HIP(hipMemcpyAsync(A, B, sizeof(double) * lev->dev.mf_state, hipMemcpyDeviceToDevice, s));
    } else if (scheme == 1) {
        // Advance.cpp:123-139; the last stage writes A in place (it reads A only at its own cell)
        if ((rc = ocfd_rk_stage(lev, A, nullptr, B, 1.0, 0.0, dt, parm, flags, s))) return rc;
        if ((rc = ocfd_fill_boundary(lev, B, time + dt, s))) return rc;
        if ((rc = ocfd_rk_stage(lev, B, A, A, 0.5, 0.5, 0.5 * dt, parm, flags, s))) return rc;
    } else {
        // SSP-RK3, stage times t, t+dt, t+dt/2 (Advance.cpp:141-148)
        if ((rc = ocfd_rk_stage(lev, A, nullptr, B, 1.0, 0.0, dt, parm, flags, s))) return rc;
        if ((rc = ocfd_fill_boundary(lev, B, time + dt, s))) return rc;
        if ((rc = ocfd_rk_stage(lev, B, A, C, 0.25, 0.75, 0.25 * dt, parm, flags, s))) return rc;
        if ((rc = ocfd_fill_boundary(lev, C, time + 0.5 * dt, s))) return rc;
        if ((rc = ocfd_rk_stage(lev, C, A, A, 2.0 / 3.0, 1.0 / 3.0, (2.0 / 3.0) * dt, parm, flags, s))) return rc;
    }
    return OCFD_OK;
}

int ocfd_profile_stage(ocfd_level lev, const double* S_cur, const double* S_old, double* S_new, double c_cur,
                       double c_old, double c_dt, const ocfd_parm* parm, int flags, int reps, double* ms_out,
                       void* stream)
{
    if (!lev || !S_cur || !S_new || !parm || !ms_out || reps <= 0 || !valid_flags(lev->dev, flags) || S_cur == S_new ||
        (flags & OCFD_PLANAR)) {
        set_error("ocfd_profile_stage: invalid arguments");
        return OCFD_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    StageArgs a{};
    a.scur = S_cur; a.sold = S_old; a.snew = S_new; a.acc = lev->dev.d_acc;
    a.c_cur = c_cur; a.c_old = c_old; a.c_dt = c_dt;
    a.flags = flags; a.ph = phys_of(*parm);
    hipEvent_t e0, e1;
    OCFD_HIP(hipEventCreate(&e0));
    OCFD_HIP(hipEventCreate(&e1));
    launch_sweep_one(lev->dev, a, 4, s);   // primitives of S_cur for the sweeps
    for (int which = 0; which < 5; ++which) {
        OCFD_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) {
            if (which < 3 || which == 4) launch_sweep_one(lev->dev, a, which, s);
            else hipLaunchKernelGGL(k_fill, dim3(lev->dev.ntile_g), dim3(256), 0, s, lev->dev.d_boxes,
                                    lev->dev.d_geom, lev->dev.d_tile_g, (double*)S_cur, 0.0,
                                    (const ocfd_step_ctl*)nullptr, 0.0);
        }
        OCFD_HIP(hipEventRecord(e1, s));
        OCFD_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        OCFD_HIP(hipEventElapsedTime(&ms, e0, e1));
        ms_out[which] = ms / reps;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return check_launch("ocfd_profile_stage");
}

int ocfd_level_set_march_chunk(ocfd_level lev, int ch)
{
    if (!lev || ch < 1 || ch > 2047) {
        set_error("ocfd_level_set_march_chunk: invalid arguments");
        return OCFD_EINVAL;
    }
    LevelDev& L = lev->dev;
    if (L.live_graphs > 0) {
        // a captured graph holds the current lists' addresses and grid sizes
        set_error("ocfd_level_set_march_chunk: step graphs of this level exist; destroy them before changing the "
                  "tile shape");
        return OCFD_EINVAL;
    }
    std::vector<int4> ty, tz;
    int ny3 = 0;
    march_tiles(L, ch, ty, tz, ny3);
    int4 *dy = nullptr, *dz = nullptr;
    int rc;
    if ((rc = upload(&dy, ty)) || (rc = upload(&dz, tz))) {
        dev_free({dy, dz});
        return rc;
    }
    OCFD_HIP(hipDeviceSynchronize());   // no launch may still read the old lists
    if (L.own_tile_yz)   // else the old lists live in the level's table allocation
        dev_free({L.d_tile_y, L.d_tile_z});
    L.d_tile_y = dy; L.ntile_y = ny3; L.ntile_yp = (int)ty.size() - ny3;
    L.d_tile_z = dz; L.ntile_z = (int)tz.size();
    L.own_tile_yz = true;
    L.march_ch = ch;
    return OCFD_OK;
}

int ocfd_level_march_chunk(ocfd_level lev) { return lev ? lev->dev.march_ch : -1; }

int ocfd_level_autotune(ocfd_level lev, const double* S, const ocfd_parm* parm, int flags, int reps, int* ch_out,
                        double* ms_out, void* stream)
{
    if (!lev || !S || !parm || !ch_out || reps <= 0 || !valid_flags(lev->dev, flags)) {
        set_error("ocfd_level_autotune: invalid arguments");
        return OCFD_EINVAL;
    }
    static const int cand[6] = {2, 4, 8, 16, 32, 64};
    hipStream_t s = (hipStream_t)stream;
    LevelDev& L = lev->dev;
    if (!L.ntile_y || !L.ntile_p) {   // no regular boxes: nothing to tune
        *ch_out = L.march_ch;
        return OCFD_OK;
    }
    if (L.live_graphs > 0) {
        set_error("ocfd_level_autotune: step graphs of this level exist; destroy them before tuning");
        return OCFD_EINVAL;
    }
    StageArgs a{};
    a.scur = S; a.acc = L.d_acc;   // plain dSdt into level scratch: timing only
    a.c_cur = 1.0; a.flags = flags & ~OCFD_PLANAR; a.ph = phys_of(*parm);
    const bool planar = (flags & OCFD_PLANAR) != 0;
    // a z-uniform level with a one-layer fill runs the packed first-layer y
    // march, which is what is timed for it
    const bool zrep = planar && L.fill_gz == 1;
    a.zrep = zrep;
    hipEvent_t e0, e1;
    OCFD_HIP(hipEventCreate(&e0));
    OCFD_HIP(hipEventCreate(&e1));
    launch_sweep_one(L, a, 4, s);   // primitives the sweeps read
    int best = L.march_ch;
    double best_ms = 1e300;
    int rc = OCFD_OK;
    const int ymarch = zrep ? 5 : 1;
    for (int c = 0; c < 6 && rc == OCFD_OK; ++c) {
        if ((rc = ocfd_level_set_march_chunk(lev, cand[c]))) break;
        launch_sweep_one(L, a, ymarch, s);   // warm-up of this tile list
        if (hipEventRecord(e0, s) != hipSuccess) { rc = OCFD_EHIP; break; }
        for (int r = 0; r < reps; ++r) {
            launch_sweep_one(L, a, ymarch, s);
            if (!planar) launch_sweep_one(L, a, 2, s);
        }
        float ms = 0.f;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = OCFD_EHIP;
            break;
        }
        const double t = ms / reps;
        if (ms_out) ms_out[c] = t;
        if (t < best_ms) { best_ms = t; best = cand[c]; }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) {
        set_error("ocfd_level_autotune: timing failed");
        return rc;
    }
    if ((rc = ocfd_level_set_march_chunk(lev, best))) return rc;
    *ch_out = best;
    return check_launch("ocfd_level_autotune");
}

// ---------------------------------------------------------------------------
// device-driven time steps and their graphs
// ---------------------------------------------------------------------------
} // extern "C"

namespace ocfd {

void launch_fill_ctl(const LevelDev& L, double* S, const ocfd_step_ctl* ctl, double toff, hipStream_t s, double time)
{
    if (const auto* gl = ghost_list(L, 0, s)) {   // a level with coarse/fine boundaries: its own ghost cells only
        if (gl->cap)
            hipLaunchKernelGGL(k_fill_list, dim3((gl->cap + 255) / 256), dim3(256), 0, s, L.d_boxes, L.d_geom, gl->d,
                               gl->cnt, L.fill_gz, S, time, ctl, toff);
        return;
    }
    int ng;
    const int4* tg = fill_tiles(L, ng);
    if (ng) hipLaunchKernelGGL(k_fill, dim3(ng), dim3(256), 0, s, L.d_boxes, L.d_geom, tg, S, time, ctl, toff);
}

void launch_ctl_start(const LevelDev& L, const double* A, const PhysArgs& ph, ocfd_step_ctl* ctl, double time,
                      double cfl, double stop_time, int max_steps, double fixed_est, hipStream_t s)
{
    hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(1), 0, s, ctl, time, cfl, stop_time, max_steps);
    hipLaunchKernelGGL(k_estdt, dim3(L.ntile_c), dim3(256), 0, s, L.d_boxes, L.d_tile_c, A, ph, &ctl->red, &ctl->bad,
                       (const ocfd_step_ctl*)nullptr);
    if (fixed_est > 0.0) hipLaunchKernelGGL(k_ctl_fixed_est, dim3(1), dim3(1), 0, s, ctl, fixed_est);
    hipLaunchKernelGGL(k_ctl_first, dim3(1), dim3(1), 0, s, ctl);
}

void launch_step_end(const LevelDev& L, const double* A, const PhysArgs& ph, ocfd_step_ctl* ctl, double fixed_est,
                     hipStream_t s)
{
    hipLaunchKernelGGL(k_estdt, dim3(L.ntile_c), dim3(256), 0, s, L.d_boxes, L.d_tile_c, A, ph, &ctl->red, &ctl->bad,
                       (const ocfd_step_ctl*)ctl);
    if (fixed_est > 0.0) hipLaunchKernelGGL(k_ctl_fixed_est, dim3(1), dim3(1), 0, s, ctl, fixed_est);
    hipLaunchKernelGGL(k_ctl_update, dim3(1), dim3(1), 0, s, ctl);
}

} // namespace ocfd

extern "C" {

int ocfd_ctl_start(ocfd_level lev, const double* A, ocfd_step_ctl* ctl, double time, double cfl, double stop_time,
                   int max_steps, const ocfd_parm* parm, void* stream)
{
    if (!lev || !A || !ctl || !parm || !(cfl > 0.0)) { set_error("ocfd_ctl_start: invalid arguments"); return OCFD_EINVAL; }
    launch_ctl_start(lev->dev, A, phys_of(*parm), ctl, time, cfl, stop_time, max_steps, 0.0, (hipStream_t)stream);
    return check_launch("ocfd_ctl_start");
}

static int step_device(ocfd_level lev, double* A, double* B, double* C, ocfd_step_ctl* ctl, int scheme,
                       const ocfd_parm* parm, int flags, hipStream_t s)
{
    const LevelDev& L = lev->dev;
    StageArgs a{};
    a.flags = flags; a.ph = phys_of(*parm); a.ctl = ctl; a.acc = L.d_acc;
    auto fill = [&](double* S, double toff) { launch_fill_ctl(L, S, ctl, toff, s); };
    // the CFL reduction of the new state rides on the stage that finishes
    // the step when every box is regular (the finishing sweep sees each new
    // cell once); levels with cut or covered boxes reduce in a pass of their own
    const bool fold = L.n_cut == 0 && L.n_covered == 0 && L.ntile_c > 0;
    auto stage = [&](const double* cur, const double* old, double* nw, double cc, double co, double cdt,
                     bool last = false) {
        a.scur = cur; a.sold = old; a.snew = nw; a.c_cur = cc; a.c_old = co; a.c_dt = cdt;
        a.red = last && fold ? &ctl->red : nullptr;
        a.bad = last && fold ? &ctl->bad : nullptr;
        launch_sweeps(L, a, s);
    };
    // the same stage sequence as ocfd_advance, with dt and time read from ctl
    fill(A, 0.0);
    if (scheme == 0) {
        stage(A, nullptr, B, 1.0, 0.0, 1.0, true);   // B is copied to A whole
        const long n = L.mf_state;
        const int nb = (int)std::min<long>((n + 255) / 256, 4096);
        hipLaunchKernelGGL(k_copy_ctl, dim3(nb), dim3(256), 0, s, A, B, n, (const ocfd_step_ctl*)ctl);
    } else if (scheme == 1) {
        stage(A, nullptr, B, 1.0, 0.0, 1.0);
        fill(B, 1.0);
        stage(B, A, A, 0.5, 0.5, 0.5, true);
    } else {
        stage(A, nullptr, B, 1.0, 0.0, 1.0);
        fill(B, 1.0);
        stage(B, A, C, 0.25, 0.75, 0.25);
        fill(C, 0.5);
        stage(C, A, A, 2.0 / 3.0, 1.0 / 3.0, 2.0 / 3.0, true);
    }
    if (!fold)
        launch_step_end(L, A, a.ph, ctl, 0.0, s);
    else
        hipLaunchKernelGGL(k_ctl_update, dim3(1), dim3(1), 0, s, ctl);
    return check_launch("ocfd_step_device");
}

static bool step_args_ok(ocfd_level lev, double* A, double* B, double* C, ocfd_step_ctl* ctl, int scheme,
                         const ocfd_parm* parm, int flags)
{
    return lev && A && B && ctl && parm && valid_flags(lev->dev, flags) && scheme >= 0 && scheme <= 2 && (scheme != 2 || C) &&
           A != B && (scheme != 2 || (C != A && C != B));
}

int ocfd_step_device(ocfd_level lev, double* A, double* B, double* C, ocfd_step_ctl* ctl, int scheme,
                     const ocfd_parm* parm, int flags, void* stream)
{
    if (!step_args_ok(lev, A, B, C, ctl, scheme, parm, flags)) {
        set_error("ocfd_step_device: invalid arguments");
        return OCFD_EINVAL;
    }
    return step_device(lev, A, B, C, ctl, scheme, parm, flags, (hipStream_t)stream);
}

} // extern "C"

struct ocfd_graph_s {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ocfd_level lev = nullptr;   // its work-lists and launch sizes are baked into the graph
};

namespace ocfd {

int capture_graph(ocfd_level lev, hipStream_t s, int (*body)(void*, hipStream_t), void* arg, ocfd_graph* out,
                  const char* what)
{
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        set_error(std::string(what) + ": hipStreamBeginCapture: " + hipGetErrorString(e));
        return OCFD_EHIP;
    }
    int rc = body(arg, s);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s, &g);
    if (rc != OCFD_OK || e != hipSuccess) {
        if (g) hipGraphDestroy(g);
        if (rc == OCFD_OK) set_error(std::string(what) + ": hipStreamEndCapture: " + hipGetErrorString(e));
        return rc != OCFD_OK ? rc : OCFD_EHIP;
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        hipGraphDestroy(g);
        set_error(std::string(what) + ": hipGraphInstantiate: " + hipGetErrorString(e));
        return OCFD_EHIP;
    }
    ocfd_graph G = new ocfd_graph_s;
    G->graph = g;
    G->exec = x;
    G->lev = lev;
    ++lev->dev.live_graphs;
    *out = G;
    return OCFD_OK;
}

} // namespace ocfd

namespace {
struct StepsArgs {
    ocfd_level lev;
    double *A, *B, *C;
    ocfd_step_ctl* ctl;
    int scheme, flags, nsteps;
    const ocfd_parm* parm;
};
int steps_body(void* p, hipStream_t s)
{
    const StepsArgs& a = *(const StepsArgs*)p;
    int rc = OCFD_OK;
    for (int n = 0; n < a.nsteps && rc == OCFD_OK; ++n)
        rc = step_device(a.lev, a.A, a.B, a.C, a.ctl, a.scheme, a.parm, a.flags, s);
    return rc;
}
} // namespace

extern "C" {

int ocfd_graph_capture_steps(ocfd_level lev, double* A, double* B, double* C, ocfd_step_ctl* ctl, int scheme,
                             const ocfd_parm* parm, int flags, int nsteps, void* stream, ocfd_graph* out)
{
    if (!step_args_ok(lev, A, B, C, ctl, scheme, parm, flags) || nsteps <= 0 || !out || !stream) {
        set_error("ocfd_graph_capture_steps: invalid arguments (a non-null stream is required)");
        return OCFD_EINVAL;
    }
    StepsArgs a{lev, A, B, C, ctl, scheme, flags, nsteps, parm};
    return capture_graph(lev, (hipStream_t)stream, steps_body, &a, out, "ocfd_graph_capture_steps");
}

int ocfd_graph_launch(ocfd_graph g, void* stream)
{
    if (!g) { set_error("ocfd_graph_launch: null graph"); return OCFD_EINVAL; }
    OCFD_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return OCFD_OK;
}

int ocfd_graph_destroy(ocfd_graph g)
{
    if (!g) return OCFD_OK;
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    if (g->lev) --g->lev->dev.live_graphs;
    delete g;
    return OCFD_OK;
}

} // extern "C"
