// Host-side level object shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstring>
#include <initializer_list>
#include <string>
#include <vector>
#include "ocfd_device.hpp"

namespace ocfd {

void set_error(const std::string& msg);

// A level: box list + geometry, with the device work-lists of every kernel.
struct LevelDev {
    int nbox = 0;
    std::vector<BoxDesc> boxes;     // host copy
    GeomDesc geom{};                // host copy (map pointer is the device map)
    std::vector<int> owner_map;     // host copy of the owner map (tiles of 8^3 cells)
    std::vector<int> owner_list;    // host copy of its box lists
    void* d_arena = nullptr;        // one allocation holding the read-only tables below
    void* d_bufs = nullptr;         // one allocation holding the stage scratch (d_acc, d_prim32, d_prim64, d_red)
    bool own_tile_yz = false;       // d_tile_y / d_tile_z allocated on their own (after a chunk change)
    BoxDesc* d_boxes = nullptr;
    GeomDesc* d_geom = nullptr;
    int* d_map = nullptr;
    int* d_map_list = nullptr;
    int4* d_tile_x = nullptr; int ntile_x = 0;     // x sweep: (box, first window, z row, z growth 3)
    int4* d_tile_xp = nullptr; int ntile_xp = 0;   // planar x sweep (z growth 1)
    int ntile_xz = 0;                               // its first ntile_xz tiles: the first z layer and its neighbours
    int4* d_tile_y = nullptr; int ntile_y = 0;   // 3D y march; then ntile_yp planar first-layer tiles
    int ntile_yp = 0;
    int4* d_tile_z = nullptr; int ntile_z = 0;
    int4* d_tile_g = nullptr; int ntile_g = 0;   // ghost fill: (box, first ghost cell, count, z depth NG)
    int4* d_tile_gp = nullptr; int ntile_gp = 0; // planar ghost fill: z depth 1
    int fill_gz = NG;               // z depth of the level's ghost fills (1: planar levels only)
    // ghost lists of a level that does not cover its domain (one with
    // coarse/fine boundaries, gl_need), per fill depth (0: NG, 1: planar):
    // [0] the cells the level fill writes, [1] the cells the coarse/fine fill
    // may write (k_ghost_lists), each sized for every ghost cell of that depth
    // (cap) with its length in device memory (cnt).  A depth's lists are built
    // on the stream of the first fill at that depth (ghost_list).
    struct GhostList { int2* d = nullptr; const unsigned* cnt = nullptr; int cap = 0; };
    bool gl_need = false;
    mutable GhostList gl[2][2];
    mutable void* d_gl[2] = {nullptr, nullptr};   // per depth: the two lists and their lengths
    mutable hipStream_t gl_stream[2] = {nullptr, nullptr};   // the stream that built them
    mutable hipEvent_t gl_event[2] = {nullptr, nullptr};     // recorded there after the build
    int4* d_tile_c = nullptr; int ntile_c = 0;   // valid-cell chunks: (box, first cell, count, 0)
    int4* d_tile_c1 = nullptr; int ntile_c1 = 0; // valid-cell chunks of 256, one cell per thread (per-cell kernels)
    int4* d_tile_v = nullptr; int ntile_v = 0;   // valid-cell chunks of 64 (chemistry)
    int4* d_tile_p = nullptr; int ntile_p = 0;   // primitive region grow(bx,3): (box, first, count, 3)
    int4* d_tile_pp = nullptr; int ntile_pp = 0; // planar: z grown by 1 only
    float* d_prim32 = nullptr;      // primitives, FP32, 6 comps, grown layout
    double* d_prim64 = nullptr;     // primitives, FP64
    long mf_prim = 0;               // elements of a primitive MultiFab
    double* d_acc = nullptr;        // dSdt accumulator, valid layout, 5 comps
    unsigned long long* d_red = nullptr;  // reduction result slot
    long mf_valid = 0;              // doubles in a 0-ghost 5-comp MultiFab
    long mf_state = 0;              // doubles in a 4-ghost 5-comp MultiFab
    long dev_bytes = 0;             // device memory held by the level
    int march_ch = 32;              // cells per march chunk of the y and z sweeps
    // The level's state combined in time and ghost-filled at a stage time of
    // a finer level (one entry per stage time, reused by the finer stages at
    // the same time; ocfd_amr.hip).  Valid while the level's old and current
    // states are unchanged: cleared when the level advances, is refluxed or
    // averaged down, or gets new data.
    struct TimeState {
        double ts = 0.0, t_old = 0.0, t_new = 0.0;
        const double* old = nullptr;
        const double* cur = nullptr;
        double* buf = nullptr;
        bool valid = false;
        unsigned long use = 0;
    };
    std::vector<TimeState> tcache;
    unsigned long tcache_clock = 0;
    int live_graphs = 0;            // instantiated step graphs holding this level's work-lists and sizes
    // embedded boundary (cut boxes only, grown layout, box order)
    bool has_eb = false;
    ocfd_eb_spec eb_spec{};
    int n_cut = 0, n_covered = 0;
    long mf_eb = 0;                 // grown cells over all cut boxes
    uint32_t* d_ebflag = nullptr;   // EBCellFlag words
    double* d_ebgeo = nullptr;      // vfrac, apx, apy, apz
    double* d_ebq = nullptr;        // primitives, FP64, 7 comps (rho,u,v,w,p,T,c)
    double* d_ebF[3] = {nullptr, nullptr, nullptr};   // face fluxes per direction, 5 comps
    int4* d_tile_eq = nullptr; int ntile_eq = 0;      // cut boxes: grow(bx,2) cells
    int4* d_tile_ef[3] = {nullptr, nullptr, nullptr}; int ntile_ef[3] = {0, 0, 0};   // faces
    int4* d_tile_ed = nullptr; int ntile_ed = 0;      // cut and covered boxes: valid cells
    // multispecies flow (regular boxes)
    int4* d_tile_mf[3] = {nullptr, nullptr, nullptr}; int ntile_mf[3] = {0, 0, 0};   // faces
    int4* d_tile_md = nullptr; int ntile_md = 0;      // valid cells
    double* d_mq = nullptr;         // 19 comps, grown layout (allocated on first use)
    double* d_mF[3] = {nullptr, nullptr, nullptr};    // 14-comp face fluxes per direction
};

void launch_eb(const LevelDev& L, const StageArgs& a, hipStream_t s);

// drops the level's time-combined states (their buffers are kept for reuse)
inline void tcache_clear(LevelDev& L)
{
    for (auto& e : L.tcache) e.valid = false;
}

// chemistry of a level (ocfd_chem.hip): the linearised-implicit step of dt
// (or, under device time control, of (0.5 ctl->dt)/iter) and the species
// renormalisation
void launch_chem_level(const LevelDev& L, double* S, double* Spec, double dt, double smallr,
                       const ocfd_step_ctl* ctl, int iter, hipStream_t s);
void launch_renorm(const LevelDev& L, const double* S, double* Spec, const ocfd_step_ctl* ctl, hipStream_t s);

// device time control (ocfd_level.hip): the ghost fill at ctl->time +
// toff ctl->dt; the end of a step (non-finite check of A, the CFL estimate --
// or fixed_est when > 0 -- and the dt update); the start of a run
void launch_fill_ctl(const LevelDev& L, double* S, const ocfd_step_ctl* ctl, double toff, hipStream_t s,
                     double time = 0.0);
void launch_step_end(const LevelDev& L, const double* A, const PhysArgs& ph, ocfd_step_ctl* ctl, double fixed_est,
                     hipStream_t s);
void launch_ctl_start(const LevelDev& L, const double* A, const PhysArgs& ph, ocfd_step_ctl* ctl, double time,
                      double cfl, double stop_time, int max_steps, double fixed_est, hipStream_t s);
// hipGraph of the launches `body` enqueues on stream s (a level's graph:
// while it exists the level keeps its work-lists)
int capture_graph(ocfd_level lev, hipStream_t s, int (*body)(void*, hipStream_t), void* arg, ocfd_graph* out,
                  const char* what);

constexpr int kMapShift = 3;   // owner-map tiles of 8^3 cells

// ghost-fill work-list of a level at its fill depth
inline const int4* fill_tiles(const LevelDev& L, int& n)
{
    n = L.fill_gz == 1 ? L.ntile_gp : L.ntile_g;
    return L.fill_gz == 1 ? L.d_tile_gp : L.d_tile_g;
}

// Builds the ghost lists of depth d (0: NG, 1: planar) on stream s; false
// when they cannot be had (no device memory), the fills then use the tiles.
bool build_ghost_lists(const LevelDev& L, int d, hipStream_t s);

// the level's ghost list `which` (0 level fill, 1 coarse/fine fill) at its
// fill depth, or null when the level keeps none (fill by the tile list).  The
// lists of a depth are built at its first use, on the stream of that fill
// (ordered before it); a fill on another stream waits for the build's event.
// A fill under stream capture that would need either takes the tile list.
inline const LevelDev::GhostList* ghost_list(const LevelDev& L, int which, hipStream_t s)
{
    if (!L.gl_need) return nullptr;
    const int d = L.fill_gz == 1 ? 1 : 0;
    if (L.d_gl[d] && s == L.gl_stream[d]) return &L.gl[d][which];
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return nullptr;
    if (!L.d_gl[d]) {
        if (!build_ghost_lists(L, d, s)) return nullptr;
    } else if (hipStreamWaitEvent(s, L.gl_event[d], 0) != hipSuccess) {
        return nullptr;
    }
    return &L.gl[d][which];
}

// A non-blocking stream of the current device for copies of freshly allocated
// tables: a copy on the legacy default stream would wait for every kernel
// already queued there (the AMR driver builds levels while the previous ones
// still run).
hipStream_t upload_stream();

// Device blocks for the library's tables and scratch from a per-device
// caching pool (ocfd_pool.hip): dev_alloc false when out of device memory;
// dev_free waits for the device (as hipFree would) and keeps the blocks for
// reuse.  upload_bytes copies host bytes to the device through a pinned
// staging buffer on upload_stream and waits for that copy only.
bool dev_alloc(void** out, size_t bytes);
void dev_free(std::initializer_list<void*> ps);
int upload_bytes(void* dst, const void* src, size_t n);

// Host tables gathered into one device allocation filled by one copy; each
// table starts on a 256-byte boundary.  allocate() sets every registered
// pointer into the allocation (null for empty tables); reserved space can be
// written (put) between allocate() and copy(), for tables that hold device
// pointers into the allocation itself.
struct Staging {
    std::vector<char> host;
    std::vector<std::pair<void**, size_t>> slots;
    size_t place(size_t bytes)
    {
        const size_t off = (host.size() + 255) & ~(size_t)255;
        host.resize(off + bytes);
        return off;
    }
    template <class T>
    void add(T** slot, const std::vector<T>& v)
    {
        *slot = nullptr;
        if (v.empty()) return;
        const size_t off = place(sizeof(T) * v.size());
        std::memcpy(host.data() + off, v.data(), sizeof(T) * v.size());
        slots.push_back({(void**)slot, off});
    }
    template <class T>
    size_t reserve(T** slot, size_t count)
    {
        *slot = nullptr;
        const size_t off = place(sizeof(T) * count);
        slots.push_back({(void**)slot, off});
        return off;
    }
    template <class T>
    void put(size_t off, const T& v) { std::memcpy(host.data() + off, &v, sizeof(T)); }
    int allocate(void** base)
    {
        *base = nullptr;
        if (host.empty()) return OCFD_OK;
        if (!dev_alloc(base, host.size())) {
            *base = nullptr;
            set_error("out of device memory for the level tables");
            return OCFD_ENOMEM;
        }
        for (auto& sl : slots) *sl.first = (char*)*base + sl.second;
        return OCFD_OK;
    }
    int copy(void* base)
    {
        if (host.empty()) return OCFD_OK;
        return upload_bytes(base, host.data(), host.size());
    }
    int commit(void** base)
    {
        int rc = allocate(base);
        return rc ? rc : copy(*base);
    }
};

// Tile map of a box list over the domain [dlo, dhi] (tiles of 2^shift cells,
// the lookup of tile_owner): a tile that one box (grown by `grow`, clipped to
// the domain) wholly covers and no other box meets maps to that box, an empty
// tile to -1, any other to a list of the boxes meeting it in ascending order.
// Returns false when two boxes overlap (checked for grow 0).
inline bool build_tile_map(const int* dlo, const int* dhi, int shift, const std::vector<BoxDesc>& boxes, int grow,
                           int* mn, std::vector<int>& map, std::vector<int>& list)
{
    for (int d = 0; d < 3; ++d) mn[d] = ((dhi[d] - dlo[d] + 1) + (1 << shift) - 1) >> shift;
    const size_t nt = (size_t)mn[0] * mn[1] * mn[2];
    auto range = [&](const BoxDesc& B, int* t0, int* t1, int* blo, int* bhi) {
        for (int d = 0; d < 3; ++d) {
            blo[d] = std::max(B.lo[d] - grow, dlo[d]);
            bhi[d] = std::min(B.hi[d] + grow, dhi[d]);
            t0[d] = (blo[d] - dlo[d]) >> shift;
            t1[d] = (bhi[d] - dlo[d]) >> shift;
        }
    };
    std::vector<int> cnt(nt + 1, 0);
    for (const BoxDesc& B : boxes) {
        int t0[3], t1[3], blo[3], bhi[3];
        range(B, t0, t1, blo, bhi);
        for (int k = t0[2]; k <= t1[2]; ++k)
            for (int j = t0[1]; j <= t1[1]; ++j)
                for (int i = t0[0]; i <= t1[0]; ++i) ++cnt[((size_t)k * mn[1] + j) * mn[0] + i];
    }
    std::vector<int> off(nt + 1, 0);
    for (size_t t = 0; t < nt; ++t) off[t + 1] = off[t] + cnt[t];
    std::vector<int> ent(off[nt]);
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int b = 0; b < (int)boxes.size(); ++b) {
        int t0[3], t1[3], blo[3], bhi[3];
        range(boxes[b], t0, t1, blo, bhi);
        for (int k = t0[2]; k <= t1[2]; ++k)
            for (int j = t0[1]; j <= t1[1]; ++j)
                for (int i = t0[0]; i <= t1[0]; ++i) ent[fill[((size_t)k * mn[1] + j) * mn[0] + i]++] = b;
    }
    map.assign(nt, -1);
    list.clear();
    for (size_t t = 0; t < nt; ++t) {
        const int c = cnt[t];
        if (c == 0) continue;
        const int ti[3] = {(int)(t % mn[0]), (int)((t / mn[0]) % mn[1]), (int)(t / ((size_t)mn[0] * mn[1]))};
        if (c == 1) {
            int t0[3], t1[3], blo[3], bhi[3];
            range(boxes[ent[off[t]]], t0, t1, blo, bhi);
            bool whole = true;
            for (int d = 0; d < 3; ++d) {
                const int tlo = dlo[d] + (ti[d] << shift), thi = std::min(tlo + (1 << shift) - 1, dhi[d]);
                whole = whole && blo[d] <= tlo && bhi[d] >= thi;
            }
            if (whole) { map[t] = ent[off[t]]; continue; }
        } else if (grow == 0) {
            for (int a = off[t]; a < off[t + 1]; ++a)
                for (int b = a + 1; b < off[t + 1]; ++b) {
                    const BoxDesc& A = boxes[ent[a]];
                    const BoxDesc& C = boxes[ent[b]];
                    bool meet = true;
                    for (int d = 0; d < 3; ++d) meet = meet && A.lo[d] <= C.hi[d] && C.lo[d] <= A.hi[d];
                    if (meet) return false;
                }
        }
        map[t] = -(int)list.size() - 2;
        list.push_back(c);
        list.insert(list.end(), ent.begin() + off[t], ent.begin() + off[t + 1]);
    }
    return true;
}

inline PhysArgs phys_of(const ocfd_parm& p)
{
    PhysArgs a;
    a.gamma = p.eos_gamma; a.smallr = p.smallr; a.smallp = p.smallp; a.Rg = p.Rg;
    a.C_s = p.C_s; a.T_s = p.T_s; a.kOverMu = p.kOverMu; a.twoThirds = p.twoThirds;
    a.sw_t1 = 1.0 / (2.0 * a.gamma);
    a.sw_t3 = (3.0 - a.gamma) / (2.0 * (a.gamma - 1.0));
    const float gf = (float)a.gamma;
    a.sw_t1f = 1.0f / (2.0f * gf);
    a.sw_t3f = (3.0f - gf) / (2.0f * (gf - 1.0f));
    a.Rinv = 1.0 / a.Rg;
    return a;
}

void launch_sweeps(const LevelDev& L, const StageArgs& a, hipStream_t s);
void launch_sweep_one(const LevelDev& L, const StageArgs& a, int which, hipStream_t s);
int tile_xseg();
int tile_chunk();

} // namespace ocfd

struct ocfd_level_s {
    ocfd::LevelDev dev;
};
