// Host-side level object shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include "ocfd_device.hpp"

namespace ocfd {

// A level: box list + geometry, with the device work-lists of every kernel.
struct LevelDev {
    int nbox = 0;
    std::vector<BoxDesc> boxes;     // host copy
    GeomDesc geom{};                // host copy (map pointer is the device map)
    std::vector<int> owner_map;     // host copy of the owner map
    BoxDesc* d_boxes = nullptr;
    GeomDesc* d_geom = nullptr;
    int* d_map = nullptr;
    int4* d_tile_x = nullptr; int ntile_x = 0;     // x sweep: (box, segment, row group, z growth 3)
    int4* d_tile_xp = nullptr; int ntile_xp = 0;   // planar x sweep (z growth 1)
    int4* d_tile_y = nullptr; int ntile_y = 0;
    int4* d_tile_z = nullptr; int ntile_z = 0;
    int4* d_tile_g = nullptr; int ntile_g = 0;   // ghost fill: (box, first ghost cell, count, 0)
    int4* d_tile_c = nullptr; int ntile_c = 0;   // valid-cell chunks: (box, first cell, count, 0)
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

inline PhysArgs phys_of(const ocfd_parm& p)
{
    PhysArgs a;
    a.gamma = p.eos_gamma; a.smallr = p.smallr; a.smallp = p.smallp; a.Rg = p.Rg;
    a.C_s = p.C_s; a.T_s = p.T_s; a.kOverMu = p.kOverMu; a.twoThirds = p.twoThirds;
    return a;
}

void launch_sweeps(const LevelDev& L, const StageArgs& a, hipStream_t s);
void launch_sweep_one(const LevelDev& L, const StageArgs& a, int which, hipStream_t s);
int tile_xseg();
int tile_chunk();

void set_error(const std::string& msg);

} // namespace ocfd

struct ocfd_level_s {
    ocfd::LevelDev dev;
};
