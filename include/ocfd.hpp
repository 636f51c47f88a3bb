// ocfd.hpp -- C++17 host facade over the C ABI of ocfd.h.
//
// The shapes of the reference's per-level advance interface, on the
// library's own types: Box / BoxArray / Geometry (amrex::Box, BoxArray,
// Geometry), MultiFab with Array4 views (one device allocation per MultiFab,
// FABs back to back in box order, each component-major over its grown box,
// i fastest: the amrex::Array4 convention), FluxRegister (EBR's flux_reg,
// an EBFluxRegister per coarse/fine interface) and EBR, whose members mirror
// EBR::compute_dSdt(S, dSdt, dt, fine, current) (Source/EBR.H:167-169),
// EBR::estTimeStep (Source/EBR.H:122), EBR::flow_advance (Source/Advance.cpp:81-150),
// EBR::chemical_advance (Source/EBR.H:133) and the driver loop of
// Source/main.cpp:61-70 with computeInitialDt / computeNewDt
// (Source/EBR.cpp:189-306).  A C++ host built with HIP (as an AMReX HIP
// build is) includes this header and links libocfd.so; every call goes
// through the plain C entry points, and every failure becomes an
// ocfd::Error carrying the library's message (where the reference calls
// amrex::Abort).
#ifndef OCFD_HPP
#define OCFD_HPP

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "ocfd.h"

namespace ocfd {

using Real = double;

struct Error : std::runtime_error {
    int status;
    Error(int rc, const std::string& what) : std::runtime_error(what), status(rc) {}
};

inline void check(int rc, const char* what)
{
    if (rc != OCFD_OK) throw Error(rc, std::string(what) + ": " + ocfd_last_error());
}

inline void check_hip(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw Error(OCFD_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Cell-centred index box, inclusive corners (amrex::Box).
struct Box {
    std::array<int, 3> lo{0, 0, 0}, hi{-1, -1, -1};
    Box() = default;
    Box(std::array<int, 3> l, std::array<int, 3> h) : lo(l), hi(h) {}
    int smallEnd(int d) const { return lo[d]; }
    int bigEnd(int d) const { return hi[d]; }
    int length(int d) const { return hi[d] - lo[d] + 1; }
    bool ok() const { return length(0) > 0 && length(1) > 0 && length(2) > 0; }
    long numPts() const { return ok() ? (long)length(0) * length(1) * length(2) : 0; }
    bool contains(int i, int j, int k) const
    {
        return i >= lo[0] && i <= hi[0] && j >= lo[1] && j <= hi[1] && k >= lo[2] && k <= hi[2];
    }
    Box grow(int n) const { return Box({lo[0] - n, lo[1] - n, lo[2] - n}, {hi[0] + n, hi[1] + n, hi[2] + n}); }
    Box refine(int r) const
    {
        return Box({lo[0] * r, lo[1] * r, lo[2] * r}, {hi[0] * r + r - 1, hi[1] * r + r - 1, hi[2] * r + r - 1});
    }
    Box coarsen(int r) const
    {
        auto fd = [r](int v) { return v >= 0 ? v / r : -((-v + r - 1) / r); };
        return Box({fd(lo[0]), fd(lo[1]), fd(lo[2])}, {fd(hi[0]), fd(hi[1]), fd(hi[2])});
    }
    bool operator==(const Box& o) const { return lo == o.lo && hi == o.hi; }
    bool operator!=(const Box& o) const { return !(*this == o); }
};

// List of disjoint boxes of one level (amrex::BoxArray).
struct BoxArray {
    std::vector<Box> boxes;
    BoxArray() = default;
    explicit BoxArray(std::vector<Box> b) : boxes(std::move(b)) {}
    // the domain [0, n) cut into boxes of at most max_grid cells per side, as
    // equal as possible (amr.max_grid_size)
    static BoxArray chop(std::array<int, 3> n, std::array<int, 3> max_grid)
    {
        std::array<std::vector<int>, 3> cuts;
        for (int d = 0; d < 3; ++d) {
            const int m = std::max(1, (n[d] + max_grid[d] - 1) / max_grid[d]);
            for (int q = 0; q <= m; ++q) cuts[d].push_back((int)((long)n[d] * q / m));
        }
        BoxArray ba;
        for (size_t k = 0; k + 1 < cuts[2].size(); ++k)
            for (size_t j = 0; j + 1 < cuts[1].size(); ++j)
                for (size_t i = 0; i + 1 < cuts[0].size(); ++i)
                    ba.boxes.push_back(Box({cuts[0][i], cuts[1][j], cuts[2][k]},
                                           {cuts[0][i + 1] - 1, cuts[1][j + 1] - 1, cuts[2][k + 1] - 1}));
        return ba;
    }
    int size() const { return (int)boxes.size(); }
    const Box& operator[](int b) const { return boxes[b]; }
    long numPts() const
    {
        long n = 0;
        for (const Box& b : boxes) n += b.numPts();
        return n;
    }
};

// Physical BC codes of ebr.lo_bc / ebr.hi_bc (Exec/<case>/inputs).
enum PhysBC { Interior = 0, Inflow = 1, Outflow = 2, Symmetry = 3, SlipWall = 4, NoSlipWall = 5 };

// Problem domain, cell size, periodicity and physical BCs of one level
// (amrex::Geometry + the BCRec set-up of Source/Setup.cpp:17-92).
struct Geometry {
    ocfd_geom g{};
    Geometry() = default;
    Geometry(std::array<int, 3> n_cell, std::array<Real, 3> prob_lo, std::array<Real, 3> prob_hi,
             std::array<int, 3> periodic, std::array<int, 3> lo_bc, std::array<int, 3> hi_bc, int probtype = 0)
    {
        for (int d = 0; d < 3; ++d) {
            g.dom_lo[d] = 0;
            g.dom_hi[d] = n_cell[d] - 1;
            g.prob_lo[d] = prob_lo[d];
            g.dx[d] = (prob_hi[d] - prob_lo[d]) / n_cell[d];
            g.periodic[d] = periodic[d];
            g.lo_bc[d] = lo_bc[d];
            g.hi_bc[d] = hi_bc[d];
        }
        g.probtype = probtype;
    }
    Box Domain() const { return Box({g.dom_lo[0], g.dom_lo[1], g.dom_lo[2]}, {g.dom_hi[0], g.dom_hi[1], g.dom_hi[2]}); }
    Real CellSize(int d) const { return g.dx[d]; }
    Real ProbLo(int d) const { return g.prob_lo[d]; }
    bool isPeriodic(int d) const { return g.periodic[d] != 0; }
    // the same physical box refined by 2 (a finer AMR level)
    Geometry refine2() const
    {
        Geometry f = *this;
        for (int d = 0; d < 3; ++d) {
            f.g.dom_lo[d] = 2 * g.dom_lo[d];
            f.g.dom_hi[d] = 2 * g.dom_hi[d] + 1;
            f.g.dx[d] = g.dx[d] / 2;
        }
        return f;
    }
};

// struct Parm (Source/Constants.H:7-64, non-CHEM members) with the
// reference's defaults and derived members (Parm::Initialize).
struct Parm {
    ocfd_parm p{};
    Parm() { check(ocfd_parm_initialize(&p), "ocfd_parm_initialize"); }
    // recompute Rg, cv, cp, kOverMu after changing eos_gamma, eos_m, Pr ...
    void initialize() { check(ocfd_parm_initialize(&p), "ocfd_parm_initialize"); }
};

// View of one FAB (amrex::Array4): data over the box grown by ngrow, component
// major, i fastest.  The pointer is a device pointer for MultiFab data.
template <class T>
struct Array4 {
    T* p = nullptr;
    std::array<int, 3> begin{0, 0, 0};   // lower corner of the grown box
    long jstride = 0, kstride = 0, nstride = 0;
    int ncomp = 0;
    T& operator()(int i, int j, int k, int n = 0) const
    {
        return p[(i - begin[0]) + (j - begin[1]) * jstride + (k - begin[2]) * kstride + n * nstride];
    }
};

// One level's box list + geometry with its device work-lists (an ocfd_level).
class Level {
public:
    Level(const BoxArray& ba, const Geometry& geom, const ocfd_eb_spec* eb = nullptr) : ba_(ba), geom_(geom)
    {
        std::vector<int> lo, hi;
        for (const Box& b : ba.boxes)
            for (int d = 0; d < 3; ++d) {
                lo.push_back(b.lo[d]);
                hi.push_back(b.hi[d]);
            }
        check(ocfd_level_create_eb(ba.size(), lo.data(), hi.data(), &geom.g, eb, &h_), "ocfd_level_create");
    }
    ~Level() { if (h_) ocfd_level_destroy(h_); }
    Level(const Level&) = delete;
    Level& operator=(const Level&) = delete;
    ocfd_level handle() const { return h_; }
    const BoxArray& boxArray() const { return ba_; }
    const Geometry& Geom() const { return geom_; }
    int nbox() const { return ba_.size(); }
    long mfSize(int ngrow, int ncomp) const { return ocfd_level_mf_size(h_, ngrow, ncomp); }
    long fabOffset(int box, int ngrow, int ncomp) const { return ocfd_level_fab_offset(h_, box, ngrow, ncomp); }
    // z-uniform stepping with a one-layer z ghost fill (OCFD_PLANAR stages)
    void setPlanar(bool on) { check(ocfd_level_set_planar(h_, on ? 1 : 0), "ocfd_level_set_planar"); }
    int marchChunk() const { return ocfd_level_march_chunk(h_); }
    void setMarchChunk(int ch) { check(ocfd_level_set_march_chunk(h_, ch), "ocfd_level_set_march_chunk"); }

private:
    ocfd_level h_ = nullptr;
    BoxArray ba_;
    Geometry geom_;
};

// Device MultiFab in the library's layout: one hipMalloc per MultiFab, the
// FABs of the level's boxes back to back (amrex::MultiFab over one arena
// block).  Move-only.
class MultiFab {
public:
    MultiFab(const Level& lev, int ncomp, int ngrow) : lev_(&lev), ncomp_(ncomp), ngrow_(ngrow)
    {
        n_ = lev.mfSize(ngrow, ncomp);
        if (n_ < 0) throw Error(OCFD_EINVAL, "MultiFab: invalid level");
        if (n_ > 0) check_hip(hipMalloc((void**)&d_, sizeof(Real) * n_), "MultiFab: hipMalloc");
    }
    ~MultiFab() { if (d_) (void)hipFree(d_); }
    MultiFab(MultiFab&& o) noexcept
        : lev_(o.lev_), d_(std::exchange(o.d_, nullptr)), n_(o.n_), ncomp_(o.ncomp_), ngrow_(o.ngrow_) {}
    MultiFab& operator=(MultiFab&& o) noexcept
    {
        std::swap(lev_, o.lev_); std::swap(d_, o.d_); std::swap(n_, o.n_);
        std::swap(ncomp_, o.ncomp_); std::swap(ngrow_, o.ngrow_);
        return *this;
    }
    MultiFab(const MultiFab&) = delete;
    MultiFab& operator=(const MultiFab&) = delete;

    Real* dataPtr() { return d_; }
    const Real* dataPtr() const { return d_; }
    int nComp() const { return ncomp_; }
    int nGrow() const { return ngrow_; }
    long size() const { return n_; }
    const Level& level() const { return *lev_; }
    Real* fabPtr(int b) { return d_ + lev_->fabOffset(b, ngrow_, ncomp_); }
    Array4<Real> array(int b)
    {
        const Box g = lev_->boxArray()[b].grow(ngrow_);
        Array4<Real> a;
        a.p = fabPtr(b);
        a.begin = g.lo;
        a.jstride = g.length(0);
        a.kstride = (long)g.length(0) * g.length(1);
        a.nstride = g.numPts();
        a.ncomp = ncomp_;
        return a;
    }
    void setVal(int byte, hipStream_t s = nullptr)
    {
        if (n_) check_hip(hipMemsetAsync(d_, byte, sizeof(Real) * n_, s), "MultiFab::setVal");
    }
    // whole-MultiFab copies between host and device (synchronous)
    void copyToHost(std::vector<Real>& h) const
    {
        h.resize(n_);
        if (n_) check_hip(hipMemcpy(h.data(), d_, sizeof(Real) * n_, hipMemcpyDeviceToHost), "MultiFab::copyToHost");
    }
    void copyFromHost(const std::vector<Real>& h)
    {
        if ((long)h.size() != n_) throw Error(OCFD_EINVAL, "MultiFab::copyFromHost: size mismatch");
        if (n_) check_hip(hipMemcpy(d_, h.data(), sizeof(Real) * n_, hipMemcpyHostToDevice), "MultiFab::copyFromHost");
    }

private:
    const Level* lev_;
    Real* d_ = nullptr;
    long n_ = 0;
    int ncomp_, ngrow_;
};

// The flux register of one coarse/fine interface (EBR's flux_reg, an
// EBFluxRegister: CrseAdd/FineAdd are fed by EBR::compute_dSdt's register
// arguments, Reflux and reset as in Source/EBR.cpp:322-330 and
// Source/Advance.cpp:98-103).
class FluxRegister {
public:
    FluxRegister(const Level& coarse, const Level& fine)
    {
        check(ocfd_cf_create(coarse.handle(), fine.handle(), &h_), "ocfd_cf_create");
    }
    ~FluxRegister() { if (h_) ocfd_cf_destroy(h_); }
    FluxRegister(const FluxRegister&) = delete;
    FluxRegister& operator=(const FluxRegister&) = delete;
    ocfd_cf handle() const { return h_; }
    int nfaces() const { return ocfd_cf_nfaces(h_); }
    void reset(hipStream_t s = nullptr) { check(ocfd_cf_reset(h_, s), "ocfd_cf_reset"); }
    void Reflux(MultiFab& S_coarse, hipStream_t s = nullptr)
    {
        check(ocfd_cf_reflux(h_, S_coarse.dataPtr(), s), "ocfd_cf_reflux");
    }
    // EBR::avgDown (Source/EBR.cpp:535-559)
    void avgDown(const MultiFab& S_fine, MultiFab& S_coarse, hipStream_t s = nullptr)
    {
        check(ocfd_cf_avg_down(h_, S_fine.dataPtr(), S_coarse.dataPtr(), S_fine.nComp(), s), "ocfd_cf_avg_down");
    }
    // fine ghosts no fine box owns, from the coarse state (FillPatchTwoLevels)
    void fillFine(MultiFab& S_fine, const MultiFab& S_coarse, hipStream_t s = nullptr)
    {
        check(ocfd_cf_fill_fine(h_, S_fine.dataPtr(), S_coarse.dataPtr(), s), "ocfd_cf_fill_fine");
    }

private:
    ocfd_cf h_ = nullptr;
};

// The per-level advance of the reference's EBR class on one Level: the
// regular/EB box dispatch of compute_dSdt, the CFL estimate, the RK driver
// and the chemistry source, on the caller's stream.
class EBR {
public:
    EBR(Level& lev, const Parm& parm, int flags = 0, hipStream_t s = nullptr)
        : lev_(lev), parm_(parm), flags_(flags), s_(s) {}

    Level& level() { return lev_; }
    int flags() const { return flags_; }
    hipStream_t stream() const { return s_; }

    // a state MultiFab of this level: 5 comps, 4 ghosts (NUM_GROW)
    MultiFab newState() const { return MultiFab(lev_, OCFD_NCONS, OCFD_NGROW); }

    // case initial data (amrex_probinit + EBR::initData, Source/EBR.cpp:111-170)
    void initData(int kind, MultiFab& S, unsigned long long seed = 0)
    {
        check(ocfd_init_state(lev_.handle(), kind, S.dataPtr(), &parm_.p, seed), "ocfd_init_state");
    }

    // FillPatch of one level at `time` (Source/Advance.cpp:119,126,134)
    void FillPatch(MultiFab& S, Real time)
    {
        check(ocfd_fill_boundary(lev_.handle(), S.dataPtr(), time, s_), "ocfd_fill_boundary");
    }

    // EBR::compute_dSdt (Source/EBR.H:167-169, Source/Advance.cpp:182-418):
    // dSdt = -div F(S) for every box; with registers, dt times this level's
    // interface fluxes goes to `fine` (CrseAdd) and `current` (FineAdd)
    void compute_dSdt(const MultiFab& S, MultiFab& dSdt, Real dt, FluxRegister* fine = nullptr,
                      FluxRegister* current = nullptr)
    {
        if (fine || current)
            check(ocfd_compute_dsdt_fr(lev_.handle(), S.dataPtr(), dSdt.dataPtr(), dt, fine ? fine->handle() : nullptr,
                                       current ? current->handle() : nullptr, &parm_.p, flags_, s_),
                  "ocfd_compute_dsdt_fr");
        else
            check(ocfd_compute_dsdt(lev_.handle(), S.dataPtr(), dSdt.dataPtr(), dt, &parm_.p, flags_, s_),
                  "ocfd_compute_dsdt");
    }

    // EBR::estTimeStep (Source/Advance.cpp:10-79); synchronises the stream
    Real estTimeStep(const MultiFab& S, Real cfl)
    {
        Real dt = 0.0;
        check(ocfd_estdt(lev_.handle(), S.dataPtr(), cfl, &parm_.p, &dt, s_), "ocfd_estdt");
        return dt;
    }

    // EBR::flow_advance (Source/Advance.cpp:81-150): scheme 0 Euler, 1 RK2,
    // 2 SSP-RK3; A holds U^n (ghosts filled) and gets U^(n+1); B, C scratch
    void flow_advance(MultiFab& A, MultiFab& B, MultiFab& C, Real time, Real dt, int scheme)
    {
        check(ocfd_advance(lev_.handle(), A.dataPtr(), B.dataPtr(), C.dataPtr(), time, dt, scheme, &parm_.p, flags_,
                           s_),
              "ocfd_advance");
    }

    // EBR::chemical_advance (Source/CHEM/ChemAdvance.cpp:7-85) on S and the
    // species MultiFab (9 comps, 4 ghosts), then the species renormalisation
    // of EBR::post_timestep (Source/EBR.cpp:333-357)
    void chemical_advance(MultiFab& S, MultiFab& Spec, Real dt)
    {
        check(ocfd_chem_advance(lev_.handle(), S.dataPtr(), Spec.dataPtr(), dt, &parm_.p, s_), "ocfd_chem_advance");
    }
    void renormalize_species(const MultiFab& S, MultiFab& Spec)
    {
        check(ocfd_species_renorm(lev_.handle(), S.dataPtr(), Spec.dataPtr(), s_), "ocfd_species_renorm");
    }

    // The single-level time loop of Source/main.cpp:61-70: initial dt from
    // estTimeStep (computeInitialDt, Source/EBR.cpp:189-229), then per step
    // computeNewDt (change_max 1.1, the stop_time clip, Source/EBR.cpp:231-306).
    // Runs until stop_time (< 0: none) or nsteps (< 0: none); returns the steps
    // taken and leaves time and dt for the next step in *time_io / *dt_io.
    int run(MultiFab& A, MultiFab& B, MultiFab& C, Real cfl, int scheme, Real stop_time, int nsteps,
            Real* time_io, Real* dt_io)
    {
        Real time = time_io ? *time_io : 0.0;
        Real dt = dt_io ? *dt_io : 0.0;
        if (dt <= 0.0) {
            dt = estTimeStep(A, cfl);
            if (stop_time >= 0.0 && time + dt > stop_time - 0.001 * dt) dt = stop_time - time;
        }
        int steps = 0;
        while ((nsteps < 0 || steps < nsteps) && (stop_time < 0.0 || time < stop_time)) {
            flow_advance(A, B, C, time, dt, scheme);
            time += dt;
            ++steps;
            const Real est = estTimeStep(A, cfl);
            Real dt_new = std::min(est, 1.1 * dt);
            if (stop_time >= 0.0 && time + dt_new > stop_time - 0.001 * dt_new) dt_new = stop_time - time;
            dt = dt_new;
        }
        if (time_io) *time_io = time;
        if (dt_io) *dt_io = dt;
        return steps;
    }

private:
    Level& lev_;
    Parm parm_;
    int flags_;
    hipStream_t s_;
};

}  // namespace ocfd

#endif  // OCFD_HPP
