// Chemistry kernels: the per-cell linearised-implicit step over a level
// (EBR::chemical_advance, Source/CHEM/ChemAdvance.cpp:7-85), the species
// renormalisation (EBR::post_timestep, Source/EBR.cpp:333-357) and batched
// constant-volume reactors (0D ignition) with fixed or adaptive substeps.
//
// The adaptive reactor kernel is persistent: every lane takes reactors from a
// device work counter until the batch is exhausted, so lanes that finish an
// easy reactor pick up the next one instead of idling behind a stiff one.
#include <algorithm>
#include <string>

#include "ocfd_chem_device.hpp"
#include "ocfd_device.hpp"
#include "ocfd_internal.hpp"

namespace ocfd {

using chem::NS;

// reactor batch layout: SoA of OCFD_REACTOR_NCOMP rows of n doubles:
// row 0 rho, row 1 sensible energy density E, rows 2..10 partial densities
constexpr int RROW_RHO = 0, RROW_E = 1, RROW_Y = 2;

// dt: the chemistry step; under device time control (ctl not null) it is
// (0.5 ctl->dt)/iter instead, the Strang substep of EBR::advance
// (Source/Advance.cpp:160-176)
__global__ __launch_bounds__(64) void k_chem_level(const BoxDesc* __restrict__ boxes, const int4* __restrict__ tiles,
                                                   double* __restrict__ S, double* __restrict__ Spec, double dt,
                                                   double smallr, const ocfd_step_ctl* __restrict__ ctl, int iter)
{
    if (ctl_done(ctl)) return;
    if (ctl) dt = (0.5 * ctl->dt) / iter;
    const int4 t = tiles[blockIdx.x];
    if ((int)threadIdx.x >= t.z) return;
    const BoxDesc B = boxes[t.x];
    if (B.type == 2) return;   // covered boxes are skipped (ChemAdvance.cpp:20)
    const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1;
    const int g = t.y + threadIdx.x;
    const int i = B.lo[0] + g % nx, j = B.lo[1] + (g / nx) % ny, k = B.lo[2] + g / (nx * ny);
    const long s0 = sidx(B, i, j, k);
    const long st = B.nptg;
    const long p0 = s0 - B.off_s + B.off_s / NC * NS;   // species FAB: same grown layout, 9 comps
    // Source/CHEM/ChemAdvance.cpp:36-46
    const double rho = fmax(S[s0], smallr);
    const double rhoinv = 1.0 / rho;
    const double ux = S[s0 + st] * rhoinv, uy = S[s0 + 2 * st] * rhoinv, uz = S[s0 + 3 * st] * rhoinv;
    const double kineng = 0.5 * (ux * ux + uy * uy + uz * uz);
    const double ei = S[s0 + 4 * st] - rho * kineng;
    double rhoi[NS];
#pragma unroll
    for (int n = 0; n < NS; ++n) rhoi[n] = Spec[p0 + n * st];
    const double dei = chem::cell_step(ei, rhoi, dt);
#pragma unroll
    for (int n = 0; n < NS; ++n) Spec[p0 + n * st] = rhoi[n];
    S[s0 + 4 * st] += dei;
}

__global__ __launch_bounds__(256) void k_renorm_level(const BoxDesc* __restrict__ boxes, const int4* __restrict__ tiles,
                                                      const double* __restrict__ S, double* __restrict__ Spec,
                                                      const ocfd_step_ctl* __restrict__ ctl)
{
    if (ctl_done(ctl)) return;
    const int4 t = tiles[blockIdx.x];
    const BoxDesc B = boxes[t.x];
    const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1;
    for (int c = threadIdx.x; c < t.z; c += 256) {
        const int g = t.y + c;
        const int i = B.lo[0] + g % nx, j = B.lo[1] + (g / nx) % ny, k = B.lo[2] + g / (nx * ny);
        const long s0 = sidx(B, i, j, k);
        const long st = B.nptg;
        const long p0 = s0 - B.off_s + B.off_s / NC * NS;
        // Source/EBR.cpp:344-354
        double r[NS];
        double rho0 = 0;
#pragma unroll
        for (int n = 0; n < NS; ++n) {
            r[n] = Spec[p0 + n * st];
            if (r[n] < 0) r[n] = 0.0;
            rho0 += r[n];
        }
        const double tmp = S[s0] / rho0;
#pragma unroll
        for (int n = 0; n < NS; ++n) Spec[p0 + n * st] = r[n] * tmp;
    }
}

// constant-volume H2/air reactor at (T0, p0, phi): X(H2):X(O2):X(N2) = 2phi:1:3.76
__global__ void k_reactor_init(long n, const double* __restrict__ T0, const double* __restrict__ p0,
                               const double* __restrict__ phi, double* __restrict__ st)
{
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double X[NS] = {0}, Y[NS];
    const double tot = 2.0 * phi[r] + 1.0 + 3.76;
    X[0] = 2.0 * phi[r] / tot;
    X[1] = 1.0 / tot;
    X[8] = 3.76 / tot;
    // CKXTY, CKRHOY (Source/CHEM/LiDryer.H:21-66)
    double XW = 0.0;
#pragma unroll
    for (int i = 0; i < NS; ++i) XW += X[i] * chem::mw[i];
#pragma unroll
    for (int i = 0; i < NS; ++i) Y[i] = X[i] * chem::mw[i] / XW;
    double YOW = 0, tmp[NS];
#pragma unroll
    for (int i = 0; i < NS; ++i) tmp[i] = Y[i] / chem::mw[i];
#pragma unroll
    for (int i = 0; i < NS; ++i) YOW += tmp[i];
    const double rho = p0[r] / (chem::Ru * T0[r] * YOW);
    double rhoi[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) rhoi[k] = Y[k] * rho;
    st[RROW_RHO * n + r] = rho;
    st[RROW_E * n + r] = chem::ckubms(T0[r], rhoi);
#pragma unroll
    for (int k = 0; k < NS; ++k) st[(RROW_Y + k) * n + r] = rhoi[k];
}

__global__ __launch_bounds__(64) void k_reactor_fixed(long n, double* __restrict__ st, double dt, int nsteps)
{
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double rhoi[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) rhoi[k] = st[(RROW_Y + k) * n + r];
    double E = st[RROW_E * n + r];
    for (int s = 0; s < nsteps; ++s) E += chem::cell_step(E, rhoi, dt);
#pragma unroll
    for (int k = 0; k < NS; ++k) st[(RROW_Y + k) * n + r] = rhoi[k];
    st[RROW_E * n + r] = E;
}

__global__ void k_counter_init(unsigned long long* c) { *c = 0ull; }

// Adaptive integration with step doubling on the linearised-implicit step:
// one step of h (y1) against two of h/2 (y2), error max_k |y2-y1| /
// (atol rho + rtol |y2|); on acceptance keep the Richardson value 2 y2 - y1
// (second order, densities clipped at 0); rescale h by clamp(0.9/sqrt(err), 0.2, 4).
//
// One attempt of an adaptive substep of a reactor: step doubling on the
// linearised-implicit step; on acceptance t advances and the Richardson value
// is kept, either way h is rescaled.  LPR = 2: the two lanes of a pair share
// the work (the even lane takes the step of h, the odd lane the two of h/2)
// and end with the same values.
template <int LPR>
__device__ __forceinline__ void adaptive_substep(bool odd, double* rhoi, double& E, double rho, double& t, double& h,
                                                 int& acc, int& rej, double rtol, double atol, double t_end)
{
    if (t + h > t_end) h = t_end - t;
    double r1[NS], r2[NS], E1, E2;
    if (LPR == 2) {
        // even: y1 = one step of h; odd: y2 = two steps of h/2 (one
        // code path for both, so the step is inlined once)
        double rr[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) rr[k] = rhoi[k];
        double Ex = E;
        const double hs = odd ? 0.5 * h : h;
        for (int pass = 0; pass < (odd ? 2 : 1); ++pass) Ex += chem::cell_step(Ex, rr, hs);
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const double other = __shfl_xor(rr[k], 1);
            r1[k] = odd ? other : rr[k];
            r2[k] = odd ? rr[k] : other;
        }
        const double Eo = __shfl_xor(Ex, 1);
        E1 = odd ? Eo : Ex;
        E2 = odd ? Ex : Eo;
    } else {
#pragma unroll
        for (int k = 0; k < NS; ++k) { r1[k] = rhoi[k]; r2[k] = rhoi[k]; }
        E1 = E + chem::cell_step(E, r1, h);
        E2 = E + chem::cell_step(E, r2, 0.5 * h);
        E2 += chem::cell_step(E2, r2, 0.5 * h);
    }
    double err = 0.0;
#pragma unroll
    for (int k = 0; k < NS; ++k) err = fmax(err, fabs(r2[k] - r1[k]) / (atol * rho + rtol * fabs(r2[k])));
    if (err <= 1.0) {
#pragma unroll
        for (int k = 0; k < NS; ++k) rhoi[k] = fmax(2.0 * r2[k] - r1[k], 0.0);
        E = 2.0 * E2 - E1;
        t += h;
        ++acc;
        const double f = err > 0.0 ? 0.9 / sqrt(err) : 4.0;
        h *= fmin(4.0, fmax(0.2, f));
    } else {
        ++rej;
        h *= fmax(0.2, 0.9 / sqrt(err));
    }
}

// LPR = 2 (batches too small to fill the chip): two lanes per reactor, the
// even lane takes the step of h, the odd lane the two steps of h/2 (its
// first one alongside the even lane's step), then the pair swaps results and
// both make the same decision on the same values -- a substep costs two
// implicit steps of wall time instead of three.  LPR = 1 (full batches): one
// lane takes all three.  Both give the same results bit for bit.
template <int LPR>
__global__ __launch_bounds__(64) void k_reactor_adaptive(long n, double* __restrict__ st, double t_end, double h0,
                                                         double rtol, double atol, int max_steps,
                                                         int* __restrict__ steps_out, int* __restrict__ rej_out,
                                                         unsigned long long* __restrict__ counter)
{
    const int lane = threadIdx.x & 63;
    const bool odd = LPR == 2 && (lane & 1);
    // the pair's reactor: taken by the even lane, shared with the odd one
    auto next = [&]() {
        long q = odd ? 0 : (long)atomicAdd(counter, 1ull);
        return LPR == 2 ? (long)__shfl(q, lane & ~1) : q;
    };
    long r = next();
    double rhoi[NS], E = 0, rho = 0, t = 0, h = h0;
    int acc = 0, rej = 0;
    auto load = [&](long q) {
#pragma unroll
        for (int k = 0; k < NS; ++k) rhoi[k] = st[(RROW_Y + k) * n + q];
        E = st[RROW_E * n + q];
        rho = st[RROW_RHO * n + q];
        t = 0.0; h = h0; acc = 0; rej = 0;
    };
    if (r < n) load(r);
    while (r < n) {
        bool done = false, failed = false;
        if (t >= t_end) done = true;
        else if (acc + rej >= max_steps) { done = true; failed = true; }
        if (!done) adaptive_substep<LPR>(odd, rhoi, E, rho, t, h, acc, rej, rtol, atol, t_end);
        if (done) {
            if (!odd) {
#pragma unroll
                for (int k = 0; k < NS; ++k) st[(RROW_Y + k) * n + r] = rhoi[k];
                st[RROW_E * n + r] = E;
                if (steps_out) steps_out[r] = failed ? -1 : acc;
                if (rej_out) rej_out[r] = rej;
            }
            r = next();
            if (r < n) load(r);
        }
    }
}

// Resumable form of the adaptive integration, for batches that must not hold
// a hardware queue for their whole integration (a served reactor batch next
// to flow cases): every reactor (a lane, or a pair of lanes) takes at most
// `budget` substep attempts per launch from its saved time and step size
// (th: t in the first n entries, h in the next n) and saves them again; the
// reactors not yet at t_end are counted in *remaining.  The substeps are
// those of k_reactor_adaptive, so the results are identical.
template <int LPR>
__global__ __launch_bounds__(64) void k_reactor_chunk(long n, double* __restrict__ st, double* __restrict__ th,
                                                      double t_end, double rtol, double atol, int max_steps,
                                                      int budget, int* __restrict__ acc_io, int* __restrict__ rej_io,
                                                      unsigned int* __restrict__ remaining)
{
    const int lane = threadIdx.x & 63;
    const bool odd = LPR == 2 && (lane & 1);
    const long r = ((long)blockIdx.x * 64 + threadIdx.x) / LPR;   // the same for both lanes of a pair
    if (r >= n) return;
    int acc = acc_io[r], rej = rej_io[r];
    double t = th[r], h = th[n + r];
    if (acc < 0 || t >= t_end) return;   // stopped at max_steps, or done
    double rhoi[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) rhoi[k] = st[(RROW_Y + k) * n + r];
    double E = st[RROW_E * n + r];
    const double rho = st[RROW_RHO * n + r];
    bool failed = false;
    for (int it = 0; it < budget; ++it) {
        if (t >= t_end) break;
        if (acc + rej >= max_steps) { failed = true; break; }
        adaptive_substep<LPR>(odd, rhoi, E, rho, t, h, acc, rej, rtol, atol, t_end);
    }
    if (!odd) {
#pragma unroll
        for (int k = 0; k < NS; ++k) st[(RROW_Y + k) * n + r] = rhoi[k];
        st[RROW_E * n + r] = E;
        th[r] = t;
        th[n + r] = h;
        acc_io[r] = failed ? -1 : acc;
        rej_io[r] = rej;
        if (!failed && t < t_end) atomicAdd(remaining, 1u);
    }
}

__global__ void k_reactor_chunk_init(long n, double* __restrict__ th, double h0, int* __restrict__ acc,
                                     int* __restrict__ rej)
{
    for (long r = (long)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x) {
        th[r] = 0.0;
        th[n + r] = h0;
        acc[r] = 0;
        rej[r] = 0;
    }
}

__global__ void k_zero_u32(unsigned int* p) { *p = 0u; }

// One attempt of an adaptive RODAS3 step of a reactor: on acceptance the
// densities are clipped at 0 (the energy follows the clipped change) and
// h <- h clamp(0.9 err^(-1/3), 0.2, 6), not above h after a rejection; on
// rejection h <- h max(0.2, 0.9 err^(-1/3)).  Tg: the temperature guess of the
// next step (the last accepted stage's temperature; 0 before the first).
__device__ __forceinline__ void rosenbrock_substep(double* rhoi, double& E, double rho, double& t, double& h, double& Tg,
                                                   bool& last_rej, int& acc, int& rej, double rtol, double atol,
                                                   double t_end, chem::LaneMat M)
{
#pragma clang fp contract(off)
    if (t + h > t_end) h = t_end - t;
    double yn[NS], En, Tl;
    const double err = chem::rodas3_step(rhoi, E, Tg, h, rho, rtol, atol, yn, En, Tl, M);
    const double fac = 0.9 / cbrt(err);
    if (err <= 1.0) {
        double dy[NS];
#pragma unroll
        for (int n = 0; n < NS; ++n) {
            const double v = fmax(yn[n], 0.0);
            dy[n] = v - rhoi[n];
            rhoi[n] = v;
        }
        E += chem::energy_of(dy);
        Tg = Tl;
        t += h;
        ++acc;
        double f = fmin(6.0, fmax(0.2, fac));
        if (last_rej) f = fmin(f, 1.0);
        h *= f;
        last_rej = false;
    } else {
        ++rej;
        h *= fmax(0.2, fac);
        last_rej = true;
    }
}

// Persistent RODAS3 integration, one lane per reactor taken from a device
// work counter (lanes that finish an easy reactor take the next one)
__global__ __launch_bounds__(64) void k_reactor_rosenbrock(long n, double* __restrict__ st, double t_end, double h0,
                                                           double rtol, double atol, int max_steps,
                                                           int* __restrict__ steps_out, int* __restrict__ rej_out,
                                                           unsigned long long* __restrict__ counter)
{
    __shared__ double mat[NS * NS * 64];   // the iteration matrix of each lane
    const chem::LaneMat M{mat + threadIdx.x};
    long r = (long)atomicAdd(counter, 1ull);
    while (r < n) {
        double rhoi[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) rhoi[k] = st[(RROW_Y + k) * n + r];
        double E = st[RROW_E * n + r];
        const double rho = st[RROW_RHO * n + r];
        double t = 0.0, h = h0, Tg = 0.0;
        bool last_rej = false, failed = false;
        int acc = 0, rej = 0;
        while (t < t_end) {
            if (acc + rej >= max_steps) { failed = true; break; }
            rosenbrock_substep(rhoi, E, rho, t, h, Tg, last_rej, acc, rej, rtol, atol, t_end, M);
        }
#pragma unroll
        for (int k = 0; k < NS; ++k) st[(RROW_Y + k) * n + r] = rhoi[k];
        st[RROW_E * n + r] = E;
        if (steps_out) steps_out[r] = failed ? -1 : acc;
        if (rej_out) rej_out[r] = rej;
        r = (long)atomicAdd(counter, 1ull);
    }
}

// Resumable RODAS3 integration (see k_reactor_chunk): th holds t, h and the
// temperature guess (n entries each); a rejected last attempt is kept as a
// negative h.  The substeps are those of k_reactor_rosenbrock.
__global__ __launch_bounds__(64) void k_rosenbrock_chunk(long n, double* __restrict__ st, double* __restrict__ th,
                                                         double t_end, double rtol, double atol, int max_steps,
                                                         int budget, int* __restrict__ acc_io, int* __restrict__ rej_io,
                                                         unsigned int* __restrict__ remaining)
{
    __shared__ double mat[NS * NS * 64];   // the iteration matrix of each lane
    const chem::LaneMat M{mat + threadIdx.x};
    const long r = (long)blockIdx.x * 64 + threadIdx.x;
    if (r >= n) return;
    int acc = acc_io[r], rej = rej_io[r];
    double t = th[r], h = th[n + r], Tg = th[2 * n + r];
    if (acc < 0 || t >= t_end) return;
    bool last_rej = h < 0.0;
    h = fabs(h);
    double rhoi[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) rhoi[k] = st[(RROW_Y + k) * n + r];
    double E = st[RROW_E * n + r];
    const double rho = st[RROW_RHO * n + r];
    bool failed = false;
    for (int it = 0; it < budget; ++it) {
        if (t >= t_end) break;
        if (acc + rej >= max_steps) { failed = true; break; }
        rosenbrock_substep(rhoi, E, rho, t, h, Tg, last_rej, acc, rej, rtol, atol, t_end, M);
    }
#pragma unroll
    for (int k = 0; k < NS; ++k) st[(RROW_Y + k) * n + r] = rhoi[k];
    st[RROW_E * n + r] = E;
    th[r] = t;
    th[n + r] = last_rej ? -h : h;
    th[2 * n + r] = Tg;
    acc_io[r] = failed ? -1 : acc;
    rej_io[r] = rej;
    if (!failed && t < t_end) atomicAdd(remaining, 1u);
}

__global__ void k_rosenbrock_chunk_init(long n, double* __restrict__ th, double h0, int* __restrict__ acc,
                                        int* __restrict__ rej)
{
    for (long r = (long)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x) {
        th[r] = 0.0;
        th[n + r] = h0;
        th[2 * n + r] = 0.0;
        acc[r] = 0;
        rej[r] = 0;
    }
}

__global__ void k_reactor_T(long n, const double* __restrict__ st, double* __restrict__ T)
{
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double rhoi[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) rhoi[k] = st[(RROW_Y + k) * n + r];
    T[r] = chem::T_given_ey(st[RROW_E * n + r], rhoi);
}

void launch_chem_level(const LevelDev& L, double* S, double* Spec, double dt, double smallr,
                       const ocfd_step_ctl* ctl, int iter, hipStream_t s)
{
    if (L.ntile_v)
        hipLaunchKernelGGL(k_chem_level, dim3(L.ntile_v), dim3(64), 0, s, L.d_boxes, L.d_tile_v, S, Spec, dt, smallr,
                           ctl, iter);
}

void launch_renorm(const LevelDev& L, const double* S, double* Spec, const ocfd_step_ctl* ctl, hipStream_t s)
{
    if (L.ntile_c)
        hipLaunchKernelGGL(k_renorm_level, dim3(L.ntile_c), dim3(256), 0, s, L.d_boxes, L.d_tile_c, S, Spec, ctl);
}

static int launch_check(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string(what) + ": " + hipGetErrorString(e));
        return OCFD_EHIP;
    }
    return OCFD_OK;
}

} // namespace ocfd

using namespace ocfd;

extern "C" {

int ocfd_chem_advance(ocfd_level lev, double* S, double* Spec, double dt, const ocfd_parm* parm, void* stream)
{
    if (!lev || !S || !Spec || !parm) { set_error("ocfd_chem_advance: invalid arguments"); return OCFD_EINVAL; }
    launch_chem_level(lev->dev, S, Spec, dt, parm->smallr, nullptr, 1, (hipStream_t)stream);
    return launch_check("ocfd_chem_advance");
}

int ocfd_species_renorm(ocfd_level lev, const double* S, double* Spec, void* stream)
{
    if (!lev || !S || !Spec) { set_error("ocfd_species_renorm: invalid arguments"); return OCFD_EINVAL; }
    launch_renorm(lev->dev, S, Spec, nullptr, (hipStream_t)stream);
    return launch_check("ocfd_species_renorm");
}

int ocfd_reactor_init(long n, const double* T0, const double* p0, const double* phi, double* state, void* stream)
{
    if (n <= 0 || !T0 || !p0 || !phi || !state) { set_error("ocfd_reactor_init: invalid arguments"); return OCFD_EINVAL; }
    hipLaunchKernelGGL(k_reactor_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, n, T0, p0,
                       phi, state);
    return launch_check("ocfd_reactor_init");
}

int ocfd_reactor_advance_fixed(long n, double* state, double dt, int nsteps, void* stream)
{
    if (n <= 0 || !state || nsteps < 0) { set_error("ocfd_reactor_advance_fixed: invalid arguments"); return OCFD_EINVAL; }
    hipLaunchKernelGGL(k_reactor_fixed, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, (hipStream_t)stream, n, state,
                       dt, nsteps);
    return launch_check("ocfd_reactor_advance_fixed");
}

int ocfd_reactor_advance_adaptive(long n, double* state, double t_end, double h0, double rtol, double atol,
                                  int max_steps, int* steps_out, int* rejects_out, void* work, void* stream)
{
    if (n <= 0 || !state || !work || t_end < 0 || h0 <= 0 || rtol <= 0 || atol < 0 || max_steps <= 0) {
        set_error("ocfd_reactor_advance_adaptive: invalid arguments");
        return OCFD_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    unsigned long long* ctr = (unsigned long long*)work;
    hipLaunchKernelGGL(k_counter_init, dim3(1), dim3(1), 0, s, ctr);
    int dev = 0, ncu = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    // persistent grid of 64-lane workgroups; one lane per reactor once the
    // batch fills one wave per SIMD (4 per CU), else two lanes per reactor
    const long lanes_full = 64L * 4 * std::max(ncu, 1);
    if (n >= lanes_full) {
        const long wgs = std::min<long>((n + 63) / 64, (long)std::max(ncu, 1) * 8);
        hipLaunchKernelGGL(k_reactor_adaptive<1>, dim3((unsigned)wgs), dim3(64), 0, s, n, state, t_end, h0, rtol,
                           atol, max_steps, steps_out, rejects_out, ctr);
    } else {
        const long wgs = (n + 31) / 32;
        hipLaunchKernelGGL(k_reactor_adaptive<2>, dim3((unsigned)wgs), dim3(64), 0, s, n, state, t_end, h0, rtol,
                           atol, max_steps, steps_out, rejects_out, ctr);
    }
    return launch_check("ocfd_reactor_advance_adaptive");
}

int ocfd_reactor_adaptive_begin(long n, double* th, double h0, int* steps, int* rejects, void* stream)
{
    if (n <= 0 || !th || !steps || !rejects || !(h0 > 0)) {
        set_error("ocfd_reactor_adaptive_begin: invalid arguments");
        return OCFD_EINVAL;
    }
    const long nb = std::min<long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_reactor_chunk_init, dim3((unsigned)nb), dim3(256), 0, (hipStream_t)stream, n, th, h0, steps,
                       rejects);
    return launch_check("ocfd_reactor_adaptive_begin");
}

int ocfd_reactor_adaptive_chunk(long n, double* state, double* th, double t_end, double rtol, double atol,
                                int max_steps, int budget, int* steps, int* rejects, unsigned int* remaining,
                                void* stream)
{
    if (n <= 0 || !state || !th || !steps || !rejects || !remaining || t_end < 0 || rtol <= 0 || atol < 0 ||
        max_steps <= 0 || budget <= 0) {
        set_error("ocfd_reactor_adaptive_chunk: invalid arguments");
        return OCFD_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, remaining);
    int dev = 0, ncu = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const long lanes_full = 64L * 4 * std::max(ncu, 1);
    if (n >= lanes_full)
        hipLaunchKernelGGL(k_reactor_chunk<1>, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, n, state, th, t_end,
                           rtol, atol, max_steps, budget, steps, rejects, remaining);
    else
        hipLaunchKernelGGL(k_reactor_chunk<2>, dim3((unsigned)((n + 31) / 32)), dim3(64), 0, s, n, state, th, t_end,
                           rtol, atol, max_steps, budget, steps, rejects, remaining);
    return launch_check("ocfd_reactor_adaptive_chunk");
}

int ocfd_reactor_advance_rosenbrock(long n, double* state, double t_end, double h0, double rtol, double atol,
                                    int max_steps, int* steps_out, int* rejects_out, void* work, void* stream)
{
    if (n <= 0 || !state || !work || t_end < 0 || h0 <= 0 || rtol <= 0 || atol < 0 || max_steps <= 0) {
        set_error("ocfd_reactor_advance_rosenbrock: invalid arguments");
        return OCFD_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    unsigned long long* ctr = (unsigned long long*)work;
    hipLaunchKernelGGL(k_counter_init, dim3(1), dim3(1), 0, s, ctr);
    int dev = 0, ncu = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const long wgs = std::min<long>((n + 63) / 64, (long)std::max(ncu, 1) * 8);
    hipLaunchKernelGGL(k_reactor_rosenbrock, dim3((unsigned)wgs), dim3(64), 0, s, n, state, t_end, h0, rtol, atol,
                       max_steps, steps_out, rejects_out, ctr);
    return launch_check("ocfd_reactor_advance_rosenbrock");
}

int ocfd_reactor_rosenbrock_begin(long n, double* th, double h0, int* steps, int* rejects, void* stream)
{
    if (n <= 0 || !th || !steps || !rejects || !(h0 > 0)) {
        set_error("ocfd_reactor_rosenbrock_begin: invalid arguments");
        return OCFD_EINVAL;
    }
    const long nb = std::min<long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_rosenbrock_chunk_init, dim3((unsigned)nb), dim3(256), 0, (hipStream_t)stream, n, th, h0,
                       steps, rejects);
    return launch_check("ocfd_reactor_rosenbrock_begin");
}

int ocfd_reactor_rosenbrock_chunk(long n, double* state, double* th, double t_end, double rtol, double atol,
                                  int max_steps, int budget, int* steps, int* rejects, unsigned int* remaining,
                                  void* stream)
{
    if (n <= 0 || !state || !th || !steps || !rejects || !remaining || t_end < 0 || rtol <= 0 || atol < 0 ||
        max_steps <= 0 || budget <= 0) {
        set_error("ocfd_reactor_rosenbrock_chunk: invalid arguments");
        return OCFD_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, remaining);
    hipLaunchKernelGGL(k_rosenbrock_chunk, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, n, state, th, t_end, rtol,
                       atol, max_steps, budget, steps, rejects, remaining);
    return launch_check("ocfd_reactor_rosenbrock_chunk");
}

int ocfd_reactor_temperature(long n, const double* state, double* T, void* stream)
{
    if (n <= 0 || !state || !T) { set_error("ocfd_reactor_temperature: invalid arguments"); return OCFD_EINVAL; }
    hipLaunchKernelGGL(k_reactor_T, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, n, state, T);
    return launch_check("ocfd_reactor_temperature");
}

} // extern "C"
