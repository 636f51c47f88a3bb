// Case initial data of the configured workloads, computed on the host and
// uploaded into a level's state MultiFab (valid cells; ghosts are filled by
// ocfd_fill_boundary).
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "ocfd_internal.hpp"

namespace {

using ocfd::BoxDesc;
using ocfd::NC;
using ocfd::NG;

inline long hidx(const BoxDesc& B, int i, int j, int k, int n)
{
    return B.off_s + n * B.nptg + ((long)(k - B.lo[2] + NG) * B.nyg + (j - B.lo[1] + NG)) * B.nxg + (i - B.lo[0] + NG);
}

// Counter-based uniform variate in [0,1): splitmix64 finaliser of
// seed * 0x9E3779B97F4A7C15 + counter, top 53 bits.
inline double rng_uniform(uint64_t seed, uint64_t counter)
{
    uint64_t z = seed * 0x9E3779B97F4A7C15ull + counter;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// Exec/0_SOD/prob.H:11-38, prob_parm.H:9-12
void init_sod(const ocfd::LevelDev& L, const ocfd_parm& parm, std::vector<double>& h)
{
    const double p_l = 1.0, p_r = 0.1, rho_l = 1.0, rho_r = 0.125;
    for (const BoxDesc& B : L.boxes)
        for (int k = B.lo[2]; k <= B.hi[2]; ++k)
            for (int j = B.lo[1]; j <= B.hi[1]; ++j)
                for (int i = B.lo[0]; i <= B.hi[0]; ++i) {
                    const double x = L.geom.prob_lo[0] + (i + 0.5) * L.geom.dx[0];
                    const bool left = x < 0.5;
                    const double Pt = left ? p_l : p_r, rhot = left ? rho_l : rho_r;
                    h[hidx(B, i, j, k, 0)] = rhot;
                    h[hidx(B, i, j, k, 1)] = 0.0;
                    h[hidx(B, i, j, k, 2)] = 0.0;
                    h[hidx(B, i, j, k, 3)] = 0.0;
                    h[hidx(B, i, j, k, 4)] = Pt / (parm.eos_gamma - 1.0);
                }
}

// Exec/1_DMR/prob.H:11-45, prob_parm.H:10-17
void init_dmr(const ocfd::LevelDev& L, const ocfd_parm& parm, std::vector<double>& h)
{
    const double p_l = 116.5, p_r = 1.0, rho_l = 8.0, rho_r = 1.4;
    const double u_l = 8.25 * std::cos(M_PI * (1.0 / 6.0)), u_r = 0.0;
    const double v_l = -8.25 * std::sin(M_PI * (1.0 / 6.0)), v_r = 0.0;
    for (const BoxDesc& B : L.boxes)
        for (int k = B.lo[2]; k <= B.hi[2]; ++k)
            for (int j = B.lo[1]; j <= B.hi[1]; ++j)
                for (int i = B.lo[0]; i <= B.hi[0]; ++i) {
                    const double x = L.geom.prob_lo[0] + (i + 0.5) * L.geom.dx[0];
                    const double y = L.geom.prob_lo[1] + (j + 0.5) * L.geom.dx[1];
                    const bool left = x < 1.0 / 6.0 + std::sqrt(1 / 3.0) * y;
                    const double Pt = left ? p_l : p_r, rhot = left ? rho_l : rho_r;
                    const double ux = left ? u_l : u_r, vy = left ? v_l : v_r;
                    h[hidx(B, i, j, k, 0)] = rhot;
                    h[hidx(B, i, j, k, 1)] = rhot * ux;
                    h[hidx(B, i, j, k, 2)] = rhot * vy;
                    h[hidx(B, i, j, k, 3)] = 0.0;
                    h[hidx(B, i, j, k, 4)] = Pt / (parm.eos_gamma - 1.0) + 0.5 * rhot * (ux * ux + vy * vy);
                }
}

// 3D viscous Kelvin-Helmholtz shear layer on a periodic unit cube: air at
// p0 = 101325 Pa, outer stream at 300 K, density ratio 2 inside
// 1/4 < y < 3/4, tanh profiles of thickness 1/32, streams at +-Mach 0.5 of the
// outer sound speed, seeded uniform velocity perturbation of amplitude 1e-2 of
// the stream speed keyed by (seed, 3*cell + component).
void init_kh3d(const ocfd::LevelDev& L, const ocfd_parm& parm, uint64_t seed, std::vector<double>& h)
{
    const ocfd::GeomDesc& G = L.geom;
    const double gamma = parm.eos_gamma;
    const double p0 = 101325.0;
    const double rho_out = p0 / (parm.Rg * 300.0);
    const double rho_in = 2.0 * rho_out;
    const double c_out = std::sqrt(gamma * p0 / rho_out);
    const double U = 0.5 * c_out;
    const double delta = 1.0 / 32.0;
    const double amp = 1.0e-2 * U;
    const long nx = G.dom_hi[0] - G.dom_lo[0] + 1, ny = G.dom_hi[1] - G.dom_lo[1] + 1;
    for (const BoxDesc& B : L.boxes)
        for (int k = B.lo[2]; k <= B.hi[2]; ++k)
            for (int j = B.lo[1]; j <= B.hi[1]; ++j) {
                const double y = G.prob_lo[1] + (j + 0.5) * G.dx[1];
                const double sh = 0.5 * (std::tanh((y - 0.25) / delta) - std::tanh((y - 0.75) / delta));
                const double rho = rho_out + (rho_in - rho_out) * sh;
                for (int i = B.lo[0]; i <= B.hi[0]; ++i) {
                    const uint64_t cell = ((uint64_t)(k - G.dom_lo[2]) * ny + (j - G.dom_lo[1])) * nx + (i - G.dom_lo[0]);
                    const double u = U * (2.0 * sh - 1.0) + amp * (2.0 * rng_uniform(seed, 3 * cell + 0) - 1.0);
                    const double v = amp * (2.0 * rng_uniform(seed, 3 * cell + 1) - 1.0);
                    const double w = amp * (2.0 * rng_uniform(seed, 3 * cell + 2) - 1.0);
                    h[hidx(B, i, j, k, 0)] = rho;
                    h[hidx(B, i, j, k, 1)] = rho * u;
                    h[hidx(B, i, j, k, 2)] = rho * v;
                    h[hidx(B, i, j, k, 3)] = rho * w;
                    h[hidx(B, i, j, k, 4)] = p0 / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w);
                }
            }
}

// Exec/3_FSTEP/prob.H:11-40 with prob_parm.H defaults: the inflow state
// everywhere, covered cells with zero momentum and the same energy
int init_eb_inflow(const ocfd::LevelDev& L, const ocfd_parm& parm, std::vector<double>& h)
{
    if (!L.has_eb) return OCFD_EINVAL;
    const double Pt = 2026.5, rhot = 0.08938, uxt = 534.49;
    ocfd_geom g{};
    for (int d = 0; d < 3; ++d) {
        g.dom_lo[d] = L.geom.dom_lo[d]; g.dom_hi[d] = L.geom.dom_hi[d];
        g.prob_lo[d] = L.geom.prob_lo[d]; g.dx[d] = L.geom.dx[d];
    }
    for (const BoxDesc& B : L.boxes) {
        const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1, nz = B.hi[2] - B.lo[2] + 1;
        std::vector<unsigned> fl((size_t)nx * ny * nz);
        ocfd_eb_geometry(&L.eb_spec, &g, B.lo, B.hi, 0, fl.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                         nullptr, nullptr);
        for (int k = B.lo[2]; k <= B.hi[2]; ++k)
            for (int j = B.lo[1]; j <= B.hi[1]; ++j)
                for (int i = B.lo[0]; i <= B.hi[0]; ++i) {
                    const bool cov = (fl[((size_t)(k - B.lo[2]) * ny + (j - B.lo[1])) * nx + (i - B.lo[0])] & 3u) == 3u;
                    const double et = Pt / (parm.eos_gamma - 1.0);
                    h[hidx(B, i, j, k, 0)] = rhot;
                    h[hidx(B, i, j, k, 1)] = cov ? 0.0 : rhot * uxt;
                    h[hidx(B, i, j, k, 2)] = 0.0;
                    h[hidx(B, i, j, k, 3)] = 0.0;
                    h[hidx(B, i, j, k, 4)] = et + 0.5 * rhot * uxt * uxt;
                }
    }
    return OCFD_OK;
}

} // namespace

extern "C" {

// inflow_state of Exec/3_FSTEP/prob.cpp:24-29 (literal 0.4)
int ocfd_inflow_state(double p, double rho, double u, double* state)
{
    if (!state) return OCFD_EINVAL;
    if (p < 0) p = 2026.5;
    if (rho < 0) rho = 0.08938;
    if (u < 0) u = 534.49;
    state[0] = rho;
    state[1] = u * rho;
    state[2] = 0;
    state[3] = 0;
    state[4] = p / 0.4 + 0.5 * rho * u * u;
    return OCFD_OK;
}

// Initial data of case `kind` (0 Sod, 1 double Mach reflection, 2 3D KH,
// 3 inflow over an embedded body) into
// the valid cells of state MultiFab S (4 ghosts, 5 comps) of level lev;
// ghost cells are zeroed.  Synchronous.
int ocfd_init_state(ocfd_level lev, int kind, double* S, const ocfd_parm* parm, unsigned long long seed)
{
    if (!lev || !S || !parm || kind < 0 || kind > 3) {
        ocfd::set_error("ocfd_init_state: invalid arguments");
        return OCFD_EINVAL;
    }
    ocfd::tcache_clear(lev->dev);   // new data in the level's states
    const ocfd::LevelDev& L = lev->dev;
    std::vector<double> h((size_t)L.mf_state, 0.0);
    if (kind == 0) init_sod(L, *parm, h);
    else if (kind == 1) init_dmr(L, *parm, h);
    else if (kind == 2) init_kh3d(L, *parm, seed, h);
    else if (init_eb_inflow(L, *parm, h)) {
        ocfd::set_error("ocfd_init_state: the inflow-over-body case needs a level with an embedded boundary");
        return OCFD_EINVAL;
    }
    hipError_t e = hipMemcpy(S, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        ocfd::set_error(std::string("ocfd_init_state: ") + hipGetErrorString(e));
        return OCFD_EHIP;
    }
    return OCFD_OK;
}

// DMR inflow states (Exec/1_DMR/prob.cpp:28-40, with its literal 0.4)
int ocfd_dmr_states(double* left, double* right)
{
    const double p_l = 116.5, p_r = 1.0, rho_l = 8.0, rho_r = 1.4;
    const double u_l = 8.25 * std::cos(M_PI * (1.0 / 6.0)), u_r = 0.0;
    const double v_l = -8.25 * std::sin(M_PI * (1.0 / 6.0)), v_r = 0.0;
    left[0] = rho_l; left[1] = u_l * rho_l; left[2] = v_l * rho_l; left[3] = 0;
    left[4] = p_l / 0.4 + 0.5 * rho_l * (u_l * u_l + v_l * v_l);
    right[0] = rho_r; right[1] = u_r * rho_r; right[2] = v_r * rho_r; right[3] = 0;
    right[4] = p_r / 0.4 + 0.5 * rho_r * (u_r * u_r + v_r * v_r);
    return OCFD_OK;
}

} // extern "C"
