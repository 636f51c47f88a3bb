// On-disk plotfiles (AMReX "HyperCLaw-V1.1", the plot file type of EBR,
// Source/EBR.H:33-35) and the plotfile comparison of fcompare
// (Util/Plotfile/fcompare.cpp).
//
// Plotfile layout (the AmrLevel::writePlotFile format restated in
// Source/IO.cpp:32-262): a text Header with the variable names, time, level
// count, physical box, refinement ratios, index domains, level steps and cell
// sizes, then per level the grids' physical extents and the path of the
// level's MultiFab; per level a VisMF header Level_<l>/Cell_H (version,
// how, ncomp, nghost, BoxArray, FabOnDisk entries, per-FAB minima and maxima)
// and one data file Level_<l>/Cell_D_00000 holding every FAB as a one-line
// FAB header (real descriptor, box, ncomp) followed by the raw little-endian
// IEEE doubles, component-major with i fastest.
//
// Plot variables: the 5 conserved components, Cost (written as 0), the
// derived velocity ux uy uz (ebr_dervel) and pressure (ebr_derpres,
// Source/Setup.cpp:94-179) of derive_plot_vars = "pressure velocity"
// (Exec/1_DMR/inputs), and vfrac.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <map>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "ocfd_device.hpp"
#include "ocfd_internal.hpp"

namespace ocfd {

constexpr int NPLOT = OCFD_PLOT_NVARS;

__device__ __forceinline__ long gl_index(const BoxDesc& B, int i, int j, int k)
{
    return ((long)(k - B.lo[2] + NG) * B.nyg + (j - B.lo[1] + NG)) * B.nxg + (i - B.lo[0] + NG);
}

// plot variables of every valid cell; out: 0-ghost layout with NPLOT comps
__global__ __launch_bounds__(256) void k_plot_vars(const BoxDesc* __restrict__ boxes, const int4* __restrict__ tiles,
                                                   const double* __restrict__ S, const double* __restrict__ ebgeo,
                                                   double gamma, double* __restrict__ out)
{
#pragma clang fp contract(off)
    const int4 t = tiles[blockIdx.x];
    const BoxDesc B = boxes[t.x];
    const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1;
    for (int c = threadIdx.x; c < t.z; c += 256) {
        const int cc = t.y + c;
        const int i = B.lo[0] + cc % nx, j = B.lo[1] + (cc / nx) % ny, k = B.lo[2] + cc / (nx * ny);
        const long s = sidx(B, i, j, k), st = B.nptg;
        const double r = S[s], mx = S[s + st], my = S[s + 2 * st], mz = S[s + 3 * st], E = S[s + 4 * st];
        double vf = 1.0;
        if (B.type == 2) vf = 0.0;
        else if (B.type == 1) vf = ebgeo[B.off_e * 4 + gl_index(B, i, j, k)];
        const double v[NPLOT] = {r, mx, my, mz, E, 0.0, mx / r, my / r, mz / r,
                                 (gamma - 1.) * (E - 0.5 / r * (mx * mx + my * my + mz * mz)), vf};
        const long d0 = B.off_d / NC * NPLOT + (didx(B, i, j, k) - B.off_d);
#pragma unroll
        for (int n = 0; n < NPLOT; ++n) out[d0 + n * B.nptv] = v[n];
    }
}

} // namespace ocfd

using namespace ocfd;

namespace {

const char* kPlotNames[OCFD_PLOT_NVARS] = {"rho", "xmom", "ymom", "zmom", "Eden", "Cost",
                                           "ux",  "uy",   "uz",   "pressure", "vfrac"};
const char* kRealDesc = "(8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1))";

std::string box_str(const int* lo, const int* hi)
{
    std::ostringstream o;
    o << "((" << lo[0] << ',' << lo[1] << ',' << lo[2] << ") (" << hi[0] << ',' << hi[1] << ',' << hi[2]
      << ") (0,0,0))";
    return o.str();
}

bool make_dir(const std::string& p)
{
    if (mkdir(p.c_str(), 0755) == 0 || errno == EEXIST) return true;
    return false;
}

int fail(const std::string& m)
{
    set_error(m);
    return OCFD_EINVAL;
}

} // namespace

namespace {

// A MultiFab on disk in VisMF form: <dir>/<name>_D_00000 holds every FAB (a
// one-line FAB header, then the raw doubles, component-major, i fastest) and
// <dir>/<name>_H the header (version, how, ncomp, nghost, BoxArray, FabOnDisk
// entries, per-FAB minima and maxima).  h: the level's 0-ghost layout with
// ncomp components.
int write_vismf(const std::string& dir, const std::string& name, const LevelDev& L, const double* h, int ncomp)
{
    const std::string dname = name + "_D_00000";
    std::ofstream df(dir + "/" + dname, std::ios::binary);
    if (!df) return fail("cannot write " + dir + "/" + dname);
    std::vector<long> offs;
    std::vector<double> mn, mx;
    for (const BoxDesc& B : L.boxes) {
        offs.push_back((long)df.tellp());
        const std::string hdr = std::string("FAB (") + kRealDesc + ")" + box_str(B.lo, B.hi) + " " +
                                std::to_string(ncomp) + "\n";
        df.write(hdr.data(), (std::streamsize)hdr.size());
        const double* p = h + B.off_d / NC * ncomp;
        df.write((const char*)p, (std::streamsize)(sizeof(double) * ncomp * B.nptv));
        for (int n = 0; n < ncomp; ++n) {
            const double* q = p + n * B.nptv;
            mn.push_back(*std::min_element(q, q + B.nptv));
            mx.push_back(*std::max_element(q, q + B.nptv));
        }
    }
    if (!df) return fail("write error in " + dir + "/" + dname);
    df.close();
    std::ofstream hf(dir + "/" + name + "_H");
    if (!hf) return fail("cannot write " + dir + "/" + name + "_H");
    hf << std::setprecision(17);
    hf << 1 << '\n' << 1 << '\n' << ncomp << '\n' << 0 << '\n';
    hf << '(' << L.nbox << ' ' << 0 << '\n';
    for (const BoxDesc& B : L.boxes) hf << box_str(B.lo, B.hi) << '\n';
    hf << ')' << '\n';
    hf << L.nbox << '\n';
    for (int b = 0; b < L.nbox; ++b) hf << "FabOnDisk: " << dname << ' ' << offs[b] << '\n';
    hf << '\n';
    for (const std::vector<double>* v : {&mn, &mx}) {
        hf << L.nbox << ',' << ncomp << '\n';
        for (int b = 0; b < L.nbox; ++b) {
            for (int n = 0; n < ncomp; ++n) hf << (*v)[(size_t)b * ncomp + n] << ',';
            hf << '\n';
        }
        hf << '\n';
    }
    if (!hf) return fail("write error in " + dir + "/" + name + "_H");
    return OCFD_OK;
}

} // namespace

extern "C" int ocfd_plot_vars(ocfd_level lev, const double* S, double* out, const ocfd_parm* parm, void* stream)
{
    if (!lev || !S || !out || !parm) return fail("ocfd_plot_vars: invalid arguments");
    const LevelDev& L = lev->dev;
    if (L.ntile_c)
        hipLaunchKernelGGL(k_plot_vars, dim3(L.ntile_c), dim3(256), 0, (hipStream_t)stream, L.d_boxes, L.d_tile_c, S,
                           L.d_ebgeo, parm->eos_gamma, out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string("ocfd_plot_vars: ") + hipGetErrorString(e));
        return OCFD_EHIP;
    }
    return OCFD_OK;
}

extern "C" int ocfd_plotfile_write(const char* dir, int nlev, const ocfd_level* levels, const double* const* S,
                                   double time, const int* level_steps, const ocfd_parm* parm, void* stream)
{
    if (!dir || nlev < 1 || !levels || !S || !level_steps || !parm) return fail("ocfd_plotfile_write: invalid arguments");
    for (int l = 0; l < nlev; ++l) {
        if (!levels[l] || !S[l]) return fail("ocfd_plotfile_write: null level or state");
        if (l > 0) {
            const GeomDesc& C = levels[l - 1]->dev.geom;
            const GeomDesc& F = levels[l]->dev.geom;
            for (int d = 0; d < 3; ++d)
                if (F.dom_lo[d] != 2 * C.dom_lo[d] || F.dom_hi[d] != 2 * C.dom_hi[d] + 1)
                    return fail("ocfd_plotfile_write: each level must refine the one below by 2");
        }
    }
    const std::string root(dir);
    if (!make_dir(root)) return fail("ocfd_plotfile_write: cannot create " + root);
    hipStream_t st = (hipStream_t)stream;
    for (int l = 0; l < nlev; ++l) {
        const LevelDev& L = levels[l]->dev;
        const std::string ldir = root + "/Level_" + std::to_string(l);
        if (!make_dir(ldir)) return fail("ocfd_plotfile_write: cannot create " + ldir);
        const long ncell = L.mf_valid / NC;
        double* d_out = nullptr;
        if (hipMalloc((void**)&d_out, sizeof(double) * NPLOT * ncell) != hipSuccess) {
            set_error("ocfd_plotfile_write: out of device memory");
            return OCFD_ENOMEM;
        }
        int rc = ocfd_plot_vars(levels[l], S[l], d_out, parm, stream);
        std::vector<double> h((size_t)NPLOT * ncell);
        if (rc == OCFD_OK &&
            (hipMemcpyAsync(h.data(), d_out, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st) != hipSuccess ||
             hipStreamSynchronize(st) != hipSuccess)) {
            set_error("ocfd_plotfile_write: copy to host failed");
            rc = OCFD_EHIP;
        }
        hipFree(d_out);
        if (rc) return rc;
        if ((rc = write_vismf(ldir, "Cell", L, h.data(), NPLOT))) return rc;
    }
    // plotfile header
    std::ofstream H(root + "/Header");
    if (!H) return fail("ocfd_plotfile_write: cannot write " + root + "/Header");
    H << std::setprecision(17);
    const GeomDesc& G0 = levels[0]->dev.geom;
    H << "HyperCLaw-V1.1" << '\n' << NPLOT << '\n';
    for (const char* nm : kPlotNames) H << nm << '\n';
    H << 3 << '\n' << time << '\n' << nlev - 1 << '\n';
    for (int d = 0; d < 3; ++d) H << G0.prob_lo[d] << ' ';
    H << '\n';
    for (int d = 0; d < 3; ++d) H << G0.prob_lo[d] + (G0.dom_hi[d] - G0.dom_lo[d] + 1) * G0.dx[d] << ' ';
    H << '\n';
    for (int l = 0; l < nlev - 1; ++l) H << 2 << ' ';
    H << '\n';
    for (int l = 0; l < nlev; ++l) H << box_str(levels[l]->dev.geom.dom_lo, levels[l]->dev.geom.dom_hi) << ' ';
    H << '\n';
    for (int l = 0; l < nlev; ++l) H << level_steps[l] << ' ';
    H << '\n';
    for (int l = 0; l < nlev; ++l) {
        for (int d = 0; d < 3; ++d) H << levels[l]->dev.geom.dx[d] << ' ';
        H << '\n';
    }
    H << 0 << '\n' << 0 << '\n';
    for (int l = 0; l < nlev; ++l) {
        const LevelDev& L = levels[l]->dev;
        H << l << ' ' << L.nbox << ' ' << time << '\n' << level_steps[l] << '\n';
        for (const BoxDesc& B : L.boxes)
            for (int d = 0; d < 3; ++d)
                H << L.geom.prob_lo[d] + B.lo[d] * L.geom.dx[d] << ' '
                  << L.geom.prob_lo[d] + (B.hi[d] + 1) * L.geom.dx[d] << '\n';
        H << "Level_" << l << "/Cell" << '\n';
    }
    if (!H) return fail("ocfd_plotfile_write: write error in " + root + "/Header");
    return OCFD_OK;
}

// ---------------------------------------------------------------------------
// reading and comparing plotfiles (fcompare)
// ---------------------------------------------------------------------------
namespace {

struct PBox {
    int lo[3], hi[3];
    long npts() const { return (long)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1); }
    bool operator==(const PBox& o) const
    {
        return std::equal(lo, lo + 3, o.lo) && std::equal(hi, hi + 3, o.hi);
    }
};

struct PLevel {
    int dom_lo[3], dom_hi[3];
    double dx[3];
    std::vector<PBox> boxes;
    std::vector<std::vector<double>> fabs;   // ncomp * npts each
};

struct Plotfile {
    std::vector<std::string> names;
    int dim = 0, finest = 0;
    double time = 0.0;
    std::vector<PLevel> levels;
};

bool parse_box(std::istream& is, PBox& b)
{
    // ((l0,l1,l2) (h0,h1,h2) (t0,t1,t2))
    char c;
    int t[3];
    if (!(is >> c) || c != '(') return false;
    if (!(is >> c) || c != '(') return false;
    if (!(is >> b.lo[0] >> c >> b.lo[1] >> c >> b.lo[2] >> c)) return false;
    if (!(is >> c) || c != '(') return false;
    if (!(is >> b.hi[0] >> c >> b.hi[1] >> c >> b.hi[2] >> c)) return false;
    if (!(is >> c) || c != '(') return false;
    if (!(is >> t[0] >> c >> t[1] >> c >> t[2] >> c)) return false;
    if (!(is >> c) || c != ')') return false;
    return true;
}

int read_plotfile(const std::string& dir, Plotfile& pf, std::string& err)
{
    std::ifstream H(dir + "/Header");
    if (!H) { err = "cannot open " + dir + "/Header"; return 1; }
    std::string line;
    std::getline(H, line);
    if (line != "HyperCLaw-V1.1") { err = dir + ": not a HyperCLaw-V1.1 plotfile"; return 1; }
    int nvar;
    H >> nvar;
    std::getline(H, line);
    pf.names.resize(nvar);
    for (int n = 0; n < nvar; ++n) std::getline(H, pf.names[n]);
    H >> pf.dim >> pf.time >> pf.finest;
    if (pf.dim != 3) { err = dir + ": only 3D plotfiles are read"; return 1; }
    double plo[3], phi[3];
    for (double& v : plo) H >> v;
    for (double& v : phi) H >> v;
    std::vector<int> rr(pf.finest);
    for (int& v : rr) H >> v;
    pf.levels.resize(pf.finest + 1);
    for (PLevel& L : pf.levels) {
        PBox d;
        if (!parse_box(H, d)) { err = dir + ": bad domain box"; return 1; }
        std::copy(d.lo, d.lo + 3, L.dom_lo);
        std::copy(d.hi, d.hi + 3, L.dom_hi);
    }
    std::vector<int> steps(pf.finest + 1);
    for (int& v : steps) H >> v;
    for (PLevel& L : pf.levels)
        for (double& v : L.dx) H >> v;
    int coord, bndry;
    H >> coord >> bndry;
    for (int l = 0; l <= pf.finest; ++l) {
        int lev, ngrid, st;
        double t;
        H >> lev >> ngrid >> t >> st;
        for (int g = 0; g < ngrid * 3; ++g) {
            double a, b;
            H >> a >> b;
        }
        std::string path;
        H >> path;
        if (!H) { err = dir + ": truncated Header"; return 1; }
        // VisMF header of the level
        const std::string base = dir + "/" + path;
        std::ifstream C(base + "_H");
        if (!C) { err = "cannot open " + base + "_H"; return 1; }
        int vers, how, nc, ng, nb, hash;
        char c;
        C >> vers >> how >> nc >> ng >> c >> nb >> hash;
        if (nc != nvar) { err = base + "_H: component count differs from the Header"; return 1; }
        PLevel& L = pf.levels[l];
        L.boxes.resize(nb);
        for (PBox& b : L.boxes)
            if (!parse_box(C, b)) { err = base + "_H: bad box"; return 1; }
        C >> c;
        int nfod;
        C >> nfod;
        std::vector<std::string> files(nfod);
        std::vector<long> offs(nfod);
        for (int f = 0; f < nfod; ++f) {
            std::string tag;
            C >> tag >> files[f] >> offs[f];
        }
        if (!C || nfod != nb) { err = base + "_H: bad FabOnDisk list"; return 1; }
        const std::string ldir = base.substr(0, base.rfind('/'));
        L.fabs.resize(nb);
        for (int b = 0; b < nb; ++b) {
            std::ifstream D(ldir + "/" + files[b], std::ios::binary);
            if (!D) { err = "cannot open " + ldir + "/" + files[b]; return 1; }
            D.seekg(offs[b]);
            std::string fh;
            std::getline(D, fh);
            if (fh.compare(0, 5, "FAB (") != 0 || fh.find(kRealDesc) == std::string::npos) {
                err = files[b] + ": FAB header is not little-endian IEEE double";
                return 1;
            }
            L.fabs[b].resize((size_t)nc * L.boxes[b].npts());
            D.read((char*)L.fabs[b].data(), (std::streamsize)(sizeof(double) * L.fabs[b].size()));
            if (!D) { err = files[b] + ": truncated FAB data"; return 1; }
        }
    }
    return 0;
}

// dense copy of component n of level L over its domain (NaN where no FAB)
std::vector<double> dense(const PLevel& L, int n, int ncomp, std::vector<char>& cov)
{
    const long nx = L.dom_hi[0] - L.dom_lo[0] + 1, ny = L.dom_hi[1] - L.dom_lo[1] + 1,
               nz = L.dom_hi[2] - L.dom_lo[2] + 1;
    std::vector<double> out((size_t)(nx * ny * nz), std::numeric_limits<double>::quiet_NaN());
    cov.assign(out.size(), 0);
    for (size_t b = 0; b < L.boxes.size(); ++b) {
        const PBox& B = L.boxes[b];
        const double* p = L.fabs[b].data() + (size_t)n * B.npts();
        long q = 0;
        for (int k = B.lo[2]; k <= B.hi[2]; ++k)
            for (int j = B.lo[1]; j <= B.hi[1]; ++j)
                for (int i = B.lo[0]; i <= B.hi[0]; ++i, ++q) {
                    const size_t c = ((size_t)(k - L.dom_lo[2]) * ny + (j - L.dom_lo[1])) * nx + (i - L.dom_lo[0]);
                    out[c] = p[q];
                    cov[c] = 1;
                }
    }
    (void)ncomp;
    return out;
}

} // namespace

extern "C" int ocfd_plotfile_compare(const char* file_a, const char* file_b, int norm, double rtol, double atol,
                                     int allow_diff_grids, int* result, char* report, long report_cap)
{
    if (!file_a || !file_b || !result || norm < 0 || norm > 2) return fail("ocfd_plotfile_compare: invalid arguments");
    Plotfile A, B;
    std::string err;
    if (read_plotfile(file_a, A, err) || read_plotfile(file_b, B, err)) return fail("ocfd_plotfile_compare: " + err);
    if (A.finest != B.finest) return fail("ocfd_plotfile_compare: number of levels do not match");
    for (size_t l = 0; l < A.levels.size(); ++l)
        for (int d = 0; d < 3; ++d)
            if (A.levels[l].dx[d] != B.levels[l].dx[d]) return fail("ocfd_plotfile_compare: grid dx does not match");
    std::ostringstream o;
    const int nA = (int)A.names.size(), nB = (int)B.names.size();
    if (nA != nB) o << "\n WARNING: number of variables do not match\n";
    std::vector<int> ivar_b(nA, -1);
    bool all_found = true;
    for (int a = 0; a < nA; ++a) {
        auto r = std::find(B.names.begin(), B.names.end(), A.names[a]);
        if (r == B.names.end()) {
            o << " WARNING: variable " << A.names[a] << " not found in plotfile 2\n";
            all_found = false;
        } else {
            ivar_b[a] = (int)(r - B.names.begin());
        }
    }
    for (int b = 0; b < nB; ++b)
        if (std::find(A.names.begin(), A.names.end(), B.names[b]) == A.names.end()) {
            o << " WARNING: variable " << B.names[b] << " not found in plotfile 1\n";
            all_found = false;
        }
    o << "\n " << std::setw(24) << std::right << "variable name" << "  " << std::setw(24) << "absolute error" << "  "
      << std::setw(24) << "relative error" << "\n " << std::setw(24) << " " << "  " << std::setw(24) << "(||A - B||)"
      << "  " << std::setw(24) << "(||A - B||/||A||)" << "\n " << std::string(76, '-') << "\n";
    double global_error = 0.0;
    bool any_nans = false, all_passed = true;
    for (size_t l = 0; l < A.levels.size(); ++l) {
        const PLevel& La = A.levels[l];
        const PLevel& Lb = B.levels[l];
        if (La.boxes.empty() && Lb.boxes.empty()) continue;
        const bool grids_match = La.boxes.size() == Lb.boxes.size() && std::equal(La.boxes.begin(), La.boxes.end(), Lb.boxes.begin());
        if (!grids_match && !allow_diff_grids) return fail("ocfd_plotfile_compare: grids do not match");
        double dv = 1.0;
        for (int d = 0; d < 3; ++d) dv *= La.dx[d];
        o << " level = " << l << "\n";
        std::vector<double> aerr(nA, 0.0), rerr(nA, 0.0);
        std::vector<int> nan_a(nA, 0), nan_b(nA, 0);
        for (int a = 0; a < nA; ++a) {
            if (ivar_b[a] < 0) continue;
            std::vector<char> ca, cb;
            const std::vector<double> da = dense(La, a, nA, ca), db = dense(Lb, ivar_b[a], nB, cb);
            if (ca != cb) return fail("ocfd_plotfile_compare: grids do not cover same domain");
            double mx = 0.0, s1 = 0.0, s2 = 0.0, amx = 0.0, a1 = 0.0, a2 = 0.0;
            for (size_t c = 0; c < da.size(); ++c) {
                if (!ca[c]) continue;
                if (std::isnan(da[c])) nan_a[a] = 1;
                if (std::isnan(db[c])) nan_b[a] = 1;
                const double e = std::fabs(db[c] - da[c]);
                mx = std::max(mx, e);
                s1 += e;
                s2 += e * e;
                amx = std::max(amx, std::fabs(da[c]));
                a1 += std::fabs(da[c]);
                a2 += da[c] * da[c];
            }
            if (norm == 0) {
                aerr[a] = mx;
                rerr[a] = mx / amx;
            } else if (norm == 1) {
                aerr[a] = s1 * dv;
                rerr[a] = s1 / a1;
            } else {
                aerr[a] = std::sqrt(s2) * std::sqrt(dv);
                rerr[a] = std::sqrt(s2) / std::sqrt(a2);
            }
        }
        for (int a = 0; a < nA; ++a) {
            if (ivar_b[a] < 0) {
                o << " " << std::setw(24) << std::left << A.names[a] << "  " << std::setw(50)
                  << "< variable not present in both files > \n";
            } else if (nan_a[a] && nan_b[a]) {
                o << " " << std::setw(24) << std::left << A.names[a] << "  " << std::setw(50)
                  << "< NaN present in both A and B > \n";
            } else if (nan_a[a]) {
                o << " " << std::setw(24) << std::left << A.names[a] << "  " << std::setw(50)
                  << "< NaN present in A > \n";
            } else if (nan_b[a]) {
                o << " " << std::setw(24) << std::left << A.names[a] << "  " << std::setw(50)
                  << "< NaN present in B > \n";
            } else {
                const double lo = std::numeric_limits<double>::min(), hi = std::numeric_limits<double>::max();
                const double ae = aerr[a] > 0. ? std::min(std::max(aerr[a], lo), hi) : 0.;
                const double re = rerr[a] > 0. ? std::min(std::max(rerr[a], lo), hi) : 0.;
                o << " " << std::setw(24) << std::left << A.names[a] << std::right << "  " << std::setw(24)
                  << std::setprecision(10) << ae << "  " << std::setw(24) << std::setprecision(10) << re << "\n";
            }
            global_error = std::max(global_error, aerr[a]);
            any_nans = any_nans || nan_a[a] || nan_b[a];
            all_passed = all_passed && (aerr[a] <= atol || rerr[a] <= rtol);
        }
    }
    if (!all_found) o << " WARNING: not all variables present in both files\n";
    if (any_nans) {
        *result = -1;
    } else if (global_error == 0.0) {
        o << " PLOTFILE AGREE\n";
        *result = 1;
    } else if (all_passed) {
        o << " PLOTFILE AGREE to specified tolerances: absolute = " << atol << " relative = " << rtol << "\n";
        *result = 2;
    } else {
        *result = 0;
    }
    if (report && report_cap > 0) {
        const std::string s = o.str();
        const size_t n = std::min((size_t)(report_cap - 1), s.size());
        std::memcpy(report, s.data(), n);
        report[n] = '\0';
    }
    return OCFD_OK;
}

// ---------------------------------------------------------------------------
// checkpoints: a level's MultiFab (valid cells) in VisMF form
// ---------------------------------------------------------------------------
extern "C" int ocfd_checkpoint_write_level(const char* dir, const char* name, ocfd_level lev, const double* S,
                                           int ncomp, void* stream)
{
    if (!dir || !name || !lev || !S || ncomp <= 0) return fail("ocfd_checkpoint_write_level: invalid arguments");
    const LevelDev& L = lev->dev;
    if (!make_dir(dir)) return fail(std::string("ocfd_checkpoint_write_level: cannot create ") + dir);
    const long ngrown = L.mf_state / NC, nvalid = L.mf_valid / NC;
    std::vector<double> g((size_t)ncomp * ngrown), v((size_t)ncomp * nvalid);
    hipStream_t st = (hipStream_t)stream;
    if (hipMemcpyAsync(g.data(), S, sizeof(double) * g.size(), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        set_error("ocfd_checkpoint_write_level: copy to host failed");
        return OCFD_EHIP;
    }
    for (const BoxDesc& B : L.boxes) {
        const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1, nz = B.hi[2] - B.lo[2] + 1;
        const double* gs = g.data() + B.off_s / NC * ncomp;
        double* vs = v.data() + B.off_d / NC * ncomp;
        for (int n = 0; n < ncomp; ++n)
            for (int k = 0; k < nz; ++k)
                for (int j = 0; j < ny; ++j)
                    std::memcpy(vs + n * B.nptv + ((long)k * ny + j) * nx,
                                gs + n * B.nptg + ((long)(k + NG) * B.nyg + (j + NG)) * B.nxg + NG,
                                sizeof(double) * nx);
    }
    return write_vismf(dir, name, L, v.data(), ncomp);
}

extern "C" int ocfd_checkpoint_read_level(const char* dir, const char* name, ocfd_level lev, double* S, int ncomp,
                                          void* stream)
{
    if (!dir || !name || !lev || !S || ncomp <= 0) return fail("ocfd_checkpoint_read_level: invalid arguments");
    tcache_clear(lev->dev);   // new data in the level's states
    const LevelDev& L = lev->dev;
    const std::string base = std::string(dir) + "/" + name;
    std::ifstream C(base + "_H");
    if (!C) return fail("ocfd_checkpoint_read_level: cannot open " + base + "_H");
    int vers, how, nc, ng, nb, hash;
    char c;
    C >> vers >> how >> nc >> ng >> c >> nb >> hash;
    if (!C || nc != ncomp || nb != L.nbox) return fail("ocfd_checkpoint_read_level: " + base + "_H does not match the level");
    for (int b = 0; b < nb; ++b) {
        PBox x;
        if (!parse_box(C, x) || !std::equal(x.lo, x.lo + 3, L.boxes[b].lo) || !std::equal(x.hi, x.hi + 3, L.boxes[b].hi))
            return fail("ocfd_checkpoint_read_level: box " + std::to_string(b) + " differs from the level's");
    }
    C >> c;
    int nfod;
    C >> nfod;
    std::vector<std::string> files(nfod);
    std::vector<long> offs(nfod);
    for (int f = 0; f < nfod; ++f) {
        std::string tag;
        C >> tag >> files[f] >> offs[f];
    }
    if (!C || nfod != nb) return fail("ocfd_checkpoint_read_level: bad FabOnDisk list in " + base + "_H");
    const long ngrown = L.mf_state / NC;
    std::vector<double> g((size_t)ncomp * ngrown, 0.0);
    for (int b = 0; b < nb; ++b) {
        const BoxDesc& B = L.boxes[b];
        std::ifstream D(std::string(dir) + "/" + files[b], std::ios::binary);
        if (!D) return fail("ocfd_checkpoint_read_level: cannot open " + files[b]);
        D.seekg(offs[b]);
        std::string fh;
        std::getline(D, fh);
        if (fh.compare(0, 5, "FAB (") != 0 || fh.find(kRealDesc) == std::string::npos)
            return fail("ocfd_checkpoint_read_level: " + files[b] + " is not little-endian IEEE double");
        const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1, nz = B.hi[2] - B.lo[2] + 1;
        std::vector<double> v((size_t)ncomp * B.nptv);
        D.read((char*)v.data(), (std::streamsize)(sizeof(double) * v.size()));
        if (!D) return fail("ocfd_checkpoint_read_level: truncated data in " + files[b]);
        double* gs = g.data() + B.off_s / NC * ncomp;
        for (int n = 0; n < ncomp; ++n)
            for (int k = 0; k < nz; ++k)
                for (int j = 0; j < ny; ++j)
                    std::memcpy(gs + n * B.nptg + ((long)(k + NG) * B.nyg + (j + NG)) * B.nxg + NG,
                                v.data() + n * B.nptv + ((long)k * ny + j) * nx, sizeof(double) * nx);
    }
    hipStream_t st = (hipStream_t)stream;
    if (hipMemcpyAsync(S, g.data(), sizeof(double) * g.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        set_error("ocfd_checkpoint_read_level: copy to the device failed");
        return OCFD_EHIP;
    }
    return OCFD_OK;
}
