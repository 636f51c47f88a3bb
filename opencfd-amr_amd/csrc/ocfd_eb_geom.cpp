// Cut-cell geometry of z-extruded bodies: the build's own replacement for
// AMReX EB2::Build (Source/Geometry.cpp:9-29), producing the fields the EB box
// path reads (Source/EB/AdvanceBox.cpp:15-31): cell flags with neighbour
// connectivity, volume fractions, face area fractions, face and boundary
// centroids.
//
// The body is given by an implicit function phi(x, y) (> 0 inside the body),
// the same for every z.  The boundary inside a cell is the straight segment
// between the roots of phi on the cell's edges (the piecewise-linear boundary
// EB2 also uses), so the cut cell is the fluid polygon of the cell square.
// Host code, run once per level set-up.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "ocfd.h"

namespace {

constexpr uint32_t T_REGULAR = 0, T_SINGLE = 1, T_COVERED = 3;

// EBCellFlag layout of AMReX: bits 0-1 cell type, bits 2-4 number of volumes,
// bit 5 + (i+1) + 3(j+1) + 9(k+1) neighbour (i,j,k) connected.
inline uint32_t ngbr_bit(int i, int j, int k) { return 1u << (5 + (i + 1) + 3 * (j + 1) + 9 * (k + 1)); }

struct Body {
    ocfd_eb_spec s;
    double phi(double x, double y) const
    {
        double v = -1.0;
        switch (s.kind) {
        case OCFD_EB_CYLINDER: v = s.radius - std::hypot(x - s.center[0], y - s.center[1]); break;
        case OCFD_EB_BOX:
            v = std::min(std::min(x - s.lo[0], s.hi[0] - x), std::min(y - s.lo[1], s.hi[1] - y));
            break;
        case OCFD_EB_PLANE: v = (x - s.point[0]) * s.normal[0] + (y - s.point[1]) * s.normal[1]; break;
        default: v = -1.0;
        }
        return s.fluid_inside ? -v : v;
    }
    bool body(double x, double y) const { return phi(x, y) > 0.0; }
    // root of phi on the segment a->b (one endpoint body, one fluid), by bisection
    void root(double ax, double ay, double bx, double by, double& rx, double& ry) const
    {
        const bool ba = body(ax, ay);
        double t0 = 0.0, t1 = 1.0;
        for (int it = 0; it < 80; ++it) {
            const double tm = 0.5 * (t0 + t1);
            const bool bm = body(ax + tm * (bx - ax), ay + tm * (by - ay));
            if (bm == ba) t0 = tm; else t1 = tm;
        }
        const double t = 0.5 * (t0 + t1);
        rx = ax + t * (bx - ax);
        ry = ay + t * (by - ay);
    }
};

struct Grid2 {
    int ilo, jlo, ni, nj;   // cell range [ilo, ilo+ni) x [jlo, jlo+nj)
    template <class T>
    T& at(std::vector<T>& v, int i, int j, int stride_i) const { return v[(size_t)(j - jlo) * stride_i + (i - ilo)]; }
};

} // namespace

extern "C" int ocfd_eb_geometry(const ocfd_eb_spec* spec, const ocfd_geom* g, const int* lo, const int* hi, int ng,
                                unsigned* flag, double* vfrac, double* apx, double* apy, double* apz, double* fcx,
                                double* fcy, double* fcz, double* bcent)
{
    if (!spec || !g || !lo || !hi || ng < 0) return OCFD_EINVAL;
    Body B{*spec};
    const double small = spec->small_volfrac > 0.0 ? spec->small_volfrac : 1e-14;
    const double dx = g->dx[0], dy = g->dx[1];
    const double x0 = g->prob_lo[0], y0 = g->prob_lo[1];
    // output range and an internal range two cells wider (neighbours of the
    // outermost cells decide their faces and connectivity)
    const int oil = lo[0] - ng, ojl = lo[1] - ng, oih = hi[0] + ng, ojh = hi[1] + ng;
    const int il = oil - 2, jl = ojl - 2, ih = oih + 2, jh = ojh + 2;
    const int ni = ih - il + 1, nj = jh - jl + 1;
    auto X = [&](int i) { return x0 + i * dx; };
    auto Y = [&](int j) { return y0 + j * dy; };

    // nodes
    std::vector<char> nb((size_t)(ni + 1) * (nj + 1));
    auto node = [&](int i, int j) -> char& { return nb[(size_t)(j - jl) * (ni + 1) + (i - il)]; };
    for (int j = jl; j <= jh + 1; ++j)
        for (int i = il; i <= ih + 1; ++i) node(i, j) = B.body(X(i), Y(j));

    // faces: x face (i, j) is the segment x = X(i), y in [Y(j), Y(j+1)]
    std::vector<double> ax((size_t)(ni + 1) * nj), cx((size_t)(ni + 1) * nj);
    std::vector<double> ay((size_t)ni * (nj + 1)), cy((size_t)ni * (nj + 1));
    auto AX = [&](int i, int j) -> double& { return ax[(size_t)(j - jl) * (ni + 1) + (i - il)]; };
    auto CX = [&](int i, int j) -> double& { return cx[(size_t)(j - jl) * (ni + 1) + (i - il)]; };
    auto AY = [&](int i, int j) -> double& { return ay[(size_t)(j - jl) * ni + (i - il)]; };
    auto CY = [&](int i, int j) -> double& { return cy[(size_t)(j - jl) * ni + (i - il)]; };
    for (int j = jl; j <= jh; ++j)
        for (int i = il; i <= ih + 1; ++i) {
            const bool b0 = node(i, j), b1 = node(i, j + 1);
            if (!b0 && !b1) { AX(i, j) = 1.0; CX(i, j) = 0.0; }
            else if (b0 && b1) { AX(i, j) = 0.0; CX(i, j) = 0.0; }
            else {
                double rx, ry;
                B.root(X(i), Y(j), X(i), Y(j + 1), rx, ry);
                const double a = b0 ? Y(j + 1) - ry : ry - Y(j);
                const double mid = b0 ? 0.5 * (ry + Y(j + 1)) : 0.5 * (Y(j) + ry);
                AX(i, j) = a / dy;
                CX(i, j) = (mid - (Y(j) + 0.5 * dy)) / dy;
            }
        }
    for (int j = jl; j <= jh + 1; ++j)
        for (int i = il; i <= ih; ++i) {
            const bool b0 = node(i, j), b1 = node(i + 1, j);
            if (!b0 && !b1) { AY(i, j) = 1.0; CY(i, j) = 0.0; }
            else if (b0 && b1) { AY(i, j) = 0.0; CY(i, j) = 0.0; }
            else {
                double rx, ry;
                B.root(X(i), Y(j), X(i + 1), Y(j), rx, ry);
                const double a = b0 ? X(i + 1) - rx : rx - X(i);
                const double mid = b0 ? 0.5 * (rx + X(i + 1)) : 0.5 * (X(i) + rx);
                AY(i, j) = a / dx;
                CY(i, j) = (mid - (X(i) + 0.5 * dx)) / dx;
            }
        }

    // cells: fluid polygon of the square (corners counter-clockwise, roots on
    // sign-changing edges), shoelace area and centroid
    std::vector<double> vf((size_t)ni * nj), cen((size_t)2 * ni * nj), bc((size_t)2 * ni * nj);
    auto VF = [&](int i, int j) -> double& { return vf[(size_t)(j - jl) * ni + (i - il)]; };
    for (int j = jl; j <= jh; ++j)
        for (int i = il; i <= ih; ++i) {
            const double px[4] = {X(i), X(i + 1), X(i + 1), X(i)};
            const double py[4] = {Y(j), Y(j), Y(j + 1), Y(j + 1)};
            const bool bb[4] = {(bool)node(i, j), (bool)node(i + 1, j), (bool)node(i + 1, j + 1), (bool)node(i, j + 1)};
            double qx[12], qy[12];
            int m = 0;
            double rxs[4], rys[4];
            int nr = 0;
            for (int c = 0; c < 4; ++c) {
                const int e = (c + 1) & 3;
                if (!bb[c]) { qx[m] = px[c]; qy[m] = py[c]; ++m; }
                if (bb[c] != bb[e]) {
                    double rx, ry;
                    if (bb[c]) B.root(px[e], py[e], px[c], py[c], rx, ry);   // root from the fluid end
                    else B.root(px[c], py[c], px[e], py[e], rx, ry);
                    qx[m] = rx; qy[m] = ry; ++m;
                    if (nr < 4) { rxs[nr] = rx; rys[nr] = ry; ++nr; }
                }
            }
            double A = 0.0, Cx = 0.0, Cy = 0.0;
            for (int a = 0; a < m; ++a) {
                const int b = (a + 1) % m;
                const double cr = qx[a] * qy[b] - qx[b] * qy[a];
                A += cr;
                Cx += (qx[a] + qx[b]) * cr;
                Cy += (qy[a] + qy[b]) * cr;
            }
            A *= 0.5;
            const size_t id = (size_t)(j - jl) * ni + (i - il);
            if (m == 0 || A <= 0.0) {
                vf[id] = 0.0; cen[2 * id] = cen[2 * id + 1] = 0.0;
            } else {
                vf[id] = std::min(1.0, A / (dx * dy));
                if (nr == 0) vf[id] = 1.0;
                cen[2 * id] = (Cx / (6.0 * A) - (X(i) + 0.5 * dx)) / dx;
                cen[2 * id + 1] = (Cy / (6.0 * A) - (Y(j) + 0.5 * dy)) / dy;
                if (nr == 0) cen[2 * id] = cen[2 * id + 1] = 0.0;
            }
            if (nr >= 2) {
                bc[2 * id] = (0.5 * (rxs[0] + rxs[1]) - (X(i) + 0.5 * dx)) / dx;
                bc[2 * id + 1] = (0.5 * (rys[0] + rys[1]) - (Y(j) + 0.5 * dy)) / dy;
            } else {
                bc[2 * id] = bc[2 * id + 1] = 0.0;
            }
        }
    // small cells become covered; faces touching a covered cell close
    for (int j = jl; j <= jh; ++j)
        for (int i = il; i <= ih; ++i)
            if (VF(i, j) < small) VF(i, j) = 0.0;
    for (int j = jl; j <= jh; ++j)
        for (int i = il + 1; i <= ih; ++i)
            if (VF(i - 1, j) == 0.0 || VF(i, j) == 0.0) { AX(i, j) = 0.0; CX(i, j) = 0.0; }
    for (int j = jl + 1; j <= jh; ++j)
        for (int i = il; i <= ih; ++i)
            if (VF(i, j - 1) == 0.0 || VF(i, j) == 0.0) { AY(i, j) = 0.0; CY(i, j) = 0.0; }

    // flags
    auto step = [&](int i, int j, int d, int s) -> bool {   // one face step is open
        if (VF(i, j) == 0.0) return false;
        if (d == 0) return (s > 0 ? AX(i + 1, j) : AX(i, j)) > 0.0 && VF(i + s, j) > 0.0;
        return (s > 0 ? AY(i, j + 1) : AY(i, j)) > 0.0 && VF(i, j + s) > 0.0;
    };
    auto cellflag = [&](int i, int j) -> uint32_t {
        if (VF(i, j) == 0.0) return T_COVERED;
        const bool reg = VF(i, j) == 1.0 && AX(i, j) == 1.0 && AX(i + 1, j) == 1.0 && AY(i, j) == 1.0 &&
                         AY(i, j + 1) == 1.0;
        uint32_t f = (reg ? T_REGULAR : T_SINGLE) | (1u << 2);
        for (int b = -1; b <= 1; ++b)
            for (int a = -1; a <= 1; ++a) {
                bool c;
                if (a == 0 && b == 0) c = true;
                else if (b == 0) c = step(i, j, 0, a);
                else if (a == 0) c = step(i, j, 1, b);
                else c = (step(i, j, 0, a) && step(i + a, j, 1, b)) || (step(i, j, 1, b) && step(i, j + b, 0, a));
                if (c)   // z steps stay in the column: open for every non-covered cell
                    for (int kk = -1; kk <= 1; ++kk) f |= ngbr_bit(a, b, kk);
            }
        return f;
    };

    const int nxg = oih - oil + 1, nyg = ojh - ojl + 1, nzg = hi[2] - lo[2] + 1 + 2 * ng;
    const long npt = (long)nxg * nyg * nzg;
    for (int kk = 0; kk < nzg; ++kk)
        for (int j = ojl; j <= ojh; ++j)
            for (int i = oil; i <= oih; ++i) {
                const long o = ((long)kk * nyg + (j - ojl)) * nxg + (i - oil);
                const size_t id = (size_t)(j - jl) * ni + (i - il);
                if (flag) flag[o] = cellflag(i, j);
                if (vfrac) vfrac[o] = VF(i, j);
                if (apx) apx[o] = AX(i, j);
                if (apy) apy[o] = AY(i, j);
                if (apz) apz[o] = VF(i, j);
                if (fcx) { fcx[o] = CX(i, j); fcx[o + npt] = 0.0; }
                if (fcy) { fcy[o] = CY(i, j); fcy[o + npt] = 0.0; }
                if (fcz) { fcz[o] = VF(i, j) > 0.0 ? cen[2 * id] : 0.0; fcz[o + npt] = VF(i, j) > 0.0 ? cen[2 * id + 1] : 0.0; }
                if (bcent) {
                    const bool cut = VF(i, j) > 0.0 && (cellflag(i, j) & 3u) == T_SINGLE;
                    bcent[o] = cut ? bc[2 * id] : 0.0;
                    bcent[o + npt] = cut ? bc[2 * id + 1] : 0.0;
                    bcent[o + 2 * npt] = 0.0;
                }
            }
    return OCFD_OK;
}
