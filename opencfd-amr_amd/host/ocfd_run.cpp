// ocfd_run -- a C++ host driving one level of the per-box advance through
// the ocfd.hpp facade (no Python anywhere on the path): the case set-up of
// Exec/<case>/inputs, EBR::initData, the time loop of Source/main.cpp:61-70
// with computeInitialDt / computeNewDt, and the level's valid cells written
// to a file for inspection.
//
//   ocfd_run CASE NX NY NZ MAX_GRID STEPS SCHEME FLAGS OUT [STOP_TIME]
//
// CASE: sod | dmr | kh.  SCHEME: 0 Euler, 1 RK2, 2 SSP-RK3.  FLAGS: the
// OCFD_* stage flags (1 viscous, 4 FP32 fluxes, ...).  OUT: binary file --
// int32 nbox, per box int32 lo[3], hi[3], then per box the 5 components of
// its valid cells as float64 (component, k, j, i order).  Prints one line
// with the steps taken, the final time and the wall time of the loop.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ocfd.hpp"

int main(int argc, char** argv)
{
    if (argc < 10) {
        std::fprintf(stderr, "usage: %s CASE NX NY NZ MAX_GRID STEPS SCHEME FLAGS OUT [STOP_TIME]\n", argv[0]);
        return 2;
    }
    const std::string cs = argv[1];
    const std::array<int, 3> n{std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4])};
    const int mg = std::atoi(argv[5]), nsteps = std::atoi(argv[6]), scheme = std::atoi(argv[7]),
              flags = std::atoi(argv[8]);
    const char* out = argv[9];
    const double stop_time = argc > 10 ? std::atof(argv[10]) : -1.0;
    try {
        using namespace ocfd;
        Geometry geom;
        int kind = 0;
        if (cs == "sod") {   // Exec/0_SOD/inputs: outflow in x, slip walls in y and z
            geom = Geometry(n, {0.0, 0.0, 0.0}, {1.0, (double)n[1] / n[0], (double)n[2] / n[0]}, {0, 0, 0},
                            {Outflow, SlipWall, SlipWall}, {Outflow, SlipWall, SlipWall});
            kind = 0;
        } else if (cs == "dmr") {   // Exec/1_DMR/inputs with the time-dependent top boundary
            geom = Geometry(n, {0.0, 0.0, 0.0}, {4.0, 4.0 * n[1] / n[0], 4.0 * n[2] / n[0]}, {0, 0, 0},
                            {Inflow, SlipWall, Outflow}, {Outflow, Inflow, Outflow}, 1);
            check(ocfd_dmr_states(geom.g.left_state, geom.g.right_state), "ocfd_dmr_states");
            kind = 1;
        } else if (cs == "kh") {   // periodic unit cube
            geom = Geometry(n, {0.0, 0.0, 0.0}, {1.0, 1.0, 1.0}, {1, 1, 1}, {Interior, Interior, Interior},
                            {Interior, Interior, Interior});
            kind = 2;
        } else {
            std::fprintf(stderr, "unknown case %s\n", cs.c_str());
            return 2;
        }
        const BoxArray ba = BoxArray::chop(n, {mg, mg, mg});
        Level lev(ba, geom);
        Parm parm;
        EBR ebr(lev, parm, flags);
        MultiFab A = ebr.newState(), B = ebr.newState(), C = ebr.newState();
        ebr.initData(kind, A, 42);
        double time = 0.0, dt = 0.0;
        check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
        const auto t0 = std::chrono::steady_clock::now();
        const int steps = ebr.run(A, B, C, 0.5, scheme, stop_time, nsteps, &time, &dt);
        check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        std::vector<double> h;
        A.copyToHost(h);
        FILE* f = std::fopen(out, "wb");
        if (!f) throw Error(OCFD_EINVAL, std::string("cannot open ") + out);
        const int nb = ba.size();
        std::fwrite(&nb, sizeof(int), 1, f);
        for (const Box& b : ba.boxes) {
            std::fwrite(b.lo.data(), sizeof(int), 3, f);
            std::fwrite(b.hi.data(), sizeof(int), 3, f);
        }
        for (int b = 0; b < nb; ++b) {
            const Box& bx = ba[b];
            const Box g = bx.grow(OCFD_NGROW);
            const long off = lev.fabOffset(b, OCFD_NGROW, OCFD_NCONS);
            for (int c = 0; c < OCFD_NCONS; ++c)
                for (int k = bx.lo[2]; k <= bx.hi[2]; ++k)
                    for (int j = bx.lo[1]; j <= bx.hi[1]; ++j) {
                        const long row = off + c * g.numPts() +
                                         ((long)(k - g.lo[2]) * g.length(1) + (j - g.lo[1])) * g.length(0) +
                                         (bx.lo[0] - g.lo[0]);
                        std::fwrite(h.data() + row, sizeof(double), bx.length(0), f);
                    }
        }
        std::fclose(f);
        std::printf("{\"case\": \"%s\", \"steps\": %d, \"time\": %.17g, \"dt\": %.17g, \"ms\": %.3f, \"boxes\": %d}\n",
                    cs.c_str(), steps, time, dt, ms, nb);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ocfd_run: %s\n", e.what());
        return 1;
    }
    return 0;
}
