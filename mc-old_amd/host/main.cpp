// main.cpp — `mc <io_dir> [xs_dir]`: the drop-in for the reference's
// `MC.exe <io_dir>` (Main.cpp:8-31).  Reads <io_dir>/input.xml, runs the
// fixed-source simulation on the GPU and writes <io_dir>/output.h5 in the
// reference's layout (report.cpp:9-50), or output.txt with the same content
// where no HDF5 library loads.
#include <cstdio>
#include <string>

#include "simulator.hpp"

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <io_dir> [xs_dir]\n", argv[0]);
        return 2;
    }
    std::string io_dir = argv[1];
    if (!io_dir.empty() && io_dir.back() != '/') io_dir += '/';
    const std::string xs_dir = argc > 2 ? argv[2] : "./xs_library";
    try {
        std::printf("\n%s\n", mc_version());
        mcold::Simulator S(io_dir, xs_dir);
        std::printf("Simulation: %llu histories\n", S.Nsample);
        S.start();
        S.report(io_dir);
        std::printf("Done: %llu flights in %.3f s of device time; wrote %s\n", S.Ntrack, S.device_seconds,
                    S.report_path.c_str());
        for (const auto& E : S.Estimators)
            if (E->tally_size() > 0)
                std::printf("  %-24s bin 0: %.6g +- %.3g\n", E->name().c_str(), E->tally(0).mean,
                            E->tally(0).uncer);
    } catch (const mcold::Error& e) {
        std::fprintf(stderr, "[ERROR %d] %s\n", (int)e.status, e.what());
        return 1;
    }
    return 0;
}
