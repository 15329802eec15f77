// tsw_compare REF FILE...: errors of each scalar-flux file against REF, as the reference's
// post-processor computes them (main_postproc.cpp:394-436): the difference on the common refined mesh,
// then its L1, L2 and Linf norms.  Files are in the reference's DensityFunction format.  The reference
// also groups files by the order/run/cell keys of their names and reads run times from log files; this
// tool prints one row per file instead.  Exit status 1 on a read or domain error.
#include <cstdio>
#include <exception>

#include "tsw_host.hpp"

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s REF FILE...\n", argv[0]);
        return 2;
    }
    try {
        const tsw_host::ScalarFlux ref = tsw_host::ScalarFlux::ReadFromDisk(argv[1]);
        std::printf("%-40s %14s %14s %14s\n", "file", "L1", "L2", "Linf");
        for (int k = 2; k < argc; ++k) {
            const tsw_host::ScalarFlux approx = tsw_host::ScalarFlux::ReadFromDisk(argv[k]);
            const tsw_host::ScalarFlux diff = tsw_host::ScalarFlux::ComputeDifference(ref, approx);
            std::printf("%-40s %14.6e %14.6e %14.6e\n", argv[k], diff.L1Norm(), diff.L2Norm(), diff.LinfNorm());
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "tsw_compare: %s\n", e.what());
        return 1;
    }
    return 0;
}
