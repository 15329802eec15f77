// tsw_init DECK [DIR]: reads a deck by the reference's rules and writes the projections Init computes
// (init.cpp:741-860) in the reference's file format and names: sigma_t_dg0_nx<nx>.cs2,
// sigma_s_dg0_nx<nx>.cs2, initial_dg<p>_nx<nx>.sd2, source_dg<p>_nx<nx>.sd2 (source before the inflow is
// added), plus source_af_dg<p>_nx<nx>.af2, the angular source with inflow halos the sweeps take.
#include <cinttypes>
#include <cstdio>
#include <exception>

#include "tsw_host.hpp"

int main(int argc, char **argv) {
    if (argc < 2 || argc > 3) {
        std::fprintf(stderr, "usage: %s DECK [DIR]\n", argv[0]);
        return 2;
    }
    try {
        const std::string dir = argc == 3 ? argv[2] : ".";
        tsw_host::ParameterList params;
        params.ReadFromFile(argv[1]);
        params.Set("output_init", "true");
        const tsw_host::ProblemInputs in = tsw_host::Init(params, dir);
        char nxs[32];
        std::snprintf(nxs, sizeof nxs, "%07" PRId64, in.dd.nx);
        const std::string af = dir + "/source_af_dg" + std::to_string(in.DG_degree) + "_nx" + nxs + ".af2";
        in.source.WriteToDisk(af);
        std::printf("%s: %" PRId64 " x %" PRId64 " cells, DG degree %d, %s order %" PRId64 " (%" PRId64
                    " ordinates)\n",
                    in.problem.c_str(), in.dd.nx, in.dd.ny, in.DG_degree, in.os.type().c_str(), in.os.ang_order(),
                    in.os.nq());
    } catch (const std::exception &e) {
        std::fprintf(stderr, "tsw_init: %s\n", e.what());
        return 1;
    }
    return 0;
}
