// dem_main.cpp -- the reference's driver (main.cpp:65-640) written against
// the engine's C ABI (include/dem_capi.h): the program a maintainer builds
// to run a heyi.in deck.  It reads the deck, allocates one context, feeds
// the particles with the reference's sampling law, and steps to the packing
// time, writing a Tecplot zone (WrtTec, main.cpp:587) and a log record
// (main.cpp:591-618) at each of the deck's outputs and the run's total at
// the end (main.cpp:631-634).
//
//   dem_main <heyi.in> <genfile> [state_out]
//
// writes <genfile>.dat (Tecplot) and <genfile>_Log.dat; with state_out, the
// final positions, velocities and spins as raw doubles [3][total][3].
#include <math.h>
#include <stdio.h>

#include <chrono>
#include <string>
#include <vector>

#include "dem_capi.h"

static int fail(const char *what) {
    fprintf(stderr, "dem_main: %s: %s\n", what, dem_last_error());
    return 1;
}

int main(int argc, char **argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: dem_main <heyi.in> <genfile> [state_out]\n");
        return 2;
    }
    const std::string gen = argv[2], tec = gen + ".dat", log = gen + "_Log.dat";
    dem_input in;
    if (dem_read_deck(argv[1], &in) != DEM_OK) return fail("deck");
    dem_ctx *ctx = nullptr;
    if (dem_create(&in, nullptr, DEM_PREC_FP64, &ctx) != DEM_OK) return fail("create");
    dem_setup S;
    const int NP = in.NP;
    std::vector<double> rad(NP), pos(3 * (size_t)NP, 0.0), vel(3 * (size_t)NP), angv(3 * (size_t)NP),
        force(3 * (size_t)NP), mom(3 * (size_t)NP), rmom(NP);
    std::vector<int> level(NP), nidx(NP);
    if (dem_derive(&in, nullptr, &S, rad.data(), level.data(), nidx.data()) != DEM_OK)
        return fail("derive");
    // ParticleFeed (main.cpp:439-448), seeded
    unsigned long long seed = 1;
    int total = 0;
    while (total < NP) {
        const int n = dem_particle_feed(&S, rad.data(), pos.data(), total, NP, &seed);
        if (n <= total) break;
        total = n;
    }
    if (dem_set_particles(ctx, total, pos.data(), nullptr, nullptr) != DEM_OK) return fail("particles");
    if (dem_set_store_forces(ctx, 1) != DEM_OK) return fail("store forces");
    if (dem_log_header(log.c_str(), gen.c_str(), 0, NP, 1, total) != DEM_OK) return fail("log");
    const int per_out = (int)ceil(S.tPacking / S.dt / (in.nOutPacking > 0 ? in.nOutPacking : 1));
    const auto t0 = std::chrono::steady_clock::now();
    int outs = 0;
    while (dem_ctime(ctx) < S.tPacking) {
        const int rb0 = dem_n_rebuild(ctx);
        if (dem_step(ctx, per_out) != DEM_OK) return fail("step");
        if (dem_get_particles(ctx, pos.data(), vel.data(), angv.data()) != DEM_OK ||
            dem_get_forces(ctx, force.data(), mom.data(), rmom.data()) != DEM_OK)
            return fail("export");
        const double ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (dem_write_tec(tec.c_str(), outs, dem_ctime(ctx) * S.rtunit, total, pos.data(), rad.data(),
                          vel.data(), angv.data(), force.data()) != DEM_OK ||
            dem_log_output(log.c_str(), outs, dem_ctime(ctx) * S.rtunit, ms,
                           dem_n_rebuild(ctx) - rb0, total, NP, dem_iter(ctx)) != DEM_OK)
            return fail("output");
        ++outs;
    }
    const double s =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (dem_log_total(log.c_str(), s) != DEM_OK) return fail("log");
    if (argc > 3) {
        FILE *f = fopen(argv[3], "wb");
        if (!f) return fail("state file");
        const size_t m = 3 * (size_t)total;
        bool ok = fwrite(pos.data(), sizeof(double), m, f) == m &&
                  fwrite(vel.data(), sizeof(double), m, f) == m &&
                  fwrite(angv.data(), sizeof(double), m, f) == m;
        fclose(f);
        if (!ok) return fail("state file");
    }
    printf("dem_main: %d particles, %d outputs, %d steps, %d rebuilds\n", total, outs, dem_iter(ctx),
           dem_n_rebuild(ctx));
    dem_destroy(ctx);
    return 0;
}
