// ParticleFeed.cpp -- generation of new particles in the feed zone.
//
// The sampling law of ParticleFeed.cpp:8-72: an angle uniform in [0, 2pi),
// a radius fraction from the reflected sum of two uniforms (uniform over the
// disc of radius R - r), a height uniform in [Lmin, Lmax], rejection when a
// new sphere comes closer than r_i + r_new + 0.002 to any placed sphere,
// giving up after 500 consecutive rejections or 60000 placements per call.
// Two changes: the generator is seeded (the reference reseeds from
// time(NULL)), and the overlap test uses a uniform hash grid instead of a
// scan over all placed particles, which gives the same accept/reject
// decision in O(1) per candidate.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <unordered_map>
#include <vector>

#include "dem_host.h"

namespace {
// splitmix64: small, seedable, and identical on every platform
struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
};

struct Grid {
    double h, ox, oy, oz;
    std::unordered_map<int64_t, std::vector<int>> cells;
    int64_t key(long ix, long iy, long iz) const {
        return ((int64_t)(ix & 0x1FFFFF) << 42) | ((int64_t)(iy & 0x1FFFFF) << 21) | (int64_t)(iz & 0x1FFFFF);
    }
    long c(double v, double o) const { return (long)floor((v - o) / h); }
    void insert(int i, const double *p) { cells[key(c(p[0], ox), c(p[1], oy), c(p[2], oz))].push_back(i); }
};
}  // namespace

extern "C" int dem_particle_feed(const dem_setup *S, const double *rad, double *pos, int numParticle,
                                 int NP, unsigned long long *seed) {
    if (!S || !rad || !pos || !seed || numParticle < 0 || numParticle > NP) {
        dem_fail(DEM_ERR_ARG, "bad feed argument");
        return -1;
    }
    const double clear = 0.002;
    double rmax = 0.0;
    for (int i = 0; i < NP; ++i) rmax = rad[i] > rmax ? rad[i] : rmax;
    Grid g;
    g.h = 2.0 * rmax + clear;
    g.ox = S->cirX - S->R;
    g.oy = S->cirY - S->R;
    g.oz = S->feedLmin - g.h;
    for (int i = 0; i < numParticle; ++i) g.insert(i, &pos[3 * (size_t)i]);
    Rng rng{*seed};
    int reject = 0, placed = 0;
    while (reject < 500 && numParticle < NP && placed < 60000) {
        const double r = rad[numParticle];
        double th = 2.0 * 3.1415926 * rng.uniform();
        double u = rng.uniform() + rng.uniform();
        double mag = u > 1 ? 2 - u : u;
        double p[3];
        p[0] = mag * (S->R - r) * cos(th) + S->cirX;
        p[1] = mag * (S->R - r) * sin(th) + S->cirY;
        p[2] = rng.uniform() * (S->feedLmax - S->feedLmin) + S->feedLmin;
        bool hit = false;
        long cx = g.c(p[0], g.ox), cy = g.c(p[1], g.oy), cz = g.c(p[2], g.oz);
        for (long dz = -1; dz <= 1 && !hit; ++dz)
            for (long dy = -1; dy <= 1 && !hit; ++dy)
                for (long dx = -1; dx <= 1 && !hit; ++dx) {
                    auto it = g.cells.find(g.key(cx + dx, cy + dy, cz + dz));
                    if (it == g.cells.end()) continue;
                    for (int j : it->second) {
                        const double *q = &pos[3 * (size_t)j];
                        double ax = p[0] - q[0], ay = p[1] - q[1], az = p[2] - q[2];
                        if (sqrt(ax * ax + ay * ay + az * az) < (rad[j] + r + clear)) {
                            hit = true;
                            break;
                        }
                    }
                }
        if (hit) {
            ++reject;
            continue;
        }
        memcpy(&pos[3 * (size_t)numParticle], p, sizeof p);
        g.insert(numParticle, p);
        ++numParticle;
        ++placed;
        reject = 0;
    }
    *seed = rng.s;
    return numParticle;
}
