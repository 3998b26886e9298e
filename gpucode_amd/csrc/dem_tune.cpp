// dem_tune.cpp -- autotuner of the skin, step kernel and storage order.
//
// Each candidate runs the packing from the same initial state in a fresh
// context for a fixed number of steps, rebuilds included, timed with HIP
// events on the context's stream after one untimed warm-up step.
#include <math.h>
#include <string.h>

#include <vector>

#include "../../include/dem_tune.h"
#include "dem_host.h"

// largest rupture distance of a bridge in reduced units: two particles of
// the largest size, each with the liquid of 12 bridges (kernel.cu:429-434,
// 511-526)
static double max_rupture(const dem_setup &S) {
    const double V = 2.0 * S.vol / 12.0;
    return S.ruptureMode == 1 ? (1.0 + 0.5 * S.theta) * cbrt(V) : 0.99 * pow(V, 0.33);
}

static int time_candidate(const dem_input &in, const double *dia, int n, const double *pos,
                          const double *vel, int steps, int precision, int kernel, int morton,
                          double *ms) {
    dem_ctx *c = nullptr;
    int rc = dem_create(&in, dia, precision, &c);
    if (rc) return rc;
    if ((rc = dem_set_kernel(c, kernel)) || (rc = dem_set_order(c, morton)) ||
        (rc = dem_set_particles(c, n, pos, vel, nullptr)) || (rc = dem_step(c, 1))) {
        dem_destroy(c);
        return rc;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        dem_destroy(c);
        return dem_fail(DEM_ERR_HIP, "hipEventCreate failed");
    }
    (void)hipEventRecord(e0, c->stream);
    rc = dem_step(c, steps);
    (void)hipEventRecord(e1, c->stream);
    float t = 0.f;
    if (rc == DEM_OK && hipEventSynchronize(e1) == hipSuccess &&
        hipEventElapsedTime(&t, e0, e1) == hipSuccess)
        *ms = (double)t / steps;
    else if (rc == DEM_OK)
        rc = dem_fail(DEM_ERR_HIP, "event timing failed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dem_destroy(c);
    return rc;
}

extern "C" int dem_autotune(const dem_input *in, const double *dia, int n, const double *pos,
                            const double *vel, int steps, const double *cutGaps, int ncut, int precision,
                            dem_tune_result *best, double *table) {
    static const double kDefault[5] = {0.05, 0.1, 0.2, 0.3, 0.5};
    if (!in || !pos || n < 1 || n > in->NP || steps < 1 || !best)
        return dem_fail(DEM_ERR_ARG, "bad argument");
    if (!cutGaps) {
        cutGaps = kDefault;
        ncut = 5;
    }
    if (ncut < 1) return dem_fail(DEM_ERR_ARG, "no candidate skin");
    memset(best, 0, sizeof(*best));
    best->ms_per_step = HUGE_VAL;
    if (table)
        for (int q = 0; q < ncut * 6; ++q) table[q] = -1.0;
    const int kernels[3] = {DEM_KERNEL_LOOP, DEM_KERNEL_COMPACT, DEM_KERNEL_WAVE};
    // the input's own configuration
    int rc = time_candidate(*in, dia, n, pos, vel, steps, precision, DEM_KERNEL_AUTO, 1,
                            &best->ms_default);
    if (rc) return rc;
    for (int s = 0; s < ncut; ++s) {
        dem_input ci = *in;
        ci.cutGap = cutGaps[s];
        dem_setup S;
        std::vector<double> rad(in->NP);
        std::vector<int> lvl(in->NP), idx(in->NP);
        if (dem_derive(&ci, dia, &S, rad.data(), lvl.data(), idx.data()) != DEM_OK) continue;
        if (S.forceCap == 1 && S.searchGap < max_rupture(S)) continue;
        int K = 0;
        for (int l = 0; l < S.nLevels; ++l) K = S.levelNNebors[l] > K ? S.levelNNebors[l] : K;
        for (int kq = 0; kq < 3; ++kq) {
            if (kernels[kq] == DEM_KERNEL_COMPACT && (S.forceCap == 1 || K > 64)) continue;
            for (int mo = 0; mo < 2; ++mo) {
                double ms = 0.0;
                rc = time_candidate(ci, dia, n, pos, vel, steps, precision, kernels[kq], mo, &ms);
                if (rc == DEM_ERR_OVERFLOW) break;  // this skin's lists are too short
                if (rc) return rc;
                best->n_tried++;
                if (table) table[(s * 3 + kq) * 2 + mo] = ms;
                if (ms < best->ms_per_step) {
                    best->ms_per_step = ms;
                    best->cutGap = cutGaps[s];
                    best->kernel = kernels[kq];
                    best->morton = mo;
                }
            }
            if (rc == DEM_ERR_OVERFLOW) break;
        }
    }
    if (best->n_tried == 0) return dem_fail(DEM_ERR_ARG, "no candidate could run");
    return DEM_OK;
}
