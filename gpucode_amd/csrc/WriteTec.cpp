// WriteTec.cpp -- Tecplot ASCII output of the particle state.
//
// The file format of WriteTec.cpp:9-38 is kept as is, so existing Tecplot
// layouts and post-processing read it unchanged: a TITLE and VARIABLES
// header written when outs == 0, then per call one ZONE whose title is the
// time in seconds, and one row per particle of X Y Z R VX VY VZ |omega| |F|
// in reduced units.
#include <math.h>
#include <stdio.h>

#include "dem_host.h"

static double norm3(const double *v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

extern "C" int dem_write_tec(const char *path, int outs, double time_s, int total, const double *pos,
                             const double *rad, const double *vel, const double *angv,
                             const double *force) {
    if (!path || !pos || !rad || !vel || !angv || !force || total < 0)
        return dem_fail(DEM_ERR_ARG, "bad WriteTec argument");
    if (outs == 0) {
        FILE *f = fopen(path, "w");
        if (!f) return dem_fail(DEM_ERR_ARG, "cannot create %s", path);
        fprintf(f, "TITLE = \" PARTICLE INFORMATION \" \n");
        fprintf(f, "VARIABLES = X   Y   Z   R   VX   VY   VZ   W   F\n");
        fclose(f);
    }
    FILE *f = fopen(path, "a");
    if (!f) return dem_fail(DEM_ERR_ARG, "cannot append to %s", path);
    fprintf(f, "ZONE T= \" %12.6lf s \" \n", time_s);
    for (int i = 0; i < total; ++i) {
        const double *p = pos + 3 * (size_t)i, *v = vel + 3 * (size_t)i;
        fprintf(f, "%11.5lf   %11.5lf   %11.5lf   %11.5f  %11.5lf   %11.5lf   %11.5lf   %11.5lf   %11.5lf\n",
                p[0], p[1], p[2], rad[i], v[0], v[1], v[2], norm3(angv + 3 * (size_t)i),
                norm3(force + 3 * (size_t)i));
    }
    fclose(f);
    return DEM_OK;
}
