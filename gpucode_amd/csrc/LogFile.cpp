// LogFile.cpp -- the run log of the reference's driver, <genfile>_Log.dat.
//
// main.cpp:93-139 writes a header (a new run or a restart, the wall-clock
// time, the project name, the capacity and, on a restart, the feed batch
// and particle count), main.cpp:591-618 one record per output (output
// number, simulated time, elapsed milliseconds, rebuilds since the last
// output, particles while feeding, iteration), and main.cpp:631-634 the
// run's total.  The same text, written by three calls; each opens the file
// itself (the header truncates it, the others append).
#include <stdio.h>
#include <time.h>

#include "dem_host.h"

static const char *kRule = "-------------------------------------------------------\n";
static const char *kBar = "=======================================================\n";

extern "C" int dem_log_header(const char *path, const char *genfile, int restart, int NP, int nBatch,
                              int total) {
    if (!path || !genfile) return dem_fail(DEM_ERR_ARG, "null argument");
    FILE *f = fopen(path, "wt");
    if (!f) return dem_fail(DEM_ERR_ARG, "cannot create %s", path);
    time_t now;
    time(&now);
    if (!restart) {
        fprintf(f, "==================|*--- PROJECT---*|===================\n");
        fprintf(f, "<File>: new\n");
    } else {
        fprintf(f, "\n\n==============|*--- PROJECT ---*|===============\n");
        fprintf(f, "<File>: restarted\n");
    }
    fprintf(f, "<Time>: %s", ctime(&now));
    fprintf(f, "<Name>: %s\n", genfile);
    fprintf(f, "%s\n\n", kBar);
    fprintf(f, "\n\n=================================\n");
    fprintf(f, "<*--- Simulation Parameters ---*>\n");
    fprintf(f, "%s", kBar);
    if (!restart) {
        fprintf(f, "<Particle>: %d \n", NP);
        fprintf(f, "%s", kRule);
    } else {
        fprintf(f, "<Particle>: %d \t", NP);
        fprintf(f, "%s", kRule);
        fprintf(f, "<Feed batch>: %d \t", nBatch);
        fprintf(f, "<Current particle>: %d \n", total);
        fprintf(f, "%s", kRule);
    }
    int bad = ferror(f);
    fclose(f);
    return bad ? dem_fail(DEM_ERR_ARG, "write error on %s", path) : DEM_OK;
}

extern "C" int dem_log_output(const char *path, int outs, double runtime_s, double elapsed_ms,
                              int nrebuild, int total, int NP, int iter) {
    if (!path) return dem_fail(DEM_ERR_ARG, "null argument");
    FILE *f = fopen(path, "at");
    if (!f) return dem_fail(DEM_ERR_ARG, "cannot open %s", path);
    fprintf(f, "%-5d\t", outs);
    fprintf(f, "runtime:%-7.4lf\t", runtime_s);
    fprintf(f, "elapsed:%-20.5fms\t", elapsed_ms);
    fprintf(f, "rebuild:%-d\t", nrebuild);
    if (total < NP) fprintf(f, "particle:%-d\t", total);  // while feeding
    fprintf(f, "Iterate:%-d\n", iter);
    fprintf(f, "%s", kRule);
    int bad = ferror(f);
    fclose(f);
    return bad ? dem_fail(DEM_ERR_ARG, "write error on %s", path) : DEM_OK;
}

extern "C" int dem_log_total(const char *path, double seconds) {
    if (!path) return dem_fail(DEM_ERR_ARG, "null argument");
    FILE *f = fopen(path, "at");
    if (!f) return dem_fail(DEM_ERR_ARG, "cannot open %s", path);
    fprintf(f, "\nGPU DEM Time used: %-12.2lf s \n\n", seconds);
    int bad = ferror(f);
    fclose(f);
    return bad ? dem_fail(DEM_ERR_ARG, "write error on %s", path) : DEM_OK;
}
