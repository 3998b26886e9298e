// DumpLoad.cpp -- restart files.
//
// dem_load_para keeps LoadPara (DumpLoad.cpp:7-18): it seeks the line
// starting "SYSTEM Info" and reads NP, nBatch and TotalParticle from the
// following text.  The reference has no writer and restores no particle
// state; dem_dump_write / dem_dump_read add both, in the same keyword-text
// style, carrying positions, velocities, spins, neighbour and wall lists,
// the full contact history, bridge counts and the clock, with %.17g so a
// round trip is exact.
#include <stdio.h>
#include <string.h>

#include <vector>

#include "dem_host.h"

extern "C" int dem_set_lqdcnt(dem_ctx *c, const int *lqd);
extern "C" int dem_set_clock(dem_ctx *c, int iter, double ctime, double Tw);

static bool seek_key(FILE *f, const char *key) {
    rewind(f);
    char line[512];
    size_t n = strlen(key);
    while (fgets(line, sizeof line, f))
        if (strncmp(line, key, n) == 0) return true;
    return false;
}

extern "C" int dem_load_para(const char *path, int *NP, int *nBatch, int *TotalParticle) {
    if (!path || !NP || !nBatch || !TotalParticle) return dem_fail(DEM_ERR_ARG, "null argument");
    FILE *f = fopen(path, "r");
    if (!f) return dem_fail(DEM_ERR_ARG, "cannot open %s", path);
    int rc = DEM_OK;
    if (!seek_key(f, "SYSTEM Info") || fscanf(f, "%d %d %d", NP, nBatch, TotalParticle) != 3)
        rc = dem_fail(DEM_ERR_ARG, "%s: no SYSTEM Info record", path);
    fclose(f);
    return rc;
}

extern "C" int dem_dump_write(dem_ctx *c, const char *path, int nBatch) {
    if (!c || !path) return dem_fail(DEM_ERR_ARG, "null argument");
    const int n = c->total, K = c->K, W = DEM_WALLS;
    std::vector<double> pos(3 * (size_t)n), vel(3 * (size_t)n), w(3 * (size_t)n);
    std::vector<int> num(n), list((size_t)n * K), wnum(n), wlist((size_t)n * W), lqd(n);
    std::vector<double> hd((size_t)n * K * 3), hn((size_t)n * K), hm((size_t)n * K),
        hr((size_t)n * K), hp((size_t)n * K);
    std::vector<int> hs((size_t)n * K), hb((size_t)n * K);
    std::vector<double> wd((size_t)n * W * 3), wn((size_t)n * W), wm((size_t)n * W),
        wr((size_t)n * W), wp((size_t)n * W);
    std::vector<int> ws((size_t)n * W), wb((size_t)n * W);
    int rc;
    if (c->rebuild_pending && c->has_lists == false && n > 0 && (rc = dem_rebuild(c))) return rc;
    if ((rc = dem_get_particles(c, pos.data(), vel.data(), w.data())) ||
        (rc = dem_get_nebor(c, num.data(), list.data(), K)) ||
        (rc = dem_get_hist_pp(c, K, hd.data(), hn.data(), hm.data(), hr.data(), hp.data(), hs.data(),
                              hb.data())) ||
        (rc = dem_get_walls(c, wnum.data(), wlist.data())) ||
        (rc = dem_get_hist_pw(c, wd.data(), wn.data(), wm.data(), wr.data(), wp.data(), ws.data(),
                              wb.data())) ||
        (rc = dem_get_lqdcnt(c, lqd.data())))
        return rc;
    FILE *f = fopen(path, "w");
    if (!f) return dem_fail(DEM_ERR_ARG, "cannot create %s", path);
    fprintf(f, "SYSTEM Info\n%d %d %d\n", c->cap, nBatch, n);
    fprintf(f, "CLOCK\n%d %.17g %.17g %d\n", c->iter, c->ctime, c->Tw, c->n_rebuild);
    fprintf(f, "PARTICLE STATE\n");
    for (int i = 0; i < n; ++i) {
        const double *p = &pos[3 * (size_t)i], *v = &vel[3 * (size_t)i], *o = &w[3 * (size_t)i];
        fprintf(f, "%.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %d\n", p[0], p[1], p[2],
                v[0], v[1], v[2], o[0], o[1], o[2], lqd[i]);
    }
    fprintf(f, "NEIGHBOR LIST\n%d\n", K);
    for (int i = 0; i < n; ++i) {
        fprintf(f, "%d", num[i]);
        for (int k = 0; k < num[i]; ++k) {
            size_t s = (size_t)i * K + k;
            fprintf(f, " %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %d %d", list[s], hd[3 * s],
                    hd[3 * s + 1], hd[3 * s + 2], hn[s], hm[s], hr[s], hp[s], hs[s], hb[s]);
        }
        fprintf(f, "\n");
    }
    fprintf(f, "WALL LIST\n");
    for (int i = 0; i < n; ++i) {
        fprintf(f, "%d", wnum[i]);
        for (int k = 0; k < wnum[i]; ++k) {
            size_t s = (size_t)i * W + k;
            fprintf(f, " %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %d %d", wlist[s], wd[3 * s],
                    wd[3 * s + 1], wd[3 * s + 2], wn[s], wm[s], wr[s], wp[s], ws[s], wb[s]);
        }
        fprintf(f, "\n");
    }
    fprintf(f, "END\n");
    int bad = ferror(f);
    fclose(f);
    return bad ? dem_fail(DEM_ERR_ARG, "write error on %s", path) : DEM_OK;
}

extern "C" int dem_dump_read(dem_ctx *c, const char *path, int *nBatch) {
    if (!c || !path) return dem_fail(DEM_ERR_ARG, "null argument");
    int NP = 0, nb = 0, n = 0;
    int rc = dem_load_para(path, &NP, &nb, &n);
    if (rc) return rc;
    if (NP != c->cap || n < 0 || n > c->cap)
        return dem_fail(DEM_ERR_ARG, "%s: NP %d does not match the context (%d)", path, NP, c->cap);
    FILE *f = fopen(path, "r");
    if (!f) return dem_fail(DEM_ERR_ARG, "cannot open %s", path);
    const int W = DEM_WALLS;
    int iter = 0, nreb = 0, K = 0;
    double ctime = 0, Tw = 0;
    std::vector<double> pos(3 * (size_t)n), vel(3 * (size_t)n), w(3 * (size_t)n);
    std::vector<int> lqd(n), num(n), wnum(n), wlist((size_t)n * W, -1);
    std::vector<int> list, hs, hb;
    std::vector<double> hd, hn, hm, hr, hp;
    std::vector<double> wd((size_t)n * W * 3), wn((size_t)n * W), wm((size_t)n * W),
        wr((size_t)n * W), wp((size_t)n * W);
    std::vector<int> ws((size_t)n * W), wb((size_t)n * W);
    bool ok = seek_key(f, "CLOCK") && fscanf(f, "%d %lf %lf %d", &iter, &ctime, &Tw, &nreb) == 4;
    ok = ok && seek_key(f, "PARTICLE STATE");
    for (int i = 0; ok && i < n; ++i) {
        double *p = &pos[3 * (size_t)i], *v = &vel[3 * (size_t)i], *o = &w[3 * (size_t)i];
        ok = fscanf(f, "%lf %lf %lf %lf %lf %lf %lf %lf %lf %d", &p[0], &p[1], &p[2], &v[0], &v[1],
                    &v[2], &o[0], &o[1], &o[2], &lqd[i]) == 10;
    }
    ok = ok && seek_key(f, "NEIGHBOR LIST") && fscanf(f, "%d", &K) == 1 && K > 0 && K <= c->K;
    if (ok) {
        size_t m = (size_t)n * K;
        list.assign(m, -1); hs.assign(m, 0); hb.assign(m, 0);
        hd.assign(3 * m, 0.0); hn.assign(m, 0.0); hm.assign(m, 0.0); hr.assign(m, 0.0); hp.assign(m, 0.0);
    }
    for (int i = 0; ok && i < n; ++i) {
        ok = fscanf(f, "%d", &num[i]) == 1 && num[i] >= 0 && num[i] <= K;
        for (int k = 0; ok && k < num[i]; ++k) {
            size_t s = (size_t)i * K + k;
            ok = fscanf(f, "%d %lf %lf %lf %lf %lf %lf %lf %d %d", &list[s], &hd[3 * s], &hd[3 * s + 1],
                        &hd[3 * s + 2], &hn[s], &hm[s], &hr[s], &hp[s], &hs[s], &hb[s]) == 10;
        }
    }
    ok = ok && seek_key(f, "WALL LIST");
    for (int i = 0; ok && i < n; ++i) {
        ok = fscanf(f, "%d", &wnum[i]) == 1 && wnum[i] >= 0 && wnum[i] <= W;
        for (int k = 0; ok && k < wnum[i]; ++k) {
            size_t s = (size_t)i * W + k;
            ok = fscanf(f, "%d %lf %lf %lf %lf %lf %lf %lf %d %d", &wlist[s], &wd[3 * s], &wd[3 * s + 1],
                        &wd[3 * s + 2], &wn[s], &wm[s], &wr[s], &wp[s], &ws[s], &wb[s]) == 10;
        }
    }
    fclose(f);
    if (!ok) return dem_fail(DEM_ERR_ARG, "%s: malformed dump", path);
    if ((rc = dem_set_particles(c, n, pos.data(), vel.data(), w.data())) ||
        (rc = dem_set_lqdcnt(c, lqd.data())) ||
        (rc = dem_set_contacts(c, K, num.data(), list.data(), hd.data(), hn.data(), hm.data(), hr.data(),
                               hp.data(), hs.data(), hb.data(), wnum.data(), wlist.data(), wd.data(),
                               wn.data(), wm.data(), wr.data(), wp.data(), ws.data(), wb.data())) ||
        (rc = dem_set_clock(c, iter, ctime, Tw)))
        return rc;
    c->n_rebuild = nreb;
    if (nBatch) *nBatch = nb;
    return DEM_OK;
}

extern "C" int dem_set_lqdcnt(dem_ctx *c, const int *lqd) {
    if (!c || !lqd) return dem_fail(DEM_ERR_ARG, "null argument");
    std::vector<int> ids(c->total);
    HIPCHK(hipMemcpyAsync(ids.data(), c->id[c->pcur], c->total * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<int> v(c->total);
    for (int i = 0; i < c->total; ++i) v[i] = lqd[ids[i]];
    HIPCHK(hipMemcpyAsync(c->lq[c->Lidx], v.data(), c->total * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return DEM_OK;
}

extern "C" int dem_set_clock(dem_ctx *c, int iter, double ctime, double Tw) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null argument");
    c->iter = iter;
    c->ctime = ctime;
    c->Tw = Tw;
    return DEM_OK;
}
