// DumpLoad.cpp -- restart files.
//
// dem_load_para keeps LoadPara (DumpLoad.cpp:7-18): it seeks the line
// starting "SYSTEM Info" and reads NP, nBatch and TotalParticle from the
// following text.  The reference has no writer and restores no particle
// state; dem_dump_write / dem_dump_read add both, in the same keyword-text
// style, carrying positions, velocities, spins, neighbour and wall lists,
// the full contact history, bridge counts and the clock, with %.17g so a
// round trip is exact.  dem_dump_write_binary writes the same records after
// the same text header (so LoadPara still reads it) as raw little-endian
// arrays, about a tenth of the size and written at disk speed: the form a
// multi-million-particle restart needs.
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

// everything a restart carries, as the getters return it
struct DumpState {
    int n = 0, K = 0;
    std::vector<double> pos, vel, w;
    std::vector<int> num, list, wnum, wlist, lqd;
    std::vector<double> hd, hn, hm, hr, hp;
    std::vector<int> hs, hb;
    std::vector<double> wd, wn, wm, wr, wp;
    std::vector<int> ws, wb;
    void size(int n_, int K_) {
        n = n_; K = K_;
        const size_t N = n, m = (size_t)n * K, W = (size_t)n * DEM_WALLS;
        pos.assign(3 * N, 0.0); vel.assign(3 * N, 0.0); w.assign(3 * N, 0.0);
        num.assign(N, 0); list.assign(m, -1); wnum.assign(N, 0); wlist.assign(W, -1);
        lqd.assign(N, 0);
        hd.assign(3 * m, 0.0); hn.assign(m, 0.0); hm.assign(m, 0.0); hr.assign(m, 0.0);
        hp.assign(m, 0.0); hs.assign(m, 0); hb.assign(m, 0);
        wd.assign(3 * W, 0.0); wn.assign(W, 0.0); wm.assign(W, 0.0); wr.assign(W, 0.0);
        wp.assign(W, 0.0); ws.assign(W, 0); wb.assign(W, 0);
    }
};

static int export_state(dem_ctx *c, DumpState &d) {
    int rc;
    if (c->rebuild_pending && c->has_lists == false && c->total > 0 && (rc = dem_rebuild(c)))
        return rc;
    d.size(c->total, c->K);
    const int K = c->K;
    if ((rc = dem_get_particles(c, d.pos.data(), d.vel.data(), d.w.data())) ||
        (rc = dem_get_nebor(c, d.num.data(), d.list.data(), K)) ||
        (rc = dem_get_hist_pp(c, K, d.hd.data(), d.hn.data(), d.hm.data(), d.hr.data(),
                              d.hp.data(), d.hs.data(), d.hb.data())) ||
        (rc = dem_get_walls(c, d.wnum.data(), d.wlist.data())) ||
        (rc = dem_get_hist_pw(c, d.wd.data(), d.wn.data(), d.wm.data(), d.wr.data(), d.wp.data(),
                              d.ws.data(), d.wb.data())) ||
        (rc = dem_get_lqdcnt(c, d.lqd.data())))
        return rc;
    return DEM_OK;
}

static int import_state(dem_ctx *c, DumpState &d, int iter, double ctime, double Tw, int nreb) {
    int rc;
    if ((rc = dem_set_particles(c, d.n, d.pos.data(), d.vel.data(), d.w.data())) ||
        (rc = dem_set_lqdcnt(c, d.lqd.data())) ||
        (rc = dem_set_contacts(c, d.K, d.num.data(), d.list.data(), d.hd.data(), d.hn.data(),
                               d.hm.data(), d.hr.data(), d.hp.data(), d.hs.data(), d.hb.data(),
                               d.wnum.data(), d.wlist.data(), d.wd.data(), d.wn.data(), d.wm.data(),
                               d.wr.data(), d.wp.data(), d.ws.data(), d.wb.data())) ||
        (rc = dem_set_clock(c, iter, ctime, Tw)))
        return rc;
    c->n_rebuild = nreb;
    return DEM_OK;
}

static void write_header(FILE *f, dem_ctx *c, int nBatch) {
    fprintf(f, "SYSTEM Info\n%d %d %d\n", c->cap, nBatch, c->total);
    fprintf(f, "CLOCK\n%d %.17g %.17g %d\n", c->iter, c->ctime, c->Tw, c->n_rebuild);
}

extern "C" int dem_dump_write(dem_ctx *c, const char *path, int nBatch) {
    if (!c || !path) return dem_fail(DEM_ERR_ARG, "null argument");
    DumpState d;
    int rc = export_state(c, d);
    if (rc) return rc;
    const int n = d.n, K = d.K, W = DEM_WALLS;
    FILE *f = fopen(path, "w");
    if (!f) return dem_fail(DEM_ERR_ARG, "cannot create %s", path);
    write_header(f, c, nBatch);
    fprintf(f, "PARTICLE STATE\n");
    for (int i = 0; i < n; ++i) {
        const double *p = &d.pos[3 * (size_t)i], *v = &d.vel[3 * (size_t)i], *o = &d.w[3 * (size_t)i];
        fprintf(f, "%.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %d\n", p[0], p[1], p[2],
                v[0], v[1], v[2], o[0], o[1], o[2], d.lqd[i]);
    }
    fprintf(f, "NEIGHBOR LIST\n%d\n", K);
    for (int i = 0; i < n; ++i) {
        fprintf(f, "%d", d.num[i]);
        for (int k = 0; k < d.num[i]; ++k) {
            size_t s = (size_t)i * K + k;
            fprintf(f, " %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %d %d", d.list[s], d.hd[3 * s],
                    d.hd[3 * s + 1], d.hd[3 * s + 2], d.hn[s], d.hm[s], d.hr[s], d.hp[s], d.hs[s],
                    d.hb[s]);
        }
        fprintf(f, "\n");
    }
    fprintf(f, "WALL LIST\n");
    for (int i = 0; i < n; ++i) {
        fprintf(f, "%d", d.wnum[i]);
        for (int k = 0; k < d.wnum[i]; ++k) {
            size_t s = (size_t)i * W + k;
            fprintf(f, " %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %d %d", d.wlist[s], d.wd[3 * s],
                    d.wd[3 * s + 1], d.wd[3 * s + 2], d.wn[s], d.wm[s], d.wr[s], d.wp[s], d.ws[s],
                    d.wb[s]);
        }
        fprintf(f, "\n");
    }
    fprintf(f, "END\n");
    int bad = ferror(f);
    fclose(f);
    return bad ? dem_fail(DEM_ERR_ARG, "write error on %s", path) : DEM_OK;
}

// Binary body: after "BINARY <K>\n", in this order, with m = sum of num:
// pos, vel, angv [n][3] f64; lqdCnt [n] i32; num [n] i32; the live slots of
// each particle in turn: partner [m] i32, dispt [m][3], nDsp, nDsp_mx,
// plstRad, plstDfm [m] f64, stage, lqdBrg [m] i32; then walls: wnum [n]
// i32, wall [n][3] i32, dispt [n][3][3], nDsp, nDsp_mx, plstRad, plstDfm
// [n][3] f64, stage, lqdBrg [n][3] i32; then "END\n".
template <class T> static bool put(FILE *f, const std::vector<T> &v) {
    return v.empty() || fwrite(v.data(), sizeof(T), v.size(), f) == v.size();
}
template <class T> static bool get(FILE *f, std::vector<T> &v) {
    return v.empty() || fread(v.data(), sizeof(T), v.size(), f) == v.size();
}

extern "C" int dem_dump_write_binary(dem_ctx *c, const char *path, int nBatch) {
    if (!c || !path) return dem_fail(DEM_ERR_ARG, "null argument");
    DumpState d;
    int rc = export_state(c, d);
    if (rc) return rc;
    const int n = d.n, K = d.K;
    size_t m = 0;
    for (int i = 0; i < n; ++i) m += (size_t)d.num[i];
    std::vector<int> L(m), S(m), B(m);
    std::vector<double> D(3 * m), N(m), X(m), R(m), P(m);
    size_t q = 0;
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < d.num[i]; ++k, ++q) {
            const size_t s = (size_t)i * K + k;
            L[q] = d.list[s]; S[q] = d.hs[s]; B[q] = d.hb[s];
            D[3 * q] = d.hd[3 * s]; D[3 * q + 1] = d.hd[3 * s + 1]; D[3 * q + 2] = d.hd[3 * s + 2];
            N[q] = d.hn[s]; X[q] = d.hm[s]; R[q] = d.hr[s]; P[q] = d.hp[s];
        }
    FILE *f = fopen(path, "wb");
    if (!f) return dem_fail(DEM_ERR_ARG, "cannot create %s", path);
    write_header(f, c, nBatch);
    fprintf(f, "BINARY %d\n", K);
    bool ok = put(f, d.pos) && put(f, d.vel) && put(f, d.w) && put(f, d.lqd) && put(f, d.num) &&
              put(f, L) && put(f, D) && put(f, N) && put(f, X) && put(f, R) && put(f, P) &&
              put(f, S) && put(f, B) && put(f, d.wnum) && put(f, d.wlist) && put(f, d.wd) &&
              put(f, d.wn) && put(f, d.wm) && put(f, d.wr) && put(f, d.wp) && put(f, d.ws) &&
              put(f, d.wb);
    ok = ok && fprintf(f, "END\n") > 0;
    ok = ok && !ferror(f);
    fclose(f);
    return ok ? DEM_OK : dem_fail(DEM_ERR_ARG, "write error on %s", path);
}

static int read_binary(dem_ctx *c, FILE *f, const char *path, int n, int K, DumpState &d) {
    if (K <= 0 || K > c->K) return dem_fail(DEM_ERR_ARG, "%s: list width %d", path, K);
    d.size(n, K);
    bool ok = get(f, d.pos) && get(f, d.vel) && get(f, d.w) && get(f, d.lqd) && get(f, d.num);
    size_t m = 0;
    for (int i = 0; ok && i < n; ++i) {
        ok = d.num[i] >= 0 && d.num[i] <= K;
        m += (size_t)d.num[i];
    }
    if (!ok) return dem_fail(DEM_ERR_ARG, "%s: malformed dump", path);
    std::vector<int> L(m), S(m), B(m);
    std::vector<double> D(3 * m), N(m), X(m), R(m), P(m);
    ok = get(f, L) && get(f, D) && get(f, N) && get(f, X) && get(f, R) && get(f, P) &&
         get(f, S) && get(f, B) && get(f, d.wnum) && get(f, d.wlist) && get(f, d.wd) &&
         get(f, d.wn) && get(f, d.wm) && get(f, d.wr) && get(f, d.wp) && get(f, d.ws) &&
         get(f, d.wb);
    for (int i = 0; ok && i < n; ++i) ok = d.wnum[i] >= 0 && d.wnum[i] <= DEM_WALLS;
    char tail[8] = {0};
    ok = ok && fgets(tail, sizeof tail, f) && strncmp(tail, "END", 3) == 0;
    if (!ok) return dem_fail(DEM_ERR_ARG, "%s: malformed dump", path);
    size_t q = 0;
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < d.num[i]; ++k, ++q) {
            const size_t s = (size_t)i * K + k;
            d.list[s] = L[q]; d.hs[s] = S[q]; d.hb[s] = B[q];
            d.hd[3 * s] = D[3 * q]; d.hd[3 * s + 1] = D[3 * q + 1]; d.hd[3 * s + 2] = D[3 * q + 2];
            d.hn[s] = N[q]; d.hm[s] = X[q]; d.hr[s] = R[q]; d.hp[s] = P[q];
        }
    return DEM_OK;
}

extern "C" int dem_dump_read(dem_ctx *c, const char *path, int *nBatch) {
    if (!c || !path) return dem_fail(DEM_ERR_ARG, "null argument");
    int NP = 0, nb = 0, n = 0;
    int rc = dem_load_para(path, &NP, &nb, &n);
    if (rc) return rc;
    if (NP != c->cap || n < 0 || n > c->cap)
        return dem_fail(DEM_ERR_ARG, "%s: NP %d does not match the context (%d)", path, NP, c->cap);
    FILE *f = fopen(path, "rb");
    if (!f) return dem_fail(DEM_ERR_ARG, "cannot open %s", path);
    const int W = DEM_WALLS;
    int iter = 0, nreb = 0, K = 0;
    double ctime = 0, Tw = 0;
    DumpState d;
    bool ok = seek_key(f, "CLOCK") && fscanf(f, "%d %lf %lf %d", &iter, &ctime, &Tw, &nreb) == 4;
    // the line after the clock: "BINARY K" or the text records
    char line[128] = {0};
    ok = ok && fgets(line, sizeof line, f) && fgets(line, sizeof line, f);
    if (ok && strncmp(line, "BINARY", 6) == 0) {
        if (sscanf(line + 6, "%d", &K) != 1) ok = false;
        if (ok) {
            rc = read_binary(c, f, path, n, K, d);
            fclose(f);
            if (rc) return rc;
            if ((rc = import_state(c, d, iter, ctime, Tw, nreb))) return rc;
            if (nBatch) *nBatch = nb;
            return DEM_OK;
        }
    }
    ok = ok && seek_key(f, "PARTICLE STATE");
    d.n = n;
    d.pos.assign(3 * (size_t)n, 0.0); d.vel.assign(3 * (size_t)n, 0.0); d.w.assign(3 * (size_t)n, 0.0);
    d.lqd.assign(n, 0);
    for (int i = 0; ok && i < n; ++i) {
        double *p = &d.pos[3 * (size_t)i], *v = &d.vel[3 * (size_t)i], *o = &d.w[3 * (size_t)i];
        ok = fscanf(f, "%lf %lf %lf %lf %lf %lf %lf %lf %lf %d", &p[0], &p[1], &p[2], &v[0], &v[1],
                    &v[2], &o[0], &o[1], &o[2], &d.lqd[i]) == 10;
    }
    ok = ok && seek_key(f, "NEIGHBOR LIST") && fscanf(f, "%d", &K) == 1 && K > 0 && K <= c->K;
    if (ok) {
        std::vector<double> pos = d.pos, vel = d.vel, w = d.w;
        std::vector<int> lqd = d.lqd;
        d.size(n, K);
        d.pos = pos; d.vel = vel; d.w = w; d.lqd = lqd;
    }
    for (int i = 0; ok && i < n; ++i) {
        ok = fscanf(f, "%d", &d.num[i]) == 1 && d.num[i] >= 0 && d.num[i] <= K;
        for (int k = 0; ok && k < d.num[i]; ++k) {
            size_t s = (size_t)i * K + k;
            ok = fscanf(f, "%d %lf %lf %lf %lf %lf %lf %lf %d %d", &d.list[s], &d.hd[3 * s],
                        &d.hd[3 * s + 1], &d.hd[3 * s + 2], &d.hn[s], &d.hm[s], &d.hr[s], &d.hp[s],
                        &d.hs[s], &d.hb[s]) == 10;
        }
    }
    ok = ok && seek_key(f, "WALL LIST");
    for (int i = 0; ok && i < n; ++i) {
        ok = fscanf(f, "%d", &d.wnum[i]) == 1 && d.wnum[i] >= 0 && d.wnum[i] <= W;
        for (int k = 0; ok && k < d.wnum[i]; ++k) {
            size_t s = (size_t)i * W + k;
            ok = fscanf(f, "%d %lf %lf %lf %lf %lf %lf %lf %d %d", &d.wlist[s], &d.wd[3 * s],
                        &d.wd[3 * s + 1], &d.wd[3 * s + 2], &d.wn[s], &d.wm[s], &d.wr[s], &d.wp[s],
                        &d.ws[s], &d.wb[s]) == 10;
        }
    }
    fclose(f);
    if (!ok) return dem_fail(DEM_ERR_ARG, "%s: malformed dump", path);
    if ((rc = import_state(c, d, iter, ctime, Tw, nreb))) return rc;
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
