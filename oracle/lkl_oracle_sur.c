/*
 * lkl_oracle_sur.c — CPU restatement of the surrogate's pose features.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP feature kernel).
 *
 * The features have no counterpart in the reference: they are the new pair
 * histograms of the learned screening step.  Definition restated here by brute
 * force over every receptor x ligand pair:
 *   x = f32(R(q) v + t)   (f64 transform, orc_transform's order, qt.rs:67-74)
 *   d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz)) in f32, dx = x - rx
 *   counted when d2 < 900; bin = number of k in 1..23 with 1.5625 k^2 <= d2
 *   feature index (g_rec * 16 + g_lig) * 24 + bin, groups taken modulo 16.
 * The bin is found by a linear threshold scan (no square root), an
 * independent route from the kernel's root-and-correct estimate.
 *
 * Pinned by tests against a numpy restatement on small inputs (same formula,
 * vectorised) and by invariants: total count = number of pairs with d2 < 900;
 * counts are invariant to atom order and group relabelling moves counts
 * between features without changing the per-bin totals.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "lkl_oracle.h"

#define NG 16
#define NB 24
#define NF (NG * NG * NB)

static int bin_scan(float d2) {
    int b = 0;
    for (int k = 1; k < NB; ++k)
        if (1.5625f * (float)(k * k) <= d2) b = k;
    return b;
}

static void feat_one(const double* rxyz, const unsigned char* rg, int nr, const double* lxyz,
                     const unsigned char* lg, int nl, const double* pose7, uint32_t* cnt,
                     double* lc) {
    memset(cnt, 0, sizeof(uint32_t) * NF);
    orc_transform(lxyz, nl, pose7, lc);
    for (int j = 0; j < nl; ++j) {
        const float x = (float)lc[3 * j], y = (float)lc[3 * j + 1], z = (float)lc[3 * j + 2];
        const int gl = lg[j] % NG;
        for (int i = 0; i < nr; ++i) {
            const float dx = x - (float)rxyz[3 * i];
            const float dy = y - (float)rxyz[3 * i + 1];
            const float dz = z - (float)rxyz[3 * i + 2];
            const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
            if (d2 < 900.0f) cnt[((rg[i] % NG) * NG + gl) * NB + bin_scan(d2)] += 1u;
        }
    }
}

typedef struct {
    const double *rxyz, *lxyz, *poses;
    const unsigned char *rg, *lg;
    int nr, nl, n, t, nt;
    uint32_t* counts;
} feat_job;

static void* feat_worker(void* arg) {
    feat_job* J = (feat_job*)arg;
    double* lc = (double*)malloc(sizeof(double) * 3 * (size_t)J->nl);
    for (int k = J->t; k < J->n; k += J->nt)
        feat_one(J->rxyz, J->rg, J->nr, J->lxyz, J->lg, J->nl, J->poses + 7 * (size_t)k,
                 J->counts + (size_t)k * NF, lc);
    free(lc);
    return NULL;
}

/* counts: n x 6144 u32. */
void orc_feat_counts(const double* rxyz, const unsigned char* rg, int nr, const double* lxyz,
                     const unsigned char* lg, int nl, const double* poses, int n, uint32_t* counts,
                     int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    pthread_t th[64];
    feat_job jobs[64];
    for (int t = 0; t < nthreads; ++t) {
        feat_job j = {rxyz, lxyz, poses, rg, lg, nr, nl, n, t, nthreads, counts};
        jobs[t] = j;
        pthread_create(&th[t], NULL, feat_worker, &jobs[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
}

/* ── grid features ─────────────────────────────────────────────────────────
 * Receptor counts on nodes p = o + h * (ix, iy, iz) of the box grown by
 * 30 A + h (origin floor(min - 30 - h), upper corner ceil(max + 30 + h),
 * spacing h a power of two), each ligand atom replaced by its nearest node
 * ix = floor(fmaf(f32(x) - ox, 1/h, 0.5)) per axis (atoms whose node falls
 * outside contribute nothing).  Node counts are recomputed here by brute force
 * over the receptor for every (pose, atom). */
int64_t orc_feat_grid_dims(const double* rxyz, int nr, double h, double origin[3], int dims[3]) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = 0; i < nr; ++i)
        for (int a = 0; a < 3; ++a) {
            double v = (double)(float)rxyz[3 * i + a];
            if (v < lo[a]) lo[a] = v;
            if (v > hi[a]) hi[a] = v;
        }
    int64_t n = 1;
    for (int a = 0; a < 3; ++a) {
        origin[a] = floor(lo[a] - 30.0 - h);
        dims[a] = (int)((ceil(hi[a] + 30.0 + h) - origin[a]) / h) + 1;
        n *= dims[a];
    }
    return n;
}

/* counts (24 per receptor group, 384 total) of node (ix, iy, iz) */
void orc_feat_grid_node(const double* rxyz, const unsigned char* rg, int nr, const double origin[3],
                        double h, int ix, int iy, int iz, uint32_t* out384) {
    const float px = (float)origin[0] + (float)ix * (float)h;
    const float py = (float)origin[1] + (float)iy * (float)h;
    const float pz = (float)origin[2] + (float)iz * (float)h;
    memset(out384, 0, sizeof(uint32_t) * NG * NB);
    for (int i = 0; i < nr; ++i) {
        const float dx = px - (float)rxyz[3 * i];
        const float dy = py - (float)rxyz[3 * i + 1];
        const float dz = pz - (float)rxyz[3 * i + 2];
        const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        if (d2 < 900.0f) out384[(rg[i] % NG) * NB + bin_scan(d2)] += 1u;
    }
}

typedef struct {
    const double *rxyz, *lxyz, *poses;
    const unsigned char *rg, *lg;
    int nr, nl, n, t, nt;
    double h, origin[3];
    int dims[3];
    uint32_t* counts;
} grid_job;

static void* grid_worker(void* arg) {
    grid_job* J = (grid_job*)arg;
    double* lc = (double*)malloc(sizeof(double) * 3 * (size_t)J->nl);
    uint32_t node[NG * NB];
    const float inv = (float)(1.0 / J->h);
    for (int k = J->t; k < J->n; k += J->nt) {
        uint32_t* cnt = J->counts + (size_t)k * NF;
        memset(cnt, 0, sizeof(uint32_t) * NF);
        orc_transform(J->lxyz, J->nl, J->poses + 7 * (size_t)k, lc);
        for (int j = 0; j < J->nl; ++j) {
            int ix[3], in = 1;
            for (int a = 0; a < 3; ++a) {
                ix[a] = (int)floorf(fmaf((float)lc[3 * j + a] - (float)J->origin[a], inv, 0.5f));
                if (ix[a] < 0 || ix[a] >= J->dims[a]) in = 0;
            }
            if (!in) continue;
            orc_feat_grid_node(J->rxyz, J->rg, J->nr, J->origin, J->h, ix[0], ix[1], ix[2], node);
            const int gl = J->lg[j] % NG;
            for (int gr = 0; gr < NG; ++gr)
                for (int b = 0; b < NB; ++b) cnt[(gr * NG + gl) * NB + b] += node[gr * NB + b];
        }
    }
    free(lc);
    return NULL;
}

void orc_feat_grid_counts(const double* rxyz, const unsigned char* rg, int nr, const double* lxyz,
                          const unsigned char* lg, int nl, const double* poses, int n, double h,
                          uint32_t* counts, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    pthread_t th[64];
    grid_job jobs[64];
    double origin[3];
    int dims[3];
    orc_feat_grid_dims(rxyz, nr, h, origin, dims);
    for (int t = 0; t < nthreads; ++t) {
        grid_job j = {rxyz, lxyz, poses, rg, lg, nr, nl, n, t, nthreads, h,
                      {origin[0], origin[1], origin[2]}, {dims[0], dims[1], dims[2]}, counts};
        jobs[t] = j;
        pthread_create(&th[t], NULL, grid_worker, &jobs[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
}
