/*
 * lkl_oracle_gso.c — CPU restatement of LKlight's GSO (glowworm swarm
 * optimisation) with CPU-grid scoring.  TEST INFRASTRUCTURE ONLY (the checker
 * of the product's GPU-scored GSO driver).
 *
 *   rng        rand 0.7.3 StdRng = rand_chacha 0.2.2 ChaCha20 (20 rounds, 64-bit
 *              block counter, zero stream), seeded by rand_core 0.5.1
 *              seed_from_u64 (PCG32 expansion); f64 = (next_u64 >> 11) * 2^-53.
 *              Pinned by the KAT of qt.rs:474-487.
 *   quaternion qt.rs:29-116 (normalize, dot, slerp with the linear threshold
 *              0.9995 of constants.rs:11, random).  Pinned by qt.rs:224-488 KATs.
 *   GSO        lib.rs:66-78 (run), swarm.rs:70-174 (update_luciferin, movement
 *              phase), glowworm.rs:61-166 (luciferin, neighbours, probabilities,
 *              neighbour choice, move, vision range).
 */
#include "lkl_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ── ChaCha20 StdRng ─────────────────────────────────────────────────────── */
typedef struct {
    uint32_t key[8];
    uint64_t ctr;
    uint32_t buf[16];
    int idx;
} orc_rng;

static uint32_t rotl32(uint32_t v, int c) { return (v << c) | (v >> (32 - c)); }

#define QR(a, b, c, d)                          \
    x[a] += x[b]; x[d] = rotl32(x[d] ^ x[a], 16); \
    x[c] += x[d]; x[b] = rotl32(x[b] ^ x[c], 12); \
    x[a] += x[b]; x[d] = rotl32(x[d] ^ x[a], 8);  \
    x[c] += x[d]; x[b] = rotl32(x[b] ^ x[c], 7);

static void chacha_block(orc_rng* r) {
    uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
    for (int i = 0; i < 8; ++i) s[4 + i] = r->key[i];
    s[12] = (uint32_t)r->ctr;
    s[13] = (uint32_t)(r->ctr >> 32);
    s[14] = s[15] = 0;
    uint32_t x[16];
    memcpy(x, s, sizeof x);
    for (int i = 0; i < 10; ++i) {
        QR(0, 4, 8, 12) QR(1, 5, 9, 13) QR(2, 6, 10, 14) QR(3, 7, 11, 15)
        QR(0, 5, 10, 15) QR(1, 6, 11, 12) QR(2, 7, 8, 13) QR(3, 4, 9, 14)
    }
    for (int i = 0; i < 16; ++i) r->buf[i] = x[i] + s[i];
    r->ctr++;
    r->idx = 0;
}

static void rng_seed(orc_rng* r, uint64_t state) {
    const uint64_t MUL = 6364136223846793005ULL, INC = 11634580027462260723ULL;
    for (int i = 0; i < 8; ++i) {
        state = state * MUL + INC;
        uint32_t xs = (uint32_t)(((state >> 18) ^ state) >> 27);
        uint32_t rot = (uint32_t)(state >> 59);
        r->key[i] = (xs >> rot) | (xs << ((32 - rot) & 31));
    }
    r->ctr = 0;
    r->idx = 16;
}

static uint32_t rng_u32(orc_rng* r) {
    if (r->idx >= 16) chacha_block(r);
    return r->buf[r->idx++];
}

static double rng_f64(orc_rng* r) {
    uint64_t lo = rng_u32(r), hi = rng_u32(r);
    return (double)(((hi << 32) | lo) >> 11) * (1.0 / 9007199254740992.0);
}

void orc_rng_f64s(uint64_t seed, int n, double* out) {
    orc_rng r;
    rng_seed(&r, seed);
    for (int i = 0; i < n; ++i) out[i] = rng_f64(&r);
}

/* ── quaternions (w, x, y, z) ─────────────────────────────────────────────── */
static void q_normalize(double* q) {
    double n = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    for (int i = 0; i < 4; ++i) q[i] /= n;
}

static double q_dot(const double* a, const double* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}

/* qt.rs:80-104 */
void orc_slerp(const double* a, const double* b, double t, double* out) {
    double q1[4], q2[4];
    memcpy(q1, a, sizeof q1);
    memcpy(q2, b, sizeof q2);
    q_normalize(q1);
    q_normalize(q2);
    double d = q_dot(q1, q2);
    if (d < 0.0) {
        for (int i = 0; i < 4; ++i) q1[i] = -q1[i];
        d *= -1.0;
    }
    if (d > 0.9995) {
        for (int i = 0; i < 4; ++i) out[i] = q1[i] + (q2[i] - q1[i]) * t;
        q_normalize(out);
    } else {
        d = fmax(fmin(d, 1.0), -1.0);
        double omega = acos(d), so = sin(omega);
        double s1 = sin((1.0 - t) * omega) / so, s2 = sin(t * omega) / so;
        for (int i = 0; i < 4; ++i) out[i] = q1[i] * s1 + q2[i] * s2;
    }
}

/* qt.rs:106-116 */
void orc_quat_random(uint64_t seed, int n, double* out) {
    orc_rng r;
    rng_seed(&r, seed);
    for (int k = 0; k < n; ++k) {
        double u1 = rng_f64(&r), u2 = rng_f64(&r), u3 = rng_f64(&r);
        double* q = out + 4 * k;
        q[0] = sqrt(1.0 - u1) * sin(2.0 * M_PI * u2);
        q[1] = sqrt(1.0 - u1) * cos(2.0 * M_PI * u2);
        q[2] = sqrt(u1) * sin(2.0 * M_PI * u3);
        q[3] = sqrt(u1) * cos(2.0 * M_PI * u3);
    }
}

/* ── glowworms ─────────────────────────────────────────────────────────────── */
typedef struct {
    double t[3], q[4];
    double luciferin, vision, scoring;
    int moved, step;
    int* nb;
    int n_nb;
    double* prob;
} gw;

#define RHO 0.5
#define GAMMA 0.4
#define BETA 0.08
#define MAX_VISION 5.0
#define MAX_NEIGHBORS 5

typedef struct {
    int family;
    const orc_mol *rec, *lig;
    const orc_field* f;
    gw* g;
    int n, tid, nthreads;
} score_job;

static void* score_worker(void* arg) {
    score_job* s = (score_job*)arg;
    for (int i = s->tid; i < s->n; i += s->nthreads) {
        gw* w = &s->g[i];
        if (!(w->moved || w->step == 0)) continue;
        double pose[7] = {w->q[0], w->q[1], w->q[2], w->q[3], w->t[0], w->t[1], w->t[2]};
        orc_parts p;
        if (s->family == ORC_DNA) orc_dna_energy_grid(s->rec, s->lig, s->f, pose, &p);
        else if (s->family == ORC_VDW) orc_vdw_energy_grid(s->rec, s->lig, pose, &p);
        else if (s->family == ORC_PYDOCK) orc_pydock_energy_grid(s->rec, s->lig, s->f, pose, &p);
        else orc_cpydock_energy_grid(s->rec, s->lig, s->f, pose, &p);
        w->scoring = orc_fold_score(s->family, p.e_raw, p.v, p.s);
    }
    return NULL;
}

/* glowworm.rs:61-72 for every glowworm (the per-glowworm path of swarm.rs:102). */
static void update_luciferin(int family, const orc_mol* rec, const orc_mol* lig,
                             const orc_field* f, gw* g, int n, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    score_job jobs[64];
    pthread_t th[64];
    for (int t = 0; t < nthreads; ++t) {
        score_job j = {family, rec, lig, f, g, n, t, nthreads};
        jobs[t] = j;
    }
    for (int t = 1; t < nthreads; ++t) pthread_create(&th[t], NULL, score_worker, &jobs[t]);
    score_worker(&jobs[0]);
    for (int t = 1; t < nthreads; ++t) pthread_join(th[t], NULL);
    for (int i = 0; i < n; ++i) {
        g[i].luciferin = (1.0 - RHO) * g[i].luciferin + GAMMA * g[i].scoring;
        g[i].step += 1;
    }
}

static double dist2(const gw* a, const gw* b) {
    double dx = a->t[0] - b->t[0], dy = a->t[1] - b->t[1], dz = a->t[2] - b->t[2];
    return dx * dx + dy * dy + dz * dz;
}

/* swarm.rs:105-174 + glowworm.rs:84-166 */
static void movement_phase(gw* g, int n, orc_rng* rng) {
    double(*pos)[3] = malloc(sizeof(double) * 3 * (size_t)n);
    double(*rot)[4] = malloc(sizeof(double) * 4 * (size_t)n);
    double* luc = malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; ++i) {
        memcpy(pos[i], g[i].t, sizeof pos[i]);
        memcpy(rot[i], g[i].q, sizeof rot[i]);
        luc[i] = g[i].luciferin;
    }
    for (int i = 0; i < n; ++i) {  /* neighbours: id order, higher luciferin, within vision */
        gw* a = &g[i];
        a->n_nb = 0;
        double vr2 = a->vision * a->vision;
        for (int k = 0; k < n; ++k)
            if (k != i && a->luciferin < g[k].luciferin && dist2(a, &g[k]) < vr2) a->nb[a->n_nb++] = k;
        double total = 0.0;
        for (int m = 0; m < a->n_nb; ++m) {
            a->prob[m] = luc[a->nb[m]] - a->luciferin;
            total += a->prob[m];
        }
        for (int m = 0; m < a->n_nb; ++m) a->prob[m] /= total;
    }
    double* rnd = malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; ++i) rnd[i] = rng_f64(rng);
    for (int i = 0; i < n; ++i) {
        gw* a = &g[i];
        int nid = i;
        if (a->n_nb > 0) {
            double sum = 0.0;
            int m = 0;
            while (sum < rnd[i] && m < a->n_nb) sum += a->prob[m++];
            nid = a->nb[m > 0 ? m - 1 : 0];
        }
        a->moved = nid != i;
        if (a->moved) {
            double dx = pos[nid][0] - a->t[0], dy = pos[nid][1] - a->t[1], dz = pos[nid][2] - a->t[2];
            double norm = sqrt(dx * dx + dy * dy + dz * dz);
            double coef = 0.5 / norm;
            a->t[0] += dx * coef;
            a->t[1] += dy * coef;
            a->t[2] += dz * coef;
            double q[4];
            orc_slerp(a->q, rot[nid], 0.5, q);
            q_normalize(q);
            memcpy(a->q, q, sizeof q);
        }
        a->vision = fmin(MAX_VISION, fmax(0.0, a->vision + BETA * (double)(MAX_NEIGHBORS - a->n_nb)));
    }
    free(rnd);
    free(luc);
    free(rot);
    free(pos);
}

/* lib.rs:47-78: positions are n rows (tx, ty, tz, w, x, y, z).  After each of
 * `steps` steps, row s of the outputs holds the post-move pose (w, x, y, z, tx,
 * ty, tz), luciferin, scoring, vision range and neighbour count. */
void orc_gso_run(int family, const orc_mol* rec, const orc_mol* lig, const orc_field* f,
                 const double* positions, int n, uint64_t seed, int steps, int nthreads,
                 double* out_pose, double* out_luc, double* out_score, double* out_vision,
                 int* out_nnb) {
    gw* g = calloc((size_t)n, sizeof(gw));
    for (int i = 0; i < n; ++i) {
        memcpy(g[i].t, positions + 7 * i, sizeof g[i].t);
        memcpy(g[i].q, positions + 7 * i + 3, sizeof g[i].q);
        g[i].luciferin = 5.0;
        g[i].vision = 0.2;
        g[i].nb = malloc(sizeof(int) * (size_t)n);
        g[i].prob = malloc(sizeof(double) * (size_t)n);
    }
    orc_rng rng;
    rng_seed(&rng, seed);
    for (int s = 0; s < steps; ++s) {
        update_luciferin(family, rec, lig, f, g, n, nthreads);
        movement_phase(g, n, &rng);
        for (int i = 0; i < n; ++i) {
            size_t k = (size_t)s * n + i;
            double* p = out_pose + 7 * k;
            p[0] = g[i].q[0]; p[1] = g[i].q[1]; p[2] = g[i].q[2]; p[3] = g[i].q[3];
            p[4] = g[i].t[0]; p[5] = g[i].t[1]; p[6] = g[i].t[2];
            out_luc[k] = g[i].luciferin;
            out_score[k] = g[i].scoring;
            out_vision[k] = g[i].vision;
            out_nnb[k] = g[i].n_nb;
        }
    }
    for (int i = 0; i < n; ++i) { free(g[i].nb); free(g[i].prob); }
    free(g);
}
