/*
 * lkl_oracle.c — CPU restatement (f64) of the LKlight reference CPU grid path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, the smoke() entry and
 * bench.py's cpu_baseline leg load this library; the product never does.
 *
 * Written from the reference's source text; each function cites the file:line
 * it follows.  Built with -ffp-contract=off so every a*b+c is a rounded multiply
 * followed by a rounded add, as the Rust source evaluates it.
 *
 * Parity status of each function (how it is pinned, see tests/test_oracle_*.py):
 *   orc_field_*             closed form of grid_dna.rs:216-232 + brute-force node sums
 *   orc_crec_build          27-cell set == all pairs (invariant) + brute force
 *   orc_nearcells_*         nearcell.rs:119-159 invariant (xorshift seed 7, 300 pts)
 *   orc_transform           qt.rs KATs (rotation of (1,0,0)), identity exactness
 *   orc_dna_energy_grid     1AZP KAT dna.rs:1149-1151 reproduced to its last digit
 *   orc_pydock_energy_exact 1AZP KAT pydock.rs:578-582 (1e-8)
 *   orc_*_energy_grid       cell path == brute force (orc_bruteforce_parts) to 1e-12,
 *                           grid vs exact 30 A sum (dna.rs:1154-1179 tolerance 2 %)
 *                           vdw grid == brute-force exact (vdw.rs:351-378, 1e-9)
 *   orc_cpydock desolvation brute-force min over all pairs, closed form of g()
 */
#include "lkl_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ── constants (dna.rs:20-42, pydock.rs:14-24, vdw.rs:10-12, cpydock.rs:36-45,
 *    grid_dna.rs:25-31, constants.rs:14-15) ─────────────────────────────────── */
#define EPSILON 4.0
#define FACTOR 332.0
#define ELEC_MAX_CUTOFF (1.0 * EPSILON / FACTOR)
#define ELEC_MIN_CUTOFF (-1.0 * EPSILON / FACTOR)
#define VDW_CUTOFF 1.0
#define CLASH_PENALTY_W 6.0
#define CLASH_PENALTY_FRAC 0.75
#define CLOSE_DIST2 100.0
#define VDW_DIST_CUTOFF2 100.0
#define ELEC_DIST_CUTOFF 30.0
#define ELEC_DIST_CUTOFF2 900.0
#define FIELD_RMAX 30.0
#define FIELD_RMIN 10.0
#define SPACING 0.5
#define VDW_WEIGHT 0.1
#define HUGE_DIST 10000.0

static double clampd(double x, double lo, double hi) {
    /* Rust f64::clamp */
    if (x < lo) return lo;
    if (x > hi) return hi;
    return x;
}

static double powi6(double x) { double x2 = x * x; return x2 * x2 * x2; }
static double powi3(double x) { return x * x * x; }

/* ── far field ─────────────────────────────────────────────────────────────── */

/* grid_dna.rs:44-61 */
int64_t orc_field_dims(const double* xyz, int n, double origin[3], int64_t dims[3]) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < n; ++k)
        for (int a = 0; a < 3; ++a) {
            double c = xyz[3 * k + a];
            lo[a] = fmin(lo[a], c);
            hi[a] = fmax(hi[a], c);
        }
    for (int a = 0; a < 3; ++a) {
        lo[a] = floor(lo[a] - FIELD_RMAX - 2.0 * SPACING);
        hi[a] = ceil(hi[a] + FIELD_RMAX + 2.0 * SPACING);
        origin[a] = lo[a];
        dims[a] = (int64_t)((hi[a] - lo[a]) / SPACING) + 1;
    }
    return dims[0] * dims[1] * dims[2];
}

/* grid_dna.rs:62-148 — per-atom shell scatter with the x-slab pruning; the
 * write set is { node : dyz2 < r2max and r2min < d2 <= r2max }, accumulated in
 * f32 in atom order. */
void orc_field_fill(const double* xyz, const double* q, int n,
                    const double origin[3], const int64_t dims[3], float* phi) {
    const double s = SPACING;
    const double* lo = origin;
    const int64_t spread = (int64_t)ceil(FIELD_RMAX / s) + 2;
    const double r2min = FIELD_RMIN * FIELD_RMIN;
    const double r2max = FIELD_RMAX * FIELD_RMAX;
    const int64_t nx = dims[0], ny = dims[1], nz = dims[2];
    memset(phi, 0, sizeof(float) * (size_t)(nx * ny * nz));
    for (int k = 0; k < n; ++k) {
        const double* c = xyz + 3 * k;
        double qk = q[k];
        if (qk == 0.0) continue;
        int64_t g[3], gi[3], gh[3];
        int64_t nn[3] = {nx, ny, nz};
        for (int a = 0; a < 3; ++a) {
            g[a] = (int64_t)floor((c[a] - lo[a]) / s);
            gi[a] = g[a] - spread; if (gi[a] < 0) gi[a] = 0;
            gh[a] = g[a] + spread; if (gh[a] > nn[a] - 1) gh[a] = nn[a] - 1;
        }
        double cxi_f = (c[0] - lo[0]) / s;
        for (int64_t iz = gi[2]; iz <= gh[2]; ++iz) {
            double zz = lo[2] + (double)iz * s;
            double dz = zz - c[2];
            for (int64_t iy = gi[1]; iy <= gh[1]; ++iy) {
                double yy = lo[1] + (double)iy * s;
                double dy = yy - c[1];
                double dyz2 = dy * dy + dz * dz;
                if (dyz2 >= r2max) continue;
                double b = sqrt(r2max - dyz2);
                int64_t xl = (int64_t)floor(cxi_f - b / s) - 1;
                if (xl < gi[0]) xl = gi[0];
                int64_t xh = (int64_t)ceil(cxi_f + b / s) + 1;
                if (xh > gh[0]) xh = gh[0];
                if (xl > gh[0] || xh < gi[0]) continue;
                int64_t base = (iz * ny + iy) * nx;
                for (int64_t ix = xl; ix <= xh; ++ix) {
                    double dx = lo[0] + (double)ix * s - c[0];
                    double d2 = dx * dx + dyz2;
                    if (d2 > r2min && d2 <= r2max) phi[base + ix] += (float)(qk / d2);
                }
            }
        }
    }
}

static float phi_at(const orc_field* f, int64_t ix, int64_t iy, int64_t iz) {
    return f->phi[(iz * f->n[1] + iy) * f->n[0] + ix];
}

/* grid_dna.rs:159-197 — f64 trilinear with the fractional offsets rounded to f32. */
double orc_field_sample(const orc_field* f, double x, double y, double z) {
    double s = f->spacing;
    double fx = (x - f->origin[0]) / s;
    double fy = (y - f->origin[1]) / s;
    double fz = (z - f->origin[2]) / s;
    double ix = floor(fx), iy = floor(fy), iz = floor(fz);
    if (ix < 0.0 || iy < 0.0 || iz < 0.0) return 0.0;
    int64_t ix0 = (int64_t)ix, iy0 = (int64_t)iy, iz0 = (int64_t)iz;
    if (ix0 + 1 >= f->n[0] || iy0 + 1 >= f->n[1] || iz0 + 1 >= f->n[2]) return 0.0;
    double tx = (double)(float)(fx - ix);
    double ty = (double)(float)(fy - iy);
    double tz = (double)(float)(fz - iz);
    double c000 = phi_at(f, ix0, iy0, iz0);
    double c100 = phi_at(f, ix0 + 1, iy0, iz0);
    double c010 = phi_at(f, ix0, iy0 + 1, iz0);
    double c110 = phi_at(f, ix0 + 1, iy0 + 1, iz0);
    double c001 = phi_at(f, ix0, iy0, iz0 + 1);
    double c101 = phi_at(f, ix0 + 1, iy0, iz0 + 1);
    double c011 = phi_at(f, ix0, iy0 + 1, iz0 + 1);
    double c111 = phi_at(f, ix0 + 1, iy0 + 1, iz0 + 1);
    double c00 = c000 * (1.0 - tx) + c100 * tx;
    double c10 = c010 * (1.0 - tx) + c110 * tx;
    double c01 = c001 * (1.0 - tx) + c101 * tx;
    double c11 = c011 * (1.0 - tx) + c111 * tx;
    double c0 = c00 * (1.0 - ty) + c10 * ty;
    double c1 = c01 * (1.0 - ty) + c11 * ty;
    return c0 * (1.0 - tz) + c1 * tz;
}

static int field_in_box(const orc_field* f, double x, double y, double z) {
    double s = f->spacing;
    double ix = floor((x - f->origin[0]) / s), iy = floor((y - f->origin[1]) / s),
           iz = floor((z - f->origin[2]) / s);
    if (ix < 0.0 || iy < 0.0 || iz < 0.0) return 0;
    return (int64_t)ix + 1 < f->n[0] && (int64_t)iy + 1 < f->n[1] && (int64_t)iz + 1 < f->n[2];
}

/* grid_dna.rs:201-209 */
static double far_field_energy(const orc_field* f, const double* lig_c, const double* q,
                               int nl, int64_t* n_far, double* f_abs) {
    double e = 0.0;
    for (int j = 0; j < nl; ++j) {
        if (q[j] != 0.0) {
            const double* c = lig_c + 3 * j;
            const double t = q[j] * orc_field_sample(f, c[0], c[1], c[2]);
            e += t;
            if (f_abs) *f_abs += fabs(t);
            if (n_far && field_in_box(f, c[0], c[1], c[2])) ++*n_far;
        }
    }
    return e;
}

/* ── cell lists ────────────────────────────────────────────────────────────── */

/* gpu_score.rs:73-145 (identical logic in gpu_family.rs:82-153). */
int orc_crec_build(const double* xyz, int n, int* ncxyz, float* c_origin,
                   int* cell_start, int* cell_atoms) {
    const double cell = 10.0;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < n; ++k)
        for (int a = 0; a < 3; ++a) {
            lo[a] = fmin(lo[a], xyz[3 * k + a]);
            hi[a] = fmax(hi[a], xyz[3 * k + a]);
        }
    for (int a = 0; a < 3; ++a) lo[a] = floor(lo[a] - 0.01);
    int nc[3];
    for (int a = 0; a < 3; ++a) {
        nc[a] = (int)floor((hi[a] - lo[a]) / cell) + 1;
        if (nc[a] < 1) nc[a] = 1;
    }
    int ncells = (nc[0] + 1) * (nc[1] + 1) * (nc[2] + 1);
    for (int a = 0; a < 3; ++a) { ncxyz[a] = nc[a]; c_origin[a] = (float)lo[a]; }
    if (!cell_start) return ncells + 1;
    int* cell_of = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    memset(cell_start, 0, sizeof(int) * (size_t)(ncells + 1));
    for (int k = 0; k < n; ++k) {
        int64_t ii[3];
        for (int a = 0; a < 3; ++a) {
            ii[a] = (int64_t)floor((xyz[3 * k + a] - lo[a]) / cell);
            if (ii[a] < 0) ii[a] = 0;
            if (ii[a] > nc[a] - 1) ii[a] = nc[a] - 1;
        }
        int c = (int)((ii[2] * nc[1] + ii[1]) * nc[0] + ii[0]);
        cell_of[k] = c;
        cell_start[c] += 1;
    }
    int acc = 0;
    for (int c = 0; c < ncells; ++c) { int m = cell_start[c]; cell_start[c] = acc; acc += m; }
    cell_start[ncells] = acc;
    int* cursor = (int*)malloc(sizeof(int) * (size_t)(ncells + 1));
    memcpy(cursor, cell_start, sizeof(int) * (size_t)(ncells + 1));
    for (int k = 0; k < n; ++k) cell_atoms[cursor[cell_of[k]]++] = k;
    free(cursor);
    free(cell_of);
    return ncells + 1;
}

/* nearcell.rs:22-79 */
typedef struct {
    double cell, lo[3];
    int64_t n[3];
    int64_t* cell_start;
    int64_t* cell_atoms;
    int n_atoms;
} nearcells;

static void nearcells_build(nearcells* nc, const double* xyz, int n, double cell) {
    nc->n_atoms = n;
    if (n == 0 || cell <= 0.0) {
        nc->cell = cell > 1e-6 ? cell : 1e-6;
        nc->lo[0] = nc->lo[1] = nc->lo[2] = 0.0;
        nc->n[0] = nc->n[1] = nc->n[2] = 1;
        nc->cell_start = (int64_t*)calloc(2, sizeof(int64_t));
        nc->cell_atoms = NULL;
        return;
    }
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < n; ++k)
        for (int d = 0; d < 3; ++d) {
            double c = xyz[3 * k + d];
            if (c < mn[d]) mn[d] = c;
            if (c > mx[d]) mx[d] = c;
        }
    nc->cell = cell;
    for (int d = 0; d < 3; ++d) {
        nc->lo[d] = mn[d] - 1.0;
        double span = mx[d] - nc->lo[d] + 1.0;
        nc->n[d] = (int64_t)ceil(span / cell);
        if (nc->n[d] < 1) nc->n[d] = 1;
    }
    int64_t total = nc->n[0] * nc->n[1] * nc->n[2];
    nc->cell_start = (int64_t*)calloc((size_t)total + 1, sizeof(int64_t));
    int64_t* key = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    for (int k = 0; k < n; ++k) {
        int64_t ii[3];
        for (int d = 0; d < 3; ++d) {
            ii[d] = (int64_t)floor((xyz[3 * k + d] - nc->lo[d]) / cell);
            if (ii[d] < 0) ii[d] = 0;
            if (ii[d] > nc->n[d] - 1) ii[d] = nc->n[d] - 1;
        }
        key[k] = (ii[2] * nc->n[1] + ii[1]) * nc->n[0] + ii[0];
        nc->cell_start[key[k] + 1] += 1;
    }
    for (int64_t i = 0; i < total; ++i) nc->cell_start[i + 1] += nc->cell_start[i];
    int64_t* cursor = (int64_t*)malloc(sizeof(int64_t) * (size_t)(total + 1));
    memcpy(cursor, nc->cell_start, sizeof(int64_t) * (size_t)(total + 1));
    nc->cell_atoms = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    for (int k = 0; k < n; ++k) nc->cell_atoms[cursor[key[k]]++] = k;
    free(cursor);
    free(key);
}

static void nearcells_free(nearcells* nc) {
    free(nc->cell_start);
    free(nc->cell_atoms);
}

/* nearcell.rs:86-112 — fills out[] with the candidate atom ids of the 27-cell
 * window around (x, y, z) in enumeration order; returns the count. */
static int nearcells_window(const nearcells* nc, double x, double y, double z, int* out) {
    if (nc->n_atoms == 0) return 0;
    int64_t c[3];
    double p[3] = {x, y, z};
    for (int d = 0; d < 3; ++d) {
        c[d] = (int64_t)floor((p[d] - nc->lo[d]) / nc->cell);
        if (c[d] < 0) c[d] = 0;
        if (c[d] > nc->n[d] - 1) c[d] = nc->n[d] - 1;
    }
    int m = 0;
    for (int64_t dz = -1; dz <= 1; ++dz) {
        int64_t iz = c[2] + dz;
        if (iz < 0 || iz >= nc->n[2]) continue;
        for (int64_t dy = -1; dy <= 1; ++dy) {
            int64_t iy = c[1] + dy;
            if (iy < 0 || iy >= nc->n[1]) continue;
            for (int64_t dx = -1; dx <= 1; ++dx) {
                int64_t ix = c[0] + dx;
                if (ix < 0 || ix >= nc->n[0]) continue;
                int64_t cell = (iz * nc->n[1] + iy) * nc->n[0] + ix;
                for (int64_t p2 = nc->cell_start[cell]; p2 < nc->cell_start[cell + 1]; ++p2)
                    out[m++] = (int)nc->cell_atoms[p2];
            }
        }
    }
    return m;
}

int64_t orc_nearcells_pair_count(const double* xyz, int n, double cell) {
    nearcells nc;
    nearcells_build(&nc, xyz, n, cell);
    double cut2 = cell * cell;
    unsigned char* seen = (unsigned char*)calloc((size_t)n * (size_t)n, 1);
    int* win = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int64_t count = 0;
    for (int j = 0; j < n; ++j) {
        const double* c = xyz + 3 * j;
        int m = nearcells_window(&nc, c[0], c[1], c[2], win);
        for (int t = 0; t < m; ++t) {
            int i = win[t];
            double dx = c[0] - xyz[3 * i], dy = c[1] - xyz[3 * i + 1], dz = c[2] - xyz[3 * i + 2];
            if (i != j && dx * dx + dy * dy + dz * dz <= cut2) {
                int a = i < j ? i : j, b = i < j ? j : i;
                if (!seen[(size_t)a * n + b]) { seen[(size_t)a * n + b] = 1; ++count; }
            }
        }
    }
    free(win);
    free(seen);
    nearcells_free(&nc);
    return count;
}

/* ── pose transform (qt.rs:67-74, qt.rs:122-128, dna.rs:868-873) ─────────── */
void orc_transform(const double* base, int nl, const double* pose7, double* out) {
    double w = pose7[0], x = pose7[1], y = pose7[2], z = pose7[3];
    double m[3][3] = {
        {1. - 2. * (y * y + z * z), 2. * (x * y - w * z), 2. * (x * z + w * y)},
        {2. * (x * y + w * z), 1. - 2. * (x * x + z * z), 2. * (y * z - w * x)},
        {2. * (x * z - w * y), 2. * (y * z + w * x), 1. - 2. * (x * x + y * y)},
    };
    for (int j = 0; j < nl; ++j) {
        const double* v = base + 3 * j;
        double r0 = m[0][0] * v[0] + m[0][1] * v[1] + m[0][2] * v[2];
        double r1 = m[1][0] * v[0] + m[1][1] * v[1] + m[1][2] * v[2];
        double r2 = m[2][0] * v[0] + m[2][1] * v[1] + m[2][2] * v[2];
        out[3 * j] = r0 + pose7[4];
        out[3 * j + 1] = r1 + pose7[5];
        out[3 * j + 2] = r2 + pose7[6];
    }
}

/* ── per-pair terms ─────────────────────────────────────────────────────────── */

/* dna.rs:951-969: clamped electrostatics + capped LJ (sqrt-product epsilon) +
 * heavy-atom clash. */
/* Rounding magnitudes of one pair (m[0] += |elec|, m[1] += LJ magnitude,
 * m[2] += clash weight * distance): not part of the reference's sums, they
 * size the FP32 error bound of a device evaluation (orc_parts). */
static void mag_elec(double* m, double e) { if (m) m[0] += fabs(e); }
static void mag_lj(double* m, double eps, double p6, double k) {
    if (m && k <= 2.0 * VDW_CUTOFF) m[1] += fabs(eps) * (p6 * p6 + 2.0 * p6);
}

static void dna_pair(double d2, double qj, double qi, double svj, double svi, double rj,
                     double ri, int hj, int hi, double* e, double* v, double* m) {
    double el = clampd(qj * qi / d2, ELEC_MIN_CUTOFF, ELEC_MAX_CUTOFF);
    *e += el;
    mag_elec(m, el);
    double vdw_e = svj * svi;
    double vdw_r = rj + ri;
    double p6 = powi6(vdw_r) / powi3(d2);
    double k = vdw_e * (p6 * p6 - 2.0 * p6);
    mag_lj(m, vdw_e, p6, k);
    double vp = fmin(k, VDW_CUTOFF);
    if (hj && hi) {
        double d = sqrt(d2);
        double dmin = CLASH_PENALTY_FRAC * vdw_r;
        if (d < dmin) vp += CLASH_PENALTY_W * (dmin - d);
        if (m && d < dmin * (1.0 + 1e-6)) m[2] += CLASH_PENALTY_W * dmin;
    }
    *v += vp;
}

/* vdw.rs:131-137: epsilon = sqrt(eps_i * eps_j), no clash, no electrostatics. */
static void vdw_pair(double d2, double epsj, double epsi, double rj, double ri, double* v, double* m) {
    double vdw_energy = sqrt(epsi * epsj);
    double vdw_radius = ri + rj;
    double p6 = powi6(vdw_radius) / powi3(d2);
    double k = vdw_energy * (p6 * p6 - 2.0 * p6);
    mag_lj(m, vdw_energy, p6, k);
    if (k > VDW_CUTOFF) k = VDW_CUTOFF;
    *v += k;
}

/* pydock.rs:421-428 */
static void pydock_pair(double d2, double qj, double qi, double svj, double svi, double rj,
                        double ri, double* e, double* v, double* m) {
    double el = clampd(qj * qi / d2, ELEC_MIN_CUTOFF, ELEC_MAX_CUTOFF);
    *e += el;
    mag_elec(m, el);
    double vdw_e = svj * svi;
    double vdw_r = rj + ri;
    double p6 = powi6(vdw_r) / powi3(d2);
    double k = vdw_e * (p6 * p6 - 2.0 * p6);
    mag_lj(m, vdw_e, p6, k);
    *v += fmin(k, VDW_CUTOFF);
}

/* cpydock.rs:628-637 (powf rather than powi, epsilon = sqrt(eps_i eps_j)). */
static void cpydock_pair(double d2, double qj, double qi, double epsj, double epsi, double rj,
                         double ri, double* e, double* v, double* m) {
    double el = clampd(qj * qi / d2, ELEC_MIN_CUTOFF, ELEC_MAX_CUTOFF);
    *e += el;
    mag_elec(m, el);
    double vdw_e = sqrt(epsj * epsi);
    double vdw_r = rj + ri;
    double p6 = pow(vdw_r, 6.0) / pow(d2, 3.0);
    double k = vdw_e * (p6 * p6 - 2.0 * p6);
    mag_lj(m, vdw_e, p6, k);
    *v += k > VDW_CUTOFF ? VDW_CUTOFF : k;
}

/* cpydock.rs:600-601 — C-binary flag view of the int64 hydrogen array. */
static int cpy_flag(const int* hyd, int i) { return (i % 2 == 0) ? (hyd[i / 2] != 0) : 0; }

/* cpydock.rs:648-665 — per-atom desolvation with the clamp outside the window
 * test (an asa < 0 atom contributes asa * des). */
static double solv_term(double min_d2, double asa, double des) {
    double solv = 0.0;
    if (min_d2 <= 6.4 * 6.4 && min_d2 > 0.0 && asa > 0.0) solv = -10.0 * sqrt(min_d2) + 65.0;
    if (solv > asa) solv = asa;
    return solv * des;
}

/* ── families ─────────────────────────────────────────────────────────────── */

/* The receptor's 10 A CudaReceptor cell list, built once per call site. */
typedef struct {
    int nc[3];
    float c_origin[3];
    int* cell_start;
    int* cell_atoms;
} crec;

static void crec_build(crec* c, const double* xyz, int n) {
    int len = orc_crec_build(xyz, n, c->nc, c->c_origin, NULL, NULL);
    c->cell_start = (int*)malloc(sizeof(int) * (size_t)len);
    c->cell_atoms = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    orc_crec_build(xyz, n, c->nc, c->c_origin, c->cell_start, c->cell_atoms);
}

static void crec_free(crec* c) { free(c->cell_start); free(c->cell_atoms); }

/* dna.rs:912-980 — the CPU near loop over the CudaReceptor cell list with
 * clamped centre cell. */
static void dna_grid_with(const crec* cr, const orc_mol* rec, const orc_mol* lig,
                          const orc_field* f, const double* pose7, orc_parts* out) {
    int nl = lig->n;
    double* lc = (double*)malloc(sizeof(double) * 3 * (size_t)(nl > 0 ? nl : 1));
    orc_transform(lig->xyz, nl, pose7, lc);
    double csp = 10.0, cox = cr->c_origin[0], coy = cr->c_origin[1], coz = cr->c_origin[2];
    int64_t ncx = cr->nc[0], ncy = cr->nc[1], ncz = cr->nc[2];
    double e_sum = 0.0, v_sum = 0.0, mag[3] = {0.0, 0.0, 0.0};
    int64_t n_near = 0, n_cand = 0;
    for (int j = 0; j < nl; ++j) {
        double x = lc[3 * j], y = lc[3 * j + 1], z = lc[3 * j + 2];
        int64_t cxi = (int64_t)floor((x - cox) / csp), cyi = (int64_t)floor((y - coy) / csp),
                czi = (int64_t)floor((z - coz) / csp);
        cxi = cxi < 0 ? 0 : (cxi > ncx - 1 ? ncx - 1 : cxi);
        cyi = cyi < 0 ? 0 : (cyi > ncy - 1 ? ncy - 1 : cyi);
        czi = czi < 0 ? 0 : (czi > ncz - 1 ? ncz - 1 : czi);
        for (int64_t dz = -1; dz <= 1; ++dz) {
            int64_t iz = czi + dz;
            if (iz < 0 || iz >= ncz) continue;
            for (int64_t dy = -1; dy <= 1; ++dy) {
                int64_t iy = cyi + dy;
                if (iy < 0 || iy >= ncy) continue;
                for (int64_t dx = -1; dx <= 1; ++dx) {
                    int64_t ix = cxi + dx;
                    if (ix < 0 || ix >= ncx) continue;
                    int64_t cell = (iz * ncy + iy) * ncx + ix;
                    for (int p = cr->cell_start[cell]; p < cr->cell_start[cell + 1]; ++p) {
                        int i = cr->cell_atoms[p];
                        const double* r = rec->xyz + 3 * i;
                        double ddx = x - r[0], ddy = y - r[1], ddz = z - r[2];
                        double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
                        ++n_cand;
                        if (d2 <= CLOSE_DIST2) {
                            ++n_near;
                            dna_pair(d2, lig->ele[j], rec->ele[i], lig->seps[j], rec->seps[i],
                                     lig->rad[j], rec->rad[i], lig->heavy[j], rec->heavy[i],
                                     &e_sum, &v_sum, mag);
                        }
                    }
                }
            }
        }
    }
    int64_t n_far = 0;
    double f_abs = 0.0;
    double e_field = f ? far_field_energy(f, lc, lig->ele, nl, &n_far, &f_abs) : 0.0;
    out->e_raw = e_sum + e_field;
    out->v = v_sum;
    out->s = 0.0;
    out->n_near = n_near;
    out->n_far = n_far;
    out->n_cand = n_cand;
    out->e_abs = mag[0];
    out->f_abs = f_abs;
    out->v_abs = mag[1];
    out->c_abs = mag[2];
    out->s_abs = 0.0;
    free(lc);
}

void orc_dna_energy_grid(const orc_mol* rec, const orc_mol* lig, const orc_field* f,
                         const double* pose7, orc_parts* out) {
    crec cr;
    crec_build(&cr, rec->xyz, rec->n);
    dna_grid_with(&cr, rec, lig, f, pose7, out);
    crec_free(&cr);
}

/* dna.rs:885-995 (and cpydock.rs:509,639): the interface flags of one pose.
 * The CPU near pass visits every pair of the 27 10-A cells around a ligand
 * atom and flags both atoms when the f64 squared distance is at most
 * INTERFACE_CUTOFF^2 (constants.rs:14-15); 3.9 A pairs always lie in that
 * window, so this is the set of all pairs within 3.9 A. */
void orc_interface_flags(const orc_mol* rec, const orc_mol* lig, const double* pose7,
                         unsigned char* iface_r, unsigned char* iface_l) {
    crec cr;
    crec_build(&cr, rec->xyz, rec->n);
    int nl = lig->n;
    double* lc = (double*)malloc(sizeof(double) * 3 * (size_t)(nl > 0 ? nl : 1));
    orc_transform(lig->xyz, nl, pose7, lc);
    memset(iface_r, 0, (size_t)rec->n);
    memset(iface_l, 0, (size_t)nl);
    const double icut2 = 3.9 * 3.9;
    double csp = 10.0, cox = cr.c_origin[0], coy = cr.c_origin[1], coz = cr.c_origin[2];
    int64_t ncx = cr.nc[0], ncy = cr.nc[1], ncz = cr.nc[2];
    for (int j = 0; j < nl; ++j) {
        double x = lc[3 * j], y = lc[3 * j + 1], z = lc[3 * j + 2];
        int64_t cxi = (int64_t)floor((x - cox) / csp), cyi = (int64_t)floor((y - coy) / csp),
                czi = (int64_t)floor((z - coz) / csp);
        cxi = cxi < 0 ? 0 : (cxi > ncx - 1 ? ncx - 1 : cxi);
        cyi = cyi < 0 ? 0 : (cyi > ncy - 1 ? ncy - 1 : cyi);
        czi = czi < 0 ? 0 : (czi > ncz - 1 ? ncz - 1 : czi);
        for (int64_t dz = -1; dz <= 1; ++dz) {
            int64_t iz = czi + dz;
            if (iz < 0 || iz >= ncz) continue;
            for (int64_t dy = -1; dy <= 1; ++dy) {
                int64_t iy = cyi + dy;
                if (iy < 0 || iy >= ncy) continue;
                for (int64_t dx = -1; dx <= 1; ++dx) {
                    int64_t ix = cxi + dx;
                    if (ix < 0 || ix >= ncx) continue;
                    int64_t cell = (iz * ncy + iy) * ncx + ix;
                    for (int p = cr.cell_start[cell]; p < cr.cell_start[cell + 1]; ++p) {
                        int i = cr.cell_atoms[p];
                        const double* r = rec->xyz + 3 * i;
                        double ddx = x - r[0], ddy = y - r[1], ddz = z - r[2];
                        if (ddx * ddx + ddy * ddy + ddz * ddz <= icut2) {
                            iface_r[i] = 1;
                            iface_l[j] = 1;
                        }
                    }
                }
            }
        }
    }
    free(lc);
    crec_free(&cr);
}

/* scoring.rs:83-98: fraction of residues with at least one interface atom
 * (0 without residues); residue k = atoms[off[k] .. off[k+1]). */
double orc_satisfied_restraints(const unsigned char* iface, const int* off, const int* atoms, int n_res) {
    if (n_res <= 0) return 0.0;
    int num = 0;
    for (int k = 0; k < n_res; ++k)
        for (int q = off[k]; q < off[k + 1]; ++q)
            if (iface[atoms[q]] == 1) {
                ++num;
                break;
            }
    return (double)num / (double)n_res;
}

/* scoring.rs:100-110: fraction of membrane beads at the interface. */
double orc_membrane_intersection(const unsigned char* iface, const int* atoms, int n) {
    if (n <= 0) return 0.0;
    int num = 0;
    for (int k = 0; k < n; ++k) num += iface[atoms[k]];
    return (double)num / (double)n;
}

/* dna.rs:989-995 (cpydock.rs:545-549): score += perc_r*score + perc_l*score - penalty. */
double orc_restraint_fold(double score, double perc_r, double perc_l, double intersection) {
    double penalty = intersection > 0.0 ? 999.0 * intersection : 0.0;
    score += perc_r * score + perc_l * score - penalty;
    return score;
}

/* vdw.rs / pydock.rs / cpydock.rs share the NearCells near loop. */
static void nc_grid_with(int family, const nearcells* nc, const orc_mol* rec, const orc_mol* lig,
                         const orc_field* f, const double* pose7, orc_parts* out) {
    int nl = lig->n, nr = rec->n;
    double* lc = (double*)malloc(sizeof(double) * 3 * (size_t)(nl > 0 ? nl : 1));
    orc_transform(lig->xyz, nl, pose7, lc);
    int* win = (int*)malloc(sizeof(int) * (size_t)(nr > 0 ? nr : 1));
    double* min_rec = NULL;
    double* min_lig = NULL;
    if (family == ORC_CPYDOCK) {
        min_rec = (double*)malloc(sizeof(double) * (size_t)(nr > 0 ? nr : 1));
        min_lig = (double*)malloc(sizeof(double) * (size_t)(nl > 0 ? nl : 1));
        for (int i = 0; i < nr; ++i) min_rec[i] = HUGE_DIST;
        for (int j = 0; j < nl; ++j) min_lig[j] = HUGE_DIST;
    }
    double e_near = 0.0, v_sum = 0.0, mag[3] = {0.0, 0.0, 0.0};
    int64_t n_near = 0, n_cand = 0;
    for (int j = 0; j < nl; ++j) {
        double x = lc[3 * j], y = lc[3 * j + 1], z = lc[3 * j + 2];
        int m = nearcells_window(nc, x, y, z, win);
        int lf = family == ORC_CPYDOCK ? cpy_flag(lig->hyd, j) : 0;
        for (int t = 0; t < m; ++t) {
            int i = win[t];
            const double* r = rec->xyz + 3 * i;
            double ddx = x - r[0], ddy = y - r[1], ddz = z - r[2];
            double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
            ++n_cand;
            if (family == ORC_CPYDOCK && !cpy_flag(rec->hyd, i) && !lf) {
                if (d2 < min_rec[i]) min_rec[i] = d2;
                if (d2 < min_lig[j]) min_lig[j] = d2;
            }
            if (d2 <= VDW_DIST_CUTOFF2) {
                ++n_near;
                if (family == ORC_VDW)
                    vdw_pair(d2, lig->eps[j], rec->eps[i], lig->rad[j], rec->rad[i], &v_sum, mag);
                else if (family == ORC_PYDOCK)
                    pydock_pair(d2, lig->ele[j], rec->ele[i], lig->seps[j], rec->seps[i],
                                lig->rad[j], rec->rad[i], &e_near, &v_sum, mag);
                else
                    cpydock_pair(d2, lig->ele[j], rec->ele[i], lig->eps[j], rec->eps[i],
                                 lig->rad[j], rec->rad[i], &e_near, &v_sum, mag);
            }
        }
    }
    int64_t n_far = 0;
    double e_field = 0.0, f_abs = 0.0;
    if (family != ORC_VDW && f) e_field = far_field_energy(f, lc, lig->ele, nl, &n_far, &f_abs);
    double s = 0.0, s_abs = 0.0;
    if (family == ORC_CPYDOCK) {
        for (int i = 0; i < nr; ++i) {
            s += solv_term(min_rec[i], rec->asa[i], rec->des[i]);
            if (min_rec[i] <= 6.4 * 6.4) s_abs += fabs(rec->des[i]);
        }
        for (int j = 0; j < nl; ++j) s += solv_term(min_lig[j], lig->asa[j], lig->des[j]);
    }
    out->e_raw = e_near + e_field;
    out->v = v_sum;
    out->s = s;
    out->n_near = n_near;
    out->n_far = n_far;
    out->n_cand = n_cand;
    out->e_abs = mag[0];
    out->f_abs = f_abs;
    out->v_abs = mag[1];
    out->c_abs = mag[2];
    out->s_abs = s_abs;
    free(min_rec);
    free(min_lig);
    free(win);
    free(lc);
}

void orc_vdw_energy_grid(const orc_mol* rec, const orc_mol* lig, const double* pose7,
                         orc_parts* out) {
    nearcells nc;
    nearcells_build(&nc, rec->xyz, rec->n, 10.0);
    nc_grid_with(ORC_VDW, &nc, rec, lig, NULL, pose7, out);
    nearcells_free(&nc);
}

void orc_pydock_energy_grid(const orc_mol* rec, const orc_mol* lig, const orc_field* f,
                            const double* pose7, orc_parts* out) {
    nearcells nc;
    nearcells_build(&nc, rec->xyz, rec->n, 10.0);
    nc_grid_with(ORC_PYDOCK, &nc, rec, lig, f, pose7, out);
    nearcells_free(&nc);
}

void orc_cpydock_energy_grid(const orc_mol* rec, const orc_mol* lig, const orc_field* f,
                             const double* pose7, orc_parts* out) {
    nearcells nc;
    nearcells_build(&nc, rec->xyz, rec->n, 10.0);
    nc_grid_with(ORC_CPYDOCK, &nc, rec, lig, f, pose7, out);
    nearcells_free(&nc);
}

/* gpu_score.rs:326-331 / gpu_family.rs:283-288 fold; cpydock.rs:669 /
 * cpydock.rs:789-791 for CPYDOCK. */
double orc_fold_score(int family, double e_raw, double v, double s) {
    double e = e_raw * FACTOR / EPSILON;
    if (family == ORC_CPYDOCK) return (e + VDW_WEIGHT * v - s) * -1.0;
    return -(e + v);
}

/* ── batched driver ─────────────────────────────────────────────────────────── */
typedef struct {
    int family;
    const orc_mol *rec, *lig;
    const orc_field* f;
    const crec* cr;
    const nearcells* nc;
    const double* poses;
    int n_pose, tid, nthreads;
    double *oe, *ov, *os;
    int64_t *onear, *ofar;
    double* omag;
} batch_job;

static void* batch_worker(void* arg) {
    batch_job* b = (batch_job*)arg;
    for (int k = b->tid; k < b->n_pose; k += b->nthreads) {
        orc_parts p;
        const double* pose = b->poses + 7 * (size_t)k;
        if (b->family == ORC_DNA) dna_grid_with(b->cr, b->rec, b->lig, b->f, pose, &p);
        else nc_grid_with(b->family, b->nc, b->rec, b->lig, b->f, pose, &p);
        if (b->oe) b->oe[k] = p.e_raw;
        if (b->ov) b->ov[k] = p.v;
        if (b->os) b->os[k] = p.s;
        if (b->onear) b->onear[k] = p.n_near;
        if (b->ofar) b->ofar[k] = p.n_far;
        if (b->omag) {
            double* m = b->omag + 5 * (size_t)k;
            m[0] = p.e_abs;
            m[1] = p.f_abs;
            m[2] = p.v_abs;
            m[3] = p.c_abs;
            m[4] = p.s_abs;
        }
    }
    return NULL;
}

void orc_batch_energy_grid(int family, const orc_mol* rec, const orc_mol* lig,
                           const orc_field* f, const double* poses, int n_pose,
                           double* out_e, double* out_v, double* out_s,
                           int64_t* out_near, int64_t* out_far, int nthreads) {
    orc_batch_energy_grid_mag(family, rec, lig, f, poses, n_pose, out_e, out_v, out_s, out_near, out_far, NULL,
                              nthreads);
}

void orc_batch_energy_grid_mag(int family, const orc_mol* rec, const orc_mol* lig,
                               const orc_field* f, const double* poses, int n_pose,
                               double* out_e, double* out_v, double* out_s,
                               int64_t* out_near, int64_t* out_far, double* out_mag, int nthreads) {
    crec cr;
    nearcells nc;
    memset(&cr, 0, sizeof(cr));
    memset(&nc, 0, sizeof(nc));
    if (family == ORC_DNA) crec_build(&cr, rec->xyz, rec->n);
    else nearcells_build(&nc, rec->xyz, rec->n, 10.0);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    batch_job jobs[256];
    pthread_t th[256];
    for (int t = 0; t < nthreads; ++t) {
        batch_job b = {family, rec, lig, f, &cr, &nc, poses, n_pose, t, nthreads,
                       out_e, out_v, out_s, out_near, out_far, out_mag};
        jobs[t] = b;
    }
    for (int t = 1; t < nthreads; ++t) pthread_create(&th[t], NULL, batch_worker, &jobs[t]);
    batch_worker(&jobs[0]);
    for (int t = 1; t < nthreads; ++t) pthread_join(th[t], NULL);
    if (family == ORC_DNA) crec_free(&cr);
    else nearcells_free(&nc);
}

/* ── brute-force FP64 re-evaluation of the same sum (no cells) ─────────────── */
void orc_bruteforce_parts(int family, const orc_mol* rec, const orc_mol* lig,
                          const orc_field* f, const double* pose7, orc_parts* out) {
    int nl = lig->n, nr = rec->n;
    double* lc = (double*)malloc(sizeof(double) * 3 * (size_t)(nl > 0 ? nl : 1));
    orc_transform(lig->xyz, nl, pose7, lc);
    double e = 0.0, v = 0.0;
    int64_t n_near = 0;
    double* min_rec = (double*)malloc(sizeof(double) * (size_t)(nr > 0 ? nr : 1));
    double* min_lig = (double*)malloc(sizeof(double) * (size_t)(nl > 0 ? nl : 1));
    for (int i = 0; i < nr; ++i) min_rec[i] = HUGE_DIST;
    for (int j = 0; j < nl; ++j) min_lig[j] = HUGE_DIST;
    for (int j = 0; j < nl; ++j) {
        double x = lc[3 * j], y = lc[3 * j + 1], z = lc[3 * j + 2];
        for (int i = 0; i < nr; ++i) {
            const double* r = rec->xyz + 3 * i;
            double ddx = x - r[0], ddy = y - r[1], ddz = z - r[2];
            double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
            if (family == ORC_CPYDOCK && !cpy_flag(rec->hyd, i) && !cpy_flag(lig->hyd, j) &&
                d2 <= 100.0) {
                if (d2 < min_rec[i]) min_rec[i] = d2;
                if (d2 < min_lig[j]) min_lig[j] = d2;
            }
            if (d2 > 100.0) continue;
            ++n_near;
            if (family == ORC_DNA)
                dna_pair(d2, lig->ele[j], rec->ele[i], lig->seps[j], rec->seps[i], lig->rad[j],
                         rec->rad[i], lig->heavy[j], rec->heavy[i], &e, &v, NULL);
            else if (family == ORC_VDW)
                vdw_pair(d2, lig->eps[j], rec->eps[i], lig->rad[j], rec->rad[i], &v, NULL);
            else if (family == ORC_PYDOCK)
                pydock_pair(d2, lig->ele[j], rec->ele[i], lig->seps[j], rec->seps[i],
                            lig->rad[j], rec->rad[i], &e, &v, NULL);
            else
                cpydock_pair(d2, lig->ele[j], rec->ele[i], lig->eps[j], rec->eps[i],
                             lig->rad[j], rec->rad[i], &e, &v, NULL);
        }
    }
    int64_t n_far = 0;
    if (family != ORC_VDW && f) e += far_field_energy(f, lc, lig->ele, nl, &n_far, NULL);
    double s = 0.0;
    if (family == ORC_CPYDOCK) {
        for (int i = 0; i < nr; ++i) s += solv_term(min_rec[i], rec->asa[i], rec->des[i]);
        for (int j = 0; j < nl; ++j) s += solv_term(min_lig[j], lig->asa[j], lig->des[j]);
    }
    out->e_raw = e;
    out->v = v;
    out->s = s;
    out->n_near = n_near;
    out->n_far = n_far;
    out->n_cand = (int64_t)nr * nl;
    free(min_rec);
    free(min_lig);
    free(lc);
}

/* dna.rs:668-819 (restraint-free): every pair within 30 A contributes clamped
 * electrostatics, pairs within 10 A the capped LJ + clash. */
double orc_dna_energy_exact(const orc_mol* rec, const orc_mol* lig, const double* pose7) {
    int nl = lig->n, nr = rec->n;
    double* lc = (double*)malloc(sizeof(double) * 3 * (size_t)(nl > 0 ? nl : 1));
    orc_transform(lig->xyz, nl, pose7, lc);
    double te = 0.0, tv = 0.0;
    for (int i = 0; i < nr; ++i) {
        const double* r = rec->xyz + 3 * i;
        double ei = 0.0, vi = 0.0;
        for (int j = 0; j < nl; ++j) {
            double dx = r[0] - lc[3 * j], dy = r[1] - lc[3 * j + 1], dz = r[2] - lc[3 * j + 2];
            double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 <= ELEC_DIST_CUTOFF2)
                ei += clampd(rec->ele[i] * lig->ele[j] / d2, ELEC_MIN_CUTOFF, ELEC_MAX_CUTOFF);
            if (d2 <= VDW_DIST_CUTOFF2) {
                double vdw_e = rec->seps[i] * lig->seps[j];
                double vdw_r = rec->rad[i] + lig->rad[j];
                double p6 = powi6(vdw_r) / powi3(d2);
                double vp = fmin(vdw_e * (p6 * p6 - 2.0 * p6), VDW_CUTOFF);
                if (rec->heavy[i] && lig->heavy[j]) {
                    double d = sqrt(d2);
                    double dmin = CLASH_PENALTY_FRAC * vdw_r;
                    if (d < dmin) vp += CLASH_PENALTY_W * (dmin - d);
                }
                vi += vp;
            }
        }
        te += ei;
        tv += vi;
    }
    free(lc);
    return -(te * FACTOR / EPSILON + tv);
}

/* pydock.rs:280-313 (restraint-free): receptor-major pair loop, every pair within
 * 30 A contributes clamped electrostatics, pairs within 10 A the capped LJ; the
 * per-receptor-atom partial sums are then added in receptor order. */
double orc_pydock_energy_exact(const orc_mol* rec, const orc_mol* lig, const double* pose7) {
    int nl = lig->n, nr = rec->n;
    double* lc = (double*)malloc(sizeof(double) * 3 * (size_t)(nl > 0 ? nl : 1));
    orc_transform(lig->xyz, nl, pose7, lc);
    double te = 0.0, tv = 0.0;
    for (int i = 0; i < nr; ++i) {
        const double* r = rec->xyz + 3 * i;
        double ei = 0.0, vi = 0.0;
        for (int j = 0; j < nl; ++j) {
            double dx = r[0] - lc[3 * j], dy = r[1] - lc[3 * j + 1], dz = r[2] - lc[3 * j + 2];
            double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 <= ELEC_DIST_CUTOFF2)
                ei += clampd(rec->ele[i] * lig->ele[j] / d2, ELEC_MIN_CUTOFF, ELEC_MAX_CUTOFF);
            if (d2 <= VDW_DIST_CUTOFF2) {
                double vdw_e = rec->seps[i] * lig->seps[j];
                double vdw_r = rec->rad[i] + lig->rad[j];
                double p6 = powi6(vdw_r) / powi3(d2);
                vi += fmin(vdw_e * (p6 * p6 - 2.0 * p6), VDW_CUTOFF);
            }
        }
        te += ei;
        tv += vi;
    }
    free(lc);
    return -(te * FACTOR / EPSILON + tv);
}
