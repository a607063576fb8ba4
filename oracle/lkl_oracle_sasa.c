/*
 * lkl_oracle_sasa.c — CPU restatement of the reference's Lee & Richards SASA
 * (lr_sasa.rs:1-291, itself following freesasa 2.2's nb.c / sasa_lr.c) and of
 * cpydock's exact all-pairs energy (cpydock.rs:386-553).
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, the smoke() entry and
 * bench.py's cpu_baseline leg load this library; the product never does.
 *
 * Neighbour lists are found by an all-pairs scan in index order instead of the
 * reference's cell traversal.  The order of a neighbour list does not change an
 * atom's area: the buried arcs of a slice are sorted by start angle before the
 * exposed length is summed, and arcs with equal starts give the same sum in
 * either order.
 *
 * Parity: orc_lr_sasa is pinned by closed forms (a lone sphere, a sphere inside
 * another, the two-sphere cap at many slices) and by an independent dense-angle
 * restatement of the 20-slice sum (tests/test_sasa_oracle.py);
 * orc_cpydock_energy_exact by the 1AZP golden test of cpydock.rs:861-875.
 */
#include <math.h>
#include <stdlib.h>

#include "lkl_oracle.h"

#define TWOPI (2.0 * M_PI)

typedef struct {
    int* off;     /* n + 1 */
    int* idx;     /* neighbour index */
    double* xyd;  /* sqrt(dx^2 + dy^2), dx = x_other - x_self */
    double* xd;
    double* yd;
} nblist;

/* lr_sasa.rs:28-172: pairs with squared distance < (Ri + Rj)^2, R with probe. */
static void nb_build(const double* xyz, const double* R, int n, nblist* nb) {
    int* cnt = (int*)calloc((size_t)n + 1, sizeof(int));
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            double dx = xyz[3 * j] - xyz[3 * i], dy = xyz[3 * j + 1] - xyz[3 * i + 1],
                   dz = xyz[3 * j + 2] - xyz[3 * i + 2];
            double c = (R[i] + R[j]) * (R[i] + R[j]);
            if (dx * dx + dy * dy + dz * dz < c) {
                cnt[i]++;
                cnt[j]++;
            }
        }
    nb->off = (int*)malloc(sizeof(int) * ((size_t)n + 1));
    nb->off[0] = 0;
    for (int i = 0; i < n; ++i) nb->off[i + 1] = nb->off[i] + cnt[i];
    size_t tot = (size_t)nb->off[n] + 1;
    nb->idx = (int*)malloc(sizeof(int) * tot);
    nb->xyd = (double*)malloc(sizeof(double) * tot);
    nb->xd = (double*)malloc(sizeof(double) * tot);
    nb->yd = (double*)malloc(sizeof(double) * tot);
    for (int i = 0; i < n; ++i) cnt[i] = nb->off[i];
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            double dx = xyz[3 * j] - xyz[3 * i], dy = xyz[3 * j + 1] - xyz[3 * i + 1],
                   dz = xyz[3 * j + 2] - xyz[3 * i + 2];
            double c = (R[i] + R[j]) * (R[i] + R[j]);
            if (dx * dx + dy * dy + dz * dz < c) {
                double dxy = sqrt(dx * dx + dy * dy);
                int a = cnt[i]++, b = cnt[j]++;
                nb->idx[a] = j; nb->xyd[a] = dxy; nb->xd[a] = dx; nb->yd[a] = dy;
                nb->idx[b] = i; nb->xyd[b] = dxy; nb->xd[b] = -dx; nb->yd[b] = -dy;
            }
        }
    free(cnt);
}

/* lr_sasa.rs:176-191: insertion sort of (start, end) pairs by start. */
static void sort_arcs(double* arc, int n) {
    for (int i = 2; i < 2 * n; i += 2) {
        double t0 = arc[i], t1 = arc[i + 1];
        int j = i;
        while (j > 0 && arc[j - 2] > t0) {
            arc[j] = arc[j - 2];
            arc[j + 1] = arc[j - 1];
            j -= 2;
        }
        arc[j] = t0;
        arc[j + 1] = t1;
    }
}

/* lr_sasa.rs:194-213: the exposed length of a circle given its buried arcs. */
static double exposed_arc_length(double* arc, int n) {
    if (n == 0) return TWOPI;
    sort_arcs(arc, n);
    double sum = arc[0], sup = arc[1];
    for (int i2 = 2; i2 < 2 * n; i2 += 2) {
        if (sup < arc[i2]) sum += arc[i2] - sup;
        if (arc[i2 + 1] > sup) sup = arc[i2 + 1];
    }
    return sum + TWOPI - sup;
}

/* lr_sasa.rs:217-274: slices at z_i - R_i + (k - 1/2) delta, k = 1..n_slices. */
static double atom_area(int i, const double* xyz, const double* R, const nblist* nb, int ns,
                        double* arc) {
    int b = nb->off[i], nni = nb->off[i + 1] - b;
    double zi = xyz[3 * i + 2], ri = R[i];
    double delta = 2.0 * ri / (double)ns;
    double z = zi - ri - 0.5 * delta;
    double sasa = 0.0;
    for (int s = 0; s < ns; ++s) {
        z += delta;
        double di = fabs(zi - z);
        double ri_p2 = ri * ri - di * di;
        if (ri_p2 < 0.0) continue;
        double ri_p = sqrt(ri_p2);
        if (ri_p <= 0.0) continue;
        int n_arcs = 0, buried = 0;
        for (int k = 0; k < nni; ++k) {
            int j = nb->idx[b + k];
            double dj = fabs(xyz[3 * j + 2] - z), rj = R[j];
            if (!(dj < rj)) continue;
            double rj_p2 = rj * rj - dj * dj;
            double rj_p = sqrt(rj_p2);
            double dij = nb->xyd[b + k];
            if (dij >= ri_p + rj_p) continue;
            if (dij + ri_p < rj_p) { buried = 1; break; }
            if (dij + rj_p < ri_p) continue;
            double alpha = acos((ri_p2 + dij * dij - rj_p2) / (2.0 * ri_p * dij));
            double beta = atan2(nb->yd[b + k], nb->xd[b + k]) + M_PI;
            double inf = beta - alpha, sup = beta + alpha;
            if (inf < 0.0) inf += TWOPI;
            if (sup > 2.0 * M_PI) sup -= TWOPI;
            int a2 = 2 * n_arcs;
            if (sup < inf) {
                arc[a2] = 0.0; arc[a2 + 1] = sup; arc[a2 + 2] = inf; arc[a2 + 3] = TWOPI;
                n_arcs += 2;
            } else {
                arc[a2] = inf; arc[a2 + 1] = sup;
                n_arcs += 1;
            }
        }
        if (!buried) sasa += delta * ri * exposed_arc_length(arc, n_arcs);
    }
    return sasa;
}

/* lr_sasa.rs:277-291 (probe and slice count as arguments; the reference uses
 * probe 1.4, 20 slices). */
void orc_lr_sasa(const double* xyz, const double* radii, int n, double probe, int n_slices,
                 double* out) {
    if (n <= 0) return;
    double* R = (double*)malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; ++i) R[i] = radii[i] + probe;
    nblist nb;
    nb_build(xyz, R, n, &nb);
    int maxnb = 0;
    for (int i = 0; i < n; ++i)
        if (nb.off[i + 1] - nb.off[i] > maxnb) maxnb = nb.off[i + 1] - nb.off[i];
    double* arc = (double*)malloc(sizeof(double) * 4 * ((size_t)maxnb + 1));
    for (int i = 0; i < n; ++i) out[i] = atom_area(i, xyz, R, &nb, n_slices, arc);
    free(arc);
    free(nb.off); free(nb.idx); free(nb.xyd); free(nb.xd); free(nb.yd);
    free(R);
}

/* cpydock.rs:386-553 (restraint-free): one receptor-major pass gives the
 * clamped electrostatics within 30 A, the capped LJ within 10 A (sqrt(eps eps),
 * powf) and the per-atom minimum squared distances over pairs outside the
 * C-binary hydrogen mask; then the desolvation of both sides. */
double orc_cpydock_energy_exact(const orc_mol* rec, const orc_mol* lig, const double* pose7) {
    int nl = lig->n, nr = rec->n;
    double* lc = (double*)malloc(sizeof(double) * 3 * (size_t)(nl > 0 ? nl : 1));
    double* min_rec = (double*)malloc(sizeof(double) * (size_t)(nr > 0 ? nr : 1));
    double* min_lig = (double*)malloc(sizeof(double) * (size_t)(nl > 0 ? nl : 1));
    orc_transform(lig->xyz, nl, pose7, lc);
    for (int i = 0; i < nr; ++i) min_rec[i] = 10000.0;
    for (int j = 0; j < nl; ++j) min_lig[j] = 10000.0;
    double te = 0.0, tv = 0.0;
    for (int i = 0; i < nr; ++i) {
        const double* r = rec->xyz + 3 * i;
        int fi = (i % 2 == 0) ? (rec->hyd[i / 2] != 0) : 0;
        for (int j = 0; j < nl; ++j) {
            double dx = r[0] - lc[3 * j], dy = r[1] - lc[3 * j + 1], dz = r[2] - lc[3 * j + 2];
            double d2 = dx * dx + dy * dy + dz * dz;
            int fj = (j % 2 == 0) ? (lig->hyd[j / 2] != 0) : 0;
            if (!fi && !fj) {
                if (d2 < min_rec[i]) min_rec[i] = d2;
                if (d2 < min_lig[j]) min_lig[j] = d2;
            }
            if (d2 <= 900.0) {
                double ae = rec->ele[i] * lig->ele[j] / d2;
                const double lo = -1.0 * 4.0 / 332.0, hi = 1.0 * 4.0 / 332.0;
                te += ae < lo ? lo : (ae > hi ? hi : ae);
            }
            if (d2 <= 100.0) {
                double vdw_e = sqrt(rec->eps[i] * lig->eps[j]);
                double vdw_r = rec->rad[i] + lig->rad[j];
                double p6 = pow(vdw_r, 6.0) / pow(d2, 3.0);
                double k = vdw_e * (p6 * p6 - 2.0 * p6);
                tv += k > 1.0 ? 1.0 : k;
            }
        }
    }
    double ts = 0.0;
    for (int i = 0; i < nr; ++i) {
        double s = 0.0;
        if (min_rec[i] <= 6.4 * 6.4 && min_rec[i] > 0.0 && rec->asa[i] > 0.0)
            s = -10.0 * sqrt(min_rec[i]) + 65.0;
        if (s > rec->asa[i]) s = rec->asa[i];
        ts += s * rec->des[i];
    }
    for (int j = 0; j < nl; ++j) {
        double s = 0.0;
        if (min_lig[j] <= 6.4 * 6.4 && min_lig[j] > 0.0 && lig->asa[j] > 0.0)
            s = -10.0 * sqrt(min_lig[j]) + 65.0;
        if (s > lig->asa[j]) s = lig->asa[j];
        ts += s * lig->des[j];
    }
    free(lc);
    free(min_rec);
    free(min_lig);
    return (te * 332.0 / 4.0 + 0.1 * tv - ts) * -1.0;
}
