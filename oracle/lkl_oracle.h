/*
 * lkl_oracle.h — CPU restatement (f64) of the LKlight reference CPU grid path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/, the
 * smoke() entry and bench.py's cpu_baseline leg may load it.  The product
 * (lklight-gpu_amd/) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  Coordinates and
 * parameters are f64, exactly as the reference Rust models hold them; poses
 * are N x 7 f64 rows (w, x, y, z, tx, ty, tz) as packed by gpu_score.rs:292-301.
 */
#ifndef LKL_ORACLE_H
#define LKL_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One molecule in the reference's model layout (dna.rs DNADockingModel,
 * pydock.rs PYDOCKDockingModel, cpydock.rs CPYDOCKDockingModel). */
typedef struct {
    int n;
    const double* xyz;           /* n*3 */
    const double* ele;           /* ele_charges */
    const double* eps;           /* vdw_charges (epsilon) */
    const double* seps;          /* sqrt_vdw_charges */
    const double* rad;           /* vdw_radii */
    const unsigned char* heavy;  /* dna heavy mask */
    const int* hyd;              /* cpydock hydrogens: 1 = heavy (cpydock.rs:168) */
    const double* des;           /* cpydock des_energy */
    const double* asa;           /* cpydock reference SASA */
} orc_mol;

/* Receptor far field (grid_dna.rs:33-40). phi is caller-owned storage. */
typedef struct {
    double origin[3];
    int64_t n[3];
    double spacing;
    float* phi;
} orc_field;

/* Family ids. */
enum { ORC_DNA = 0, ORC_VDW = 1, ORC_PYDOCK = 2, ORC_CPYDOCK = 3 };

/* grid_dna.rs:44-61 — box of the field; returns number of grid points. */
int64_t orc_field_dims(const double* xyz, int n, double origin[3], int64_t dims[3]);
/* grid_dna.rs:62-148 — shell scatter; phi must hold dims[0]*dims[1]*dims[2] floats. */
void orc_field_fill(const double* xyz, const double* q, int n,
                    const double origin[3], const int64_t dims[3], float* phi);
/* grid_dna.rs:159-197 */
double orc_field_sample(const orc_field* f, double x, double y, double z);

/* gpu_score.rs:73-145 (CudaReceptor::build): 10 A cells.
 * cell_start must hold (ncx+1)(ncy+1)(ncz+1)+1 ints; call with cell_start ==
 * NULL first to get the grid (return value = length of cell_start). */
int orc_crec_build(const double* xyz, int n, int* ncxyz, float* c_origin,
                   int* cell_start, int* cell_atoms);

/* qt.rs:67-74 + qt.rs:122-128 + the translation add of dna.rs:868-873. */
void orc_transform(const double* base, int nl, const double* pose7, double* out);

/* One pose of each family's energy_grid, returning the components the GPU
 * bridge returns: E_raw (raw electrostatics, x332/4 for kcal), V (LJ [+clash]),
 * S (cpydock desolvation).  n_near / n_far count pairs with d2 <= 100 and
 * charged ligand atoms whose far-field sample lands in the grid. */
typedef struct {
    double e_raw, v, s;
    int64_t n_near, n_far, n_cand;
    /* Magnitudes for the FP32 rounding bound of a device evaluation of the
     * same sum: sum over the pairs of |clamped electrostatic term| (e_abs),
     * sum over charged ligand atoms of |q phi| (f_abs), sum over the pairs
     * whose f64 LJ term is not clearly capped of |eps_ij| (p6^2 + 2 p6)
     * (v_abs), sum over pairs within 1e-6 of the clash distance or closer of
     * the clash weight times that distance (c_abs), and the sum of |des| over
     * receptor atoms inside the desolvation window (s_abs). */
    double e_abs, f_abs, v_abs, c_abs, s_abs;
} orc_parts;

/* dna.rs:834-998 */
void orc_dna_energy_grid(const orc_mol* rec, const orc_mol* lig, const orc_field* f,
                         const double* pose7, orc_parts* out);
/* dna.rs:885-995 interface flags of one pose (pairs with f64 d2 <= 3.9^2). */
void orc_interface_flags(const orc_mol* rec, const orc_mol* lig, const double* pose7,
                         unsigned char* iface_r, unsigned char* iface_l);
/* scoring.rs:83-98 / 100-110 and the fold of dna.rs:989-995. */
double orc_satisfied_restraints(const unsigned char* iface, const int* off, const int* atoms, int n_res);
double orc_membrane_intersection(const unsigned char* iface, const int* atoms, int n);
double orc_restraint_fold(double score, double perc_r, double perc_l, double intersection);
/* vdw.rs:73-151 */
void orc_vdw_energy_grid(const orc_mol* rec, const orc_mol* lig,
                         const double* pose7, orc_parts* out);
/* pydock.rs:353-447 */
void orc_pydock_energy_grid(const orc_mol* rec, const orc_mol* lig, const orc_field* f,
                            const double* pose7, orc_parts* out);
/* cpydock.rs:556-676 */
void orc_cpydock_energy_grid(const orc_mol* rec, const orc_mol* lig, const orc_field* f,
                             const double* pose7, orc_parts* out);

/* Score folds of the reference (gpu_score.rs:326-331, cpydock.rs:669). */
double orc_fold_score(int family, double e_raw, double v, double s);

/* Batched driver: one energy_grid per pose, poses spread over nthreads
 * pthreads (mirrors rayon's per-glowworm par_iter_mut, swarm.rs:102).
 * Any of out_e/out_v/out_s/out_near/out_far may be NULL. */
void orc_batch_energy_grid(int family, const orc_mol* rec, const orc_mol* lig,
                           const orc_field* f, const double* poses, int n_pose,
                           double* out_e, double* out_v, double* out_s,
                           int64_t* out_near, int64_t* out_far, int nthreads);
/* The same batch, also returning per pose the five rounding magnitudes of
 * orc_parts (out_mag: n_pose x 5, e_abs f_abs v_abs c_abs s_abs). */
void orc_batch_energy_grid_mag(int family, const orc_mol* rec, const orc_mol* lig,
                               const orc_field* f, const double* poses, int n_pose,
                               double* out_e, double* out_v, double* out_s,
                               int64_t* out_near, int64_t* out_far, double* out_mag, int nthreads);

/* FP64 brute-force re-evaluation of the same grid-split sum (no cell list):
 * every (i, j) pair, same terms, same far field.  Pins the cell enumeration. */
void orc_bruteforce_parts(int family, const orc_mol* rec, const orc_mol* lig,
                          const orc_field* f, const double* pose7, orc_parts* out);

/* dna.rs:668-819 energy_exact without restraints: 30 A all-pairs, no grid. */
double orc_dna_energy_exact(const orc_mol* rec, const orc_mol* lig, const double* pose7);
double orc_pydock_energy_exact(const orc_mol* rec, const orc_mol* lig, const double* pose7);
/* cpydock.rs:386-553 energy_exact without restraints (all pairs; desolvation). */
double orc_cpydock_energy_exact(const orc_mol* rec, const orc_mol* lig, const double* pose7);

/* lr_sasa.rs:277-291 Lee & Richards SASA per atom; radii without the probe. */
void orc_lr_sasa(const double* xyz, const double* radii, int n, double probe, int n_slices,
                 double* out);

/* nearcell.rs:22-112 self-test support: count pairs within cut2 found by the
 * 27-cell enumeration (NearCells with the given cell edge) over a point set. */
int64_t orc_nearcells_pair_count(const double* xyz, int n, double cell);

/* ── GSO (lkl_oracle_gso.c) ─────────────────────────────────────────────── */
/* rand 0.7.3 StdRng (ChaCha20, seed_from_u64) draws of gen::<f64>(). */
void orc_rng_f64s(uint64_t seed, int n, double* out);
/* qt.rs:106-116 Quaternion::random, n draws from one seeded StdRng. */
void orc_quat_random(uint64_t seed, int n, double* out);
/* qt.rs:80-104 */
void orc_slerp(const double* a, const double* b, double t, double* out);
/* lib.rs:47-78 GSO over one swarm with CPU-grid scoring of `family`. */
void orc_gso_run(int family, const orc_mol* rec, const orc_mol* lig, const orc_field* f,
                 const double* positions, int n, uint64_t seed, int steps, int nthreads,
                 double* out_pose, double* out_luc, double* out_score, double* out_vision,
                 int* out_nnb);

/* Surrogate pair-histogram features (lkl_oracle_sur.c): counts n x 6144. */
void orc_feat_counts(const double* rxyz, const unsigned char* rg, int nr, const double* lxyz,
                     const unsigned char* lg, int nl, const double* poses, int n, uint32_t* counts,
                     int nthreads);
/* Grid-mode features (lkl_oracle_sur.c): box of the node grid, counts of one
 * node (384 = group x bin), and per-pose features, n x 6144. */
int64_t orc_feat_grid_dims(const double* rxyz, int nr, double h, double origin[3], int dims[3]);
void orc_feat_grid_node(const double* rxyz, const unsigned char* rg, int nr, const double origin[3],
                        double h, int ix, int iy, int iz, uint32_t* out384);
void orc_feat_grid_counts(const double* rxyz, const unsigned char* rg, int nr, const double* lxyz,
                          const unsigned char* lg, int nl, const double* poses, int n, double h,
                          uint32_t* counts, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
