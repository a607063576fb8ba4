/*
 * lklight_gpu.h — C ABI of the MI355X (gfx950) batched GSO pose-scoring path.
 *
 * Drop-in boundary for LKlight's GPU bridge.  The four stateless entry points
 * keep the argument lists, argument meaning and 0 / -1 status convention of the
 * reference's extern "C" functions (new lkl_ names, no cuda* symbols).  The
 * context API underneath them is explicit: create (upload constants once),
 * score (per GSO step), destroy.
 *
 * Conventions (all entry points):
 *   - the caller owns every buffer passed in; nothing is retained after return
 *     except device copies held by a context;
 *   - return 0 on success, -1 on failure; lkl_last_error() returns the
 *     message of the calling thread's last failure (also printed to stderr);
 *   - results are deterministic: per-pose sums are reduced in a fixed order,
 *     independent of launch shape and scheduling;
 *   - N (poses per call) is bounded only by memory (no 65,535 limit);
 *   - one context must not be used by two threads at once; distinct contexts
 *     are independent.  The stateless wrappers serialise on an internal mutex.
 *
 * Poses are N x 7 f64 rows (w, x, y, z, tx, ty, tz): unit quaternion then
 * translation, as packed by gpu_score.rs:292-301.  Score outputs are N x 2 f64
 * rows (E_raw, V): raw electrostatics (x 332/4 for kcal/mol) and the capped
 * Lennard-Jones sum (+ heavy-atom clash when F_CLASH is set).
 */
#ifndef LKLIGHT_GPU_H
#define LKLIGHT_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* Family flag word (full_score.cu:23-25, gpu_score.rs:15-21).  LJ is always on. */
#define LKL_F_FAR 1u    /* 10-30 A far field  q_j * phi(x_j)          */
#define LKL_F_ELEC 2u   /* <= 10 A clamped electrostatics             */
#define LKL_F_CLASH 4u  /* heavy-atom linear clash penalty            */
#define LKL_F_DESOLV 16u /* cpydock contact-SASA desolvation (S)       */
#define LKL_FLAGS_DNA 7u
#define LKL_FLAGS_PYDOCK 3u
#define LKL_FLAGS_VDW 0u

/* ── stateless entry points (reference signatures) ─────────────────────────── */

/* Replaces cuda_batch_score (full_score.cu:359-371), called from
 * gpu_score.rs:252-261 and gpu_family.rs:211-221.  The receptor cell list
 * arguments are accepted for signature compatibility; the library builds its
 * own cell-sorted layout from r_coords.  out must hold N*2 doubles. */
int lkl_batch_score(const float* phi, int nx, int ny, int nz, float ox, float oy, float oz,
                    float sp, const float* r_coords, const float* r_ele, const float* r_svdw,
                    const float* r_vdwr, const unsigned char* r_heavy, int nr,
                    const int* cell_start, const int* cell_atoms, int ncx, int ncy, int ncz,
                    float c_ox, float c_oy, float c_oz, float c_sp, const float* l_base,
                    const double* poses, const float* l_ele, const float* l_svdw,
                    const float* l_vdwr, const unsigned char* l_heavy, int nl, int N,
                    unsigned flags, double* out);

/* Replaces cuda_cpydock_solv (full_score.cu:571-581), called from
 * gpu_family.rs:360-372.  outS[k] = desolvation S of pose k, the term of
 * cpydock's score = -(E*332/4 + 0.1*V - S).  l_ele is unused (as in the
 * reference). */
int lkl_cpydock_solv(const float* r_coords, const unsigned char* r_flag, int nr,
                     const int* cell_start, const int* cell_atoms, int ncx, int ncy, int ncz,
                     float c_ox, float c_oy, float c_oz, float c_sp, const float* l_base,
                     const double* poses, const unsigned char* l_flag, const float* l_ele,
                     int nl, int N, const float* r_des, const float* r_asa,
                     const float* l_des, const float* l_asa, double* outS);

/* Replaces cuda_full_score (full_score.cu:140-150), called from
 * gpu_score.rs:159-168: one pre-transformed pose, per-ligand-atom outputs
 * (out_elec[j] raw electrostatics near + far, out_vdw[j] LJ + clash), with
 * the dna flag word (far + elec + clash). */
int lkl_full_score(const float* phi, int nx, int ny, int nz, float ox, float oy, float oz,
                   float sp, const float* r_coords, const float* r_ele, const float* r_svdw,
                   const float* r_vdwr, const unsigned char* r_heavy, int nr,
                   const int* cell_start, const int* cell_atoms, int ncx, int ncy, int ncz,
                   float c_ox, float c_oy, float c_oz, float c_sp, const float* l_coords,
                   const float* l_ele, const float* l_svdw, const float* l_vdwr,
                   const unsigned char* l_heavy, int nl, float* out_elec, float* out_vdw);

/* Replaces cuda_far_field (far_field.cu:59-63), called from gpu_field.rs:20-35:
 * *result = sum_j q_j * phi(x_j) over pre-transformed coordinates. */
int lkl_far_field(const float* phi, int nx, int ny, int nz, float ox, float oy, float oz,
                  float sp, const float* coords, const float* charges, int n, double* result);

/* Replaces the direct cudaGetDeviceCount probe of gpu_score.rs:226-232
 * (cuda_available) and gpu_family.rs:322-337: 1 if a usable gfx950 device is
 * present, else 0.  Never fails. */
int lkl_device_available(void);

/* Message of the calling thread's last failed call ("" if none).  No
 * reference counterpart: the reference prints to stderr only
 * (full_score.cu:463). */
const char* lkl_last_error(void);

/* Drop every context cached by the stateless entry points (their device copies
 * are keyed on host-pointer identity, as in full_score.cu:380-387).  Call
 * after mutating a cached host buffer in place.  No reference counterpart. */
void lkl_cache_clear(void);

/* ── context API (no reference counterpart; the stateless entry points above
 *    are implemented on it) ──────────────────────────────────────────────── */

typedef struct lkl_ctx lkl_ctx;

/* One complex: receptor far field, receptor and ligand atoms (reference
 * order), optional cpydock desolvation sides.  Unused pointers may be NULL
 * (phi == NULL or nx == 0 means "no far field": F_FAR is then rejected;
 * r_flag == NULL means "no desolvation": F_DESOLV is then rejected). */
typedef struct {
    const float* phi;  /* nx*ny*nz, x fastest (grid_dna.rs:151-154) */
    int nx, ny, nz;
    double ox, oy, oz, sp;
    int nr;
    const float* r_xyz; /* nr*3 */
    const float* r_ele;
    const float* r_svdw; /* sqrt(epsilon) */
    const float* r_vdwr;
    const unsigned char* r_heavy;
    int nl;
    const float* l_xyz; /* nl*3, reference (untransformed) coordinates */
    const float* l_ele;
    const float* l_svdw;
    const float* l_vdwr;
    const unsigned char* l_heavy;
    const unsigned char* r_flag; /* cpydock exclusion masks (cpydock.rs:808-819) */
    const float* r_des;
    const float* r_asa;
    const unsigned char* l_flag;
    const float* l_des;
    const float* l_asa;
    float cell_size; /* receptor cell edge in A; 0 = library default */
} lkl_complex_desc;

/* Upload one complex.  Returns NULL on failure (see lkl_last_error). */
lkl_ctx* lkl_ctx_create(const lkl_complex_desc* desc);
void lkl_ctx_destroy(lkl_ctx* ctx);

/* Blocking, host buffers: poses N*7, out N*2 (E_raw, V), outS N (needed only
 * with LKL_F_DESOLV, may be NULL otherwise).  The poses are copied into the
 * context's pinned, device-mapped staging buffer, which the kernels read over
 * the bus and the scoring kernel writes the scores into (no DMA copies, one
 * stream synchronisation).  One call at a time per context. */
int lkl_ctx_score(lkl_ctx* ctx, const double* poses, int N, unsigned flags, double* out,
                  double* outS);

/* Asynchronous, device buffers, enqueued on `stream` (a hipStream_t; NULL =
 * the context's own stream).  d_out / d_outS as above; d_outS may be NULL
 * without LKL_F_DESOLV.  Launches the scoring kernel, preceded, when the
 * launch is dispatched longest first (lkl_ctx_set_pose_order), by two short
 * kernels that estimate each pose's cost and order the poses; no copies or
 * memsets, so the call can be captured in a graph. */
int lkl_ctx_score_async(lkl_ctx* ctx, const double* d_poses, int N, unsigned flags,
                        double* d_out, double* d_outS, void* stream);

/* lkl_ctx_score_async without any allocation (no counterpart in the
 * reference).  For callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * Same arguments, results and stream semantics as lkl_ctx_score_async; the
 * launch's scratch (per-pose dispatch order, tile accumulators, global
 * desolvation minima) lives in one static 16 MiB device array per GPU instead
 * of buffers the context grows on demand.  At most about 400,000 poses per
 * call (fewer with desolvation minima too large for LDS: 4 bytes per receptor
 * atom per pose); larger calls return -1 with a message.  Every call on a GPU
 * shares that array, so calls that overlap on one GPU (two streams, two
 * threads) race: serialise them.  The context's device must be the current
 * one.  Allocates nothing (the array's address is looked up once per device,
 * on the first call there). */
int lkl_ctx_score_async_noalloc(lkl_ctx* ctx, const double* d_poses, int N, unsigned flags,
                                double* d_out, double* d_outS, void* stream);

/* Longest-first dispatch (default on): one-tile-per-pose launches of 257 to
 * 4,096 poses over neighbour lists first estimate each pose's near-field
 * cost (list lengths at its atoms' nodes) and start the costly poses first,
 * so costly tiles share CUs with cheap ones and cheap poses fill the last
 * round of workgroup slots.  Scores are the same
 * bits either way.  0 turns it off (one kernel per launch).  0 / -1. */
int lkl_ctx_set_pose_order(lkl_ctx* ctx, int enable);
/* Current setting: 1 / 0, or -1 for a NULL context. */
int lkl_ctx_pose_order(const lkl_ctx* ctx);

/* Kernel timing (no reference counterpart; measurement only).  With timing on,
 * every launch of the context (blocking or asynchronous) records HIP events on
 * its stream around the dispatch-order pre-pass and around the scoring
 * kernel.  lkl_ctx_kernel_ms waits for the last launch and returns both
 * durations in ms (prepass_ms is 0 when that launch ran no pre-pass).  0 / -1. */
int lkl_ctx_set_timing(lkl_ctx* ctx, int on);
int lkl_ctx_kernel_ms(lkl_ctx* ctx, double* prepass_ms, double* score_ms);

/* Scores ligands already placed: coords N*nl*3 f32 (pose-major, the
 * descriptor's ligand order), one pose per coordinate set, the context's
 * charges and parameters.  This is the input form of the reference's
 * dump_metal_dataset poses.bin (dna.rs:473-663), which stores transformed
 * coordinates rather than rotations; a rigidly placed set scores as the
 * batched kernel scores the f32 coordinates of that placement.  Coordinates
 * must be finite with |x| < 1e6.  Blocking; out / outS as lkl_ctx_score. */
int lkl_ctx_score_coords(lkl_ctx* ctx, const float* coords, int N, unsigned flags, double* out,
                         double* outS);

/* The pose set of DNA::dump_metal_dataset (dna.rs:497-519) for a seed, from
 * rand 0.7.3's StdRng: n x 7 (w, x, y, z, tx, ty, tz).  Returns 0 / -1. */
int lkl_dataset_poses(unsigned long long seed, int n, double* out);

/* Diagnostic: f64 transformed ligand coordinates of N poses (N*nl*3, ligand
 * order of the descriptor), computed by the device transform. */
int lkl_ctx_transform(lkl_ctx* ctx, const double* poses, int N, double* out_xyz);

/* Set the receptor cell edge (A, in (0.5, 20]), workgroup size (64, 128, 256,
 * 512 or 1024) and ligand atoms per workgroup tile (0 = the whole ligand) used
 * by later score calls.  Scores are bitwise identical for every setting. */
int lkl_ctx_set_tiles(lkl_ctx* ctx, float cell_size, int block, int chunk);

/* Receptor cell edge along x (A, in (0.5, 20]; 0 = cell_size, cubic cells).
 * Narrower x cells trim each cell row's candidate range.  Scores are bitwise
 * identical for every setting. */
int lkl_ctx_set_cell_x(lkl_ctx* ctx, float cell_x);

/* Work ordering of the near field (default on): each tile's ligand atoms are
 * bucketed by the receptor-atom count near them (a 2 A receptor grid built at
 * creation) so a wavefront's lanes walk windows of similar size, and atoms with
 * no receptor atom in reach skip the near-field walk.  Scores are bitwise
 * identical either way. */
int lkl_ctx_set_work_sort(lkl_ctx* ctx, int on);

/* Neighbour lists of the near field (default on, spacing 1 A): for each node
 * of a receptor grid of that spacing, the receptor atoms within the 10.025 A
 * window plus the node's half-diagonal, built on the device once per receptor
 * layout.  A ligand atom then scans its nearest node's list instead of walking
 * cell rows.  spacing 0 switches them off (cell-row walk); they are also off
 * for receptors above 65,535 atoms or when the tables would exceed 2^30
 * entries.  Scores are bitwise identical either way.  Library-internal layout
 * with no counterpart in the reference (cells are built by
 * src/gpu_score.rs:61-112 there). */
int lkl_ctx_set_list_spacing(lkl_ctx* ctx, float spacing);

/* Interface flags for restraints / membrane (dna.rs:885-995, cpydock.rs:509,
 * 639, scoring.rs:83-110), which keep the reference's restraint runs off the
 * GPU batch (dna.rs:1047-1056).  Residues are CSR over atom indices in the
 * caller's (reference) order: residue k holds atoms res_atoms[res_off[k] ..
 * res_off[k+1]); membrane beads are receptor atom indices.  Counts replace
 * the previous data; all zero clears it.  0 / -1. */
int lkl_ctx_set_interface(lkl_ctx* ctx, const int* rec_res_off, const int* rec_res_atoms, int n_rec_res,
                          const int* lig_res_off, const int* lig_res_atoms, int n_lig_res,
                          const int* membrane_atoms, int n_membrane);
/* 1 when the context holds restraints or a membrane, else 0. */
int lkl_ctx_has_interface(const lkl_ctx* ctx);
/* Numbers of receptor residues, ligand residues and membrane beads set. */
int lkl_ctx_interface_sizes(const lkl_ctx* ctx, int* n_rec_res, int* n_lig_res, int* n_membrane);
/* Per pose: counts[3k] = receptor residues with an atom at the interface
 * (f64 squared distance <= 3.9^2 to some ligand atom), counts[3k+1] = the
 * same for ligand residues, counts[3k+2] = membrane beads at the interface.
 * The score fold is then score += (r/n_rec_res)*score + (l/n_lig_res)*score
 * - 999*(m/n_membrane) (a fraction with no residues counts 0; the penalty only
 * when m > 0), dna.rs:989-995.  Needs the neighbour lists.  Blocking on host
 * buffers, or async on device buffers (`stream` NULL = the context's). */
int lkl_ctx_interface(lkl_ctx* ctx, const double* poses, int N, int* counts);
int lkl_ctx_interface_async(lkl_ctx* ctx, const double* d_poses, int N, int* d_counts, void* stream);

/* Far-field layout (default on): besides phi, keep a corner-packed copy (the
 * 8 corners of every grid cell together, 32 bytes per cell, 8x the memory of
 * phi) so a trilinear sample is one aligned read.  Samples are bitwise those
 * of phi; when the copy cannot be allocated phi is sampled directly.
 * Library-internal layout with no counterpart in the reference (grid_dna.rs
 * keeps phi only). */
int lkl_ctx_set_phi_packed(lkl_ctx* ctx, int on);

/* 1 when the context holds the corner-packed far field, else 0. */
int lkl_ctx_phi_packed(const lkl_ctx* ctx);

/* Parameter types of the list walk (level, default 2): when the receptor's
 * per-atom (sqrt(eps), vdw radius, heavy) records take at most 256 distinct
 * values, the neighbour-list walk copies them and one type byte per atom to
 * LDS and reads a candidate's parameters there instead of gathering them with
 * its coordinates (level 1).  At level 2, when the ligand's records take at
 * most 256 values too and the table has at most 2,048 entries, each workgroup
 * instead forms a table over (ligand type, receptor type) pairs of the LJ
 * factor, (r_i + r_j)^2 and the clash threshold, with the walk's own f32
 * operations, and the walk reads a pair's entry (fewer operations per
 * candidate).  Level 0 turns the tables off.  Scores are bitwise the same at
 * every level.  Library-internal layout with no counterpart in the reference
 * (full_score.cu reads each receptor atom's parameters from global memory).
 * lkl_ctx_param_types returns the number of receptor types in use (0 = off,
 * or more than 256 records). */
int lkl_ctx_set_param_types(lkl_ctx* ctx, int on);
int lkl_ctx_param_types(const lkl_ctx* ctx);

/* Neighbour-list state: returns the number of list entries (0 = lists off,
 * -1 = null context) and the node grid dimensions in dims3 (may be NULL). */
long long lkl_ctx_list_info(const lkl_ctx* ctx, int* dims3);

/* Autotuner: time every (cell edge y/z, cell edge x, workgroup size, ligand
 * tile) candidate on N host poses with `flags`, keep the fastest in the context
 * and report it.  Scores are bitwise identical for every candidate.  table
 * (may be NULL) receives up to 320 rows of (cell A, cell_x A, block, chunk,
 * median ms); *n_table their count.  No reference counterpart
 * (full_score.cu:442 hard-codes one 256-thread shape). */
int lkl_ctx_autotune(lkl_ctx* ctx, const double* poses, int N, unsigned flags,
                     float* best_cell, float* best_cell_x, int* best_block, int* best_chunk,
                     float* table, int* n_table);

/* Build the context's far field on the device from its receptor coordinates
 * and charges (ReceptorField::build, grid_dna.rs:44-149, the reference's values
 * bit for bit) without a host copy; replaces any grid given at creation. */
int lkl_ctx_build_field(lkl_ctx* ctx);
/* Origin and node counts of the context's far field (-1 if it has none). */
int lkl_ctx_field_info(const lkl_ctx* ctx, double* origin3, int* dims3);

/* ── multi-problem batch (no reference counterpart) ───────────────────────────
 * Many complexes (contexts) of uneven sizes, each with its own flag word and
 * pose count, scored by ONE kernel launch per step: one workgroup per
 * (problem, pose, ligand chunk) tile, largest problems first, the hardware
 * dispatcher back-filling freed slots.  Every problem's scores are bitwise those of its own
 * lkl_ctx_score call.  Contexts must outlive the batch; a context re-tiled
 * later is picked up automatically.  Desolvation is not batched. */
typedef struct lkl_multi lkl_multi;
lkl_multi* lkl_multi_create(lkl_ctx* const* ctxs, const unsigned* flags, const int* n_poses,
                            int n_problems);
void lkl_multi_destroy(lkl_multi* m);
/* Total poses P (the batch rows: problem 0's poses first). */
int lkl_multi_poses(const lkl_multi* m);
/* Launch shape: block in {64,128,256,512,1024}, ligand atoms per lane (0 =
 * whole ligand per tile). */
int lkl_multi_set_tiles(lkl_multi* m, int block, int chunk_mult);
/* Host buffers poses (P x 7) -> out (P x 2 of E_raw, V).  One step is H2D,
 * the batch kernel, D2H; with the graph on (default) the three are captured
 * once into a hipGraph and replayed. */
int lkl_multi_score(lkl_multi* m, const double* poses, double* out);
/* Device buffers, one kernel launch on `stream` (NULL = the batch's stream). */
int lkl_multi_score_async(lkl_multi* m, const double* d_poses, double* d_out, void* stream);
int lkl_multi_use_graph(lkl_multi* m, int enable);
/* Autotune (block, chunk_mult) on host poses; table (may be NULL) receives up
 * to 64 rows (block, chunk_mult, median ms). */
int lkl_multi_autotune(lkl_multi* m, const double* poses, int* best_block, int* best_mult,
                       float* table, int* n_table);
/* Load-balance trace (no counterpart in the reference; instrumentation for
 * the batch launch): with on, every following launch records per tile
 * {start, end} on the device's constant wall clock, the CU it ran on
 * (XCC, shader engine, CU as HIP's __smid) and its problem (batch order).
 * lkl_multi_trace waits for the device (the launch may be on any stream) and
 * copies the last launch's
 * records, 4 x uint64 per tile in launch order (at most cap tiles), stores
 * the clock's rate in kHz in *clock_khz (if not NULL) and returns the
 * launch's tile count. */
int lkl_multi_set_trace(lkl_multi* m, int on);
int lkl_multi_trace(lkl_multi* m, unsigned long long* out, int cap, int* clock_khz);

/* ── receptor far field and synthetic inputs ──────────────────────────────── */

/* grid_dna.rs:44-61: origin (3 doubles) and node counts (3 ints) of the 0.5 A
 * field box around the receptor; returns the node count or -1. */
long long lkl_field_dims(const float* r_xyz, int nr, double* origin3, int* dims3);

/* Replaces ReceptorField::build (grid_dna.rs:44-149), run on the GPU: phi_out
 * (lkl_field_dims() floats, x fastest) receives, per node, the f32 sum in
 * atom order of q_i/d^2 over 10 < d <= 30 A -- the reference's values bit for
 * bit.  Returns 0 / -1. */
int lkl_field_build(const float* r_xyz, const float* r_ele, int nr, float* phi_out);

/* Replaces lee_richards_sasa (lr_sasa.rs:277-291; cpydock.rs:287-332 calls it
 * on each unbound monomer's heavy atoms), run on the GPU: out[i] = the Lee &
 * Richards solvent-accessible area (A^2) of atom i, xyz n x 3 f64, radii
 * without the probe.  The reference uses probe 1.4 and 20 slices.  f64 in the
 * reference's operation order; neighbour order does not change an area (arcs
 * are sorted before summing).  Host buffers, blocking.  Returns 0 / -1. */
int lkl_lr_sasa(const double* xyz, const double* radii, int n, double probe, int n_slices,
                double* out);

/* Seeded random sequential addition of n atoms in the ellipsoid with semi-axes
 * (a, b, c) at minimum separation dmin, coordinates quantised to 2^-10 A
 * (synthetic benchmark inputs; no reference counterpart).  Returns the number
 * of atoms placed; *attempts receives the candidates drawn. */
int lkl_synth_rsa(int n, double a, double b, double c, double dmin, unsigned long long seed,
                  long long max_attempts, double* xyz, long long* attempts);

/* ── GSO driver (lib.rs / swarm.rs / glowworm.rs) ──────────────────────────── */

typedef struct lkl_gso lkl_gso;

/* Caller's CPU scoring path (the reference's per-pose energy_grid fallback,
 * dna.rs:1081-1085): poses n x 7 (w, x, y, z, tx, ty, tz) -> scores n.
 * Returns 0 on success. */
typedef int (*lkl_score_fallback)(void* user, const double* poses, int n, double* scores);

/* GSO over n_swarms independent swarms of one complex (GSO::new, lib.rs:47-64;
 * every swarm seeds its own StdRng with `seed`, simulator.rs:184).  positions:
 * sum(swarm_sizes) rows (tx, ty, tz, w, x, y, z) as in swarm.rs:38-42.  family:
 * 0 dna, 1 vdw, 2 pydock, 3 cpydock (flag word of gpu_family.rs:187-194 and the
 * score fold).  ctx scores on the GPU; fallback (may be NULL) is called only if
 * the GPU call fails or ctx is NULL.  Returns NULL on bad arguments. */
lkl_gso* lkl_gso_create(lkl_ctx* ctx, int family, int n_swarms, const int* swarm_sizes,
                        const double* positions, unsigned long long seed,
                        lkl_score_fallback fallback, void* fallback_user);
void lkl_gso_destroy(lkl_gso* gso);
/* Swarm::update_luciferin (swarm.rs:70-103) for all swarms: the glowworms that
 * moved (or all, at step 0) are scored in one launch.  0 / -1. */
int lkl_gso_update_luciferin(lkl_gso* gso);
/* Swarm::movement_phase (swarm.rs:105-174) for every swarm.  0 / -1. */
int lkl_gso_movement_phase(lkl_gso* gso);
/* With restraints or a membrane on the context (lkl_ctx_set_interface) the GSO
 * scores carry the restraint fold of dna.rs:989-995 (interface counts from
 * the same GPU pass); the reference keeps such runs off the batch path.
 * GSO::run (lib.rs:66-78): `steps` x (update_luciferin, movement_phase). */
int lkl_gso_run(lkl_gso* gso, int steps);
/* GSO::run's steps with the glowworm state resident on the device (no
 * reference counterpart; replaces the per-step host round trip of
 * swarm.rs:70-174): each step is the scoring launch over all poses, the
 * luciferin update and the movement as kernels on `stream` (NULL = a
 * stream of the driver's own), with every swarm's StdRng draws made on the host in the
 * reference order and uploaded once.  The movement's slerp uses correctly
 * rounded acos/sin, so poses can differ from lkl_gso_run's (host libm) in the
 * last bit; decisions (neighbours, moves) are the same unless a comparison
 * sits within that rounding.  The state is copied back to the host at the end
 * (neighbour ids are not kept, only their count).  Needs a GPU context; no
 * fallback.  0 / -1. */
int lkl_gso_device_run(lkl_gso* gso, int steps, void* stream);
/* Number of glowworms over all swarms. */
int lkl_gso_size(const lkl_gso* gso);
/* Per glowworm (swarm order): pose (w, x, y, z, tx, ty, tz), luciferin, last
 * score, vision range, neighbour count, moved flag; any pointer may be NULL. */
int lkl_gso_state(const lkl_gso* gso, double* pose7, double* luciferin, double* scoring,
                  double* vision, int* n_neighbors, unsigned char* moved);
/* Replace the poses of the glowworms selected by mask (NULL = all) and mark
 * them moved (no reference counterpart: used by surrogate-screened moves). */
int lkl_gso_set_poses(lkl_gso* gso, const double* pose7, const unsigned char* mask);
/* Surrogate-screened movement (no reference counterpart), first half: plan
 * every swarm's GSO move exactly as lkl_gso_movement_phase would (same draws,
 * same vision update) and write the planned poses (rows w, x, y, z, tx, ty, tz;
 * the current pose where a glowworm has no brighter neighbour) into moves.
 * Nothing moves until lkl_gso_commit.  0 / -1. */
int lkl_gso_plan(lkl_gso* gso, double* moves);
/* Second half: glowworms with accept[k] != 0 take pose7[k] and its exact score
 * scores[k]; no glowworm is marked moved (the next update_luciferin scores
 * nothing).  0 / -1. */
int lkl_gso_commit(lkl_gso* gso, const double* pose7, const double* scores,
                   const unsigned char* accept);
/* GPU launches, fallback calls and poses scored on the GPU so far. */
int lkl_gso_stats(const lkl_gso* gso, long long* gpu_launches, long long* fallback_calls,
                  long long* scored);
/* Swarm::save (swarm.rs:232-271): <directory>/gso_<step>.out of one swarm. */
int lkl_gso_save(const lkl_gso* gso, int swarm, int step, const char* directory);
/* rand 0.7.3 StdRng gen::<f64>() stream of seed_from_u64(seed) (qt.rs:474-487 KAT). */
void lkl_rng_f64s(unsigned long long seed, int n, double* out);
/* Quaternion::slerp (qt.rs:80-104) of (w, x, y, z) quaternions. */
void lkl_quat_slerp(const double* a, const double* b, double t, double* out);
/* The device GSO step's math on the host (for tests): correctly rounded sin
 * (0 <= x <= 1.6) and acos (0 <= d <= 1) by double-double evaluation, and
 * the slerp of qt.rs:80-104 built on them. */
double lkl_sin_cr(double x);
double lkl_acos_cr(double d);
void lkl_quat_slerp_cr(const double* a, const double* b, double t, double* out);

/* ── learned surrogate (no reference counterpart: the screening step that
 * ranks candidate moves before exact scoring) ────────────────────────────────
 * Features per pose: 16 receptor type groups x 16 ligand type groups x 24
 * distance bins (1.25 A, 0-30 A) of atom-pair counts, index
 * (g_rec * 16 + g_lig) * 24 + bin, log1p-scaled to bf16 (6,144 values).  The
 * pair distance is f32: x = f32(R(q) v + t) (f64 transform of the scoring
 * path), d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz)); counted when d2 < 900,
 * bin = largest b with 1.5625 b^2 <= d2. */
typedef struct lkl_feat lkl_feat;
/* Receptor/ligand f32 coordinates (n x 3, ligand in its reference frame) and
 * type groups (modulo 16).  NULL on error (lkl_last_error). */
lkl_feat* lkl_feat_create(const float* r_xyz, const unsigned char* r_group, int nr,
                          const float* l_xyz, const unsigned char* l_group, int nl);
void lkl_feat_destroy(lkl_feat* f);
/* Grid mode: the receptor's per-(group, bin) counts are tabulated on nodes of
 * spacing h (0.25, 0.5, 1 or 2 A) over the receptor box grown by 30 A + h (u8
 * per count, 384 bytes per node; creation fails if a count exceeds 255), and a
 * pose's features sum, per ligand atom, the counts of its nearest node
 * (ix = floor(fmaf(f32(x) - ox, 1/h, 0.5)) per axis).  Requires nr < 65,536
 * and nl <= 16,384. */
lkl_feat* lkl_feat_create_grid(const float* r_xyz, const unsigned char* r_group, int nr,
                               const float* l_xyz, const unsigned char* l_group, int nl,
                               double spacing);
/* Origin, node counts and spacing of a grid-mode object; returns the node
 * count (0 for an exact-mode object, -1 on error). */
long long lkl_feat_grid_info(const lkl_feat* f, double* origin3, int* dims3, double* spacing);
/* Copy nodes [first, first + count) of the count table (384 bytes per node,
 * order g_rec * 24 + bin, node index (iz * ny + iy) * nx + ix).  0 / -1. */
int lkl_feat_grid_read(const lkl_feat* f, long long first, long long count, unsigned char* out);
/* Device buffers: poses N x 7 (w, x, y, z, tx, ty, tz) -> bf16 bits N x 6144
 * and/or u32 counts N x 6144 (either may be NULL), enqueued on `stream`
 * (hipStream_t; NULL = the object's stream).  0 / -1. */
int lkl_feat_compute(lkl_feat* f, const double* d_poses, int N, unsigned short* d_out,
                     unsigned* d_counts, void* stream);
/* Host buffers, blocking.  0 / -1. */
int lkl_feat_compute_host(lkl_feat* f, const double* poses, int N, unsigned short* out,
                          unsigned* counts);

/* MLP 6144 -> 512 -> 512 -> 512 -> 1 with ReLU; bf16 weights/activations,
 * fp32 accumulation on the matrix cores, fp32 master weights and Adam state.
 * Flat parameter layout (fp32): W1 [6144][512] (input-major), b1 [512],
 * W2 [512][512], b2, W3 [512][512], b3, W4 [512], b4 [1]: 3,672,065 values. */
typedef struct lkl_mlp lkl_mlp;
long long lkl_mlp_param_count(void);
/* He-uniform weights from a splitmix64 stream of `seed`, zero biases. */
lkl_mlp* lkl_mlp_create(unsigned long long seed);
void lkl_mlp_destroy(lkl_mlp* n);
int lkl_mlp_get_params(lkl_mlp* n, float* out);
int lkl_mlp_set_params(lkl_mlp* n, const float* in);
int lkl_mlp_get_grads(lkl_mlp* n, float* out);
/* Device buffers: x (B x 6144 bf16 bits) -> y (B fp32) on `stream` (NULL = the
 * object's stream).  B must be a positive multiple of 128.  0 / -1. */
int lkl_mlp_forward(lkl_mlp* n, const void* d_x, int B, float* d_y, void* stream);
/* Loss 0.5 * mean((y - t)^2) and its gradient into the object's gradient
 * buffer (device x, t; B multiple of 128); *loss_out (host, may be NULL)
 * receives the loss after a synchronize.  0 / -1. */
int lkl_mlp_grad(lkl_mlp* n, const void* d_x, const float* d_t, int B, float* loss_out);
/* One bias-corrected Adam step on the stored gradient.  0 / -1. */
int lkl_mlp_adam(lkl_mlp* n, float lr, float beta1, float beta2, float eps);
int lkl_mlp_sync(lkl_mlp* n);
/* Host-buffer variants (blocking). */
int lkl_mlp_forward_host(lkl_mlp* n, const unsigned short* x, int B, float* y);
int lkl_mlp_grad_host(lkl_mlp* n, const unsigned short* x, const float* t, int B, float* loss);
/* C (M x N fp32) = sum_k A(m, k) B(k, n) on the matrix cores, bf16 operands
 * given as bits with element strides A(m, k) = A[m*sam + k*sak],
 * B(k, n) = B[k*sbk + n*sbn]; M, N multiples of 128, K of 64.  0 / -1. */
int lkl_gemm_bf16_host(const unsigned short* A, long a_len, long sam, long sak,
                       const unsigned short* B, long b_len, long sbk, long sbn, int M, int N,
                       int K, float* C);

/* Screening step (no reference counterpart).  Candidates: n base poses
 * (device, rows w, x, y, z, tx, ty, tz) -> n * n_cand poses, glowworm-major;
 * candidate 0 is the base pose, candidate c > 0 adds trans_radius * U(-1, 1) to
 * each translation component and rot_sigma * U(-1, 1) to each quaternion
 * component (then renormalised).  Draws are counter-based:
 * u = (splitmix64(seed ^ splitmix64(8 i + d)) >> 11) * 2^-53 for candidate i,
 * draw d (0-3 quaternion, 4-6 translation).  0 / -1. */
int lkl_screen_candidates(const double* d_base, int n, int n_cand, double trans_radius,
                          double rot_sigma, unsigned long long seed, double* d_cand, void* stream);
/* Per glowworm, the k (1..8) candidates with the highest prediction (ties:
 * lower index first): indices (n x k) and poses (n x k x 7), device buffers on
 * `stream`.  0 / -1. */
int lkl_screen_select(const float* d_pred, const double* d_cand, int n, int n_cand, int k,
                      int* d_pick, double* d_sel, void* stream);

#ifdef __cplusplus
}
#endif
#endif
