// lkl_device.hpp — device-side building blocks of the gfx950 pose-scoring path.
//
// Numerics contract (what makes the FP32 kernel agree with the f64 CPU grid
// path of dna.rs:834-998 / pydock.rs:353-447 / vdw.rs:73-151 /
// cpydock.rs:556-676 to ~1e-7 relative):
//   * the pose transform is f64 with contraction disabled, evaluated in the
//     order of qt.rs:67-74 + qt.rs:122-128 + dna.rs:868-873, so the
//     transformed coordinates equal the CPU's bit for bit;
//   * pair distances are filtered in f32 from a hi/lo split of the f64
//     ligand coordinate (dx = (x_hi - r_x) + x_lo, relative error ~1e-7);
//   * pairs within 1e-3 A^2 of the 10 A cutoff, and every pair entering the
//     6.4 A desolvation window, are re-measured in exact f64 (same expression
//     order as the CPU), so cutoff and window membership match the CPU exactly;
//   * per-pair terms are f32, accumulated as int64 fixed point (E and S at
//     2^-37, V at 2^-26): integer sums are associative, so per-pose results
//     are bitwise identical for every launch shape, cell size, pose batching
//     and multi-problem packing; the per-term rounding (<= 7e-12 for E,
//     <= 7.5e-9 for V) is far below the f32 term error;
//   * the far field is sampled in f64 exactly like grid_dna.rs:159-197
//     (fractional offsets rounded to f32), from the same f32 phi values.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lkl {

// Constants of full_score.cu:15-19 / dna.rs:20-42 / cpydock.rs:36-45.
constexpr double ES_CAP = 1.0 * 4.0 / 332.0;  // clamp of the raw electrostatics
constexpr float ES_CAPF = (float)ES_CAP;
constexpr float LJ_CAP = 1.0f;
constexpr float CP_W = 6.0f;
constexpr float CP_F = 0.75f;
constexpr double NEAR2 = 100.0;        // (10 A)^2, inclusive
constexpr float NEAR_HI = 100.001f;    // f32 filter: pairs above never count
constexpr float NEAR_LO = 99.999f;     // f32 filter: pairs above are re-measured in f64
constexpr double SOLV2 = 6.4 * 6.4;    // 40.96000000000001 (cpydock.rs:44)
constexpr float SOLV_HI = 40.97f;      // f32 filter of the desolvation window
constexpr float WIN2 = 100.5f;         // cell window radius^2 (10.025 A, > 10 A + rounding)
constexpr float SOLV_KEY_T = 40.96f;   // largest float <= SOLV2
constexpr double MIN_INIT = 10000.0;   // HUGE_DISTANCE of cpydock.rs:592
// neighbour lists are padded to whole blocks of LIST_PAD entries; each block is
// one step of a 4-lane subgroup of the grouped walk, stored lane-major (entry
// 4q + r of the block, q < 8, r < 4, at position 8r + q) so a lane loads its 8
// entries of the step with one 16-byte read
constexpr int LIST_PAD = 32;
constexpr float DUMMY_XYZ = 1.0e15f;   // coordinate of the padding record (d2 ~ 3e30, finite)

// Family flag word (full_score.cu:23-25) + desolvation bit.
constexpr unsigned F_FAR = 1u, F_ELEC = 2u, F_CLASH = 4u, F_DESOLV = 16u;

// One complex as seen by the kernels (passed by value).
struct DevComplex {
    // receptor, sorted by cell (x fastest, then y, then z)
    const float4* rxyzq;  // x, y, z, q
    const float4* rpar;   // sqrt(eps), vdw radius, heavy (0/1), 0
    const int* cstart;    // ncx*ncy*ncz + 1 prefix offsets into the sorted atoms
    const unsigned char* rflag;
    const float* rdes;
    const float* rasa;
    int nr, ncx, ncy, ncz;
    double cox, coy, coz;           // cell-grid origin (A)
    float csyz, inv_csyz, inv_csx;  // cell edge along y and z, 1/edge (y, z) and 1/edge (x)
    float rwin, rwin2;              // window radius (A) and its square
    // work grid: receptor atoms within rwin + (node half-diagonal) of each
    // node (NULL = unsorted near field); nodes at wo + wh * i
    const unsigned short* wgrid;
    int wnx, wny, wnz;
    float wox, woy, woz, inv_wh;
    // neighbour lists (NULL = cell-row walk): for each node of a grid of
    // spacing 1/inv_nh at no + i/inv_nh, the sorted receptor indices of the
    // atoms within rwin + (node half-diagonal) of the node, in sorted order,
    // each list padded to a multiple of LIST_PAD with index nr (a dummy record
    // far from everything) and its blocks stored lane-major.  noff holds
    // nodes + 1 offsets into nidx.
    const unsigned* noff;
    const unsigned short* nidx;
    // the same lists with 32-bit indices (receptors of more than 65,535
    // atoms, whose dummy index nr does not fit 16 bits); when set, nidx is
    // NULL and the walks read these instead
    const unsigned* nidx32;
    // the sorted receptor records interleaved for the list walk: rrec[2p] =
    // rxyzq[p], rrec[2p + 1] = rpar[p] (one 32-byte record per atom, nr + 1)
    // except that the radius carries the sign bit where the atom is excluded
    // from desolvation (rflag[p] != 0, and the padding record)
    const float4* rrec;
    // position in nidx of a block of LIST_PAD entries holding nr, past the
    // last list: a lane whose list has ended loads its indices from there
    unsigned ndum;
    // parameter types (NULL = none): when the receptor's rpar records take at
    // most 256 distinct values, rtab[ntype] holds them and rtyp holds one type
    // byte per sorted atom and the dummy, zero-padded to whole 32-bit words;
    // the list walk then reads the parameters from an LDS copy of both
    const unsigned* rtyp;
    const float4* rtab;
    int ntype;
    // ligand parameter types (NULL = none): the distinct lpar records (at most
    // 256) and one type byte per sorted ligand atom, zero-padded to whole
    // words; with the receptor's types the list walk may read each pair's LJ
    // and clash factors from an LDS table over (ligand type, receptor type)
    const unsigned char* lty;
    const float4* lttab;
    int nltype;
    int nnx, nny, nnz;
    float nox, noy, noz, inv_nh;
    // ligand, spatially sorted (Morton order of the reference coordinates)
    const float4* lxyzq;
    const float4* lpar;
    const unsigned char* lflag;
    const float* ldes;
    const float* lasa;
    int nl;
    // records per pose when each pose brings its own ligand coordinates
    // (lxyzq then holds N x nl records, pose-major); 0 = one shared set
    long long lstride;
    // dispatch order of the poses (NULL = pose = tile): tile L scores pose
    // porder[L]; set for one-tile-per-pose launches ordered by expected cost
    const int* porder;
    // far field (grid_dna.rs:33-40), x fastest
    const float* phi;
    int nx, ny, nz;
    double ox, oy, oz, sp;
    double inv_sp;  // 1 / sp, exact when sp_pow2
    int sp_pow2;    // sp is a power of two: x / sp == x * inv_sp
    // corner-packed copy (NULL = sample phi): per cell (ix, iy, iz) < (nx-1,
    // ny-1, nz-1), x fastest, two float4 {c000, c100, c010, c110} and
    // {c001, c101, c011, c111} -- the cell's 8 corners in one 32-byte read
    const float4* phi8;
};

struct Mat {
    double m00, m01, m02, m10, m11, m12, m20, m21, m22, tx, ty, tz;
};

// qt.rs:67-74 — rotation matrix of a unit quaternion, row-major.
__device__ __forceinline__ Mat pose_matrix(const double* __restrict__ p) {
#pragma clang fp contract(off)
    const double w = p[0], x = p[1], y = p[2], z = p[3];
    Mat M;
    M.m00 = 1. - 2. * (y * y + z * z);
    M.m01 = 2. * (x * y - w * z);
    M.m02 = 2. * (x * z + w * y);
    M.m10 = 2. * (x * y + w * z);
    M.m11 = 1. - 2. * (x * x + z * z);
    M.m12 = 2. * (y * z - w * x);
    M.m20 = 2. * (x * z - w * y);
    M.m21 = 2. * (y * z + w * x);
    M.m22 = 1. - 2. * (x * x + y * y);
    M.tx = p[4];
    M.ty = p[5];
    M.tz = p[6];
    return M;
}

// qt.rs:122-128 then + translation (dna.rs:868-873).
__device__ __forceinline__ void transform(const Mat& M, float vxf, float vyf, float vzf,
                                          double& x, double& y, double& z) {
#pragma clang fp contract(off)
    const double vx = vxf, vy = vyf, vz = vzf;
    x = M.m00 * vx + M.m01 * vy + M.m02 * vz + M.tx;
    y = M.m10 * vx + M.m11 * vy + M.m12 * vz + M.ty;
    z = M.m20 * vx + M.m21 * vy + M.m22 * vz + M.tz;
}

// f64 squared distance in the CPU's expression order (dna.rs:940-943).
__device__ __forceinline__ double exact_d2(double x, double y, double z, float4 r) {
#pragma clang fp contract(off)
    const double dx = x - (double)r.x;
    const double dy = y - (double)r.y;
    const double dz = z - (double)r.z;
    return dx * dx + dy * dy + dz * dz;
}

// grid_dna.rs:159-197 — f64 trilinear interpolation, offsets rounded to f32.
// Branch-free: outside the grid (or for NaN input) the corner reads go to cell
// 0 and the result is selected to 0, so the loads of several samples can be
// issued together.  With a power-of-two spacing the division is a
// multiplication by the exact reciprocal (the same IEEE result).  phi_fetch
// issues a sample's corner reads and phi_lerp interpolates them, so a caller
// can put other work between the two while the reads are in flight.
struct PhiTap {
    float c[8];       // corners 000, 100, 010, 110, 001, 101, 011, 111
    float tx, ty, tz;  // offsets in the cell, rounded to f32
    bool in;
};

__device__ __forceinline__ void phi_fetch(const DevComplex& C, double x, double y, double z, PhiTap& T) {
#pragma clang fp contract(off)
    double fx, fy, fz;
    if (C.sp_pow2) {
        fx = (x - C.ox) * C.inv_sp;
        fy = (y - C.oy) * C.inv_sp;
        fz = (z - C.oz) * C.inv_sp;
    } else {
        fx = (x - C.ox) / C.sp;
        fy = (y - C.oy) / C.sp;
        fz = (z - C.oz) / C.sp;
    }
    const double ix = floor(fx), iy = floor(fy), iz = floor(fz);
    const bool in = ix >= 0.0 && iy >= 0.0 && iz >= 0.0 && ix + 1.0 < (double)C.nx &&
                    iy + 1.0 < (double)C.ny && iz + 1.0 < (double)C.nz;
    const size_t jx = in ? (size_t)ix : 0, jy = in ? (size_t)iy : 0, jz = in ? (size_t)iz : 0;
    T.in = in;
    T.tx = in ? (float)(fx - ix) : 0.0f;
    T.ty = in ? (float)(fy - iy) : 0.0f;
    T.tz = in ? (float)(fz - iz) : 0.0f;
    if (C.phi8) {
        const size_t c = (jz * (size_t)(C.ny - 1) + jy) * (size_t)(C.nx - 1) + jx;
        const float4 lo = C.phi8[2 * c], hi = C.phi8[2 * c + 1];
        T.c[0] = lo.x; T.c[1] = lo.y; T.c[2] = lo.z; T.c[3] = lo.w;
        T.c[4] = hi.x; T.c[5] = hi.y; T.c[6] = hi.z; T.c[7] = hi.w;
    } else {
        const size_t sy = (size_t)C.nx, sz = (size_t)C.nx * (size_t)C.ny;
        const size_t b = jz * sz + jy * sy + jx;
        const float* __restrict__ g = C.phi;
        T.c[0] = g[b]; T.c[1] = g[b + 1];
        T.c[2] = g[b + sy]; T.c[3] = g[b + sy + 1];
        T.c[4] = g[b + sz]; T.c[5] = g[b + sz + 1];
        T.c[6] = g[b + sz + sy]; T.c[7] = g[b + sz + sy + 1];
    }
}

__device__ __forceinline__ double phi_lerp(const PhiTap& T) {
#pragma clang fp contract(off)
    const double tx = T.tx, ty = T.ty, tz = T.tz;
    const double c000 = T.c[0], c100 = T.c[1], c010 = T.c[2], c110 = T.c[3];
    const double c001 = T.c[4], c101 = T.c[5], c011 = T.c[6], c111 = T.c[7];
    const double c00 = c000 * (1.0 - tx) + c100 * tx;
    const double c10 = c010 * (1.0 - tx) + c110 * tx;
    const double c01 = c001 * (1.0 - tx) + c101 * tx;
    const double c11 = c011 * (1.0 - tx) + c111 * tx;
    const double c0 = c00 * (1.0 - ty) + c10 * ty;
    const double c1 = c01 * (1.0 - ty) + c11 * ty;
    const double r = c0 * (1.0 - tz) + c1 * tz;
    return T.in ? r : 0.0;
}

__device__ __forceinline__ double sample_phi(const DevComplex& C, double x, double y, double z) {
    PhiTap T;
    phi_fetch(C, x, y, z, T);
    return phi_lerp(T);
}

// Monotone f32 key of an f64 squared distance that keeps membership of the
// desolvation window exact: key <= bits(40.96f) iff d2 <= 6.4*6.4 (f64).
// Non-negative floats order like their bit patterns, so ds_min_u32 on keys
// is a min on distances.
__device__ __forceinline__ unsigned solv_key(double d2) {
    float f = (float)d2;
    if (d2 <= SOLV2) f = fminf(f, SOLV_KEY_T);
    else f = fmaxf(f, __uint_as_float(__float_as_uint(SOLV_KEY_T) + 1u));
    return __float_as_uint(f);
}

// cpydock.rs:648-665 — desolvation term of one atom, clamp placed as on the CPU.
__device__ __forceinline__ double solv_term(double m, float asaf, float desf) {
    const double asa = asaf;
    double solv = 0.0;
    if (m <= SOLV2 && m > 0.0 && asa > 0.0) solv = -10.0 * sqrt(m) + 65.0;
    if (solv > asa) solv = asa;
    return solv * (double)desf;
}

// Fixed-point scales of the per-pose accumulators.
constexpr double E_SCALE = 137438953472.0;  // 2^37: raw electrostatics, desolvation
constexpr double V_SCALE = 67108864.0;      // 2^26: LJ + clash
constexpr float E_SCALE_F = 137438953472.0f;
constexpr float V_SCALE_F = 67108864.0f;
constexpr float ES_CAP_SF = (float)(ES_CAP * E_SCALE);
constexpr long long ES_CAP_I = 1655891006LL;  // round(4/332 * 2^37)

__device__ __forceinline__ long long to_fix(double x, double scale) {
    return __double2ll_rn(x * scale);
}

// Butterfly sum over a 64-lane wavefront (integer: exact, order-free).
__device__ __forceinline__ long long wave_sum(long long x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// XCD-aware tile order: the hardware places workgroup b on XCD b % 8, so give
// each XCD a contiguous range of tiles (neighbouring tiles then share one L2).
// Bijective on [0, total).
__device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned total) {
    const unsigned q = total / 8u, r = total % 8u, x = b % 8u, slot = b / 8u;
    return x < r ? x * (q + 1u) + slot : r * (q + 1u) + (x - r) * q + slot;
}

}  // namespace lkl
