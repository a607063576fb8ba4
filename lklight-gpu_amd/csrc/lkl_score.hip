// lkl_score.hip — gfx950 kernels + C ABI of the batched GSO pose-scoring path.
//
// One workgroup scores one pose: its lanes walk the (spatially sorted) ligand
// atoms, transform each one in f64 (device twin of qt.rs:67-74,122-128), gather
// the far field, and scan the receptor cell rows within 10 A.  Receptor atoms
// are stored cell-sorted as packed float4 records, so a cell row (x-adjacent
// cells) is one contiguous run and every candidate is one 16-byte load.  The
// per-pose sums are int64 fixed point, reduced by wave butterflies and then
// across waves — exact and order-free, replacing the per-block double atomics
// of full_score.cu:332-335.
//
// cpydock's two-stage desolvation (full_score.cu:477-569: global atomicMin
// scratch, 0x7F memset, second launch) is fused into the same pass: the
// receptor-side minima live in LDS (ds_min_u32 on monotone f32 keys) and are
// reduced by the same workgroup after a barrier.  Receptors too large for LDS
// use a per-pose global scratch slice instead.
#include "lkl_device.hpp"
#include "../../include/lklight_gpu.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

using namespace lkl;

#define LKL_API extern "C" __attribute__((visibility("default")))

// ── error reporting + fault injection ─────────────────────────────────────────
static thread_local std::string t_err;

static void set_err(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    t_err = buf;
    fprintf(stderr, "[lklight-gpu] %s\n", buf);
}

// Library-internal: lets the other translation units report through
// lkl_last_error().
void lkl_set_error_text(const char* msg) { set_err("%s", msg); }

// LKL_FAIL_AT=<stage> makes the named stage of every call fail, so host-only
// tests can drive each failure path (stages: probe, alloc, upload, launch,
// sync, readback).
static bool fail_at(const char* stage) {
    const char* e = getenv("LKL_FAIL_AT");
    return e && strcmp(e, stage) == 0;
}

#define HIPCK(stage, expr)                                                            \
    do {                                                                              \
        hipError_t _e = (expr);                                                       \
        if (_e != hipSuccess || fail_at(stage)) {                                     \
            set_err("%s stage=%s: %s", __func__, stage,                               \
                    _e != hipSuccess ? hipGetErrorString(_e) : "injected failure");   \
            return -1;                                                                \
        }                                                                             \
    } while (0)

// ── kernels ───────────────────────────────────────────────────────────────────
enum DesMode { DES_NONE = 0, DES_LDS = 1, DES_GLOBAL = 2 };

// Near field of one ligand atom: scan the receptor cell rows whose cells meet
// the 10.025 A ball, accumulate the pair terms of full_score.cu:296-315 /
// dna.rs:951-969 and (with desolvation) the per-atom minima of
// cpydock.rs:620-627.
#ifndef LKL_NU
#define LKL_NU 4
#endif
#ifndef SQRT_W
#define SQRT_W(x) __builtin_amdgcn_sqrtf(x)
#endif
#ifndef SQRT_D
#define SQRT_D(x) __builtin_amdgcn_sqrtf(x)
#endif
constexpr int NU = LKL_NU;
// one-tile-per-pose launches of POSE_ORDER_MIN + 1 to POSE_ORDER_MAX poses
// are dispatched longest first: beyond one tile per CU, tiles share CUs (and,
// further on, wait for a free slot), and starting the costly ones first pairs
// them with cheap ones.  Above POSE_ORDER_MAX the launch runs many rounds of
// workgroup slots, its last round is a small share of it, and the order's
// pre-pass (a rank sort, quadratic in the range) costs more than it saves.
constexpr int POSE_ORDER_MIN = 256;
constexpr int POSE_ORDER_MAX = 4096;

// The pair-type entry of a ligand record lp and a receptor record rp: the
// walk's LJ factor sqrt(eps_j) V_SCALE sqrt(eps_i), (r_i + r_j)^2 and the clash
// threshold 0.75 (r_i + r_j) when both atoms are heavy (else 0: no clash),
// each formed by the f32 operations pair_term applies to the two records, so a
// walk over these entries gives the same bits.
__device__ __forceinline__ float4 pair_entry(float4 lp, float4 rp) {
    const float svjs = lp.x * V_SCALE_F;
    const float rr = lp.y + fabsf(rp.y);
    const bool clash = lp.z != 0.0f && rp.z != 0.0f;
    return make_float4(svjs * rp.x, rr * rr, clash ? CP_F * rr : 0.0f, 0.0f);
}

// The clash gap dmin - d, rounded on its own: not fused with the product that
// forms dmin, so a walk that reads dmin from the pair-type table and one that
// forms it from the two radii give the same bits.
__device__ __forceinline__ float clash_gap(float dmin, float d) {
#pragma clang fp contract(off)
    return dmin - d;
}

// One receptor candidate r (sorted index p) against ligand atom (x64, y64,
// z64): the pair terms of full_score.cu:296-315 / dna.rs:951-969 formed
// directly in fixed-point units, and (with desolvation) the per-atom minima of
// cpydock.rs:620-627.  (xh + xl) is the hi/lo f32 split of the f64 coordinate.
// With PRE the caller has already loaded the candidate's parameters (rp_pre);
// otherwise they are loaded for the pairs within the cutoff only.  With PAIR
// the preloaded rp_pre is the pair-type entry (pair_entry) instead of the
// receptor record (no desolvation).
template <bool ELEC, bool CLASH, int DES, bool PRE = false, bool PAIR = false>
__device__ __forceinline__ void pair_term(const DevComplex& C, const float4 r, int p, double x64, double y64,
                                          double z64, float xh, float yh, float zh, float xl, float yl,
                                          float zl, float qjs, float svjs, float rj, bool hj, bool lf,
                                          long long& e, long long& v, double& minL,
                                          unsigned* __restrict__ minR, float4 rp_pre = float4()) {
    const float dx = (xh - r.x) + xl;
    const float dy = (yh - r.y) + yl;
    const float dz = (zh - r.z) + zl;
    const float d2 = dx * dx + dy * dy + dz * dz;
    if (DES != DES_NONE && d2 <= SOLV_HI && !lf) {
        // in the interleaved records the receptor atom's desolvation flag is
        // the sign bit of its radius (set: excluded), so the walk needs no
        // extra gather for it
        if (PRE ? (__float_as_uint(rp_pre.y) >> 31) == 0u : C.rflag[p] == 0) {
            const double d2e = exact_d2(x64, y64, z64, r);
            minL = fmin(minL, d2e);
            atomicMin(&minR[p], solv_key(d2e));
        }
    }
    if (d2 <= NEAR_HI && !(d2 > NEAR_LO && exact_d2(x64, y64, z64, r) > NEAR2)) {
        const float4 rp = PRE ? rp_pre : C.rpar[p];
        const float inv = __builtin_amdgcn_rcpf(d2);
        if (ELEC) {  // clamp to +-ES_CAP in fixed-point units (one integer median)
            const float ae = qjs * r.w * inv;
            e += min(max(__float2int_rn(ae), -(int)ES_CAP_I), (int)ES_CAP_I);
        }
        if constexpr (PAIR) {  // rp: (LJ factor, (r_i + r_j)^2, clash threshold or 0, unused)
            const float t = rp.y * inv;
            const float p6 = t * t * t;
            float vp = rp.x * (p6 * p6 - 2.0f * p6);
            vp = fminf(vp, LJ_CAP * V_SCALE_F);
            if (CLASH) {
                const float d = SQRT_D(d2);
                if (d < rp.z) vp += (CP_W * V_SCALE_F) * clash_gap(rp.z, d);
            }
            v += (long long)__float2int_rn(vp);
        } else {
            const float svs = svjs * rp.x;
            const float rr = rj + fabsf(rp.y);
            const float t = (rr * rr) * inv;  // (r_i + r_j)^2 / d^2
            const float p6 = t * t * t;
            float vp = svs * (p6 * p6 - 2.0f * p6);
            vp = fminf(vp, LJ_CAP * V_SCALE_F);
            if (CLASH && hj && rp.z != 0.0f) {
                const float d = SQRT_D(d2);
                const float dmin = CP_F * rr;
                if (d < dmin) vp += (CP_W * V_SCALE_F) * clash_gap(dmin, d);
            }
            v += (long long)__float2int_rn(vp);
        }
    }
}

// Near field of one ligand atom by the cell-row walk: scan the receptor cell
// rows whose cells meet the 10.025 A ball.
template <bool ELEC, bool CLASH, int DES>
__device__ __forceinline__ void near_atom(const DevComplex& C, double x64, double y64, double z64,
                                          float qj, float svj, float rj, bool hj, bool lf,
                                          long long& e, long long& v, double& minL,
                                          unsigned* __restrict__ minR) {
    // terms are formed directly in fixed-point units (power-of-two scales are exact)
    const float qjs = qj * E_SCALE_F;
    const float svjs = svj * V_SCALE_F;
    const float xh = (float)x64, yh = (float)y64, zh = (float)z64;
    const float xl = (float)(x64 - (double)xh);
    const float yl = (float)(y64 - (double)yh);
    const float zl = (float)(z64 - (double)zh);
    // position relative to the cell-grid origin, in A (window arithmetic only;
    // the window radius's margin covers f32 error).  Cells are cs x cs in y/z
    // and may be narrower along x, which trims each row's x-range.
    const float px = fminf(fmaxf((float)(x64 - C.cox), -1.0e4f), 1.0e4f);
    const float py = fminf(fmaxf((float)(y64 - C.coy), -1.0e4f), 1.0e4f);
    const float pz = fminf(fmaxf((float)(z64 - C.coz), -1.0e4f), 1.0e4f);
    const float rw = C.rwin, rw2 = C.rwin2, cs = C.csyz, ics = C.inv_csyz;
    const int izlo = max(0, (int)floorf((pz - rw) * ics)), izhi = min(C.ncz - 1, (int)floorf((pz + rw) * ics));
    const int iylo = max(0, (int)floorf((py - rw) * ics)), iyhi = min(C.ncy - 1, (int)floorf((py + rw) * ics));
    for (int iz = izlo; iz <= izhi; ++iz) {
        const float gz = fmaxf(0.0f, fmaxf((float)iz * cs - pz, pz - (float)(iz + 1) * cs));
        const float gz2 = gz * gz;
        for (int iy = iylo; iy <= iyhi; ++iy) {
            const float gy = fmaxf(0.0f, fmaxf((float)iy * cs - py, py - (float)(iy + 1) * cs));
            const float g2 = gz2 + gy * gy;
            if (g2 > rw2) continue;
            const float b = SQRT_W(rw2 - g2);
            const int xlo = max(0, (int)floorf((px - b) * C.inv_csx));
            const int xhi = min(C.ncx - 1, (int)floorf((px + b) * C.inv_csx));
            if (xlo > xhi) continue;
            const int row = (iz * C.ncy + iy) * C.ncx;
            const int pb = C.cstart[row + xlo];
            const int pe = C.cstart[row + xhi + 1];
            // NU candidates per step: their records are requested together
            // (indices clamped into the row, so every load is in bounds) and
            // the distance tests then run on data already in flight.
            for (int p0 = pb; p0 < pe; p0 += NU) {
                float4 rq[NU];
#pragma unroll
                for (int k = 0; k < NU; ++k) rq[k] = C.rxyzq[min(p0 + k, pe - 1)];
#pragma unroll
                for (int k = 0; k < NU; ++k) {
                    if (p0 + k >= pe) break;
                    pair_term<ELEC, CLASH, DES>(C, rq[k], p0 + k, x64, y64, z64, xh, yh, zh, xl, yl, zl, qjs,
                                                svjs, rj, hj, lf, e, v, minL, minR);
                }
            }
        }
    }
}

// Node of the neighbour-list grid nearest to (x, y, z), or -1 outside the grid
// (then no receptor atom is within reach).  NaN coordinates give -1.
__device__ __forceinline__ int list_node(const DevComplex& C, double x, double y, double z) {
    const float fx = ((float)x - C.nox) * C.inv_nh + 0.5f;
    const float fy = ((float)y - C.noy) * C.inv_nh + 0.5f;
    const float fz = ((float)z - C.noz) * C.inv_nh + 0.5f;
    if (!(fx >= 0.0f && fx < (float)C.nnx && fy >= 0.0f && fy < (float)C.nny && fz >= 0.0f &&
          fz < (float)C.nnz))
        return -1;
    return ((int)fz * C.nny + (int)fy) * C.nnx + (int)fx;
}

// Near field of one ligand atom by its node's neighbour list [b, e): one
// aligned 8-byte load gives LIST_PAD receptor indices, whose records are then
// requested together; padding entries point at the dummy record, which fails
// every distance test.
// Near field of one ligand atom by its node's neighbour list [b, e): each step
// takes two aligned 8-byte loads of LIST_PAD receptor indices (the second
// replaced by padding past e) and then requests all 8 interleaved records
// (coordinates and parameters, 32 bytes each) together, so a wavefront keeps
// 8 gathers per lane in flight; padding entries point at the dummy record,
// which fails every distance test.
template <bool ELEC, bool CLASH, int DES>
__device__ __forceinline__ void near_list(const DevComplex& C, double x64, double y64, double z64,
                                          float qj, float svj, float rj, bool hj, bool lf, unsigned b,
                                          unsigned en, long long& e, long long& v, double& minL,
                                          unsigned* __restrict__ minR) {
    static_assert(LIST_PAD % 8 == 0, "lists are whole pairs of ushort4 vectors");
    constexpr int U = 2;  // list vectors per step
    const float qjs = qj * E_SCALE_F;
    const float svjs = svj * V_SCALE_F;
    const float xh = (float)x64, yh = (float)y64, zh = (float)z64;
    const float xl = (float)(x64 - (double)xh);
    const float yl = (float)(y64 - (double)yh);
    const float zl = (float)(z64 - (double)zh);
    const unsigned dn = (unsigned)C.nr;  // the dummy record
    for (unsigned p0 = b; p0 < en; p0 += 4u * U) {
        int id[4 * U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            uint4 q = make_uint4(dn, dn, dn, dn);
            if (u == 0 || p0 + 4u * u < en) {
                if (C.nidx32) {
                    q = *reinterpret_cast<const uint4*>(C.nidx32 + p0 + 4u * u);
                } else {
                    const ushort4 h = *reinterpret_cast<const ushort4*>(C.nidx + p0 + 4u * u);
                    q = make_uint4(h.x, h.y, h.z, h.w);
                }
            }
            id[4 * u] = (int)q.x;
            id[4 * u + 1] = (int)q.y;
            id[4 * u + 2] = (int)q.z;
            id[4 * u + 3] = (int)q.w;
        }
        float4 rq[4 * U], rp[4 * U];
#pragma unroll
        for (int k = 0; k < 4 * U; ++k) {
            rq[k] = C.rrec[2 * id[k]];
            rp[k] = C.rrec[2 * id[k] + 1];
        }
#pragma unroll
        for (int k = 0; k < 4 * U; ++k)
            pair_term<ELEC, CLASH, DES, true>(C, rq[k], id[k], x64, y64, z64, xh, yh, zh, xl, yl, zl, qjs, svjs,
                                              rj, hj, lf, e, v, minL, minR, rp[k]);
    }
}

// Near field of one ligand atom: neighbour lists (LIST; the context has them)
// or the cell-row walk.  Both visit a superset of the pairs within 10 A and
// apply the same tests, so the fixed-point sums are identical.  LIST is a
// template parameter so each kernel carries one walk (fewer registers).
template <bool LIST, bool ELEC, bool CLASH, int DES>
__device__ __forceinline__ void near_any(const DevComplex& C, double x64, double y64, double z64, float qj,
                                         float svj, float rj, bool hj, bool lf, long long& e, long long& v,
                                         double& minL, unsigned* __restrict__ minR) {
    if constexpr (LIST) {
        const int n = list_node(C, x64, y64, z64);
        if (n >= 0)
            near_list<ELEC, CLASH, DES>(C, x64, y64, z64, qj, svj, rj, hj, lf, C.noff[n], C.noff[n + 1], e, v,
                                        minL, minR);
    } else {
        near_atom<ELEC, CLASH, DES>(C, x64, y64, z64, qj, svj, rj, hj, lf, e, v, minL, minR);
    }
}

// ── work grid ────────────────────────────────────────────────────────────────
// Per node (spacing WORK_H over the receptor box grown by WORK_PAD): the number
// of receptor atoms within the window radius plus the node's half-diagonal
// (+ margin), so a ligand atom whose nearest node counts 0 has no receptor atom
// in its window.  Used to order each tile's ligand atoms by expected work and
// to skip those with none; it never changes a sum.
constexpr float WORK_H = 2.0f, WORK_PAD = 12.0f;
constexpr int WORK_TILE_MAX = 16384;  // ligand atoms per work-ordered tile (LDS arrays)
constexpr unsigned IFACE_PAD = 4;     // entries per padding unit of the interface pass's lists
constexpr int RTYPE_MAX = 256;        // receptor parameter types of the LDS table (one byte each)
constexpr int PAIR_TAB_MAX = 2048;    // entries of the pair-type LDS table (32 KB)

__global__ void workgrid_kernel(const float4* __restrict__ rec, int nr, float ox, float oy, float oz,
                                int nx, int ny, long long nodes, float r2, unsigned short* __restrict__ out) {
    const long long node = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (node >= nodes) return;
    const int ix = (int)(node % nx);
    const long long t = node / nx;
    const int iy = (int)(t % ny), iz = (int)(t / ny);
    const float px = ox + (float)ix * WORK_H, py = oy + (float)iy * WORK_H, pz = oz + (float)iz * WORK_H;
    unsigned cnt = 0;
    for (int i = 0; i < nr; ++i) {
        const float4 r = rec[i];
        const float dx = px - r.x, dy = py - r.y, dz = pz - r.z;
        cnt += (dx * dx + dy * dy + dz * dz <= r2) ? 1u : 0u;
    }
    out[node] = (unsigned short)min(cnt, 65535u);
}

// ── neighbour lists ──────────────────────────────────────────────────────────
// One workgroup owns an 8x8x4 brick of list nodes (one lane per node).  It
// streams the sorted receptor in index order, keeps the atoms within reach of
// the brick's box (order-preserving compaction into LDS) and every lane then
// tests them against its node: |node - r|^2 <= rad2 in f32 (rad carries a
// 0.01 A margin over the window radius plus the half-diagonal).  COUNT writes
// each node's list length rounded up to pad (a power of two); FILL writes the
// indices at the node's offset, in receptor order, then the padding index nr,
// with every block of LIST_PAD entries stored lane-major when lanemajor is set.
// Both instantiations evaluate the same test, so they agree on every list.
constexpr int NLX = 8, NLY = 8, NLZ = 4, NLT = NLX * NLY * NLZ;
constexpr int NLCAP = 2048;  // LDS candidate buffer (32 KB)

// position of list entry k in its block of LIST_PAD when blocks are lane-major
__device__ __forceinline__ unsigned lane_major(unsigned k) {
    static_assert(LIST_PAD == 32, "blocks of 8 steps x 4 lanes");
    return (k & ~31u) | ((k & 3u) << 3) | ((k >> 2) & 7u);
}

template <bool FILL>
__global__ void __launch_bounds__(NLT)
nlist_kernel(const float4* __restrict__ rec, int nr, float ox, float oy, float oz, float h, int nx, int ny,
             int nz, float rad, unsigned pad, int lanemajor, unsigned* __restrict__ cnt_off,
             unsigned short* __restrict__ idx, unsigned* __restrict__ idx32) {
#pragma clang fp contract(off)
    __shared__ float4 buf[NLCAP];
    __shared__ int wave_cnt[NLT / 64];
    __shared__ int count;
    const int tid = threadIdx.x;
    const int ix = blockIdx.x * NLX + (tid % NLX);
    const int iy = blockIdx.y * NLY + ((tid / NLX) % NLY);
    const int iz = blockIdx.z * NLZ + (tid / (NLX * NLY));
    const bool valid = ix < nx && iy < ny && iz < nz;
    const size_t node = valid ? ((size_t)iz * ny + iy) * nx + ix : 0;
    const float px = ox + (float)ix * h, py = oy + (float)iy * h, pz = oz + (float)iz * h;
    // brick box over its node positions (the same f32 expressions, monotone in i)
    const int bx = blockIdx.x * NLX, by = blockIdx.y * NLY, bz = blockIdx.z * NLZ;
    const float bx0 = ox + (float)bx * h, bx1 = ox + (float)(bx + NLX - 1) * h;
    const float by0 = oy + (float)by * h, by1 = oy + (float)(by + NLY - 1) * h;
    const float bz0 = oz + (float)bz * h, bz1 = oz + (float)(bz + NLZ - 1) * h;
    const float rad2 = rad * rad, pre2 = (rad + 0.01f) * (rad + 0.01f);
    unsigned k = 0, base_out = 0, end_out = 0;
    if (FILL && valid) {
        base_out = cnt_off[node];
        end_out = cnt_off[node + 1];  // the COUNT pass's padded length bounds every write
    }
    if (tid == 0) count = 0;
    __syncthreads();
    for (int base = 0; base < nr; base += NLT) {
        const int i = base + tid;
        bool take = false;
        float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
        if (i < nr) {
            a = rec[i];
            const float gx = fmaxf(0.f, fmaxf(bx0 - a.x, a.x - bx1));
            const float gy = fmaxf(0.f, fmaxf(by0 - a.y, a.y - by1));
            const float gz = fmaxf(0.f, fmaxf(bz0 - a.z, a.z - bz1));
            take = gx * gx + gy * gy + gz * gz <= pre2;
        }
        const unsigned long long m = __ballot(take);
        const int lane = tid & 63, w = tid >> 6;
        const int before = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wave_cnt[w] = __popcll(m);
        __syncthreads();
        int off = count;
        for (int q = 0; q < w; ++q) off += wave_cnt[q];
        if (take) buf[off + before] = make_float4(a.x, a.y, a.z, __int_as_float(i));
        __syncthreads();
        if (tid == 0) {
            int t = count;
            for (int q = 0; q < NLT / 64; ++q) t += wave_cnt[q];
            count = t;
        }
        __syncthreads();
        const int n = count;
        if (n > NLCAP - NLT || base + NLT >= nr) {
            if (valid) {
                for (int q = 0; q < n; ++q) {
                    const float4 b = buf[q];
                    const float dx = px - b.x, dy = py - b.y, dz = pz - b.z;
                    if (dx * dx + dy * dy + dz * dz <= rad2) {
                        if (FILL && base_out + k < end_out) {
                            const unsigned at = base_out + (lanemajor ? lane_major(k) : k);
                            if (idx32) idx32[at] = (unsigned)__float_as_int(b.w);
                            else idx[at] = (unsigned short)__float_as_int(b.w);
                        }
                        ++k;
                    }
                }
            }
            __syncthreads();
            if (tid == 0) count = 0;
            __syncthreads();
        }
    }
    if (!valid) return;
    const unsigned padded = (k + pad - 1u) / pad * pad;
    if (FILL) {
        for (; base_out + k < end_out; ++k) {
            const unsigned at = base_out + (lanemajor ? lane_major(k) : k);
            if (idx32) idx32[at] = (unsigned)nr;
            else idx[at] = (unsigned short)nr;
        }
    } else {
        cnt_off[node] = padded;
    }
}

// Bucket of a work count: NBKT (last) for no work, else descending
// quarter-octaves of the count (larger counts first).
#ifndef LKL_NBKT
#define LKL_NBKT 16
#endif
constexpr int NBKT = LKL_NBKT;
static_assert(NBKT >= 1 && NBKT <= 254, "bucket ids are stored in one byte");
__device__ __forceinline__ int work_bucket(unsigned w) {
    if (w == 0u) return NBKT;
    const int q = (int)(__log2f((float)w) * (float)(NBKT / 16));
    return NBKT - 1 - min(NBKT - 1, max(0, q));
}

__device__ __forceinline__ unsigned work_at(const DevComplex& C, double x, double y, double z) {
    const int ix = (int)floorf(((float)x - C.wox) * C.inv_wh + 0.5f);
    const int iy = (int)floorf(((float)y - C.woy) * C.inv_wh + 0.5f);
    const int iz = (int)floorf(((float)z - C.woz) * C.inv_wh + 0.5f);
    if (ix < 0 || iy < 0 || iz < 0 || ix >= C.wnx || iy >= C.wny || iz >= C.wnz) return 0u;
    return C.wgrid[((size_t)iz * C.wny + iy) * C.wnx + ix];
}

// Expected near-field work of a ligand atom: its neighbour-list length when
// the context has lists (exact), else the work-grid count.  0 = nothing in reach.
template <bool LIST>
__device__ __forceinline__ unsigned work_of(const DevComplex& C, double x, double y, double z) {
    if constexpr (LIST) {
        const int n = list_node(C, x, y, z);
        return n < 0 ? 0u : C.noff[n + 1] - C.noff[n];
    }
    return work_at(C, x, y, z);
}

// Near field (+ far field) of ligand atoms [j0, j1) of one pose, work-ordered:
// pass 1 transforms every atom, adds its far field and files it in a bucket
// (descending work-grid count; NBKT = none); pass 2 walks the windows of the
// atoms with work in bucket order, so the lanes of a wavefront walk windows of
// similar size.  With desolvation (whole-ligand tiles) the per-atom ligand term
// is added in whichever pass finishes the atom.  All threads of the block must
// call it (barriers); lds holds (j1 - j0) * 3 bytes.
template <bool LIST, bool FAR, bool ELEC, bool CLASH, int DES>
__device__ __forceinline__ void near_sorted(const DevComplex& C, const Mat& M, int j0, int j1, void* lds,
                                            unsigned* __restrict__ minR, long long& e, long long& v,
                                            long long& s) {
    __shared__ unsigned bcnt[NBKT + 1], boff[NBKT + 1];
    const int tid = (int)threadIdx.x, bs = (int)blockDim.x;
    const int nT = j1 - j0;
    unsigned short* list = reinterpret_cast<unsigned short*>(lds);
    unsigned char* bkt = reinterpret_cast<unsigned char*>(list + nT);
    for (int k = tid; k <= NBKT; k += bs) bcnt[k] = 0u;
    __syncthreads();
    for (int t = tid; t < nT; t += bs) {
        const float4 lb = C.lxyzq[j0 + t];
        double x, y, z;
        transform(M, lb.x, lb.y, lb.z, x, y, z);
        if (FAR && lb.w != 0.0f) e += to_fix((double)lb.w * sample_phi(C, x, y, z), E_SCALE);
        const int bb = work_bucket(work_of<LIST>(C, x, y, z));
        bkt[t] = (unsigned char)bb;
        atomicAdd(&bcnt[bb], 1u);
        // no receptor atom in reach: the desolvation minimum stays at its initial value
        if (DES != DES_NONE && bb == NBKT)
            s += to_fix(solv_term(MIN_INIT, C.lasa[j0 + t], C.ldes[j0 + t]), E_SCALE);
    }
    __syncthreads();
    if (tid == 0) {
        unsigned acc0 = 0u;
        for (int k = 0; k <= NBKT; ++k) {
            boff[k] = acc0;
            acc0 += bcnt[k];
        }
    }
    __syncthreads();
    for (int t = tid; t < nT; t += bs) list[atomicAdd(&boff[bkt[t]], 1u)] = (unsigned short)t;
    __syncthreads();
    const int nact = nT - (int)bcnt[NBKT];
    for (int k = tid; k < nact; k += bs) {
        const int j = j0 + list[k];
        const float4 lb = C.lxyzq[j];
        const float4 lp = C.lpar[j];
        double x, y, z;
        transform(M, lb.x, lb.y, lb.z, x, y, z);
        double minL = MIN_INIT;
        const bool lf = (DES != DES_NONE) ? (C.lflag[j] != 0) : false;
        near_any<LIST, ELEC, CLASH, DES>(C, x, y, z, lb.w, lp.x, lp.y, lp.z != 0.0f, lf, e, v, minL, minR);
        if (DES != DES_NONE) s += to_fix(solv_term(minL, C.lasa[j], C.ldes[j]), E_SCALE);
    }
    __syncthreads();  // the LDS lists may be reused by the caller
}

// Near field over neighbour lists, lanes grouped for coalescing: pass 1
// transforms every atom of [j0, j1), adds its far field and looks up its list
// [b, e) (P1 atoms per thread per round, loads issued together); the atoms
// with a non-empty list are then compacted, bucketed by the number of steps
// their list takes and in (Morton) order within a bucket, and the wavefronts
// take groups of GA consecutive active atoms from a shared counter.
// Within a wavefront each GL-lane subgroup walks one atom's list, lane r taking
// entries r, r + GL, ..., r + (GK-1) GL of each GK*GL-entry step (one 16-byte
// index load from the step's lane-major block): the records that a gather
// instruction requests for a subgroup are neighbours in the sorted receptor,
// the records of neighbouring atoms' lists (which overlap) share cache lines,
// and a lane keeps GK gathers in flight.  With desolvation the ligand-side minimum of an atom is reduced over
// its subgroup by shuffles (exact) and the per-atom term is added once.  With
// TYPED the receptor's parameter types (rtab, rtyp) are copied to LDS first and
// each candidate's parameters are read there (type byte, then table entry)
// while its record gather is in flight: one gather per candidate instead of
// two.  All threads of the block must call it (barriers).  LDS: [b u32 x nT]
// [e u32 x nT][active u16 x nT], then, with TYPED, from the next 16-byte
// boundary [rtab float4 x ntype][rtyp].
constexpr int GL = 4;        // lanes per list
constexpr int GA = 64 / GL;  // atoms per group (one wavefront)
constexpr int GK = 8;        // entries per lane per step
constexpr int GB = 4;        // step-count buckets of the active atoms

template <bool FAR, bool ELEC, bool CLASH, int DES, int TYPED = 0, bool WIDE = false>
__device__ __forceinline__ void near_groups(const DevComplex& C, const Mat& M, int j0, int j1, void* lds,
                                            unsigned* __restrict__ minR, long long& e, long long& v,
                                            long long& s) {
    __shared__ unsigned next_group;
    const int tid = (int)threadIdx.x, bs = (int)blockDim.x;
    const int lane = tid & 63, w = tid >> 6, nw = (bs + 63) >> 6;
    const int nT = j1 - j0;
    unsigned* ab = reinterpret_cast<unsigned*>(lds);
    unsigned* ae = ab + nT;
    unsigned short* act = reinterpret_cast<unsigned short*>(ae + nT);
    float4* ltab = reinterpret_cast<float4*>(reinterpret_cast<char*>(lds) + ((10u * (unsigned)nT + 15u) & ~15u));
    // TYPED 1: the receptor's type table; 2: the pair-type table, row
    // ligand type, column receptor type (pair_entry)
    const int ntab = TYPED == 2 ? C.nltype * C.ntype : TYPED ? C.ntype : 0;
    unsigned* ltyp = reinterpret_cast<unsigned*>(ltab + ntab);
    if constexpr (TYPED != 0) {  // visible to the walk after pass 1's barrier
        if constexpr (TYPED == 2) {
            for (int k = tid; k < ntab; k += bs) ltab[k] = pair_entry(C.lttab[k / C.ntype], C.rtab[k % C.ntype]);
        } else {
            for (int k = tid; k < C.ntype; k += bs) ltab[k] = C.rtab[k];
        }
        const int nwd = (C.nr + 4) >> 2;  // (nr + 1) type bytes in whole words
        for (int k = tid; k < nwd; k += bs) ltyp[k] = C.rtyp[k];
    }
    if (tid == 0) next_group = 0u;
    // The far field is sampled after the compaction by the tile's last quarter
    // of wavefronts (at least one), which then join the walk: the random phi
    // reads wait on those waves while the others walk.
    constexpr int PHI_U = 2;  // atoms per lane in flight (more spill the walk's registers, and gain nothing)
    constexpr int P1 = 3;
    for (int t0 = tid; t0 < nT; t0 += P1 * bs) {
        double px[P1], py[P1], pz[P1];
        float q[P1];
        unsigned nb[P1], ne[P1];
#pragma unroll
        for (int u = 0; u < P1; ++u) {
            const int t = t0 + u * bs;
            const float4 lb = C.lxyzq[j0 + (t < nT ? t : t0)];
            transform(M, lb.x, lb.y, lb.z, px[u], py[u], pz[u]);
            q[u] = lb.w;
        }
#pragma unroll
        for (int u = 0; u < P1; ++u) {
            const int n = list_node(C, px[u], py[u], pz[u]);
            const int m = n < 0 ? 0 : n;
            const unsigned b0 = C.noff[m], b1 = C.noff[m + 1];
            nb[u] = n < 0 ? 0u : b0;
            ne[u] = n < 0 ? 0u : b1;
        }
#pragma unroll
        for (int u = 0; u < P1; ++u) {
            const int t = t0 + u * bs;
            if (t >= nT) break;
            ab[t] = nb[u];
            ae[t] = ne[u];
            // no receptor atom in reach: the desolvation minimum stays at its initial value
            if (DES != DES_NONE && ne[u] == nb[u])
                s += to_fix(solv_term(MIN_INIT, C.lasa[j0 + t], C.ldes[j0 + t]), E_SCALE);
        }
    }
    __syncthreads();
    // stable compaction of the active atoms into GB buckets by the number of
    // GL*GK-entry steps their list takes (most steps first, 4 or more share
    // one), each bucket in (Morton) atom order: a group's atoms then take the
    // same number of steps and stay spatially close
    unsigned nact = 0;
    {
        __shared__ unsigned wb[GB][16], bbase[GB], brun[GB];
        constexpr unsigned SE = (unsigned)(GL * GK);
        if (tid < GB) bbase[tid] = 0u;
        __syncthreads();
        for (int t = tid; t < nT; t += bs) {
            const unsigned len = ae[t] - ab[t];
            if (len) atomicAdd(&bbase[GB - (int)min((len + SE - 1u) / SE, (unsigned)GB)], 1u);
        }
        __syncthreads();
        if (tid == 0) {
            unsigned acc0 = 0u;
            for (int q = 0; q < GB; ++q) {
                const unsigned c = bbase[q];
                bbase[q] = acc0;
                brun[q] = 0u;
                acc0 += c;
            }
        }
        __syncthreads();
        for (int c0 = 0; c0 < nT; c0 += bs) {
            const int t = c0 + tid;
            const unsigned len = t < nT ? ae[t] - ab[t] : 0u;
            const int bk = len ? GB - (int)min((len + SE - 1u) / SE, (unsigned)GB) : -1;
            unsigned long long mine = 0ull;
#pragma unroll
            for (int q = 0; q < GB; ++q) {
                const unsigned long long m = __ballot(bk == q);
                if (bk == q) mine = m;
                if (lane == 0) wb[q][w] = (unsigned)__popcll(m);
            }
            __syncthreads();
            if (bk >= 0) {
                unsigned off = bbase[bk] + brun[bk];
                for (int q = 0; q < w; ++q) off += wb[bk][q];
                act[off + (unsigned)__popcll(mine & ((1ull << lane) - 1ull))] = (unsigned short)t;
            }
            __syncthreads();
            if (tid < GB) {
                unsigned tot = 0u;
                for (int q = 0; q < nw; ++q) tot += wb[tid][q];
                brun[tid] += tot;
            }
            __syncthreads();
        }
        for (int q = 0; q < GB; ++q) nact += brun[q];
    }
    if constexpr (FAR) {
        const int pw = max(1, nw / 4);
        if (w >= nw - pw) {  // PHI_U atoms per lane in flight, transformed again from their records
            for (int t0 = lane + (w - (nw - pw)) * 64 * PHI_U; t0 < nT; t0 += PHI_U * 64 * pw) {
                PhiTap tp[PHI_U];
                float qq[PHI_U];
#pragma unroll
                for (int u = 0; u < PHI_U; ++u) {
                    const int t = t0 + 64 * u;
                    const float4 lb = C.lxyzq[j0 + (t < nT ? t : t0)];
                    double x, y, z;
                    transform(M, lb.x, lb.y, lb.z, x, y, z);
                    qq[u] = t < nT ? lb.w : 0.0f;
                    phi_fetch(C, x, y, z, tp[u]);
                }
#pragma unroll
                for (int u = 0; u < PHI_U; ++u)
                    if (qq[u] != 0.0f) e += to_fix((double)qq[u] * phi_lerp(tp[u]), E_SCALE);
            }
        }
    }
    const unsigned ngroups = (nact + (unsigned)GA - 1u) / (unsigned)GA;
    const int sub = lane / GL, r = lane % GL;
    for (;;) {
        unsigned g = 0;
        if (lane == 0) g = atomicAdd(&next_group, 1u);
        g = __shfl(g, 0, 64);
        if (g >= ngroups) break;
        const unsigned ai = g * (unsigned)GA + (unsigned)sub;
        const bool has = ai < nact;
        const int t = act[has ? ai : g * (unsigned)GA];
        const int j = j0 + t;
        const float4 lb = C.lxyzq[j];
        const float4 lp = C.lpar[j];
        double x, y, z;
        transform(M, lb.x, lb.y, lb.z, x, y, z);
        const bool lf = (DES != DES_NONE) ? (C.lflag[j] != 0) : false;
        const unsigned b = has ? ab[t] : 0u, en = has ? ae[t] : 0u;
        const float qjs = lb.w * E_SCALE_F, svjs = lp.x * V_SCALE_F, rj = lp.y;
        const bool hj = lp.z != 0.0f;
        const float xh = (float)x, yh = (float)y, zh = (float)z;
        const float xl = (float)(x - (double)xh), yl = (float)(y - (double)yh), zl = (float)(z - (double)zh);
        const float4* prow = TYPED == 2 ? ltab + (int)C.lty[j] * C.ntype : ltab;  // this atom's row of pairs
        double mloc = MIN_INIT;
        // a lane's GK indices of a step are one 16-byte load from the
        // lane-major block (the dummy block once its list has ended); the next
        // step's are requested before this step's pairs are formed, so a step
        // waits for its records only
        static_assert(GL * GK == LIST_PAD && GK == 8, "one lane-major block per step");
        auto step_ids = [&](unsigned p, int* out) {
            const unsigned at = (p < en ? p : C.ndum) + (unsigned)(r * GK);
            if constexpr (WIDE) {  // 32-bit lists: the lane's 8 entries are two 16-byte loads
                const uint4 a = *reinterpret_cast<const uint4*>(C.nidx32 + at);
                const uint4 c = *reinterpret_cast<const uint4*>(C.nidx32 + at + 4u);
                out[0] = (int)a.x; out[1] = (int)a.y; out[2] = (int)a.z; out[3] = (int)a.w;
                out[4] = (int)c.x; out[5] = (int)c.y; out[6] = (int)c.z; out[7] = (int)c.w;
                return;
            }
            const uint4 q = *reinterpret_cast<const uint4*>(C.nidx + at);
            out[0] = (int)(q.x & 0xffffu);
            out[1] = (int)(q.x >> 16);
            out[2] = (int)(q.y & 0xffffu);
            out[3] = (int)(q.y >> 16);
            out[4] = (int)(q.z & 0xffffu);
            out[5] = (int)(q.z >> 16);
            out[6] = (int)(q.w & 0xffffu);
            out[7] = (int)(q.w >> 16);
        };
        int id[GK];
        step_ids(b, id);
        const char* __restrict__ rbase = reinterpret_cast<const char*>(TYPED ? C.rxyzq : C.rrec);
        for (unsigned p = b; __any(p < en); p += (unsigned)(GK * GL)) {
            float4 rq[GK], rp[GK];
#pragma unroll
            for (int k = 0; k < GK; ++k) {
                // with TYPED only the coordinates are gathered, from the dense
                // (x, y, z, q) array, and the two LDS reads complete meanwhile.
                // Addresses are a uniform base plus a 32-bit byte offset (one
                // shift per record; both halves of a record share it).
                const unsigned off = (unsigned)id[k] << (TYPED ? 4 : 5);
                rq[k] = *reinterpret_cast<const float4*>(rbase + off);
                if constexpr (TYPED != 0) rp[k] = prow[reinterpret_cast<const unsigned char*>(ltyp)[id[k]]];
                else rp[k] = *reinterpret_cast<const float4*>(rbase + off + 16u);
            }
            int idn[GK];
            step_ids(p + (unsigned)(GK * GL), idn);
#pragma unroll
            for (int k = 0; k < GK; ++k)
                pair_term<ELEC, CLASH, DES, true, TYPED == 2>(C, rq[k], id[k], x, y, z, xh, yh, zh, xl, yl, zl, qjs,
                                                              svjs, rj, hj, lf, e, v, mloc, minR, rp[k]);
#pragma unroll
            for (int k = 0; k < GK; ++k) id[k] = idn[k];
        }
        if (DES != DES_NONE) {
#pragma unroll
            for (int o = GL / 2; o > 0; o >>= 1) mloc = fmin(mloc, __shfl_xor(mloc, o, 64));
            if (has && r == 0) s += to_fix(solv_term(mloc, C.lasa[j], C.ldes[j]), E_SCALE);
        }
    }
    __syncthreads();  // the LDS arrays may be reused by the caller
}

// Cross-tile exchange without cache maintenance.  Tiles of one pose meet only
// through device-scope atomics (read-modify-writes, and loads and stores of
// the atomic kind), which are coherent across the chip's L2 caches without
// the L2 write-back and invalidation that a device-scope fence costs.  What a
// fence would add here is ordering: vm_drain() waits until every global memory
// operation this wave issued is performed (s_waitcnt vmcnt(0); a workgroup
// barrier does not wait for them), so a tile's merges and partial sums are in
// place before the wave that counts the tile in issues that count.  The tile
// that sees the count complete reads them with device-scope atomic loads,
// issued after its count returned.
__device__ __forceinline__ void vm_drain() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// One workgroup per (pose, chunk of `chunk` ligand atoms) tile.  out[2k] = E_raw,
// out[2k+1] = V, outS[k] = S.  With several chunks per pose the tiles add their
// fixed-point partials into acc[3k..3k+2] (integer atomics: exact, order-free);
// the last tile of a pose (counter cnt[k]) writes the result and re-zeroes acc
// and cnt for the next launch.  Desolvation with global minima (DES_GLOBAL)
// needs the whole ligand in one tile.  With minima in LDS a pose may be split:
// each tile's ligand atoms take their own minima whole, each tile merges its
// receptor minima into the pose's row of gmin (global atomic minima, only the
// entries it lowered), and the pose's last tile sums the receptor terms from
// gmin and restores its initial keys (gmin holds them between launches).
// PERATOM also writes each ligand atom's (E, V) (sorted order) to pa_e/pa_v.
#ifndef LKL_SCORE_ATTR
#define LKL_SCORE_ATTR
#endif
// Launch bounds of 1024 threads also cap every variant at 128 VGPRs, so two
// 512-thread tiles always share a CU (the desolvation variants would otherwise
// take up to 148 and run one tile per CU).
template <bool LIST, bool FAR, bool ELEC, bool CLASH, int DES, bool PERATOM, int TYPED = 0>
__global__ void __launch_bounds__(1024) LKL_SCORE_ATTR
score_kernel(DevComplex C, const double* __restrict__ poses, int N, double* __restrict__ out,
             double* __restrict__ outS, unsigned* __restrict__ gmin, float* __restrict__ pa_e,
             float* __restrict__ pa_v, long long* __restrict__ acc, unsigned* __restrict__ cnt,
             int chunk, int nchunks) {
    extern __shared__ __attribute__((aligned(16))) unsigned lds_min[];
    __shared__ long long red[3][16];
    __shared__ unsigned last_tile;
    const bool split = DES == DES_LDS && nchunks > 1;
    const unsigned total = (unsigned)N * (unsigned)nchunks;
    if (blockIdx.x >= total) return;
    const unsigned L = xcd_tile(blockIdx.x, total);
    const int pose = C.porder ? C.porder[L] : (int)(L / (unsigned)nchunks);
    const int j0 = (int)(L % (unsigned)nchunks) * chunk;
    const int j1 = min(C.nl, j0 + chunk);
    const int tid = threadIdx.x, bs = blockDim.x;
    const Mat M = pose_matrix(poses + (size_t)pose * 7);
    if (C.lstride) C.lxyzq += (size_t)pose * (size_t)C.lstride;

    unsigned* minR = nullptr;
    if (DES == DES_LDS) minR = lds_min;
    if (DES == DES_GLOBAL) minR = gmin + (size_t)pose * (size_t)C.nr;
    if (DES != DES_NONE) {
        const unsigned init = __float_as_uint((float)MIN_INIT);
        for (int i = tid; i < C.nr; i += bs) minR[i] = init;
        if (DES == DES_GLOBAL) vm_drain();  // the stores land before any wave's atomics
        __syncthreads();
    }

    long long e = 0, v = 0, s = 0;
    if (!PERATOM && C.wgrid) {
        // lists after the receptor minima when those live in LDS
        void* lists = DES == DES_LDS ? (void*)(((uintptr_t)(lds_min + C.nr) + 7u) & ~(uintptr_t)7u)
                                     : (void*)lds_min;
        // every launch over lists takes the grouped walk, desolvation included
        // (its receptor minima are LDS atomics, its ligand minima a shuffle
        // reduction over the atom's lanes)
        // (the index width is chosen per launch, outside the walk's loop)
        if constexpr (LIST) {
            if (C.nidx32) near_groups<FAR, ELEC, CLASH, DES, TYPED, true>(C, M, j0, j1, lists, minR, e, v, s);
            else near_groups<FAR, ELEC, CLASH, DES, TYPED, false>(C, M, j0, j1, lists, minR, e, v, s);
        } else {
            near_sorted<LIST, FAR, ELEC, CLASH, DES>(C, M, j0, j1, lists, minR, e, v, s);
        }
    } else if constexpr (DES == DES_NONE && !PERATOM) {
        {
            for (int j = j0 + tid; j < j1; j += bs) {
                const float4 lb = C.lxyzq[j];
                const float4 lp = C.lpar[j];
                double x, y, z;
                transform(M, lb.x, lb.y, lb.z, x, y, z);
                if (FAR && lb.w != 0.0f) e += to_fix((double)lb.w * sample_phi(C, x, y, z), E_SCALE);
                double minL = MIN_INIT;
                near_any<LIST, ELEC, CLASH, DES_NONE>(C, x, y, z, lb.w, lp.x, lp.y, lp.z != 0.0f, false, e, v,
                                                 minL, nullptr);
            }
        }
    } else
    for (int j = j0 + tid; j < j1; j += bs) {
        const float4 lb = C.lxyzq[j];
        const float4 lp = C.lpar[j];
        double x, y, z;
        transform(M, lb.x, lb.y, lb.z, x, y, z);
        long long ea = 0, va = 0;
        double minL = MIN_INIT;
        if (FAR && lb.w != 0.0f) ea += to_fix((double)lb.w * sample_phi(C, x, y, z), E_SCALE);
        const bool lf = (DES != DES_NONE) ? (C.lflag[j] != 0) : false;
        near_any<LIST, ELEC, CLASH, DES>(C, x, y, z, lb.w, lp.x, lp.y, lp.z != 0.0f, lf, ea, va, minL,
                                    minR);
        if (PERATOM) {
            pa_e[(size_t)pose * C.nl + j] = (float)((double)ea / E_SCALE);
            pa_v[(size_t)pose * C.nl + j] = (float)((double)va / V_SCALE);
        }
        e += ea;
        v += va;
        if (DES != DES_NONE) s += to_fix(solv_term(minL, C.lasa[j], C.ldes[j]), E_SCALE);
    }
    if (split) {
        __syncthreads();
        unsigned* __restrict__ g = gmin + (size_t)pose * (size_t)C.nr;
        const unsigned init = __float_as_uint((float)MIN_INIT);
        for (int p = tid; p < C.nr; p += bs) {
            const unsigned k = minR[p];
            if (k < init) atomicMin(&g[p], k);
        }
        vm_drain();  // this wave's merges are performed before the tile is counted
    } else if (DES != DES_NONE) {
        if (DES == DES_GLOBAL) vm_drain();  // the atomics land before any wave reads the minima
        __syncthreads();
        for (int p = tid; p < C.nr; p += bs) {
            unsigned k;
            if (DES == DES_GLOBAL)
                k = __hip_atomic_load(&minR[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            else
                k = minR[p];
            s += to_fix(solv_term((double)__uint_as_float(k), C.rasa[p], C.rdes[p]), E_SCALE);
        }
    }
    e = wave_sum(e);
    v = wave_sum(v);
    if (DES != DES_NONE) s = wave_sum(s);
    const int w = tid >> 6, nw = (bs + 63) >> 6;
    if ((tid & 63) == 0) {
        red[0][w] = e;
        red[1][w] = v;
        red[2][w] = s;
    }
    __syncthreads();
    if (split) {
        if (tid == 0) {
            long long te = 0, tv = 0, ts = 0;
            for (int k = 0; k < nw; ++k) {
                te += red[0][k];
                tv += red[1][k];
                ts += red[2][k];
            }
            unsigned long long* a = (unsigned long long*)acc + 3 * (size_t)pose;
            atomicAdd(&a[0], (unsigned long long)te);
            atomicAdd(&a[1], (unsigned long long)tv);
            atomicAdd(&a[2], (unsigned long long)ts);
            vm_drain();
            last_tile = atomicAdd(&cnt[pose], 1u) == (unsigned)nchunks - 1u ? 1u : 0u;
        }
        __syncthreads();
        if (!last_tile) return;
        unsigned* __restrict__ g = gmin + (size_t)pose * (size_t)C.nr;
        const unsigned init = __float_as_uint((float)MIN_INIT);
        long long sr = 0;
        for (int p = tid; p < C.nr; p += bs) {
            const unsigned k = __hip_atomic_load(&g[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (k != init) __hip_atomic_store(&g[p], init, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            sr += to_fix(solv_term((double)__uint_as_float(k), C.rasa[p], C.rdes[p]), E_SCALE);
        }
        sr = wave_sum(sr);
        if ((tid & 63) == 0) red[2][w] = sr;
        __syncthreads();
        if (tid != 0) return;
        long long tsr = 0;
        for (int k = 0; k < nw; ++k) tsr += red[2][k];
        unsigned long long* a = (unsigned long long*)acc + 3 * (size_t)pose;
        const long long te = (long long)atomicExch(&a[0], 0ull);
        const long long tv = (long long)atomicExch(&a[1], 0ull);
        const long long ts = (long long)atomicExch(&a[2], 0ull) + tsr;
        atomicExch(&cnt[pose], 0u);
        out[2 * (size_t)pose] = (double)te / E_SCALE;
        out[2 * (size_t)pose + 1] = (double)tv / V_SCALE;
        outS[pose] = (double)ts / E_SCALE;
        return;
    }
    if (tid != 0) return;
    long long te = 0, tv = 0, ts = 0;
    for (int k = 0; k < nw; ++k) {
        te += red[0][k];
        tv += red[1][k];
        ts += red[2][k];
    }
    if (nchunks > 1) {
        unsigned long long* a = (unsigned long long*)acc + 3 * (size_t)pose;
        atomicAdd(&a[0], (unsigned long long)te);
        atomicAdd(&a[1], (unsigned long long)tv);
        vm_drain();
        if (atomicAdd(&cnt[pose], 1u) != (unsigned)nchunks - 1u) return;
        te = (long long)atomicExch(&a[0], 0ull);
        tv = (long long)atomicExch(&a[1], 0ull);
        atomicExch(&cnt[pose], 0u);
    }
    out[2 * (size_t)pose] = (double)te / E_SCALE;
    out[2 * (size_t)pose + 1] = (double)tv / V_SCALE;
    if (DES != DES_NONE) outS[pose] = (double)ts / E_SCALE;
}

// ── multi-problem batch ───────────────────────────────────────────────────────
// Fixed-point (E, V) sums of ligand atoms [j0, j1) of one pose over this
// thread's share (no desolvation, no per-atom output).  A problem whose view
// carries ligand and receptor parameter types walks its lists with the
// pair-type table, one with receptor types only with the receptor's table
// (the host clears the types that do not fit the launch's LDS).
template <bool LIST, bool FAR, bool ELEC, bool CLASH>
__device__ __forceinline__ void tile_sums(const DevComplex& C, const Mat& M, int j0, int j1, void* lds,
                                          long long& e, long long& v) {
    if (C.wgrid) {
        long long s_unused = 0;
        if constexpr (LIST) {
            if (C.nidx32) {
                if (C.ntype && C.nltype) near_groups<FAR, ELEC, CLASH, DES_NONE, 2, true>(C, M, j0, j1, lds, nullptr, e, v, s_unused);
                else if (C.ntype) near_groups<FAR, ELEC, CLASH, DES_NONE, 1, true>(C, M, j0, j1, lds, nullptr, e, v, s_unused);
                else near_groups<FAR, ELEC, CLASH, DES_NONE, 0, true>(C, M, j0, j1, lds, nullptr, e, v, s_unused);
            } else {
                if (C.ntype && C.nltype) near_groups<FAR, ELEC, CLASH, DES_NONE, 2>(C, M, j0, j1, lds, nullptr, e, v, s_unused);
                else if (C.ntype) near_groups<FAR, ELEC, CLASH, DES_NONE, 1>(C, M, j0, j1, lds, nullptr, e, v, s_unused);
                else near_groups<FAR, ELEC, CLASH, DES_NONE, 0>(C, M, j0, j1, lds, nullptr, e, v, s_unused);
            }
        } else
            near_sorted<false, FAR, ELEC, CLASH, DES_NONE>(C, M, j0, j1, lds, nullptr, e, v, s_unused);
        return;
    }
    for (int j = j0 + (int)threadIdx.x; j < j1; j += (int)blockDim.x) {
        const float4 lb = C.lxyzq[j];
        const float4 lp = C.lpar[j];
        double x, y, z;
        transform(M, lb.x, lb.y, lb.z, x, y, z);
        double minL = MIN_INIT;
        if (FAR && lb.w != 0.0f) e += to_fix((double)lb.w * sample_phi(C, x, y, z), E_SCALE);
        near_any<LIST, ELEC, CLASH, DES_NONE>(C, x, y, z, lb.w, lp.x, lp.y, lp.z != 0.0f, false, e, v,
                                         minL, nullptr);
    }
}

struct ProblemDesc {
    int tile_begin;  // first batch tile of this problem (problems sorted by cost)
    int nchunks;     // ligand chunks per pose
    int chunk;       // ligand atoms per chunk
    int pose_off;    // first pose row of this problem in the batch arrays
    unsigned flags;
    int pad[3];
};

// One launch scores every (problem, pose, ligand chunk) tile of a batch of
// complexes: one workgroup per tile, tiles numbered largest problem first so
// the hardware dispatcher starts the expensive work early and back-fills the
// small problems as workgroup slots free up.  Per-pose partials meet in
// acc/cnt exactly as in score_kernel, so each problem's results are bitwise
// those of its own launch.  (No loop around barriers: a claim-loop variant
// with a divergent tid==0 claim was restructured by the compiler into nested
// loops that cycle the barriers without the claiming lane, which never ends.)
// With trace set, each tile records {start, end} on the device's constant
// 100 MHz wall clock, the CU it ran on (__smid: XCC, shader engine, CU) and
// its problem, four 64-bit words per tile, for load-balance measurement.
template <bool LIST>
__global__ void __launch_bounds__(1024)
multi_kernel(const DevComplex* __restrict__ cxs, const ProblemDesc* __restrict__ probs, int K,
             int total_tiles, const double* __restrict__ poses, double* __restrict__ out,
             long long* __restrict__ acc, unsigned* __restrict__ cnt, unsigned long long* __restrict__ trace) {
    __shared__ long long red[2][16];
    extern __shared__ __attribute__((aligned(16))) unsigned char multi_lds[];  // work-ordered lists
    const int t = (int)blockIdx.x;
    if (t >= total_tiles) return;
    const int tid = threadIdx.x, bs = blockDim.x;
    int lo = 0, hi = K - 1;  // last problem with tile_begin <= t
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (probs[mid].tile_begin <= t) lo = mid;
        else hi = mid - 1;
    }
    const ProblemDesc P = probs[lo];
    const DevComplex C = cxs[lo];
    if (trace && tid == 0) {
        trace[4 * (size_t)t] = (unsigned long long)wall_clock64();
        trace[4 * (size_t)t + 2] = __smid();
        trace[4 * (size_t)t + 3] = (unsigned long long)lo;
    }
    const int local = t - P.tile_begin;
    const int pose = P.pose_off + local / P.nchunks;
    const int j0 = (local % P.nchunks) * P.chunk;
    const int j1 = min(C.nl, j0 + P.chunk);
    const Mat M = pose_matrix(poses + (size_t)pose * 7);
    long long e = 0, v = 0;
    switch (P.flags & 7u) {
        case 0: tile_sums<LIST, false, false, false>(C, M, j0, j1, multi_lds, e, v); break;
        case 1: tile_sums<LIST, true, false, false>(C, M, j0, j1, multi_lds, e, v); break;
        case 2: tile_sums<LIST, false, true, false>(C, M, j0, j1, multi_lds, e, v); break;
        case 3: tile_sums<LIST, true, true, false>(C, M, j0, j1, multi_lds, e, v); break;
        case 4: tile_sums<LIST, false, false, true>(C, M, j0, j1, multi_lds, e, v); break;
        case 5: tile_sums<LIST, true, false, true>(C, M, j0, j1, multi_lds, e, v); break;
        case 6: tile_sums<LIST, false, true, true>(C, M, j0, j1, multi_lds, e, v); break;
        default: tile_sums<LIST, true, true, true>(C, M, j0, j1, multi_lds, e, v); break;
    }
    e = wave_sum(e);
    v = wave_sum(v);
    const int w = tid >> 6, nw = (bs + 63) >> 6;
    if ((tid & 63) == 0) { red[0][w] = e; red[1][w] = v; }
    __syncthreads();
    if (tid != 0) return;
    if (trace) trace[4 * (size_t)t + 1] = (unsigned long long)wall_clock64();
    long long te = 0, tv = 0;
    for (int k = 0; k < nw; ++k) { te += red[0][k]; tv += red[1][k]; }
    if (P.nchunks > 1) {
        unsigned long long* a = (unsigned long long*)acc + 3 * (size_t)pose;
        atomicAdd(&a[0], (unsigned long long)te);
        atomicAdd(&a[1], (unsigned long long)tv);
        vm_drain();  // as in score_kernel: partials performed before the count
        if (atomicAdd(&cnt[pose], 1u) != (unsigned)P.nchunks - 1u) return;
        te = (long long)atomicExch(&a[0], 0ull);
        tv = (long long)atomicExch(&a[1], 0ull);
        atomicExch(&cnt[pose], 0u);
    }
    out[2 * (size_t)pose] = (double)te / E_SCALE;
    out[2 * (size_t)pose + 1] = (double)tv / V_SCALE;
}

// ── interface flags (restraints / membrane) ───────────────────────────────────
// dna.rs:885-995 / cpydock.rs:509,639: a receptor and a ligand atom are at the
// interface when their f64 squared distance is <= INTERFACE_CUTOFF^2
// (constants.rs:14-15); satisfied_restraints / membrane_intersection
// (scoring.rs:83-110) then count residues with an interface atom and membrane
// beads at the interface.  One workgroup per pose: the ligand atoms walk their
// nodes' short lists (a superset of the 3.9 A pairs), pairs within 1e-3 A^2 of
// the cutoff are re-measured in f64 as the CPU measures them, and the flags
// are bits in LDS.  out[3k..3k+2] = satisfied receptor residues, satisfied
// ligand residues, membrane beads at the interface (integers: exact).
constexpr double IFACE2 = 3.9 * 3.9;
constexpr float IFACE_HI = 15.211f, IFACE_LO = 15.209f;

struct IfaceDev {
    const unsigned* noff;          // short lists over the main lists' nodes
    const unsigned short* nidx;
    const unsigned* nidx32;        // the same with 32-bit indices (NULL unless nidx is)
    const int* r_off;
    const int* r_atoms;  // sorted receptor indices
    const int* l_off;
    const int* l_atoms;  // sorted ligand indices
    const int* mem;      // sorted receptor indices
    int nres_r, nres_l, nmem;
};

__global__ void __launch_bounds__(512)
iface_kernel(DevComplex C, IfaceDev I, const double* __restrict__ poses, int N, int* __restrict__ out) {
    extern __shared__ __attribute__((aligned(16))) unsigned ibits[];
    __shared__ int icnt[3];
    const int pose = (int)blockIdx.x;
    if (pose >= N) return;
    const int tid = (int)threadIdx.x, bs = (int)blockDim.x;
    const int wr = (C.nr + 31) / 32, wl = (C.nl + 31) / 32;
    unsigned* rb = ibits;
    unsigned* lbits = ibits + wr;
    for (int k = tid; k < wr + wl; k += bs) ibits[k] = 0u;
    if (tid < 3) icnt[tid] = 0;
    __syncthreads();
    const Mat M = pose_matrix(poses + (size_t)pose * 7);
    for (int t = tid; t < C.nl; t += bs) {
        const float4 lb = C.lxyzq[t];
        double x, y, z;
        transform(M, lb.x, lb.y, lb.z, x, y, z);
        const int n = list_node(C, x, y, z);
        if (n < 0) continue;
        const unsigned b = I.noff[n], en = I.noff[n + 1];
        const float xh = (float)x, yh = (float)y, zh = (float)z;
        const float xl = (float)(x - (double)xh), yl = (float)(y - (double)yh), zl = (float)(z - (double)zh);
        bool hit = false;
        for (unsigned p = b; p < en; ++p) {
            const int id = I.nidx32 ? (int)I.nidx32[p] : (int)I.nidx[p];
            const float4 r = C.rrec[2 * id];
            const float dx = (xh - r.x) + xl, dy = (yh - r.y) + yl, dz = (zh - r.z) + zl;
            const float d2 = dx * dx + dy * dy + dz * dz;
            if (d2 <= IFACE_HI && !(d2 > IFACE_LO && exact_d2(x, y, z, r) > IFACE2)) {
                atomicOr(&rb[id >> 5], 1u << (id & 31));
                hit = true;
            }
        }
        if (hit) atomicOr(&lbits[t >> 5], 1u << (t & 31));
    }
    __syncthreads();
    for (int k = tid; k < I.nres_r; k += bs) {
        for (int q = I.r_off[k]; q < I.r_off[k + 1]; ++q) {
            const int a = I.r_atoms[q];
            if ((rb[a >> 5] >> (a & 31)) & 1u) {
                atomicAdd(&icnt[0], 1);
                break;
            }
        }
    }
    for (int k = tid; k < I.nres_l; k += bs) {
        for (int q = I.l_off[k]; q < I.l_off[k + 1]; ++q) {
            const int a = I.l_atoms[q];
            if ((lbits[a >> 5] >> (a & 31)) & 1u) {
                atomicAdd(&icnt[1], 1);
                break;
            }
        }
    }
    for (int k = tid; k < I.nmem; k += bs) {
        const int a = I.mem[k];
        if ((rb[a >> 5] >> (a & 31)) & 1u) atomicAdd(&icnt[2], 1);
    }
    __syncthreads();
    if (tid < 3) out[3 * (size_t)pose + tid] = icnt[tid];
}

// Longest-first dispatch order, one launch.  xcd_tile gives XCD x the
// contiguous range of tiles it dispatches in ascending order; within each
// range the order sorts the poses by expected near-field cost, descending
// (ties by pose), so the costly poses start first and the cheap ones fill the
// last round.  The cost of a pose is the sum of the neighbour-list lengths at
// the nodes of every eighth ligand atom (the walk's work is proportional to
// the candidates it visits): one wavefront per pose, each lane placing U atoms
// and requesting their offsets together.  Block b serves range b % 8 (ORDER_W
// poses per block and round); the last block of a range to finish (counter
// cnt[x]) reads the range's costs and writes its order by a rank sort in LDS,
// then re-zeroes the counter for the next launch.  Blocks meet only through
// device-scope atomics (see vm_drain), so the placement of blocks on XCDs
// does not matter.  The wavefronts also keep a device copy of each pose
// (pose_copy, when given): the poses may be read from host memory.  Ranges
// of up to LPT_MAX poses (launches of up to 8 LPT_MAX).
constexpr int COST_STRIDE = 8;
constexpr int LPT_MAX = POSE_ORDER_MAX / 8;
constexpr int ORDER_W = 16;  // wavefronts (poses) per block: 1,024 threads sort a range
__global__ void __launch_bounds__(64 * ORDER_W) pose_order_kernel(DevComplex C, const double* __restrict__ poses,
                                                                  int N, unsigned* __restrict__ cost,
                                                                  int* __restrict__ order, unsigned* __restrict__ cnt,
                                                                  double* __restrict__ pose_copy) {
    __shared__ unsigned long long key[LPT_MAX];
    __shared__ unsigned last;
    const unsigned q = (unsigned)N / 8u, r = (unsigned)N % 8u, x = blockIdx.x % 8u;
    const unsigned start = x < r ? x * (q + 1u) : r * (q + 1u) + (x - r) * q;
    const unsigned m = q + (x < r ? 1u : 0u);
    const unsigned nblk = gridDim.x / 8u;  // blocks per range
    const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
    for (unsigned i = (blockIdx.x / 8u) * ORDER_W + (unsigned)w; i < m; i += nblk * ORDER_W) {
        const unsigned pose = start + i;
        const Mat M = pose_matrix(poses + (size_t)pose * 7);
        if (pose_copy && lane < 7) pose_copy[(size_t)pose * 7 + lane] = poses[(size_t)pose * 7 + lane];
        const float4* lx = C.lxyzq + (C.lstride ? (size_t)pose * (size_t)C.lstride : 0);
        constexpr int U = 4;
        unsigned c = 0;
        for (int j0 = lane * COST_STRIDE; j0 < C.nl; j0 += U * 64 * COST_STRIDE) {
            int n[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int j = j0 + u * 64 * COST_STRIDE;
                n[u] = -1;
                if (j < C.nl) {
                    const float4 lb = lx[j];
                    double px, py, pz;
                    transform(M, lb.x, lb.y, lb.z, px, py, pz);
                    n[u] = list_node(C, px, py, pz);
                }
            }
            unsigned lo[U], hi[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int qn = n[u] >= 0 ? n[u] : 0;
                lo[u] = C.noff[qn];
                hi[u] = C.noff[qn + 1];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) c += n[u] >= 0 ? hi[u] - lo[u] : 0u;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
        if (lane == 0) __hip_atomic_store(&cost[pose], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    vm_drain();  // this wave's cost stores are performed before the block counts itself
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(&cnt[x], 1u) == nblk - 1u ? 1u : 0u;
    __syncthreads();
    if (!last) return;
    for (unsigned t = threadIdx.x; t < m; t += blockDim.x) {
        // descending cost first, then ascending pose: key = ~cost : pose (unique)
        const unsigned ct = __hip_atomic_load(&cost[start + t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        key[t] = ((unsigned long long)(~ct) << 32) | (start + t);
    }
    if (threadIdx.x == 0) __hip_atomic_store(&cnt[x], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    // rank sort: a key's position is the number of smaller keys (every lane
    // reads the same key at once, an LDS broadcast)
    for (unsigned t = threadIdx.x; t < m; t += blockDim.x) {
        const unsigned long long k = key[t];
        unsigned rank = 0;
        for (unsigned i = 0; i < m; ++i) rank += key[i] < k ? 1u : 0u;
        order[start + rank] = (int)(k & 0xffffffffu);
    }
}

// Diagnostic: device transform of every (pose, ligand atom), sorted order.
__global__ void transform_kernel(DevComplex C, const double* __restrict__ poses, int N,
                                 double* __restrict__ outxyz) {
    const int pose = blockIdx.x;
    if (pose >= N) return;
    const Mat M = pose_matrix(poses + (size_t)pose * 7);
    for (int j = threadIdx.x; j < C.nl; j += blockDim.x) {
        const float4 lb = C.lxyzq[j];
        double x, y, z;
        transform(M, lb.x, lb.y, lb.z, x, y, z);
        double* o = outxyz + ((size_t)pose * C.nl + j) * 3;
        o[0] = x;
        o[1] = y;
        o[2] = z;
    }
}

// ── context ───────────────────────────────────────────────────────────────────
struct lkl_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t lds_max = 65536;
    int block = 256;
    int chunk = 0;  // ligand atoms per workgroup tile (0 = the whole ligand)
    float cell_size = 5.0f;  // cell edge along y and z (A)
    float cell_x = 0.0f;     // cell edge along x (A); 0 = cell_size
    bool work_sort = true;   // order each tile's ligand atoms by near-field work
    unsigned short* d_wgrid = nullptr;
    int wn[3] = {0, 0, 0};
    double wo[3] = {0, 0, 0};
    float list_h = 1.0f;  // neighbour-list node spacing (A); 0 = cell-row walk
    unsigned* d_noff = nullptr;
    unsigned short* d_nidx = nullptr;
    unsigned* d_nidx32 = nullptr;  // 32-bit indices instead (receptors above 65,535 atoms)
    int nn[3] = {0, 0, 0};
    float no[3] = {0.f, 0.f, 0.f};
    size_t list_entries = 0;  // 0 = no lists (row walk)
    unsigned list_dummy = 0;  // position of the block of LIST_PAD dummy entries in d_nidx
    // host copies in reference order (kept to rebuild the cell layout)
    int nr = 0, nl = 0;
    std::vector<float> r_xyz, r_ele, r_svdw, r_vdwr, r_des, r_asa;
    std::vector<unsigned char> r_heavy, r_flag;
    bool has_phi = false, has_desolv = false;
    // device
    float* d_phi = nullptr;
    float4* d_phi8 = nullptr;  // corner-packed copy of phi (NULL = off / not built)
    bool phi_packed = true;
    int nx = 0, ny = 0, nz = 0;
    double ox = 0, oy = 0, oz = 0, sp = 1;
    float4 *d_rxyzq = nullptr, *d_rpar = nullptr, *d_rrec = nullptr;
    // receptor parameter types (ntype 0: more than RTYPE_MAX distinct records)
    unsigned* d_rtyp = nullptr;
    float4* d_rtab = nullptr;
    int ntype = 0;
    // ligand parameter types (nltype 0: more than RTYPE_MAX distinct records)
    unsigned* d_lty = nullptr;
    float4* d_lttab = nullptr;
    int nltype = 0;
    // LDS type tables of the list walk: 0 none, 1 the receptor's, 2 (when the
    // ligand has types and the table fits) the pair types
    int param_types = 2;
    int* d_cstart = nullptr;
    unsigned char* d_rflag = nullptr;
    float *d_rdes = nullptr, *d_rasa = nullptr;
    float4 *d_lxyzq = nullptr, *d_lpar = nullptr;
    unsigned char* d_lflag = nullptr;
    float *d_ldes = nullptr, *d_lasa = nullptr;
    std::vector<int> lperm;  // sorted position -> reference index
    std::vector<int> rperm;  // receptor: sorted position -> reference index
    // interface flags (restraints / membrane), reference atom order, CSR
    bool has_iface = false;
    std::vector<int> if_roff, if_ratoms, if_loff, if_latoms, if_mem;
    int* d_iface = nullptr;      // translated to sorted indices: r_off, r_atoms, l_off, l_atoms, mem
    size_t cap_iface = 0;
    unsigned iface_version = ~0u;  // layout version the device copy was translated for
    int* d_icounts = nullptr;
    size_t cap_icounts = 0;
    // short neighbour lists for the interface pass (same nodes as the main
    // lists, radius 3.9 A + half-diagonal + margin), built on first use
    unsigned* d_noff_if = nullptr;
    unsigned short* d_nidx_if = nullptr;
    unsigned* d_nidx32_if = nullptr;
    unsigned iflist_version = ~0u;
    int ncx = 1, ncy = 1, ncz = 1;
    double cox = 0, coy = 0, coz = 0;
    // per-call buffers (grown on demand, never shrunk)
    double* d_poses = nullptr;
    size_t cap_poses = 0;
    // per-pose ligand records of lkl_ctx_score_coords (N x nl, sorted order),
    // and the ligand charges in sorted order that fill their w
    float4* d_lcoords = nullptr;
    size_t cap_lcoords = 0;
    // per-pose cost and dispatch order of one-tile-per-pose launches
    unsigned* d_pcost = nullptr;
    int* d_porder = nullptr;
    unsigned* d_ocnt = nullptr;  // the dispatch-order counters (8, zero between launches)
    size_t cap_pcost = 0, cap_porder = 0;
    bool pose_order = true;
    std::vector<float> lq;
    double* d_out = nullptr;
    size_t cap_out = 0;
    double* d_outS = nullptr;
    size_t cap_outS = 0;
    unsigned* d_gmin = nullptr;
    size_t cap_gmin = 0;
    unsigned* d_smin = nullptr;  // receptor minima of split desolvation poses (held at the initial key)
    size_t cap_smin = 0;
    float *d_pa_e = nullptr, *d_pa_v = nullptr;
    size_t cap_pa = 0, cap_pa_v = 0;
    double* d_xyz = nullptr;
    size_t cap_xyz = 0;
    long long* d_acc = nullptr;  // per-pose tile accumulators (kept zero between launches)
    unsigned* d_cnt = nullptr;
    size_t cap_acc = 0;
    // pinned, device-mapped staging of the blocking call (poses in, scores
    // out), read and written by the kernels over the bus: no DMA copies and a
    // single stream synchronisation per call
    double* h_stage = nullptr;
    double* d_stage = nullptr;  // its device address
    size_t cap_stage = 0;
    // optional kernel timing of every launch: events before the dispatch-order
    // pre-pass, before the scoring kernel and after it, on the launch's stream
    bool timing = false;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    bool ev_prepass = false;  // the last timed launch ran the pre-pass
    unsigned version = 0;  // bumped whenever a device array is replaced
};

int lkl_field_build_device(const float* r_xyz, const float* r_ele, int nr, float* d_phi,
                           hipStream_t st);
extern "C" long long lkl_field_dims(const float* r_xyz, int nr, double* origin3, int* dims3);

// Host-to-device copy ordered on `st` and waited for.  A plain hipMemcpy from
// pageable memory may return before the data lands, and the kernels run on
// non-blocking streams that do not wait for the null stream.
static hipError_t h2d(void* dst, const void* src, size_t bytes, hipStream_t st) {
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
    return e != hipSuccess ? e : hipStreamSynchronize(st);
}

template <typename T>
static int dev_upload(T** dst, const T* src, size_t n, hipStream_t st) {
    if (*dst) { (void)hipFree(*dst); *dst = nullptr; }
    if (n == 0) return 0;
    HIPCK("alloc", hipMalloc((void**)dst, n * sizeof(T)));
    HIPCK("upload", h2d(*dst, src, n * sizeof(T), st));
    return 0;
}

template <typename T>
static int dev_reserve(T** buf, size_t* cap, size_t n) {
    if (n <= *cap && *buf) return 0;
    if (*buf) { (void)hipFree(*buf); *buf = nullptr; *cap = 0; }
    HIPCK("alloc", hipMalloc((void**)buf, std::max<size_t>(n, 1) * sizeof(T)));
    *cap = n;
    return 0;
}

// Receptor layout: uniform cells of edge cell_size over the receptor's box,
// atoms stably counting-sorted by cell (x fastest), packed float4 records.
static float cell_x_of(const lkl_ctx* c) { return c->cell_x > 0.0f ? c->cell_x : c->cell_size; }

// Corner-packed far field: for every cell of phi, its 8 corner values in the
// order sample_phi reads them, so one 32-byte read serves a sample.  Values
// are copied, not recomputed: samples are bitwise those of phi.
__global__ void phi_pack_kernel(const float* __restrict__ phi, int nx, int ny, int nz, float4* __restrict__ out) {
    const long long mx = nx - 1, my = ny - 1, mz = nz - 1;
    const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (c >= mx * my * mz) return;
    const long long ix = c % mx, t = c / mx, iy = t % my, iz = t / my;
    const size_t sy = (size_t)nx, sz = (size_t)nx * (size_t)ny;
    const size_t b = (size_t)iz * sz + (size_t)iy * sy + (size_t)ix;
    out[2 * c] = make_float4(phi[b], phi[b + 1], phi[b + sy], phi[b + sy + 1]);
    out[2 * c + 1] = make_float4(phi[b + sz], phi[b + sz + 1], phi[b + sz + sy], phi[b + sz + sy + 1]);
}

// (Re)build the packed copy of the context's phi when packing is on; an
// allocation failure leaves it off (sample_phi then reads phi itself).
static int build_phi8(lkl_ctx* c) {
    if (c->d_phi8) { (void)hipFree(c->d_phi8); c->d_phi8 = nullptr; }
    if (!c->phi_packed || !c->has_phi || !c->d_phi) return 0;
    const long long cells = (long long)(c->nx - 1) * (c->ny - 1) * (c->nz - 1);
    if (cells <= 0) return 0;
    if (hipMalloc((void**)&c->d_phi8, sizeof(float4) * 2 * (size_t)cells) != hipSuccess) {
        (void)hipGetLastError();
        c->d_phi8 = nullptr;
        return 0;
    }
    hipLaunchKernelGGL(phi_pack_kernel, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, c->stream, c->d_phi,
                       c->nx, c->ny, c->nz, c->d_phi8);
    HIPCK("launch", hipGetLastError());
    HIPCK("sync", hipStreamSynchronize(c->stream));
    c->version++;
    return 0;
}

// Neighbour lists over the sorted receptor (indices into d_rxyzq).  Lists are
// skipped (row walk) when switched off, when an index would not fit 16 bits, or
// when the node grid or the entry count exceeds the caps below.
constexpr long long LIST_NODES_MAX = 1ll << 27;
constexpr long long LIST_ENTRIES_MAX = 1ll << 30;  // 2 GiB of 16-bit indices (4 GiB of 32-bit)

// Count and fill one list set over the node grid (o, n, h) with radius rad,
// each list padded to a multiple of pad (blocks lane-major if lanemajor);
// outputs a fresh offsets array (nodes + 1) and index array.  Returns the
// entry count, 0 when it exceeds LIST_ENTRIES_MAX (nothing allocated), or -1.
static long long fill_lists(lkl_ctx* c, const float o[3], const int n[3], float h, float rad, unsigned pad,
                            bool lanemajor, unsigned** d_off, unsigned short** d_idx, unsigned** d_idx32,
                            unsigned* dummy_pos = nullptr) {
    // 16-bit indices while the dummy index nr fits them, else 32-bit
    const bool wide = c->nr > 65535;
    const long long nodes = (long long)n[0] * n[1] * n[2];
    HIPCK("alloc", hipMalloc((void**)d_off, sizeof(unsigned) * (size_t)(nodes + 1)));
    dim3 grid((n[0] + NLX - 1) / NLX, (n[1] + NLY - 1) / NLY, (n[2] + NLZ - 1) / NLZ);
    hipLaunchKernelGGL(nlist_kernel<false>, grid, dim3(NLT), 0, c->stream, c->d_rxyzq, c->nr, o[0], o[1], o[2], h,
                       n[0], n[1], n[2], rad, pad, lanemajor ? 1 : 0, *d_off, (unsigned short*)nullptr,
                       (unsigned*)nullptr);
    HIPCK("launch", hipGetLastError());
    HIPCK("sync", hipStreamSynchronize(c->stream));
    std::vector<unsigned> off((size_t)nodes + 1);
    HIPCK("readback", hipMemcpy(off.data(), *d_off, sizeof(unsigned) * (size_t)nodes, hipMemcpyDeviceToHost));
    unsigned long long acc = 0;
    for (long long k = 0; k < nodes; ++k) {
        const unsigned cnt = off[k];
        off[k] = (unsigned)acc;
        acc += cnt;
        if (acc > (unsigned long long)LIST_ENTRIES_MAX) break;
    }
    if (acc > (unsigned long long)LIST_ENTRIES_MAX) {
        (void)hipFree(*d_off);
        *d_off = nullptr;
        return 0;
    }
    off[nodes] = (unsigned)acc;
    HIPCK("upload", h2d(*d_off, off.data(), sizeof(unsigned) * off.size(), c->stream));
    // a block of pad dummy entries (index nr) after the last list
    if (wide) {
        HIPCK("alloc", hipMalloc((void**)d_idx32, sizeof(unsigned) * (size_t)(acc + pad)));
        const std::vector<unsigned> dum(pad, (unsigned)c->nr);
        HIPCK("upload", h2d(*d_idx32 + acc, dum.data(), sizeof(unsigned) * pad, c->stream));
    } else {
        HIPCK("alloc", hipMalloc((void**)d_idx, sizeof(unsigned short) * (size_t)(acc + pad)));
        const std::vector<unsigned short> dum(pad, (unsigned short)c->nr);
        HIPCK("upload", h2d(*d_idx + acc, dum.data(), sizeof(unsigned short) * pad, c->stream));
    }
    if (dummy_pos) *dummy_pos = (unsigned)acc;
    hipLaunchKernelGGL(nlist_kernel<true>, grid, dim3(NLT), 0, c->stream, c->d_rxyzq, c->nr, o[0], o[1], o[2], h,
                       n[0], n[1], n[2], rad, pad, lanemajor ? 1 : 0, *d_off, wide ? nullptr : *d_idx,
                       wide ? *d_idx32 : nullptr);
    HIPCK("launch", hipGetLastError());
    HIPCK("sync", hipStreamSynchronize(c->stream));
    return (long long)std::max(acc, 1ull);
}

static void free_iface_lists(lkl_ctx* c) {
    if (c->d_noff_if) { (void)hipFree(c->d_noff_if); c->d_noff_if = nullptr; }
    if (c->d_nidx_if) { (void)hipFree(c->d_nidx_if); c->d_nidx_if = nullptr; }
    if (c->d_nidx32_if) { (void)hipFree(c->d_nidx32_if); c->d_nidx32_if = nullptr; }
    c->iflist_version = ~0u;
}

static int build_lists(lkl_ctx* c) {
    if (c->d_noff) { (void)hipFree(c->d_noff); c->d_noff = nullptr; }
    if (c->d_nidx) { (void)hipFree(c->d_nidx); c->d_nidx = nullptr; }
    if (c->d_nidx32) { (void)hipFree(c->d_nidx32); c->d_nidx32 = nullptr; }
    free_iface_lists(c);
    c->list_entries = 0;
    const int nr = c->nr;
    if (!(c->list_h > 0.0f) || nr <= 0) return 0;
    const float h = c->list_h;
    const float rad = std::sqrt(WIN2) + 0.5f * std::sqrt(3.0f) * h + 0.01f;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = 0; i < nr; ++i)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], (double)c->r_xyz[3 * i + a]);
            hi[a] = std::max(hi[a], (double)c->r_xyz[3 * i + a]);
        }
    long long nodes = 1;
    int n[3];
    float o[3];
    for (int a = 0; a < 3; ++a) {
        o[a] = (float)std::floor(lo[a] - rad - h);
        n[a] = (int)std::ceil((hi[a] + rad + h - (double)o[a]) / h) + 1;
        nodes *= n[a];
    }
    if (nodes > LIST_NODES_MAX || (n[1] + NLY - 1) / NLY > 65535 || (n[2] + NLZ - 1) / NLZ > 65535) return 0;
    const long long entries = fill_lists(c, o, n, h, rad, LIST_PAD, true, &c->d_noff, &c->d_nidx, &c->d_nidx32,
                                         &c->list_dummy);
    if (entries <= 0) return (int)entries;
    for (int a = 0; a < 3; ++a) { c->nn[a] = n[a]; c->no[a] = o[a]; }
    c->list_entries = (size_t)entries;
    c->version++;
    return 0;
}

// The interface pass's short lists, on the main lists' nodes (built on first
// use after each receptor layout change).  Radius: every receptor atom whose
// f32 squared distance can reach the pass's filter (15.211) from a point
// within the node's half-diagonal, plus the same 0.01 A margin.
static int ensure_iface_lists(lkl_ctx* c) {
    if (c->iflist_version == c->version && c->d_noff_if) return 0;
    free_iface_lists(c);
    const float h = c->list_h;
    const float rad = std::sqrt(15.211f) + 0.5f * std::sqrt(3.0f) * h + 0.01f;
    const long long entries = fill_lists(c, c->no, c->nn, h, rad, IFACE_PAD, false, &c->d_noff_if, &c->d_nidx_if,
                                         &c->d_nidx32_if);
    if (entries < 0) return -1;
    if (entries == 0) { set_err("interface lists exceed the entry cap"); return -1; }
    c->iflist_version = c->version;
    return 0;
}

static int build_receptor(lkl_ctx* c) {
    const int nr = c->nr;
    const double edge[3] = {(double)cell_x_of(c), (double)c->cell_size, (double)c->cell_size};
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    if (nr > 0) {
        for (int a = 0; a < 3; ++a) { lo[a] = INFINITY; hi[a] = -INFINITY; }
        for (int i = 0; i < nr; ++i)
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], (double)c->r_xyz[3 * i + a]);
                hi[a] = std::max(hi[a], (double)c->r_xyz[3 * i + a]);
            }
    }
    int n[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = std::floor(lo[a]) - 1.0;
        n[a] = std::max(1, (int)std::floor((hi[a] - lo[a]) / edge[a]) + 1);
    }
    const size_t ncell = (size_t)n[0] * n[1] * n[2];
    if (ncell > (size_t)1 << 28) { set_err("receptor cell grid too large (%zu cells)", ncell); return -1; }
    std::vector<int> cell(nr), start(ncell + 1, 0);
    for (int i = 0; i < nr; ++i) {
        int ii[3];
        for (int a = 0; a < 3; ++a) {
            ii[a] = (int)std::floor(((double)c->r_xyz[3 * i + a] - lo[a]) / edge[a]);
            ii[a] = std::min(std::max(ii[a], 0), n[a] - 1);
        }
        cell[i] = (ii[2] * n[1] + ii[1]) * n[0] + ii[0];
        start[cell[i] + 1]++;
    }
    for (size_t k = 0; k < ncell; ++k) start[k + 1] += start[k];
    std::vector<int> cur(start.begin(), start.end() - 1), order(nr);
    for (int i = 0; i < nr; ++i) order[cur[cell[i]]++] = i;
    c->rperm = order;
    // one extra record past the sorted atoms: the neighbour lists' padding
    // entry, far from every atom (excluded from desolvation by its flag)
    std::vector<float4> xyzq(nr + 1), par(nr + 1);
    std::vector<unsigned char> flag(nr + 1, 1);
    xyzq[nr] = make_float4(DUMMY_XYZ, DUMMY_XYZ, DUMMY_XYZ, 0.f);
    par[nr] = make_float4(0.f, 0.f, 0.f, 0.f);
    std::vector<float> des(nr, 0.f), asa(nr, 0.f);
    for (int p = 0; p < nr; ++p) {
        const int i = order[p];
        xyzq[p] = make_float4(c->r_xyz[3 * i], c->r_xyz[3 * i + 1], c->r_xyz[3 * i + 2],
                              c->r_ele.empty() ? 0.f : c->r_ele[i]);
        par[p] = make_float4(c->r_svdw.empty() ? 0.f : c->r_svdw[i],
                             c->r_vdwr.empty() ? 0.f : c->r_vdwr[i],
                             (!c->r_heavy.empty() && c->r_heavy[i]) ? 1.f : 0.f, 0.f);
        flag[p] = 0;
        if (c->has_desolv) {
            flag[p] = c->r_flag[i];
            des[p] = c->r_des[i];
            asa[p] = c->r_asa[i];
        }
    }
    if (dev_upload(&c->d_rxyzq, xyzq.data(), (size_t)nr + 1, c->stream)) return -1;
    if (dev_upload(&c->d_rpar, par.data(), (size_t)nr + 1, c->stream)) return -1;
    {
        std::vector<float4> rec(2 * ((size_t)nr + 1));
        // (x, y, z, q) then (svdw, radius, heavy, 0); a receptor atom excluded
        // from desolvation (and the padding record) carries its radius with
        // the sign bit set, which the pair terms take the magnitude of
        for (size_t p = 0; p <= (size_t)nr; ++p) {
            rec[2 * p] = xyzq[p];
            rec[2 * p + 1] = par[p];
            if (flag[p]) rec[2 * p + 1].y = -std::fabs(par[p].y);
        }
        if (dev_upload(&c->d_rrec, rec.data(), rec.size(), c->stream)) return -1;
    }
    {
        // parameter types: the distinct rpar records, when there are at most
        // RTYPE_MAX of them (the dummy takes type 0: it fails every pair test)
        std::map<std::array<unsigned, 3>, int> ids;
        std::vector<float4> tab;
        std::vector<unsigned char> typ(((size_t)nr + 4) / 4 * 4, 0);
        bool fits = true;
        for (int p = 0; p < nr && fits; ++p) {
            std::array<unsigned, 3> key;
            std::memcpy(key.data(), &par[p], sizeof key);
            auto it = ids.find(key);
            if (it == ids.end()) {
                if ((int)tab.size() == RTYPE_MAX) { fits = false; break; }
                it = ids.emplace(key, (int)tab.size()).first;
                tab.push_back(par[p]);
            }
            typ[p] = (unsigned char)it->second;
        }
        c->ntype = 0;
        if (fits && !tab.empty()) {
            if (dev_upload(&c->d_rtab, tab.data(), tab.size(), c->stream)) return -1;
            if (dev_upload(&c->d_rtyp, reinterpret_cast<const unsigned*>(typ.data()), typ.size() / 4, c->stream))
                return -1;
            c->ntype = (int)tab.size();
        }
    }
    if (dev_upload(&c->d_cstart, start.data(), ncell + 1, c->stream)) return -1;
    if (c->has_desolv) {
        if (dev_upload(&c->d_rflag, flag.data(), (size_t)nr + 1, c->stream)) return -1;
        if (dev_upload(&c->d_rdes, des.data(), (size_t)nr, c->stream)) return -1;
        if (dev_upload(&c->d_rasa, asa.data(), (size_t)nr, c->stream)) return -1;
    }
    if (!c->d_wgrid && nr > 0) {  // receptor-only: built once
        double wlo[3], whi[3];
        long long nodes = 1;
        for (int a = 0; a < 3; ++a) {
            wlo[a] = INFINITY;
            whi[a] = -INFINITY;
        }
        for (int i = 0; i < nr; ++i)
            for (int a = 0; a < 3; ++a) {
                wlo[a] = std::min(wlo[a], (double)c->r_xyz[3 * i + a]);
                whi[a] = std::max(whi[a], (double)c->r_xyz[3 * i + a]);
            }
        for (int a = 0; a < 3; ++a) {
            c->wo[a] = std::floor(wlo[a] - WORK_PAD);
            c->wn[a] = (int)std::ceil((whi[a] + WORK_PAD - c->wo[a]) / WORK_H) + 1;
            nodes *= c->wn[a];
        }
        const float rw = std::sqrt(WIN2) + 0.5f * std::sqrt(3.0f) * WORK_H + 0.01f;
        HIPCK("alloc", hipMalloc((void**)&c->d_wgrid, sizeof(unsigned short) * (size_t)nodes));
        hipLaunchKernelGGL(workgrid_kernel, dim3((unsigned)((nodes + 255) / 256)), dim3(256), 0, c->stream,
                           c->d_rxyzq, nr, (float)c->wo[0], (float)c->wo[1], (float)c->wo[2], c->wn[0],
                           c->wn[1], nodes, rw * rw, c->d_wgrid);
        HIPCK("launch", hipGetLastError());
        HIPCK("sync", hipStreamSynchronize(c->stream));
    }
    c->version++;
    c->ncx = n[0];
    c->ncy = n[1];
    c->ncz = n[2];
    c->cox = lo[0];
    c->coy = lo[1];
    c->coz = lo[2];
    return build_lists(c);
}

// Morton code of a 10-bit-per-axis quantised point.
static uint32_t spread3(uint32_t v) {
    v &= 0x3ff;
    v = (v | (v << 16)) & 0x30000ff;
    v = (v | (v << 8)) & 0x300f00f;
    v = (v | (v << 4)) & 0x30c30c3;
    v = (v | (v << 2)) & 0x9249249;
    return v;
}

// Ligand layout: Morton order of the reference coordinates (1 A quanta), so a
// wavefront's 64 atoms are spatially compact after any rigid transform.
static int build_ligand(lkl_ctx* c, const float* xyz, const float* ele, const float* svdw,
                        const float* vdwr, const unsigned char* heavy, const unsigned char* flag,
                        const float* des, const float* asa) {
    const int nl = c->nl;
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    for (int j = 0; j < nl; ++j)
        for (int a = 0; a < 3; ++a) lo[a] = std::min(lo[a], xyz[3 * j + a]);
    std::vector<uint32_t> code(nl);
    for (int j = 0; j < nl; ++j) {
        uint32_t q[3];
        for (int a = 0; a < 3; ++a)
            q[a] = (uint32_t)std::min(1023.0f, std::max(0.0f, std::floor(xyz[3 * j + a] - lo[a])));
        code[j] = spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2);
    }
    c->lperm.resize(nl);
    std::iota(c->lperm.begin(), c->lperm.end(), 0);
    std::stable_sort(c->lperm.begin(), c->lperm.end(),
                     [&](int a, int b) { return code[a] < code[b]; });
    std::vector<float4> lx(nl), lp(nl);
    std::vector<unsigned char> lf(nl, 0);
    std::vector<float> ld(nl, 0.f), la(nl, 0.f);
    for (int s = 0; s < nl; ++s) {
        const int j = c->lperm[s];
        lx[s] = make_float4(xyz[3 * j], xyz[3 * j + 1], xyz[3 * j + 2], ele ? ele[j] : 0.f);
        lp[s] = make_float4(svdw ? svdw[j] : 0.f, vdwr ? vdwr[j] : 0.f,
                            (heavy && heavy[j]) ? 1.f : 0.f, 0.f);
        if (flag) {
            lf[s] = flag[j];
            ld[s] = des[j];
            la[s] = asa[j];
        }
    }
    c->lq.resize(nl);
    for (int s = 0; s < nl; ++s) c->lq[s] = lx[s].w;
    if (dev_upload(&c->d_lxyzq, lx.data(), (size_t)nl, c->stream)) return -1;
    if (dev_upload(&c->d_lpar, lp.data(), (size_t)nl, c->stream)) return -1;
    {
        // ligand parameter types: the distinct lpar records (sorted order),
        // when there are at most RTYPE_MAX of them
        std::map<std::array<unsigned, 3>, int> ids;
        std::vector<float4> tab;
        std::vector<unsigned char> typ(((size_t)nl + 3) / 4 * 4, 0);
        bool fits = true;
        for (int j = 0; j < nl && fits; ++j) {
            std::array<unsigned, 3> key;
            std::memcpy(key.data(), &lp[j], sizeof key);
            auto it = ids.find(key);
            if (it == ids.end()) {
                if ((int)tab.size() == RTYPE_MAX) { fits = false; break; }
                it = ids.emplace(key, (int)tab.size()).first;
                tab.push_back(lp[j]);
            }
            typ[j] = (unsigned char)it->second;
        }
        c->nltype = 0;
        if (fits && !tab.empty()) {
            if (dev_upload(&c->d_lttab, tab.data(), tab.size(), c->stream)) return -1;
            if (dev_upload(&c->d_lty, reinterpret_cast<const unsigned*>(typ.data()), typ.size() / 4, c->stream))
                return -1;
            c->nltype = (int)tab.size();
        }
    }
    if (flag) {
        if (dev_upload(&c->d_lflag, lf.data(), (size_t)nl, c->stream)) return -1;
        if (dev_upload(&c->d_ldes, ld.data(), (size_t)nl, c->stream)) return -1;
        if (dev_upload(&c->d_lasa, la.data(), (size_t)nl, c->stream)) return -1;
    }
    c->version++;
    return 0;
}

static DevComplex dev_view(const lkl_ctx* c) {
    DevComplex C;
    memset(&C, 0, sizeof C);
    C.rxyzq = c->d_rxyzq;
    C.rpar = c->d_rpar;
    C.rrec = c->d_rrec;
    if (c->ntype) {
        C.rtyp = c->d_rtyp;
        C.rtab = c->d_rtab;
        C.ntype = c->ntype;
    }
    if (c->nltype) {
        C.lty = reinterpret_cast<const unsigned char*>(c->d_lty);
        C.lttab = c->d_lttab;
        C.nltype = c->nltype;
    }
    C.cstart = c->d_cstart;
    C.rflag = c->d_rflag;
    C.rdes = c->d_rdes;
    C.rasa = c->d_rasa;
    C.nr = c->nr;
    C.ncx = c->ncx;
    C.ncy = c->ncy;
    C.ncz = c->ncz;
    C.cox = c->cox;
    C.coy = c->coy;
    C.coz = c->coz;
    C.csyz = c->cell_size;
    C.inv_csyz = (float)(1.0 / (double)c->cell_size);
    C.inv_csx = (float)(1.0 / (double)cell_x_of(c));
    C.rwin = std::sqrt(WIN2);
    C.rwin2 = C.rwin * C.rwin;
    C.wgrid = c->work_sort ? c->d_wgrid : nullptr;
    C.wnx = c->wn[0];
    C.wny = c->wn[1];
    C.wnz = c->wn[2];
    C.wox = (float)c->wo[0];
    C.woy = (float)c->wo[1];
    C.woz = (float)c->wo[2];
    C.inv_wh = 1.0f / WORK_H;
    if (c->list_entries) {
        C.noff = c->d_noff;
        C.nidx = c->d_nidx;
        C.nidx32 = c->d_nidx32;
        C.ndum = c->list_dummy;
        C.nnx = c->nn[0];
        C.nny = c->nn[1];
        C.nnz = c->nn[2];
        C.nox = c->no[0];
        C.noy = c->no[1];
        C.noz = c->no[2];
        C.inv_nh = 1.0f / c->list_h;
    }
    C.lxyzq = c->d_lxyzq;
    C.lpar = c->d_lpar;
    C.lflag = c->d_lflag;
    C.ldes = c->d_ldes;
    C.lasa = c->d_lasa;
    C.nl = c->nl;
    C.phi = c->d_phi;
    C.phi8 = c->d_phi8;
    C.nx = c->nx;
    C.ny = c->ny;
    C.nz = c->nz;
    C.ox = c->ox;
    C.oy = c->oy;
    C.oz = c->oz;
    C.sp = c->sp;
    {
        int ex = 0;
        C.sp_pow2 = c->sp > 0.0 && std::frexp(c->sp, &ex) == 0.5 ? 1 : 0;
        C.inv_sp = C.sp_pow2 ? std::ldexp(1.0, 1 - ex) : 1.0 / c->sp;
    }
    return C;
}

static void ctx_free(lkl_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    void* bufs[] = {c->d_phi, c->d_phi8, c->d_rxyzq, c->d_rpar, c->d_rrec, c->d_rtyp, c->d_rtab, c->d_lty, c->d_lttab, c->d_cstart, c->d_rflag, c->d_rdes,
                    c->d_rasa, c->d_lxyzq, c->d_lpar, c->d_lflag, c->d_ldes, c->d_lasa,
                    c->d_poses, c->d_out, c->d_outS, c->d_gmin, c->d_pa_e, c->d_pa_v, c->d_xyz,
                    c->d_acc, c->d_cnt, c->d_wgrid, c->d_noff, c->d_nidx, c->d_iface, c->d_icounts,
                    c->d_noff_if, c->d_nidx_if, c->d_lcoords, c->d_pcost, c->d_porder, c->d_ocnt, c->d_nidx32,
                    c->d_nidx32_if, c->d_smin};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    for (hipEvent_t e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

static int probe_device() {
    if (fail_at("probe")) return 0;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
    return strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
}

static lkl_ctx* ctx_create(const lkl_complex_desc* d) {
    if (!d) { set_err("lkl_ctx_create: null descriptor"); return nullptr; }
    if (d->nr < 0 || d->nl <= 0) { set_err("lkl_ctx_create: need nl > 0 and nr >= 0 (nr=%d nl=%d)", d->nr, d->nl); return nullptr; }
    if (d->nr > 0 && !d->r_xyz) { set_err("lkl_ctx_create: r_xyz is null"); return nullptr; }
    // the list walk addresses receptor records by 32-bit byte offsets (32 B each)
    if (d->nr >= (1 << 27)) { set_err("lkl_ctx_create: nr=%d exceeds the 134,217,727-atom receptor limit", d->nr); return nullptr; }
    if (!d->l_xyz) { set_err("lkl_ctx_create: l_xyz is null"); return nullptr; }
    if (!probe_device()) { set_err("lkl_ctx_create: no gfx950 device available"); return nullptr; }
    lkl_ctx* c = new lkl_ctx();
    auto fail = [&]() -> lkl_ctx* { ctx_free(c); return nullptr; };
    if (hipGetDevice(&c->device) != hipSuccess) { set_err("hipGetDevice failed"); return fail(); }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) c->lds_max = prop.sharedMemPerBlock;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        set_err("hipStreamCreate failed");
        return fail();
    }
    c->nr = d->nr;
    c->nl = d->nl;
    if (d->cell_size > 0.0f) c->cell_size = d->cell_size;
    const size_t nr = (size_t)d->nr;
    auto cp = [&](std::vector<float>& v, const float* s, size_t n) { if (s) v.assign(s, s + n); };
    cp(c->r_xyz, d->r_xyz, 3 * nr);
    cp(c->r_ele, d->r_ele, nr);
    cp(c->r_svdw, d->r_svdw, nr);
    cp(c->r_vdwr, d->r_vdwr, nr);
    if (d->r_heavy) c->r_heavy.assign(d->r_heavy, d->r_heavy + nr);
    c->has_desolv = d->r_flag && d->r_des && d->r_asa && d->l_flag && d->l_des && d->l_asa;
    if (c->has_desolv) {
        c->r_flag.assign(d->r_flag, d->r_flag + nr);
        cp(c->r_des, d->r_des, nr);
        cp(c->r_asa, d->r_asa, nr);
    }
    c->has_phi = d->phi && d->nx > 1 && d->ny > 1 && d->nz > 1 && d->sp > 0.0;
    if (c->has_phi) {
        c->nx = d->nx; c->ny = d->ny; c->nz = d->nz;
        c->ox = d->ox; c->oy = d->oy; c->oz = d->oz; c->sp = d->sp;
        if (dev_upload(&c->d_phi, d->phi, (size_t)d->nx * d->ny * d->nz, c->stream)) return fail();
        if (build_phi8(c)) return fail();
    }
    if (build_receptor(c)) return fail();
    if (build_ligand(c, d->l_xyz, d->l_ele, d->l_svdw, d->l_vdwr, d->l_heavy,
                     c->has_desolv ? d->l_flag : nullptr, d->l_des, d->l_asa))
        return fail();
    return c;
}

// Kernel dispatch over the flag word.
typedef void (*score_fn)(DevComplex, const double*, int, double*, double*, unsigned*, float*,
                         float*, long long*, unsigned*, int, int);

// Zeroed per-pose tile accumulators (zeroed once here; every launch leaves them zero).
static int reserve_acc(lkl_ctx* c, size_t N) {
    if (N <= c->cap_acc && c->d_acc) return 0;
    if (c->d_acc) { (void)hipFree(c->d_acc); c->d_acc = nullptr; }
    if (c->d_cnt) { (void)hipFree(c->d_cnt); c->d_cnt = nullptr; }
    c->cap_acc = 0;
    HIPCK("alloc", hipMalloc((void**)&c->d_acc, sizeof(long long) * 3 * N));
    HIPCK("alloc", hipMalloc((void**)&c->d_cnt, sizeof(unsigned) * N));
    // cleared on the context's stream and waited for: the first launch may be
    // enqueued on another stream, which must not overtake the clear
    HIPCK("alloc", hipMemsetAsync(c->d_acc, 0, sizeof(long long) * 3 * N, c->stream));
    HIPCK("alloc", hipMemsetAsync(c->d_cnt, 0, sizeof(unsigned) * N, c->stream));
    HIPCK("sync", hipStreamSynchronize(c->stream));
    c->cap_acc = N;
    return 0;
}

template <bool LIST, int DES, bool PA, int TY = 0>
static score_fn pick3(unsigned f) {
    switch (f & 7u) {
        case 0: return score_kernel<LIST, false, false, false, DES, PA, TY>;
        case 1: return score_kernel<LIST, true, false, false, DES, PA, TY>;
        case 2: return score_kernel<LIST, false, true, false, DES, PA, TY>;
        case 3: return score_kernel<LIST, true, true, false, DES, PA, TY>;
        case 4: return score_kernel<LIST, false, false, true, DES, PA, TY>;
        case 5: return score_kernel<LIST, true, false, true, DES, PA, TY>;
        case 6: return score_kernel<LIST, false, true, true, DES, PA, TY>;
        default: return score_kernel<LIST, true, true, true, DES, PA, TY>;
    }
}

static int validate_flags(const lkl_ctx* c, unsigned flags) {
    if (flags & ~(F_FAR | F_ELEC | F_CLASH | F_DESOLV)) { set_err("unknown flag bits 0x%x", flags); return -1; }
    if ((flags & F_FAR) && !c->has_phi) { set_err("F_FAR requested but the context has no far field"); return -1; }
    if ((flags & F_DESOLV) && !c->has_desolv) { set_err("F_DESOLV requested but the context has no desolvation sides"); return -1; }
    return 0;
}

// LDS bytes of a work-ordered tile of nT ligand atoms: near_groups' arrays
// (list bounds and the active list) when grouped, else near_sorted's (index
// list and bucket ids).
static size_t tile_lds(size_t nT, bool grouped) { return grouped ? 10 * nT + 16 : 3 * nT + 16; }
// LDS bytes of the receptor's parameter types beside a grouped tile: the
// table and one type byte per sorted atom and the dummy, in whole words.
static size_t type_lds(const lkl_ctx* c) { return 16 * (size_t)c->ntype + ((size_t)c->nr + 4) / 4 * 4; }
// LDS of the pair-type table and the receptor's type bytes; 0 when the
// context cannot use it (no ligand types, or a table over PAIR_TAB_MAX)
static size_t pair_lds(const lkl_ctx* c) {
    if (!c->ntype || !c->nltype || c->ntype * c->nltype > PAIR_TAB_MAX) return 0;
    return 16 * (size_t)c->ntype * (size_t)c->nltype + ((size_t)c->nr + 4) / 4 * 4;
}

// Device scratch of one launch given by the caller instead of the context's
// own grown-on-demand buffers: per-pose cost and dispatch order, desolvation
// minima of the global variant, and the tile accumulators of multi-tile
// poses (cleared by the launch itself).
struct LaunchScratch {
    unsigned* pcost = nullptr;
    int* porder = nullptr;
    unsigned* ocnt = nullptr;
    unsigned* gmin = nullptr;
    long long* acc = nullptr;
    unsigned* cnt = nullptr;
};

// Clears n 64-bit words (a fill kernel: capturable, unlike a memset node).
__global__ void zero_u64_kernel(unsigned long long* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__global__ void fill_u32_kernel(unsigned* __restrict__ p, size_t n, unsigned v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Desolvation minima of one pose fit in LDS beside the kernel's own arrays.
static bool des_in_lds(const lkl_ctx* c) {
    return (size_t)c->nr * sizeof(unsigned) + 1024 <= c->lds_max && !getenv("LKL_FORCE_GLOBAL_MIN");
}

// Largest global receptor-minima buffer a split desolvation launch may use
// (N x nr keys); larger launches keep one tile per pose, which balances well
// at that size anyway.
constexpr size_t SPLIT_MIN_BYTES = (size_t)64 << 20;

// The split launches' receptor minima: every key at the initial distance's
// key when allocated (the last tile of each pose restores them after use).
static int reserve_split_min(lkl_ctx* c, size_t n) {
    if (n <= c->cap_smin && c->d_smin) return 0;
    if (c->d_smin) { (void)hipFree(c->d_smin); c->d_smin = nullptr; }
    c->cap_smin = 0;
    HIPCK("alloc", hipMalloc((void**)&c->d_smin, sizeof(unsigned) * n));
    const float init_f = (float)MIN_INIT;
    unsigned init;
    std::memcpy(&init, &init_f, sizeof init);
    // on the context's stream and waited for: the first launch may be
    // enqueued on another stream, which must not overtake the fill
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->d_smin, n,
                       init);
    HIPCK("alloc", hipGetLastError());
    HIPCK("sync", hipStreamSynchronize(c->stream));
    c->cap_smin = n;
    return 0;
}

// Enqueue one score launch; all pointers are device pointers.
static int launch_score(lkl_ctx* c, const double* d_poses, int N, unsigned flags, double* d_out,
                        double* d_outS, hipStream_t st, bool peratom, const float4* d_lcoords = nullptr,
                        const LaunchScratch* ext = nullptr, double* pose_copy = nullptr) {
    if (N <= 0) return 0;
    if (validate_flags(c, flags)) return -1;
    DevComplex C = dev_view(c);
    if (d_lcoords) {  // per-pose ligand records (lkl_ctx_score_coords)
        C.lxyzq = d_lcoords;
        C.lstride = c->nl;
    }
    int des = DES_NONE;
    int dchunk = c->nl;  // ligand atoms per desolvation tile
    size_t shm = 0;
    // work-ordered near field: tile-local lists of (chunk x 3) bytes in LDS
    const bool want_sort = C.wgrid != nullptr && !peratom;
    const bool list = C.noff != nullptr;
    if (flags & F_DESOLV) {
        if (!d_outS) { set_err("F_DESOLV requires an outS buffer"); return -1; }
        // static LDS of the kernel (reduction slots) needs headroom too
        if (des_in_lds(c)) {
            des = DES_LDS;
            shm = (size_t)c->nr * sizeof(unsigned);
        } else {
            des = DES_GLOBAL;
            if (!ext && dev_reserve(&c->d_gmin, &c->cap_gmin, (size_t)N * c->nr)) return -1;
        }
        // a tile holds the whole ligand, or (minima in LDS, a ligand tile
        // set, a launch small enough for the global minima) one chunk of it;
        // the near field is ordered if the tile's arrays fit (8 bytes of
        // alignment after the receptor minima)
        if (des == DES_LDS && !ext && !peratom && c->chunk > 0 && c->chunk < c->nl &&
            (size_t)N * c->nr * sizeof(unsigned) <= SPLIT_MIN_BYTES)
            dchunk = c->chunk;
        const size_t tl = tile_lds((size_t)dchunk, list) + 8;
        if (want_sort && dchunk <= WORK_TILE_MAX && shm + tl + 1024 <= c->lds_max) shm += tl;
        else C.wgrid = nullptr;
    }
    if (peratom && dev_reserve(&c->d_pa_e, &c->cap_pa, (size_t)N * c->nl)) return -1;
    if (peratom && dev_reserve(&c->d_pa_v, &c->cap_pa_v, (size_t)N * c->nl)) return -1;
    score_fn fn;
    if (peratom) fn = list ? pick3<true, DES_NONE, true>(flags) : pick3<false, DES_NONE, true>(flags);
    else if (des == DES_LDS) fn = list ? pick3<true, DES_LDS, false>(flags) : pick3<false, DES_LDS, false>(flags);
    else if (des == DES_GLOBAL)
        fn = list ? pick3<true, DES_GLOBAL, false>(flags) : pick3<false, DES_GLOBAL, false>(flags);
    else fn = list ? pick3<true, DES_NONE, false>(flags) : pick3<false, DES_NONE, false>(flags);
    int chunk = des != DES_NONE ? dchunk : c->chunk > 0 ? std::min(c->chunk, c->nl) : c->nl;
    if (des == DES_NONE && want_sort) {
        const int tmax = (int)std::min<size_t>(WORK_TILE_MAX, (c->lds_max - 2048) / (list ? 10 : 3));
        if (chunk > tmax) chunk = tmax;  // LDS bound of the tile's arrays
        shm = tile_lds((size_t)chunk, list);
        // the list walk reads receptor parameters from an LDS type table when
        // the receptor has one and the table fits beside the tile's arrays
        // (ligand x receptor pair types first, when they fit)
        const size_t typed = shm + type_lds(c), paired = shm + pair_lds(c);
        if (list && !peratom && c->param_types >= 2 && pair_lds(c) &&
            paired + 1024 <= c->lds_max) {
            fn = pick3<true, DES_NONE, false, 2>(flags);
            shm = paired;
        } else if (list && !peratom && c->param_types && c->ntype && typed + 1024 <= c->lds_max) {
            fn = pick3<true, DES_NONE, false, 1>(flags);
            shm = typed;
        }
    }
    if (peratom) C.wgrid = nullptr;
    const int nchunks = (c->nl + chunk - 1) / chunk;
    if ((size_t)N * nchunks >= ((size_t)1 << 31)) { set_err("too many tiles (N=%d)", N); return -1; }
    if (!ext && nchunks > 1 && reserve_acc(c, (size_t)N)) return -1;
    const bool split = des == DES_LDS && nchunks > 1;
    if (split && reserve_split_min(c, (size_t)N * c->nr)) return -1;
    unsigned* gmin = ext ? ext->gmin : split ? c->d_smin : c->d_gmin;
    long long* acc = ext ? ext->acc : c->d_acc;
    unsigned* cnt = ext ? ext->cnt : c->d_cnt;
    if (fail_at("launch")) { set_err("launch_score stage=launch: injected failure"); return -1; }
    if (c->timing) (void)hipEventRecord(c->ev[0], st);
    c->ev_prepass = false;
    // longest-first dispatch when the launch takes a few rounds of the chip's
    // workgroup slots: a short pre-pass estimates each pose's cost and orders
    // each XCD's range (results do not depend on the order)
    if (ext && nchunks > 1) {  // the caller's accumulators start at zero
        // one fill over acc (3N words) and cnt, which follows it directly
        if ((void*)cnt != (void*)(acc + 3 * (size_t)N)) { set_err("scratch: cnt must follow acc"); return -1; }
        const size_t w = 3 * (size_t)N + ((size_t)N + 1) / 2;
        hipLaunchKernelGGL(zero_u64_kernel, dim3((unsigned)((w + 255) / 256)), dim3(256), 0, st,
                           (unsigned long long*)acc, w);
    }
    if (c->pose_order && list && nchunks == 1 && !peratom && N > POSE_ORDER_MIN && N <= POSE_ORDER_MAX) {
        if (!ext && dev_reserve(&c->d_pcost, &c->cap_pcost, (size_t)N)) return -1;
        if (!ext && dev_reserve(&c->d_porder, &c->cap_porder, (size_t)N)) return -1;
        if (!ext && !c->d_ocnt) {  // zeroed once; every launch leaves them zero
            HIPCK("alloc", hipMalloc((void**)&c->d_ocnt, 8 * sizeof(unsigned)));
            hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, st, (unsigned long long*)c->d_ocnt, (size_t)4);
        }
        unsigned* pcost = ext ? ext->pcost : c->d_pcost;
        int* porder = ext ? ext->porder : c->d_porder;
        unsigned* ocnt = ext ? ext->ocnt : c->d_ocnt;
        if (ext)  // the shared scratch holds other launches' data: clear the counters
            hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, st, (unsigned long long*)ocnt, (size_t)4);
        const unsigned per_range = ((unsigned)N / 8u + 1u + ORDER_W - 1u) / ORDER_W;
        hipLaunchKernelGGL(pose_order_kernel, dim3(8u * per_range), dim3(64 * ORDER_W), 0, st, C, d_poses, N, pcost,
                           porder, ocnt, pose_copy);
        HIPCK("launch", hipGetLastError());
        C.porder = porder;
        c->ev_prepass = true;
        if (pose_copy) d_poses = pose_copy;
    }
    if (c->timing) (void)hipEventRecord(c->ev[1], st);
    hipLaunchKernelGGL(fn, dim3((unsigned)(N * nchunks)), dim3((unsigned)c->block), shm, st, C,
                       d_poses, N, d_out, d_outS, gmin, c->d_pa_e, c->d_pa_v, acc, cnt, chunk, nchunks);
    HIPCK("launch", hipGetLastError());
    if (c->timing) (void)hipEventRecord(c->ev[2], st);
    return 0;
}

// Blocking host-buffer score.  The poses go into a pinned, device-mapped
// staging buffer that the kernels read over the bus: the dispatch-order
// pre-pass keeps a device copy of each pose it reads for the scoring kernel
// (without the pre-pass the scoring kernel reads the staging buffer itself),
// and the scoring kernel stores the scores straight into the staging buffer.
// No DMA copies: one launch sequence and one stream synchronisation per call.
// Staging layout: N x 7 poses, then N x 2 scores and N desolvation terms.
static int score_host(lkl_ctx* c, const double* poses, int N, unsigned flags, double* out,
                      double* outS, float* pa_e, float* pa_v) {
    if (N <= 0) { set_err("N must be positive (N=%d)", N); return -1; }
    if (!poses || !out) { set_err("null poses/out buffer"); return -1; }
    if ((flags & F_DESOLV) && !outS) { set_err("F_DESOLV requires outS"); return -1; }
    HIPCK("probe", hipSetDevice(c->device));
    (void)hipGetLastError();  // drain a stale sticky error (full_score.cu:375)
    if (dev_reserve(&c->d_poses, &c->cap_poses, (size_t)N * 7)) return -1;
    const size_t n = (size_t)N;
    const size_t stage = n * 10;
    if (stage > c->cap_stage) {
        if (c->h_stage) (void)hipHostFree(c->h_stage);
        c->h_stage = c->d_stage = nullptr;
        c->cap_stage = 0;
        HIPCK("alloc", hipHostMalloc((void**)&c->h_stage, sizeof(double) * stage,
                                     hipHostMallocMapped | hipHostMallocCoherent));
        const hipError_t e = hipHostGetDevicePointer((void**)&c->d_stage, c->h_stage, 0);
        if (e != hipSuccess || !c->d_stage) {
            (void)hipHostFree(c->h_stage);
            c->h_stage = c->d_stage = nullptr;
            HIPCK("alloc", e != hipSuccess ? e : hipErrorInvalidValue);
        }
        c->cap_stage = stage;
    }
    memcpy(c->h_stage, poses, sizeof(double) * 7 * n);
    HIPCK("upload", hipSuccess);  // the failure-injection stage of the pose hand-over
    double* d_res = c->d_stage + 7 * n;
    if (launch_score(c, c->d_stage, N, flags, d_res, (flags & F_DESOLV) ? d_res + 2 * n : nullptr, c->stream,
                     pa_e != nullptr, nullptr, nullptr, c->d_poses)) {
        (void)hipStreamSynchronize(c->stream);  // nothing may still read the staging buffer
        return -1;
    }
    HIPCK("sync", hipStreamSynchronize(c->stream));
    HIPCK("readback", hipSuccess);
    const double* h_res = c->h_stage + 7 * n;
    memcpy(out, h_res, sizeof(double) * 2 * n);
    if (flags & F_DESOLV) memcpy(outS, h_res + 2 * n, sizeof(double) * n);
    if (pa_e) {
        std::vector<float> e(n * c->nl), v(n * c->nl);
        HIPCK("readback", hipMemcpy(e.data(), c->d_pa_e, sizeof(float) * e.size(), hipMemcpyDeviceToHost));
        HIPCK("readback", hipMemcpy(v.data(), c->d_pa_v, sizeof(float) * v.size(), hipMemcpyDeviceToHost));
        for (int k = 0; k < N; ++k)
            for (int s = 0; s < c->nl; ++s) {
                pa_e[(size_t)k * c->nl + c->lperm[s]] = e[(size_t)k * c->nl + s];
                pa_v[(size_t)k * c->nl + c->lperm[s]] = v[(size_t)k * c->nl + s];
            }
    }
    return 0;
}

// ── C ABI: context API ────────────────────────────────────────────────────────
LKL_API int lkl_device_available(void) { return probe_device(); }

LKL_API const char* lkl_last_error(void) { return t_err.c_str(); }

LKL_API lkl_ctx* lkl_ctx_create(const lkl_complex_desc* desc) { return ctx_create(desc); }

LKL_API void lkl_ctx_destroy(lkl_ctx* ctx) { ctx_free(ctx); }

LKL_API int lkl_ctx_score(lkl_ctx* ctx, const double* poses, int N, unsigned flags, double* out,
                          double* outS) {
    if (!ctx) { set_err("lkl_ctx_score: null context"); return -1; }
    return score_host(ctx, poses, N, flags, out, outS, nullptr, nullptr);
}

LKL_API int lkl_ctx_score_async(lkl_ctx* ctx, const double* d_poses, int N, unsigned flags,
                                double* d_out, double* d_outS, void* stream) {
    if (!ctx) { set_err("lkl_ctx_score_async: null context"); return -1; }
    if (N < 0) { set_err("N must be >= 0"); return -1; }
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    return launch_score(ctx, d_poses, N, flags, d_out, d_outS, st, false);
}

// ── no-allocation variant of lkl_ctx_score_async ─────────────────────────────
// For callers that cannot allocate device memory (code under graph capture,
// serving loops, harnesses that forbid it): the launch's scratch lives in one
// static device array of SCRATCH_BYTES, shared by every call on a device.
// Calls that overlap on one device race on it.
constexpr size_t SCRATCH_BYTES = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];
constexpr int SCRATCH_DEVICES = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_addr[SCRATCH_DEVICES];  // each device's copy, found once

// The address of the current device's copy of `scratch` (NULL and an error
// message on failure).  Host code never uses the symbol itself as an address.
static unsigned char* scratch_of(int dev) {
    if (dev < 0 || dev >= SCRATCH_DEVICES) { set_err("scratch: device %d out of range", dev); return nullptr; }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_err("scratch: hipGetSymbolAddress failed on device %d: %s", dev, hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_addr[dev] = (unsigned char*)p;
    }
    return g_scratch_addr[dev];
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Scratch bytes of one no-allocation launch of N poses: cost and order (4 B
// each), the 8 dispatch-order counters, accumulators (24 B) with the counters
// right after them (4 B, a whole number of 8-byte words), then the global
// desolvation minima (4 B per receptor atom per pose) when they do not fit in
// LDS.
static size_t noalloc_layout(const lkl_ctx* c, int N, unsigned flags, unsigned char* base, LaunchScratch* L) {
    size_t off = 0;
    const size_t n = (size_t)N;
    auto take = [&](size_t bytes) { const size_t o = off; off += align256(bytes); return base ? base + o : nullptr; };
    L->pcost = (unsigned*)take(4 * n);
    L->porder = (int*)take(4 * n);
    L->ocnt = (unsigned*)take(32);
    unsigned char* a = take(24 * n + 8 * ((n + 1) / 2));
    L->acc = (long long*)a;
    L->cnt = a ? (unsigned*)(a + 24 * n) : nullptr;
    L->gmin = ((flags & F_DESOLV) && !des_in_lds(c)) ? (unsigned*)take(4 * n * (size_t)c->nr) : nullptr;
    return off;
}

LKL_API int lkl_ctx_score_async_noalloc(lkl_ctx* ctx, const double* d_poses, int N, unsigned flags,
                                        double* d_out, double* d_outS, void* stream) {
    if (!ctx) { set_err("lkl_ctx_score_async_noalloc: null context"); return -1; }
    if (N < 0) { set_err("N must be >= 0"); return -1; }
    if (N == 0) return 0;
    LaunchScratch L;
    const size_t need = noalloc_layout(ctx, N, flags, nullptr, &L);
    if (need > SCRATCH_BYTES) {
        set_err("lkl_ctx_score_async_noalloc: %d poses need %zu bytes of scratch, more than its %zu-byte "
                "maximum (about %zu poses without global desolvation minima)",
                N, need, SCRATCH_BYTES, SCRATCH_BYTES / 40);
        return -1;
    }
    int dev = 0;
    HIPCK("probe", hipGetDevice(&dev));
    if (dev != ctx->device) { set_err("lkl_ctx_score_async_noalloc: the context lives on device %d, current is %d", ctx->device, dev); return -1; }
    unsigned char* base = scratch_of(dev);
    if (!base) return -1;
    noalloc_layout(ctx, N, flags, base, &L);
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    return launch_score(ctx, d_poses, N, flags, d_out, d_outS, st, false, nullptr, &L);
}

// The ligand already placed, per pose: coords N x nl x 3 f32 in the ligand's
// input order.  Each pose gets its own records (sorted order, charge in w) and
// the identity pose, whose f64 transform returns the f32 inputs exactly, so a
// pose scores as the batched kernel scores those coordinates.  Blocking.
LKL_API int lkl_ctx_score_coords(lkl_ctx* c, const float* coords, int N, unsigned flags, double* out,
                                 double* outS) {
    if (!c) { set_err("lkl_ctx_score_coords: null context"); return -1; }
    if (N <= 0 || !coords || !out) { set_err("lkl_ctx_score_coords: bad arguments (N=%d)", N); return -1; }
    if ((flags & F_DESOLV) && !outS) { set_err("F_DESOLV requires outS"); return -1; }
    const size_t nl = (size_t)c->nl, n = (size_t)N * nl;
    std::vector<float4> rec(n);
    for (size_t k = 0; k < (size_t)N; ++k)
        for (size_t s = 0; s < nl; ++s) {
            const float* p = coords + (k * nl + (size_t)c->lperm[s]) * 3;
            if (!(std::fabs(p[0]) < 1e6f && std::fabs(p[1]) < 1e6f && std::fabs(p[2]) < 1e6f)) {
                set_err("lkl_ctx_score_coords: pose %zu atom %d has a non-finite or |x| >= 1e6 coordinate", k,
                        c->lperm[s]);
                return -1;
            }
            rec[k * nl + s] = make_float4(p[0], p[1], p[2], c->lq[s]);
        }
    std::vector<double> ident((size_t)N * 7, 0.0);
    for (size_t k = 0; k < (size_t)N; ++k) ident[k * 7] = 1.0;
    HIPCK("probe", hipSetDevice(c->device));
    (void)hipGetLastError();
    if (dev_reserve(&c->d_lcoords, &c->cap_lcoords, n)) return -1;
    if (dev_reserve(&c->d_poses, &c->cap_poses, (size_t)N * 7)) return -1;
    if (dev_reserve(&c->d_out, &c->cap_out, (size_t)N * 2)) return -1;
    if ((flags & F_DESOLV) && dev_reserve(&c->d_outS, &c->cap_outS, (size_t)N)) return -1;
    HIPCK("upload", h2d(c->d_lcoords, rec.data(), sizeof(float4) * n, c->stream));
    HIPCK("upload", h2d(c->d_poses, ident.data(), sizeof(double) * ident.size(), c->stream));
    if (launch_score(c, c->d_poses, N, flags, c->d_out, (flags & F_DESOLV) ? c->d_outS : nullptr, c->stream,
                     false, c->d_lcoords))
        return -1;
    HIPCK("sync", hipStreamSynchronize(c->stream));
    HIPCK("readback", hipMemcpy(out, c->d_out, sizeof(double) * 2 * (size_t)N, hipMemcpyDeviceToHost));
    if (flags & F_DESOLV)
        HIPCK("readback", hipMemcpy(outS, c->d_outS, sizeof(double) * (size_t)N, hipMemcpyDeviceToHost));
    return 0;
}

LKL_API int lkl_ctx_set_timing(lkl_ctx* c, int on) {
    if (!c) { set_err("lkl_ctx_set_timing: null context"); return -1; }
    if (on)
        for (hipEvent_t& e : c->ev)
            if (!e) HIPCK("alloc", hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    c->timing = on != 0;
    return 0;
}

LKL_API int lkl_ctx_kernel_ms(lkl_ctx* c, double* prepass_ms, double* score_ms) {
    if (!c || !c->timing) { set_err("lkl_ctx_kernel_ms: timing is not enabled on this context"); return -1; }
    HIPCK("sync", hipEventSynchronize(c->ev[2]));
    float a = 0.0f, b = 0.0f;
    HIPCK("sync", hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
    HIPCK("sync", hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
    if (prepass_ms) *prepass_ms = c->ev_prepass ? a : 0.0;
    if (score_ms) *score_ms = b;
    return 0;
}

LKL_API int lkl_ctx_set_pose_order(lkl_ctx* c, int enable) {
    if (!c) { set_err("lkl_ctx_set_pose_order: null context"); return -1; }
    c->pose_order = enable != 0;
    return 0;
}

LKL_API int lkl_ctx_pose_order(const lkl_ctx* c) { return c ? (c->pose_order ? 1 : 0) : -1; }

LKL_API int lkl_ctx_build_field(lkl_ctx* c) {
    if (!c) { set_err("lkl_ctx_build_field: null context"); return -1; }
    if (c->nr <= 0 || c->r_ele.empty()) { set_err("lkl_ctx_build_field: receptor charges required"); return -1; }
    HIPCK("probe", hipSetDevice(c->device));
    double o[3];
    int d[3];
    const long long total = lkl_field_dims(c->r_xyz.data(), c->nr, o, d);
    if (total <= 0) { set_err("lkl_ctx_build_field: bad field box"); return -1; }
    if (c->d_phi) { (void)hipFree(c->d_phi); c->d_phi = nullptr; }
    if (c->d_phi8) { (void)hipFree(c->d_phi8); c->d_phi8 = nullptr; }
    c->has_phi = false;
    HIPCK("alloc", hipMalloc((void**)&c->d_phi, sizeof(float) * (size_t)total));
    if (lkl_field_build_device(c->r_xyz.data(), c->r_ele.data(), c->nr, c->d_phi, c->stream)) {
        set_err("lkl_ctx_build_field: field kernel failed");
        return -1;
    }
    c->nx = d[0]; c->ny = d[1]; c->nz = d[2];
    c->ox = o[0]; c->oy = o[1]; c->oz = o[2]; c->sp = 0.5;
    c->has_phi = true;
    c->version++;
    return build_phi8(c);
}

LKL_API int lkl_ctx_field_info(const lkl_ctx* c, double* origin3, int* dims3) {
    if (!c || !c->has_phi) return -1;
    if (origin3) { origin3[0] = c->ox; origin3[1] = c->oy; origin3[2] = c->oz; }
    if (dims3) { dims3[0] = c->nx; dims3[1] = c->ny; dims3[2] = c->nz; }
    return 0;
}

LKL_API int lkl_ctx_transform(lkl_ctx* c, const double* poses, int N, double* out_xyz) {
    if (!c || !poses || !out_xyz || N <= 0) { set_err("lkl_ctx_transform: bad arguments"); return -1; }
    HIPCK("probe", hipSetDevice(c->device));
    if (dev_reserve(&c->d_poses, &c->cap_poses, (size_t)N * 7)) return -1;
    if (dev_reserve(&c->d_xyz, &c->cap_xyz, (size_t)N * c->nl * 3)) return -1;
    HIPCK("upload", h2d(c->d_poses, poses, sizeof(double) * 7 * (size_t)N, c->stream));
    hipLaunchKernelGGL(transform_kernel, dim3((unsigned)N), dim3(256), 0, c->stream, dev_view(c),
                       c->d_poses, N, c->d_xyz);
    HIPCK("launch", hipGetLastError());
    HIPCK("sync", hipStreamSynchronize(c->stream));
    std::vector<double> tmp((size_t)N * c->nl * 3);
    HIPCK("readback", hipMemcpy(tmp.data(), c->d_xyz, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost));
    for (int k = 0; k < N; ++k)
        for (int s = 0; s < c->nl; ++s)
            for (int a = 0; a < 3; ++a)
                out_xyz[((size_t)k * c->nl + c->lperm[s]) * 3 + a] = tmp[((size_t)k * c->nl + s) * 3 + a];
    return 0;
}

LKL_API int lkl_ctx_set_tiles(lkl_ctx* c, float cell_size, int block, int chunk) {
    if (!c) { set_err("lkl_ctx_set_tiles: null context"); return -1; }
    if (!(cell_size > 0.5f && cell_size <= 20.0f)) { set_err("cell_size %.3f out of (0.5, 20]", cell_size); return -1; }
    if (block != 64 && block != 128 && block != 256 && block != 512 && block != 1024) {
        set_err("block %d not in {64,128,256,512,1024}", block);
        return -1;
    }
    if (chunk < 0) { set_err("chunk %d must be >= 0", chunk); return -1; }
    HIPCK("probe", hipSetDevice(c->device));
    c->block = block;
    c->chunk = chunk;
    if (cell_size != c->cell_size) {
        c->cell_size = cell_size;
        if (build_receptor(c)) return -1;
    }
    return 0;
}

LKL_API int lkl_ctx_set_work_sort(lkl_ctx* c, int on) {
    if (!c) { set_err("lkl_ctx_set_work_sort: null context"); return -1; }
    if (c->work_sort != (on != 0)) {
        c->work_sort = on != 0;
        c->version++;  // batches built on this context refresh their views
    }
    return 0;
}

// Interface data: receptor and ligand restraint residues as CSR over atom
// indices (reference order) and the membrane beads (receptor atom indices).
LKL_API int lkl_ctx_set_interface(lkl_ctx* c, const int* rec_res_off, const int* rec_res_atoms, int n_rec_res,
                                  const int* lig_res_off, const int* lig_res_atoms, int n_lig_res,
                                  const int* membrane_atoms, int n_membrane) {
    if (!c) { set_err("lkl_ctx_set_interface: null context"); return -1; }
    if (n_rec_res < 0 || n_lig_res < 0 || n_membrane < 0 || (n_rec_res && (!rec_res_off || !rec_res_atoms)) ||
        (n_lig_res && (!lig_res_off || !lig_res_atoms)) || (n_membrane && !membrane_atoms)) {
        set_err("lkl_ctx_set_interface: bad arguments");
        return -1;
    }
    auto check_csr = [](const int* off, const int* atoms, int n, int bound) {
        if (n == 0) return true;
        if (off[0] != 0) return false;
        for (int k = 0; k < n; ++k)
            if (off[k + 1] < off[k]) return false;
        for (int q = 0; q < off[n]; ++q)
            if (atoms[q] < 0 || atoms[q] >= bound) return false;
        return true;
    };
    if (!check_csr(rec_res_off, rec_res_atoms, n_rec_res, c->nr) ||
        !check_csr(lig_res_off, lig_res_atoms, n_lig_res, c->nl)) {
        set_err("lkl_ctx_set_interface: residue offsets must start at 0, not decrease, and index atoms in range");
        return -1;
    }
    for (int k = 0; k < n_membrane; ++k)
        if (membrane_atoms[k] < 0 || membrane_atoms[k] >= c->nr) {
            set_err("lkl_ctx_set_interface: membrane atom %d out of range", membrane_atoms[k]);
            return -1;
        }
    c->if_roff.assign(rec_res_off ? rec_res_off : nullptr, rec_res_off ? rec_res_off + n_rec_res + 1 : nullptr);
    c->if_ratoms.assign(rec_res_atoms ? rec_res_atoms : nullptr, rec_res_atoms && n_rec_res ? rec_res_atoms + rec_res_off[n_rec_res] : rec_res_atoms);
    c->if_loff.assign(lig_res_off ? lig_res_off : nullptr, lig_res_off ? lig_res_off + n_lig_res + 1 : nullptr);
    c->if_latoms.assign(lig_res_atoms ? lig_res_atoms : nullptr, lig_res_atoms && n_lig_res ? lig_res_atoms + lig_res_off[n_lig_res] : lig_res_atoms);
    c->if_mem.assign(membrane_atoms ? membrane_atoms : nullptr, membrane_atoms ? membrane_atoms + n_membrane : nullptr);
    if (c->if_roff.empty()) c->if_roff.assign(1, 0);
    if (c->if_loff.empty()) c->if_loff.assign(1, 0);
    c->has_iface = n_rec_res > 0 || n_lig_res > 0 || n_membrane > 0;
    c->iface_version = ~0u;
    return 0;
}

// Device copy of the interface data in the current sorted layouts (rebuilt
// after the receptor is re-sorted).  Returns the view, or -1.
static int iface_view(lkl_ctx* c, IfaceDev* I) {
    if (c->iface_version != c->version) {
        std::vector<int> inv_r(c->nr), inv_l(c->nl);
        for (int p = 0; p < c->nr; ++p) inv_r[c->rperm[p]] = p;
        for (int s = 0; s < c->nl; ++s) inv_l[c->lperm[s]] = s;
        std::vector<int> blob;
        blob.insert(blob.end(), c->if_roff.begin(), c->if_roff.end());
        for (int a : c->if_ratoms) blob.push_back(inv_r[a]);
        blob.insert(blob.end(), c->if_loff.begin(), c->if_loff.end());
        for (int a : c->if_latoms) blob.push_back(inv_l[a]);
        for (int a : c->if_mem) blob.push_back(inv_r[a]);
        if (dev_reserve(&c->d_iface, &c->cap_iface, blob.size())) return -1;
        HIPCK("upload", h2d(c->d_iface, blob.data(), sizeof(int) * blob.size(), c->stream));
        c->iface_version = c->version;
    }
    int* p = c->d_iface;
    I->nres_r = (int)c->if_roff.size() - 1;
    I->nres_l = (int)c->if_loff.size() - 1;
    I->nmem = (int)c->if_mem.size();
    I->r_off = p; p += c->if_roff.size();
    I->r_atoms = p; p += c->if_ratoms.size();
    I->l_off = p; p += c->if_loff.size();
    I->l_atoms = p; p += c->if_latoms.size();
    I->mem = p;
    return 0;
}

// Per-pose interface counts (see iface_kernel) for N device poses into
// d_counts (N x 3 ints) on `stream` (NULL = the context's stream).
LKL_API int lkl_ctx_interface_async(lkl_ctx* c, const double* d_poses, int N, int* d_counts, void* stream) {
    if (!c) { set_err("lkl_ctx_interface_async: null context"); return -1; }
    if (N < 0 || (N > 0 && (!d_poses || !d_counts))) { set_err("lkl_ctx_interface_async: bad arguments"); return -1; }
    if (!c->has_iface) { set_err("lkl_ctx_interface_async: no restraints or membrane set"); return -1; }
    if (!c->list_entries) {
        set_err("lkl_ctx_interface_async: interface flags need the neighbour lists (on, receptor <= 65,535 atoms)");
        return -1;
    }
    if (N == 0) return 0;
    IfaceDev I;
    if (ensure_iface_lists(c) || iface_view(c, &I)) return -1;
    I.noff = c->d_noff_if;
    I.nidx = c->d_nidx_if;
    I.nidx32 = c->d_nidx32_if;
    const size_t shm = sizeof(unsigned) * (size_t)((c->nr + 31) / 32 + (c->nl + 31) / 32);
    if (shm + 1024 > c->lds_max) { set_err("lkl_ctx_interface_async: interface bitsets exceed LDS"); return -1; }
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    hipLaunchKernelGGL(iface_kernel, dim3((unsigned)N), dim3(256), shm, st, dev_view(c), I, d_poses, N, d_counts);
    HIPCK("launch", hipGetLastError());
    return 0;
}

// Blocking host-buffer variant: counts (N x 3 ints).
LKL_API int lkl_ctx_interface(lkl_ctx* c, const double* poses, int N, int* counts) {
    if (!c || !poses || !counts || N <= 0) { set_err("lkl_ctx_interface: bad arguments"); return -1; }
    HIPCK("probe", hipSetDevice(c->device));
    if (dev_reserve(&c->d_poses, &c->cap_poses, (size_t)N * 7)) return -1;
    if (dev_reserve(&c->d_icounts, &c->cap_icounts, (size_t)N * 3)) return -1;
    HIPCK("upload", h2d(c->d_poses, poses, sizeof(double) * 7 * (size_t)N, c->stream));
    if (lkl_ctx_interface_async(c, c->d_poses, N, c->d_icounts, nullptr)) return -1;
    HIPCK("sync", hipStreamSynchronize(c->stream));
    HIPCK("readback", hipMemcpy(counts, c->d_icounts, sizeof(int) * 3 * (size_t)N, hipMemcpyDeviceToHost));
    return 0;
}

LKL_API int lkl_ctx_has_interface(const lkl_ctx* c) { return c && c->has_iface ? 1 : 0; }

LKL_API int lkl_ctx_interface_sizes(const lkl_ctx* c, int* n_rec_res, int* n_lig_res, int* n_membrane) {
    if (!c) return -1;
    if (n_rec_res) *n_rec_res = c->has_iface ? (int)c->if_roff.size() - 1 : 0;
    if (n_lig_res) *n_lig_res = c->has_iface ? (int)c->if_loff.size() - 1 : 0;
    if (n_membrane) *n_membrane = c->has_iface ? (int)c->if_mem.size() : 0;
    return 0;
}

LKL_API int lkl_ctx_set_phi_packed(lkl_ctx* c, int on) {
    if (!c) { set_err("lkl_ctx_set_phi_packed: null context"); return -1; }
    HIPCK("probe", hipSetDevice(c->device));
    if (c->phi_packed != (on != 0)) {
        c->phi_packed = on != 0;
        if (build_phi8(c)) return -1;
        c->version++;
    }
    return 0;
}

LKL_API int lkl_ctx_phi_packed(const lkl_ctx* c) { return c && c->d_phi8 ? 1 : 0; }

LKL_API int lkl_ctx_set_param_types(lkl_ctx* c, int on) {
    if (!c) { set_err("lkl_ctx_set_param_types: null context"); return -1; }
    const int level = on <= 0 ? 0 : on == 1 ? 1 : 2;
    if (level != c->param_types) {
        c->param_types = level;
        c->version++;  // batches built on this context refresh their views
    }
    return 0;
}

LKL_API int lkl_ctx_param_types(const lkl_ctx* c) { return c && c->param_types ? c->ntype : 0; }

LKL_API int lkl_ctx_set_list_spacing(lkl_ctx* c, float spacing) {
    if (!c) { set_err("lkl_ctx_set_list_spacing: null context"); return -1; }
    if (!(spacing == 0.0f || (spacing >= 0.25f && spacing <= 4.0f))) {
        set_err("list spacing %.3f not 0 or in [0.25, 4]", spacing);
        return -1;
    }
    HIPCK("probe", hipSetDevice(c->device));
    if (spacing != c->list_h) {
        c->list_h = spacing;
        if (build_lists(c)) return -1;
        c->version++;
    }
    return 0;
}

LKL_API long long lkl_ctx_list_info(const lkl_ctx* c, int* dims3) {
    if (!c) return -1;
    if (dims3) { dims3[0] = c->nn[0]; dims3[1] = c->nn[1]; dims3[2] = c->nn[2]; }
    return c->list_entries ? (long long)c->list_entries : 0;
}

LKL_API int lkl_ctx_set_cell_x(lkl_ctx* c, float cell_x) {
    if (!c) { set_err("lkl_ctx_set_cell_x: null context"); return -1; }
    if (!(cell_x == 0.0f || (cell_x > 0.5f && cell_x <= 20.0f))) { set_err("cell_x %.3f not 0 or in (0.5, 20]", cell_x); return -1; }
    HIPCK("probe", hipSetDevice(c->device));
    if (cell_x != c->cell_x) {
        c->cell_x = cell_x;
        if (build_receptor(c)) return -1;
    }
    return 0;
}

// ── autotuner ────────────────────────────────────────────────────────────────
// Times every (cell edge y/z, cell edge x, workgroup size, ligand tile)
// candidate on the caller's poses with HIP events and keeps the fastest.
// Per-pose results are bitwise identical for every candidate (fixed-point
// sums), so tuning never changes a score.  table (optional) receives rows of
// (cell, cell_x, block, chunk, median ms).
static const float k_cells[] = {3.4f, 4.0f, 5.0f, 6.0f, 7.5f};
static const float k_cell_x[] = {0.0f, 2.5f, 1.25f};  // 0 = cubic cells
static const int k_blocks[] = {64, 128, 256, 512, 1024};  // workgroup sizes the autotuners try
static const int k_chunk_mult[] = {1, 2, 4, 0};  // ligand atoms per lane (0 = whole ligand)
constexpr int TUNE_ROWS = 320;

LKL_API int lkl_ctx_autotune(lkl_ctx* c, const double* poses, int N, unsigned flags,
                             float* best_cell, float* best_cell_x, int* best_block, int* best_chunk,
                             float* table, int* n_table) {
    if (!c || !poses || N <= 0) { set_err("lkl_ctx_autotune: bad arguments"); return -1; }
    if (validate_flags(c, flags)) return -1;
    HIPCK("probe", hipSetDevice(c->device));
    if (dev_reserve(&c->d_poses, &c->cap_poses, (size_t)N * 7)) return -1;
    if (dev_reserve(&c->d_out, &c->cap_out, (size_t)N * 2)) return -1;
    if (dev_reserve(&c->d_outS, &c->cap_outS, (size_t)N)) return -1;
    HIPCK("upload", h2d(c->d_poses, poses, sizeof(double) * 7 * (size_t)N, c->stream));
    hipEvent_t e0, e1;
    HIPCK("alloc", hipEventCreate(&e0));
    HIPCK("alloc", hipEventCreate(&e1));
    float bc = c->cell_size, bcx = c->cell_x, bms = INFINITY;
    int bb = c->block, bch = c->chunk, k = 0;
    int rc = 0;
    for (float cell : k_cells) {
        for (float cx : k_cell_x) {
            if (cx != 0.0f && cx >= cell) continue;  // same as (or coarser than) cubic
            if (cell != c->cell_size || cx != c->cell_x) {
                c->cell_size = cell;
                c->cell_x = cx;
                if (build_receptor(c)) { rc = -1; break; }
            }
            for (int block : k_blocks) {
                for (int mult : k_chunk_mult) {
                    const int chunk = mult ? block * mult : 0;
                    if (mult && chunk >= c->nl) continue;  // same as the whole-ligand tile
                    // desolvation tiles split a pose only with minima in LDS
                    // and a launch within the global minima's bound
                    if ((flags & F_DESOLV) && mult &&
                        !(des_in_lds(c) && (size_t)N * c->nr * sizeof(unsigned) <= SPLIT_MIN_BYTES))
                        continue;
                    c->block = block;
                    c->chunk = chunk;
                    double* dS = (flags & F_DESOLV) ? c->d_outS : nullptr;
                    if (launch_score(c, c->d_poses, N, flags, c->d_out, dS, c->stream, false)) { rc = -1; break; }
                    float ms[3];
                    for (int r = 0; r < 3 && rc == 0; ++r) {
                        (void)hipEventRecord(e0, c->stream);
                        if (launch_score(c, c->d_poses, N, flags, c->d_out, dS, c->stream, false)) rc = -1;
                        (void)hipEventRecord(e1, c->stream);
                        (void)hipEventSynchronize(e1);
                        (void)hipEventElapsedTime(&ms[r], e0, e1);
                    }
                    if (rc) break;
                    std::sort(ms, ms + 3);
                    if (table && k < TUNE_ROWS) {
                        float* row = table + 5 * k;
                        row[0] = cell; row[1] = cell_x_of(c); row[2] = (float)block;
                        row[3] = (float)chunk; row[4] = ms[1];
                    }
                    ++k;
                    if (ms[1] < bms) { bms = ms[1]; bc = cell; bcx = cx; bb = block; bch = chunk; }
                }
                if (rc) break;
            }
            if (rc) break;
        }
        if (rc) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return -1;
    if (n_table) *n_table = std::min(k, TUNE_ROWS);
    c->block = bb;
    c->chunk = bch;
    if (bc != c->cell_size || bcx != c->cell_x) {
        c->cell_size = bc;
        c->cell_x = bcx;
        if (build_receptor(c)) return -1;
    }
    if (best_cell) *best_cell = bc;
    if (best_cell_x) *best_cell_x = cell_x_of(c);
    if (best_block) *best_block = bb;
    if (best_chunk) *best_chunk = bch;
    return 0;
}

// ── batch object ─────────────────────────────────────────────────────────────
struct lkl_multi {
    std::vector<lkl_ctx*> ctxs;  // not owned
    std::vector<unsigned> flags;
    std::vector<int> npose, pose_off, order;  // order: problems by descending cost
    std::vector<unsigned> versions;
    int K = 0, P = 0, total_tiles = 0, device = 0, cus = 256;
    int block = 256, chunk_mult = 2;
    size_t lds = 0;  // dynamic LDS of the work-ordered tiles (largest chunk x 3 bytes)
    bool list = false;  // every problem has neighbour lists (or no receptor atoms)
    hipStream_t stream = nullptr;
    DevComplex* d_cx = nullptr;
    ProblemDesc* d_prob = nullptr;
    double *d_poses = nullptr, *d_out = nullptr;
    long long* d_acc = nullptr;
    unsigned* d_cnt = nullptr;
    unsigned long long* d_trace = nullptr;  // per-tile records of the last launch (lkl_multi_set_trace)
    size_t cap_trace = 0;
    bool trace_on = false;
    double *h_poses = nullptr, *h_out = nullptr;  // pinned staging
    bool use_graph = true, graph_ok = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static void batch_drop_graph(lkl_multi* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_ok = false;
}

static void batch_free(lkl_multi* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    batch_drop_graph(b);
    void* bufs[] = {b->d_cx, b->d_prob, b->d_poses, b->d_out, b->d_acc, b->d_cnt, b->d_trace};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    if (b->h_poses) (void)hipHostFree(b->h_poses);
    if (b->h_out) (void)hipHostFree(b->h_out);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

// (Re)build the problem descriptors from the contexts' current layouts.
static int batch_refresh(lkl_multi* b) {
    std::vector<DevComplex> cx(b->K);
    std::vector<ProblemDesc> pd(b->K);
    int tile = 0;
    b->lds = 0;
    b->list = true;  // one kernel variant for the batch: lists unless a problem lacks them
    for (int r = 0; r < b->K; ++r) {
        cx[r] = dev_view(b->ctxs[b->order[r]]);
        if (!cx[r].noff && b->ctxs[b->order[r]]->nr > 0) b->list = false;
    }
    for (int r = 0; r < b->K; ++r) {
        const int k = b->order[r];
        lkl_ctx* c = b->ctxs[k];
        int chunk = b->chunk_mult > 0 ? std::min(c->nl, b->block * b->chunk_mult) : c->nl;
        if (cx[r].wgrid) {  // LDS bound of the work-ordered tile's arrays
            const size_t tmax = std::min<size_t>(WORK_TILE_MAX, (c->lds_max - 2048) / (b->list ? 10 : 3));
            chunk = std::min(chunk, (int)tmax);
            size_t need = tile_lds((size_t)chunk, b->list);
            // the pair-type table, else the receptor's parameter-type table,
            // beside the tile's arrays, when it fits
            if (b->list && c->param_types >= 2 && pair_lds(c) && need + pair_lds(c) + 1024 <= c->lds_max) {
                need += pair_lds(c);
            } else if (b->list && c->param_types && c->ntype && need + type_lds(c) + 1024 <= c->lds_max) {
                need += type_lds(c);
                cx[r].nltype = 0;
            } else {
                cx[r].ntype = 0;
                cx[r].nltype = 0;
            }
            b->lds = std::max(b->lds, need);
        } else {
            cx[r].ntype = 0;
            cx[r].nltype = 0;
        }
        pd[r].tile_begin = tile;
        pd[r].nchunks = (c->nl + chunk - 1) / chunk;
        pd[r].chunk = chunk;
        pd[r].pose_off = b->pose_off[k];
        pd[r].flags = b->flags[k];
        tile += b->npose[k] * pd[r].nchunks;
        b->versions[k] = c->version;
    }
    b->total_tiles = tile;
    if (b->trace_on && dev_reserve(&b->d_trace, &b->cap_trace, 4 * (size_t)std::max(tile, 1))) return -1;
    HIPCK("upload", h2d(b->d_cx, cx.data(), sizeof(DevComplex) * b->K, b->stream));
    HIPCK("upload", h2d(b->d_prob, pd.data(), sizeof(ProblemDesc) * b->K, b->stream));
    batch_drop_graph(b);
    return 0;
}

static int batch_check(lkl_multi* b) {
    for (int k = 0; k < b->K; ++k)
        if (b->versions[k] != b->ctxs[k]->version) return batch_refresh(b);
    return 0;
}

typedef void (*multi_fn)(const DevComplex*, const ProblemDesc*, int, int, const double*, double*, long long*,
                         unsigned*, unsigned long long*);

static int batch_launch(lkl_multi* b, const double* d_poses, double* d_out, hipStream_t st) {
    if (b->total_tiles <= 0) return 0;
    multi_fn fn = b->list ? multi_kernel<true> : multi_kernel<false>;
    hipLaunchKernelGGL(fn, dim3((unsigned)b->total_tiles), dim3((unsigned)b->block), b->lds, st, b->d_cx,
                       b->d_prob, b->K, b->total_tiles, d_poses, d_out, b->d_acc, b->d_cnt, b->d_trace);
    HIPCK("launch", hipGetLastError());
    return 0;
}

LKL_API lkl_multi* lkl_multi_create(lkl_ctx* const* ctxs, const unsigned* flags,
                                    const int* n_poses, int n_problems) {
    if (!ctxs || !flags || !n_poses || n_problems <= 0) { set_err("lkl_multi_create: bad arguments"); return nullptr; }
    lkl_multi* b = new lkl_multi();
    auto fail = [&]() -> lkl_multi* { batch_free(b); return nullptr; };
    b->K = n_problems;
    double cost_total = 0;
    std::vector<double> cost(n_problems);
    for (int k = 0; k < n_problems; ++k) {
        lkl_ctx* c = ctxs[k];
        if (!c || n_poses[k] <= 0) { set_err("lkl_multi_create: problem %d has no context or poses", k); return fail(); }
        if (flags[k] & F_DESOLV) { set_err("lkl_multi_create: desolvation is not batched (problem %d)", k); return fail(); }
        if (validate_flags(c, flags[k])) return fail();
        b->ctxs.push_back(c);
        b->flags.push_back(flags[k]);
        b->npose.push_back(n_poses[k]);
        b->pose_off.push_back(b->P);
        b->P += n_poses[k];
        cost[k] = (double)n_poses[k] * c->nl * std::max(1, c->nr);
        cost_total += cost[k];
    }
    b->order.resize(n_problems);
    std::iota(b->order.begin(), b->order.end(), 0);
    std::stable_sort(b->order.begin(), b->order.end(), [&](int a, int c) { return cost[a] > cost[c]; });
    b->versions.assign(n_problems, ~0u);
    b->device = ctxs[0]->device;
    if (hipSetDevice(b->device) != hipSuccess) { set_err("hipSetDevice failed"); return fail(); }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, b->device) == hipSuccess) b->cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) { set_err("stream"); return fail(); }
    const size_t P = (size_t)b->P;
    if (hipMalloc((void**)&b->d_cx, sizeof(DevComplex) * b->K) != hipSuccess ||
        hipMalloc((void**)&b->d_prob, sizeof(ProblemDesc) * b->K) != hipSuccess ||
        hipMalloc((void**)&b->d_poses, sizeof(double) * 7 * P) != hipSuccess ||
        hipMalloc((void**)&b->d_out, sizeof(double) * 2 * P) != hipSuccess ||
        hipMalloc((void**)&b->d_acc, sizeof(long long) * 3 * P) != hipSuccess ||
        hipMalloc((void**)&b->d_cnt, sizeof(unsigned) * P) != hipSuccess ||
        hipHostMalloc((void**)&b->h_poses, sizeof(double) * 7 * P, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&b->h_out, sizeof(double) * 2 * P, hipHostMallocDefault) != hipSuccess) {
        set_err("lkl_multi_create: allocation failed (%zu poses)", P);
        return fail();
    }
    // cleared on the batch's stream and waited for before any launch
    if (hipMemsetAsync(b->d_acc, 0, sizeof(long long) * 3 * P, b->stream) != hipSuccess ||
        hipMemsetAsync(b->d_cnt, 0, sizeof(unsigned) * P, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess) {
        set_err("lkl_multi_create: clearing counters failed");
        return fail();
    }
    if (batch_refresh(b)) return fail();
    return b;
}

LKL_API void lkl_multi_destroy(lkl_multi* b) { batch_free(b); }

LKL_API int lkl_multi_poses(const lkl_multi* b) { return b ? b->P : -1; }

LKL_API int lkl_multi_set_tiles(lkl_multi* b, int block, int chunk_mult) {
    if (!b) { set_err("lkl_multi_set_tiles: null batch"); return -1; }
    if ((block != 64 && block != 128 && block != 256 && block != 512 && block != 1024) || chunk_mult < 0) {
        set_err("lkl_multi_set_tiles: block %d / chunk_mult %d out of range", block, chunk_mult);
        return -1;
    }
    HIPCK("probe", hipSetDevice(b->device));
    b->block = block;
    b->chunk_mult = chunk_mult;
    return batch_refresh(b);
}

LKL_API int lkl_multi_use_graph(lkl_multi* b, int enable) {
    if (!b) return -1;
    b->use_graph = enable != 0;
    return 0;
}

// Per-tile trace of every following launch (on) or none (off, the default).
// Switching drops the captured graph (its launch carries the trace pointer).
LKL_API int lkl_multi_set_trace(lkl_multi* b, int on) {
    if (!b) { set_err("lkl_multi_set_trace: null batch"); return -1; }
    HIPCK("probe", hipSetDevice(b->device));
    if ((on != 0) == b->trace_on) return 0;
    b->trace_on = on != 0;
    if (b->trace_on) {
        if (dev_reserve(&b->d_trace, &b->cap_trace, 4 * (size_t)std::max(b->total_tiles, 1))) return -1;
    } else {
        (void)hipFree(b->d_trace);
        b->d_trace = nullptr;
        b->cap_trace = 0;
    }
    batch_drop_graph(b);
    return 0;
}

// The last traced launch's records (blocking: waits for the whole device, as
// the launch may have been enqueued on any stream): 4 words per tile, tiles
// in launch order; returns the tile count (at most cap records copied) or -1.
// clock_khz (may be NULL) receives the wall clock's rate.
LKL_API int lkl_multi_trace(lkl_multi* b, unsigned long long* out, int cap, int* clock_khz) {
    if (!b || !out || cap < 0) { set_err("lkl_multi_trace: bad arguments"); return -1; }
    if (!b->trace_on) { set_err("lkl_multi_trace: tracing is off (lkl_multi_set_trace)"); return -1; }
    HIPCK("probe", hipSetDevice(b->device));
    if (clock_khz) HIPCK("probe", hipDeviceGetAttribute(clock_khz, hipDeviceAttributeWallClockRate, b->device));
    HIPCK("sync", hipDeviceSynchronize());
    const int n = std::min(cap, b->total_tiles);
    HIPCK("readback", hipMemcpy(out, b->d_trace, sizeof(unsigned long long) * 4 * (size_t)n, hipMemcpyDeviceToHost));
    return b->total_tiles;
}

// Device buffers: poses (P x 7, batch order), out (P x 2).  One kernel launch.
LKL_API int lkl_multi_score_async(lkl_multi* b, const double* d_poses, double* d_out, void* stream) {
    if (!b || !d_poses || !d_out) { set_err("lkl_multi_score_async: bad arguments"); return -1; }
    HIPCK("probe", hipSetDevice(b->device));
    if (batch_check(b)) return -1;
    return batch_launch(b, d_poses, d_out, stream ? (hipStream_t)stream : b->stream);
}

// Host buffers: one step = H2D poses, the batch kernel, D2H scores.  With
// the graph enabled the three are captured once into a hipGraph and replayed.
LKL_API int lkl_multi_score(lkl_multi* b, const double* poses, double* out) {
    if (!b || !poses || !out) { set_err("lkl_multi_score: bad arguments"); return -1; }
    HIPCK("probe", hipSetDevice(b->device));
    if (batch_check(b)) return -1;
    const size_t P = (size_t)b->P;
    memcpy(b->h_poses, poses, sizeof(double) * 7 * P);
    if (b->use_graph && !b->graph_ok) {
        HIPCK("capture", hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e1 = hipMemcpyAsync(b->d_poses, b->h_poses, sizeof(double) * 7 * P,
                                       hipMemcpyHostToDevice, b->stream);
        if (b->total_tiles > 0) {
            multi_fn fn = b->list ? multi_kernel<true> : multi_kernel<false>;
            hipLaunchKernelGGL(fn, dim3((unsigned)b->total_tiles), dim3((unsigned)b->block), b->lds, b->stream,
                               b->d_cx, b->d_prob, b->K, b->total_tiles, b->d_poses, b->d_out, b->d_acc,
                               b->d_cnt, b->d_trace);
        }
        hipError_t e2 = hipGetLastError();
        hipError_t e3 = hipMemcpyAsync(b->h_out, b->d_out, sizeof(double) * 2 * P,
                                       hipMemcpyDeviceToHost, b->stream);
        hipError_t e4 = hipStreamEndCapture(b->stream, &b->graph);
        if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess || e4 != hipSuccess) {
            set_err("lkl_multi_score: graph capture failed");
            batch_drop_graph(b);
            return -1;
        }
        HIPCK("capture", hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->graph_ok = true;
    }
    if (b->use_graph) {
        HIPCK("launch", hipGraphLaunch(b->exec, b->stream));
    } else {
        HIPCK("upload", hipMemcpyAsync(b->d_poses, b->h_poses, sizeof(double) * 7 * P,
                                       hipMemcpyHostToDevice, b->stream));
        if (batch_launch(b, b->d_poses, b->d_out, b->stream)) return -1;
        HIPCK("readback", hipMemcpyAsync(b->h_out, b->d_out, sizeof(double) * 2 * P,
                                         hipMemcpyDeviceToHost, b->stream));
    }
    HIPCK("sync", hipStreamSynchronize(b->stream));
    memcpy(out, b->h_out, sizeof(double) * 2 * P);
    return 0;
}

// Autotuner of the batch launch shape (block, ligand atoms per lane) on host
// poses; results are bitwise shape-invariant.
LKL_API int lkl_multi_autotune(lkl_multi* b, const double* poses, int* best_block, int* best_mult,
                               float* table, int* n_table) {
    if (!b || !poses) { set_err("lkl_multi_autotune: bad arguments"); return -1; }
    HIPCK("probe", hipSetDevice(b->device));
    HIPCK("upload", h2d(b->d_poses, poses, sizeof(double) * 7 * (size_t)b->P, b->stream));
    hipEvent_t e0, e1;
    HIPCK("alloc", hipEventCreate(&e0));
    HIPCK("alloc", hipEventCreate(&e1));
    int bb = b->block, bm = b->chunk_mult, k = 0, rc = 0;
    float bms = INFINITY;
    for (int block : k_blocks)
        for (int mult : {1, 2, 4, 0}) {
            if (rc) break;
            b->block = block; b->chunk_mult = mult;
            if (batch_refresh(b) || batch_launch(b, b->d_poses, b->d_out, b->stream)) { rc = -1; break; }
            float ms[3];
            for (int r = 0; r < 3 && rc == 0; ++r) {
                (void)hipEventRecord(e0, b->stream);
                if (batch_launch(b, b->d_poses, b->d_out, b->stream)) rc = -1;
                (void)hipEventRecord(e1, b->stream);
                (void)hipEventSynchronize(e1);
                (void)hipEventElapsedTime(&ms[r], e0, e1);
            }
            std::sort(ms, ms + 3);
            if (table && k < 64) {
                table[3 * k] = (float)block; table[3 * k + 1] = (float)mult; table[3 * k + 2] = ms[1];
            }
            ++k;
            if (ms[1] < bms) { bms = ms[1]; bb = block; bm = mult; }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return -1;
    b->block = bb; b->chunk_mult = bm;
    if (batch_refresh(b)) return -1;
    if (best_block) *best_block = bb;
    if (best_mult) *best_mult = bm;
    if (n_table) *n_table = std::min(k, 64);
    return 0;
}

// ── C ABI: stateless wrappers with the reference's signatures ─────────────────
// Contexts are cached per distinct host-buffer identity (the reference's
// BatchCache keys, full_score.cu:380-387; the cpydock cache now keys on
// pointers too, not only (nr, nl) as full_score.cu:594 did).
static std::mutex g_mu;
static std::map<std::vector<uint64_t>, lkl_ctx*> g_cache;
static std::vector<std::vector<uint64_t>> g_lru;

static uint64_t bits(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static uint64_t ptr(const void* p) { return (uint64_t)(uintptr_t)p; }

// Least recently used first in g_lru: a hit moves its key to the back, so the
// complexes a caller alternates between stay resident.
static lkl_ctx* cache_get(const std::vector<uint64_t>& key, const lkl_complex_desc& d) {
    auto it = g_cache.find(key);
    if (it != g_cache.end()) {
        auto pos = std::find(g_lru.begin(), g_lru.end(), key);
        if (pos != g_lru.end() && pos + 1 != g_lru.end()) {
            std::vector<uint64_t> k = std::move(*pos);
            g_lru.erase(pos);
            g_lru.push_back(std::move(k));
        }
        return it->second;
    }
    lkl_ctx* c = ctx_create(&d);
    if (!c) return nullptr;
    if (g_lru.size() >= 8) {  // bound the number of resident complexes
        auto old = g_cache.find(g_lru.front());
        if (old != g_cache.end()) { ctx_free(old->second); g_cache.erase(old); }
        g_lru.erase(g_lru.begin());
    }
    g_cache[key] = c;
    g_lru.push_back(key);
    return c;
}

LKL_API void lkl_cache_clear(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto& kv : g_cache) ctx_free(kv.second);
    g_cache.clear();
    g_lru.clear();
}

LKL_API int lkl_batch_score(const float* phi, int nx, int ny, int nz, float ox, float oy, float oz,
                            float sp, const float* r_coords, const float* r_ele,
                            const float* r_svdw, const float* r_vdwr, const unsigned char* r_heavy,
                            int nr, const int* cell_start, const int* cell_atoms, int ncx, int ncy,
                            int ncz, float c_ox, float c_oy, float c_oz, float c_sp,
                            const float* l_base, const double* poses, const float* l_ele,
                            const float* l_svdw, const float* l_vdwr,
                            const unsigned char* l_heavy, int nl, int N, unsigned flags,
                            double* out) {
    (void)cell_start; (void)cell_atoms; (void)ncx; (void)ncy; (void)ncz;
    (void)c_ox; (void)c_oy; (void)c_oz; (void)c_sp;
    std::lock_guard<std::mutex> lk(g_mu);
    if (N <= 0 || nl <= 0 || nr < 0) { set_err("lkl_batch_score: need N > 0, nl > 0 (N=%d nl=%d nr=%d)", N, nl, nr); return -1; }
    if (flags & ~(F_FAR | F_ELEC | F_CLASH)) { set_err("lkl_batch_score: unknown flag bits 0x%x", flags); return -1; }
    lkl_complex_desc d;
    memset(&d, 0, sizeof d);
    d.phi = phi; d.nx = nx; d.ny = ny; d.nz = nz; d.ox = ox; d.oy = oy; d.oz = oz; d.sp = sp;
    d.nr = nr; d.r_xyz = r_coords; d.r_ele = r_ele; d.r_svdw = r_svdw; d.r_vdwr = r_vdwr; d.r_heavy = r_heavy;
    d.nl = nl; d.l_xyz = l_base; d.l_ele = l_ele; d.l_svdw = l_svdw; d.l_vdwr = l_vdwr; d.l_heavy = l_heavy;
    std::vector<uint64_t> key = {1, ptr(phi), (uint64_t)nx, (uint64_t)ny, (uint64_t)nz,
                                 bits(ox), bits(oy), bits(oz), bits(sp), ptr(r_coords), ptr(r_ele),
                                 ptr(r_svdw), ptr(r_vdwr), ptr(r_heavy), (uint64_t)nr, ptr(l_base),
                                 ptr(l_ele), ptr(l_svdw), ptr(l_vdwr), ptr(l_heavy), (uint64_t)nl};
    lkl_ctx* c = cache_get(key, d);
    if (!c) return -1;
    return score_host(c, poses, N, flags, out, nullptr, nullptr, nullptr);
}

LKL_API int lkl_cpydock_solv(const float* r_coords, const unsigned char* r_flag, int nr,
                             const int* cell_start, const int* cell_atoms, int ncx, int ncy,
                             int ncz, float c_ox, float c_oy, float c_oz, float c_sp,
                             const float* l_base, const double* poses,
                             const unsigned char* l_flag, const float* l_ele, int nl, int N,
                             const float* r_des, const float* r_asa, const float* l_des,
                             const float* l_asa, double* outS) {
    (void)cell_start; (void)cell_atoms; (void)ncx; (void)ncy; (void)ncz;
    (void)c_ox; (void)c_oy; (void)c_oz; (void)c_sp; (void)l_ele;
    std::lock_guard<std::mutex> lk(g_mu);
    if (N <= 0 || nl <= 0 || nr <= 0) { set_err("lkl_cpydock_solv: need N, nl, nr > 0 (N=%d nl=%d nr=%d)", N, nl, nr); return -1; }
    if (!r_flag || !l_flag || !r_des || !r_asa || !l_des || !l_asa || !outS) { set_err("lkl_cpydock_solv: null desolvation array"); return -1; }
    lkl_complex_desc d;
    memset(&d, 0, sizeof d);
    d.nr = nr; d.r_xyz = r_coords; d.nl = nl; d.l_xyz = l_base;
    d.r_flag = r_flag; d.r_des = r_des; d.r_asa = r_asa;
    d.l_flag = l_flag; d.l_des = l_des; d.l_asa = l_asa;
    std::vector<uint64_t> key = {2, ptr(r_coords), ptr(r_flag), (uint64_t)nr, ptr(l_base),
                                 ptr(l_flag), (uint64_t)nl, ptr(r_des), ptr(r_asa), ptr(l_des), ptr(l_asa)};
    lkl_ctx* c = cache_get(key, d);
    if (!c) return -1;
    std::vector<double> ev((size_t)N * 2);
    return score_host(c, poses, N, F_DESOLV, ev.data(), outS, nullptr, nullptr);
}

// Single pre-transformed pose: the identity pose reproduces the f32 input
// coordinates exactly through the f64 transform.
static const double k_identity[7] = {1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};

LKL_API int lkl_full_score(const float* phi, int nx, int ny, int nz, float ox, float oy, float oz,
                           float sp, const float* r_coords, const float* r_ele,
                           const float* r_svdw, const float* r_vdwr, const unsigned char* r_heavy,
                           int nr, const int* cell_start, const int* cell_atoms, int ncx, int ncy,
                           int ncz, float c_ox, float c_oy, float c_oz, float c_sp,
                           const float* l_coords, const float* l_ele, const float* l_svdw,
                           const float* l_vdwr, const unsigned char* l_heavy, int nl,
                           float* out_elec, float* out_vdw) {
    (void)cell_start; (void)cell_atoms; (void)ncx; (void)ncy; (void)ncz;
    (void)c_ox; (void)c_oy; (void)c_oz; (void)c_sp;
    std::lock_guard<std::mutex> lk(g_mu);
    if (nl <= 0 || nr < 0 || !out_elec || !out_vdw) { set_err("lkl_full_score: bad arguments (nl=%d nr=%d)", nl, nr); return -1; }
    lkl_complex_desc d;
    memset(&d, 0, sizeof d);
    d.phi = phi; d.nx = nx; d.ny = ny; d.nz = nz; d.ox = ox; d.oy = oy; d.oz = oz; d.sp = sp;
    d.nr = nr; d.r_xyz = r_coords; d.r_ele = r_ele; d.r_svdw = r_svdw; d.r_vdwr = r_vdwr; d.r_heavy = r_heavy;
    d.nl = nl; d.l_xyz = l_coords; d.l_ele = l_ele; d.l_svdw = l_svdw; d.l_vdwr = l_vdwr; d.l_heavy = l_heavy;
    // receptor + far field cached on pointer identity; the (pre-transformed)
    // ligand is replaced on every call, as the reference uploaded it per call
    // (full_score.cu:169-188)
    std::vector<uint64_t> key = {3, ptr(phi), (uint64_t)nx, (uint64_t)ny, (uint64_t)nz,
                                 bits(ox), bits(oy), bits(oz), bits(sp), ptr(r_coords), ptr(r_ele),
                                 ptr(r_svdw), ptr(r_vdwr), ptr(r_heavy), (uint64_t)nr};
    lkl_ctx* c = cache_get(key, d);
    if (!c) return -1;
    if (!c->has_phi) { set_err("lkl_full_score: a far-field grid is required"); return -1; }
    c->nl = nl;
    if (build_ligand(c, l_coords, l_ele, l_svdw, l_vdwr, l_heavy, nullptr, nullptr, nullptr)) return -1;
    double out2[2];
    return score_host(c, k_identity, 1, F_FAR | F_ELEC | F_CLASH, out2, nullptr, out_elec, out_vdw);
}

LKL_API int lkl_far_field(const float* phi, int nx, int ny, int nz, float ox, float oy, float oz,
                          float sp, const float* coords, const float* charges, int n,
                          double* result) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (n <= 0 || !coords || !charges || !result) { set_err("lkl_far_field: bad arguments (n=%d)", n); return -1; }
    lkl_complex_desc d;
    memset(&d, 0, sizeof d);
    d.phi = phi; d.nx = nx; d.ny = ny; d.nz = nz; d.ox = ox; d.oy = oy; d.oz = oz; d.sp = sp;
    d.nr = 0; d.nl = n; d.l_xyz = coords; d.l_ele = charges;
    std::vector<uint64_t> key = {4, ptr(phi), (uint64_t)nx, (uint64_t)ny, (uint64_t)nz,
                                 bits(ox), bits(oy), bits(oz), bits(sp)};
    lkl_ctx* c = cache_get(key, d);
    if (!c) return -1;
    if (!c->has_phi) { set_err("lkl_far_field: empty grid"); return -1; }
    c->nl = n;
    if (build_ligand(c, coords, charges, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr)) return -1;
    double out2[2];
    int rc = score_host(c, k_identity, 1, F_FAR, out2, nullptr, nullptr, nullptr);
    if (rc == 0) *result = out2[0];
    return rc;
}
