// lkl_surrogate.hip — learned surrogate of the exact pose score (gfx950).
//
// Features: per pose, 16 receptor type groups x 16 ligand type groups x 24
// distance bins (1.25 A wide, 0-30 A) of atom-pair counts, log1p-scaled and
// stored as bf16 (6,144 values, index (g_rec * 16 + g_lig) * 24 + bin).  The
// pair distance is defined in f32 from the f32-rounded transformed ligand atom
// (the coordinate the scoring kernel works with):
//   d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz)),  dx = f32(x) - rx, ...
// and the bin is the largest b with 1.5625 * b^2 <= d2 (thresholds exact in
// f32), counted when d2 < 900.  Counts are exact integers; log1p is evaluated
// in f64 and rounded once to bf16 (round to nearest even).
//
// Model: MLP 6144 -> 512 -> 512 -> 512 -> 1, ReLU, bf16 weights and
// activations, fp32 accumulation (v_mfma_f32_32x32x16_bf16), fp32 master
// weights + Adam state.  Training minimises 0.5 * mean((y - t)^2).  Forward,
// backward and Adam run in this file's kernels.
#include "lkl_device.hpp"
#include "../../include/lklight_gpu.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#define LKL_API extern "C" __attribute__((visibility("default")))

using namespace lkl;

void lkl_set_error_text(const char* msg);  // lkl_score.hip

// Host-to-device copy ordered on `st` and waited for.  A plain hipMemcpy from
// pageable memory may return before the data lands, and the kernels run on
// non-blocking streams that do not wait for the null stream.
static hipError_t h2d(void* dst, const void* src, size_t bytes, hipStream_t st) {
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
    return e != hipSuccess ? e : hipStreamSynchronize(st);
}

namespace {

constexpr int NG = 16, NB = 24, NF = NG * NG * NB;  // 6144 features
constexpr int D1 = 512;                              // hidden width
constexpr int LUT_N = 4096;                          // log1p table for small counts

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

void serr(const char* msg) { lkl_set_error_text(msg); }

// f32 -> bf16 bits, round to nearest even (finite inputs).
__host__ __device__ __forceinline__ unsigned short f2bf_bits(float x) {
    unsigned u;
    memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

// positive normal f64 -> bf16 bits, one rounding (nearest even).
__host__ __device__ __forceinline__ unsigned short d2bf_bits(double x) {
    unsigned long long u;
    memcpy(&u, &x, 8);
    unsigned long long e = (u >> 52) - 1023 + 127, keep = (u >> 45) & 0x7F;
    const unsigned long long rest = u & ((1ull << 45) - 1), half = 1ull << 44;
    if (rest > half || (rest == half && (keep & 1))) ++keep;
    if (keep == 128) { keep = 0; ++e; }
    return (unsigned short)((e << 7) | keep);
}

__device__ __forceinline__ bf16 to_bf16(float x) {
    const unsigned short b = f2bf_bits(x);
    bf16 r;
    memcpy(&r, &b, 2);
    return r;
}

inline float bits_as_float(int v) {
    float f;
    memcpy(&f, &v, 4);
    return f;
}

__device__ __forceinline__ unsigned short log1p_bits(unsigned c, const unsigned short* lut) {
    return c < (unsigned)LUT_N ? lut[c] : d2bf_bits(log1p((double)c));
}

// ── features ─────────────────────────────────────────────────────────────────
struct FeatDev {
    const float4* rec;          // x, y, z, bits(g_rec * 24 * 16)
    const float4* lig;          // reference x, y, z, bits(g_lig)
    const unsigned short* lut;  // bf16 bits of log1p(c), c < LUT_N
    int nr, nl;
};

// Largest b in [0, 23] with 1.5625 b^2 <= d2 (0 <= d2 < 900).  The estimate
// from the hardware square root is off by at most one bin; the exact f32
// threshold compares correct it (for b = 23 the upper threshold is 900 > d2).
__device__ __forceinline__ int dist_bin(float d2) {
    const int b = min((int)(__builtin_amdgcn_sqrtf(d2) * 0.8f), NB - 1);
    const float lo = 1.5625f * (float)(b * b), hi = 1.5625f * (float)((b + 1) * (b + 1));
    return b + (d2 >= hi ? 1 : 0) - (d2 < lo ? 1 : 0);
}

// One workgroup per pose.  Lanes own ligand atoms; the receptor is streamed in
// index order with uniform (scalar) loads; counts go to an LDS histogram laid
// out [g_rec][bin][g_lig] so lanes of one receptor atom spread over banks.
__global__ void __launch_bounds__(256)
feat_kernel(FeatDev F, const double* __restrict__ poses, int N, unsigned short* __restrict__ out,
            unsigned* __restrict__ counts) {
    __shared__ unsigned hist[NF];
    const int pose = blockIdx.x;
    if (pose >= N) return;
    const int tid = threadIdx.x, bs = blockDim.x;
    for (int i = tid; i < NF; i += bs) hist[i] = 0u;
    __syncthreads();
    const Mat M = pose_matrix(poses + (size_t)pose * 7);
    for (int j = tid; j < F.nl; j += bs) {
        const float4 lb = F.lig[j];
        double x, y, z;
        transform(M, lb.x, lb.y, lb.z, x, y, z);
        const float xf = (float)x, yf = (float)y, zf = (float)z;
        const int lg = __float_as_int(lb.w);
#pragma unroll 4
        for (int i = 0; i < F.nr; ++i) {
            const float4 r = F.rec[i];
            const float dx = xf - r.x, dy = yf - r.y, dz = zf - r.z;
            const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
            if (d2 < 900.0f) atomicAdd(&hist[__float_as_int(r.w) + dist_bin(d2) * NG + lg], 1u);
        }
    }
    __syncthreads();
    for (int f = tid; f < NF; f += bs) {
        // canonical index f = (g_rec * 16 + g_lig) * 24 + bin
        const int gr = f / (NG * NB), gl = (f / NB) % NG, b = f % NB;
        const unsigned c = hist[(gr * NB + b) * NG + gl];
        if (out) out[(size_t)pose * NF + f] = log1p_bits(c, F.lut);
        if (counts) counts[(size_t)pose * NF + f] = c;
    }
}

// ── grid features ─────────────────────────────────────────────────────────────
// The receptor's pair counts are tabulated once on a node grid (spacing h, a
// power of two; origin integral so node coordinates are exact f32):
//   C[node][g_rec][bin] = #{receptor atoms of g_rec: d2(node, r) < 900, bin}
// with the pair-distance definition above (the node in place of the ligand
// atom), stored as u8 (384 bytes per node).  A pose's grid features replace
// each ligand atom by its nearest node, ix = floor(fmaf(x - ox, 1/h, 0.5)) per
// axis (no contribution outside the grid, which extends 30 A + h beyond the
// receptor), and sum that node's counts into the atom's ligand group.
constexpr int ROW = NG * NB;  // 384 bytes per node

struct GridDev {
    const unsigned char* grid;
    float ox, oy, oz, inv_h;
    int nx, ny, nz;
};

// One thread per (node, receptor group): 24 bin counters in 16-bit halves of
// 12 registers (a group holds < 65,536 atoms), then 24 bytes out.
__global__ void __launch_bounds__(256)
grid_build_kernel(const float4* __restrict__ rec, const int* __restrict__ gstart, float ox, float oy,
                  float oz, float h, int nx, int ny, long long nodes, unsigned char* __restrict__ grid,
                  unsigned* __restrict__ overflow) {
    const long long gid = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long node = gid >> 4;
    const int gr = (int)(gid & 15);
    if (node >= nodes) return;
    const int ix = (int)(node % nx);
    const long long t = node / nx;
    const int iy = (int)(t % ny), iz = (int)(t / ny);
    const float px = ox + (float)ix * h, py = oy + (float)iy * h, pz = oz + (float)iz * h;
    unsigned acc[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = 0u;
    const int i1 = gstart[gr + 1];
    for (int i = gstart[gr]; i < i1; ++i) {
        const float4 r = rec[i];
        const float dx = px - r.x, dy = py - r.y, dz = pz - r.z;
        const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        if (d2 < 900.0f) {
            const int b = dist_bin(d2);
            const unsigned inc = 1u << ((b & 1) * 16);
#pragma unroll
            for (int k = 0; k < 12; ++k) acc[k] += (b >> 1) == k ? inc : 0u;
        }
    }
    unsigned* out = (unsigned*)(grid + node * ROW + gr * NB);
    bool over = false;
#pragma unroll
    for (int w = 0; w < 6; ++w) {
        unsigned word = 0u;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int b = 4 * w + q;
            unsigned c = (acc[b >> 1] >> ((b & 1) * 16)) & 0xFFFFu;
            if (c > 255u) { over = true; c = 255u; }
            word |= c << (8 * q);
        }
        out[w] = word;
    }
    if (over) atomicOr(overflow, 1u);
}

// One workgroup (4 waves) per pose.  Phase 1 maps every ligand atom (sorted by
// group) to its node; phase 2 runs 8 atom streams (half-waves): lane l of a
// stream reads 12 bytes (g_rec = l / 2, bins 12 * (l % 2) ..) of each atom's
// node row and sums them in 16-bit halves, flushing to the LDS histogram when
// the ligand group changes or after 256 atoms (255 * 256 < 65,536).
__global__ void __launch_bounds__(256)
feat_grid_kernel(GridDev G, const float4* __restrict__ lig, int nl, const unsigned short* __restrict__ lut,
                 const double* __restrict__ poses, int N, unsigned short* __restrict__ out,
                 unsigned* __restrict__ counts) {
    extern __shared__ int node_of[];  // nl entries
    __shared__ unsigned hist[NF];     // canonical order
    const int pose = (int)xcd_tile(blockIdx.x, (unsigned)N);
    const int tid = threadIdx.x;
    for (int i = tid; i < NF; i += 256) hist[i] = 0u;
    const Mat M = pose_matrix(poses + (size_t)pose * 7);
    for (int j = tid; j < nl; j += 256) {
        const float4 a = lig[j];
        double x, y, z;
        transform(M, a.x, a.y, a.z, x, y, z);
        const int ix = (int)floorf(__builtin_fmaf((float)x - G.ox, G.inv_h, 0.5f));
        const int iy = (int)floorf(__builtin_fmaf((float)y - G.oy, G.inv_h, 0.5f));
        const int iz = (int)floorf(__builtin_fmaf((float)z - G.oz, G.inv_h, 0.5f));
        const bool in = ix >= 0 && iy >= 0 && iz >= 0 && ix < G.nx && iy < G.ny && iz < G.nz;
        node_of[j] = in ? (iz * G.ny + iy) * G.nx + ix : -1;
    }
    __syncthreads();
    const int lane = tid & 63, stream = (tid >> 6) * 2 + (lane >> 5), sub = lane & 31;
    const int gr = sub >> 1, b0 = 12 * (sub & 1);
    unsigned lo[3] = {0u, 0u, 0u}, hi[3] = {0u, 0u, 0u};
    int cur = -1, cnt = 0;
    auto flush = [&]() {
        if (cur >= 0) {
            unsigned* h = hist + (gr * NG + cur) * NB + b0;
#pragma unroll
            for (int w = 0; w < 3; ++w) {
                const unsigned v0 = lo[w] & 0xFFFFu, v2 = lo[w] >> 16, v1 = hi[w] & 0xFFFFu, v3 = hi[w] >> 16;
                if (v0) atomicAdd(h + 4 * w + 0, v0);
                if (v1) atomicAdd(h + 4 * w + 1, v1);
                if (v2) atomicAdd(h + 4 * w + 2, v2);
                if (v3) atomicAdd(h + 4 * w + 3, v3);
                lo[w] = hi[w] = 0u;
            }
        }
        cnt = 0;
    };
    for (int j = stream; j < nl; j += 8) {
        const int g = __float_as_int(lig[j].w);
        if (g != cur || cnt == 256) {
            flush();
            cur = g;
        }
        const int nd = node_of[j];
        if (nd >= 0) {
            const unsigned* row = (const unsigned*)(G.grid + (size_t)nd * ROW + gr * NB + b0);
#pragma unroll
            for (int w = 0; w < 3; ++w) {
                const unsigned v = row[w];
                lo[w] += v & 0x00FF00FFu;
                hi[w] += (v >> 8) & 0x00FF00FFu;
            }
        }
        ++cnt;
    }
    flush();
    __syncthreads();
    for (int f = tid; f < NF; f += 256) {
        const unsigned c = hist[f];
        if (out) out[(size_t)pose * NF + f] = log1p_bits(c, lut);
        if (counts) counts[(size_t)pose * NF + f] = c;
    }
}

// ── screening-step kernels ────────────────────────────────────────────────────
// Counter-based draws: u = (splitmix64(key) >> 11) * 2^-53 with key derived
// from (seed, candidate index, draw index), so a candidate set depends only on
// its inputs (reproducible, no generator state on the device).
__host__ __device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double draw_u01(unsigned long long seed, long long i, int d) {
    return (double)(mix64(seed ^ mix64((unsigned long long)i * 8ull + (unsigned long long)d)) >> 11) *
           (1.0 / 9007199254740992.0);
}

// cand[i], i = g * C + c: c = 0 copies base[g]; c > 0 adds trans_radius * U(-1, 1)
// to each translation component and rot_sigma * U(-1, 1) to each quaternion
// component, then renormalises the quaternion.
__global__ void cand_kernel(const double* __restrict__ base, long long total, int C, double tr, double rs,
                            unsigned long long seed, double* __restrict__ cand) {
#pragma clang fp contract(off)
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= total) return;
    const double* b = base + (i / C) * 7;
    double* o = cand + i * 7;
    if (i % C == 0) {
        for (int k = 0; k < 7; ++k) o[k] = b[k];
        return;
    }
    double q[4];
    for (int k = 0; k < 4; ++k) q[k] = b[k] + rs * (2.0 * draw_u01(seed, i, k) - 1.0);
    const double nrm = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    for (int k = 0; k < 4; ++k) o[k] = q[k] / nrm;
    for (int k = 0; k < 3; ++k) o[4 + k] = b[4 + k] + tr * (2.0 * draw_u01(seed, i, 4 + k) - 1.0);
}

// Per glowworm g: the k candidates with the highest prediction (ties: lower
// index first), their indices in pick[g * k ..] and poses in sel.
constexpr int SEL_MAX = 8;
__global__ void select_kernel(const float* __restrict__ pred, const double* __restrict__ cand, int n, int C,
                              int k, int* __restrict__ pick, double* __restrict__ sel) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    int best[SEL_MAX];
    float val[SEL_MAX];
    int m = 0;
    for (int c = 0; c < C; ++c) {
        const float v = pred[(size_t)g * C + c];
        int pos = m;
        while (pos > 0 && !(val[pos - 1] >= v)) --pos;  // keep earlier index on ties
        if (pos >= k) continue;
        if (m < k) ++m;
        for (int q = m - 1; q > pos; --q) { val[q] = val[q - 1]; best[q] = best[q - 1]; }
        val[pos] = v;
        best[pos] = c;
    }
    for (int r = 0; r < k; ++r) {
        const int c = best[r < m ? r : m - 1];
        pick[(size_t)g * k + r] = c;
        const double* src = cand + ((size_t)g * C + c) * 7;
        double* dst = sel + ((size_t)g * k + r) * 7;
        for (int q = 0; q < 7; ++q) dst[q] = src[q];
    }
}

// ── GEMM: C[M x N] = sum_k A(m, k) * B(k, n), bf16 in, fp32 accumulate ────────
// A(m, k) = A[m * sam + k * sak], B(k, n) = B[k * sbk + n * sbn] (one of each
// pair of strides is 1).  M, N multiples of 128, K of TK (64).  Epilogues:
//   EPI_F32:       C32[m, n] = acc
//   EPI_BIAS_RELU: C16[m, n] = bf16(relu(acc + bias[n]))
//   EPI_MASK:      C32[m, n] = acc * (mask16[m, n] > 0), C16 = bf16 of the same
enum { EPI_F32 = 0, EPI_BIAS_RELU = 1, EPI_MASK = 2 };
#ifndef LKL_GEMM_TK
#define LKL_GEMM_TK 64
#endif
constexpr int TM = 128, TN = 128, TK = LKL_GEMM_TK, LDK = TK + 8;
constexpr int NCH = TM * TK / 8 / 256;  // 16-byte chunks per thread per operand tile

struct GemmArgs {
    const bf16* A;
    const bf16* B;
    long sam, sak, sbk, sbn;
    int M, N, K;
    float* C32;
    bf16* C16;
    long ldc;
    const float* bias;
    const bf16* mask;
};

// Global -> register staging of a 128 x TK tile of X(r, k) = X[r * sr + k * sk]
// (NCH 16-byte loads per thread), then register -> LDS lds[r][k] (k
// contiguous; a k-major source is transposed on the way in).
struct Stage {
    uint4 v[NCH];
};

__device__ __forceinline__ void stage_load(const bf16* __restrict__ X, long sr, long sk, int r0, int k0,
                                           Stage& st) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int i = 0; i < NCH; ++i) {
        const int c = tid + 256 * i;
        if (sk == 1) {
            const int r = c / (TK / 8), kc = (c % (TK / 8)) * 8;
            st.v[i] = *(const uint4*)(X + (long)(r0 + r) * sr + k0 + kc);
        } else {
            const int k = c >> 4, rc = (c & 15) * 8;
            st.v[i] = *(const uint4*)(X + (r0 + rc) + (long)(k0 + k) * sk);
        }
    }
}

__device__ __forceinline__ void stage_store(long sk, const Stage& st, bf16 (*lds)[LDK]) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int i = 0; i < NCH; ++i) {
        const int c = tid + 256 * i;
        if (sk == 1) {
            const int r = c / (TK / 8), kc = (c % (TK / 8)) * 8;
            *(uint4*)&lds[r][kc] = st.v[i];
        } else {
            const int k = c >> 4, rc = (c & 15) * 8;
            const bf16* e = (const bf16*)&st.v[i];
#pragma unroll
            for (int q = 0; q < 8; ++q) lds[rc + q][k] = e[q];
        }
    }
}

// 128 x 128 output tile per workgroup, 4 waves of 64 x 64 (2 x 2 MFMA tiles of
// 32 x 32), K in steps of TK (64).  Double-buffered: the next step's tiles are read
// from global memory into registers while the MFMAs consume the current LDS
// buffer, then stored into the other buffer; one barrier per step.  The 1-D
// grid is mapped so the N-tiles of one M-tile share an XCD (and its L2).
template <int EPI>
__global__ void __launch_bounds__(256) gemm_kernel(GemmArgs g) {
    __shared__ __attribute__((aligned(16))) bf16 As[2][TM][LDK];
    __shared__ __attribute__((aligned(16))) bf16 Bs[2][TN][LDK];
    const unsigned ntn = (unsigned)(g.N / TN), total = ntn * (unsigned)(g.M / TM);
    const unsigned L = xcd_tile(blockIdx.x, total);
    const int m0 = (int)(L / ntn) * TM, n0 = (int)(L % ntn) * TN;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = (w >> 1) * 64, wn = (w & 1) * 64;
    f32x16 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.0f;
    const int nk = g.K / TK;
    Stage sa, sb;
    stage_load(g.A, g.sam, g.sak, m0, 0, sa);
    stage_load(g.B, g.sbn, g.sbk, n0, 0, sb);  // B^T tile: rows n, k contiguous
    stage_store(g.sak, sa, As[0]);
    stage_store(g.sbk, sb, Bs[0]);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < nk) {
            stage_load(g.A, g.sam, g.sak, m0, (kt + 1) * TK, sa);
            stage_load(g.B, g.sbn, g.sbk, n0, (kt + 1) * TK, sb);
        }
#pragma unroll
        for (int kk = 0; kk < TK; kk += 16) {
            bf16x8 a[2], b[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                a[t] = *(const bf16x8*)&As[cur][wm + 32 * t + (lane & 31)][kk + 8 * (lane >> 5)];
                b[t] = *(const bf16x8*)&Bs[cur][wn + 32 * t + (lane & 31)][kk + 8 * (lane >> 5)];
            }
#pragma unroll
            for (int ti = 0; ti < 2; ++ti)
#pragma unroll
                for (int tj = 0; tj < 2; ++tj)
                    acc[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[ti], b[tj], acc[ti][tj], 0, 0, 0);
        }
        if (kt + 1 < nk) {  // the other buffer was last read before the previous barrier
            stage_store(g.sak, sa, As[cur ^ 1]);
            stage_store(g.sbk, sb, Bs[cur ^ 1]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = m0 + wm + 32 * ti + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                const int n = n0 + wn + 32 * tj + (lane & 31);
                const float v = acc[ti][tj][r];
                const long o = (long)m * g.ldc + n;
                if (EPI == EPI_F32) {
                    g.C32[o] = v;
                } else if (EPI == EPI_BIAS_RELU) {
                    g.C16[o] = to_bf16(fmaxf(v + g.bias[n], 0.0f));
                } else {
                    const float z = (float)g.mask[o] > 0.0f ? v : 0.0f;
                    g.C32[o] = z;
                    g.C16[o] = to_bf16(z);
                }
            }
}

// ── GEMM, K-contiguous operands: 256 x 128 tiles, split-K ─────────────────────
// C[M x N] = sum_k A[m * lda + k] * B[n * ldb + k] (both operands K-contiguous:
// the forward reads the transposed weights, the weight gradients transposed
// activations).  A workgroup of 4 waves owns a 256 x 128 tile, each wave
// 128 x 64 = 4 x 2 MFMA tiles of 32 x 32, so a k-step of 16 reads 6 operand
// fragments from LDS for 8 MFMAs (the 64 x 64 wave tile of gemm_kernel reads
// 4 for 4, which LDS bandwidth caps near half the MFMA rate).  K advances in
// steps of 32, double-buffered through registers as in gemm_kernel.
// Split-K: blockIdx.y selects a chunk of kc columns of K; with more than one
// chunk each writes its fp32 partial tile to P[z][M][N] and gemm_reduce sums
// the chunks in ascending z (a fixed order: deterministic) and applies the
// epilogue.  Rows past M (M a multiple of 128, not of 256) read row M - 1 and
// are never stored.
constexpr int BM2 = 256, BN2 = 128, BK2 = 32, LD2 = BK2 + 8;
constexpr int CH_A2 = BM2 * BK2 / 8 / 256, CH_B2 = BN2 * BK2 / 8 / 256;  // 16-byte chunks per thread
constexpr int EP_PITCH = 36;  // fp32 pitch of a wave's 32 x 32 epilogue block in LDS

struct Gemm2Args {
    const bf16* A;
    const bf16* B;
    long lda, ldb;
    int M, N, K, kc;  // kc: K columns per split chunk (a multiple of BK2)
    float* P;         // split partials [z][M][N], or NULL: epilogue in place
    float* C32;
    bf16* C16;
    long ldc;
    const float* bias;
    const bf16* mask;
};

// Epilogue of 16 consecutive columns n..n+15 of row m (acc values v[16]).
template <int EPI>
__device__ __forceinline__ void epi16(const Gemm2Args& g, int m, int n, const float* v) {
    const long o = (long)m * g.ldc + n;
    if (EPI == EPI_F32) {
        float4* d = (float4*)(g.C32 + o);
#pragma unroll
        for (int q = 0; q < 4; ++q) d[q] = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
    } else if (EPI == EPI_BIAS_RELU) {
        bf16x8 h[2];
#pragma unroll
        for (int q = 0; q < 16; ++q) h[q >> 3][q & 7] = to_bf16(fmaxf(v[q] + g.bias[n + q], 0.0f));
        *(bf16x8*)(g.C16 + o) = h[0];
        *(bf16x8*)(g.C16 + o + 8) = h[1];
    } else {
        const bf16x8 k0 = *(const bf16x8*)(g.mask + o), k1 = *(const bf16x8*)(g.mask + o + 8);
        float z[16];
        bf16x8 h[2];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const float mk = (float)(q < 8 ? k0[q] : k1[q - 8]);
            z[q] = mk > 0.0f ? v[q] : 0.0f;
            h[q >> 3][q & 7] = to_bf16(z[q]);
        }
        float4* d = (float4*)(g.C32 + o);
#pragma unroll
        for (int q = 0; q < 4; ++q) d[q] = make_float4(z[4 * q], z[4 * q + 1], z[4 * q + 2], z[4 * q + 3]);
        *(bf16x8*)(g.C16 + o) = h[0];
        *(bf16x8*)(g.C16 + o + 8) = h[1];
    }
}

template <int EPI>
__global__ void __launch_bounds__(256) gemm2_kernel(Gemm2Args g) {
    __shared__ __attribute__((aligned(16))) bf16 As[2][BM2][LD2];
    __shared__ __attribute__((aligned(16))) bf16 Bs[2][BN2][LD2];
    const unsigned ntn = (unsigned)(g.N / BN2), total = ntn * (unsigned)((g.M + BM2 - 1) / BM2);
    const unsigned L = xcd_tile(blockIdx.x, total);
    const int m0 = (int)(L / ntn) * BM2, n0 = (int)(L % ntn) * BN2;
    const int kbeg = (int)blockIdx.y * g.kc;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = (w >> 1) * 128, wn = (w & 1) * 64;
    f32x16 acc[4][2];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.0f;
    // staging: thread t moves 16 bytes of rows t / 4 + 64 i (A: i < 4, B:
    // i < 2) at column (t % 4) * 8 of each k-step; A rows past M read row M - 1
    // (named registers: an array here would be kept in scratch memory)
    static_assert(CH_A2 == 4 && CH_B2 == 2, "four A and two B chunks per thread");
    const int sr = tid >> 2, sq = (tid & 3) * 8;
    const bf16* pa0 = g.A + (long)min(m0 + sr, g.M - 1) * g.lda + sq;
    const bf16* pa1 = g.A + (long)min(m0 + sr + 64, g.M - 1) * g.lda + sq;
    const bf16* pa2 = g.A + (long)min(m0 + sr + 128, g.M - 1) * g.lda + sq;
    const bf16* pa3 = g.A + (long)min(m0 + sr + 192, g.M - 1) * g.lda + sq;
    const bf16* pb0 = g.B + (long)(n0 + sr) * g.ldb + sq;
    const bf16* pb1 = g.B + (long)(n0 + sr + 64) * g.ldb + sq;
    uint4 a0, a1, a2, a3, b0, b1;
#define G2_LOAD(k0)                            \
    do {                                       \
        a0 = *(const uint4*)(pa0 + (k0));      \
        a1 = *(const uint4*)(pa1 + (k0));      \
        a2 = *(const uint4*)(pa2 + (k0));      \
        a3 = *(const uint4*)(pa3 + (k0));      \
        b0 = *(const uint4*)(pb0 + (k0));      \
        b1 = *(const uint4*)(pb1 + (k0));      \
    } while (0)
#define G2_STORE(buf)                          \
    do {                                       \
        *(uint4*)&As[buf][sr][sq] = a0;        \
        *(uint4*)&As[buf][sr + 64][sq] = a1;   \
        *(uint4*)&As[buf][sr + 128][sq] = a2;  \
        *(uint4*)&As[buf][sr + 192][sq] = a3;  \
        *(uint4*)&Bs[buf][sr][sq] = b0;        \
        *(uint4*)&Bs[buf][sr + 64][sq] = b1;   \
    } while (0)
    const int nk = g.kc / BK2;
    G2_LOAD(kbeg);
    G2_STORE(0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < nk) G2_LOAD(kbeg + (kt + 1) * BK2);
#pragma unroll
        for (int kk = 0; kk < BK2; kk += 16) {
            bf16x8 a[4], b[2];
#pragma unroll
            for (int t = 0; t < 4; ++t) a[t] = *(const bf16x8*)&As[cur][wm + 32 * t + (lane & 31)][kk + 8 * (lane >> 5)];
#pragma unroll
            for (int t = 0; t < 2; ++t) b[t] = *(const bf16x8*)&Bs[cur][wn + 32 * t + (lane & 31)][kk + 8 * (lane >> 5)];
#pragma unroll
            for (int ti = 0; ti < 4; ++ti)
#pragma unroll
                for (int tj = 0; tj < 2; ++tj)
                    acc[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[ti], b[tj], acc[ti][tj], 0, 0, 0);
        }
        if (kt + 1 < nk) G2_STORE(cur ^ 1);  // the other buffer was last read before the previous barrier
        __syncthreads();
    }
#undef G2_LOAD
#undef G2_STORE
    // Epilogue through LDS: each wave turns one 32 x 32 block at a time from
    // the MFMA layout (lane: column lane % 32, rows 8 (r / 4) + 4 (lane / 32) +
    // r % 4) into rows, so a lane stores 16 consecutive columns of one row.
    float* eb = reinterpret_cast<float*>(&As[0][0][0]) + w * 32 * EP_PITCH;
    const int er = lane >> 1, ec = (lane & 1) * 16;
#pragma unroll
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj) {
#pragma unroll
            for (int r = 0; r < 16; ++r)
                eb[(8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)) * EP_PITCH + (lane & 31)] = acc[ti][tj][r];
            __syncthreads();
            float v[16];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float4 t4 = *(const float4*)&eb[er * EP_PITCH + ec + 4 * q];
                v[4 * q] = t4.x;
                v[4 * q + 1] = t4.y;
                v[4 * q + 2] = t4.z;
                v[4 * q + 3] = t4.w;
            }
            __syncthreads();
            const int m = m0 + wm + 32 * ti + er, n = n0 + wn + 32 * tj + ec;
            if (m < g.M) {
                if (g.P) {
                    float4* d = (float4*)(g.P + ((long)blockIdx.y * g.M + m) * g.N + n);
#pragma unroll
                    for (int q = 0; q < 4; ++q) d[q] = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
                } else {
                    epi16<EPI>(g, m, n, v);
                }
            }
        }
}

// Sum of the S split partials in ascending z, then the epilogue; one thread
// per 16 consecutive columns of a row.
template <int EPI>
__global__ void __launch_bounds__(256) gemm_reduce(Gemm2Args g, int S) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int per_row = g.N / 16;
    if (t >= (long)g.M * per_row) return;
    const int m = (int)(t / per_row), n = (int)(t % per_row) * 16;
    float v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = 0.0f;
    for (int z = 0; z < S; ++z) {
        const float4* p = (const float4*)(g.P + ((long)z * g.M + m) * g.N + n);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float4 x = p[q];
            v[4 * q] += x.x;
            v[4 * q + 1] += x.y;
            v[4 * q + 2] += x.z;
            v[4 * q + 3] += x.w;
        }
    }
    epi16<EPI>(g, m, n, v);
}

// Split-K factor of a K-contiguous GEMM: double it while the launch has fewer
// than 512 workgroups (two per CU) and each chunk keeps at least 128 columns.
static int split_of(int M, int N, int K) {
    const long tiles = (long)((M + BM2 - 1) / BM2) * (N / BN2);
    int S = 1;
    while (tiles * S < 512 && K % (2 * S * BK2) == 0 && K / (2 * S) >= 128) S *= 2;
    return S;
}

// fp32 workspace (floats) of the split partials of an M x N x K GEMM.
static size_t split_ws(int M, int N, int K) {
    const int S = split_of(M, N, K);
    return S > 1 ? (size_t)S * M * N : 0;
}

static bool k_contiguous(long sak, long sbk, int M, int N, int K) {
    return sak == 1 && sbk == 1 && M % 128 == 0 && N % BN2 == 0 && K % BK2 == 0;
}

// ── small kernels of the MLP ──────────────────────────────────────────────────
// y[b] = sum_k H[b, k] W4[k] + b4   (512 -> 1 layer)
__global__ void head_forward(const bf16* __restrict__ H, const bf16* __restrict__ w4, const float* b4,
                             int B, float* __restrict__ y) {
    const int b = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (b >= B) return;
    float s = 0.0f;
    for (int k = lane; k < D1; k += 64) s += (float)H[(size_t)b * D1 + k] * (float)w4[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) y[b] = s + b4[0];
}

// dy[b] = (y[b] - t[b]) / B; loss partial sums per block
__global__ void loss_grad(const float* __restrict__ y, const float* __restrict__ t, int B,
                          float* __restrict__ dy, float* __restrict__ loss) {
    __shared__ float red[256];
    float s = 0.0f;
    for (int b = blockIdx.x * 256 + threadIdx.x; b < B; b += gridDim.x * 256) {
        const float d = y[b] - t[b];
        dy[b] = d / (float)B;
        s += 0.5f * d * d / (float)B;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// dZ3[b, k] = dy[b] W4[k] (H3[b, k] > 0), elementwise.
__global__ void head_dz(const bf16* __restrict__ H, const bf16* __restrict__ w4,
                        const float* __restrict__ dy, long total, float* __restrict__ dz32,
                        bf16* __restrict__ dz16) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int k = (int)(i % D1);
        const float d = (float)H[i] > 0.0f ? dy[i / D1] * (float)w4[k] : 0.0f;
        dz32[i] = d;
        dz16[i] = to_bf16(d);
    }
}

// Column sums in a fixed order (deterministic): pass 1 sums each 256-row
// chunk c into part[c][n] (value A32[r, n], or A16[r, n] * w[r]); pass 2 adds
// the chunks in ascending order.
constexpr int CS_ROWS = 256;
__global__ void __launch_bounds__(256)
colsum_partial(const float* __restrict__ A32, const bf16* __restrict__ A16, const float* __restrict__ w,
               int B, int N, float* __restrict__ part) {
    __shared__ float red[4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int n = blockIdx.x * 64 + lane, c = blockIdx.y;
    const int r1 = min(B, (c + 1) * CS_ROWS);
    float s = 0.0f;
    if (n < N)
        for (int r = c * CS_ROWS + wv; r < r1; r += 4)
            s += A32 ? A32[(size_t)r * N + n] : (float)A16[(size_t)r * N + n] * w[r];
    red[wv][lane] = s;
    __syncthreads();
    if (wv == 0 && n < N) part[(size_t)c * N + n] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}

__global__ void colsum_final(const float* __restrict__ part, int nchunk, int N, float* __restrict__ out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.0f;
    for (int c = 0; c < nchunk; ++c) s += part[(size_t)c * N + n];
    out[n] = s;
}

// Adam (bias-corrected), then refresh the bf16 working copy of the weights.
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, bf16* __restrict__ p16, long n, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        p[i] = pi;
        p16[i] = to_bf16(pi);
    }
}

// dst[c][r] = src[r][c] (bf16, rows x cols, both multiples of 64), through a
// 64 x 64 LDS tile so reads and writes are coalesced.
__global__ void __launch_bounds__(256) transpose_kernel(const bf16* __restrict__ src, int rows, int cols,
                                                        bf16* __restrict__ dst) {
    __shared__ bf16 t[64][66];
    const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
    for (int i = threadIdx.x; i < 64 * 64; i += 256) {
        const int r = i >> 6, c = i & 63;
        t[r][c] = src[(size_t)(r0 + r) * cols + c0 + c];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 64 * 64; i += 256) {
        const int c = i >> 6, r = i & 63;
        dst[(size_t)(c0 + c) * rows + r0 + r] = t[r][c];
    }
}

__global__ void to_bf16_kernel(const float* __restrict__ p, bf16* __restrict__ p16, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        p16[i] = to_bf16(p[i]);
}

// Parameter layout (flat fp32): W1 [6144 x 512] (row k: input feature), b1,
// W2 [512 x 512], b2, W3 [512 x 512], b3, W4 [512], b4.
constexpr long OW1 = 0, OB1 = OW1 + (long)NF * D1, OW2 = OB1 + D1, OB2 = OW2 + (long)D1 * D1,
               OW3 = OB2 + D1, OB3 = OW3 + (long)D1 * D1, OW4 = OB3 + D1, OB4 = OW4 + D1,
               NPARAM = OB4 + 1;

}  // namespace

// ── feature object ────────────────────────────────────────────────────────────
struct lkl_feat {
    int device = 0, nr = 0, nl = 0;
    float4 *d_rec = nullptr, *d_lig = nullptr;  // grid mode: ligand sorted by group
    unsigned short* d_lut = nullptr;
    hipStream_t stream = nullptr;
    bool grid_mode = false;
    unsigned char* d_grid = nullptr;
    double origin[3] = {0, 0, 0}, h = 0;
    int dims[3] = {0, 0, 0};
    long long nodes = 0;
};

LKL_API void lkl_feat_destroy(lkl_feat* f) {
    if (!f) return;
    if (f->d_rec) (void)hipFree(f->d_rec);
    if (f->d_lig) (void)hipFree(f->d_lig);
    if (f->d_lut) (void)hipFree(f->d_lut);
    if (f->d_grid) (void)hipFree(f->d_grid);
    if (f->stream) (void)hipStreamDestroy(f->stream);
    delete f;
}

static std::vector<unsigned short> log1p_table() {
    std::vector<unsigned short> lut(LUT_N);
    lut[0] = 0;
    for (int c = 1; c < LUT_N; ++c) lut[c] = d2bf_bits(std::log1p((double)c));
    return lut;
}

constexpr int GRID_MAX_NL = 16384;  // node indices of one pose live in LDS

// Grid mode (see grid_build_kernel): receptor counts tabulated on nodes of
// spacing h (0.25, 0.5, 1 or 2 A) over the receptor box grown by 30 A + h.
// Fails if a node's count for one (group, bin) exceeds 255.
LKL_API lkl_feat* lkl_feat_create_grid(const float* r_xyz, const unsigned char* r_group, int nr,
                                       const float* l_xyz, const unsigned char* l_group, int nl,
                                       double spacing) {
    if (!r_xyz || !r_group || !l_xyz || !l_group || nr <= 0 || nl <= 0 || nr >= 65536 || nl > GRID_MAX_NL ||
        !(spacing == 0.25 || spacing == 0.5 || spacing == 1.0 || spacing == 2.0)) {
        serr("lkl_feat_create_grid: bad arguments (need 0 < nr < 65536, 0 < nl <= 16384, spacing in {0.25, 0.5, 1, 2})");
        return nullptr;
    }
    if (!lkl_device_available()) {
        serr("lkl_feat_create_grid: no gfx950 device");
        return nullptr;
    }
    lkl_feat* f = new lkl_feat();
    f->grid_mode = true;
    f->nr = nr;
    f->nl = nl;
    f->h = spacing;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = 0; i < nr; ++i)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::fmin(lo[a], (double)r_xyz[3 * i + a]);
            hi[a] = std::fmax(hi[a], (double)r_xyz[3 * i + a]);
        }
    f->nodes = 1;
    for (int a = 0; a < 3; ++a) {
        f->origin[a] = std::floor(lo[a] - 30.0 - spacing);
        const double top = std::ceil(hi[a] + 30.0 + spacing);
        f->dims[a] = (int)((top - f->origin[a]) / spacing) + 1;
        f->nodes *= f->dims[a];
    }
    // receptor by group (stable), group starts; ligand by group (stable)
    std::vector<float4> r;
    std::vector<int> gstart(NG + 1, 0);
    r.reserve(nr);
    for (int g = 0; g < NG; ++g) {
        gstart[g] = (int)r.size();
        for (int i = 0; i < nr; ++i)
            if (r_group[i] % NG == g) r.push_back(make_float4(r_xyz[3 * i], r_xyz[3 * i + 1], r_xyz[3 * i + 2], 0.f));
    }
    gstart[NG] = nr;
    std::vector<float4> l;
    l.reserve(nl);
    for (int g = 0; g < NG; ++g)
        for (int j = 0; j < nl; ++j)
            if (l_group[j] % NG == g)
                l.push_back(make_float4(l_xyz[3 * j], l_xyz[3 * j + 1], l_xyz[3 * j + 2], bits_as_float(g)));
    const std::vector<unsigned short> lut = log1p_table();
    int* d_gs = nullptr;
    unsigned* d_over = nullptr;
    unsigned over = 0;
    const size_t gbytes = (size_t)f->nodes * ROW;
    bool ok = hipGetDevice(&f->device) == hipSuccess &&
              hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking) == hipSuccess &&
              hipMalloc((void**)&f->d_rec, sizeof(float4) * nr) == hipSuccess &&
              hipMalloc((void**)&f->d_lig, sizeof(float4) * nl) == hipSuccess &&
              hipMalloc((void**)&f->d_lut, 2 * LUT_N) == hipSuccess &&
              hipMalloc((void**)&f->d_grid, gbytes) == hipSuccess &&
              hipMalloc((void**)&d_gs, sizeof(int) * (NG + 1)) == hipSuccess &&
              hipMalloc((void**)&d_over, 4) == hipSuccess &&
              h2d(f->d_rec, r.data(), sizeof(float4) * nr, f->stream) == hipSuccess &&
              h2d(f->d_lig, l.data(), sizeof(float4) * nl, f->stream) == hipSuccess &&
              h2d(f->d_lut, lut.data(), 2 * LUT_N, f->stream) == hipSuccess &&
              h2d(d_gs, gstart.data(), sizeof(int) * (NG + 1), f->stream) == hipSuccess &&
              h2d(d_over, &over, 4, f->stream) == hipSuccess;
    if (ok) {
        const long long threads = f->nodes * 16;
        hipLaunchKernelGGL(grid_build_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, f->stream,
                           f->d_rec, d_gs, (float)f->origin[0], (float)f->origin[1], (float)f->origin[2],
                           (float)spacing, f->dims[0], f->dims[1], f->nodes, f->d_grid, d_over);
        ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(f->stream) == hipSuccess &&
             hipMemcpy(&over, d_over, 4, hipMemcpyDeviceToHost) == hipSuccess;
    }
    if (d_gs) (void)hipFree(d_gs);
    if (d_over) (void)hipFree(d_over);
    if (!ok || over) {
        serr(over ? "lkl_feat_create_grid: a node count exceeds 255 (too many atoms of one group in one shell)"
                  : "lkl_feat_create_grid: device setup failed");
        lkl_feat_destroy(f);
        return nullptr;
    }
    return f;
}

// Grid box of a grid-mode object: origin, node counts, spacing; returns the
// number of nodes (0 for an exact-mode object, -1 on error).
LKL_API long long lkl_feat_grid_info(const lkl_feat* f, double* origin3, int* dims3, double* spacing) {
    if (!f) return -1;
    if (!f->grid_mode) return 0;
    for (int a = 0; a < 3; ++a) {
        if (origin3) origin3[a] = f->origin[a];
        if (dims3) dims3[a] = f->dims[a];
    }
    if (spacing) *spacing = f->h;
    return f->nodes;
}

// Copy nodes [first, first + count) of the count grid (384 bytes each, order
// g_rec * 24 + bin) to host memory.
LKL_API int lkl_feat_grid_read(const lkl_feat* f, long long first, long long count, unsigned char* out) {
    if (!f || !f->grid_mode || !out || first < 0 || count < 0 || first + count > f->nodes) {
        serr("lkl_feat_grid_read: bad arguments");
        return -1;
    }
    return hipMemcpy(out, f->d_grid + (size_t)first * ROW, (size_t)count * ROW, hipMemcpyDeviceToHost) == hipSuccess
               ? 0
               : -1;
}

// Receptor / ligand coordinates (f32, n x 3; the ligand in its reference
// frame) and type groups (taken modulo 16).
LKL_API lkl_feat* lkl_feat_create(const float* r_xyz, const unsigned char* r_group, int nr,
                                  const float* l_xyz, const unsigned char* l_group, int nl) {
    if (!r_xyz || !r_group || !l_xyz || !l_group || nr <= 0 || nl <= 0) {
        serr("lkl_feat_create: bad arguments");
        return nullptr;
    }
    if (!lkl_device_available()) {
        serr("lkl_feat_create: no gfx950 device");
        return nullptr;
    }
    lkl_feat* f = new lkl_feat();
    f->nr = nr;
    f->nl = nl;
    std::vector<float4> r(nr), l(nl);
    for (int i = 0; i < nr; ++i)
        r[i] = make_float4(r_xyz[3 * i], r_xyz[3 * i + 1], r_xyz[3 * i + 2],
                           bits_as_float((r_group[i] % NG) * NB * NG));
    for (int j = 0; j < nl; ++j)
        l[j] = make_float4(l_xyz[3 * j], l_xyz[3 * j + 1], l_xyz[3 * j + 2], bits_as_float(l_group[j] % NG));
    const std::vector<unsigned short> lut = log1p_table();
    if (hipGetDevice(&f->device) != hipSuccess ||
        hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void**)&f->d_rec, sizeof(float4) * nr) != hipSuccess ||
        hipMalloc((void**)&f->d_lig, sizeof(float4) * nl) != hipSuccess ||
        hipMalloc((void**)&f->d_lut, 2 * LUT_N) != hipSuccess ||
        h2d(f->d_rec, r.data(), sizeof(float4) * nr, f->stream) != hipSuccess ||
        h2d(f->d_lig, l.data(), sizeof(float4) * nl, f->stream) != hipSuccess ||
        h2d(f->d_lut, lut.data(), 2 * LUT_N, f->stream) != hipSuccess) {
        serr("lkl_feat_create: device setup failed");
        lkl_feat_destroy(f);
        return nullptr;
    }
    return f;
}

// Device buffers: poses N x 7 (w, x, y, z, tx, ty, tz) -> out (N x 6144 bf16
// bits, may be NULL) and/or raw counts (N x 6144 u32, may be NULL), enqueued
// on `stream` (NULL = the object's own stream).
LKL_API int lkl_feat_compute(lkl_feat* f, const double* d_poses, int N, unsigned short* d_out,
                             unsigned* d_counts, void* stream) {
    if (!f || !d_poses || N < 0 || (!d_out && !d_counts)) {
        serr("lkl_feat_compute: bad arguments");
        return -1;
    }
    if (N == 0) return 0;
    hipStream_t st = stream ? (hipStream_t)stream : f->stream;
    if (f->grid_mode) {
        GridDev G = {f->d_grid, (float)f->origin[0], (float)f->origin[1], (float)f->origin[2],
                     (float)(1.0 / f->h), f->dims[0], f->dims[1], f->dims[2]};
        hipLaunchKernelGGL(feat_grid_kernel, dim3((unsigned)N), dim3(256), sizeof(int) * f->nl, st, G,
                           f->d_lig, f->nl, f->d_lut, d_poses, N, d_out, d_counts);
    } else {
        FeatDev F = {f->d_rec, f->d_lig, f->d_lut, f->nr, f->nl};
        hipLaunchKernelGGL(feat_kernel, dim3((unsigned)N), dim3(256), 0, st, F, d_poses, N, d_out, d_counts);
    }
    if (hipGetLastError() != hipSuccess) {
        serr("lkl_feat_compute: launch failed");
        return -1;
    }
    return 0;
}

// Host buffers variant (blocking); either output may be NULL.
LKL_API int lkl_feat_compute_host(lkl_feat* f, const double* poses, int N, unsigned short* out,
                                  unsigned* counts) {
    if (!f || !poses || N <= 0 || (!out && !counts)) {
        serr("lkl_feat_compute_host: bad arguments");
        return -1;
    }
    double* dp = nullptr;
    unsigned short* dout = nullptr;
    unsigned* dc = nullptr;
    int rc = 0;
    if (hipMalloc((void**)&dp, sizeof(double) * 7 * N) != hipSuccess) rc = -1;
    if (!rc && out && hipMalloc((void**)&dout, 2 * (size_t)N * NF) != hipSuccess) rc = -1;
    if (!rc && counts && hipMalloc((void**)&dc, 4 * (size_t)N * NF) != hipSuccess) rc = -1;
    if (!rc && h2d(dp, poses, sizeof(double) * 7 * N, f->stream) != hipSuccess) rc = -1;
    if (!rc) rc = lkl_feat_compute(f, dp, N, dout, dc, f->stream);
    if (!rc && hipStreamSynchronize(f->stream) != hipSuccess) rc = -1;
    if (!rc && out && hipMemcpy(out, dout, 2 * (size_t)N * NF, hipMemcpyDeviceToHost) != hipSuccess) rc = -1;
    if (!rc && counts && hipMemcpy(counts, dc, 4 * (size_t)N * NF, hipMemcpyDeviceToHost) != hipSuccess) rc = -1;
    if (dp) (void)hipFree(dp);
    if (dout) (void)hipFree(dout);
    if (dc) (void)hipFree(dc);
    if (rc) serr("lkl_feat_compute_host failed");
    return rc;
}

// ── MLP object ────────────────────────────────────────────────────────────────
struct lkl_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    bf16* p16 = nullptr;   // bf16 working copy, parameter layout
    bf16* p16t = nullptr;  // W1, W2, W3 transposed ([out][in]): the forward GEMMs' K-contiguous operand
    long step = 0;
    size_t cap = 0;  // batch capacity of the activation buffers
    bf16 *h1 = nullptr, *h2 = nullptr, *h3 = nullptr, *dza = nullptr, *dzb = nullptr;  // dz ping-pong
    float *y = nullptr, *dy = nullptr, *dz32 = nullptr, *part = nullptr, *loss = nullptr;
    bf16 *xt = nullptr, *ht = nullptr, *dzt = nullptr;  // transposed operands of the weight-gradient GEMMs
    float* ws = nullptr;  // split-K partials of the GEMMs
    size_t ws_cap = 0;    // floats
};

// Offsets of the transposed W1, W2, W3 in p16t.
constexpr long T1 = 0, T2 = T1 + (long)NF * D1, T3 = T2 + (long)D1 * D1, NT_W = T3 + (long)D1 * D1;

static void refresh_transposed(lkl_mlp* n, hipStream_t st) {
    hipLaunchKernelGGL(transpose_kernel, dim3(D1 / 64, NF / 64), dim3(256), 0, st, n->p16 + OW1, NF, D1, n->p16t + T1);
    hipLaunchKernelGGL(transpose_kernel, dim3(D1 / 64, D1 / 64), dim3(256), 0, st, n->p16 + OW2, D1, D1, n->p16t + T2);
    hipLaunchKernelGGL(transpose_kernel, dim3(D1 / 64, D1 / 64), dim3(256), 0, st, n->p16 + OW3, D1, D1, n->p16t + T3);
}

static void mlp_free_acts(lkl_mlp* n) {
    void* b[] = {n->h1, n->h2, n->h3, n->dza, n->dzb, n->y, n->dy, n->dz32, n->part, n->xt, n->ht, n->dzt};
    for (void* x : b)
        if (x) (void)hipFree(x);
    n->h1 = n->h2 = n->h3 = n->dza = n->dzb = nullptr;
    n->xt = n->ht = n->dzt = nullptr;
    n->y = n->dy = n->dz32 = n->part = nullptr;
    n->cap = 0;
}

static int mlp_reserve(lkl_mlp* n, size_t B) {
    if (B <= n->cap) return 0;
    mlp_free_acts(n);
    const size_t h = B * D1;
    if (hipMalloc((void**)&n->h1, h * 2) || hipMalloc((void**)&n->h2, h * 2) ||
        hipMalloc((void**)&n->h3, h * 2) || hipMalloc((void**)&n->dza, h * 2) ||
        hipMalloc((void**)&n->dzb, h * 2) ||
        hipMalloc((void**)&n->y, B * 4) || hipMalloc((void**)&n->dy, B * 4) ||
        hipMalloc((void**)&n->dz32, h * 4) ||
        hipMalloc((void**)&n->part, ((B + CS_ROWS - 1) / CS_ROWS) * (size_t)D1 * 4) ||
        hipMalloc((void**)&n->xt, B * (size_t)NF * 2) || hipMalloc((void**)&n->ht, h * 2) ||
        hipMalloc((void**)&n->dzt, h * 2)) {
        mlp_free_acts(n);
        serr("lkl_mlp: activation allocation failed");
        return -1;
    }
    n->cap = B;
    return 0;
}

// Split-K workspace for the GEMMs of a batch of B rows (grown on demand; the
// old buffer is released only after the device is idle).
static int mlp_ws(lkl_mlp* n, int B) {
    const size_t need = std::max(std::max(split_ws(B, D1, NF), split_ws(B, D1, D1)),
                                 std::max(split_ws(NF, D1, B), split_ws(D1, D1, B)));
    if (need <= n->ws_cap) return 0;
    if (n->ws) {
        (void)hipDeviceSynchronize();
        (void)hipFree(n->ws);
    }
    n->ws = nullptr;
    n->ws_cap = 0;
    if (hipMalloc((void**)&n->ws, need * 4) != hipSuccess) {
        n->ws = nullptr;
        serr("lkl_mlp: split-K workspace allocation failed");
        return -1;
    }
    n->ws_cap = need;
    return 0;
}

LKL_API long long lkl_mlp_param_count(void) { return NPARAM; }

LKL_API void lkl_mlp_destroy(lkl_mlp* n) {
    if (!n) return;
    mlp_free_acts(n);
    void* b[] = {n->p, n->g, n->m, n->v, n->p16, n->p16t, n->loss, n->ws};
    for (void* x : b)
        if (x) (void)hipFree(x);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
}

// Weights: He-uniform (limit sqrt(6 / fan_in)) from a splitmix64 stream; biases 0.
LKL_API lkl_mlp* lkl_mlp_create(unsigned long long seed) {
    if (!lkl_device_available()) {
        serr("lkl_mlp_create: no gfx950 device");
        return nullptr;
    }
    lkl_mlp* n = new lkl_mlp();
    std::vector<float> p(NPARAM, 0.0f);
    uint64_t s = seed;
    auto next = [&]() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](long off, long cnt, int fan_in) {
        const double lim = std::sqrt(6.0 / fan_in);
        for (long i = 0; i < cnt; ++i) p[off + i] = (float)((2.0 * next() - 1.0) * lim);
    };
    fill(OW1, (long)NF * D1, NF);
    fill(OW2, (long)D1 * D1, D1);
    fill(OW3, (long)D1 * D1, D1);
    fill(OW4, D1, D1);
    if (hipGetDevice(&n->device) != hipSuccess || hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void**)&n->p, NPARAM * 4) || hipMalloc((void**)&n->g, NPARAM * 4) ||
        hipMalloc((void**)&n->m, NPARAM * 4) || hipMalloc((void**)&n->v, NPARAM * 4) ||
        hipMalloc((void**)&n->p16, NPARAM * 2) || hipMalloc((void**)&n->p16t, NT_W * 2) ||
        hipMalloc((void**)&n->loss, 4) ||
        h2d(n->p, p.data(), NPARAM * 4, n->stream) ||
        hipMemsetAsync(n->m, 0, NPARAM * 4, n->stream) || hipMemsetAsync(n->v, 0, NPARAM * 4, n->stream) ||
        hipMemsetAsync(n->g, 0, NPARAM * 4, n->stream)) {
        serr("lkl_mlp_create: device setup failed");
        lkl_mlp_destroy(n);
        return nullptr;
    }
    hipLaunchKernelGGL(to_bf16_kernel, dim3(1024), dim3(256), 0, n->stream, n->p, n->p16, NPARAM);
    refresh_transposed(n, n->stream);
    if (hipStreamSynchronize(n->stream) != hipSuccess) {
        serr("lkl_mlp_create: init failed");
        lkl_mlp_destroy(n);
        return nullptr;
    }
    return n;
}

// The object's stream is non-blocking: copies of its buffers first wait for
// the work queued on it (an Adam step may still be running).
LKL_API int lkl_mlp_get_params(lkl_mlp* n, float* out) {
    if (!n || !out) return -1;
    if (hipStreamSynchronize(n->stream) != hipSuccess) return -1;
    return hipMemcpy(out, n->p, NPARAM * 4, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
}

LKL_API int lkl_mlp_set_params(lkl_mlp* n, const float* in) {
    if (!n || !in) return -1;
    if (hipStreamSynchronize(n->stream) != hipSuccess) return -1;
    if (h2d(n->p, in, NPARAM * 4, n->stream) != hipSuccess) return -1;
    hipLaunchKernelGGL(to_bf16_kernel, dim3(1024), dim3(256), 0, n->stream, n->p, n->p16, NPARAM);
    refresh_transposed(n, n->stream);
    return hipStreamSynchronize(n->stream) == hipSuccess ? 0 : -1;
}

LKL_API int lkl_mlp_get_grads(lkl_mlp* n, float* out) {
    if (!n || !out) return -1;
    if (hipStreamSynchronize(n->stream) != hipSuccess) return -1;
    return hipMemcpy(out, n->g, NPARAM * 4, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
}

// C = op(A) op(B) with an epilogue.  K-contiguous operands take gemm2_kernel
// (split over K when the tile count is small; ws holds the partials),
// other strides the general gemm_kernel.
template <int EPI>
static void gemm(hipStream_t st, const bf16* A, long sam, long sak, const bf16* B, long sbk, long sbn,
                 int M, int N, int K, float* C32, bf16* C16, long ldc, const float* bias,
                 const bf16* mask, float* ws = nullptr) {
    if (k_contiguous(sak, sbk, M, N, K)) {
        const int S = split_of(M, N, K);
        Gemm2Args a = {A, B, sam, sbn, M, N, K, K / S, S > 1 ? ws : nullptr, C32, C16, ldc, bias, mask};
        const unsigned tiles = (unsigned)(((M + BM2 - 1) / BM2) * (N / BN2));
        hipLaunchKernelGGL(gemm2_kernel<EPI>, dim3(tiles, (unsigned)S), dim3(256), 0, st, a);
        if (S > 1) {
            const long thr = (long)M * (N / 16);
            hipLaunchKernelGGL(gemm_reduce<EPI>, dim3((unsigned)((thr + 255) / 256)), dim3(256), 0, st, a, S);
        }
        return;
    }
    GemmArgs a = {A, B, sam, sak, sbk, sbn, M, N, K, C32, C16, ldc, bias, mask};
    hipLaunchKernelGGL(gemm_kernel<EPI>, dim3((unsigned)((N / TN) * (M / TM))), dim3(256), 0, st, a);
}

static int check_batch(int B) {
    if (B <= 0 || B % 128) { serr("lkl_mlp: batch size must be a positive multiple of 128"); return -1; }
    return 0;
}

// Forward on device buffers: x (B x 6144 bf16) -> y (B fp32).  B % 128 == 0.
LKL_API int lkl_mlp_forward(lkl_mlp* n, const void* d_x, int B, float* d_y, void* stream) {
    if (!n || !d_x || !d_y || check_batch(B) || mlp_reserve(n, B) || mlp_ws(n, B)) return -1;
    hipStream_t st = stream ? (hipStream_t)stream : n->stream;
    const bf16* X = (const bf16*)d_x;
    const bf16* W = n->p16;
    const bf16* Wt = n->p16t;  // B(k, n) = Wt[n][k]: both operands K-contiguous
    gemm<EPI_BIAS_RELU>(st, X, NF, 1, Wt + T1, 1, NF, B, D1, NF, nullptr, n->h1, D1, n->p + OB1, nullptr, n->ws);
    gemm<EPI_BIAS_RELU>(st, n->h1, D1, 1, Wt + T2, 1, D1, B, D1, D1, nullptr, n->h2, D1, n->p + OB2, nullptr, n->ws);
    gemm<EPI_BIAS_RELU>(st, n->h2, D1, 1, Wt + T3, 1, D1, B, D1, D1, nullptr, n->h3, D1, n->p + OB3, nullptr, n->ws);
    hipLaunchKernelGGL(head_forward, dim3((B + 3) / 4), dim3(256), 0, st, n->h3, W + OW4, n->p + OB4, B, d_y);
    if (hipGetLastError() != hipSuccess) { serr("lkl_mlp_forward: launch failed"); return -1; }
    return 0;
}

static void colsum(hipStream_t st, const float* A32, const bf16* A16, const float* w, int B, int N,
                   float* part, float* out) {
    const int nchunk = (B + CS_ROWS - 1) / CS_ROWS;
    hipLaunchKernelGGL(colsum_partial, dim3((N + 63) / 64, nchunk), dim3(256), 0, st, A32, A16, w, B, N, part);
    hipLaunchKernelGGL(colsum_final, dim3((N + 255) / 256), dim3(256), 0, st, part, nchunk, N, out);
}

// Loss + gradients (into the object's gradient buffer) on device buffers:
// x (B x 6144 bf16), t (B fp32 targets).  *loss_out (host) = 0.5 mean (y-t)^2.
LKL_API int lkl_mlp_grad(lkl_mlp* n, const void* d_x, const float* d_t, int B, float* loss_out) {
    if (!n || !d_x || !d_t || check_batch(B) || mlp_reserve(n, B) || mlp_ws(n, B)) return -1;
    hipStream_t st = n->stream;
    if (lkl_mlp_forward(n, d_x, B, n->y, st)) return -1;
    const bf16* X = (const bf16*)d_x;
    const bf16* W = n->p16;
    float* G = n->g;
    if (hipMemsetAsync(n->loss, 0, 4, st) != hipSuccess) return -1;
    hipLaunchKernelGGL(loss_grad, dim3(std::min(64, (B + 255) / 256)), dim3(256), 0, st, n->y, d_t, B, n->dy, n->loss);
    // layer 4 (512 -> 1): dZ3 = dy W4^T (H3 > 0)
    hipLaunchKernelGGL(head_dz, dim3(1024), dim3(256), 0, st, n->h3, W + OW4, n->dy, (long)B * D1,
                       n->dz32, n->dza);
    colsum(st, nullptr, n->h3, n->dy, B, D1, n->part, G + OW4);  // dW4 = H3^T dy
    colsum(st, n->dy, nullptr, nullptr, B, 1, n->part, G + OB4);  // db4 = sum dy
    // Weight gradients dW = H^T dZ run on transposed copies (H^T and dZ^T, batch
    // contiguous), so both GEMM operands are K-contiguous.
    auto tr = [&](const bf16* src, int rows, int cols, bf16* dst) {
        hipLaunchKernelGGL(transpose_kernel, dim3(cols / 64, rows / 64), dim3(256), 0, st, src, rows, cols, dst);
    };
    auto dw = [&](const bf16* Ht, int nin, float* out) {
        gemm<EPI_F32>(st, Ht, B, 1, n->dzt, 1, B, nin, D1, B, out, nullptr, D1, nullptr, nullptr, n->ws);
    };
    // layer 3: dW3 = H2^T dZ3, db3, dZ2 = dZ3 W3^T (H2 > 0)
    tr(n->h2, B, D1, n->ht);
    tr(n->dza, B, D1, n->dzt);
    dw(n->ht, D1, G + OW3);
    colsum(st, n->dz32, nullptr, nullptr, B, D1, n->part, G + OB3);
    gemm<EPI_MASK>(st, n->dza, D1, 1, W + OW3, 1, D1, B, D1, D1, n->dz32, n->dzb, D1, nullptr, n->h2, n->ws);
    // layer 2: dW2 = H1^T dZ2, db2, dZ1 = dZ2 W2^T (H1 > 0)
    tr(n->h1, B, D1, n->ht);
    tr(n->dzb, B, D1, n->dzt);
    dw(n->ht, D1, G + OW2);
    colsum(st, n->dz32, nullptr, nullptr, B, D1, n->part, G + OB2);
    gemm<EPI_MASK>(st, n->dzb, D1, 1, W + OW2, 1, D1, B, D1, D1, n->dz32, n->dza, D1, nullptr, n->h1, n->ws);
    // layer 1: dW1 = X^T dZ1, db1
    tr(X, B, NF, n->xt);
    tr(n->dza, B, D1, n->dzt);
    dw(n->xt, NF, G + OW1);
    colsum(st, n->dz32, nullptr, nullptr, B, D1, n->part, G + OB1);
    if (hipGetLastError() != hipSuccess) { serr("lkl_mlp_grad: launch failed"); return -1; }
    if (loss_out) {
        if (hipMemcpyAsync(loss_out, n->loss, 4, hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
        if (hipStreamSynchronize(st) != hipSuccess) return -1;
    }
    return 0;
}

// One Adam step on the stored gradients (Kingma & Ba, bias-corrected).
LKL_API int lkl_mlp_adam(lkl_mlp* n, float lr, float beta1, float beta2, float eps) {
    if (!n) return -1;
    n->step += 1;
    const float c1 = 1.0f - std::pow(beta1, (float)n->step), c2 = 1.0f - std::pow(beta2, (float)n->step);
    hipLaunchKernelGGL(adam_kernel, dim3(2048), dim3(256), 0, n->stream, n->p, n->g, n->m, n->v, n->p16,
                       NPARAM, lr, beta1, beta2, eps, c1, c2);
    refresh_transposed(n, n->stream);
    if (hipGetLastError() != hipSuccess) { serr("lkl_mlp_adam: launch failed"); return -1; }
    return 0;
}

LKL_API int lkl_mlp_sync(lkl_mlp* n) {
    if (!n) return -1;
    return hipStreamSynchronize(n->stream) == hipSuccess ? 0 : -1;
}

// Host-buffer forward (blocking): x (B x 6144 bf16 bits), y (B floats).
LKL_API int lkl_mlp_forward_host(lkl_mlp* n, const unsigned short* x, int B, float* y) {
    if (!n || !x || !y || check_batch(B)) return -1;
    void* dx = nullptr;
    float* dy = nullptr;
    int rc = 0;
    if (hipMalloc(&dx, (size_t)B * NF * 2) || hipMalloc((void**)&dy, (size_t)B * 4)) rc = -1;
    if (!rc && h2d(dx, x, (size_t)B * NF * 2, n->stream)) rc = -1;
    if (!rc) rc = lkl_mlp_forward(n, dx, B, dy, n->stream);
    if (!rc && hipStreamSynchronize(n->stream)) rc = -1;
    if (!rc && hipMemcpy(y, dy, (size_t)B * 4, hipMemcpyDeviceToHost)) rc = -1;
    if (dx) (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    return rc;
}

// Host-buffer gradient (blocking): returns the loss; gradients stay on device
// (lkl_mlp_get_grads).
LKL_API int lkl_mlp_grad_host(lkl_mlp* n, const unsigned short* x, const float* t, int B, float* loss) {
    if (!n || !x || !t || check_batch(B)) return -1;
    void* dx = nullptr;
    float* dt = nullptr;
    int rc = 0;
    if (hipMalloc(&dx, (size_t)B * NF * 2) || hipMalloc((void**)&dt, (size_t)B * 4)) rc = -1;
    if (!rc && (h2d(dx, x, (size_t)B * NF * 2, n->stream) ||
                h2d(dt, t, (size_t)B * 4, n->stream))) rc = -1;
    if (!rc) rc = lkl_mlp_grad(n, dx, dt, B, loss);
    if (hipStreamSynchronize(n->stream) != hipSuccess) rc = -1;  // before the inputs are freed
    if (dx) (void)hipFree(dx);
    if (dt) (void)hipFree(dt);
    return rc;
}

// Raw GEMM on host buffers for tests: C (M x N fp32) = op(A) op(B) with the
// element strides given (A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn]).
LKL_API int lkl_gemm_bf16_host(const unsigned short* A, long a_len, long sam, long sak,
                               const unsigned short* B, long b_len, long sbk, long sbn, int M, int N,
                               int K, float* C) {
    if (M % TM || N % TN || (K % TK && !k_contiguous(sak, sbk, M, N, K)) || M <= 0 || N <= 0 || K <= 0) {
        serr("lkl_gemm_bf16_host: shape");
        return -1;
    }
    // one unit stride per operand, the other a multiple of 8 (16-byte vector
    // loads), and every element the kernel reads inside the given lengths
    auto strides_ok = [](long s_row, long s_k) {
        return (s_k == 1 && s_row > 0 && s_row % 8 == 0) || (s_row == 1 && s_k > 0 && s_k % 8 == 0);
    };
    if (!A || !B || !C || !strides_ok(sam, sak) || !strides_ok(sbn, sbk) ||
        (long)(M - 1) * sam + (long)(K - 1) * sak >= a_len ||
        (long)(K - 1) * sbk + (long)(N - 1) * sbn >= b_len) {
        serr("lkl_gemm_bf16_host: strides or buffer lengths do not cover the shape");
        return -1;
    }
    void *dA = nullptr, *dB = nullptr;
    float *dC = nullptr, *dW = nullptr;
    int rc = 0;
    const size_t wsn = k_contiguous(sak, sbk, M, N, K) ? split_ws(M, N, K) : 0;
    if (hipMalloc(&dA, a_len * 2) || hipMalloc(&dB, b_len * 2) || hipMalloc((void**)&dC, (size_t)M * N * 4) ||
        (wsn && hipMalloc((void**)&dW, wsn * 4)))
        rc = -1;
    if (!rc && (hipMemcpy(dA, A, a_len * 2, hipMemcpyHostToDevice) || hipMemcpy(dB, B, b_len * 2, hipMemcpyHostToDevice))) rc = -1;
    if (!rc) {
        gemm<EPI_F32>(nullptr, (const bf16*)dA, sam, sak, (const bf16*)dB, sbk, sbn, M, N, K, dC, nullptr, N, nullptr,
                      nullptr, dW);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = -1;
    }
    if (!rc && hipMemcpy(C, dC, (size_t)M * N * 4, hipMemcpyDeviceToHost)) rc = -1;
    if (dA) (void)hipFree(dA);
    if (dB) (void)hipFree(dB);
    if (dC) (void)hipFree(dC);
    if (dW) (void)hipFree(dW);
    return rc;
}

// ── screening step C ABI ──────────────────────────────────────────────────────
// n base poses (device, rows w, x, y, z, tx, ty, tz) -> n * n_cand candidates
// (glowworm-major) on `stream`.  0 / -1.
LKL_API int lkl_screen_candidates(const double* d_base, int n, int n_cand, double trans_radius,
                                  double rot_sigma, unsigned long long seed, double* d_cand, void* stream) {
    if (!d_base || !d_cand || n < 0 || n_cand <= 0 || trans_radius < 0.0 || rot_sigma < 0.0) {
        serr("lkl_screen_candidates: bad arguments");
        return -1;
    }
    const long long total = (long long)n * n_cand;
    if (total == 0) return 0;
    hipLaunchKernelGGL(cand_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       d_base, total, n_cand, trans_radius, rot_sigma, seed, d_cand);
    if (hipGetLastError() != hipSuccess) {
        serr("lkl_screen_candidates: launch failed");
        return -1;
    }
    return 0;
}

// Per glowworm, the k (1..8) best of its n_cand predictions: indices into
// d_pick (n x k) and poses into d_sel (n x k x 7), on `stream`.  0 / -1.
LKL_API int lkl_screen_select(const float* d_pred, const double* d_cand, int n, int n_cand, int k,
                              int* d_pick, double* d_sel, void* stream) {
    if (!d_pred || !d_cand || !d_pick || !d_sel || n < 0 || n_cand <= 0 || k < 1 || k > SEL_MAX || k > n_cand) {
        serr("lkl_screen_select: bad arguments (need 1 <= k <= min(8, n_cand))");
        return -1;
    }
    if (n == 0) return 0;
    hipLaunchKernelGGL(select_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       d_pred, d_cand, n, n_cand, k, d_pick, d_sel);
    if (hipGetLastError() != hipSuccess) {
        serr("lkl_screen_select: launch failed");
        return -1;
    }
    return 0;
}
