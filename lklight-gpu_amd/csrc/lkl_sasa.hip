// lkl_sasa.hip — Lee & Richards solvent-accessible surface area per atom on
// the GPU (the reference's lr_sasa.rs:1-291, freesasa 2.2's slice algorithm),
// used to give cpydock's desolvation its reference areas (cpydock.rs:287-332).
//
// Each atom's sphere (radius + probe) is cut by n_slices planes at
// z_i - R_i + (k - 1/2) delta, delta = 2 R_i / n_slices.  On each plane the
// neighbour spheres bury arcs of the atom's circle; the exposed length times
// delta * R_i is summed over the planes.  All arithmetic is f64 with
// contraction off, in the reference's operation order.
//
// Three launches:
//   count  one lane per atom; the atoms stream through LDS in tiles of 256 and
//          each lane counts those within R_i + R_j (strict, squared distance,
//          lr_sasa.rs:150-151).  The host turns the counts into CSR offsets.
//   fill   the same stream writes each neighbour as (z, R, |dxy|, angle), in
//          index order; angle = atan2(dy, dx) + pi seen from the atom, which the
//          reference recomputes per slice with the same operands.
//   area   one wavefront per atom, slices in order: on each slice the lanes
//          form the buried arcs of the atom's neighbours in parallel (lane k:
//          neighbours k, k + 64), a wave vote takes the slice out when one of
//          them buries it, the arcs are placed in LDS in order of their start
//          by a rank count (ties by slot), and one lane sweeps the sorted arcs
//          (lr_sasa.rs:176-213, the reference's additions in its order).
//          Atoms with more than 128 neighbours take one lane and a scratch
//          region of their own, insertion-sorted (the same sweep).  The
//          sweep's result does not depend on the order of arcs with equal
//          starts, so both give the same bits, and the reference's areas.
// The device acos/atan2 are within an ulp or two of the host libm's, so areas
// agree with a CPU evaluation to ~1e-15 relative, not bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <vector>

#define LKL_API extern "C" __attribute__((visibility("default")))

namespace {

constexpr int TILE = 256;
constexpr double PI = 3.141592653589793;
constexpr double TWOPI = 2.0 * PI;

template <bool FILL>
__global__ void __launch_bounds__(TILE)
sasa_pairs_kernel(const double4* __restrict__ at, int n, const long long* __restrict__ off,
                  int* __restrict__ cnt, double4* __restrict__ nbr) {
#pragma clang fp contract(off)
    __shared__ double4 tile[TILE];
    const int i = blockIdx.x * TILE + threadIdx.x;
    const bool valid = i < n;
    const double4 a = valid ? at[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    long long w = (FILL && valid) ? off[i] : 0;
    int c = 0;
    for (int base = 0; base < n; base += TILE) {
        const int j0 = base + (int)threadIdx.x;
        tile[threadIdx.x] = j0 < n ? at[j0] : make_double4(0.0, 0.0, 0.0, 0.0);
        __syncthreads();
        const int m = min(TILE, n - base);
        if (valid) {
            for (int t = 0; t < m; ++t) {
                const int j = base + t;
                if (j == i) continue;
                const double4 b = tile[t];
                const double dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
                const double rr = (a.w + b.w) * (a.w + b.w);
                if (dx * dx + dy * dy + dz * dz < rr) {
                    if (FILL) {
                        const double dxy = sqrt(dx * dx + dy * dy);
                        nbr[w++] = make_double4(b.z, b.w, dxy, atan2(dy, dx) + PI);
                    } else {
                        ++c;
                    }
                }
            }
        }
        __syncthreads();
    }
    if (!FILL && valid) cnt[i] = c;
}

// lr_sasa.rs:176-213 on one atom's scratch arcs (n pairs of start, end).
__device__ double exposed_arc_length(double* arc, int n) {
    if (n == 0) return TWOPI;
    for (int i = 2; i < 2 * n; i += 2) {
        const double t0 = arc[i], t1 = arc[i + 1];
        int j = i;
        while (j > 0 && arc[j - 2] > t0) {
            arc[j] = arc[j - 2];
            arc[j + 1] = arc[j - 1];
            j -= 2;
        }
        arc[j] = t0;
        arc[j + 1] = t1;
    }
    double sum = arc[0], sup = arc[1];
    for (int i2 = 2; i2 < 2 * n; i2 += 2) {
        if (sup < arc[i2]) sum += arc[i2] - sup;
        if (arc[i2 + 1] > sup) sup = arc[i2 + 1];
    }
    return sum + TWOPI - sup;
}

// The arc (or two, when it wraps past 2 pi) that neighbour q buries on the
// slice circle of radius ri_p at height z (lr_sasa.rs:225-262); returns the
// number of arcs (0: q misses the circle or lies inside it), or -1 when q
// buries the whole circle.
__device__ __forceinline__ int slice_arcs(const double4 q, double z, double ri_p, double ri_p2, double* a) {
#pragma clang fp contract(off)
    const double dj = fabs(q.x - z), rj = q.y;
    if (!(dj < rj)) return 0;
    const double rj_p2 = rj * rj - dj * dj;
    const double rj_p = sqrt(rj_p2);
    const double dij = q.z;
    if (dij >= ri_p + rj_p) return 0;
    if (dij + ri_p < rj_p) return -1;
    if (dij + rj_p < ri_p) return 0;
    const double alpha = acos((ri_p2 + dij * dij - rj_p2) / (2.0 * ri_p * dij));
    double inf = q.w - alpha, sup = q.w + alpha;
    if (inf < 0.0) inf += TWOPI;
    if (sup > TWOPI) sup -= TWOPI;
    if (sup < inf) {
        a[0] = 0.0;
        a[1] = sup;
        a[2] = inf;
        a[3] = TWOPI;
        return 2;
    }
    a[0] = inf;
    a[1] = sup;
    return 1;
}

// One atom on one lane: arcs in the atom's own scratch region.
__device__ double area_one_lane(const double4 ai, long long b, int nni, const double4* __restrict__ nbr,
                                double* __restrict__ arcs, int ns) {
#pragma clang fp contract(off)
    const double zi = ai.z, ri = ai.w;
    double* arc = arcs + 4 * b;
    const double delta = 2.0 * ri / (double)ns;
    double z = zi - ri - 0.5 * delta;
    double sasa = 0.0;
    for (int s = 0; s < ns; ++s) {
        z += delta;
        const double di = fabs(zi - z);
        const double ri_p2 = ri * ri - di * di;
        if (ri_p2 < 0.0) continue;
        const double ri_p = sqrt(ri_p2);
        if (ri_p <= 0.0) continue;
        int na = 0;
        bool buried = false;
        for (int k = 0; k < nni; ++k) {
            const int c = slice_arcs(nbr[b + k], z, ri_p, ri_p2, arc + 2 * na);
            if (c < 0) { buried = true; break; }
            na += c;
        }
        if (!buried) sasa += delta * ri * exposed_arc_length(arc, na);
    }
    return sasa;
}

// Wavefront per atom (see the file comment).  SW atoms per workgroup, each
// wave on its own LDS; no workgroup barrier (the waves take their own paths):
// lanes of one wave meet in LDS through a wave-local fence.
constexpr int SW = 4;
constexpr int ARC_CAP = 256;  // arc slots per slice: two per neighbour, 128 neighbours

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");  // this wave's LDS accesses complete
    __builtin_amdgcn_wave_barrier();
}

__global__ void __launch_bounds__(64 * SW)
sasa_area_wave_kernel(const double4* __restrict__ at, int n, const long long* __restrict__ off,
                      const double4* __restrict__ nbr, double* __restrict__ arcs, int ns, int lane_only,
                      double* __restrict__ out) {
#pragma clang fp contract(off)
    __shared__ double2 slot[SW][ARC_CAP];    // arcs by neighbour slot (start = +inf: empty)
    __shared__ double2 sorted[SW][ARC_CAP];  // the same arcs by start
    const int w = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    const int i = (int)blockIdx.x * SW + w;
    if (i >= n) return;  // uniform per wave
    const double4 ai = at[i];
    const long long b = off[i];
    const int nni = (int)(off[i + 1] - b);
    if (lane_only || 2 * nni > ARC_CAP) {
        if (lane == 0) out[i] = area_one_lane(ai, b, nni, nbr, arcs, ns);
        return;
    }
    const double zi = ai.z, ri = ai.w;
    const double delta = 2.0 * ri / (double)ns;
    double z = zi - ri - 0.5 * delta;
    double sasa = 0.0;
    const int nslot = 2 * nni;
    for (int s = 0; s < ns; ++s) {
        z += delta;
        const double di = fabs(zi - z);
        const double ri_p2 = ri * ri - di * di;
        if (ri_p2 < 0.0) continue;
        const double ri_p = sqrt(ri_p2);
        if (ri_p <= 0.0) continue;
        // arcs of neighbours lane and lane + 64, two slots each
        bool buried = false;
        for (int k = lane; k < nni; k += 64) {
            double a[4];
            const int c = slice_arcs(nbr[b + k], z, ri_p, ri_p2, a);
            buried |= c < 0;
            slot[w][2 * k] = c >= 1 ? make_double2(a[0], a[1]) : make_double2(INFINITY, INFINITY);
            slot[w][2 * k + 1] = c == 2 ? make_double2(a[2], a[3]) : make_double2(INFINITY, INFINITY);
        }
        if (__any(buried)) {
            wave_sync_lds();  // the slots are rewritten on the next slice
            continue;
        }
        wave_sync_lds();
        // rank by (start, slot): the arcs' order of start, ties by slot.  Starts
        // are +0 .. 2 pi or +inf (empty), so their bit patterns order like the
        // values; a NaN bound (a degenerate tangency) sorts after +inf and is
        // dropped with the empty slots.
        int na = 0;
        for (int e = lane; e < nslot; e += 64) {
            const double2 me = slot[w][e];
            const unsigned long long ke = (unsigned long long)__double_as_longlong(me.x);
            int rank = 0;
            for (int f = 0; f < nslot; ++f) {
                const unsigned long long kf = (unsigned long long)__double_as_longlong(slot[w][f].x);
                rank += (kf < ke || (kf == ke && f < e)) ? 1 : 0;
            }
            sorted[w][rank] = me;
            na += me.x < INFINITY ? 1 : 0;
        }
        for (int o = 32; o > 0; o >>= 1) na += __shfl_xor(na, o, 64);
        wave_sync_lds();
        double exposed = 0.0;
        if (lane == 0) {  // lr_sasa.rs:176-213 over the sorted arcs
            if (na == 0) {
                exposed = TWOPI;
            } else {
                double sum = sorted[w][0].x, sup = sorted[w][0].y;
                for (int k = 1; k < na; ++k) {
                    const double2 q = sorted[w][k];
                    if (sup < q.x) sum += q.x - sup;
                    if (q.y > sup) sup = q.y;
                }
                exposed = sum + TWOPI - sup;
            }
            sasa += delta * ri * exposed;
        }
        wave_sync_lds();  // sorted[] is rewritten on the next slice
    }
    if (lane == 0) out[i] = sasa;
}

}  // namespace

// Per-atom Lee & Richards area (A^2) of n atoms: xyz n x 3 f64, radii without
// the probe; probe 1.4 and 20 slices are the reference's values
// (lr_sasa.rs:277-291).  Host buffers in and out; blocking.  0 / -1.
LKL_API int lkl_lr_sasa(const double* xyz, const double* radii, int n, double probe, int n_slices,
                        double* out) {
    if (n == 0) return 0;
    if (n < 0 || !xyz || !radii || !out || n_slices < 1 || !(probe >= 0.0) || !std::isfinite(probe)) {
        fprintf(stderr, "[lklight-gpu] lkl_lr_sasa: bad arguments (n=%d, n_slices=%d, probe=%g)\n", n,
                n_slices, probe);
        return -1;
    }
    std::vector<double4> at(n);
    double rmax = 0.0;
    for (int k = 0; k < n; ++k) {
        const double r = radii[k] + probe;
        if (!std::isfinite(r) || r < 0.0 || !std::isfinite(xyz[3 * k]) || !std::isfinite(xyz[3 * k + 1]) ||
            !std::isfinite(xyz[3 * k + 2])) {
            fprintf(stderr, "[lklight-gpu] lkl_lr_sasa: atom %d has a non-finite coordinate or radius\n", k);
            return -1;
        }
        rmax = std::fmax(rmax, r);
        at[k] = make_double4(xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2], r);
    }
    if (!(rmax > 0.0)) {  // lr_sasa.rs:40 asserts a positive cell size
        fprintf(stderr, "[lklight-gpu] lkl_lr_sasa: every radius + probe is zero\n");
        return -1;
    }
    double4 *d_at = nullptr, *d_nbr = nullptr;
    long long* d_off = nullptr;
    int* d_cnt = nullptr;
    double *d_arc = nullptr, *d_out = nullptr;
    hipStream_t st = nullptr;
    auto release = [&]() {
        if (d_at) (void)hipFree(d_at);
        if (d_nbr) (void)hipFree(d_nbr);
        if (d_off) (void)hipFree(d_off);
        if (d_cnt) (void)hipFree(d_cnt);
        if (d_arc) (void)hipFree(d_arc);
        if (d_out) (void)hipFree(d_out);
        if (st) (void)hipStreamDestroy(st);
    };
    auto fail = [&](const char* stage, hipError_t e) {
        fprintf(stderr, "[lklight-gpu] lkl_lr_sasa stage=%s: %s\n", stage, hipGetErrorString(e));
        release();
        return -1;
    };
    hipError_t e;
    const dim3 grid((n + TILE - 1) / TILE);
    if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess) return fail("stream", e);
    if ((e = hipMalloc((void**)&d_at, sizeof(double4) * (size_t)n)) != hipSuccess) return fail("alloc", e);
    if ((e = hipMalloc((void**)&d_cnt, sizeof(int) * (size_t)n)) != hipSuccess) return fail("alloc", e);
    if ((e = hipMalloc((void**)&d_off, sizeof(long long) * ((size_t)n + 1))) != hipSuccess) return fail("alloc", e);
    if ((e = hipMalloc((void**)&d_out, sizeof(double) * (size_t)n)) != hipSuccess) return fail("alloc", e);
    if ((e = hipMemcpyAsync(d_at, at.data(), sizeof(double4) * (size_t)n, hipMemcpyHostToDevice, st)) != hipSuccess)
        return fail("upload", e);
    hipLaunchKernelGGL(sasa_pairs_kernel<false>, grid, dim3(TILE), 0, st, d_at, n, nullptr, d_cnt, nullptr);
    if ((e = hipGetLastError()) != hipSuccess) return fail("count", e);
    std::vector<int> cnt(n);
    if ((e = hipMemcpyAsync(cnt.data(), d_cnt, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, st)) != hipSuccess)
        return fail("counts", e);
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return fail("sync", e);
    std::vector<long long> off(n + 1, 0);
    for (int k = 0; k < n; ++k) off[k + 1] = off[k] + cnt[k];
    const size_t tot = (size_t)off[n] + 1;
    if ((e = hipMalloc((void**)&d_nbr, sizeof(double4) * tot)) != hipSuccess) return fail("alloc", e);
    if ((e = hipMalloc((void**)&d_arc, sizeof(double) * 4 * tot)) != hipSuccess) return fail("alloc", e);
    if ((e = hipMemcpyAsync(d_off, off.data(), sizeof(long long) * ((size_t)n + 1), hipMemcpyHostToDevice, st)) !=
        hipSuccess)
        return fail("upload", e);
    hipLaunchKernelGGL(sasa_pairs_kernel<true>, grid, dim3(TILE), 0, st, d_at, n, d_off, nullptr, d_nbr);
    if ((e = hipGetLastError()) != hipSuccess) return fail("fill", e);
    // LKL_SASA_LANE=1 runs every atom on one lane (the path of atoms with more
    // than 128 neighbours), for tests that compare the two paths
    const int lane_only = getenv("LKL_SASA_LANE") && getenv("LKL_SASA_LANE")[0] == '1' ? 1 : 0;
    hipLaunchKernelGGL(sasa_area_wave_kernel, dim3((unsigned)((n + SW - 1) / SW)), dim3(64 * SW), 0, st, d_at, n,
                       d_off, d_nbr, d_arc, n_slices, lane_only, d_out);
    if ((e = hipGetLastError()) != hipSuccess) return fail("area", e);
    if ((e = hipMemcpyAsync(out, d_out, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, st)) != hipSuccess)
        return fail("readback", e);
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return fail("sync", e);
    release();
    return 0;
}
