// lkl_field.hip — receptor far-field grid phi built on the GPU.
//
// phi(node) = sum over charged receptor atoms i (in atom-index order) of
// (float)(q_i / d2) for 10 < |node - r_i| <= 30 A, accumulated in f32 — the
// per-node sum of the reference's shell scatter (grid_dna.rs:44-149).  The
// scatter adds atom contributions to each node in atom order; this kernel
// gathers them per node in the same order with the same f64 distance
// expression and the same membership test (dyz2 < 900 and 100 < d2 <= 900),
// so every node gets the reference's f32 value bit for bit.
//
// One workgroup owns an 8x8x4 brick of nodes.  It streams the receptor in
// index order, keeps the atoms whose shell can reach the brick (order-
// preserving compaction into LDS), and every lane then walks that list for its
// node.  No atomics, no float reassociation.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#define LKL_API extern "C" __attribute__((visibility("default")))

namespace {

constexpr int BX = 8, BY = 8, BZ = 4, BT = BX * BY * BZ;  // 256 lanes
constexpr int CAP = 2048;                                  // LDS atom buffer (32 KB)
constexpr double SPACING = 0.5;                            // grid_dna.rs:29
constexpr double R2MIN = 100.0, R2MAX = 900.0;             // grid_dna.rs:91-92

__global__ void __launch_bounds__(BT)
field_kernel(const float4* __restrict__ atoms, int nr, double lo0, double lo1, double lo2,
             int nx, int ny, int nz, float* __restrict__ phi) {
    __shared__ float4 buf[CAP];
    __shared__ int wave_cnt[BT / 64];
    __shared__ int count;
    const int tid = threadIdx.x;
    const int ix = blockIdx.x * BX + (tid % BX);
    const int iy = blockIdx.y * BY + ((tid / BX) % BY);
    const int iz = blockIdx.z * BZ + (tid / (BX * BY));
    const bool valid = ix < nx && iy < ny && iz < nz;
    // brick box in A (node positions), f32 for the conservative prefilter
    const float bx0 = (float)(lo0 + (double)(blockIdx.x * BX) * SPACING);
    const float by0 = (float)(lo1 + (double)(blockIdx.y * BY) * SPACING);
    const float bz0 = (float)(lo2 + (double)(blockIdx.z * BZ) * SPACING);
    const float bx1 = bx0 + (BX - 1) * 0.5f, by1 = by0 + (BY - 1) * 0.5f, bz1 = bz0 + (BZ - 1) * 0.5f;
    const float rmax2 = 30.02f * 30.02f, rmin2 = 9.98f * 9.98f;
    if (tid == 0) count = 0;
    __syncthreads();
    float acc = 0.0f;
    for (int base = 0; base < nr; base += BT) {
        const int i = base + tid;
        bool take = false;
        float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
        if (i < nr) {
            a = atoms[i];
            if (a.w != 0.0f) {
                const float gx = fmaxf(0.f, fmaxf(bx0 - a.x, a.x - bx1));
                const float gy = fmaxf(0.f, fmaxf(by0 - a.y, a.y - by1));
                const float gz = fmaxf(0.f, fmaxf(bz0 - a.z, a.z - bz1));
                const float fx = fmaxf(fabsf(a.x - bx0), fabsf(a.x - bx1));
                const float fy = fmaxf(fabsf(a.y - by0), fabsf(a.y - by1));
                const float fz = fmaxf(fabsf(a.z - bz0), fabsf(a.z - bz1));
                take = (gx * gx + gy * gy + gz * gz <= rmax2) && (fx * fx + fy * fy + fz * fz > rmin2);
            }
        }
        // order-preserving compaction: wave ballot + per-wave prefix
        const unsigned long long m = __ballot(take);
        const int lane = tid & 63, w = tid >> 6;
        const int before = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wave_cnt[w] = __popcll(m);
        __syncthreads();
        int off = count;
        for (int k = 0; k < w; ++k) off += wave_cnt[k];
        if (take) buf[off + before] = a;
        __syncthreads();
        if (tid == 0) {
            int t = count;
            for (int k = 0; k < BT / 64; ++k) t += wave_cnt[k];
            count = t;
        }
        __syncthreads();
        const int n = count;
        if (n > CAP - BT || base + BT >= nr) {
            if (valid) {
#pragma clang fp contract(off)
                const double nxp = lo0 + (double)ix * SPACING;
                const double nyp = lo1 + (double)iy * SPACING;
                const double nzp = lo2 + (double)iz * SPACING;
                for (int k = 0; k < n; ++k) {
                    const float4 b = buf[k];
                    const double dz = nzp - (double)b.z;
                    const double dy = nyp - (double)b.y;
                    const double dyz2 = dy * dy + dz * dz;
                    if (dyz2 >= R2MAX) continue;
                    const double dx = nxp - (double)b.x;
                    const double d2 = dx * dx + dyz2;
                    if (d2 > R2MIN && d2 <= R2MAX) acc += (float)((double)b.w / d2);
                }
            }
            __syncthreads();
            if (tid == 0) count = 0;
            __syncthreads();
        }
    }
    if (valid) phi[((size_t)iz * ny + iy) * nx + ix] = acc;
}

thread_local std::string t_ferr;

}  // namespace

// grid_dna.rs:44-61 — origin (integral) and node counts of the field box over
// the receptor coordinates.  Returns the number of nodes, or -1.
LKL_API long long lkl_field_dims(const float* r_xyz, int nr, double* origin3, int* dims3) {
    if (nr <= 0 || !r_xyz || !origin3 || !dims3) return -1;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < nr; ++k)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::fmin(lo[a], (double)r_xyz[3 * k + a]);
            hi[a] = std::fmax(hi[a], (double)r_xyz[3 * k + a]);
        }
    long long total = 1;
    for (int a = 0; a < 3; ++a) {
        lo[a] = std::floor(lo[a] - 30.0 - 2.0 * SPACING);
        hi[a] = std::ceil(hi[a] + 30.0 + 2.0 * SPACING);
        origin3[a] = lo[a];
        dims3[a] = (int)((hi[a] - lo[a]) / SPACING) + 1;
        total *= dims3[a];
    }
    return total;
}

// Build phi into a device buffer d_phi (lkl_field_dims() floats) on `stream`
// (library-internal: used by lkl_ctx_build_field).  Returns 0 / -1.
int lkl_field_build_device(const float* r_xyz, const float* r_ele, int nr, float* d_phi,
                           hipStream_t st) {
    double o[3];
    int d[3];
    if (lkl_field_dims(r_xyz, nr, o, d) <= 0 || !r_ele || !d_phi) return -1;
    std::vector<float4> at(nr);
    for (int k = 0; k < nr; ++k)
        at[k] = make_float4(r_xyz[3 * k], r_xyz[3 * k + 1], r_xyz[3 * k + 2], r_ele[k]);
    float4* d_at = nullptr;
    if (hipMalloc((void**)&d_at, sizeof(float4) * (size_t)nr) != hipSuccess) return -1;
    int rc = 0;
    if (hipMemcpyAsync(d_at, at.data(), sizeof(float4) * (size_t)nr, hipMemcpyHostToDevice, st) != hipSuccess) rc = -1;
    if (rc == 0) {
        dim3 grid((d[0] + BX - 1) / BX, (d[1] + BY - 1) / BY, (d[2] + BZ - 1) / BZ);
        hipLaunchKernelGGL(field_kernel, grid, dim3(BT), 0, st, d_at, nr, o[0], o[1], o[2], d[0],
                           d[1], d[2], d_phi);
        if (hipGetLastError() != hipSuccess) rc = -1;
    }
    if (hipStreamSynchronize(st) != hipSuccess) rc = -1;
    (void)hipFree(d_at);
    return rc;
}

// Build phi for a receptor (coordinates + charges) into phi_out (host buffer of
// lkl_field_dims() floats, x fastest).  Returns 0 / -1.
LKL_API int lkl_field_build(const float* r_xyz, const float* r_ele, int nr, float* phi_out) {
    double o[3];
    int d[3];
    const long long total = lkl_field_dims(r_xyz, nr, o, d);
    if (total <= 0 || !r_ele || !phi_out) {
        fprintf(stderr, "[lklight-gpu] lkl_field_build: bad arguments (nr=%d)\n", nr);
        return -1;
    }
    std::vector<float4> at(nr);
    for (int k = 0; k < nr; ++k)
        at[k] = make_float4(r_xyz[3 * k], r_xyz[3 * k + 1], r_xyz[3 * k + 2], r_ele[k]);
    float4* d_at = nullptr;
    float* d_phi = nullptr;
    hipStream_t st = nullptr;
    auto fail = [&](const char* stage, hipError_t e) {
        fprintf(stderr, "[lklight-gpu] lkl_field_build stage=%s: %s\n", stage, hipGetErrorString(e));
        if (d_at) (void)hipFree(d_at);
        if (d_phi) (void)hipFree(d_phi);
        if (st) (void)hipStreamDestroy(st);
        return -1;
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess) return fail("stream", e);
    if ((e = hipMalloc((void**)&d_at, sizeof(float4) * (size_t)nr)) != hipSuccess) return fail("alloc", e);
    if ((e = hipMalloc((void**)&d_phi, sizeof(float) * (size_t)total)) != hipSuccess) return fail("alloc", e);
    if ((e = hipMemcpyAsync(d_at, at.data(), sizeof(float4) * (size_t)nr, hipMemcpyHostToDevice, st)) != hipSuccess)
        return fail("upload", e);
    dim3 grid((d[0] + BX - 1) / BX, (d[1] + BY - 1) / BY, (d[2] + BZ - 1) / BZ);
    hipLaunchKernelGGL(field_kernel, grid, dim3(BT), 0, st, d_at, nr, o[0], o[1], o[2], d[0], d[1], d[2], d_phi);
    if ((e = hipGetLastError()) != hipSuccess) return fail("launch", e);
    if ((e = hipMemcpyAsync(phi_out, d_phi, sizeof(float) * (size_t)total, hipMemcpyDeviceToHost, st)) != hipSuccess)
        return fail("readback", e);
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return fail("sync", e);
    (void)hipFree(d_at);
    (void)hipFree(d_phi);
    (void)hipStreamDestroy(st);
    return 0;
}
