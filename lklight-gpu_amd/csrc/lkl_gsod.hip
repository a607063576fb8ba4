// lkl_gsod.hip — device-resident GSO steps.
//
// The state of every glowworm of every swarm lives in device memory and one
// step is three launches on one stream: the scoring kernel over all poses,
// the luciferin update (swarm.rs:70-103: the moved glowworms take their new
// score, luciferin = (1 - rho) l + gamma s), and the movement
// (swarm.rs:105-174, glowworm.rs:84-166): per glowworm, the brighter
// glowworms within its vision range in id order, their probabilities, the
// choice by the step's uniform draw, the move towards the chosen neighbour
// and the vision update.  The draws come from each swarm's StdRng on the host
// (one per glowworm in id order, as swarm.rs:165), uploaded once for all steps.
// Arithmetic is the host driver's in the same order with contraction off;
// slerp's acos and sin are correctly rounded (lkl_gsomath.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <vector>

#include "../../include/lklight_gpu.h"
#include "lkl_gsomath.hpp"

#define LKL_API extern "C" __attribute__((visibility("default")))

namespace {

// glowworm.rs:45-51, constants.rs:5-8
constexpr double kRho = 0.5, kGamma = 0.4, kBeta = 0.08, kMaxVision = 5.0;
constexpr int kMaxNeighbors = 5;
constexpr double kTransStep = 0.5, kRotStep = 0.5;

struct GsoView {
    double* pose;     // n x 7 (w, x, y, z, tx, ty, tz), current
    double* pose_nx;  // n x 7, written by the movement
    double* luc;
    double* vis;
    double* score;
    unsigned char* moved;
    int* nbc;
    unsigned* step;
    const int* swarm_of;   // glowworm -> swarm
    const int* swarm_off;  // swarm -> first glowworm (n_swarms + 1)
    int n;
};

// gpu_score.rs:326-331 / gpu_family.rs:283-288; cpydock.rs:789-791.
__device__ __forceinline__ double fold(int family, double e, double v, double s) {
#pragma clang fp contract(off)
    const double ek = e * 332.0 / 4.0;
    return family == 3 ? -(ek + 0.1 * v) + s : -(ek + v);
}

// dna.rs:989-995 with scoring.rs:83-110 fractions from interface counts.
__device__ __forceinline__ double restraint_fold(double score, const int* cnt, int3 sz) {
#pragma clang fp contract(off)
    const double perc_r = sz.x ? (double)cnt[0] / (double)sz.x : 0.0;
    const double perc_l = sz.y ? (double)cnt[1] / (double)sz.y : 0.0;
    const double inter = sz.z ? (double)cnt[2] / (double)sz.z : 0.0;
    const double penalty = inter > 0.0 ? 999.0 * inter : 0.0;
    score += perc_r * score + perc_l * score - penalty;
    return score;
}

__global__ void luciferin_kernel(GsoView G, const double* __restrict__ out, const double* __restrict__ outS,
                                 const int* __restrict__ icnt, int3 isz, int family) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G.n) return;
    if (G.moved[i] || G.step[i] == 0u) {
        double sc = fold(family, out[2 * i], out[2 * i + 1], outS ? outS[i] : 0.0);
        if (icnt) sc = restraint_fold(sc, icnt + 3 * (size_t)i, isz);
        G.score[i] = sc;
    }
    G.luc[i] = (1.0 - kRho) * G.luc[i] + kGamma * G.score[i];
    G.step[i] += 1u;
}

// glowworm.rs:128-150 and the vision update: glowworm i moves towards nid
// (nid == i: stays) having seen cnt neighbours.  Reads the step's snapshot and
// writes only glowworm i's next pose, moved flag, neighbour count and vision.
__device__ void move_to(const GsoView& G, int i, int nid, int cnt) {
#pragma clang fp contract(off)
    const double* pi = G.pose + 7 * (size_t)i;
    const double vi = G.vis[i];
    const double tx = pi[4], ty = pi[5], tz = pi[6];
    double* po = G.pose_nx + 7 * (size_t)i;
    if (nid != i) {
        const double* pn = G.pose + 7 * (size_t)nid;
        const double dx = pn[4] - tx, dy = pn[5] - ty, dz = pn[6] - tz;
        const double norm = sqrt(dx * dx + dy * dy + dz * dz);
        const double coef = kTransStep / norm;
        double t0 = tx, t1 = ty, t2 = tz;
        t0 += dx * coef;
        t1 += dy * coef;
        t2 += dz * coef;
        double q[4];
        lkl::gm::slerp_cr(pi, pn, kRotStep, q);
        lkl::gm::normalize4(q[0], q[1], q[2], q[3]);
        po[0] = q[0]; po[1] = q[1]; po[2] = q[2]; po[3] = q[3];
        po[4] = t0; po[5] = t1; po[6] = t2;
    } else {
        for (int c = 0; c < 7; ++c) po[c] = pi[c];
    }
    G.moved[i] = nid != i ? 1 : 0;
    G.nbc[i] = cnt;
    G.vis[i] = fmin(kMaxVision, fmax(0.0, vi + kBeta * (double)(kMaxNeighbors - cnt)));
}

// glowworm.rs:84-126, one lane per glowworm (swarms larger than the LDS list
// of move_wave_kernel): the brighter members within the vision range in id
// order, their total, and the member chosen by walking the cumulative
// probabilities with the step's draw.
__global__ void move_kernel(GsoView G, const double* __restrict__ rnd) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G.n) return;
    const int s = G.swarm_of[i], b = G.swarm_off[s], e = G.swarm_off[s + 1];
    const double* pi = G.pose + 7 * (size_t)i;
    const double li = G.luc[i], vi = G.vis[i], vr2 = vi * vi;
    const double tx = pi[4], ty = pi[5], tz = pi[6];
    int cnt = 0, first = -1;
    double total = 0.0;
    for (int k = b; k < e; ++k) {
        if (k == i || !(li < G.luc[k])) continue;
        const double* pk = G.pose + 7 * (size_t)k;
        const double dx = tx - pk[4], dy = ty - pk[5], dz = tz - pk[6];
        if (dx * dx + dy * dy + dz * dz < vr2) {
            if (first < 0) first = k;
            ++cnt;
            total += G.luc[k] - li;
        }
    }
    int nid = i;
    if (cnt > 0) {
        const double r = rnd[i];
        double sum = 0.0;
        int last = -1;
        for (int k = b; k < e && sum < r; ++k) {
            if (k == i || !(li < G.luc[k])) continue;
            const double* pk = G.pose + 7 * (size_t)k;
            const double dx = tx - pk[4], dy = ty - pk[5], dz = tz - pk[6];
            if (!(dx * dx + dy * dy + dz * dz < vr2)) continue;
            double p = G.luc[k] - li;
            p /= total;
            sum += p;
            last = k;
        }
        nid = last >= 0 ? last : first;
    }
    move_to(G, i, nid, cnt);
}

// The same selection with one wavefront per glowworm: the 64 lanes test 64
// members of the swarm at a time and those that pass are appended to an LDS
// list in id order (ballot + prefix count).  Lane 0 then forms the total and
// walks the cumulative probabilities over that list in id order -- the sums
// and their order are move_kernel's -- and moves the glowworm.  A lane per
// glowworm would wait on a dependent load per member; here the loads of a
// whole chunk are in flight together.  Swarms of up to kSwarmList members.
constexpr int kSwarmList = 1024;
__global__ void __launch_bounds__(64) move_wave_kernel(GsoView G, const double* __restrict__ rnd) {
#pragma clang fp contract(off)
    __shared__ int nb_id[kSwarmList];
    __shared__ double nb_l[kSwarmList];
    const int i = blockIdx.x, lane = threadIdx.x;
    const int s = G.swarm_of[i], b = G.swarm_off[s], e = G.swarm_off[s + 1];
    const double* pi = G.pose + 7 * (size_t)i;
    const double li = G.luc[i], vr2 = G.vis[i] * G.vis[i];
    const double tx = pi[4], ty = pi[5], tz = pi[6];
    int cnt = 0;
    for (int k0 = b; k0 < e; k0 += 64) {
        const int k = k0 + lane;
        bool pass = false;
        double lk = 0.0;
        if (k < e && k != i) {
            lk = G.luc[k];
            const double* pk = G.pose + 7 * (size_t)k;
            const double dx = tx - pk[4], dy = ty - pk[5], dz = tz - pk[6];
            pass = li < lk && dx * dx + dy * dy + dz * dz < vr2;
        }
        const unsigned long long m = __ballot(pass);
        if (pass) {  // at < cnt + popcount(m) <= e - b - 1 < kSwarmList
            const int at = cnt + __popcll(m & ((1ull << lane) - 1ull));
            nb_id[at] = k;
            nb_l[at] = lk;
        }
        cnt += __popcll(m);
    }
    __syncthreads();
    if (lane != 0) return;
    int nid = i;
    if (cnt > 0) {
        double total = 0.0;
        for (int q = 0; q < cnt; ++q) total += nb_l[q] - li;
        const double r = rnd[i];
        double sum = 0.0;
        int last = -1;
        for (int q = 0; q < cnt && sum < r; ++q) {
            double p = nb_l[q] - li;
            p /= total;
            sum += p;
            last = nb_id[q];
        }
        nid = last >= 0 ? last : nb_id[0];
    }
    move_to(G, i, nid, cnt);
}

template <typename T>
hipError_t grow(T** p, size_t* cap, size_t n) {
    if (n <= *cap && *p) return hipSuccess;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    const hipError_t e = hipMalloc((void**)p, sizeof(T) * (n ? n : 1));
    if (e == hipSuccess) *cap = n;
    return e;
}

}  // namespace

// Device buffers of one GSO driver (grown on demand, freed by lkl_gsod_free).
struct lkl_gsod_buf {
    double *pose = nullptr, *pose_nx = nullptr, *luc = nullptr, *vis = nullptr, *score = nullptr, *rnd = nullptr,
           *out = nullptr, *outS = nullptr;
    unsigned char* moved = nullptr;
    int *nbc = nullptr, *swarm_of = nullptr, *swarm_off = nullptr, *icnt = nullptr;
    unsigned* step = nullptr;
    hipStream_t stream = nullptr;  // used when the caller passes none
    size_t c_pose = 0, c_pose_nx = 0, c_luc = 0, c_vis = 0, c_score = 0, c_rnd = 0, c_out = 0, c_outS = 0,
           c_moved = 0, c_nbc = 0, c_swarm_of = 0, c_swarm_off = 0, c_step = 0, c_icnt = 0;
};

void lkl_gsod_free(lkl_gsod_buf* d) {
    if (!d) return;
    void* p[] = {d->pose, d->pose_nx, d->luc, d->vis, d->score, d->rnd, d->out, d->outS,
                 d->moved, d->nbc, d->swarm_of, d->swarm_off, d->step, d->icnt};
    for (void* q : p)
        if (q) (void)hipFree(q);
    if (d->stream) (void)hipStreamDestroy(d->stream);
    delete d;
}

// Run `steps` device-resident steps.  Host arrays in and out (n glowworms,
// the swarms contiguous): pose7 (w, x, y, z, tx, ty, tz), luciferin, vision,
// score, moved, step, neighbour count (out only); rnd holds steps x n draws.
// Returns 0, or -1 with a message in err.
int lkl_gsod_run(lkl_gsod_buf** pd, lkl_ctx* ctx, int family, unsigned flags, int n, int n_swarms,
                 const int* swarm_off, double* pose7, double* luc, double* vis, double* score,
                 unsigned char* moved, unsigned* step, int* nbc, const double* rnd, int steps, void* stream,
                 char* err, size_t err_len) {
    auto fail = [&](const char* what, hipError_t e) {
        snprintf(err, err_len, "lkl_gso_device_run: %s: %s", what, e == hipSuccess ? "failed" : hipGetErrorString(e));
        return -1;
    };
    if (!*pd) *pd = new lkl_gsod_buf();
    lkl_gsod_buf* d = *pd;
    hipError_t e;
    if (!stream && !d->stream) {
        if ((e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking)) != hipSuccess) return fail("stream", e);
    }
    // every upload, launch and readback of the run is ordered on this one stream
    hipStream_t st = stream ? (hipStream_t)stream : d->stream;
    const size_t N = (size_t)n;
    std::vector<int> swarm_of(N);
    for (int s = 0; s < n_swarms; ++s)
        for (int k = swarm_off[s]; k < swarm_off[s + 1]; ++k) swarm_of[k] = s;
#define GROW(field, count)                                                         \
    if ((e = grow(&d->field, &d->c_##field, (count))) != hipSuccess) return fail("alloc", e);
    GROW(pose, 7 * N) GROW(pose_nx, 7 * N) GROW(luc, N) GROW(vis, N) GROW(score, N) GROW(rnd, (size_t)steps * N)
    GROW(out, 2 * N) GROW(outS, N) GROW(moved, N) GROW(nbc, N) GROW(swarm_of, N) GROW(swarm_off, (size_t)n_swarms + 1)
    GROW(step, N) GROW(icnt, 3 * N)
#undef GROW
    const bool iface = lkl_ctx_has_interface(ctx) == 1;
    int3 isz = make_int3(0, 0, 0);
    if (iface && lkl_ctx_interface_sizes(ctx, &isz.x, &isz.y, &isz.z)) return fail("interface sizes", hipSuccess);
    struct Up { void* dst; const void* src; size_t bytes; };
    const Up up[] = {{d->pose, pose7, 56 * N}, {d->luc, luc, 8 * N}, {d->vis, vis, 8 * N}, {d->score, score, 8 * N},
                     {d->moved, moved, N}, {d->step, step, 4 * N}, {d->swarm_of, swarm_of.data(), 4 * N},
                     {d->swarm_off, swarm_off, 4 * ((size_t)n_swarms + 1)}, {d->rnd, rnd, 8 * (size_t)steps * N}};
    for (const Up& u : up)
        if ((e = hipMemcpyAsync(u.dst, u.src, u.bytes, hipMemcpyHostToDevice, st)) != hipSuccess) return fail("upload", e);
    GsoView G{d->pose, d->pose_nx, d->luc, d->vis, d->score, d->moved, d->nbc, d->step, d->swarm_of, d->swarm_off, n};
    const dim3 grid((unsigned)((N + 127) / 128)), block(128);
    int smax = 0;
    for (int q = 0; q < n_swarms; ++q) smax = std::max(smax, swarm_off[q + 1] - swarm_off[q]);
    const bool listed = smax <= kSwarmList;  // every swarm fits the LDS list
    // the swarm poses' launches are short and even: the dispatch-order
    // pre-pass would cost more than it returns, so it is off for the run
    const int order = lkl_ctx_pose_order(ctx);
    lkl_ctx_set_pose_order(ctx, 0);
    struct Restore {
        lkl_ctx* c;
        int on;
        ~Restore() { lkl_ctx_set_pose_order(c, on); }
    } restore{ctx, order};
    for (int k = 0; k < steps; ++k) {
        if (lkl_ctx_score_async(ctx, G.pose, n, flags, d->out, (flags & LKL_F_DESOLV) ? d->outS : nullptr, st))
            return fail("score", hipSuccess);
        if (iface && lkl_ctx_interface_async(ctx, G.pose, n, d->icnt, st)) return fail("interface", hipSuccess);
        hipLaunchKernelGGL(luciferin_kernel, grid, block, 0, st, G, d->out,
                           (flags & LKL_F_DESOLV) ? d->outS : nullptr, iface ? d->icnt : nullptr, isz, family);
        if (listed) hipLaunchKernelGGL(move_wave_kernel, dim3((unsigned)N), dim3(64), 0, st, G, d->rnd + (size_t)k * N);
        else hipLaunchKernelGGL(move_kernel, grid, block, 0, st, G, d->rnd + (size_t)k * N);
        if ((e = hipGetLastError()) != hipSuccess) return fail("launch", e);
        std::swap(G.pose, G.pose_nx);
    }
    struct Down { void* dst; const void* src; size_t bytes; };
    const Down down[] = {{pose7, G.pose, 56 * N}, {luc, d->luc, 8 * N}, {vis, d->vis, 8 * N}, {score, d->score, 8 * N},
                         {moved, d->moved, N}, {step, d->step, 4 * N}, {nbc, d->nbc, 4 * N}};
    for (const Down& u : down)
        if ((e = hipMemcpyAsync(u.dst, u.src, u.bytes, hipMemcpyDeviceToHost, st)) != hipSuccess) return fail("readback", e);
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return fail("sync", e);
    return 0;
}

// The device step's correctly rounded sin and acos, on the host (for tests).
LKL_API double lkl_sin_cr(double x) { return lkl::gm::sin_cr(x); }
LKL_API double lkl_acos_cr(double d) { return lkl::gm::acos_cr(d); }
LKL_API void lkl_quat_slerp_cr(const double* a, const double* b, double t, double* out) {
    lkl::gm::slerp_cr(a, b, t, out);
}
