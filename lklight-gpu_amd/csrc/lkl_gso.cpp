// lkl_gso.cpp — host GSO driver over the GPU scorer (C ABI: lkl_gso_*).
//
// Restates LKlight's glowworm swarm optimisation: GSO::run (lib.rs:66-78),
// Swarm::update_luciferin's batched branch (swarm.rs:70-103), the movement
// phase (swarm.rs:105-174) and the glowworm rules (glowworm.rs:61-166), with
// rand 0.7.3's StdRng (ChaCha20, seed_from_u64) so trajectories follow the
// reference's random stream.  Several independent swarms (each with its own
// StdRng seeded like the reference seeds every swarm) are stepped together:
// the poses needing a score in all swarms go to the GPU in ONE launch.
//
// Scoring failures return -1.  A caller-supplied fallback (the host's own CPU
// grid path, as dna.rs:1081-1085 re-scores on the CPU) is used only when one is
// registered; a driver created without a context scores through it alone
// (the reference's supports_batch() == false branch, swarm.rs:102).
#include "../../include/lklight_gpu.h"

#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#define LKL_API extern "C" __attribute__((visibility("default")))

namespace {

// rand_chacha 0.2.2 ChaCha20 core with rand_core 0.5.1 seed_from_u64.
class StdRng {
  public:
    explicit StdRng(uint64_t state) {
        const uint64_t mul = 6364136223846793005ULL, inc = 11634580027462260723ULL;
        for (auto& k : key_) {
            state = state * mul + inc;
            const uint32_t xs = (uint32_t)(((state >> 18) ^ state) >> 27);
            const uint32_t rot = (uint32_t)(state >> 59);
            k = (xs >> rot) | (xs << ((32 - rot) & 31));
        }
    }
    // gen::<f64>(): top 53 bits of next_u64 (two little-endian words).
    double next_f64() {
        const uint64_t lo = next_u32(), hi = next_u32();
        return (double)(((hi << 32) | lo) >> 11) * 0x1.0p-53;
    }

  private:
    uint32_t next_u32() {
        if (pos_ == 16) refill();
        return block_[pos_++];
    }
    static uint32_t rotl(uint32_t v, int c) { return (v << c) | (v >> (32 - c)); }
    static void quarter(uint32_t* x, int a, int b, int c, int d) {
        x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 16);
        x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 12);
        x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 8);
        x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 7);
    }
    void refill() {
        uint32_t in[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
        for (int i = 0; i < 8; ++i) in[4 + i] = key_[i];
        in[12] = (uint32_t)counter_;
        in[13] = (uint32_t)(counter_ >> 32);
        in[14] = in[15] = 0;
        uint32_t x[16];
        std::memcpy(x, in, sizeof x);
        for (int r = 0; r < 10; ++r) {
            quarter(x, 0, 4, 8, 12); quarter(x, 1, 5, 9, 13);
            quarter(x, 2, 6, 10, 14); quarter(x, 3, 7, 11, 15);
            quarter(x, 0, 5, 10, 15); quarter(x, 1, 6, 11, 12);
            quarter(x, 2, 7, 8, 13); quarter(x, 3, 4, 9, 14);
        }
        for (int i = 0; i < 16; ++i) block_[i] = x[i] + in[i];
        ++counter_;
        pos_ = 0;
    }
    uint32_t key_[8];
    uint32_t block_[16];
    uint64_t counter_ = 0;
    int pos_ = 16;
};

struct Quat {
    double w, x, y, z;
    double dot(const Quat& o) const { return w * o.w + x * o.x + y * o.y + z * o.z; }
    void normalize() {  // qt.rs:41-47
        const double n = std::sqrt(w * w + x * x + y * y + z * z);
        w /= n; x /= n; y /= n; z /= n;
    }
    // qt.rs:80-104 (linear threshold 0.9995, constants.rs:11)
    Quat slerp(const Quat& other, double t) const {
        Quat a = *this, b = other;
        a.normalize();
        b.normalize();
        double d = a.dot(b);
        if (d < 0.0) { a = {-a.w, -a.x, -a.y, -a.z}; d *= -1.0; }
        if (d > 0.9995) {
            Quat r = {a.w + (b.w - a.w) * t, a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t,
                      a.z + (b.z - a.z) * t};
            r.normalize();
            return r;
        }
        d = std::fmax(std::fmin(d, 1.0), -1.0);
        const double om = std::acos(d), so = std::sin(om);
        const double s1 = std::sin((1.0 - t) * om) / so, s2 = std::sin(t * om) / so;
        return {a.w * s1 + b.w * s2, a.x * s1 + b.x * s2, a.y * s1 + b.y * s2, a.z * s1 + b.z * s2};
    }
};

// glowworm.rs:45-51
struct Glowworm {
    std::array<double, 3> t;
    Quat q;
    double luciferin = 5.0, vision = 0.2, scoring = 0.0;
    bool moved = false;
    unsigned step = 0;
    std::vector<int> nb;
    std::vector<double> prob;
};
constexpr double kRho = 0.5, kGamma = 0.4, kBeta = 0.08, kMaxVision = 5.0;
constexpr int kMaxNeighbors = 5;
constexpr double kTransStep = 0.5, kRotStep = 0.5;  // constants.rs:5-8

// A glowworm's planned move (the current pose when it has no brighter neighbour).
struct Move {
    std::array<double, 3> t;
    Quat q;
    bool moved;
};

struct Swarm {
    std::vector<Glowworm> g;
    StdRng rng;
    explicit Swarm(uint64_t seed) : rng(seed) {}

    // swarm.rs:105-174 with glowworm.rs:84-166, without applying the moves:
    // neighbours and probabilities from the step's snapshot, one draw per
    // glowworm in id order, the chosen neighbour's move, the vision update.
    std::vector<Move> plan_moves() {
        const size_t n = g.size();
        std::vector<std::array<double, 3>> pos(n);
        std::vector<Quat> rot(n);
        std::vector<double> luc(n);
        for (size_t i = 0; i < n; ++i) { pos[i] = g[i].t; rot[i] = g[i].q; luc[i] = g[i].luciferin; }
        for (size_t i = 0; i < n; ++i) {
            Glowworm& a = g[i];
            a.nb.clear();
            const double vr2 = a.vision * a.vision;
            for (size_t k = 0; k < n; ++k) {
                if (k == i || !(a.luciferin < g[k].luciferin)) continue;
                const double dx = a.t[0] - g[k].t[0], dy = a.t[1] - g[k].t[1], dz = a.t[2] - g[k].t[2];
                if (dx * dx + dy * dy + dz * dz < vr2) a.nb.push_back((int)k);
            }
            a.prob.assign(a.nb.size(), 0.0);
            double total = 0.0;
            for (size_t m = 0; m < a.nb.size(); ++m) {
                a.prob[m] = luc[a.nb[m]] - a.luciferin;
                total += a.prob[m];
            }
            for (double& p : a.prob) p /= total;
        }
        std::vector<double> rnd(n);
        for (double& r : rnd) r = rng.next_f64();  // serial draws, id order (swarm.rs:165)
        std::vector<Move> mv(n);
        for (size_t i = 0; i < n; ++i) {
            Glowworm& a = g[i];
            size_t nid = i;
            if (!a.nb.empty()) {  // glowworm.rs:114-126 (bounded walk)
                double sum = 0.0;
                size_t m = 0;
                while (sum < rnd[i] && m < a.nb.size()) sum += a.prob[m++];
                nid = (size_t)a.nb[m > 0 ? m - 1 : 0];
            }
            mv[i] = {a.t, a.q, nid != i};
            if (mv[i].moved) {  // glowworm.rs:128-150
                const double dx = pos[nid][0] - a.t[0], dy = pos[nid][1] - a.t[1],
                             dz = pos[nid][2] - a.t[2];
                const double norm = std::sqrt(dx * dx + dy * dy + dz * dz);
                const double coef = kTransStep / norm;
                mv[i].t[0] += dx * coef;
                mv[i].t[1] += dy * coef;
                mv[i].t[2] += dz * coef;
                mv[i].q = a.q.slerp(rot[nid], kRotStep);
                mv[i].q.normalize();
            }
            // glowworm.rs:91-96
            a.vision = std::fmin(kMaxVision,
                                 std::fmax(0.0, a.vision + kBeta * (double)(kMaxNeighbors - (int)a.nb.size())));
        }
        return mv;
    }

    void movement_phase() {
        const std::vector<Move> mv = plan_moves();
        for (size_t i = 0; i < g.size(); ++i) {
            g[i].moved = mv[i].moved;
            if (mv[i].moved) {
                g[i].t = mv[i].t;
                g[i].q = mv[i].q;
            }
        }
    }
};

}  // namespace

typedef int (*lkl_score_fallback)(void* user, const double* poses, int n, double* scores);

void lkl_set_error_text(const char* msg);  // lkl_score.hip: the thread's lkl_last_error()

// Device-resident steps (lkl_gsod.hip).
struct lkl_gsod_buf;
void lkl_gsod_free(lkl_gsod_buf* d);
int lkl_gsod_run(lkl_gsod_buf** pd, lkl_ctx* ctx, int family, unsigned flags, int n, int n_swarms,
                 const int* swarm_off, double* pose7, double* luc, double* vis, double* score,
                 unsigned char* moved, unsigned* step, int* nbc, const double* rnd, int steps, void* stream,
                 char* err, size_t err_len);  // restraints: from lkl_ctx_has_interface / _sizes

struct lkl_gso {
    lkl_ctx* ctx = nullptr;
    int family = 0;
    unsigned flags = 0;
    std::vector<Swarm> swarms;
    lkl_score_fallback fallback = nullptr;
    void* fallback_user = nullptr;
    long long gpu_launches = 0, fallback_calls = 0, scored = 0;
    std::string err;
    lkl_gsod_buf* dev = nullptr;  // device-resident step buffers
    ~lkl_gso() { lkl_gsod_free(dev); }
};

static unsigned family_flags(int family) {
    switch (family) {
        case 0: return LKL_FLAGS_DNA;
        case 1: return LKL_FLAGS_VDW;
        case 2: return LKL_FLAGS_PYDOCK;
        case 3: return LKL_FLAGS_PYDOCK | LKL_F_DESOLV;
        default: return LKL_FLAGS_DNA;
    }
}

// gpu_score.rs:326-331 / gpu_family.rs:283-288; cpydock.rs:789-791.
static double fold(int family, double e, double v, double s) {
#pragma clang fp contract(off)
    const double ek = e * 332.0 / 4.0;
    return family == 3 ? -(ek + 0.1 * v) + s : -(ek + v);
}

// dna.rs:989-995 / cpydock.rs:545-549 with scoring.rs:83-110 fractions from
// interface counts (satisfied residues of each side, membrane beads).
static double restraint_fold(double score, const int* cnt, const int* sizes) {
#pragma clang fp contract(off)
    const double perc_r = sizes[0] ? (double)cnt[0] / (double)sizes[0] : 0.0;
    const double perc_l = sizes[1] ? (double)cnt[1] / (double)sizes[1] : 0.0;
    const double inter = sizes[2] ? (double)cnt[2] / (double)sizes[2] : 0.0;
    const double penalty = inter > 0.0 ? 999.0 * inter : 0.0;
    score += perc_r * score + perc_l * score - penalty;
    return score;
}

// Record a failure of the driver: kept in G->err and reported through
// lkl_last_error() like every other entry point's.
static int gso_fail(lkl_gso* G, const std::string& msg) {
    G->err = msg;
    lkl_set_error_text(msg.c_str());
    return -1;
}

// Score `poses` (n x 7) into `scores` (GPU first, then the fallback if any).
static int score_batch(lkl_gso* G, const std::vector<double>& poses, int n, std::vector<double>& scores) {
    scores.assign(n, 0.0);
    if (n == 0) return 0;
    if (G->ctx) {
        std::vector<double> out((size_t)n * 2), s((size_t)n);
        const bool iface = lkl_ctx_has_interface(G->ctx) == 1;
        std::vector<int> cnt(iface ? (size_t)n * 3 : 0);
        int sizes[3] = {0, 0, 0};
        if (lkl_ctx_score(G->ctx, poses.data(), n, G->flags, out.data(),
                          (G->flags & LKL_F_DESOLV) ? s.data() : nullptr) == 0 &&
            (!iface || (lkl_ctx_interface(G->ctx, poses.data(), n, cnt.data()) == 0 &&
                        lkl_ctx_interface_sizes(G->ctx, &sizes[0], &sizes[1], &sizes[2]) == 0))) {
            for (int k = 0; k < n; ++k) {
                scores[k] = fold(G->family, out[2 * k], out[2 * k + 1], s[k]);
                if (iface) scores[k] = restraint_fold(scores[k], &cnt[3 * (size_t)k], sizes);
            }
            ++G->gpu_launches;
            G->scored += n;
            return 0;
        }
        // dna.rs:1081-1085: a failed GPU call is re-scored on the CPU grid
        // path when a fallback is registered; otherwise the GPU's message is
        // the run's error
        if (!G->fallback)
            return gso_fail(G, std::string("GPU scoring failed and no fallback is registered: ") +
                                   lkl_last_error());
    }
    if (!G->fallback) return gso_fail(G, "no scorer: no GPU context and no fallback");
    ++G->fallback_calls;
    if (G->fallback(G->fallback_user, poses.data(), n, scores.data()) != 0)
        return gso_fail(G, "the registered CPU fallback scorer failed");
    return 0;
}

// The pose set of dna.rs:497-519 (DNA::dump_metal_dataset): per pose, three
// gen::<f64>() for the translation (x 40 - 20) then quaternion candidates of
// four gen::<f64>() (x 2 - 1, order w, x, y, z) until the squared norm exceeds
// 1e-12, normalised by 1 / sqrt(norm2).  out: n x 7 (w, x, y, z, tx, ty, tz).
LKL_API int lkl_dataset_poses(unsigned long long seed, int n, double* out) {
#pragma clang fp contract(off)
    if (n < 0 || (n > 0 && !out)) return -1;
    StdRng rng(seed);
    for (int k = 0; k < n; ++k) {
        double t[3];
        for (double& v : t) v = rng.next_f64() * 40.0 - 20.0;
        double q[4];
        for (;;) {
            for (double& v : q) v = rng.next_f64() * 2.0 - 1.0;
            const double n2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
            if (n2 > 1e-12) {
                const double inv = 1.0 / std::sqrt(n2);
                for (double& v : q) v = v * inv;
                break;
            }
        }
        double* o = out + 7 * (size_t)k;
        o[0] = q[0]; o[1] = q[1]; o[2] = q[2]; o[3] = q[3];
        o[4] = t[0]; o[5] = t[1]; o[6] = t[2];
    }
    return 0;
}

LKL_API lkl_gso* lkl_gso_create(lkl_ctx* ctx, int family, int n_swarms, const int* swarm_sizes,
                                const double* positions, unsigned long long seed,
                                lkl_score_fallback fallback, void* fallback_user) {
    if (family < 0 || family > 3 || n_swarms <= 0 || !swarm_sizes || !positions) return nullptr;
    if (!ctx && !fallback) return nullptr;
    lkl_gso* G = new lkl_gso();
    G->ctx = ctx;
    G->family = family;
    G->flags = family_flags(family);
    G->fallback = fallback;
    G->fallback_user = fallback_user;
    size_t off = 0;
    for (int s = 0; s < n_swarms; ++s) {
        if (swarm_sizes[s] <= 0) { delete G; return nullptr; }
        G->swarms.emplace_back(seed);  // every swarm seeds its own StdRng (simulator.rs:184, lib.rs:58)
        for (int i = 0; i < swarm_sizes[s]; ++i, ++off) {
            const double* p = positions + 7 * off;  // tx, ty, tz, w, x, y, z (swarm.rs:40-42)
            Glowworm w;
            w.t = {p[0], p[1], p[2]};
            w.q = {p[3], p[4], p[5], p[6]};
            G->swarms.back().g.push_back(w);
        }
    }
    return G;
}

LKL_API void lkl_gso_destroy(lkl_gso* G) { delete G; }

// swarm.rs:70-103 (batched branch) for every swarm, one scoring call in total.
LKL_API int lkl_gso_update_luciferin(lkl_gso* G) {
    if (!G) return -1;
    std::vector<double> poses;
    std::vector<Glowworm*> need;
    for (Swarm& sw : G->swarms)
        for (Glowworm& w : sw.g)
            if (w.moved || w.step == 0) {
                need.push_back(&w);
                const double p[7] = {w.q.w, w.q.x, w.q.y, w.q.z, w.t[0], w.t[1], w.t[2]};
                poses.insert(poses.end(), p, p + 7);
            }
    std::vector<double> scores;
    if (score_batch(G, poses, (int)need.size(), scores)) return -1;
    for (size_t k = 0; k < need.size(); ++k) need[k]->scoring = scores[k];
    for (Swarm& sw : G->swarms)
        for (Glowworm& w : sw.g) {
            w.luciferin = (1.0 - kRho) * w.luciferin + kGamma * w.scoring;
            w.step += 1;
        }
    return 0;
}

LKL_API int lkl_gso_movement_phase(lkl_gso* G) {
    if (!G) return -1;
    for (Swarm& sw : G->swarms) sw.movement_phase();
    return 0;
}

// lib.rs:66-78 without the periodic file output (see lkl_gso_save).
LKL_API int lkl_gso_run(lkl_gso* G, int steps) {
    if (!G || steps < 0) return -1;
    for (int s = 0; s < steps; ++s) {
        if (lkl_gso_update_luciferin(G)) return -1;
        lkl_gso_movement_phase(G);
    }
    return 0;
}

// swarm.rs:70-174 for `steps` full steps (luciferin, then movement) with the
// state on the device: see lkl_gsod.hip.  The swarms' StdRng draws are made
// here, in the order the host movement makes them, and the state returns to
// the host glowworms at the end (neighbour ids are not kept, only their count).
LKL_API int lkl_gso_device_run(lkl_gso* G, int steps, void* stream) {
    if (!G || steps < 0 || !G->ctx) return -1;
    if (steps == 0) return 0;
    int n = 0;
    std::vector<int> off(1, 0);
    for (const Swarm& sw : G->swarms) {
        n += (int)sw.g.size();
        off.push_back(n);
    }
    const size_t N = (size_t)n;
    std::vector<double> pose(7 * N), luc(N), vis(N), score(N), rnd((size_t)steps * N);
    std::vector<unsigned char> moved(N);
    std::vector<unsigned> step(N);
    std::vector<int> nbc(N);
    size_t k = 0;
    for (const Swarm& sw : G->swarms)
        for (const Glowworm& w : sw.g) {
            double* p = &pose[7 * k];
            p[0] = w.q.w; p[1] = w.q.x; p[2] = w.q.y; p[3] = w.q.z;
            p[4] = w.t[0]; p[5] = w.t[1]; p[6] = w.t[2];
            luc[k] = w.luciferin;
            vis[k] = w.vision;
            score[k] = w.scoring;
            moved[k] = w.moved ? 1 : 0;
            step[k] = w.step;
            ++k;
        }
    for (int s = 0; s < steps; ++s) {  // serial draws per swarm, id order (swarm.rs:165)
        size_t b = (size_t)s * N;
        for (Swarm& sw : G->swarms)
            for (size_t i = 0; i < sw.g.size(); ++i) rnd[b++] = sw.rng.next_f64();
    }
    char err[256] = {0};
    if (lkl_gsod_run(&G->dev, G->ctx, G->family, G->flags, n, (int)G->swarms.size(), off.data(), pose.data(),
                     luc.data(), vis.data(), score.data(), moved.data(), step.data(), nbc.data(), rnd.data(), steps,
                     stream, err, sizeof err)) {
        G->err = err;
        fprintf(stderr, "[lklight-gpu] %s\n", err);
        return -1;
    }
    k = 0;
    for (Swarm& sw : G->swarms)
        for (Glowworm& w : sw.g) {
            const double* p = &pose[7 * k];
            w.q = {p[0], p[1], p[2], p[3]};
            w.t = {p[4], p[5], p[6]};
            w.luciferin = luc[k];
            w.vision = vis[k];
            w.scoring = score[k];
            w.moved = moved[k] != 0;
            w.step = step[k];
            w.nb.assign((size_t)nbc[k], -1);
            ++k;
        }
    G->gpu_launches += steps;
    G->scored += (long long)steps * n;
    return 0;
}

LKL_API int lkl_gso_size(const lkl_gso* G) {
    if (!G) return -1;
    int n = 0;
    for (const Swarm& sw : G->swarms) n += (int)sw.g.size();
    return n;
}

// Per glowworm (all swarms, in order): pose (w, x, y, z, tx, ty, tz), luciferin,
// last score, vision range, neighbour count, moved flag.  Any pointer may be NULL.
LKL_API int lkl_gso_state(const lkl_gso* G, double* pose7, double* luciferin, double* scoring,
                          double* vision, int* n_neighbors, unsigned char* moved) {
    if (!G) return -1;
    size_t k = 0;
    for (const Swarm& sw : G->swarms)
        for (const Glowworm& w : sw.g) {
            if (pose7) {
                double* p = pose7 + 7 * k;
                p[0] = w.q.w; p[1] = w.q.x; p[2] = w.q.y; p[3] = w.q.z;
                p[4] = w.t[0]; p[5] = w.t[1]; p[6] = w.t[2];
            }
            if (luciferin) luciferin[k] = w.luciferin;
            if (scoring) scoring[k] = w.scoring;
            if (vision) vision[k] = w.vision;
            if (n_neighbors) n_neighbors[k] = (int)w.nb.size();
            if (moved) moved[k] = w.moved ? 1 : 0;
            ++k;
        }
    return 0;
}

// Overwrite glowworm poses (w, x, y, z, tx, ty, tz) and mark them moved, so
// the next update_luciferin re-scores them (used by screened moves).
LKL_API int lkl_gso_set_poses(lkl_gso* G, const double* pose7, const unsigned char* mask) {
    if (!G || !pose7) return -1;
    size_t k = 0;
    for (Swarm& sw : G->swarms)
        for (Glowworm& w : sw.g) {
            if (!mask || mask[k]) {
                const double* p = pose7 + 7 * k;
                w.q = {p[0], p[1], p[2], p[3]};
                w.t = {p[4], p[5], p[6]};
                w.moved = true;
            }
            ++k;
        }
    return 0;
}

// Screened movement, first half: plan every swarm's GSO moves with the same
// draws and vision update as lkl_gso_movement_phase, and write the planned
// poses (rows w, x, y, z, tx, ty, tz; the current pose for a glowworm with no
// brighter neighbour) without moving anything.
LKL_API int lkl_gso_plan(lkl_gso* G, double* moves) {
    if (!G || !moves) return -1;
    size_t k = 0;
    for (Swarm& sw : G->swarms) {
        const std::vector<Move> mv = sw.plan_moves();
        for (size_t i = 0; i < sw.g.size(); ++i, ++k) {
            sw.g[i].moved = false;
            double* p = moves + 7 * k;
            p[0] = mv[i].q.w; p[1] = mv[i].q.x; p[2] = mv[i].q.y; p[3] = mv[i].q.z;
            p[4] = mv[i].t[0]; p[5] = mv[i].t[1]; p[6] = mv[i].t[2];
        }
    }
    return 0;
}

// Screened movement, second half: glowworms with accept[k] set take pose7[k]
// and its (exact) score; the others keep pose and score.  Nothing is marked
// moved, so the next update_luciferin scores nothing.
LKL_API int lkl_gso_commit(lkl_gso* G, const double* pose7, const double* scores, const unsigned char* accept) {
    if (!G || !pose7 || !scores || !accept) return -1;
    size_t k = 0;
    for (Swarm& sw : G->swarms)
        for (Glowworm& w : sw.g) {
            if (accept[k]) {
                const double* p = pose7 + 7 * k;
                w.q = {p[0], p[1], p[2], p[3]};
                w.t = {p[4], p[5], p[6]};
                w.scoring = scores[k];
            }
            w.moved = false;
            ++k;
        }
    return 0;
}

// Scoring statistics: GPU launches, fallback calls, poses scored on the GPU.
LKL_API int lkl_gso_stats(const lkl_gso* G, long long* gpu_launches, long long* fallback_calls,
                          long long* scored) {
    if (!G) return -1;
    if (gpu_launches) *gpu_launches = G->gpu_launches;
    if (fallback_calls) *fallback_calls = G->fallback_calls;
    if (scored) *scored = G->scored;
    return 0;
}

// swarm.rs:232-271: <directory>/gso_<step>.out for one swarm.
LKL_API int lkl_gso_save(const lkl_gso* G, int swarm, int step, const char* directory) {
    if (!G || swarm < 0 || swarm >= (int)G->swarms.size() || !directory) return -1;
    const std::string path = std::string(directory) + "/gso_" + std::to_string(step) + ".out";
    FILE* f = fopen(path.c_str(), "w");
    if (!f) return -1;
    fprintf(f, "#Coordinates  RecID  LigID  Luciferin  Neighbor's number  Vision Range  Scoring\n");
    for (const Glowworm& w : G->swarms[swarm].g)
        fprintf(f, "(%.7f, %.7f, %.7f, %.7f, %.7f, %.7f, %.7f)    0    0   %.8f  %d %.3f %.8f\n",
                w.t[0], w.t[1], w.t[2], w.q.w, w.q.x, w.q.y, w.q.z, w.luciferin, (int)w.nb.size(),
                w.vision, w.scoring);
    return fclose(f) == 0 ? 0 : -1;
}

// rand 0.7.3 StdRng gen::<f64>() draws (exposed for the qt.rs:474-487 KAT).
LKL_API void lkl_rng_f64s(unsigned long long seed, int n, double* out) {
    StdRng r(seed);
    for (int i = 0; i < n; ++i) out[i] = r.next_f64();
}

// qt.rs:80-104 slerp of two (w, x, y, z) quaternions.
LKL_API void lkl_quat_slerp(const double* a, const double* b, double t, double* out) {
    const Quat r = Quat{a[0], a[1], a[2], a[3]}.slerp(Quat{b[0], b[1], b[2], b[3]}, t);
    out[0] = r.w; out[1] = r.x; out[2] = r.y; out[3] = r.z;
}
