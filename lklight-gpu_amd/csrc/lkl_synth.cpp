// lkl_synth.cpp — seeded synthetic-complex generator (host only).
//
// Random sequential addition (RSA) of atoms inside an ellipsoid with a minimum
// pair separation, the construction the benchmark configurations specify.
// Coordinates are quantised to multiples of 2^-10 A, so the f32 arrays handed
// to the GPU and the f64 arrays the CPU grid path works on hold the same
// values.  Deterministic for a given seed (splitmix64 -> xoshiro256**).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#define LKL_API extern "C" __attribute__((visibility("default")))

namespace {

struct Xoshiro {
    uint64_t s[4];
    explicit Xoshiro(uint64_t seed) {
        for (int i = 0; i < 4; ++i) {  // splitmix64 seeding
            uint64_t z = (seed += 0x9e3779b97f4a7c15ULL);
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

double quant(double x) { return std::nearbyint(x * 1024.0) / 1024.0; }

}  // namespace

// Place n atoms uniformly at random in the ellipsoid (x/a)^2+(y/b)^2+(z/c)^2 <= 1,
// rejecting any candidate closer than dmin to an accepted atom.  Returns the
// number placed (== n on success) and the number of candidates drawn.
LKL_API int lkl_synth_rsa(int n, double a, double b, double c, double dmin, uint64_t seed,
                          int64_t max_attempts, double* xyz, int64_t* attempts_out) {
    if (n <= 0 || a <= 0 || b <= 0 || c <= 0 || dmin <= 0) return -1;
    Xoshiro rng(seed);
    const double cell = dmin;
    const int gx = (int)std::ceil(2 * a / cell) + 3, gy = (int)std::ceil(2 * b / cell) + 3,
              gz = (int)std::ceil(2 * c / cell) + 3;
    std::vector<int> head((size_t)gx * gy * gz, -1), next(n, -1);
    const double d2min = dmin * dmin;
    int placed = 0;
    int64_t tries = 0;
    auto key = [&](double x, double y, double z, int& ix, int& iy, int& iz) {
        ix = (int)std::floor((x + a) / cell) + 1;
        iy = (int)std::floor((y + b) / cell) + 1;
        iz = (int)std::floor((z + c) / cell) + 1;
    };
    while (placed < n && tries < max_attempts) {
        ++tries;
        const double x = quant((2 * rng.uniform() - 1) * a);
        const double y = quant((2 * rng.uniform() - 1) * b);
        const double z = quant((2 * rng.uniform() - 1) * c);
        if ((x / a) * (x / a) + (y / b) * (y / b) + (z / c) * (z / c) > 1.0) continue;
        int ix, iy, iz;
        key(x, y, z, ix, iy, iz);
        bool ok = true;
        for (int dz = -1; dz <= 1 && ok; ++dz)
            for (int dy = -1; dy <= 1 && ok; ++dy)
                for (int dx = -1; dx <= 1 && ok; ++dx) {
                    const int jx = ix + dx, jy = iy + dy, jz = iz + dz;
                    if (jx < 0 || jy < 0 || jz < 0 || jx >= gx || jy >= gy || jz >= gz) continue;
                    for (int k = head[((size_t)jz * gy + jy) * gx + jx]; k >= 0; k = next[k]) {
                        const double ex = xyz[3 * k] - x, ey = xyz[3 * k + 1] - y,
                                     ez = xyz[3 * k + 2] - z;
                        if (ex * ex + ey * ey + ez * ez < d2min) { ok = false; break; }
                    }
                }
        if (!ok) continue;
        xyz[3 * placed] = x;
        xyz[3 * placed + 1] = y;
        xyz[3 * placed + 2] = z;
        const size_t h = ((size_t)iz * gy + iy) * gx + ix;
        next[placed] = head[h];
        head[h] = placed;
        ++placed;
    }
    if (attempts_out) *attempts_out = tries;
    return placed;
}
