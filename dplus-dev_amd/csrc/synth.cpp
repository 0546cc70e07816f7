// synth.cpp -- seeded synthetic structures for tests and benchmarks: a
// self-avoiding random walk of fixed step confined to an ellipsoid, with a
// minimum separation between all atoms (cell list).  When a walker is stuck
// (no admissible step in 200 tries) the walk restarts from a random earlier
// atom, so the chain branches instead of failing.

#include <cmath>
#include <cstdint>
#include <unordered_map>
#include <vector>

#include "device.hpp"

namespace dplus {

namespace {
struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return ((double)(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
};
}  // namespace

void synth_walk(unsigned long long seed, int n, double step, double min_sep, double ax, double ay, double az,
                double *xyz) {
    if (n == 0) return;
    SplitMix rng{seed * 0x2545F4914F6CDD1DULL + 1};
    const double cell = min_sep > 0 ? min_sep : step;
    std::unordered_map<uint64_t, std::vector<int>> cells;
    auto key = [&](long long i, long long j, long long k) {
        return ((uint64_t)(i & 0x1FFFFF) << 42) | ((uint64_t)(j & 0x1FFFFF) << 21) | (uint64_t)(k & 0x1FFFFF);
    };
    auto cidx = [&](double v) { return (long long)std::floor(v / cell); };
    auto inside = [&](double x, double y, double z) {
        return (x / ax) * (x / ax) + (y / ay) * (y / ay) + (z / az) * (z / az) <= 1.0;
    };
    auto free_at = [&](double x, double y, double z) {
        const long long i = cidx(x), j = cidx(y), k = cidx(z);
        for (long long a = i - 1; a <= i + 1; a++)
            for (long long b = j - 1; b <= j + 1; b++)
                for (long long c = k - 1; c <= k + 1; c++) {
                    auto it = cells.find(key(a, b, c));
                    if (it == cells.end()) continue;
                    for (int o : it->second) {
                        const double dx = xyz[3 * o] - x, dy = xyz[3 * o + 1] - y, dz = xyz[3 * o + 2] - z;
                        if (dx * dx + dy * dy + dz * dz < min_sep * min_sep) return false;
                    }
                }
        return true;
    };
    auto place = [&](int a, double x, double y, double z) {
        xyz[3 * a] = x;
        xyz[3 * a + 1] = y;
        xyz[3 * a + 2] = z;
        cells[key(cidx(x), cidx(y), cidx(z))].push_back(a);
    };
    place(0, 0.0, 0.0, 0.0);
    int cur = 0;
    long long stuck = 0;
    for (int a = 1; a < n;) {
        bool ok = false;
        for (int t = 0; t < 200 && !ok; t++) {
            const double cz = 2.0 * rng.uniform() - 1.0, ph = k2Pi * rng.uniform();
            const double sz = std::sqrt(std::max(0.0, 1.0 - cz * cz));
            const double x = xyz[3 * cur] + step * sz * std::cos(ph), y = xyz[3 * cur + 1] + step * sz * std::sin(ph),
                         z = xyz[3 * cur + 2] + step * cz;
            if (inside(x, y, z) && free_at(x, y, z)) {
                place(a, x, y, z);
                cur = a;
                a++;
                ok = true;
            }
        }
        if (!ok) {
            cur = (int)(rng.next() % (uint64_t)a);
            if (++stuck > 1000LL * n + 100000) throw Error(kErrInvalidArgs, "walk cannot place atoms: volume too small");
        }
    }
}

}  // namespace dplus
