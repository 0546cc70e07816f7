// common.hpp -- geometry of D+'s Jacobian spherical reciprocal grid, the
// counter-based direction generator and small helpers shared by the HIP
// kernels and the host code of libdplus_mi355x.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#define DPLUS_HD __host__ __device__ __forceinline__

namespace dplus {

constexpr double kPi = 3.14159265358979323846;
constexpr double k2Pi = 6.28318530717958647692528676656;
// shells whose spline rows (6i nodes <= 1024) take the one-wave-per-row kernel
constexpr int kWaveShellMax = 170;

// ---------------------------------------------------------------------------
// Grid geometry (Backend/Backend/Grid.cpp:169-186, 1078-1111).
// Shell i (1..N, N = gridSize/2 + 3) has 3i+1 theta rows of 6i phi points,
// theta_j = j*pi/(3i), phi_k = 2*pi*k/(6i).  Index 0 is the origin; shells
// are stored in order, theta-major, phi-minor.
// ---------------------------------------------------------------------------
DPLUS_HD int actual_size(int grid_size) { return grid_size / 2 + 3; }

DPLUS_HD long long shell_base(long long i) { return 1 + 6 * (i - 1) * i * i; }

DPLUS_HD long long total_points(int grid_size) {
    long long n = actual_size(grid_size);
    return 1 + 6 * n * (n + 1) * (n + 1);
}

DPLUS_HD long long index_from_indices(int qi, long long ti, long long pi) {
    return qi == 0 ? 0 : shell_base(qi) + ti * 6LL * qi + pi;
}

// Canonical half-shell used by the Friedel-symmetric builders: shell i has
// 3i(3i+1) canonical points h = j*6i + k (rows j < 3i/2 and, for even i, the
// first 3i points of the equator row).  Their antipodes (3i-j, k+3i mod 6i)
// are the other half.  Shells 1..i-1 hold 3 i^2 (i-1) canonical points.
DPLUS_HD long long half_base(long long i) { return 3 * i * i * (i - 1); }

DPLUS_HD int shell_of_half(long long c) {
    long long i = (long long)cbrt((double)c / 3.0);
    if (i < 1) i = 1;
    while (i > 1 && half_base(i) > c) i--;
    while (half_base(i + 1) <= c) i++;
    return (int)i;
}

// Pair units of the builder: (shell i, row j <= 3i/2, triple t < i) owns the
// nodes (j, 3t+p) and their phi + pi partners (j, 3t+p+3i), p = 0..2.  Shell i
// holds i (floor(3i/2) + 1) units; pair_units_before(n) counts shells 1..n.
DPLUS_HD long long pair_units_before(long long n) {
    const long long m = (n + 1) / 2;
    return n * (n + 1) / 2 + (n * (n + 1) * (2 * n + 1) / 2 - m * m) / 2;
}

// rows of shells 1..n: sum of floor(3s/2) + 1
DPLUS_HD long long rows_before(long long n) { return n + (3 * n * (n + 1) / 2 - (n + 1) / 2) / 2; }

DPLUS_HD int shell_of_pair_unit(long long u) {
    long long i = (long long)cbrt(2.0 * (double)u);
    if (i < 1) i = 1;
    while (i > 1 && pair_units_before(i - 1) > u) i--;
    while (pair_units_before(i) <= u) i++;
    return (int)i;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Counter = (c0..c3),
// key = seed.  Two 53-bit uniforms in (0,1) per call.
// ---------------------------------------------------------------------------
DPLUS_HD void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = (uint32_t)p1;
        c[2] = n2;
        c[3] = (uint32_t)p0;
    }
}

DPLUS_HD void uniform_pair(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                           double &u, double &v) {
    uint32_t c[4] = {c0, c1, c2, c3};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    uint64_t a = ((uint64_t)(c[0] >> 6) << 27) | (uint64_t)(c[1] >> 5);
    uint64_t b = ((uint64_t)(c[2] >> 6) << 27) | (uint64_t)(c[3] >> 5);
    u = ((double)a + 0.5) * (1.0 / 9007199254740992.0);
    v = ((double)b + 0.5) * (1.0 / 9007199254740992.0);
}

// Stratified MC directions.  The directions of a chunk of C (a multiple of 256)
// are cut into runs of 64, one wave each; run w of the chunk's W = C / 64 fills
// patch (w / nv, w % nv) of an nu x nv grid of equal-area patches in
// (u, v) = ((1 - cos theta) / 2, phi / 2 pi), nu the largest divisor of W with
// 2 nu^2 <= W.  Every chunk covers the sphere once with equal weight per patch,
// so the chunk mean stays an unbiased estimate of the orientation average, and
// the 64 directions of a wave land on neighbouring grid cells.
DPLUS_HD void mc_strata(int chunk, int &nu, int &nv) {
    const int W = chunk / 64;
    nu = 1;
    for (int a = 1; 2 * a * a <= W; a++)
        if (W % a == 0) nu = a;
    nv = W / nu;
}

DPLUS_HD void mc_stratify(long long d, int chunk, int nu, int nv, double &u, double &v) {
    const long long w = (d % chunk) / 64;
    u = ((double)(w / nv) + u) / (double)nu;
    v = ((double)(w % nv) + v) / (double)nv;
}

constexpr uint32_t kStreamMC = 0x4D43u;
constexpr uint32_t kStreamVegas = 0x5645u;
constexpr uint32_t kStreamAzimuth = 0x415Au;

// ---------------------------------------------------------------------------
// Error codes of the C ABI (values follow Common/Common.h:184-228 where a
// reference code exists).
// ---------------------------------------------------------------------------
enum : int {
    kOk = 0,              // OK
    kErrStopped = 1,      // ERROR_STOPPED
    kErrJobRunning = 2,   // ERROR_JOBRUNNING
    kErrJobNotFound = 3,  // ERROR_JOBNOTFOUND
    kErrFileNotFound = 6, // ERROR_FILENOTFOUND
    kErrIllegalJson = 9,  // ERROR_ILLEGAL_JSON
    kErrOutOfMemory = 13, // ERROR_INSUFFICIENT_MEMORY
    kErrModelNotFound = 14,
    kErrInvalidModelType = 16,
    kErrInvalidParamTree = 19,
    kErrUnimplemented = 20,
    kErrInvalidArgs = 21,
    kErrGeneral = 24,
    kErrNoGpu = 31,       // ERROR_NO_GPU
};

}  // namespace dplus
