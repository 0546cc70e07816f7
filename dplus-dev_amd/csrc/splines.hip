// splines.hip -- periodic cubic-spline phi-derivatives of every theta row
// (JacobianSphereGrid::CalculateSplines, Backend/Backend/Grid.cpp:1113-1151,
// solver EvenlySpacedPeriodicSplineSolver, PeriodicSplineSolver.h:8-67).
//
// The reference eliminates the cyclic tridiagonal system (4 on the diagonal,
// 1 off it and in the corners) with the matrix held in single precision and
// the right-hand side in double.  The matrix part depends only on the row
// length n = 6i, so the host computes, once per shell, every multiplier that
// elimination produces (in the reference's float/double mix); the kernel
// then applies them to the data in double.  The coefficients therefore equal
// the reference's to double rounding.  Intermediate values live in the
// derivative slots (.z, .w) of the node array, so no scratch is needed.

#include <algorithm>
#include <cstdlib>
#include <vector>

#include "common.hpp"

namespace dplus {

// Per-shell table layout (doubles), n = 6i:
//   [0, n)      fwd[i]  forward multipliers (index 1..n-2 used)
//   [n, 2n)     bl[i]   bottom-row multipliers (0..n-2)
//   [2n, 3n)    rr[i]   right-column back-substitution factors (0..n-3)
//   [3n, 4n)    uu[i]   upper-diagonal back-substitution factors (0..n-3)
//   [4n, 5n)    dg[i]   final diagonal
//   5n          vp after the forward sweep
//   5n+1        fac
static void spline_table(int n, double *t) {
    std::vector<float> diag(n, 4.0f), rc(n, 0.0f);
    double *fwd = t, *bl = t + n, *rr = t + 2 * n, *uu = t + 3 * n, *dg = t + 4 * n;
    for (int i = 0; i < 5 * n + 2; i++) t[i] = 0.0;
    rc[0] = 1.0f;
    rc[n - 2] = 1.0f;
    double vp = 0.25, vc;
    for (int i = 1; i < n - 1; i++) {
        fwd[i] = vp;
        vc = 1.0 / ((double)diag[i] - vp);
        diag[i] = (float)((double)diag[i] - vp);
        rc[i] = (float)((double)rc[i] - vp * (double)rc[i - 1]);
        vp = vc;
    }
    t[5 * n] = vp;
    diag[n - 1] = (float)((double)diag[n - 1] - vp * (double)rc[n - 2]);
    float bottom_left = 1.0f;
    for (int i = 0; i < n - 1; i++) {
        float vcf = bottom_left / diag[i];
        bl[i] = (double)vcf;
        diag[n - 1] = (float)((double)diag[n - 1] - bl[i] * (double)rc[i]);
        bottom_left = (float)(-bl[i]);
    }
    t[5 * n + 1] = (double)(rc[n - 2] / diag[n - 1]);
    for (int i = n - 3; i >= 0; i--) {
        rr[i] = (double)(rc[i] / diag[n - 1]);
        uu[i] = 1.0 / (double)diag[i + 1];
    }
    for (int i = 0; i < n; i++) dg[i] = (double)diag[i];
}


// the same five tables for the one-wave kernel, lane-interleaved: with
// m = ceil(n / 64), entry e * 64 + l of each block of 64 m holds the serial
// table's element l m + e (lane l's e-th node), then vp and fac
static size_t wave_table_size(int n) { return 5 * 64 * (size_t)((n + 63) / 64) + 2; }

static void wave_table(int n, const double *t, double *w) {
    const int m = (n + 63) / 64, W = 64 * m;
    for (int b = 0; b < 5; b++)
        for (int l = 0; l < 64; l++)
            for (int e = 0; e < m; e++) {
                const int k = l * m + e;
                w[b * W + e * 64 + l] = k < n ? t[b * n + k] : 0.0;
            }
    w[5 * W] = t[5 * n];
    w[5 * W + 1] = t[5 * n + 1];
}

// tables for shells 1..nshells; off[i] = start of shell i's serial tables,
// off[nshells + 1 + i] = start of its lane-interleaved tables (shells <= 170)
void build_spline_tables(int nshells, std::vector<double> &tab, std::vector<long long> &off) {
    size_t total = 0;
    for (int i = 1; i <= nshells; i++) {
        total += 5 * (6 * (size_t)i) + 2;
        if (i <= kWaveShellMax) total += wave_table_size(6 * i);
    }
    tab.assign(total, 0.0);
    off.assign(2 * (size_t)(nshells + 1), 0);
    size_t o = 0;
    for (int i = 1; i <= nshells; i++) {
        off[i] = (long long)o;
        spline_table(6 * i, tab.data() + o);
        o += 5 * (6 * (size_t)i) + 2;
    }
    for (int i = 1; i <= std::min(nshells, kWaveShellMax); i++) {
        off[nshells + 1 + i] = (long long)o;
        wave_table(6 * i, tab.data() + off[i], tab.data() + o);
        o += wave_table_size(6 * i);
    }
}

// one thread per row (rows row0..): the serial sweeps
template <typename T4, typename T>
__global__ __launch_bounds__(64) void splines_kernel(T4 *__restrict__ g, int nshells, long long row0,
                                                     const double *__restrict__ tab,
                                                     const long long *__restrict__ off) {
    const long long row = row0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    // rows of shells 1..i-1: sum (3s+1) = (i-1)(3i+2)/2
    int i = (int)((sqrt(24.0 * (double)row + 1.0) + 1.0) / 6.0);
    if (i < 1) i = 1;
    while (i > 1 && (long long)(i - 1) * (3 * i + 2) / 2 > row) i--;
    while ((long long)i * (3 * i + 5) / 2 <= row) i++;
    if (i > nshells) return;
    const int j = (int)(row - (long long)(i - 1) * (3 * i + 2) / 2);
    const int n = 6 * i;
    T4 *p = g + shell_base(i) + (long long)j * n;
    const double *t = tab + off[i];
    const double *fwd = t, *bl = t + n, *rr = t + 2 * n, *uu = t + 3 * n, *dg = t + 4 * n;

    // right-hand side 3 (p[i+1] - p[i-1]) and the forward sweep
    double prx, pry;
    {
        T4 a = p[1], b = p[n - 1];
        prx = 3.0 * ((double)a.x - (double)b.x);
        pry = 3.0 * ((double)a.y - (double)b.y);
        p[0].z = (T)prx;
        p[0].w = (T)pry;
    }
    double x0 = prx, y0 = pry;  // as[0], used by the bottom-row sweep
    double lx = x0, ly = y0;
    double bx = 0.0, by = 0.0;  // bottom-row accumulation of bl[i] * as[i]
    bx += bl[0] * x0;
    by += bl[0] * y0;
    // the node values two apart are carried in registers; unrolled so that the
    // loads of later nodes and multipliers (independent of the recurrence)
    // are in flight while earlier steps compute
    T vxm = p[0].x, vym = p[0].y, vx0 = p[1].x, vy0 = p[1].y;
#pragma unroll 8
    for (int k = 1; k < n - 1; k++) {
        const T vx1 = p[k + 1].x, vy1 = p[k + 1].y;
        double ax = 3.0 * ((double)vx1 - (double)vxm) - lx * fwd[k];
        double ay = 3.0 * ((double)vy1 - (double)vym) - ly * fwd[k];
        p[k].z = (T)ax;
        p[k].w = (T)ay;
        bx += bl[k] * ax;
        by += bl[k] * ay;
        lx = ax;
        ly = ay;
        vxm = vx0;
        vym = vy0;
        vx0 = vx1;
        vy0 = vy1;
    }
    // last row
    double nx, ny;
    {
        T4 a = p[0], b = p[n - 2];
        nx = 3.0 * ((double)a.x - (double)b.x);
        ny = 3.0 * ((double)a.y - (double)b.y);
        nx -= t[5 * n] * lx;
        ny -= t[5 * n] * ly;
        // reference order: the bottom sweep subtracts bl[i]*as[i] one by one;
        // the sum is formed above in the same order.
        nx -= bx;
        ny -= by;
    }
    // back substitution
    double fac = t[5 * n + 1];
    double cx = lx - nx * fac, cy = ly - ny * fac;  // as[n-2]
    T4 q;
    q = p[n - 1];
    q.z = (T)(nx / dg[n - 1]);
    q.w = (T)(ny / dg[n - 1]);
    p[n - 1] = q;
    q = p[n - 2];
    q.z = (T)(cx / dg[n - 2]);
    q.w = (T)(cy / dg[n - 2]);
    p[n - 2] = q;
#pragma unroll 8
    for (int k = n - 3; k >= 0; k--) {
        T4 a = p[k];
        double ax = (double)a.z - nx * rr[k] - cx * uu[k];
        double ay = (double)a.w - ny * rr[k] - cy * uu[k];
        a.z = (T)(ax / dg[k]);
        a.w = (T)(ay / dg[k]);
        p[k] = a;
        cx = ax;
        cy = ay;
    }
}

// One wave per theta row (shells whose rows have n = 6i <= 64 M nodes): lane
// l holds the m = ceil(n / 64) consecutive nodes [l m, l m + m) in registers.
// The forward sweep a_k = r_k - fwd_k a_{k-1} (r_k = 3 (v_{k+1} - v_{k-1}),
// a_0 = r_0) and the back substitution c_k = (a_k - n_x rr_k) - uu_k c_{k+1}
// are first-order recurrences: each lane composes the affine maps of its
// chunk, the wave scans the composed maps (6 shuffle steps), and each lane
// then replays its chunk from its incoming value; the bottom-row sum
// sum_k bl_k a_k is a wave sum.  The values agree with the serial sweeps to
// rounding (|fwd_k| < 0.3, so the composition adds no error growth); every
// node is read and written once, the writes coalesced per lane chunk.
struct Aff {  // x -> A + B x, A complex (ax, ay), B real
    double ax, ay, b;
};

__device__ __forceinline__ Aff aff_shfl(const Aff &f, int src) {
    return {__shfl(f.ax, src, 64), __shfl(f.ay, src, 64), __shfl(f.b, src, 64)};
}

__device__ __forceinline__ float2 make_vector2(float a, float b) { return make_float2(a, b); }
__device__ __forceinline__ double2 make_vector2(double a, double b) { return make_double2(a, b); }

// g o f: apply f, then g
__device__ __forceinline__ Aff aff_then(const Aff &f, const Aff &g) {
    return {g.ax + g.b * f.ax, g.ay + g.b * f.ay, g.b * f.b};
}


// One wave (workgroup) per row: the row's values are staged in LDS by one
// coalesced read, the four sweeps run on LDS (lane l on its chunk), and the
// nodes go back by one coalesced write.  LDS: 2 * rowmax * sizeof(T2) bytes
// (values, then a_k overwritten by the coefficients), rowmax = the launch's
// longest row.
template <typename T4, typename T>
__global__ __launch_bounds__(64) void splines_wave(T4 *__restrict__ g, int shell0, long long row0, int rowmax,
                                                   const double *__restrict__ tab,
                                                   const long long *__restrict__ woff) {
    using T2 = decltype(make_vector2(T(0), T(0)));
    extern __shared__ unsigned char lds_raw[];
    T2 *vs = reinterpret_cast<T2 *>(lds_raw), *as = vs + rowmax;
    const int lane = threadIdx.x;
    const long long row = row0 + (long long)blockIdx.x;
    int i = (int)((sqrt(24.0 * (double)row + 1.0) + 1.0) / 6.0);
    if (i < shell0) i = shell0;
    while (i > 1 && (long long)(i - 1) * (3 * i + 2) / 2 > row) i--;
    while ((long long)i * (3 * i + 5) / 2 <= row) i++;
    const int j = (int)(row - (long long)(i - 1) * (3 * i + 2) / 2);
    const int n = 6 * i, m = (n + 63) / 64, s = lane * m;
    const int fend = min(s + m, n - 1), bend = min(s + m, n - 2), wend = min(s + m, n);
    T4 *p = g + shell_base(i) + (long long)j * n;
    // lane-interleaved tables: node k = s + e of this lane at [e * 64 + lane] of each block
    const int W = 64 * m;
    const double *tw = tab + woff[i];
    const double *fwd = tw + lane, *bl = fwd + W, *rr = fwd + 2 * W, *uu = fwd + 3 * W, *dg = fwd + 4 * W;
    const double vp = tw[5 * W], fac = tw[5 * W + 1];
#define TW(a, k) (a)[((k) - s) * 64]
    for (int k = lane; k < n; k += 64) {
        const T4 v = p[k];
        vs[k] = make_vector2(v.x, v.y);
    }
    __syncthreads();
    // r_k = 3 (v_{k+1} - v_{k-1}) (cyclic)
    auto rhs = [&](int k, double &rx, double &ry) {
        const T2 u = vs[k + 1 < n ? k + 1 : 0], v = vs[k > 0 ? k - 1 : n - 1];
        rx = 3.0 * ((double)u.x - (double)v.x);
        ry = 3.0 * ((double)u.y - (double)v.y);
    };
    // forward: the maps of k in [s, min(s + m, n - 1))
    Aff F{0.0, 0.0, 1.0};
#pragma unroll 4
    for (int k = s; k < fend; k++) {
        double rx, ry;
        rhs(k, rx, ry);
        F = aff_then(F, Aff{rx, ry, k == 0 ? 0.0 : -TW(fwd, k)});
    }
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const Aff o = aff_shfl(F, lane - d < 0 ? lane : lane - d);
        if (lane >= d) F = aff_then(o, F);
    }
    // a_{s-1}: the prefix of the lanes before (a_0's map is constant, so its value is A)
    double inx = __shfl(F.ax, lane > 0 ? lane - 1 : 0, 64), iny = __shfl(F.ay, lane > 0 ? lane - 1 : 0, 64);
    double bx = 0.0, by = 0.0;
#pragma unroll 4
    for (int k = s; k < fend; k++) {
        double rx, ry;
        rhs(k, rx, ry);
        const double f = k == 0 ? 0.0 : TW(fwd, k);
        inx = rx - f * inx;
        iny = ry - f * iny;
        as[k] = make_vector2((T)inx, (T)iny);
        bx += TW(bl, k) * inx;
        by += TW(bl, k) * iny;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        bx += __shfl_xor(bx, o, 64);
        by += __shfl_xor(by, o, 64);
    }
    // a_{n-2}: the last value of its owner lane's chunk
    const int own = (n - 2) / m;
    const double lx = __shfl(inx, own, 64), ly = __shfl(iny, own, 64);
    double nx, ny;
    {
        const T2 u = vs[0], v = vs[n - 2];
        nx = 3.0 * ((double)u.x - (double)v.x) - vp * lx - bx;
        ny = 3.0 * ((double)u.y - (double)v.y) - vp * ly - by;
    }
    const double cnx = lx - nx * fac, cny = ly - ny * fac;  // c_{n-2}
    // backward: the maps of k in [s, min(s + m, n - 2)), applied from high k to low
    Aff G{0.0, 0.0, 1.0};
#pragma unroll 4
    for (int k = bend - 1; k >= s; k--) {
        const T2 ak = as[k];
        G = aff_then(G, Aff{(double)ak.x - nx * TW(rr, k), (double)ak.y - ny * TW(rr, k), -TW(uu, k)});
    }
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const Aff o = aff_shfl(G, lane + d > 63 ? lane : lane + d);
        if (lane + d < 64) G = aff_then(o, G);
    }
    // c at the node after this lane's chunk: the maps of the lanes after, applied to c_{n-2}
    const Aff nxt = aff_shfl(G, lane < 63 ? lane + 1 : lane);
    double cx = lane < 63 ? nxt.ax + nxt.b * cnx : cnx, cy = lane < 63 ? nxt.ay + nxt.b * cny : cny;
    // coefficients into the a_k slots (each lane its own), then the nodes out
#pragma unroll 4
    for (int k = wend - 1; k >= s; k--) {
        double ox, oy;
        if (k == n - 1) {
            ox = nx;
            oy = ny;
        } else if (k == n - 2) {
            ox = cnx;
            oy = cny;
        } else {
            const T2 ak = as[k];
            cx = ((double)ak.x - nx * TW(rr, k)) - TW(uu, k) * cx;
            cy = ((double)ak.y - ny * TW(rr, k)) - TW(uu, k) * cy;
            ox = cx;
            oy = cy;
        }
        const double d = TW(dg, k);
        as[k] = make_vector2((T)(ox / d), (T)(oy / d));
    }
    __syncthreads();
    for (int k = lane; k < n; k += 64) {
        const T2 v = vs[k], c = as[k];
        T4 o;
        o.x = v.x;
        o.y = v.y;
        o.z = c.x;
        o.w = c.y;
        p[k] = o;
    }
#undef TW
}

template <typename T4>
__global__ void zero_origin_coeffs(T4 *g) {
    T4 a = g[0];
    a.z = 0;
    a.w = 0;
    g[0] = a;
}

// rows of shells 1..i
static inline long long rows_upto(long long i) { return i * (3 * i + 5) / 2; }

template <typename T4, typename T>
static void launch_rows(T4 *grid, int nshells, const double *tab, const long long *off, hipStream_t s) {
    if (getenv("DPLUS_SPLINES_SERIAL")) {
        const long long rows = rows_upto(nshells);
        splines_kernel<T4, T><<<(unsigned)((rows + 63) / 64), 64, 0, s>>>(grid, nshells, 0, tab, off);
        return;
    }
    // one wave per row for shells up to kWaveShellMax, LDS sized to the longest row
    const int s1 = std::min(nshells, kWaveShellMax);
    const long long r1 = rows_upto(s1);
    const size_t lds = 2 * (size_t)(6 * s1) * sizeof(T) * 2;
    splines_wave<T4, T><<<(unsigned)r1, 64, lds, s>>>(grid, 1, 0, 6 * s1, tab, off + nshells + 1);
    if (nshells > s1) {
        // longer rows: one thread per row
        const long long rows = rows_upto(nshells) - r1;
        splines_kernel<T4, T><<<(unsigned)((rows + 63) / 64), 64, 0, s>>>(grid, nshells, r1, tab, off);
    }
}

void launch_splines(bool fp64, void *grid, int nshells, const double *tab, const long long *off,
                    hipStream_t s) {
    if (fp64) {
        launch_rows<double4, double>((double4 *)grid, nshells, tab, off, s);
        zero_origin_coeffs<double4><<<1, 1, 0, s>>>((double4 *)grid);
    } else {
        launch_rows<float4, float>((float4 *)grid, nshells, tab, off, s);
        zero_origin_coeffs<float4><<<1, 1, 0, s>>>((float4 *)grid);
    }
}

}  // namespace dplus
