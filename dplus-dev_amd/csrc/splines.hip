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

size_t spline_tables_size(int nshells) {
    size_t s = 0;
    for (int i = 1; i <= nshells; i++) s += 5 * (6 * (size_t)i) + 2;
    return s;
}

// tables for shells 1..nshells; offsets[i] = start of shell i (offsets[0] unused)
void build_spline_tables(int nshells, std::vector<double> &tab, std::vector<long long> &off) {
    tab.assign(spline_tables_size(nshells), 0.0);
    off.assign(nshells + 1, 0);
    size_t o = 0;
    for (int i = 1; i <= nshells; i++) {
        off[i] = (long long)o;
        spline_table(6 * i, tab.data() + o);
        o += 5 * (6 * (size_t)i) + 2;
    }
}

template <typename T4, typename T>
__global__ __launch_bounds__(64) void splines_kernel(T4 *__restrict__ g, int nshells,
                                                     const double *__restrict__ tab,
                                                     const long long *__restrict__ off) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    // rows of shells 1..i-1: sum (3s+1) = (i-1)(3i+2)/2
    int i = (int)((sqrt(24.0 * row + 1.0) + 1.0) / 6.0);
    if (i < 1) i = 1;
    while (i > 1 && (long long)(i - 1) * (3 * i + 2) / 2 > row) i--;
    while ((long long)i * (3 * i + 5) / 2 <= row) i++;
    if (i > nshells) return;
    const int j = row - (i - 1) * (3 * i + 2) / 2;
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

template <typename T4>
__global__ void zero_origin_coeffs(T4 *g) {
    T4 a = g[0];
    a.z = 0;
    a.w = 0;
    g[0] = a;
}

void launch_splines(bool fp64, void *grid, int nshells, const double *tab, const long long *off,
                    hipStream_t s) {
    long long rows = (long long)nshells * (3 * nshells + 5) / 2;
    unsigned nb = (unsigned)((rows + 63) / 64);
    if (fp64) {
        splines_kernel<double4, double><<<nb, 64, 0, s>>>((double4 *)grid, nshells, tab, off);
        zero_origin_coeffs<double4><<<1, 1, 0, s>>>((double4 *)grid);
    } else {
        splines_kernel<float4, float><<<nb, 64, 0, s>>>((float4 *)grid, nshells, tab, off);
        zero_origin_coeffs<float4><<<1, 1, 0, s>>>((float4 *)grid);
    }
}

}  // namespace dplus
