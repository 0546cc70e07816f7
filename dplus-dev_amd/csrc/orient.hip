// orient.hip -- grid lookups, OA-time docking and orientation averages.
//
// Lookup: GetAmpAtPointInPlaneJacob (Backend/GPU/CommonJacobGPUMethods.cu:69-154)
//   = cubic Hermite in phi on 4 theta rows + 4-point spline in theta, and the
//   4-plane q spline (OAJacob.cu:57-113, HybridOA.cu:78-133; GetSphr
//   Backend/Backend/Grid.cpp:487-548 for single points).
// Docking: AddAmplitudeAtPoint (HybridOA.cu:17-159): for each rotation
//   (R = Rx(a)Ry(b)Rz(c), Amplitude.h:557-573) A_leaf(R^T q) * sum_t e^{i q.t} * scale.
// MC: directions (theta = acos(2u-1), phi = 2 pi v) from a counter-based Philox
//   stream keyed by (seed; direction, shell group), so any launch shape and any
//   batch composition reproduce the same numbers (the reference draws them from
//   a clock-seeded MTGP32, GPUHybridCalc.cu:196-198), stratified per chunk
//   (mc_stratify: a wave's 64 directions share one equal-area patch, so its
//   gathers hit neighbouring grid cells).
// VEGAS: 32x32 (cos theta, phi) bins, bin-volume weighted estimate, per-bin
//   variance marginals and the damped re-binning of HybridOA.cu:181-449.
//
// Node layout: one T4 per grid node = (Re A, Im A, Re dA/dphi-spline, Im ...),
// so the two phi neighbours of a lookup are one contiguous 2*sizeof(T4) read.

#include "common.hpp"
#include "device.hpp"

namespace dplus {

template <typename T>
struct cplx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cplx<T> hermite(cplx<T> p1, cplx<T> p2, cplx<T> d1, cplx<T> d2, T t) {
    T t2 = t * t, t3 = t2 * t;
    cplx<T> r;
    r.x = p1.x + d1.x * t + ((T)3 * (p2.x - p1.x) - (T)2 * d1.x - d2.x) * t2 +
          ((T)2 * (p1.x - p2.x) + d1.x + d2.x) * t3;
    r.y = p1.y + d1.y * t + ((T)3 * (p2.y - p1.y) - (T)2 * d1.y - d2.y) * t2 +
          ((T)2 * (p1.y - p2.y) + d1.y + d2.y) * t3;
    return r;
}

// PeriodicSplineSolver.h:227-234 / CommonJacobGPUMethods.cu:34-63
template <typename T>
__device__ __forceinline__ void four_point(const cplx<T> y[4], cplx<T> &d2, cplx<T> &d3) {
    const T a = (T)(7. / 15.), b = (T)(14. / 15.), c = (T)(4. / 15.), e = (T)(2. / 15.);
    d2.x = -a * (y[1].x - y[0].x) + b * (y[2].x - y[0].x) - c * (y[3].x - y[1].x) + e * (y[3].x - y[2].x);
    d2.y = -a * (y[1].y - y[0].y) + b * (y[2].y - y[0].y) - c * (y[3].y - y[1].y) + e * (y[3].y - y[2].y);
    d3.x = e * (y[1].x - y[0].x) - c * (y[2].x - y[0].x) + b * (y[3].x - y[1].x) - a * (y[3].x - y[2].x);
    d3.y = e * (y[1].y - y[0].y) - c * (y[2].y - y[0].y) + b * (y[3].y - y[1].y) - a * (y[3].y - y[2].y);
}

// one theta-phi plane (shell >= 1) at theta = ut pi, phi = up 2 pi; kZeroTiny
// reproduces the CPU variant's snapping of interpolation fractions below 1e-10
// (Grid.cpp:1242-1245).  The cell and fractions are the reference's
// theta / (pi / 3i) and phi / (2 pi / 6i) (CommonJacobGPUMethods.cu:69-154),
// formed as products with the direction's ut and up; the row and column
// wrap-arounds are branches instead of integer remainders.
template <typename T, typename T4, bool kZeroTiny>
__device__ cplx<T> plane_lookup(const T4 *__restrict__ plane, int shell, double ut, double up) {
    const int phiPoints = 6 * shell, thePoints = 3 * shell;
    const double xt = ut * (double)thePoints, xp = up * (double)phiPoints;
    const int tI = (int)xt;
    int pI = (int)xp;
    double tphd = xp - (double)pI;
    double tthd = xt - (double)tI;
    if (kZeroTiny) {
        if (fabs(tthd) < 1.0e-10) tthd = 0.0;
        if (fabs(tphd) < 1.0e-10) tphd = 0.0;
    }
    const T tph = (T)tphd, tth = (T)tthd;
    pI = (pI == phiPoints) ? 0 : pI;
    const int pI1 = (pI + 1 == phiPoints) ? 0 : pI + 1;
    const int half = phiPoints / 2;
    cplx<T> nn[4];
#pragma unroll
    for (int rl = -1; rl < 3; rl++) {
        const int v = tI + rl;  // in [-1, thePoints + 2]
        int pos1, pos2;
        if (v < 0 || v > thePoints) {
            // past a pole: the row on the other side, half a turn away in phi
            const int theInd = v < 0 ? v + thePoints + 1 : v - (thePoints + 1);
            pos1 = phiPoints * theInd + pI + ((2 * pI >= phiPoints) ? -half : half);
            pos2 = phiPoints * theInd + pI1 + ((2 * pI1 >= phiPoints) ? -half : half);
        } else {
            pos1 = phiPoints * v + pI;
            pos2 = phiPoints * v + pI1;
        }
        const T4 a = plane[pos1], b = plane[pos2];
        nn[rl + 1] = hermite<T>({a.x, a.y}, {b.x, b.y}, {a.z, a.w}, {b.z, b.w}, tph);
    }
    cplx<T> d1, d2;
    four_point<T>(nn, d1, d2);
    return hermite<T>(nn[1], nn[2], d1, d2, tth);
}

constexpr double kInvPi = 0.31830988618379067153776752674503;
constexpr double kInv2Pi = 0.15915494309189533576888376337251;

__device__ __forceinline__ void antipode(double th, double ph, double &nth, double &nph) {
    double a = kPi - th, b = ph + kPi;
    if (b < 0) b = k2Pi + (b - (int)(b / k2Pi) * k2Pi);
    if (b >= k2Pi) b = b - (int)(b / k2Pi) * k2Pi;
    nth = fmin(fmax(a, 0.0), kPi);
    nph = b;
}

// planes layer-1 .. layer+2 and the q-spline derivatives (OAJacob.cu:57-98)
template <typename T, typename T4>
__device__ __forceinline__ void group_planes(const T4 *__restrict__ g, int layer, double th, double ph,
                                             cplx<T> amps[4], cplx<T> &d1, cplx<T> &d2) {
    const double ut = th * kInvPi, up = ph * kInv2Pi;
#pragma unroll
    for (int i = -1; i <= 2; i++) {
        const int lqi = i + layer;
        if (lqi == -1) {
            double nth, nph;
            antipode(th, ph, nth, nph);
            amps[i + 1] = plane_lookup<T, T4, false>(g + 1, 1, nth * kInvPi, nph * kInv2Pi);
        } else if (lqi == 0) {
            const T4 o = g[0];
            amps[i + 1] = {o.x, o.y};
        } else {
            amps[i + 1] = plane_lookup<T, T4, false>(g + shell_base(lqi), lqi, ut, up);
        }
    }
    four_point<T>(amps, d1, d2);
}

// R^T q with R = Rx(a) Ry(b) Rz(c): HybridOA.cu:60-63
__device__ __forceinline__ void rotate_t(const RotDev &r, const double q[3], double o[3]) {
#pragma unroll
    for (int i = 0; i < 3; i++) o[i] = r.m[3 * i] * q[0] + r.m[3 * i + 1] * q[1] + r.m[3 * i + 2] * q[2];
}

// phase factor e^{i q d} for the docking translations: FP64 argument, reduced
// to revolutions in FP64 and evaluated by the hardware sin/cos in mixed mode
template <typename T>
__device__ __forceinline__ void phase_sincos(double qd, T &s, T &c);
template <>
__device__ __forceinline__ void phase_sincos<float>(double qd, float &s, float &c) {
    double r = qd * (1.0 / k2Pi);
    float f = (float)(r - floor(r));
    s = __builtin_amdgcn_sinf(f);
    c = __builtin_amdgcn_cosf(f);
}
template <>
__device__ __forceinline__ void phase_sincos<double>(double qd, double &s, double &c) {
    sincos(qd, &s, &c);
}

constexpr int kMaxM = 16;  // q-values of one shell group handled per pass (one lookup serves all)


// accumulate the docked amplitude along unit direction qv for q-values qs[0..m), m <= M.
// stepq (equally spaced q, single precision): the translations of a rotation
// with several step their phases in q, e^{i q_j d} = e^{i q_0 d} (e^{i dq d})^j:
// two range-reduced sin/cos per translation and one complex rotation per further
// q instead of one sin/cos per q (a single translation keeps the direct form)
template <typename T, typename T4, int M = kMaxM>
__device__ __forceinline__ void docked_amplitude(const T4 *const *__restrict__ grids,
                                                 const RotDev *__restrict__ rots, int nrot,
                                                 const double4 *__restrict__ trans, int layer,
                                                 double step, const double qv[3],
                                                 const double *__restrict__ qs, int m, bool stepq,
                                                 cplx<T> acc[M]) {
    const double inv_step = 1.0 / step;
    const bool rec = sizeof(T) == 4 && stepq && m > 1;
    const double dq = rec ? (qs[m - 1] - qs[0]) / (double)(m - 1) : 0.0;
    for (int r = 0; r < nrot; r++) {
        const RotDev rd = rots[r];
        double rv[3];
        rotate_t(rd, qv, rv);
        if (fabs(rv[2]) > 1.) rv[2] = (rv[2] > 0.) ? 1. : -1.;
        const double nth = acos(rv[2]);
        double nph = atan2(rv[1], rv[0]);
        if (nph < 0.0) nph += k2Pi;
        cplx<T> amps[4], d1, d2;
        group_planes<T, T4>(grids[rd.leaf], layer, nth, nph, amps, d1, d2);
        const T sc = (T)rd.scale;
        const bool trivial = (rd.nt == 1) && trans[rd.t0].x == 0.0 && trans[rd.t0].y == 0.0 &&
                             trans[rd.t0].z == 0.0;
        if (rec && rd.nt > 1) {
            cplx<T> a[M];
#pragma unroll
            for (int j = 0; j < M; j++) {
                a[j] = {(T)0, (T)0};
                if (j < m) {
                    const T t = (T)((qs[j] - layer * step) * inv_step);
                    a[j] = hermite<T>(amps[1], amps[2], d1, d2, t);
                    a[j].x *= sc;
                    a[j].y *= sc;
                }
            }
            for (int tr = 0; tr < rd.nt; tr++) {
                const double4 tt = trans[rd.t0 + tr];
                const double d = tt.x * qv[0] + tt.y * qv[1] + tt.z * qv[2];
                T cs, sn, cd, sd;
                phase_sincos<T>(qs[0] * d, sn, cs);
                phase_sincos<T>(dq * d, sd, cd);
#pragma unroll
                for (int j = 0; j < M; j++) {
                    if (j < m) {
                        acc[j].x += a[j].x * cs - a[j].y * sn;
                        acc[j].y += a[j].y * cs + a[j].x * sn;
                        const T c2 = cs * cd - sn * sd;
                        sn = sn * cd + cs * sd;
                        cs = c2;
                    }
                }
            }
            continue;
        }
#pragma unroll
        for (int j = 0; j < M; j++) {
            if (j < m) {
                const T t = (T)((qs[j] - layer * step) * inv_step);
                cplx<T> a = hermite<T>(amps[1], amps[2], d1, d2, t);
                if (trivial) {
                    acc[j].x += a.x * sc;
                    acc[j].y += a.y * sc;
                } else {
                    T sumc = 0, sums = 0;
                    for (int tr = 0; tr < rd.nt; tr++) {
                        const double4 tt = trans[rd.t0 + tr];
                        const double qd = qs[j] * (tt.x * qv[0] + tt.y * qv[1] + tt.z * qv[2]);
                        T s, c;
                        phase_sincos<T>(qd, s, c);
                        sumc += c;
                        sums += s;
                    }
                    acc[j].x += (a.x * sumc - a.y * sums) * sc;
                    acc[j].y += (a.y * sumc + a.x * sums) * sc;
                }
            }
        }
    }
}

// the same for one leaf placed once, unrotated, at the origin: the lookup takes
// the direction's angles as drawn (no rotation back to angles, no translation
// phases), a = scale * A(q, theta, phi)
template <typename T, typename T4, int M = kMaxM>
__device__ __forceinline__ void single_amplitude(const T4 *__restrict__ grid, T sc, int layer, double step,
                                                 double th, double ph, const double *__restrict__ qs, int m,
                                                 cplx<T> acc[M]) {
    const double inv_step = 1.0 / step;
    cplx<T> amps[4], d1, d2;
    group_planes<T, T4>(grid, layer, th, ph, amps, d1, d2);
#pragma unroll
    for (int j = 0; j < M; j++) {
        if (j < m) {
            const T t = (T)((qs[j] - layer * step) * inv_step);
            cplx<T> a = hermite<T>(amps[1], amps[2], d1, d2, t);
            acc[j].x += a.x * sc;
            acc[j].y += a.y * sc;
        }
    }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// ------------------------------------------------------------------ MC
// grid: x = direction block, y = active group.  mstride >= every group's m.
// partial[(gi * nblk + bx) * mstride + j] = sum over the block's directions of |A_j|^2
template <typename T, typename T4, bool kSingle, int M = kMaxM>
__global__ __launch_bounds__(256, sizeof(T) == 4 ? 1 : 2) void mc_kernel(const T4 *const *__restrict__ grids,
                                                 const RotDev *__restrict__ rots, int nrot,
                                                 const double4 *__restrict__ trans,
                                                 const GroupDev *__restrict__ groups,
                                                 const int *__restrict__ active, int mstride,
                                                 const double *__restrict__ qs, double step,
                                                 uint64_t seed, long long dir0, int nblk, int chunk, int nu,
                                                 int nv, double *__restrict__ partial) {
    const int gi = blockIdx.y;
    const GroupDev G = groups[active[gi]];
    const long long d = dir0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double u, v;
    uniform_pair(seed, (uint32_t)d, (uint32_t)(d >> 32), (uint32_t)G.layer, kStreamMC, u, v);
    mc_stratify(d, chunk, nu, nv, u, v);
    const double th = acos(2.0 * u - 1.0), ph = k2Pi * v;
    double qv[3] = {0.0, 0.0, 0.0};
    if (!kSingle) {
        double st, ct, sp, cp;
        sincos(th, &st, &ct);
        sincos(ph, &sp, &cp);
        qv[0] = cp * st;
        qv[1] = sp * st;
        qv[2] = ct;
    }
    __shared__ double red[4][kMaxM];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (kSingle) {
        // one lookup per direction serves every q of the group, each q reduced as it is formed
        cplx<T> amps[4], d1, d2;
        group_planes<T, T4>(grids[rots[0].leaf], G.layer, th, ph, amps, d1, d2);
        const T sc = (T)rots[0].scale;
        const double inv_step = 1.0 / step;
        for (int jb = 0; jb < G.m; jb += kMaxM) {
            const int m = min(kMaxM, G.m - jb);
#pragma unroll
            for (int j = 0; j < kMaxM; j++) {
                if (j < m) {
                    const T t = (T)((qs[G.q0 + jb + j] - G.layer * step) * inv_step);
                    const cplx<T> a = hermite<T>(amps[1], amps[2], d1, d2, t);
                    const T ax = a.x * sc, ay = a.y * sc;
                    double I = (double)ax * (double)ax + (double)ay * (double)ay;
                    I = wave_sum(I);
                    if (lane == 0) red[w][j] = I;
                }
            }
            __syncthreads();
            if (threadIdx.x < m) {
                const int j = threadIdx.x;
                partial[((size_t)gi * nblk + blockIdx.x) * mstride + jb + j] =
                    red[0][j] + red[1][j] + red[2][j] + red[3][j];
            }
            __syncthreads();
        }
        return;
    }
    // docked: M q per pass (8 when no group holds more: fewer accumulators)
    for (int jb = 0; jb < G.m; jb += M) {
        const int m = min(M, G.m - jb);
        cplx<T> acc[M];
#pragma unroll
        for (int j = 0; j < M; j++) acc[j] = {(T)0, (T)0};
        docked_amplitude<T, T4, M>(grids, rots, nrot, trans, G.layer, step, qv, qs + G.q0 + jb, m, G.uniform != 0,
                                   acc);
#pragma unroll
        for (int j = 0; j < M; j++) {
            double I = (double)acc[j].x * (double)acc[j].x + (double)acc[j].y * (double)acc[j].y;
            I = wave_sum(I);
            if (lane == 0) red[w][j] = I;
        }
        __syncthreads();
        if (threadIdx.x < m) {
            const int j = threadIdx.x;
            partial[((size_t)gi * nblk + blockIdx.x) * mstride + jb + j] =
                red[0][j] + red[1][j] + red[2][j] + red[3][j];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ MC, copies turned about z
// lane i's value of v, for a wave-uniform i
__device__ __forceinline__ double lane_value(double v, int i) {
    const unsigned long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, i);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), i);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// the sum over the wave's 64 lanes, in every lane: DPP exchanges within each row
// of 16 lanes (quad swaps, half-row and row mirrors), then the four row sums
__device__ __forceinline__ float wave_total(float v) {
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xF, 0xF, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x140, 0xF, 0xF, false));
    const int b = __float_as_int(v);
    return (__int_as_float(__builtin_amdgcn_readlane(b, 0)) + __int_as_float(__builtin_amdgcn_readlane(b, 16))) +
           (__int_as_float(__builtin_amdgcn_readlane(b, 32)) + __int_as_float(__builtin_amdgcn_readlane(b, 48)));
}
__device__ __forceinline__ double wave_total(double v) { return wave_sum(v); }

// weights of the 4-point spline (four_point's end derivatives, then the cubic
// Hermite at t) over values y0..y3: the spline at t is sum_r w[r] y[r]
template <typename T>
__device__ __forceinline__ void spline4_weights(T t, T w[4]) {
    const T a = (T)(7. / 15.), b = (T)(14. / 15.), c = (T)(4. / 15.), e = (T)(2. / 15.);
    const T t2 = t * t, t3 = t2 * t;
    const T h00 = (T)1 - (T)3 * t2 + (T)2 * t3, h01 = (T)3 * t2 - (T)2 * t3, h10 = t - (T)2 * t2 + t3,
            h11 = t3 - t2;
    w[0] = h10 * (a - b) + h11 * (c - e);
    w[1] = h00 + h10 * (c - a) + h11 * (e - b);
    w[2] = h01 + h10 * (b - e) + h11 * (a - c);
    w[3] = h10 * (e - c) + h11 * (b - a);
}

// the four planes' amplitudes (layer - 1 .. layer + 2) of one lookup at (th, ph),
// as group_planes takes them, without the q derivatives
template <typename T, typename T4>
__device__ __forceinline__ void direct_planes(const T4 *__restrict__ g, int layer, double th, double ph,
                                              cplx<T> amps[4]) {
#pragma unroll
    for (int p = 0; p < 4; p++) {
        const int lqi = layer - 1 + p;
        if (lqi == -1) {
            double nth, nph;
            antipode(th, ph, nth, nph);
            amps[p] = plane_lookup<T, T4, false>(g + 1, 1, nth * kInvPi, nph * kInv2Pi);
        } else if (lqi == 0) {
            const T4 o = g[0];
            amps[p] = {o.x, o.y};
        } else {
            amps[p] = plane_lookup<T, T4, false>(g + shell_base(lqi), lqi, th * kInvPi, ph * kInv2Pi);
        }
    }
}

// direct_planes kept out of line, for the copy-major kernel's rare chunks whose
// segments do not fit: its registers then cost only on that path
template <typename T, typename T4>
__device__ __noinline__ void direct_planes_call(const T4 *__restrict__ g, int layer, double th, double ph,
                                                cplx<T> *amps) {
    direct_planes<T, T4>(g, layer, th, ph, amps);
}

constexpr int kZWin = 64;  // nodes of one plane's collapsed row segment (one per lane)

// The MC average of a docking whose copies all turn about z (a helix): every
// copy of a direction looks up the direction's own theta rows, at phi - alpha.
// A block takes the same 256 directions (and writes the same partial sums) as
// mc_kernel, but a wave's lanes are copies: it walks its 64 directions one at a
// time (each broadcast from the lane that drew it) and, per direction, its
// copies 64 at a time in (leaf, alpha) order, so the copies of one chunk span a
// short segment of each theta row.  kCollapse (FP32 grids): the lookup is
// linear in the node data and theta is the same for every copy, so per plane
// the 4-row theta spline is applied once to the segment (at most kZWin nodes,
// one per lane; the row loads of two planes, 8 per lane, issued together and
// coalesced) into LDS, and
// each copy takes the phi Hermite of two collapsed nodes; a segment that does
// not fit, or a chunk spanning two leaves, takes the direct lookup, which
// every copy takes without kCollapse.  The q spline is applied with per-q
// weights.  The copies' amplitudes are summed over the lanes per direction,
// then |A_j|^2.
template <typename T, typename T4, bool kCollapse, int M>
__global__ __launch_bounds__(256, sizeof(T) == 4 ? 2 : 1) void mc_zcopy_kernel(
    const T4 *const *__restrict__ grids, const ZCopyDev *__restrict__ zc, int nz,
    const GroupDev *__restrict__ groups, const int *__restrict__ active, int mstride,
    const double *__restrict__ qs, double step, uint64_t seed, long long dir0, int nblk, int chunk, int nu, int nv,
    bool split, double *__restrict__ partial) {
    const int gi = blockIdx.y;
    const GroupDev G = groups[active[gi]];
    // split: the collapse instance takes the shell groups whose planes (up to
    // shell layer + 2) fit the window, the direct instance the others
    const bool collapse_group = G.layer + 2 <= zc[0].aux;
    if (kCollapse ? !collapse_group : (split && collapse_group)) return;
    const long long d = dir0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double u, v;
    uniform_pair(seed, (uint32_t)d, (uint32_t)(d >> 32), (uint32_t)G.layer, kStreamMC, u, v);
    mc_stratify(d, chunk, nu, nv, u, v);
    // the wave's directions (theta, phi, unit vector), read back one at a time
    __shared__ double dirs[4][5][64];
    __shared__ double red[4][M];
    __shared__ T wq[M][4];
    __shared__ T wts[4][4][4];  // per wave and plane: the theta spline weights of the current direction
    __shared__ T4 win[4][4][kCollapse ? kZWin : 1];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    {
        const double th = acos(2.0 * u - 1.0), ph = k2Pi * v;
        double st, ct, sp, cp;
        sincos(th, &st, &ct);
        sincos(ph, &sp, &cp);
        dirs[w][0][lane] = th;
        dirs[w][1][lane] = ph;
        dirs[w][2][lane] = cp * st;
        dirs[w][3][lane] = sp * st;
        dirs[w][4][lane] = ct;
    }
    const double inv_step = 1.0 / step;
    for (int jb = 0; jb < G.m; jb += M) {
        const int m = min(M, G.m - jb);
        const double *q = qs + G.q0 + jb;
        if (threadIdx.x < m) spline4_weights<T>((T)((q[threadIdx.x] - G.layer * step) * inv_step), wq[threadIdx.x]);
        __syncthreads();
        double mine = 0.0;  // lane j < m: the wave's sum over its directions of |A_j|^2
        for (int i = 0; i < 64; i++) {
            const double thi = dirs[w][0][i], phi = dirs[w][1][i];
            const double qv0 = dirs[w][2][i], qv1 = dirs[w][3][i], qv2 = dirs[w][4][i];
            // per plane: theta (the antipode's for layer -1), its row, and (lane p) its spline weights
            int tI[4];
#pragma unroll
            for (int p = 0; p < 4; p++) {
                const int lqi = G.layer - 1 + p;
                const double tp = lqi == -1 ? fmin(fmax(kPi - thi, 0.0), kPi) : thi;
                const int shell = lqi <= 0 ? 1 : lqi;
                const double xt = tp * kInvPi * (double)(3 * shell);
                tI[p] = (int)xt;
                if (kCollapse && lane == p) spline4_weights<T>((T)(xt - (double)tI[p]), wts[w][p]);
            }
            if (kCollapse) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
            cplx<T> acc[M];
#pragma unroll
            for (int j = 0; j < M; j++) acc[j] = {(T)0, (T)0};
            // lanes past the last copy of a chunk repeat it with scale 0
            ZCopyDev zn = zc[min(lane, nz - 1)];
            for (int c0 = 0; c0 < nz; c0 += 64) {
                const int last = min(63, nz - 1 - c0);
                const ZCopyDev z = zn;
                if (c0 + 64 < nz) zn = zc[min(c0 + 64 + lane, nz - 1)];
                const T sc = lane <= last ? (T)z.scale : (T)0;
                double nph = phi - z.alpha;
                if (nph < 0.0) nph += k2Pi;
                cplx<T> amps[4];
                if (kCollapse) {
                    const int leaf0 = __builtin_amdgcn_readfirstlane(z.leaf);
                    bool all_fit = __builtin_amdgcn_readlane(z.leaf, last) == leaf0;
                    int off[4], kl[4], nseg[4];
                    T tph[4];
                    // each plane's segment: from the last copy's node (largest alpha)
                    // forward to the first copy's node + 1, cyclically
#pragma unroll
                    for (int p = 0; p < 4; p++) {
                        const int lqi = G.layer - 1 + p;
                        const int shell = lqi <= 0 ? 1 : lqi;
                        const int P = 6 * shell;
                        double pp = nph;
                        if (lqi < 0) {  // antipode(): phi + pi, wrapped
                            pp = nph + kPi;
                            if (pp >= k2Pi) pp = pp - (int)(pp / k2Pi) * k2Pi;
                        }
                        const double xp = (pp * kInv2Pi) * (double)P;
                        int pi = (int)xp;
                        tph[p] = (T)(xp - (double)pi);
                        pi = (pi == P) ? 0 : pi;
                        const int kf = __builtin_amdgcn_readfirstlane(pi);
                        kl[p] = __builtin_amdgcn_readlane(pi, last);
                        int span = kf - kl[p];
                        if (span < 0) span += P;
                        int o = pi - kl[p];
                        if (o < 0) o += P;
                        off[p] = o;
                        nseg[p] = lqi == 0 ? 0 : span + 2;
                        if (lqi != 0) all_fit = all_fit && span + 2 <= kZWin && __ballot(o > span) == 0;
                    }
                    if (all_fit) {
                        const T4 *g0 = grids[leaf0];
                        // two planes at a time: their 8 row loads issued together, then
                        // the theta spline of each segment node into LDS
#pragma unroll
                        for (int h = 0; h < 2; h++) {
                            T4 nd[2][4];
#pragma unroll
                            for (int e = 0; e < 2; e++) {
                                const int p = 2 * h + e;
                                const int lqi = G.layer - 1 + p;
                                const int shell = lqi <= 0 ? 1 : lqi;
                                const int P = 6 * shell, half = 3 * shell, thePoints = 3 * shell;
                                // this lane's segment node (a valid node for lanes past the segment)
                                const bool mine_node = lane < nseg[p];
                                int k = kl[p] + lane;
                                if (k >= P) k -= P;
                                const T4 *plane = g0 + shell_base(shell);
#pragma unroll
                                for (int r = 0; r < 4; r++) {
                                    const int vr = tI[p] + r - 1;
                                    int pos;
                                    if (vr < 0 || vr > thePoints) {
                                        const int theInd = vr < 0 ? vr + thePoints + 1 : vr - (thePoints + 1);
                                        pos = P * theInd + k + ((2 * k >= P) ? -half : half);
                                    } else {
                                        pos = P * vr + k;
                                    }
                                    nd[e][r] = plane[mine_node ? pos : 0];
                                }
                            }
#pragma unroll
                            for (int e = 0; e < 2; e++) {
                                const int p = 2 * h + e;
                                if (lane < nseg[p]) {
                                    T wt[4];
#pragma unroll
                                    for (int r = 0; r < 4; r++) wt[r] = wts[w][p][r];
                                    T4 cn;
                                    cn.x = wt[0] * nd[e][0].x;
                                    cn.y = wt[0] * nd[e][0].y;
                                    cn.z = wt[0] * nd[e][0].z;
                                    cn.w = wt[0] * nd[e][0].w;
#pragma unroll
                                    for (int r = 1; r < 4; r++) {
                                        cn.x += wt[r] * nd[e][r].x;
                                        cn.y += wt[r] * nd[e][r].y;
                                        cn.z += wt[r] * nd[e][r].z;
                                        cn.w += wt[r] * nd[e][r].w;
                                    }
                                    win[w][p][lane] = cn;
                                }
                            }
                            __builtin_amdgcn_wave_barrier();
                        }
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                        __builtin_amdgcn_wave_barrier();
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
                        for (int p = 0; p < 4; p++) {
                            if (nseg[p] == 0) {
                                const T4 o = g0[0];
                                amps[p] = {o.x, o.y};
                            } else {
                                const T4 A = win[w][p][off[p]], B = win[w][p][off[p] + 1];
                                amps[p] = hermite<T>({A.x, A.y}, {B.x, B.y}, {A.z, A.w}, {B.z, B.w}, tph[p]);
                            }
                        }
                    } else {
                        direct_planes_call<T, T4>(grids[z.leaf], G.layer, thi, nph, amps);
                    }
                } else {
                    direct_planes<T, T4>(grids[z.leaf], G.layer, thi, nph, amps);
                }
#pragma unroll
                for (int p = 0; p < 4; p++) {
                    amps[p].x *= sc;
                    amps[p].y *= sc;
                }
                const double dot = z.t[0] * qv0 + z.t[1] * qv1 + z.t[2] * qv2;
#pragma unroll
                for (int j = 0; j < M; j++) {
                    if (j < m) {
                        const T w0 = wq[j][0], w1 = wq[j][1], w2 = wq[j][2], w3 = wq[j][3];
                        const T ax = w0 * amps[0].x + w1 * amps[1].x + w2 * amps[2].x + w3 * amps[3].x;
                        const T ay = w0 * amps[0].y + w1 * amps[1].y + w2 * amps[2].y + w3 * amps[3].y;
                        T s, co;
                        phase_sincos<T>(q[j] * dot, s, co);
                        acc[j].x += ax * co - ay * s;
                        acc[j].y += ay * co + ax * s;
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < M; j++) {
                if (j < m) {
                    const double ax = (double)wave_total(acc[j].x), ay = (double)wave_total(acc[j].y);
                    if (lane == j) mine += ax * ax + ay * ay;
                }
            }
        }
        if (lane < m) red[w][lane] = mine;
        __syncthreads();
        if (threadIdx.x < m) {
            const int j = threadIdx.x;
            partial[((size_t)gi * nblk + blockIdx.x) * mstride + jb + j] =
                red[0][j] + red[1][j] + red[2][j] + red[3][j];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ gather ceiling
// The node reads of the mixed MC average without its arithmetic: for every
// (direction, layer) of an MC launch (the same stratified Philox directions) and
// every rotation, the 4 planes x 4 theta rows x 2 phi nodes of the lookup
// (plane_lookup's cells; the antipode / origin planes of layers 0 and 1 read
// shell 1) are loaded and summed, and each wave writes one float.  The rotated
// directions' angles are formed in FP32.  With few registers it runs at full
// occupancy, so its rate is the ceiling of these gathers on the device.
__device__ __forceinline__ void plane_gather(const float4 *__restrict__ plane, int shell, float ut, float up,
                                             float4 &acc) {
    const int phiPoints = 6 * shell, thePoints = 3 * shell;
    const int tI = (int)(ut * (float)thePoints);
    int pI = (int)(up * (float)phiPoints);
    pI = (pI >= phiPoints) ? 0 : pI;
    const int pI1 = (pI + 1 == phiPoints) ? 0 : pI + 1;
    const int half = phiPoints / 2;
#pragma unroll
    for (int rl = -1; rl < 3; rl++) {
        const int v = tI + rl;
        int pos1, pos2;
        if (v < 0 || v > thePoints) {
            const int theInd = v < 0 ? v + thePoints + 1 : v - (thePoints + 1);
            pos1 = phiPoints * theInd + pI + ((2 * pI >= phiPoints) ? -half : half);
            pos2 = phiPoints * theInd + pI1 + ((2 * pI1 >= phiPoints) ? -half : half);
        } else {
            pos1 = phiPoints * v + pI;
            pos2 = phiPoints * v + pI1;
        }
        const float4 a = plane[pos1], b = plane[pos2];
        acc.x += a.x + b.x;
        acc.y += a.y + b.y;
        acc.z += a.z + b.z;
        acc.w += a.w + b.w;
    }
}

template <bool kRot>
__global__ __launch_bounds__(256) void gather_probe(const float4 *__restrict__ g, int N,
                                                    const RotDev *__restrict__ rots, int nrot, int layer0,
                                                    uint64_t seed, int chunk, int nu, int nv,
                                                    float *__restrict__ sink) {
    const int layer = layer0 + blockIdx.y;
    const long long d = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double u, v;
    uniform_pair(seed, (uint32_t)d, (uint32_t)(d >> 32), (uint32_t)layer, kStreamMC, u, v);
    mc_stratify(d, chunk, nu, nv, u, v);
    const float th = acosf((float)(2.0 * u - 1.0)), ph = (float)(k2Pi * v);
    float st, ct, sp, cp;
    sincosf(th, &st, &ct);
    sincosf(ph, &sp, &cp);
    const float qv[3] = {cp * st, sp * st, ct};
    float4 acc = {0.f, 0.f, 0.f, 0.f};
    const int nr = kRot ? nrot : 1;
    for (int r = 0; r < nr; r++) {
        float ut, up;
        if (kRot) {
            const double *m = rots[r].m;
            float rv[3];
#pragma unroll
            for (int i = 0; i < 3; i++)
                rv[i] = (float)m[3 * i] * qv[0] + (float)m[3 * i + 1] * qv[1] + (float)m[3 * i + 2] * qv[2];
            rv[2] = fminf(fmaxf(rv[2], -1.f), 1.f);
            float a = atan2f(rv[1], rv[0]);
            if (a < 0.f) a += (float)k2Pi;
            ut = acosf(rv[2]) * (float)kInvPi;
            up = fminf(a * (float)kInv2Pi, 1.f);
        } else {
            ut = fminf(th * (float)kInvPi, 1.f);
            up = (float)v;
        }
#pragma unroll
        for (int i = -1; i <= 2; i++) {
            const int lq = min(max(layer + i, 1), N);
            plane_gather(g + shell_base(lq), lq, ut, up, acc);
        }
    }
    float t = acc.x + acc.y + acc.z + acc.w;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    if ((threadIdx.x & 63) == 0)
        sink[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6)] = t;
}

// sum the block partials of every active group in a fixed order:
// out[gi * mstride + j]; one wave per q (lane l adds blocks l, l + 64, ... in
// turn, then the wave's fixed shuffle tree), four waves per group
__global__ __launch_bounds__(256) void mc_reduce(const GroupDev *__restrict__ groups, const int *__restrict__ active,
                                                 const double *__restrict__ partial, int nblk, int mstride,
                                                 double *__restrict__ out) {
    const int gi = blockIdx.x;
    const int m = groups[active[gi]].m;
    const int lane = threadIdx.x & 63;
    for (int j = threadIdx.x >> 6; j < m; j += blockDim.x >> 6) {
        double s = 0.0;
        for (int b = lane; b < nblk; b += 64) s += partial[((size_t)gi * nblk + b) * mstride + j];
        s = wave_sum(s);
        if (lane == 0) out[(size_t)gi * mstride + j] = s;
    }
}

// intensity of every q from the group sums of a fixed-count MC average:
// I[q0 + j] = out[gi * mstride + j] / total (one block per group)
__global__ void mc_finish(const GroupDev *__restrict__ groups, int mstride, const double *__restrict__ out,
                          double total, double *__restrict__ I) {
    const GroupDev g = groups[blockIdx.x];
    for (int j = threadIdx.x; j < g.m; j += blockDim.x) I[g.q0 + j] = out[(size_t)blockIdx.x * mstride + j] / total;
}

// chi^2 = sum_q ((scale * I_q + constant - y_q) / sigma_q)^2 in one block, in a
// fixed summation order; q with sigma <= 0 are skipped
__global__ __launch_bounds__(1024) void chi2_kernel(const double *__restrict__ I, const double *__restrict__ y,
                                                    const double *__restrict__ sig, int n, double scale,
                                                    double constant, double *__restrict__ chi2) {
    __shared__ double part[1024];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double sg = sig[i];
        if (sg > 0.0) {
            const double r = (scale * I[i] + constant - y[i]) / sg;
            acc = fma(r, r, acc);
        }
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *chi2 = part[0];
}

void launch_mc_finish(const GroupDev *groups, int ngroups, int mstride, const double *out, double total, double *I,
                      hipStream_t s) {
    mc_finish<<<ngroups, 64, 0, s>>>(groups, mstride, out, total, I);
}

void launch_chi2(const double *I, const double *y, const double *sig, int n, double scale, double constant,
                 double *chi2, hipStream_t s) {
    chi2_kernel<<<1, 1024, 0, s>>>(I, y, sig, n, scale, constant, chi2);
}

// the same with (scale, constant) read from device memory, for captured graphs
// whose fit parameters change between replays
__global__ __launch_bounds__(1024) void chi2_param_kernel(const double *__restrict__ I, const double *__restrict__ y,
                                                          const double *__restrict__ sig, int n,
                                                          const double *__restrict__ sc, double *__restrict__ chi2) {
    __shared__ double part[1024];
    const double scale = sc[0], constant = sc[1];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double sg = sig[i];
        if (sg > 0.0) {
            const double r = (scale * I[i] + constant - y[i]) / sg;
            acc = fma(r, r, acc);
        }
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *chi2 = part[0];
}

void launch_chi2_param(const double *I, const double *y, const double *sig, int n, const double *scale_constant,
                       double *chi2, hipStream_t s) {
    chi2_param_kernel<<<1, 1024, 0, s>>>(I, y, sig, n, scale_constant, chi2);
}

// ------------------------------------------------------------------ VEGAS
constexpr int kVB = 32;  // bins per axis (GPUHybridCalc.cu:30-31)

constexpr int kVM = 8;         // q-values per pass in the VEGAS kernels
constexpr int kVPts = 4;       // points per lane and bin: points per bin <= 256

// bounds[g*66 + 0..32] theta-axis boundaries in u = (1-cos)/2, [33..65] phi-axis in phi/2pi.
// A bin is served by L lanes (L a power of two, ppb <= kVPts L), a wave by 64 / L
// bins: lane l of the wave takes the points l % L + L k of bin slot l / L.  It
// writes binsum[(gi * mstride + j) * 1024 + bin] = sum over the bin's points of
// I_j and binvar[gi * 1024 + bin] = ppb * sum_p (vol * S_p - mean)^2 with
// S_p = sum_j I_j(p) (HybridVEGASBinReduceToIKernel / HybridVEGASBinReduceKernel,
// HybridOA.cu:300-449).
__device__ __forceinline__ double lanes_sum(double v, int L) {
    for (int o = L >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <typename T, typename T4, int kL, bool kSingle, int kP>  // kL: lanes per bin fixed at compile time, or 0 for L
__global__ __launch_bounds__(256, sizeof(T) == 4 ? 2 : 1) void vegas_sample(const T4 *const *__restrict__ grids,
                                                    const RotDev *__restrict__ rots, int nrot,
                                                    const double4 *__restrict__ trans,
                                                    const GroupDev *__restrict__ groups,
                                                    const int *__restrict__ active, int mstride,
                                                    const double *__restrict__ qs, double step,
                                                    uint64_t seed, int iter, int ppb, int Lrt,
                                                    const double *__restrict__ bounds,
                                                    double *__restrict__ binsum,
                                                    double *__restrict__ binvar) {
    const int L = kL > 0 ? kL : Lrt;
    const int gi = blockIdx.y;
    const int g = active[gi];
    const GroupDev G = groups[g];
    const int lane = threadIdx.x & 63, pl = lane & (L - 1);
    // every bin of the launch exists: 1024 is a multiple of the 4 (64 / L) bins of a block
    const int bin = (blockIdx.x * 4 + (threadIdx.x >> 6)) * (64 / L) + lane / L;
    const int ti = bin / kVB, pj = bin % kVB;
    const double *bt = bounds + (size_t)g * 66, *bp = bt + 33;
    const double vol = (bt[ti + 1] - bt[ti]) * (bp[pj + 1] - bp[pj]);
    double qv[kP][3];
    double so[kP];
#pragma unroll
    for (int k = 0; k < kP; k++) {
        so[k] = 0.0;
        const int p = pl + L * k;
        double u = 0.5, v = 0.5;
        if (p < ppb) uniform_pair(seed, (uint32_t)(bin * ppb + p), (uint32_t)iter, (uint32_t)G.layer, kStreamVegas, u, v);
        const double th = acos(1. - 2. * (u * (bt[ti + 1] - bt[ti]) + bt[ti]));
        const double ph = (v * (bp[pj + 1] - bp[pj]) + bp[pj]) * 2. * kPi;
        if (kSingle) {  // the lookup takes the angles themselves
            qv[k][0] = th;
            qv[k][1] = ph;
            qv[k][2] = 0.0;
        } else {
            double st, ct, sp, cp;
            sincos(th, &st, &ct);
            sincos(ph, &sp, &cp);
            qv[k][0] = cp * st;
            qv[k][1] = sp * st;
            qv[k][2] = ct;
        }
    }
    // single leaf: each point's lookup (the two inner planes and the q-spline slopes)
    // once, held for every q chunk of the group; points in a rolled loop, one
    // lookup in flight (unrolled, the points' gathers would be interleaved past
    // the register budget), their results kept by selection
    T hold[kP][8];
    if (kSingle) {
#pragma unroll 1
        for (int k = 0; k < kP; k++) {
            if (pl + L * k >= ppb) break;
            double th = qv[0][0], ph = qv[0][1];
#pragma unroll
            for (int kk = 1; kk < kP; kk++)
                if (kk == k) {
                    th = qv[kk][0];
                    ph = qv[kk][1];
                }
            cplx<T> amps[4], d1, d2;
            group_planes<T, T4>(grids[rots[0].leaf], G.layer, th, ph, amps, d1, d2);
#pragma unroll
            for (int kk = 0; kk < kP; kk++)
                if (kk == k) {
                    hold[kk][0] = amps[1].x;
                    hold[kk][1] = amps[1].y;
                    hold[kk][2] = amps[2].x;
                    hold[kk][3] = amps[2].y;
                    hold[kk][4] = d1.x;
                    hold[kk][5] = d1.y;
                    hold[kk][6] = d2.x;
                    hold[kk][7] = d2.y;
                }
        }
    }
    for (int jb = 0; jb < G.m; jb += kVM) {
        const int m = min(kVM, G.m - jb);
        double e[kVM];
#pragma unroll
        for (int j = 0; j < kVM; j++) e[j] = 0.0;
        if (kSingle) {
            const T sc = (T)rots[0].scale;
            const double inv_step = 1.0 / step;
#pragma unroll
            for (int k = 0; k < kP; k++) {
                if (pl + L * k >= ppb) continue;
                const cplx<T> p1 = {hold[k][0], hold[k][1]}, p2 = {hold[k][2], hold[k][3]};
                const cplx<T> d1 = {hold[k][4], hold[k][5]}, d2 = {hold[k][6], hold[k][7]};
#pragma unroll
                for (int j = 0; j < kVM; j++) {
                    if (j < m) {
                        const T t = (T)((qs[G.q0 + jb + j] - G.layer * step) * inv_step);
                        const cplx<T> a = hermite<T>(p1, p2, d1, d2, t);
                        cplx<T> acc = {(T)0, (T)0};
                        acc.x += a.x * sc;
                        acc.y += a.y * sc;
                        const double I = (double)acc.x * (double)acc.x + (double)acc.y * (double)acc.y;
                        e[j] += I;
                        so[k] += I;
                    }
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < kP; k++) {
                if (pl + L * k >= ppb) continue;
                cplx<T> acc[kVM];
#pragma unroll
                for (int j = 0; j < kVM; j++) acc[j] = {(T)0, (T)0};
                docked_amplitude<T, T4, kVM>(grids, rots, nrot, trans, G.layer, step, qv[k], qs + G.q0 + jb, m,
                                             G.uniform != 0, acc);
#pragma unroll
                for (int j = 0; j < kVM; j++) {
                    const double I = (double)acc[j].x * (double)acc[j].x + (double)acc[j].y * (double)acc[j].y;
                    if (j < m) {
                        e[j] += I;
                        so[k] += I;
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < kVM; j++) {
            const double t = lanes_sum(e[j], L);
            if (pl == 0 && j < m) binsum[((size_t)gi * mstride + jb + j) * (kVB * kVB) + bin] = t;
        }
    }
    // variance of vol * S over the bin's points, two-pass (equal to the reference's Welford sum)
    double sx = 0.0;
#pragma unroll
    for (int k = 0; k < kP; k++)
        if (pl + L * k < ppb) sx += vol * so[k];
    const double mean = lanes_sum(sx, L) / (double)ppb;
    double sq = 0.0;
#pragma unroll
    for (int k = 0; k < kP; k++)
        if (pl + L * k < ppb) sq += (vol * so[k] - mean) * (vol * so[k] - mean);
    sq = lanes_sum(sq, L);
    if (pl == 0) binvar[(size_t)gi * (kVB * kVB) + bin] = sq * ppb;
}

// damped re-binning weight of axis bin i from the axis's variance marginal
// (HybridOA.cu:181-263): x = total / (local mean of var), w = ((x-1)/(x ln x))^1.5
__device__ __forceinline__ double vegas_weight(const double *var, double total, int i) {
    double x;
    if (i == 0)
        x = (var[0] + var[1]) / 2.;
    else if (i == kVB - 1)
        x = (var[i] + var[i - 1]) / 2.;
    else
        x = (var[i - 1] + var[i] + var[i + 1]) / 3.;
    x = total / x;
    x = pow((x - 1.) / (x * log(x)), 1.5);
    if (x != x) x = 0.;
    return x;
}

// serial sweep of the re-binning: new interior boundaries at equal shares of
// the weights; boundaries the sweep does not reach keep their previous value.
// w, b and the scratch nb (kVB entries) are LDS arrays: the sweep's chain of
// dependent steps then waits on LDS, not on global or private memory.
__device__ void vegas_sweep(const double *w, double *b, double *nb) {
    double tw = 0.0;
    for (int i = 0; i < kVB; i++) tw += w[i];
    for (int i = 0; i < kVB; i++) nb[i] = b[i + 1];
    double xold, xnew = 0., dw = 0., pts = tw / kVB;
    int j = 0;
    for (int i = 0; i < kVB; i++) {
        dw += w[i];
        xold = xnew;
        xnew = b[i + 1];
        for (; dw > pts; j++) {
            dw -= pts;
            if (j < kVB) nb[j] = xnew - dw * (xnew - xold) / w[i];
        }
    }
    for (int i = 0; i < kVB - 1; i++) b[1 + i] = nb[i];
}

// one block (1024 threads = the 32x32 bins) per active group: the iteration's
// integral sum_bins vol * mean_bin(I_j) (wave w takes q = w, w + 16, ...), the
// theta/phi marginals of the bin variances, re-binning (weights on 64 lanes,
// then one serial sweep per axis), and the running mean over iterations
// (HybridVEGASResizeGridKernel; GPUHybridCalc.cu:654-701).
// hist[iter * nq + q] holds the running mean after iteration iter.
__global__ __launch_bounds__(1024) void vegas_reduce(const GroupDev *__restrict__ groups,
                                                     const int *__restrict__ active, int mstride,
                                                     int ppb, int iter, const double *__restrict__ binsum,
                                                     const double *__restrict__ binvar,
                                                     double *__restrict__ bounds,
                                                     double *__restrict__ hist, int nq) {
    const int gi = blockIdx.x;
    const int g = active[gi];
    const GroupDev G = groups[g];
    const int bin = threadIdx.x, ti = bin / kVB, pj = bin % kVB;
    double *bt = bounds + (size_t)g * 66, *bp = bt + 33;
    __shared__ double sv[1024];
    __shared__ double svol[1024];
    __shared__ double marg[2][kVB];
    __shared__ double wgt[2][kVB];
    __shared__ double tot[2];
    __shared__ double sb[2][kVB + 1], snb[2][kVB];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    sv[bin] = binvar[(size_t)gi * 1024 + bin];
    svol[bin] = (bt[ti + 1] - bt[ti]) * (bp[pj + 1] - bp[pj]) / (double)ppb;
    if (threadIdx.x < 2 * (kVB + 1)) {
        const int ax = threadIdx.x / (kVB + 1), i = threadIdx.x % (kVB + 1);
        sb[ax][i] = (ax == 0 ? bt : bp)[i];
    }
    __syncthreads();
    for (int j = w; j < G.m; j += 16) {
        const double *bs = binsum + ((size_t)gi * mstride + j) * 1024;
        double c = 0.0;
        for (int b = lane; b < 1024; b += 64) c += bs[b] * svol[b];
        c = wave_sum(c);
        if (lane == 0) {
            const int q = G.q0 + j;
            double e = c;
            if (iter > 0) e = (hist[(size_t)(iter - 1) * nq + q] * (double)iter + e) / (double)(iter + 1);
            hist[(size_t)iter * nq + q] = e;
        }
    }
    // every boundary is read above before any is rewritten below
    if (threadIdx.x < 2 * kVB) {
        const int ax = threadIdx.x / kVB, i = threadIdx.x % kVB;
        double a = 0.0;
        for (int k = 0; k < kVB; k++) a += ax == 0 ? sv[i * kVB + k] : sv[k * kVB + i];
        marg[ax][i] = a;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        double t = 0.0;
        for (int i = 0; i < kVB; i++) t += marg[threadIdx.x][i];
        tot[threadIdx.x] = t;
    }
    __syncthreads();
    if (threadIdx.x < 2 * kVB) {
        const int ax = threadIdx.x / kVB, i = threadIdx.x % kVB;
        wgt[ax][i] = tot[ax] == 0. ? 0. : vegas_weight(marg[ax], tot[ax], i);
    }
    __syncthreads();
    if (threadIdx.x < 2 && tot[threadIdx.x] != 0.) vegas_sweep(wgt[threadIdx.x], sb[threadIdx.x], snb[threadIdx.x]);
    __syncthreads();
    if (threadIdx.x < 2 * (kVB + 1)) {
        const int ax = threadIdx.x / (kVB + 1), i = threadIdx.x % (kVB + 1);
        (ax == 0 ? bt : bp)[i] = sb[ax][i];
    }
}

// ------------------------------------------------------------------ single points
// GetSphr (Grid.cpp:487-548): the CPU-path lookup used by getAmplitude and GK
template <typename T, typename T4>
__device__ __forceinline__ cplx<T> get_sphr(const T4 *__restrict__ g, int N, double step, double rr, double th,
                            double ph) {
    const T4 o = g[0];
    if (rr == 0.0) return {o.x, o.y};
    double frc = rr / step;
    frc = __longlong_as_double(__double_as_longlong(frc) + 8);
    const int fInd = (int)frc;
    const double frac = frc - (double)fInd;
    cplx<T> p[4];
    const double ut = th * kInvPi, up = ph * kInv2Pi;
    p[1] = fInd > 0 ? plane_lookup<T, T4, true>(g + shell_base(fInd), fInd, ut, up) : cplx<T>{o.x, o.y};
    if (frac < 1.e-9) return p[1];
    p[2] = plane_lookup<T, T4, true>(g + shell_base(fInd + 1), fInd + 1, ut, up);
    if (fInd == 0) {
        double nth, nph;
        antipode(th, ph, nth, nph);
        p[0] = plane_lookup<T, T4, true>(g + 1, 1, nth * kInvPi, nph * kInv2Pi);
    } else if (fInd == 1) {
        p[0] = {o.x, o.y};
    } else {
        p[0] = plane_lookup<T, T4, true>(g + shell_base(fInd - 1), fInd - 1, ut, up);
    }
    p[3] = (fInd + 2 < N) ? plane_lookup<T, T4, true>(g + shell_base(fInd + 2), fInd + 2, ut, up) : p[2];
    cplx<T> d1, d2;
    four_point<T>(p, d1, d2);
    return hermite<T>(p[1], p[2], d1, d2, (T)frac);
}

// GetCart (Grid.cpp:384-392)
template <typename T, typename T4>
__device__ __forceinline__ cplx<T> get_cart(const T4 *__restrict__ g, int N, double step, double x,
                                            double y, double z) {
    double r = sqrt(x * x + y * y + z * z);
    if (fabs(z) > r) r = fabs(z);
    if (r == 0.0) return get_sphr<T, T4>(g, N, step, 0.0, 0.0, 0.0);
    const double th = acos(z / r);
    double ph = atan2(y, x);
    if (ph < 0.0) ph += k2Pi;
    return get_sphr<T, T4>(g, N, step, r, th, ph);
}

// batch GetSphr at (q, theta, phi) triples
template <typename T, typename T4>
__global__ __launch_bounds__(256) void sphr_batch(const T4 *__restrict__ g, int N, double step, int n,
                           const double *__restrict__ q, const double *__restrict__ th,
                           const double *__restrict__ ph, double2 *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    cplx<T> a = get_sphr<T, T4>(g, N, step, q[i], th[i], ph[i]);
    out[i] = make_double2((double)a.x, (double)a.y);
}

// batch InterpolateThetaPhiPlane (Grid.cpp:61-134): shell ri[i] >= 1 at
// (theta, phi) with the CPU variant's snapping of tiny fractions
template <typename T, typename T4>
__global__ __launch_bounds__(256) void plane_batch(const T4 *__restrict__ g, int n, const int *__restrict__ ri,
                                                   const double *__restrict__ th, const double *__restrict__ ph,
                                                   double2 *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int r = ri[i];
    cplx<T> a = plane_lookup<T, T4, true>(g + shell_base(r), r, th[i] * kInvPi, ph[i] * kInv2Pi);
    out[i] = make_double2((double)a.x, (double)a.y);
}

// |A|^2 of JacobianSphereGrid::CalculateIntensity's samples (Grid.cpp:421-466):
// phi[i] and v2[i] in [0,2), theta = acos(v2-1), lookup through GetCart
template <typename T, typename T4>
__global__ __launch_bounds__(256) void seeded_samples(const T4 *__restrict__ g, int N, double step, double q, long long n,
                               const double *__restrict__ phi, const double *__restrict__ v2,
                               double *__restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double theta = acos(v2[i] - 1.0);
    double st, ct, sp, cp;
    sincos(theta, &st, &ct);
    sincos(phi[i], &sp, &cp);
    cplx<T> a = get_cart<T, T4>(g, N, step, q * st * cp, q * st * sp, q * ct);
    out[i] = (double)a.x * (double)a.x + (double)a.y * (double)a.y;
}

// |A(q)|^2 of docked leaves at cartesian q (Amplitude.cpp:3035-3056 composed
// with Symmetries.cpp:697-716): the integrand of Gauss-Kronrod

template <typename T, typename T4>
__device__ double docked_intensity(const T4 *const *__restrict__ grids, const LeafGeo *__restrict__ geo,
                                   const RotDev *__restrict__ rots, int nrot, const double4 *__restrict__ trans,
                                   const double q[3]) {
    double ax = 0.0, ay = 0.0;
    for (int r = 0; r < nrot; r++) {
        const RotDev rd = rots[r];
        double rv[3];
        rotate_t(rd, q, rv);
        const LeafGeo G = geo[rd.leaf];
        cplx<T> a = get_cart<T, T4>(grids[rd.leaf], G.N, G.step, rv[0], rv[1], rv[2]);
        double sc = 0.0, ss = 0.0;
        for (int tr = 0; tr < rd.nt; tr++) {
            const double4 t = trans[rd.t0 + tr];
            double s, c;
            sincos(q[0] * t.x + q[1] * t.y + q[2] * t.z, &s, &c);
            sc += c;
            ss += s;
        }
        ax += ((double)a.x * sc - (double)a.y * ss) * rd.scale;
        ay += ((double)a.y * sc + (double)a.x * ss) * rd.scale;
    }
    return ax * ax + ay * ay;
}

template <typename T, typename T4>
__global__ __launch_bounds__(256) void docked_cart_batch(const T4 *const *__restrict__ grids, const LeafGeo *__restrict__ geo,
                                  const RotDev *__restrict__ rots, int nrot,
                                  const double4 *__restrict__ trans, int n,
                                  const double *__restrict__ qxyz, double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double q[3] = {qxyz[3 * i], qxyz[3 * i + 1], qxyz[3 * i + 2]};
    out[i] = docked_intensity<T, T4>(grids, geo, rots, nrot, trans, q);
}

// Azimuthal average of the 2-D generate path (DomainModel::CalculateIntensity2DMatrix,
// Amplitude.cpp:843-993): point p = (q sin(theta), q cos(theta)) of the (q_z, q_perp)
// plane; I_p = (1 / nphi) sum_k |A(q, theta, phi_k)|^2 over stratified azimuths
// phi_k = 2 pi (k + u_k) / nphi, u_k from the Philox stream keyed by (seed, k, p).
// The reference draws phi uniformly per (q, theta) point with a per-layer
// mt19937 (AverageIntensitiesBetweenLayers2D, :2825-2897); one stratum per
// sample gives the same expectation with a fraction of the variance.  One block
// per point; the block's sum is formed in a fixed order.
template <typename T, typename T4>
__global__ __launch_bounds__(256) void azimuthal_kernel(const T4 *const *__restrict__ grids,
                                                        const LeafGeo *__restrict__ geo,
                                                        const RotDev *__restrict__ rots, int nrot,
                                                        const double4 *__restrict__ trans,
                                                        const double2 *__restrict__ pts, int nphi, uint64_t seed,
                                                        double *__restrict__ out) {
    __shared__ double red[4];
    const int p = blockIdx.x;
    const double2 pt = pts[p];
    double sum = 0.0;
    for (int k = threadIdx.x; k < nphi; k += blockDim.x) {
        double u, v;
        uniform_pair(seed, (uint32_t)k, (uint32_t)p, 0u, kStreamAzimuth, u, v);
        const double ph = k2Pi * ((double)k + u) / (double)nphi;
        double sp, cp;
        sincos(ph, &sp, &cp);
        const double q[3] = {pt.x * cp, pt.x * sp, pt.y};
        sum += docked_intensity<T, T4>(grids, geo, rots, nrot, trans, q);
    }
    sum = wave_sum(sum);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) out[p] = (red[0] + red[1] + red[2] + red[3]) / (double)nphi;
}

// Gauss-Kronrod phi intervals: task t = (q sin(theta), q cos(theta), a, b) of
// one interval; its 15 integrand points are phi = a + b x_k.  A workgroup of
// 240 threads evaluates the points of 16 tasks, then one thread per task forms
// the Kronrod and Gauss sums
//   out[2t] = sum_k I_k b kw_k,  out[2t + 1] = sum_{k odd} I_k b gw_{k/2}
// in k order with every product and sum rounded as written (no fused
// multiply-add), the arithmetic of DomainModel::GaussLegPhi
// (Amplitude.cpp:1443-1467), so accept/split decisions on the host follow it.
constexpr int kGkTasksPerBlock = 16;

template <typename T, typename T4>
__global__ __launch_bounds__(15 * kGkTasksPerBlock) void docked_gk_batch(
    const T4 *const *__restrict__ grids, const LeafGeo *__restrict__ geo, const RotDev *__restrict__ rots,
    int nrot, const double4 *__restrict__ trans, long long ntasks, const double4 *__restrict__ tasks,
    const GkNodes *__restrict__ nodes, double *__restrict__ out) {
    __shared__ double val[15 * kGkTasksPerBlock];
    const long long t0 = (long long)blockIdx.x * kGkTasksPerBlock;
    const long long t = t0 + threadIdx.x / 15;
    double I = 0.0;
    if (t < ntasks) {
        const double4 tk = tasks[t];
        const double xi = tk.z + tk.w * nodes->x[threadIdx.x % 15];
        const double q[3] = {tk.x * cos(xi), tk.x * sin(xi), tk.y};
        I = docked_intensity<T, T4>(grids, geo, rots, nrot, trans, q);
    }
    val[threadIdx.x] = I;
    __syncthreads();
    const long long ts = t0 + threadIdx.x;
    if (threadIdx.x < kGkTasksPerBlock && ts < ntasks) {
        const double hw = tasks[ts].w;
        const double *v = val + 15 * threadIdx.x;
        double kv = 0.0, gv = 0.0;
        for (int k = 0; k < 15; k++) {
            const double p = __dmul_rn(v[k], hw);
            kv = __dadd_rn(kv, __dmul_rn(p, nodes->kw[k]));
            if (k % 2 == 1) gv = __dadd_rn(gv, __dmul_rn(p, nodes->gw[k / 2]));
        }
        out[2 * ts] = kv;
        out[2 * ts + 1] = gv;
    }
}

// Gauss-Kronrod phi recursion, one level per launch: each open interval is
// accepted (its node takes its Kronrod sum) or split in two at its midpoint
// (two new nodes and intervals at slots taken from counters, one atomic per
// wave: cnt[0] next-level intervals, cnt[1] nodes, cnt[2] set on overflow).
// Nodes are numbered level by level, so after the last level gk_tree_sums
// walks the node ranges back to the roots: an inner node takes left + right,
// and the roots' values are the phi integrals.  The accept rule and the
// midpoint are GaussLegPhi's (Amplitude.cpp:1449-1467).
__global__ __launch_bounds__(256) void gk_decide(long long n, const GkInterval *__restrict__ cur,
                                                  const double *__restrict__ kvgv, double eps, GkNode *nodes,
                                                  int *cnt, int cap_int, int cap_nodes, const double4 *__restrict__ task,
                                                  double4 *__restrict__ next_task, GkInterval *__restrict__ next) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = t < n;
    GkInterval I{0.0, 0.0, 0, 0, 0, 0};
    double kv = 0.0, gv = 0.0;
    bool accept = false;
    if (live) {
        I = cur[t];
        kv = kvgv[2 * t];
        gv = kvgv[2 * t + 1];
        accept = I.mind < 0 && (I.maxd < 0 || fabs(1.0 - (gv / kv)) <= eps || gv == kv);
    }
    const bool split = live && !accept;
    const unsigned long long mask = __ballot(split);
    const int lane = threadIdx.x & 63;
    int c = 0, w = 0;
    if (mask) {
        const int leader = __ffsll((long long)mask) - 1, nsplit = __popcll(mask);
        int bc = 0, bw = 0;
        if (lane == leader) {
            bc = atomicAdd(&cnt[1], 2 * nsplit);
            bw = atomicAdd(&cnt[0], 2 * nsplit);
        }
        bc = __shfl(bc, leader, 64);
        bw = __shfl(bw, leader, 64);
        const int rank = __popcll(mask & ((1ull << lane) - 1ull));
        c = bc + 2 * rank;
        w = bw + 2 * rank;
    }
    if (live && accept) nodes[I.node].value = kv;
    if (!split) return;
    if (c + 2 > cap_nodes || w + 2 > cap_int) {
        atomicExch(&cnt[2], 1);
        return;
    }
    nodes[c] = GkNode{0.0, I.node, -1, -1, 0};
    nodes[c + 1] = GkNode{0.0, I.node, -1, -1, 0};
    nodes[I.node].child0 = c;
    nodes[I.node].child1 = c + 1;
    const double4 tk = task[t];
    const double mid = (I.pmax + I.pmin) * 0.5;
    next[w] = GkInterval{I.pmin, mid, c, I.maxd - 1, I.mind - 1, 0};
    next[w + 1] = GkInterval{mid, I.pmax, c + 1, I.maxd - 1, I.mind - 1, 0};
    next_task[w] = make_double4(tk.x, tk.y, (mid + I.pmin) * 0.5, (mid - I.pmin) * 0.5);
    next_task[w + 1] = make_double4(tk.x, tk.y, (I.pmax + mid) * 0.5, (I.pmax - mid) * 0.5);
}

// nodes [lo, hi) of one level: a split node takes left + right (its children,
// in the next range, are final); a root writes its value to phival
__global__ __launch_bounds__(256) void gk_tree_sums(GkNode *nodes, int lo, int hi, int nroots, double *phival) {
    const int v = lo + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (v >= hi) return;
    GkNode &N = nodes[v];
    if (N.child0 >= 0) N.value = nodes[N.child0].value + nodes[N.child1].value;
    if (v < nroots) phival[v] = N.value;
}

void launch_gk_decide(long long n, const GkInterval *cur, const double *kvgv, double eps, GkNode *nodes, int *cnt,
                      int cap_int, int cap_nodes, const double4 *task, double4 *next_task, GkInterval *next,
                      hipStream_t s) {
    gk_decide<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, cur, kvgv, eps, nodes, cnt, cap_int, cap_nodes, task,
                                                           next_task, next);
}

void launch_gk_tree_sums(GkNode *nodes, int lo, int hi, int nroots, double *phival, hipStream_t s) {
    if (hi > lo) gk_tree_sums<<<(unsigned)((hi - lo + 255) / 256), 256, 0, s>>>(nodes, lo, hi, nroots, phival);
}

// ------------------------------------------------------------------ launchers
void launch_mc(bool fp64, bool single, const void *const *grids, const RotDev *rots, int nrot,
               const double4 *trans, const ZCopyDev *zc, int nz, const GroupDev *groups, const int *active,
               int nactive, int mstride, const double *qs, double step, uint64_t seed, long long dir0, int nblk,
               int chunk, double *partial, double *out, hipStream_t s) {
    dim3 grid(nblk, nactive);
    int nu, nv;
    mc_strata(chunk, nu, nv);
    if (!single && zc && nz >= kZCopyMin) {
        // 8 q per pass when no shell group holds more (fewer accumulators);
        // FP32 grids: the collapse instance and the direct one over the same
        // grid, each block running in one of them
#define DPLUS_MCZ(T, T4, C, M, SPLIT)                                                                             \
    mc_zcopy_kernel<T, T4, C, M><<<grid, 256, 0, s>>>((const T4 *const *)grids, zc, nz, groups, active, mstride, \
                                                      qs, step, seed, dir0, nblk, chunk, nu, nv, SPLIT, partial)
        if (fp64 && mstride <= 8) {
            DPLUS_MCZ(double, double4, false, 8, false);
        } else if (fp64) {
            DPLUS_MCZ(double, double4, false, kMaxM, false);
        } else if (mstride <= 8) {
            DPLUS_MCZ(float, float4, true, 8, true);
            DPLUS_MCZ(float, float4, false, 8, true);
        } else {
            DPLUS_MCZ(float, float4, true, kMaxM, true);
            DPLUS_MCZ(float, float4, false, kMaxM, true);
        }
#undef DPLUS_MCZ
        mc_reduce<<<nactive, 256, 0, s>>>(groups, active, partial, nblk, mstride, out);
        return;
    }
#define DPLUS_MC(T, T4, S, M)                                                                                      \
    mc_kernel<T, T4, S, M><<<grid, 256, 0, s>>>((const T4 *const *)grids, rots, nrot, trans, groups, active, mstride, \
                                                qs, step, seed, dir0, nblk, chunk, nu, nv, partial)
    if (fp64 && single)
        DPLUS_MC(double, double4, true, kMaxM);
    else if (fp64 && mstride <= 8)
        DPLUS_MC(double, double4, false, 8);
    else if (fp64)
        DPLUS_MC(double, double4, false, kMaxM);
    else if (single)
        DPLUS_MC(float, float4, true, kMaxM);
    else if (mstride <= 8)
        DPLUS_MC(float, float4, false, 8);
    else
        DPLUS_MC(float, float4, false, kMaxM);
#undef DPLUS_MC
    mc_reduce<<<nactive, 256, 0, s>>>(groups, active, partial, nblk, mstride, out);
}

// the gather-ceiling probe over layers [layer0, layer0 + nlayers) of a mixed grid
// with N shells: nblk blocks of 256 directions per layer; sink holds
// nblk * nlayers * 4 floats
void launch_gather_probe(const void *grid, int N, const RotDev *rots, int nrot, int layer0, int nlayers, int nblk,
                         uint64_t seed, int chunk, float *sink, hipStream_t s) {
    int nu, nv;
    mc_strata(chunk, nu, nv);
    dim3 gr(nblk, nlayers);
    if (rots)
        gather_probe<true><<<gr, 256, 0, s>>>((const float4 *)grid, N, rots, nrot, layer0, seed, chunk, nu, nv, sink);
    else
        gather_probe<false><<<gr, 256, 0, s>>>((const float4 *)grid, N, nullptr, 1, layer0, seed, chunk, nu, nv,
                                               sink);
}

void launch_vegas_iter(bool fp64, bool single, const void *const *grids, const RotDev *rots, int nrot,
                       const double4 *trans, const GroupDev *groups, const int *active, int nactive,
                       int mstride, const double *qs, double step, uint64_t seed, int iter, int ppb,
                       double *bounds, double *binsum, double *binvar, double *hist, int nq, hipStream_t s) {
    // lanes per bin: the least power of two >= ppb, at most a wave (ppb <= kVPts * 64)
    int L = 1;
    while (L < ppb && L < 64) L <<= 1;
    dim3 grid(kVB * kVB / (4 * (64 / L)), nactive);
#define DPLUS_VEGAS(T, T4, KL, S, P)                                                                          \
    vegas_sample<T, T4, KL, S, P><<<grid, 256, 0, s>>>((const T4 *const *)grids, rots, nrot, trans, groups,      \
                                                       active, mstride, qs, step, seed, iter, ppb, L, bounds,   \
                                                       binsum, binvar)
    // points per lane: the single-leaf instances are sized to the bin's points
    // (fewer registers), the docked ones take up to kVPts
    const int per_lane = (ppb + L - 1) / L;
#define DPLUS_VEGAS_P(T, T4, KL)                 \
    if (!single)                                 \
        DPLUS_VEGAS(T, T4, KL, false, kVPts);    \
    else if (per_lane <= 1)                      \
        DPLUS_VEGAS(T, T4, KL, true, 1);         \
    else if (per_lane <= 2)                      \
        DPLUS_VEGAS(T, T4, KL, true, 2);         \
    else                                         \
        DPLUS_VEGAS(T, T4, KL, true, kVPts);
    if (fp64 && L == 64) {
        DPLUS_VEGAS_P(double, double4, 64)
    } else if (fp64) {
        DPLUS_VEGAS_P(double, double4, 0)
    } else if (L == 64) {
        DPLUS_VEGAS_P(float, float4, 64)
    } else {
        DPLUS_VEGAS_P(float, float4, 0)
    }
#undef DPLUS_VEGAS_P
#undef DPLUS_VEGAS
    vegas_reduce<<<nactive, 1024, 0, s>>>(groups, active, mstride, ppb, iter, binsum, binvar, bounds, hist, nq);
}

void launch_sphr_batch(bool fp64, const void *g, int N, double step, int n, const double *q,
                       const double *th, const double *ph, double2 *out, hipStream_t s) {
    unsigned nb = (n + 255) / 256;
    if (fp64)
        sphr_batch<double, double4><<<nb, 256, 0, s>>>((const double4 *)g, N, step, n, q, th, ph, out);
    else
        sphr_batch<float, float4><<<nb, 256, 0, s>>>((const float4 *)g, N, step, n, q, th, ph, out);
}

void launch_plane_batch(bool fp64, const void *g, int n, const int *ri, const double *th, const double *ph,
                        double2 *out, hipStream_t s) {
    unsigned nb = (n + 255) / 256;
    if (fp64)
        plane_batch<double, double4><<<nb, 256, 0, s>>>((const double4 *)g, n, ri, th, ph, out);
    else
        plane_batch<float, float4><<<nb, 256, 0, s>>>((const float4 *)g, n, ri, th, ph, out);
}

void launch_seeded_samples(bool fp64, const void *g, int N, double step, double q, long long n,
                           const double *phi, const double *v2, double *out, hipStream_t s) {
    unsigned nb = (unsigned)((n + 255) / 256);
    if (fp64)
        seeded_samples<double, double4><<<nb, 256, 0, s>>>((const double4 *)g, N, step, q, n, phi, v2, out);
    else
        seeded_samples<float, float4><<<nb, 256, 0, s>>>((const float4 *)g, N, step, q, n, phi, v2, out);
}

void launch_docked_gk(bool fp64, const void *const *grids, const LeafGeo *geo, const RotDev *rots, int nrot,
                      const double4 *trans, long long ntasks, const double4 *tasks, const GkNodes *nodes,
                      double *out, hipStream_t s) {
    const unsigned nb = (unsigned)((ntasks + kGkTasksPerBlock - 1) / kGkTasksPerBlock);
    const int bs = 15 * kGkTasksPerBlock;
    if (fp64)
        docked_gk_batch<double, double4><<<nb, bs, 0, s>>>((const double4 *const *)grids, geo, rots, nrot, trans,
                                                            ntasks, tasks, nodes, out);
    else
        docked_gk_batch<float, float4><<<nb, bs, 0, s>>>((const float4 *const *)grids, geo, rots, nrot, trans,
                                                          ntasks, tasks, nodes, out);
}

void launch_azimuthal(bool fp64, const void *const *grids, const LeafGeo *geo, const RotDev *rots, int nrot,
                      const double4 *trans, int npts, const double2 *pts, int nphi, uint64_t seed, double *out,
                      hipStream_t s) {
    if (npts <= 0) return;
    if (fp64)
        azimuthal_kernel<double, double4><<<npts, 256, 0, s>>>((const double4 *const *)grids, geo, rots, nrot, trans,
                                                               pts, nphi, seed, out);
    else
        azimuthal_kernel<float, float4><<<npts, 256, 0, s>>>((const float4 *const *)grids, geo, rots, nrot, trans,
                                                             pts, nphi, seed, out);
}

void launch_docked_cart(bool fp64, const void *const *grids, const LeafGeo *geo, const RotDev *rots,
                        int nrot, const double4 *trans, int n, const double *qxyz, double *out,
                        hipStream_t s) {
    unsigned nb = (n + 255) / 256;
    if (fp64)
        docked_cart_batch<double, double4><<<nb, 256, 0, s>>>((const double4 *const *)grids, geo, rots,
                                                               nrot, trans, n, qxyz, out);
    else
        docked_cart_batch<float, float4><<<nb, 256, 0, s>>>((const float4 *const *)grids, geo, rots, nrot,
                                                             trans, n, qxyz, out);
}

}  // namespace dplus

namespace dplus {

// ------------------------------------------------------------------ grid-level docking
__device__ __forceinline__ void node_indices(long long index, int &qi, long long &ti, long long &pi) {
    if (index <= 0) {
        qi = 0;
        ti = pi = 0;
        return;
    }
    long long q = (long long)cbrt((double)index / 6.0);
    if (q < 1) q = 1;
    while (q > 1 && shell_base(q) > index) q--;
    while (shell_base(q + 1) <= index) q++;
    const long long rem = index - shell_base(q);
    qi = (int)q;
    ti = rem / (6 * q);
    pi = rem % (6 * q);
}

// ManualSymmetry::calcAmplitude (Symmetries.cpp:697-716) at every node of the
// output grid (Grid.cpp:760-826 node vectors), sub-amplitudes through GetCart
// Space-filling lattice sum sum_{n<N} e^{i n d} = e^{i(N-1)d/2} sin(Nd/2)/sin(d/2)
// (GridSymmetry::calcAmplitude, DefaultModels/Symmetries.cpp:97-136; GPU
// CalcSpaceFillingKernelJcb, GPU/SymmetryGenJacob.cu:32-132).  N at d = 0 as the
// reference does; at the other zeros of sin(d/2) (d = 2 pi k) the limit
// N (-1)^{(N-1)k}, where the reference's expression is 0/0.
__device__ inline void lattice_mul(double d, double n, double &re, double &im) {
    if (fabs(d) < 100.0 * 2.220446049250313e-16 || n == 1.0) {
        re *= n;
        im *= n;
        return;
    }
    const double h = 0.5 * d;
    const double sh = sin(h);
    double f;
    if (fabs(sh) < 1e-12) {
        const double k = rint(d / k2Pi);
        f = (fmod(fabs((n - 1.0) * k), 2.0) == 1.0) ? -n : n;
    } else {
        f = sin(n * h) / sh;
    }
    double sp, cp;
    sincos((n - 1.0) * h, &sp, &cp);
    const double r = re * cp - im * sp, i = im * cp + re * sp;
    re = r * f;
    im = i * f;
}

template <typename T, typename T4>
__global__ __launch_bounds__(256) void dock_grid_kernel(const T4 *const *__restrict__ grids, const LeafGeo *__restrict__ geo,
                                 const RotDev *__restrict__ rots, int nrot, const double4 *__restrict__ trans,
                                 T4 *__restrict__ out, double step_out, long long P, double scale,
                                 LatticeDev lat) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= P) return;
    int qi;
    long long ti, pi;
    node_indices(idx, qi, ti, pi);
    double q[3] = {0.0, 0.0, 0.0};
    if (qi > 0) {
        const double qI = (double)qi * step_out;
        const double tI = kPi * (double)ti / (double)(3 * qi);
        const double pI = k2Pi * (double)pi / (double)(6 * qi);
        double st, ct, sp, cp;
        sincos(tI, &st, &ct);
        sincos(pI, &sp, &cp);
        q[0] = qI * st * cp;
        q[1] = qI * st * sp;
        q[2] = qI * ct;
    }
    double ax = 0.0, ay = 0.0;
    for (int r = 0; r < nrot; r++) {
        const RotDev rd = rots[r];
        double rv[3];
        rotate_t(rd, q, rv);
        const LeafGeo G = geo[rd.leaf];
        cplx<T> a = get_cart<T, T4>(grids[rd.leaf], G.N, G.step, rv[0], rv[1], rv[2]);
        double sc = 0.0, ss = 0.0;
        for (int tr = 0; tr < rd.nt; tr++) {
            const double4 t = trans[rd.t0 + tr];
            double s, c;
            sincos(q[0] * t.x + q[1] * t.y + q[2] * t.z, &s, &c);
            sc += c;
            ss += s;
        }
        ax += ((double)a.x * sc - (double)a.y * ss) * rd.scale;
        ay += ((double)a.y * sc + (double)a.x * ss) * rd.scale;
    }
    if (lat.on)
        for (int k = 0; k < 3; k++)
            lattice_mul(q[0] * lat.v[3 * k] + q[1] * lat.v[3 * k + 1] + q[2] * lat.v[3 * k + 2], lat.n[k], ax, ay);
    T4 o;
    o.x = (T)(ax * scale);
    o.y = (T)(ay * scale);
    o.z = (T)0;
    o.w = (T)0;
    out[idx] = o;
}

void dock_grid(const DockSet &D, Grid &out, double scale, hipStream_t s, const LatticeDev *lattice) {
    const unsigned nb = (unsigned)((out.P + 255) / 256);
    LatticeDev lat{};
    if (lattice) lat = *lattice;
    KTimer kt("dock", s);
    if (D.fp64)
        dock_grid_kernel<double, double4><<<nb, 256, 0, s>>>((const double4 *const *)D.grids.p,
                                                              (const LeafGeo *)D.geos.p, (const RotDev *)D.rotb.p,
                                                              D.nrot, (const double4 *)D.trb.p,
                                                              (double4 *)out.dev.p, out.step, out.P, scale, lat);
    else
        dock_grid_kernel<float, float4><<<nb, 256, 0, s>>>((const float4 *const *)D.grids.p,
                                                            (const LeafGeo *)D.geos.p, (const RotDev *)D.rotb.p,
                                                            D.nrot, (const double4 *)D.trb.p, (float4 *)out.dev.p,
                                                            out.step, out.P, scale, lat);
    check_hip(hipGetLastError(), "dock_grid launch");
    out.splines(s);
    check_hip(hipStreamSynchronize(s), "dock_grid");
}

}  // namespace dplus
