// amp_build.hip -- atomic amplitude grid builder for gfx950.
//
// Computes, at every node q of the Jacobian spherical grid,
//   A(q) = scale * sum_ion f~_ion(|q|) * sum_{j in ion} exp(i q.r_j)
//        + scale * contrast * sum_box prod_a sinc(q_a n_a D/2) n_a D exp(i q.c_box)
// which is PDBAmplitude::calcAmplitude (Backend/Backend/PDBAmplitude.cpp:243-317)
// evaluated on the grid (Backend/Backend/Grid.cpp:760-826), the work the
// reference GPU path splits over CombinedPDBandDummyAtomJacobForSingleIonKernel
// and CalcPDBVoxelSolventKernelJacobSphr (PDBAmplitudeCUDA.cu:15-186).
//
// Design (MI355X):
//  * Friedel symmetry: for real form factors A(-q) = conj(A(q)); every shell is
//    antipodally closed, so only its canonical half (3i(3i+1) nodes) is summed
//    and the conjugate is stored at the antipode.  Half the sincos work.
//  * Atoms are sorted by ion and cut into tiles of <= kTile atoms of one ion.
//    The atom loop index is wave-uniform, so atom coordinates arrive through
//    scalar loads (s_load) and feed v_fma as SGPR operands: no LDS traffic and
//    no barriers in the inner loop.
//  * Mixed precision: q is pre-scaled to revolutions, the phase is an FP32 FMA
//    chain (or an FP64 one, kPhase64), reduced with v_fract and fed to the
//    hardware v_sin_f32 / v_cos_f32.  Each tile is summed in FP32 and folded
//    into FP64 accumulators with the ion's form factor at the node's shell:
//    a two-level compensated sum whose FP32 partials never exceed kTile terms.
//  * FP64 mode: the same structure with double coordinates and double sincos.
//  * Each thread owns K canonical nodes, so every scalar atom load feeds K
//    phase evaluations.

#include "common.hpp"

namespace dplus {

constexpr int kTile = 256;

struct NodeInfo {
    long long idx, anti;
    int shell;
};

// canonical node c -> (storage index, antipode index, shell)
__device__ __forceinline__ NodeInfo canonical_node(long long c, double step, double &qx, double &qy,
                                                   double &qz) {
    NodeInfo n;
    int i = shell_of_half(c);
    long long h = c - half_base(i);
    long long j = h / (6LL * i), k = h % (6LL * i);
    n.shell = i;
    n.idx = index_from_indices(i, j, k);
    n.anti = index_from_indices(i, 3LL * i - j, (k + 3LL * i) % (6LL * i));
    double q = (double)i * step;
    double th = kPi * (double)j / (double)(3 * i);
    double ph = k2Pi * (double)k / (double)(6 * i);
    double st, ct, sp, cp;
    sincos(th, &st, &ct);
    sincos(ph, &sp, &cp);
    qx = q * st * cp;
    qy = q * st * sp;
    qz = q * ct;
    return n;
}

// weight of a tile of solvent boxes of one shape at node q (nm^-1):
// contrast * prod_a L_a sinc(q_a L_a / 2), L_a = n_a * vox
__device__ __forceinline__ double box_weight64(double qx, double qy, double qz, int4 d, double vox) {
    const double q3[3] = {qx, qy, qz};
    const int n3[3] = {d.x, d.y, d.z};
    double w = 1.0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const double L = (double)n3[a] * vox, x = 0.5 * q3[a] * L;
        w *= (q3[a] == 0.0 || n3[a] == 0) ? L : sin(x) / x * L;
    }
    return w;
}

__device__ __forceinline__ float box_weight32(float qx, float qy, float qz, int4 d, float vox) {
    const float q3[3] = {qx, qy, qz};
    const int n3[3] = {d.x, d.y, d.z};
    float w = 1.0f;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float L = (float)n3[a] * vox, x = 0.5f * q3[a] * L;
        // sin(x) on the hardware unit (argument in revolutions, reduced to
        // [-1/2, 1/2] by round-to-nearest so tiny negative x stay tiny);
        // below |x| = 1e-3 the series 1 - x^2/6 is exact in FP32
        float r = x * (float)(1.0 / k2Pi);
        r -= rintf(r);
        const float s = __builtin_amdgcn_sinf(r);
        w *= (fabsf(x) < 1e-3f || n3[a] == 0) ? L * (1.0f - x * x * (1.0f / 6.0f)) : s / x * L;
    }
    return w;
}

// Tiles: (first, count, index, kind).  kind 0: atoms of ion `index`, weighted by
// the ion's form factor at the node's shell; kind 1: solvent boxes of shape
// bdims[index], weighted by contrast * box_weight at the node.
template <int K, bool kPhase64>
__global__ __launch_bounds__(256) void build_atoms_mixed(
    const float4 *__restrict__ atoms, const double4 *__restrict__ atoms64,
    const int4 *__restrict__ tiles, int ntiles, const double *__restrict__ ff, int nion,
    const int4 *__restrict__ bdims, double vox, double contrast,
    double step, long long nhalf, float4 *__restrict__ grid, double scale) {
    const long long c0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * K;
    float qx[K], qy[K], qz[K];
    double dqx[K], dqy[K], dqz[K];
    NodeInfo node[K];
    double accR[K], accI[K];
    const double inv2pi = 1.0 / k2Pi;
#pragma unroll
    for (int p = 0; p < K; p++) {
        long long c = c0 + p;
        if (c >= nhalf) c = nhalf - 1;  // duplicate work, store skipped below
        double x, y, z;
        node[p] = canonical_node(c, step, x, y, z);
        dqx[p] = x * inv2pi;
        dqy[p] = y * inv2pi;
        dqz[p] = z * inv2pi;
        qx[p] = (float)dqx[p];
        qy[p] = (float)dqy[p];
        qz[p] = (float)dqz[p];
        accR[p] = 0.0;
        accI[p] = 0.0;
    }
    for (int t = 0; t < ntiles; t++) {
        const int4 tl = tiles[t];
        float sr[K], si[K];
#pragma unroll
        for (int p = 0; p < K; p++) sr[p] = si[p] = 0.f;
        const int a1 = tl.x + tl.y;
#pragma unroll 2
        for (int a = tl.x; a < a1; a++) {
            if constexpr (kPhase64) {
                const double4 at = atoms64[a];
#pragma unroll
                for (int p = 0; p < K; p++) {
                    double ph = fma(dqx[p], at.x, fma(dqy[p], at.y, dqz[p] * at.z));
                    float r = (float)(ph - floor(ph));
                    sr[p] += __builtin_amdgcn_cosf(r);
                    si[p] += __builtin_amdgcn_sinf(r);
                }
            } else {
                const float4 at = atoms[a];
#pragma unroll
                for (int p = 0; p < K; p++) {
                    float ph = fmaf(qx[p], at.x, fmaf(qy[p], at.y, qz[p] * at.z));
                    float r = __builtin_amdgcn_fractf(ph);
                    sr[p] += __builtin_amdgcn_cosf(r);
                    si[p] += __builtin_amdgcn_sinf(r);
                }
            }
        }
        if (tl.w == 0) {
#pragma unroll
            for (int p = 0; p < K; p++) {
                double f = ff[(size_t)node[p].shell * nion + tl.z];
                accR[p] = fma(f, (double)sr[p], accR[p]);
                accI[p] = fma(f, (double)si[p], accI[p]);
            }
        } else {
            const int4 d = bdims[tl.z];
            const float k2 = (float)k2Pi;
#pragma unroll
            for (int p = 0; p < K; p++) {
                double f = contrast * (double)box_weight32(qx[p] * k2, qy[p] * k2, qz[p] * k2, d, (float)vox);
                accR[p] = fma(f, (double)sr[p], accR[p]);
                accI[p] = fma(f, (double)si[p], accI[p]);
            }
        }
    }
#pragma unroll
    for (int p = 0; p < K; p++) {
        if (c0 + p < nhalf) {
            float re = (float)(accR[p] * scale), im = (float)(accI[p] * scale);
            grid[node[p].idx] = make_float4(re, im, 0.f, 0.f);
            grid[node[p].anti] = make_float4(re, -im, 0.f, 0.f);
        }
    }
}

// Paired builder (mixed).  The nodes (j, k) and (j, k + 3i) of a row lie at
// azimuths phi and phi + pi: q_z is shared and (q_x, q_y) changes sign, so with
// z = q_z z_a and u = q_x x_a + q_y y_a their phases are z + u and z - u, and
//   e^{i(z+u)} = (cz cu - sz su) + i (sz cu + cz su)
//   e^{i(z-u)} = (cz cu + sz su) + i (sz cu - cz su).
// One unit owns three such pairs; per atom it takes sin/cos of z once and of u
// once per pair, and sums the four products cz cu, sz su, sz cu, cz su: eight
// terms from eight transcendentals instead of twelve.  Box weights are even in
// (q_x, q_y), so one per pair.  Rows off the equator also write both
// antipodes; on the equator (j = 3i/2) the partner is the antipode itself.
// With kReduce false the phases (in revolutions) go to v_sin_f32 / v_cos_f32
// unreduced; the host selects that form only when every |q.r| / 2 pi of the
// structure is below 128 revolutions (|z| and |u| are both bounded by |q||r|),
// inside the units' input domain.
constexpr int kPair = 3;

template <bool kReduce>
__global__ __launch_bounds__(256) void build_atoms_pairs(
    const float4 *__restrict__ atoms, const int4 *__restrict__ tiles, int ntiles, const double *__restrict__ ff,
    int nion, const int4 *__restrict__ bdims, double vox, double contrast, double step, long long nunits,
    float4 *__restrict__ grid, double scale) {
    const long long u = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nunits) return;
    const int shell = shell_of_pair_unit(u);
    const long long r = u - pair_units_before(shell - 1);
    const long long j = r / shell, k0 = 3 * (r % shell), half = 3LL * shell;
    float qx[kPair], qy[kPair], qz;
    double aR[kPair], aI[kPair], bR[kPair], bI[kPair];
    const double inv2pi = 1.0 / k2Pi;
    {
        const double q = (double)shell * step;
        double st, ct;
        sincos(kPi * (double)j / (double)(3 * shell), &st, &ct);
        qz = (float)(q * ct * inv2pi);
#pragma unroll
        for (int p = 0; p < kPair; p++) {
            double sp, cp;
            sincos(k2Pi * (double)(k0 + p) / (double)(6 * shell), &sp, &cp);
            qx[p] = (float)(q * st * cp * inv2pi);
            qy[p] = (float)(q * st * sp * inv2pi);
            aR[p] = aI[p] = bR[p] = bI[p] = 0.0;
        }
    }
    for (int t = 0; t < ntiles; t++) {
        const int4 tl = tiles[t];
        float cc[kPair], ss[kPair], sc[kPair], cs[kPair];
#pragma unroll
        for (int p = 0; p < kPair; p++) cc[p] = ss[p] = sc[p] = cs[p] = 0.f;
        const int a1 = tl.x + tl.y;
#pragma unroll 2
        for (int a = tl.x; a < a1; a++) {
            const float4 at = atoms[a];
            const float zz = kReduce ? __builtin_amdgcn_fractf(qz * at.z) : qz * at.z;
            const float cz = __builtin_amdgcn_cosf(zz), sz = __builtin_amdgcn_sinf(zz);
#pragma unroll
            for (int p = 0; p < kPair; p++) {
                const float ph = fmaf(qx[p], at.x, qy[p] * at.y);
                const float v = kReduce ? __builtin_amdgcn_fractf(ph) : ph;
                const float cu = __builtin_amdgcn_cosf(v), su = __builtin_amdgcn_sinf(v);
                cc[p] = fmaf(cz, cu, cc[p]);
                ss[p] = fmaf(sz, su, ss[p]);
                sc[p] = fmaf(sz, cu, sc[p]);
                cs[p] = fmaf(cz, su, cs[p]);
            }
        }
#pragma unroll
        for (int p = 0; p < kPair; p++) {
            double f;
            if (tl.w == 0) {
                f = ff[(size_t)shell * nion + tl.z];
            } else {
                const float k2 = (float)k2Pi;
                f = contrast * (double)box_weight32(qx[p] * k2, qy[p] * k2, qz * k2, bdims[tl.z], (float)vox);
            }
            const double c2 = (double)cc[p], s2 = (double)ss[p], sc2 = (double)sc[p], cs2 = (double)cs[p];
            aR[p] = fma(f, c2 - s2, aR[p]);
            aI[p] = fma(f, sc2 + cs2, aI[p]);
            bR[p] = fma(f, c2 + s2, bR[p]);
            bI[p] = fma(f, sc2 - cs2, bI[p]);
        }
    }
    const bool equator = 2 * j == 3LL * shell;
#pragma unroll
    for (int p = 0; p < kPair; p++) {
        const long long k = k0 + p;
        const float ar = (float)(aR[p] * scale), ai = (float)(aI[p] * scale);
        const float br = (float)(bR[p] * scale), bi = (float)(bI[p] * scale);
        grid[index_from_indices(shell, j, k)] = make_float4(ar, ai, 0.f, 0.f);
        grid[index_from_indices(shell, j, k + half)] = make_float4(br, bi, 0.f, 0.f);
        if (!equator) {
            grid[index_from_indices(shell, 3LL * shell - j, k + half)] = make_float4(ar, -ai, 0.f, 0.f);
            grid[index_from_indices(shell, 3LL * shell - j, k)] = make_float4(br, -bi, 0.f, 0.f);
        }
    }
}

template <int K>
__global__ __launch_bounds__(256) void build_atoms_fp64(
    const double4 *__restrict__ atoms, const int4 *__restrict__ tiles, int ntiles,
    const double *__restrict__ ff, int nion, const int4 *__restrict__ bdims, double vox,
    double contrast, double step, long long nhalf, double4 *__restrict__ grid, double scale) {
    const long long c0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * K;
    double qx[K], qy[K], qz[K], accR[K], accI[K];
    NodeInfo node[K];
#pragma unroll
    for (int p = 0; p < K; p++) {
        long long c = c0 + p;
        if (c >= nhalf) c = nhalf - 1;
        node[p] = canonical_node(c, step, qx[p], qy[p], qz[p]);
        accR[p] = accI[p] = 0.0;
    }
    for (int t = 0; t < ntiles; t++) {
        const int4 tl = tiles[t];
        double sr[K], si[K];
#pragma unroll
        for (int p = 0; p < K; p++) sr[p] = si[p] = 0.0;
        const int a1 = tl.x + tl.y;
        for (int a = tl.x; a < a1; a++) {
            const double4 at = atoms[a];
#pragma unroll
            for (int p = 0; p < K; p++) {
                double ph = fma(qx[p], at.x, fma(qy[p], at.y, qz[p] * at.z));
                double s, c;
                sincos(ph, &s, &c);
                sr[p] += c;
                si[p] += s;
            }
        }
#pragma unroll
        for (int p = 0; p < K; p++) {
            double f = (tl.w == 0) ? ff[(size_t)node[p].shell * nion + tl.z]
                                   : contrast * box_weight64(qx[p], qy[p], qz[p], bdims[tl.z], vox);
            accR[p] = fma(f, sr[p], accR[p]);
            accI[p] = fma(f, si[p], accI[p]);
        }
    }
#pragma unroll
    for (int p = 0; p < K; p++) {
        if (c0 + p < nhalf) {
            double re = accR[p] * scale, im = accI[p] * scale;
            grid[node[p].idx] = make_double4(re, im, 0.0, 0.0);
            grid[node[p].anti] = make_double4(re, -im, 0.0, 0.0);
        }
    }
}

template <typename T4>
__global__ void set_origin(T4 *grid, double re) {
    // A(0) = sum of f~(0) (+ box volumes); the spline slots of the origin are 0
    T4 g;
    g.x = re;
    g.y = 0;
    g.z = 0;
    g.w = 0;
    grid[0] = g;
}

// reference layout (complex<double> per node) <-> device node layout
template <typename T4>
__global__ void pack_values(const T4 *__restrict__ grid, long long n, double2 *__restrict__ data,
                            double2 *__restrict__ coeffs) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T4 g = grid[i];
    if (data) data[i] = make_double2((double)g.x, (double)g.y);
    if (coeffs) coeffs[i] = make_double2((double)g.z, (double)g.w);
}

template <typename T4, typename T>
__global__ void unpack_values(const double2 *__restrict__ data, long long n, T4 *__restrict__ grid) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double2 d = data[i];
    T4 g;
    g.x = (T)d.x;
    g.y = (T)d.y;
    g.z = (T)0;
    g.w = (T)0;
    grid[i] = g;
}

// ----------------------------------------------------------------- launchers
static inline unsigned blocks_for(long long n, int per_block) {
    return (unsigned)((n + per_block - 1) / per_block);
}

constexpr int kNodesPerThread = 4;

void launch_build_atoms(bool fp64, bool phase64, bool reduce, const void *atoms, const void *atoms64,
                        const int4 *tiles, int ntiles, const double *ff, int nion, const int4 *bdims,
                        double vox, double contrast, double step, long long nhalf, void *grid,
                        double scale, hipStream_t s) {
    const int K = kNodesPerThread;
    unsigned nb = blocks_for(nhalf, 256 * K);
    if (!fp64 && !phase64) {
        // nhalf = half_base(n + 1) for a grid of n shells
        const int nshell = shell_of_half(nhalf) - 1;
        const long long nunits = pair_units_before(nshell);
        const unsigned nr = blocks_for(nunits, 256);
        if (reduce)
            build_atoms_pairs<true><<<nr, 256, 0, s>>>((const float4 *)atoms, tiles, ntiles, ff, nion, bdims, vox,
                                                      contrast, step, nunits, (float4 *)grid, scale);
        else
            build_atoms_pairs<false><<<nr, 256, 0, s>>>((const float4 *)atoms, tiles, ntiles, ff, nion, bdims, vox,
                                                       contrast, step, nunits, (float4 *)grid, scale);
        return;
    }
    if (fp64) {
        build_atoms_fp64<K><<<nb, 256, 0, s>>>((const double4 *)atoms64, tiles, ntiles, ff, nion, bdims,
                                               vox, contrast, step, nhalf, (double4 *)grid, scale);
    } else if (phase64) {
        build_atoms_mixed<K, true><<<nb, 256, 0, s>>>((const float4 *)atoms, (const double4 *)atoms64,
                                                      tiles, ntiles, ff, nion, bdims, vox, contrast,
                                                      step, nhalf, (float4 *)grid, scale);
    } else {
        build_atoms_mixed<K, false><<<nb, 256, 0, s>>>((const float4 *)atoms, (const double4 *)atoms64,
                                                       tiles, ntiles, ff, nion, bdims, vox, contrast,
                                                       step, nhalf, (float4 *)grid, scale);
    }
}

void launch_set_origin(bool fp64, void *grid, double re, hipStream_t s) {
    if (fp64)
        set_origin<double4><<<1, 1, 0, s>>>((double4 *)grid, re);
    else
        set_origin<float4><<<1, 1, 0, s>>>((float4 *)grid, re);
}

void launch_pack(bool fp64, const void *grid, long long n, double2 *data, double2 *coeffs,
                 hipStream_t s) {
    unsigned nb = blocks_for(n, 256);
    if (fp64)
        pack_values<double4><<<nb, 256, 0, s>>>((const double4 *)grid, n, data, coeffs);
    else
        pack_values<float4><<<nb, 256, 0, s>>>((const float4 *)grid, n, data, coeffs);
}

void launch_unpack(bool fp64, const double2 *data, long long n, void *grid, hipStream_t s) {
    unsigned nb = blocks_for(n, 256);
    if (fp64)
        unpack_values<double4, double><<<nb, 256, 0, s>>>(data, n, (double4 *)grid);
    else
        unpack_values<float4, float><<<nb, 256, 0, s>>>(data, n, (float4 *)grid);
}

}  // namespace dplus
