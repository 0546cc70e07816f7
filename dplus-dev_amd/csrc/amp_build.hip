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
//  * Atoms are sorted by ion and cut into tiles of <= kTile atoms of one ion;
//    solvent boxes into tiles of one shape.  The atom loop index is
//    wave-uniform, so coordinates arrive through scalar loads (s_load) and feed
//    v_fma as SGPR operands.
//  * Mixed precision (build_atoms_pairs): q is pre-scaled to revolutions and the
//    phases are FP32 FMA chains fed to the hardware v_sin_f32 / v_cos_f32.  The
//    nodes phi and phi + pi of a row are evaluated together from sin/cos of
//    q_z z (shared by the row, through LDS) and of q_x x + q_y y, one
//    transcendental per term.  Each tile is summed in FP32 and folded into FP64
//    accumulators with the ion's form factor at the node's shell (or the box
//    shape's weight): a two-level compensated sum whose FP32 partials never
//    exceed kTile terms.  One launch builds up to kMaxBuildJobs grids, with a
//    workgroup size (64..512 units) and atom-slice count chosen per grid
//    class by the host's autotuner or occupancy model.
//  * FP64 mode (build_atoms_pairs64): the paired builder's algebra in double --
//    double coordinates, phases, sincos (of q_x x + q_y y per node pair, of
//    q_z z once per row and atom through LDS) and sums.  The plain form
//    (build_atoms_fp64: one node at a time, K nodes per thread, one double
//    sincos per term) stays selectable as the baseline FP64 path.

#include <climits>
#include <mutex>

#include "device.hpp"

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

// L sinc(q L / 2) for a box edge of n voxels, L = n vox (q in nm^-1): sin on the
// hardware unit with the argument in revolutions reduced to [-1/2, 1/2], the
// quotient by the hardware reciprocal; below |x| = 1e-3 the series 1 - x^2/6
__device__ __forceinline__ float edge_weight(float q, int n, float vox) {
    const float L = (float)n * vox, x = 0.5f * q * L;
    float r = x * (float)(1.0 / k2Pi);
    r -= rintf(r);
    const float s = __builtin_amdgcn_sinf(r);
    return (fabsf(x) < 1e-3f || n == 0) ? L * (1.0f - x * x * (1.0f / 6.0f)) : s * __builtin_amdgcn_rcpf(x) * L;
}

// Tiles: (first, count, index, kind).  kind 0: atoms of ion `index`, weighted by
// the ion's form factor at the node's shell; kind 1: solvent boxes of shape
// bdims[index], weighted by that shape's contrast bcon[index] * box_weight at
// the node (hydration-layer and excluded-volume boxes carry their own contrast).
//
// build_atoms_phase64: FP32 grid and FP32 tile sums, but every phase q.r formed
// and reduced in FP64 -- a diagnostic form of the mixed builder (environment
// variable DPLUS_PHASE64) that isolates the FP32 phase error.
template <int K>
__global__ __launch_bounds__(256) void build_atoms_phase64(
    const double4 *__restrict__ atoms64, const int4 *__restrict__ tiles, int ntiles, const double *__restrict__ ff, int nion,
    const int4 *__restrict__ bdims, const double *__restrict__ bcon, double vox,
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
            const double4 at = atoms64[a];
#pragma unroll
            for (int p = 0; p < K; p++) {
                double ph = fma(dqx[p], at.x, fma(dqy[p], at.y, dqz[p] * at.z));
                float r = (float)(ph - floor(ph));
                sr[p] += __builtin_amdgcn_cosf(r);
                si[p] += __builtin_amdgcn_sinf(r);
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
                double f = bcon[tl.z] * (double)box_weight32(qx[p] * k2, qy[p] * k2, qz[p] * k2, d, (float)vox);
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
// One unit (thread) owns three such pairs at azimuths phi, phi + 2 pi / 3 and
// phi + 4 pi / 3 (indices k0, k0 + 2i, k0 + 4i of the row's 6i: the unit's six
// nodes are k0 + m i, m = 0..5).  Their in-plane phases sum to zero,
// u0 + u1 + u2 = 0 (cos phi + cos(phi + 2 pi/3) + cos(phi + 4 pi/3) = 0, and the
// same for sin), so e^{i u2} = conj(e^{i u0} e^{i u1}): per atom the unit takes
// sin/cos of u0 and u1 only and forms the third pair's cos/sin with four
// products, then sums the four products cz cu, sz su, sz cu, cz su of each
// pair: two transcendentals per six terms.  cz, sz depend only on the row and the atom, so a
// workgroup (whose kWg units span at most kRows rows for all but the smallest
// shells) evaluates them once per (row, atom) into LDS, chunk by chunk,
// double-buffered with one barrier per chunk; a workgroup spanning more rows
// takes them per unit.  Box weights are even in (q_x, q_y), so one per pair.
// Rows off the equator also write both antipodes; on the equator (j = 3i/2)
// the partner is the antipode itself.
// Every workgroup sums all atoms of its slice, so the atoms are cut into
// slices (tile ranges) to fill the device's last round of resident
// workgroups; slices end in FP64 partial sums added in order.
// One launch serves up to kMaxBuildJobs grids (of different structures, grid
// sizes and slice counts): workgroups are numbered job by job, slice by slice,
// and each finds its job from the launch arguments.
// With kReduce false the phases (in revolutions) go to v_sin_f32 / v_cos_f32
// unreduced; the host selects that form only when every |q.r| / 2 pi of the
// structure is below 128 revolutions (|z| and |u| are both bounded by |q||r|),
// inside the units' input domain.
constexpr int kPair = 3;
constexpr int kChunk = 64;

// pair unit u -> its row and first azimuthal index (k0 < i: the unit's pairs
// are at k0, k0 + 2i, k0 + 4i)
struct PairUnit {
    int shell;
    long long j, k0;
};

__device__ __forceinline__ PairUnit pair_unit(long long u) {
    PairUnit pu;
    pu.shell = shell_of_pair_unit(u);
    const long long r = u - pair_units_before(pu.shell - 1);
    pu.j = r / pu.shell;
    pu.k0 = r % pu.shell;
    return pu;
}

// azimuthal index of pair p of a unit
__device__ __forceinline__ long long pair_k(const PairUnit &pu, int p) { return pu.k0 + 2LL * pu.shell * p; }

// the four nodes of pair p of a unit: (j, k), (j, k + 3i) and, off the
// equator, their antipodes (3i - j, k + 3i), (3i - j, k); azimuths mod 6i
__device__ __forceinline__ void store_pair(float4 *__restrict__ grid, const PairUnit &pu, int p, double aR,
                                           double aI, double bR, double bI, double scale) {
    const long long half = 3LL * pu.shell, k = pair_k(pu, p), kb = k + half >= 2 * half ? k - half : k + half;
    const float ar = (float)(aR * scale), ai = (float)(aI * scale);
    const float br = (float)(bR * scale), bi = (float)(bI * scale);
    grid[index_from_indices(pu.shell, pu.j, k)] = make_float4(ar, ai, 0.f, 0.f);
    grid[index_from_indices(pu.shell, pu.j, kb)] = make_float4(br, bi, 0.f, 0.f);
    if (2 * pu.j != half) {
        grid[index_from_indices(pu.shell, half - pu.j, kb)] = make_float4(ar, -ai, 0.f, 0.f);
        grid[index_from_indices(pu.shell, half - pu.j, k)] = make_float4(br, -bi, 0.f, 0.f);
    }
}

__device__ __forceinline__ void store_pair64(double4 *__restrict__ grid, const PairUnit &pu, int p, double aR,
                                             double aI, double bR, double bI, double scale) {
    const long long half = 3LL * pu.shell, k = pair_k(pu, p), kb = k + half >= 2 * half ? k - half : k + half;
    const double ar = aR * scale, ai = aI * scale, br = bR * scale, bi = bI * scale;
    grid[index_from_indices(pu.shell, pu.j, k)] = make_double4(ar, ai, 0.0, 0.0);
    grid[index_from_indices(pu.shell, pu.j, kb)] = make_double4(br, bi, 0.0, 0.0);
    if (2 * pu.j != half) {
        grid[index_from_indices(pu.shell, half - pu.j, kb)] = make_double4(ar, -ai, 0.0, 0.0);
        grid[index_from_indices(pu.shell, half - pu.j, k)] = make_double4(br, -bi, 0.0, 0.0);
    }
}

// job of workgroup `wg` (jobs are numbered in order of their first workgroup)
__device__ __forceinline__ int batch_job(const BuildBatch &b, int wg, bool finishing) {
    int jb = 0;
    for (int k = 1; k < b.njobs; k++)
        if (wg >= (finishing ? b.job[k].fin_begin : b.job[k].wg_begin)) jb = k;
    return jb;
}

// one atom's terms for a unit's three pairs: sin/cos of u0 and u1 (phases in
// revolutions), e^{i u2} = conj(e^{i u0} e^{i u1}), then per pair the four
// products with the row's cz, sz.  t2 = s0 c1 + c0 s1 = -sin u2.
template <bool kReduce>
__device__ __forceinline__ void pair_terms(const float qx[kPair], const float qy[kPair], const float4 &at, float cz,
                                           float sz, float cc[kPair], float ss[kPair], float sc[kPair],
                                           float cs[kPair]) {
    float cu[2], su[2];
#pragma unroll
    for (int p = 0; p < 2; p++) {
        const float ph = fmaf(qx[p], at.x, qy[p] * at.y);
        const float v = kReduce ? __builtin_amdgcn_fractf(ph) : ph;
        cu[p] = __builtin_amdgcn_cosf(v);
        su[p] = __builtin_amdgcn_sinf(v);
        cc[p] = fmaf(cz, cu[p], cc[p]);
        ss[p] = fmaf(sz, su[p], ss[p]);
        sc[p] = fmaf(sz, cu[p], sc[p]);
        cs[p] = fmaf(cz, su[p], cs[p]);
    }
    const float c2 = fmaf(cu[0], cu[1], -su[0] * su[1]);
    const float t2 = fmaf(su[0], cu[1], cu[0] * su[1]);
    cc[2] = fmaf(cz, c2, cc[2]);
    ss[2] = fmaf(-sz, t2, ss[2]);
    sc[2] = fmaf(sz, c2, sc[2]);
    cs[2] = fmaf(-cz, t2, cs[2]);
}

__device__ __forceinline__ void pair_terms64(const double qx[kPair], const double qy[kPair], const double4 &at,
                                             double cz, double sz, double cc[kPair], double ss[kPair],
                                             double sc[kPair], double cs[kPair]) {
    double cu[2], su[2];
#pragma unroll
    for (int p = 0; p < 2; p++) {
        sincos(fma(qx[p], at.x, qy[p] * at.y), &su[p], &cu[p]);
        cc[p] = fma(cz, cu[p], cc[p]);
        ss[p] = fma(sz, su[p], ss[p]);
        sc[p] = fma(sz, cu[p], sc[p]);
        cs[p] = fma(cz, su[p], cs[p]);
    }
    const double c2 = fma(cu[0], cu[1], -su[0] * su[1]);
    const double t2 = fma(su[0], cu[1], cu[0] * su[1]);
    cc[2] = fma(cz, c2, cc[2]);
    ss[2] = fma(-sz, t2, ss[2]);
    sc[2] = fma(sz, c2, sc[2]);
    cs[2] = fma(-cz, t2, cs[2]);
}

// LDS rows of a workgroup of kWg units (kWg units of shell i span at most
// ceil(kWg / i) + 1 rows, so these cover every shell from i = 10 (64 units),
// 19 (128), 18 (256) or 17 (512) up)
template <int kWg>
constexpr int pair_rows() {
    return kWg <= 128 ? 8 : kWg == 256 ? 16 : 32;
}

// Workgroup w of job J, slice y = (w - J.wg_begin) / J.xblocks, sums the tiles
// of slice y for units kWg x .. kWg x + kWg - 1 (x the rest) and leaves its FP64
// sums, unscaled, in J.partial[(y * 12 + 4 p + m) * nunits + u] (m = aR, aI,
// bR, bI); pairs_finish adds the slices in order and stores.
template <bool kReduce, int kWg>
__global__ __launch_bounds__(kWg) void build_atoms_pairs(BuildBatch batch) {
    constexpr int kRows = pair_rows<kWg>();
    __shared__ float2 rowz[2][kRows][kChunk];
    __shared__ float rowqz[kRows];
    const BuildJob &J = batch.job[batch_job(batch, (int)blockIdx.x, false)];
    const float4 *__restrict__ atoms = J.atoms;
    const int4 *__restrict__ tiles = J.tiles;
    const long long nunits = J.nunits;
    const int local = (int)blockIdx.x - J.wg_begin;
    const int slice = local / J.xblocks;
    const long long u0 = (long long)(local - slice * J.xblocks) * kWg;
    const long long u = u0 + threadIdx.x;
    const bool live = u < nunits;
    const long long uc = live ? u : nunits - 1;
    const PairUnit pu = pair_unit(uc);
    const int shell = pu.shell;
    const long long j = pu.j;
    // rows of this workgroup: global row index of the first and last unit
    int lrow, nrows;
    {
        const int s0 = shell_of_pair_unit(u0);
        const long long g0 = rows_before(s0 - 1) + (u0 - pair_units_before(s0 - 1)) / s0;
        const long long ul = min(u0 + (long long)kWg, nunits) - 1;
        const int s1 = shell_of_pair_unit(ul);
        const long long g1 = rows_before(s1 - 1) + (ul - pair_units_before(s1 - 1)) / s1;
        lrow = (int)(rows_before(shell - 1) + j - g0);
        nrows = (int)(g1 - g0 + 1);
    }
    const bool shared_z = nrows <= kRows;  // uniform over the workgroup
    float qx[kPair], qy[kPair], qz;
    double aR[kPair], aI[kPair], bR[kPair], bI[kPair];
    const double inv2pi = 1.0 / k2Pi;
    {
        const double q = (double)shell * J.step;
        double st, ct;
        sincos(kPi * (double)j / (double)(3 * shell), &st, &ct);
        qz = (float)(q * ct * inv2pi);
#pragma unroll
        for (int p = 0; p < kPair; p++) {
            double sp, cp;
            sincos(k2Pi * (double)pair_k(pu, p) / (double)(6 * shell), &sp, &cp);
            qx[p] = (float)(q * st * cp * inv2pi);
            qy[p] = (float)(q * st * sp * inv2pi);
            aR[p] = aI[p] = bR[p] = bI[p] = 0.0;
        }
    }
    if (shared_z) {
        // any unit of a row may publish the row's q_z: all write the same value
        rowqz[lrow] = qz;
        __syncthreads();
    }
    int buf = 0;
    const int t_end = J.slice_begin[slice + 1];
    for (int t = J.slice_begin[slice]; t < t_end; t++) {
        const int4 tl = tiles[t];
        // the ion's form factor at this shell, loaded now and used after the atom loop
        const double fion = tl.w == 0 ? J.ff[(size_t)shell * J.nion + tl.z] : 0.0;
        float cc[kPair], ss[kPair], sc[kPair], cs[kPair];
#pragma unroll
        for (int p = 0; p < kPair; p++) cc[p] = ss[p] = sc[p] = cs[p] = 0.f;
        const int a1 = tl.x + tl.y;
        if (shared_z) {
            for (int c0 = tl.x; c0 < a1; c0 += kChunk) {
                const int n = min(kChunk, a1 - c0);
                // cos / sin of q_z z for every (row, atom) of the chunk, once per workgroup
                for (int e = threadIdx.x; e < nrows * kChunk; e += kWg) {
                    const int rr = e / kChunk, c = e % kChunk;
                    if (c < n) {
                        const float zz = rowqz[rr] * atoms[c0 + c].z;
                        const float v = kReduce ? __builtin_amdgcn_fractf(zz) : zz;
                        rowz[buf][rr][c] = make_float2(__builtin_amdgcn_cosf(v), __builtin_amdgcn_sinf(v));
                    }
                }
                __syncthreads();
#pragma unroll 2
                for (int c = 0; c < n; c++) {
                    const float4 at = atoms[c0 + c];
                    const float2 z = rowz[buf][lrow][c];
                    pair_terms<kReduce>(qx, qy, at, z.x, z.y, cc, ss, sc, cs);
                }
                buf ^= 1;
            }
        } else {
#pragma unroll 2
            for (int a = tl.x; a < a1; a++) {
                const float4 at = atoms[a];
                const float zz = kReduce ? __builtin_amdgcn_fractf(qz * at.z) : qz * at.z;
                const float cz = __builtin_amdgcn_cosf(zz), sz = __builtin_amdgcn_sinf(zz);
                pair_terms<kReduce>(qx, qy, at, cz, sz, cc, ss, sc, cs);
            }
        }
        // box tiles: contrast * prod_a L_a sinc(q_a L_a / 2); the z edge is common to the row
        const float k2 = (float)k2Pi;
        const float vox = (float)J.vox;
        int4 d = make_int4(0, 0, 0, 0);
        float wz = 0.f;
        if (tl.w != 0) {
            d = J.bdims[tl.z];
            wz = edge_weight(qz * k2, d.z, vox);
        }
#pragma unroll
        for (int p = 0; p < kPair; p++) {
            const double f = tl.w == 0 ? fion
                                       : J.bcon[tl.z] * (double)(wz * edge_weight(qx[p] * k2, d.x, vox) *
                                                               edge_weight(qy[p] * k2, d.y, vox));
            aR[p] = fma(f, (double)(cc[p] - ss[p]), aR[p]);
            aI[p] = fma(f, (double)(sc[p] + cs[p]), aI[p]);
            bR[p] = fma(f, (double)(cc[p] + ss[p]), bR[p]);
            bI[p] = fma(f, (double)(sc[p] - cs[p]), bI[p]);
        }
    }
    if (!live) return;
    double *dst = J.partial + (size_t)slice * 12 * nunits + u;
#pragma unroll
    for (int p = 0; p < kPair; p++) {
        dst[(4 * p + 0) * nunits] = aR[p];
        dst[(4 * p + 1) * nunits] = aI[p];
        dst[(4 * p + 2) * nunits] = bR[p];
        dst[(4 * p + 3) * nunits] = bI[p];
    }
}

// FP64 paired builder: the unit layout, slices and partials of build_atoms_pairs
// with every quantity in double -- phases in radians, sin/cos by the double
// sincos (full range reduction), cos/sin of q_z z shared by the row through
// LDS, tile sums folded with the form factor or box weight in double.  The
// values are the plain FP64 builder's up to rounding order.
template <int kWg>
__global__ __launch_bounds__(kWg) void build_atoms_pairs64(BuildBatch batch) {
    constexpr int kRows = pair_rows<kWg>();
    constexpr int kChunk64 = 32;
    __shared__ double2 rowz[2][kRows][kChunk64];
    __shared__ double rowqz[kRows];
    const BuildJob &J = batch.job[batch_job(batch, (int)blockIdx.x, false)];
    const double4 *__restrict__ atoms = J.atoms64;
    const int4 *__restrict__ tiles = J.tiles;
    const long long nunits = J.nunits;
    const int local = (int)blockIdx.x - J.wg_begin;
    const int slice = local / J.xblocks;
    const long long u0 = (long long)(local - slice * J.xblocks) * kWg;
    const long long u = u0 + threadIdx.x;
    const bool live = u < nunits;
    const long long uc = live ? u : nunits - 1;
    const PairUnit pu = pair_unit(uc);
    const int shell = pu.shell;
    const long long j = pu.j;
    int lrow, nrows;
    {
        const int s0 = shell_of_pair_unit(u0);
        const long long g0 = rows_before(s0 - 1) + (u0 - pair_units_before(s0 - 1)) / s0;
        const long long ul = min(u0 + (long long)kWg, nunits) - 1;
        const int s1 = shell_of_pair_unit(ul);
        const long long g1 = rows_before(s1 - 1) + (ul - pair_units_before(s1 - 1)) / s1;
        lrow = (int)(rows_before(shell - 1) + j - g0);
        nrows = (int)(g1 - g0 + 1);
    }
    const bool shared_z = nrows <= kRows;
    double qx[kPair], qy[kPair], qz;
    double aR[kPair], aI[kPair], bR[kPair], bI[kPair];
    {
        const double q = (double)shell * J.step;
        double st, ct;
        sincos(kPi * (double)j / (double)(3 * shell), &st, &ct);
        qz = q * ct;
#pragma unroll
        for (int p = 0; p < kPair; p++) {
            double sp, cp;
            sincos(k2Pi * (double)pair_k(pu, p) / (double)(6 * shell), &sp, &cp);
            qx[p] = q * st * cp;
            qy[p] = q * st * sp;
            aR[p] = aI[p] = bR[p] = bI[p] = 0.0;
        }
    }
    if (shared_z) {
        rowqz[lrow] = qz;
        __syncthreads();
    }
    int buf = 0;
    const int t_end = J.slice_begin[slice + 1];
    for (int t = J.slice_begin[slice]; t < t_end; t++) {
        const int4 tl = tiles[t];
        const double fion = tl.w == 0 ? J.ff[(size_t)shell * J.nion + tl.z] : 0.0;
        double cc[kPair], ss[kPair], sc[kPair], cs[kPair];
#pragma unroll
        for (int p = 0; p < kPair; p++) cc[p] = ss[p] = sc[p] = cs[p] = 0.0;
        const int a1 = tl.x + tl.y;
        if (shared_z) {
            for (int c0 = tl.x; c0 < a1; c0 += kChunk64) {
                const int n = min(kChunk64, a1 - c0);
                for (int e = threadIdx.x; e < nrows * kChunk64; e += kWg) {
                    const int rr = e / kChunk64, c = e % kChunk64;
                    if (c < n) {
                        double sz, cz;
                        sincos(rowqz[rr] * atoms[c0 + c].z, &sz, &cz);
                        rowz[buf][rr][c] = make_double2(cz, sz);
                    }
                }
                __syncthreads();
                for (int c = 0; c < n; c++) {
                    const double4 at = atoms[c0 + c];
                    const double2 z = rowz[buf][lrow][c];
                    pair_terms64(qx, qy, at, z.x, z.y, cc, ss, sc, cs);
                }
                buf ^= 1;
            }
        } else {
            for (int a = tl.x; a < a1; a++) {
                const double4 at = atoms[a];
                double sz, cz;
                sincos(qz * at.z, &sz, &cz);
                pair_terms64(qx, qy, at, cz, sz, cc, ss, sc, cs);
            }
        }
#pragma unroll
        for (int p = 0; p < kPair; p++) {
            const double f = tl.w == 0 ? fion : J.bcon[tl.z] * box_weight64(qx[p], qy[p], qz, J.bdims[tl.z], J.vox);
            aR[p] = fma(f, cc[p] - ss[p], aR[p]);
            aI[p] = fma(f, sc[p] + cs[p], aI[p]);
            bR[p] = fma(f, cc[p] + ss[p], bR[p]);
            bI[p] = fma(f, sc[p] - cs[p], bI[p]);
        }
    }
    if (!live) return;
    double *dst = J.partial + (size_t)slice * 12 * nunits + u;
#pragma unroll
    for (int p = 0; p < kPair; p++) {
        dst[(4 * p + 0) * nunits] = aR[p];
        dst[(4 * p + 1) * nunits] = aI[p];
        dst[(4 * p + 2) * nunits] = bR[p];
        dst[(4 * p + 3) * nunits] = bI[p];
    }
}

// Sum of the atom slices of every pair unit of every job, in slice order,
// scaled and stored; workgroup w serves units 256 (w - J.fin_begin) + .. of job J.
__global__ __launch_bounds__(256) void pairs_finish(BuildBatch batch) {
    const BuildJob &J = batch.job[batch_job(batch, (int)blockIdx.x, true)];
    const long long nunits = J.nunits;
    const long long u = (long long)((int)blockIdx.x - J.fin_begin) * 256 + threadIdx.x;
    if (u >= nunits) return;
    const PairUnit pu = pair_unit(u);
    const double *__restrict__ partial = J.partial;
    for (int p = 0; p < kPair; p++) {
        double v[4] = {0.0, 0.0, 0.0, 0.0};
        for (int y = 0; y < J.nslices; y++)
#pragma unroll
            for (int m = 0; m < 4; m++) v[m] += partial[((size_t)y * 12 + 4 * p + m) * nunits + u];
        if (J.fp64)
            store_pair64(J.grid64, pu, p, v[0], v[1], v[2], v[3], J.scale);
        else
            store_pair(J.grid, pu, p, v[0], v[1], v[2], v[3], J.scale);
    }
}

template <int K>
__global__ __launch_bounds__(256) void build_atoms_fp64(
    const double4 *__restrict__ atoms, const int4 *__restrict__ tiles, int ntiles,
    const double *__restrict__ ff, int nion, const int4 *__restrict__ bdims, const double *__restrict__ bcon,
    double vox, double step, long long nhalf, double4 *__restrict__ grid, double scale) {
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
                                   : bcon[tl.z] * box_weight64(qx[p], qy[p], qz[p], bdims[tl.z], vox);
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
                        const double *bcon, double vox, double step, long long nhalf, void *grid,
                        double scale, hipStream_t s) {
    const int K = kNodesPerThread;
    unsigned nb = blocks_for(nhalf, 256 * K);
    if (fp64) {
        build_atoms_fp64<K><<<nb, 256, 0, s>>>((const double4 *)atoms64, tiles, ntiles, ff, nion, bdims,
                                               bcon, vox, step, nhalf, (double4 *)grid, scale);
    } else {
        // mixed grids reach this launcher only for FP64 phases (the paired builder
        // takes every other mixed build)
        (void)phase64;
        build_atoms_phase64<K><<<nb, 256, 0, s>>>((const double4 *)atoms64, tiles, ntiles, ff, nion, bdims,
                                                  bcon, vox, step, nhalf, (float4 *)grid, scale);
    }
}

long long pair_units(long long nhalf) {
    // nhalf = half_base(n + 1) for a grid of n shells
    return pair_units_before(shell_of_half(nhalf) - 1);
}

// Resident workgroups of the paired builder (mixed or FP64) of workgroup size wg on this device.
static int pair_resident(int wg, bool fp64) {
    static int cached[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    static std::mutex mu;
    const int slot = (wg == 64 ? 0 : wg == 128 ? 1 : wg == 256 ? 2 : 3) + (fp64 ? 4 : 0);
    std::lock_guard<std::mutex> lk(mu);
    if (cached[slot] < 0) {
        const void *fn = fp64 ? (wg == 64    ? (const void *)build_atoms_pairs64<64>
                                 : wg == 128 ? (const void *)build_atoms_pairs64<128>
                                 : wg == 256 ? (const void *)build_atoms_pairs64<256>
                                             : (const void *)build_atoms_pairs64<512>)
                              : (wg == 64    ? (const void *)build_atoms_pairs<false, 64>
                                 : wg == 128 ? (const void *)build_atoms_pairs<false, 128>
                                 : wg == 256 ? (const void *)build_atoms_pairs<false, 256>
                                             : (const void *)build_atoms_pairs<false, 512>);
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, wg, 0) != hipSuccess)
            cached[slot] = 0;
        else
            cached[slot] = cus * per_cu;
    }
    return cached[slot];
}

// Number of atom slices of the paired builder (occupancy model).  Every
// workgroup sums all atoms of its slice, so a launch runs in whole rounds of
// the device's resident workgroups; S slices turn W workgroups into W S
// shorter ones.  S minimises ceil(W S / resident) / S, with 1 % per extra
// slice for its partial sums.
int pair_slices(long long nunits, int ntiles, int wg, bool fp64) {
    const int resident = pair_resident(wg, fp64);
    if (resident <= 0) return 1;
    const long long W = blocks_for(nunits, wg);
    int best = 1;
    double best_cost = 1e300;
    for (int S = 1; S <= std::min(kMaxSlices, std::max(ntiles, 1)); S++) {
        const double cost = std::ceil((double)(W * S) / resident) / S * (1.0 + 0.01 * (S - 1));
        if (cost < best_cost - 1e-12) {
            best_cost = cost;
            best = S;
        }
    }
    return best;
}

template <bool kReduce>
static void launch_pairs(int wg, const BuildBatch &b, unsigned nblocks, hipStream_t s) {
    switch (wg) {
        case 64:
            build_atoms_pairs<kReduce, 64><<<nblocks, 64, 0, s>>>(b);
            break;
        case 128:
            build_atoms_pairs<kReduce, 128><<<nblocks, 128, 0, s>>>(b);
            break;
        case 256:
            build_atoms_pairs<kReduce, 256><<<nblocks, 256, 0, s>>>(b);
            break;
        default:
            build_atoms_pairs<kReduce, 512><<<nblocks, 512, 0, s>>>(b);
            break;
    }
}

template <int kWg>
static void launch_pairs64(const BuildBatch &b, unsigned nblocks, hipStream_t s) {
    build_atoms_pairs64<kWg><<<nblocks, kWg, 0, s>>>(b);
}

static bool plain_fp64_flag = getenv("DPLUS_FP64_PLAIN") != nullptr;
bool fp64_plain_builder() { return plain_fp64_flag; }
void set_fp64_plain_builder(bool plain) { plain_fp64_flag = plain; }

void launch_build_batch(bool reduce, int wg, BuildJob *jobs, int njobs, hipStream_t s) {
    if (njobs < 1 || njobs > kMaxBuildJobs) throw Error(kErrInvalidArgs, "builder batch of 1..16 jobs");
    if (wg != 64 && wg != 128 && wg != 256 && wg != 512)
        throw Error(kErrInvalidArgs, "builder workgroup size must be 64, 128, 256 or 512");
    const bool fp64 = jobs[0].fp64 != 0;
    BuildBatch b{};
    b.njobs = njobs;
    long long wgs = 0, fins = 0;
    for (int k = 0; k < njobs; k++) {
        BuildJob &J = jobs[k];
        if ((J.fp64 != 0) != fp64) throw Error(kErrInvalidArgs, "a builder batch is all mixed or all FP64");
        if (J.nslices < 1 || J.nslices > kMaxSlices || J.nunits < 1)
            throw Error(kErrInvalidArgs, "builder job with no units or a slice count outside 1..8");
        J.xblocks = (int)blocks_for(J.nunits, wg);
        J.wg_begin = (int)wgs;
        J.fin_begin = (int)fins;
        wgs += (long long)J.xblocks * J.nslices;
        fins += blocks_for(J.nunits, 256);
        b.job[k] = J;
    }
    if (wgs > INT_MAX || fins > INT_MAX) throw Error(kErrInvalidArgs, "builder batch too large for one launch");
    if (fp64) {
        switch (wg) {
            case 64: launch_pairs64<64>(b, (unsigned)wgs, s); break;
            case 128: launch_pairs64<128>(b, (unsigned)wgs, s); break;
            case 256: launch_pairs64<256>(b, (unsigned)wgs, s); break;
            default: launch_pairs64<512>(b, (unsigned)wgs, s); break;
        }
    } else if (reduce) {
        launch_pairs<true>(wg, b, (unsigned)wgs, s);
    } else {
        launch_pairs<false>(wg, b, (unsigned)wgs, s);
    }
    pairs_finish<<<(unsigned)fins, 256, 0, s>>>(b);
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
