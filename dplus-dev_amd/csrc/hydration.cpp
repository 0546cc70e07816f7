// hydration.cpp -- voxelised hydration ("outer solvent") layer reduced to boxes.
//
// Restates PDBAmplitude::CalculateSolventSpace (Backend/Backend/VoxelizedSolvent.cpp:4-194)
// step by step on a byte grid of voxel colours:
//   range     FindPDBRanges (PDBAmplitude.cpp:710-750): atom extents +- radius,
//             widened by max(probe, thickness) + 2 voxels and rounded to whole
//             voxels, in the reference's float/double arithmetic; the z extent
//             is padded to a multiple of 16 as SolventSpace::allocate does
//             (Amplitude.cpp:4051-4069);
//   atoms     voxels whose corner point lies within an atom radius -> ATOM
//             (MarkVoxelsNeighboringAtoms, VoxelizedSolvent.cpp:365-396);
//   probe     empty voxels within radius + probe -> PROBE;
//   bulk      26-connected flood fill from voxel (0,0,0) through empty voxels
//             -> BULK; with fill-holes the remaining empty voxels (cavities)
//             become ATOM, otherwise BULK (Floodfill3D, :398-441);
//   layer     PROBE voxels within round(probe/voxel) voxels of a BULK voxel
//             that touches PROBE -> LAYER, then the layer grown into BULK (or
//             shrunk back) by round(|thickness - probe| / voxel) voxels
//             (MarkLayerOfSolvent, :196-261); leftover PROBE -> ATOM;
//   corner    the reference's slice marking of y = 1..1/voxel, z = 1..4 for
//             every x (:104-115) is kept, so a layer voxel there is dropped
//             exactly as the reference drops it;
//   boxes     LAYER voxels (the hydration layer) and ATOM voxels (the excluded
//             volume of radius types 1, 2, 3 and 5: atoms, filled cavities and
//             the probe shell the layer leaves) reduced to boxes, each set on
//             request.  Voxel (h,k,l) is the cube of
//             side `vox` centred at (xMin + h vox, yMin + k vox, zMin + l vox),
//             the reference's box-centre convention
//             (ReduceSolventSpaceToIrregularBoxes, :263-362).  The partition is
//             a greedy 3-D meshing (grow z, then y, then x); any partition of
//             the same voxels has the same Fourier transform.
// Every marking step only ever turns voxels of one colour into another, so the
// per-atom and per-seed loops run in parallel with relaxed byte stores and give
// the same grid as the reference's sequential loops.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <deque>
#include <vector>

#include "device.hpp"

namespace dplus {

namespace {

enum : char { kEmpty = 0, kAtom = 1, kProbe = 2, kBulk = 3, kLayer = 4, kCorner = 8 };

struct Space {
    long long X = 0, Y = 0, Z = 0;  // Z padded to a multiple of 16
    std::vector<char> v;
    char *at(long long i, long long j, long long k) { return &v[(size_t)((i * Y + j) * Z + k)]; }
    char get(long long i, long long j, long long k) { return __atomic_load_n(at(i, j, k), __ATOMIC_RELAXED); }
    void put(long long i, long long j, long long k, char c) { __atomic_store_n(at(i, j, k), c, __ATOMIC_RELAXED); }
};

struct Frame {
    double xmin, ymin, zmin, vox;
};

// MarkVoxelsNeighboringAtoms: voxels of colour `from` whose corner point lies
// within rad + r_sol of an atom become `to`
void mark_atoms(Space &S, const Frame &F, int n, const float *x, const float *y, const float *z, const float *r,
                double r_sol, char from, char to) {
    parallel_for(n, [&](long long a) {
        const double dd = (double)r[a] + r_sol;
        const int dist = 2 + (int)(dd / F.vox);
        const int xc = (int)((x[a] - F.xmin) / F.vox), yc = (int)((y[a] - F.ymin) / F.vox),
                  zc = (int)((z[a] - F.zmin) / F.vox);
        for (int h = std::max(0, xc - dist); h < std::min(xc + dist, (int)S.X); h++)
            for (int k = std::max(0, yc - dist); k < std::min(yc + dist, (int)S.Y); k++)
                for (int l = std::max(0, zc - dist); l < std::min(zc + dist, (int)S.Z); l++) {
                    if (S.get(h, k, l) != from) continue;
                    const double dx = F.xmin + (double)h * F.vox - x[a], dy = F.ymin + (double)k * F.vox - y[a],
                                 dz = F.zmin + (double)l * F.vox - z[a];
                    if (std::sqrt(dx * dx + dy * dy + dz * dz) <= dd) S.put(h, k, l, to);
                }
    });
}

// MarkLayerOfSolvent: seeds are interior voxels of colour `type` with a
// `neighbor` among their 27 neighbours; voxels of colour `from` within
// round(radius / vox) voxels (Euclidean, in voxel units) of a seed become `to`
void mark_layer(Space &S, double radius, double vox, char type, char neighbor, char to, char from) {
    const int d = (int)(0.5 + radius / vox);
    std::vector<std::vector<long long>> seeds(S.X);
    parallel_for(S.X, [&](long long i) {
        if (i < 1 || i >= S.X - 1) return;
        for (long long j = 1; j < S.Y - 1; j++)
            for (long long k = 1; k < S.Z - 1; k++) {
                if (*S.at(i, j, k) != type) continue;
                bool hit = false;
                for (int a = -1; a <= 1 && !hit; a++)
                    for (int b = -1; b <= 1 && !hit; b++)
                        for (int c = -1; c <= 1 && !hit; c++) hit = *S.at(i + a, j + b, k + c) == neighbor;
                if (hit) seeds[i].push_back((j << 32) | k);
            }
    });
    std::vector<long long> flat;  // (i, j, k) triples
    for (long long i = 0; i < S.X; i++)
        for (long long jk : seeds[i]) {
            flat.push_back(i);
            flat.push_back(jk >> 32);
            flat.push_back(jk & 0xffffffffLL);
        }
    parallel_for((long long)flat.size() / 3, [&](long long s) {
        const long long hh = flat[3 * s], kk = flat[3 * s + 1], ll = flat[3 * s + 2];
        for (long long m = std::max(0LL, hh - d); m <= std::min(S.X - 1, hh + d); m++)
            for (long long nn = std::max(0LL, kk - d); nn <= std::min(S.Y - 1, kk + d); nn++)
                for (long long p = std::max(0LL, ll - d); p <= std::min(S.Z - 1, ll + d); p++) {
                    if (S.get(m, nn, p) != from) continue;
                    const double r = std::sqrt((double)((m - hh) * (m - hh) + (nn - kk) * (nn - kk) + (p - ll) * (p - ll)));
                    if (r <= d) S.put(m, nn, p, to);
                }
    });
}

}  // namespace

// greedy 3-D meshing of the voxels of one colour into boxes (grow z, then y,
// then x); the meshed voxels are cleared
static void mesh_colour(Space &S, char colour, const Frame &F, std::vector<int> &dims, std::vector<double> &com) {
    auto todo = [&](long long i, long long j, long long k) { return *S.at(i, j, k) == colour; };
    for (long long i = 0; i < S.X; i++)
        for (long long j = 0; j < S.Y; j++)
            for (long long k = 0; k < S.Z; k++) {
                if (!todo(i, j, k)) continue;
                long long nk = 1;
                while (k + nk < S.Z && todo(i, j, k + nk)) nk++;
                long long nj = 1;
                while (j + nj < S.Y) {
                    bool ok = true;
                    for (long long c = 0; c < nk && ok; c++) ok = todo(i, j + nj, k + c);
                    if (!ok) break;
                    nj++;
                }
                long long ni = 1;
                while (i + ni < S.X) {
                    bool ok = true;
                    for (long long b2 = 0; b2 < nj && ok; b2++)
                        for (long long c = 0; c < nk && ok; c++) ok = todo(i + ni, j + b2, k + c);
                    if (!ok) break;
                    ni++;
                }
                for (long long a2 = 0; a2 < ni; a2++)
                    for (long long b2 = 0; b2 < nj; b2++)
                        for (long long c = 0; c < nk; c++) *S.at(i + a2, j + b2, k + c) = kEmpty;
                dims.push_back((int)ni);
                dims.push_back((int)nj);
                dims.push_back((int)nk);
                com.push_back(F.xmin + ((double)i + 0.5 * (double)(ni - 1)) * F.vox);
                com.push_back(F.ymin + ((double)j + 0.5 * (double)(nj - 1)) * F.vox);
                com.push_back(F.zmin + ((double)k + 0.5 * (double)(nk - 1)) * F.vox);
            }
}

void solvent_boxes(int n, const double *xyz, const double *rad, double probe, double thickness, double vox,
                   bool fill_holes, bool want_layer, bool want_excl, std::vector<int> &ldims, std::vector<double> &lcom,
                   std::vector<int> &edims, std::vector<double> &ecom) {
    ldims.clear();
    lcom.clear();
    edims.clear();
    ecom.clear();
    want_layer = want_layer && probe > 0.0;  // the layer is marked only with a probe radius (:56, :75)
    if (n <= 0 || (!want_layer && !want_excl)) return;
    // the reference keeps PDB coordinates and radii in float
    std::vector<float> x(n), y(n), z(n), r(n);
    for (int a = 0; a < n; a++) {
        x[a] = (float)xyz[3 * a];
        y[a] = (float)xyz[3 * a + 1];
        z[a] = (float)xyz[3 * a + 2];
        r[a] = (float)rad[a];
    }
    // FindPDBRanges
    double xmin = x[0] - r[0], xmax = x[0] + r[0], ymin = y[0] - r[0], ymax = y[0] + r[0], zmin = z[0] - r[0],
           zmax = z[0] + r[0];
    for (int a = 1; a < n; a++) {
        xmin = std::min((float)xmin, x[a] - r[a]);
        xmax = std::max((float)xmax, x[a] + r[a]);
        ymin = std::min((float)ymin, y[a] - r[a]);
        ymax = std::max((float)ymax, y[a] + r[a]);
        zmin = std::min((float)zmin, z[a] - r[a]);
        zmax = std::max((float)zmax, z[a] + r[a]);
    }
    const double dist = std::max(probe, thickness) + 2.0 * vox;
    xmin -= dist, xmax += dist, ymin -= dist, ymax += dist, zmin -= dist, zmax += dist;
    xmax = xmin + (double)(int)(2.0 + (xmax - xmin) / vox) * vox;
    ymax = ymin + (double)(int)(2.0 + (ymax - ymin) / vox) * vox;
    zmax = zmin + (double)(int)(2.0 + (zmax - zmin) / vox) * vox;
    const long long xs = (long long)((xmax - xmin) / vox), ys = (long long)((ymax - ymin) / vox),
                    zs = (long long)((zmax - zmin) / vox);
    if (xs < 3 || ys < 3 || zs < 3) return;
    Space S;
    S.X = xs;
    S.Y = ys;
    S.Z = zs + (15 - (zs + 15) % 16);
    if ((double)S.X * S.Y * S.Z > 4e9) throw Error(kErrInvalidArgs, "solvent voxel grid too large; use a coarser voxel step");
    S.v.assign((size_t)(S.X * S.Y * S.Z), kEmpty);
    const Frame F{xmin, ymin, zmin, vox};

    mark_atoms(S, F, n, x.data(), y.data(), z.data(), r.data(), 0.0, kEmpty, kAtom);
    if (probe > 0.0) mark_atoms(S, F, n, x.data(), y.data(), z.data(), r.data(), probe, kEmpty, kProbe);
    if (fill_holes) {
        std::deque<long long> q;
        if (*S.at(0, 0, 0) == kEmpty) {
            *S.at(0, 0, 0) = kBulk;
            q.push_back(0);
        }
        while (!q.empty()) {
            const long long v = q.front();
            q.pop_front();
            const long long k = v % S.Z, j = (v / S.Z) % S.Y, i = v / (S.Y * S.Z);
            for (long long a = std::max(0LL, i - 1); a <= std::min(S.X - 1, i + 1); a++)
                for (long long b = std::max(0LL, j - 1); b <= std::min(S.Y - 1, j + 1); b++)
                    for (long long c = std::max(0LL, k - 1); c <= std::min(S.Z - 1, k + 1); c++) {
                        char *p = S.at(a, b, c);
                        if (*p != kEmpty) continue;
                        *p = kBulk;
                        q.push_back((a * S.Y + b) * S.Z + c);
                    }
        }
        for (auto &c : S.v)
            if (c == kEmpty) c = kAtom;
    } else {
        for (auto &c : S.v)
            if (c == kEmpty) c = kBulk;
    }
    if (probe > 0.0) {
        mark_layer(S, probe, vox, kBulk, kProbe, kLayer, kProbe);
        if (probe < thickness)
            mark_layer(S, std::fabs(thickness - probe), vox, kLayer, kBulk, kLayer, kBulk);
        else
            mark_layer(S, std::fabs(thickness - probe), vox, kLayer, kBulk, kBulk, kLayer);
        // leftover probe voxels are excluded volume (:100-101)
        for (auto &c : S.v)
            if (c == kProbe) c = kAtom;
        const int per_nm = (int)(1.0 / vox);
        for (long long i = 0; i < S.X; i++)
            for (int t = 0; t < per_nm && 1 + t < S.Y; t++)
                for (int zk = 1; zk <= 4 && zk < S.Z; zk++) *S.at(i, 1 + t, zk) = kCorner;
    }
    if (want_excl) mesh_colour(S, kAtom, F, edims, ecom);
    if (want_layer) mesh_colour(S, kLayer, F, ldims, lcom);
}

void hydration_boxes(int n, const double *xyz, const double *rad, double probe, double thickness, double vox,
                     bool fill_holes, std::vector<int> &dims, std::vector<double> &com) {
    std::vector<int> ed;
    std::vector<double> ec;
    solvent_boxes(n, xyz, rad, probe, thickness, vox, fill_holes, true, false, dims, com, ed, ec);
}

}  // namespace dplus

using namespace dplus;

extern "C" int dplus_excluded_volume_boxes(int n, const double *xyz, const double *rad, double probe,
                                           double thickness, double vox, int fill_holes, int capacity, int *nbox,
                                           int *dims, double *com) {
    try {
        if (n < 0 || (n > 0 && (!xyz || !rad)) || !nbox || vox <= 0 || probe < 0 || thickness < 0)
            throw Error(kErrInvalidArgs, "bad excluded-volume arguments");
        std::vector<int> ld, d;
        std::vector<double> lc, c;
        solvent_boxes(n, xyz, rad, probe, thickness, vox, fill_holes != 0, false, true, ld, lc, d, c);
        *nbox = (int)(d.size() / 3);
        if (*nbox > capacity) return kOk;  // caller retries with capacity >= *nbox
        if (*nbox) {
            std::copy(d.begin(), d.end(), dims);
            std::copy(c.begin(), c.end(), com);
        }
        return kOk;
    } catch (const Error &e) {
        return set_error(e.code, e.what());
    } catch (const std::exception &e) {
        return set_error(kErrGeneral, e.what());
    }
}

extern "C" int dplus_hydration_boxes(int n, const double *xyz, const double *rad, double probe, double thickness,
                                     double vox, int fill_holes, int capacity, int *nbox, int *dims, double *com) {
    try {
        if (n < 0 || (n > 0 && (!xyz || !rad)) || !nbox || vox <= 0 || probe < 0 || thickness < 0)
            throw Error(kErrInvalidArgs, "bad hydration arguments");
        std::vector<int> d;
        std::vector<double> c;
        hydration_boxes(n, xyz, rad, probe, thickness, vox, fill_holes != 0, d, c);
        *nbox = (int)(d.size() / 3);
        if (*nbox > capacity) return kOk;  // caller retries with capacity >= *nbox
        if (*nbox) {
            std::copy(d.begin(), d.end(), dims);
            std::copy(c.begin(), c.end(), com);
        }
        return kOk;
    } catch (const Error &e) {
        return set_error(e.code, e.what());
    } catch (const std::exception &e) {
        return set_error(kErrGeneral, e.what());
    }
}
