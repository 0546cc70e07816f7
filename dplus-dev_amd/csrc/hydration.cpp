// hydration.cpp -- voxelised hydration layer reduced to solvent boxes.
//
// Follows the steps of PDBAmplitude::CalculateSolventSpace
// (Backend/Backend/VoxelizedSolvent.cpp:4-194): mark voxels inside atoms,
// mark the probe shell around them, flood-fill the bulk solvent from a corner
// of the box, take the bulk voxels within the solvation thickness of the
// probe shell as the hydration layer, and reduce the layer to boxes
// (ReduceSolventSpaceToIrregularBoxes, :263-363).  Distances are measured
// from voxel centres to atom surfaces (|c - r_atom| - radius).  The box
// decomposition is a greedy 3-D meshing (grow z, then y, then x); any
// decomposition gives the same Fourier transform, fewer boxes cost less.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <deque>
#include <vector>

#include "device.hpp"

namespace dplus {

void hydration_boxes(int n, const double *xyz, const double *rad, double probe, double thickness, double vox,
                     std::vector<int> &dims, std::vector<double> &com) {
    dims.clear();
    com.clear();
    if (n <= 0) return;
    double rmax = 0.0, lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int a = 0; a < n; a++) {
        rmax = std::max(rmax, rad[a]);
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], xyz[3 * a + k]);
            hi[k] = std::max(hi[k], xyz[3 * a + k]);
        }
    }
    const double reach = rmax + probe + thickness;
    const double margin = reach + 2 * vox;
    long long dim[3];
    double org[3];
    for (int k = 0; k < 3; k++) {
        org[k] = lo[k] - margin;
        dim[k] = (long long)std::ceil((hi[k] - lo[k] + 2 * margin) / vox);
    }
    const long long nv = dim[0] * dim[1] * dim[2];
    if (nv > 2000000000LL) throw Error(kErrInvalidArgs, "hydration voxel grid too large; use a coarser voxel step");
    auto vid = [&](long long i, long long j, long long k) { return (i * dim[1] + j) * dim[2] + k; };
    // distance to the nearest atom surface, stamped atom by atom; x-slabs are
    // owned by one thread each so the min-updates never race
    std::vector<float> dist(nv, 1e30f);
    std::vector<std::vector<int>> slab_atoms(dim[0]);
    for (int a = 0; a < n; a++) {
        const double R = rad[a] + probe + thickness;
        long long i0 = std::max(0LL, (long long)std::floor((xyz[3 * a] - R - org[0]) / vox));
        long long i1 = std::min(dim[0] - 1, (long long)std::ceil((xyz[3 * a] + R - org[0]) / vox));
        for (long long i = i0; i <= i1; i++) slab_atoms[i].push_back(a);
    }
    parallel_for(dim[0], [&](long long i) {
        const double cx = org[0] + (i + 0.5) * vox;
        for (int a : slab_atoms[i]) {
            const double R = rad[a] + probe + thickness;
            const double dx = cx - xyz[3 * a];
            const double r2 = R * R - dx * dx;
            if (r2 < 0) continue;
            const double ry = std::sqrt(r2);
            long long j0 = std::max(0LL, (long long)std::floor((xyz[3 * a + 1] - ry - org[1]) / vox));
            long long j1 = std::min(dim[1] - 1, (long long)std::ceil((xyz[3 * a + 1] + ry - org[1]) / vox));
            for (long long j = j0; j <= j1; j++) {
                const double cy = org[1] + (j + 0.5) * vox;
                const double dy = cy - xyz[3 * a + 1];
                const double r2z = R * R - dx * dx - dy * dy;
                if (r2z < 0) continue;
                const double rz = std::sqrt(r2z);
                long long k0 = std::max(0LL, (long long)std::floor((xyz[3 * a + 2] - rz - org[2]) / vox));
                long long k1 = std::min(dim[2] - 1, (long long)std::ceil((xyz[3 * a + 2] + rz - org[2]) / vox));
                for (long long k = k0; k <= k1; k++) {
                    const double cz = org[2] + (k + 0.5) * vox;
                    const double dz = cz - xyz[3 * a + 2];
                    const float d = (float)(std::sqrt(dx * dx + dy * dy + dz * dz) - rad[a]);
                    float &t = dist[vid(i, j, k)];
                    if (d < t) t = d;
                }
            }
        }
    });
    // bulk solvent: flood fill from the corner through voxels outside the probe shell
    std::vector<uint8_t> bulk(nv, 0);
    std::deque<long long> queue;
    auto outside = [&](long long v) { return dist[v] > (float)probe; };
    if (outside(0)) {
        bulk[0] = 1;
        queue.push_back(0);
    }
    while (!queue.empty()) {
        const long long v = queue.front();
        queue.pop_front();
        const long long k = v % dim[2], j = (v / dim[2]) % dim[1], i = v / (dim[1] * dim[2]);
        const long long nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
        for (auto &c : nb) {
            if (c[0] < 0 || c[1] < 0 || c[2] < 0 || c[0] >= dim[0] || c[1] >= dim[1] || c[2] >= dim[2]) continue;
            const long long w = vid(c[0], c[1], c[2]);
            if (!bulk[w] && outside(w)) {
                bulk[w] = 1;
                queue.push_back(w);
            }
        }
    }
    // hydration layer: bulk voxels within the solvation thickness of the probe shell
    const float layer_max = (float)(probe + thickness);
    auto in_layer = [&](long long v) { return bulk[v] && dist[v] <= layer_max; };
    // greedy 3-D meshing: from each unassigned layer voxel (in x, y, z order)
    // grow a box along z, then y, then x while every covered voxel is an
    // unassigned layer voxel; cover those voxels and continue
    struct Box {
        long long i0, ni, j0, nj, k0, nk;
    };
    std::vector<Box> boxes;
    std::vector<uint8_t> todo(nv, 0);
    parallel_for(dim[0], [&](long long i) {
        for (long long j = 0; j < dim[1]; j++)
            for (long long k = 0; k < dim[2]; k++) {
                const long long v = vid(i, j, k);
                todo[v] = in_layer(v) ? 1 : 0;
            }
    });
    for (long long i = 0; i < dim[0]; i++)
        for (long long j = 0; j < dim[1]; j++)
            for (long long k = 0; k < dim[2]; k++) {
                if (!todo[vid(i, j, k)]) continue;
                long long nk = 1;
                while (k + nk < dim[2] && todo[vid(i, j, k + nk)]) nk++;
                long long nj = 1;
                while (j + nj < dim[1]) {
                    bool ok = true;
                    for (long long c = 0; c < nk && ok; c++) ok = todo[vid(i, j + nj, k + c)] != 0;
                    if (!ok) break;
                    nj++;
                }
                long long ni = 1;
                while (i + ni < dim[0]) {
                    bool ok = true;
                    for (long long b2 = 0; b2 < nj && ok; b2++)
                        for (long long c = 0; c < nk && ok; c++) ok = todo[vid(i + ni, j + b2, k + c)] != 0;
                    if (!ok) break;
                    ni++;
                }
                for (long long a2 = 0; a2 < ni; a2++)
                    for (long long b2 = 0; b2 < nj; b2++)
                        for (long long c = 0; c < nk; c++) todo[vid(i + a2, j + b2, k + c)] = 0;
                boxes.push_back({i, ni, j, nj, k, nk});
            }
    dims.resize(3 * boxes.size());
    com.resize(3 * boxes.size());
    for (size_t b = 0; b < boxes.size(); b++) {
        const Box &B = boxes[b];
        dims[3 * b] = (int)B.ni;
        dims[3 * b + 1] = (int)B.nj;
        dims[3 * b + 2] = (int)B.nk;
        com[3 * b] = org[0] + (B.i0 + 0.5 * B.ni) * vox;
        com[3 * b + 1] = org[1] + (B.j0 + 0.5 * B.nj) * vox;
        com[3 * b + 2] = org[2] + (B.k0 + 0.5 * B.nk) * vox;
    }
}

}  // namespace dplus

using namespace dplus;

extern "C" int dplus_hydration_boxes(int n, const double *xyz, const double *rad, double probe, double thickness,
                                     double vox, int capacity, int *nbox, int *dims, double *com) {
    try {
        if (n < 0 || !xyz || !rad || !nbox || vox <= 0 || probe < 0 || thickness < 0)
            throw Error(kErrInvalidArgs, "bad hydration arguments");
        std::vector<int> d;
        std::vector<double> c;
        hydration_boxes(n, xyz, rad, probe, thickness, vox, d, c);
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
