// device.hpp -- host-side objects of libdplus_mi355x shared by the C ABI,
// the orientation-average drivers and the backend: device handle, device
// buffers, grids and docked-leaf sets, plus the records the kernels read.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dplus_mi355x.h"
#include "common.hpp"

namespace dplus {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

int set_error(int code, const std::string &msg);
std::string json_escape(const std::string &s);
void check_hip(hipError_t e, const char *what);

struct Device {
    bool ok = false;
    int id = 0;
    std::string name;
    int cus = 0;
    size_t mem = 0;
    hipStream_t stream = nullptr;
    static Device &get();
    void require();

   private:
    Device();
};

// owning device allocation that only grows
struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf();
    void release();
    void ensure(size_t bytes);
    void upload(const void *src, size_t bytes, hipStream_t s);
};

// Rotation record of a docked leaf: (alpha, beta, gamma, scale), translations
// [t0, t0+nt) of the translation table, leaf = index into the grid table.
struct RotDev {
    double4 ang;
    int leaf, t0, nt, pad;
};

// one shell group: q-values q0..q0+m-1 lie in (layer*step, (layer+1)*step]
struct GroupDev {
    int layer, q0, m, pad;
};

struct LeafGeo {
    int N;
    double step;
};

struct Grid {
    int grid_size = 0, N = 0, precision = DPLUS_PRECISION_MIXED;
    double qmax = 0, step = 0;
    long long P = 0;
    DevBuf dev;  // P nodes of (Re, Im, dRe, dIm) in float4 or double4
    std::vector<double> host_data, host_interp;
    Grid(int gs, double qmax, int precision);
    bool fp64() const { return precision == DPLUS_PRECISION_FP64; }
    size_t node_bytes() const { return fp64() ? sizeof(double4) : sizeof(float4); }
    void ensure_host();
    void splines(hipStream_t s);
    void download(hipStream_t s);
    void upload(hipStream_t s);
};

// leaves of one orientation average, flattened for the kernels
struct DockSet {
    int grid_size = 0, nrot = 0;
    double qmax = 0, step = 0;
    bool fp64 = false;
    DevBuf grids, geos, rotb, trb;
    DockSet(int nleaf, const dplus_leaf *leaves, hipStream_t s);
};

std::vector<GroupDev> make_groups(const double *q, int nq, double step);
void pdb_build(Grid &g, const dplus_pdb_desc *d, bool host_copy, hipStream_t s);
void orientation_average(const DockSet &D, int nq, const double *q, const dplus_oa_params *p, double *out,
                         dplus_oa_stats *st, hipStream_t s);
void gk_average(const DockSet &D, int nq, const double *q, double eps, long long max_depth, double *out,
                dplus_oa_stats *st, hipStream_t s);
void dock_grid(const DockSet &D, Grid &out, double scale, hipStream_t s);
void synth_walk(unsigned long long seed, int n, double step, double min_sep, double ax, double ay, double az,
                double *xyz);

}  // namespace dplus
