// ref_shim.cpp -- C entry points over the reference's own JacobianSphereGrid
// (Backend/Backend/Grid.cpp), compiled unmodified from the reference tree by
// oracle/Makefile into oracle/_ref/.  TEST INFRASTRUCTURE ONLY: the tests
// compare the oracle's grid, spline and lookup restatement with it; it is
// never linked into the product library.
#include "Grid.h"

#include <vector>

extern "C" {

void *refg_new(int grid_size, double qmax) { return new JacobianSphereGrid((unsigned short)grid_size, qmax); }

void refg_free(void *g) { delete static_cast<JacobianSphereGrid *>(g); }

double *refg_data(void *g) { return static_cast<JacobianSphereGrid *>(g)->GetDataPointer(); }

double *refg_interp(void *g) { return static_cast<JacobianSphereGrid *>(g)->GetInterpolantPointer(); }

long long refg_real_size(void *g) { return (long long)static_cast<JacobianSphereGrid *>(g)->GetRealSize(); }

void refg_splines(void *g) { static_cast<JacobianSphereGrid *>(g)->CalculateSplines(); }

void refg_get_sphr(void *g, double q, double th, double ph, double *out) {
    std::complex<double> r = static_cast<JacobianSphereGrid *>(g)->GetSphr(q, th, ph);
    out[0] = r.real();
    out[1] = r.imag();
}

void refg_get_cart(void *g, double x, double y, double z, double *out) {
    std::complex<double> r = static_cast<JacobianSphereGrid *>(g)->GetCart(x, y, z);
    out[0] = r.real();
    out[1] = r.imag();
}

double refg_intensity(void *g, double q, double eps, unsigned seed, unsigned long long iters) {
    return static_cast<JacobianSphereGrid *>(g)->CalculateIntensity(q, eps, seed, (uint64_t)iters);
}

void refg_amps_at_points(void *g, int n, const double *qs, double th, double ph, double *out) {
    std::vector<double> q(qs, qs + n);
    ArrayXcX r = static_cast<JacobianSphereGrid *>(g)->getAmplitudesAtPoints(q, th, ph);
    for (int i = 0; i < n; i++) {
        out[2 * i] = r(i).real();
        out[2 * i + 1] = r(i).imag();
    }
}

}  // extern "C"
