/*
 * oracle.h -- FP64 CPU restatement of D+'s scattering hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library may link, load or
 * call this code; it exists so that tests, the smoke check and the CPU
 * baseline leg of the benchmark can check the HIP path against an
 * independent double-precision statement of the reference algorithms.
 * Every function cites the reference file:line it follows.
 */
#ifndef DPLUS_ORACLE_H
#define DPLUS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* form-factor calculation flags (Backend/GPU/atomicFormDefines.h:5-9) */
#define ORC_CALC_ATOMIC_FORMFACTORS 1
#define ORC_CALC_DUMMY_SOLVENT 2

/* ---- grid geometry (Backend/Backend/Grid.cpp:169-186, 1078-1111) ---- */
int orc_actual_size(int grid_size);
long long orc_total_points(int grid_size);
double orc_step_size(int grid_size, double qmax);
long long orc_index_from_indices(int qi, long long ti, long long pi);
void orc_indices_from_index(long long index, int *qi, long long *ti, long long *pi);
void orc_qvector_from_index(int grid_size, double qmax, long long index, double *q3);

/* ---- counter-based RNG shared with the device code ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_mc_stratify(long long d, int chunk, double *u, double *v);
void orc_uniform_pair(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                      double *u, double *v);

/* ---- atomic form factors (Backend/GPU/Atomic Form Factor.cpp:11-27,91-176) ----
 * coef9: nion rows of (a1,b1,a2,b2,a3,b3,a4,b4,c); rad: ion radius in nm;
 * count: atoms of each ion (for the cube-mean radius r_m). */
void orc_form_factors(int nion, const double *coef9, const double *rad, const int *count,
                      double sol_ed, double c1, int flags, double q, double *out);

/* ---- direct amplitudes (Backend/Backend/PDBAmplitude.cpp:243-317) ---- */
void orc_amp_atoms(int natoms, const double *xyz, const int *ion, int nion, const double *coef9,
                   const double *rad, const int *count, double sol_ed, double c1, int flags,
                   double scale, int nq, const double *qvec, double *out);
void orc_amp_boxes(int nbox, const int *dims, const double *com, double vox_step, double contrast,
                   double scale, int nq, const double *qvec, double *out);

/* ---- grid fill (Grid.cpp:760-826 with PDBAmplitude::calcAmplitude) ----
 * data: 2*P doubles, complex interleaved, reference layout. Adds boxes if nbox>0. */
void orc_grid_fill_pdb(int grid_size, double qmax, int natoms, const double *xyz, const int *ion,
                       int nion, const double *coef9, const double *rad, const int *count,
                       double sol_ed, double c1, int flags, double scale, int nbox,
                       const int *box_dims, const double *box_com, double vox_step,
                       double box_contrast, double *data);

/* ---- spline coefficients (Grid.cpp:1113-1151, PeriodicSplineSolver.h:8-67) ---- */
void orc_periodic_spline_ref(int n, const double *pts, double *res);
void orc_periodic_spline_exact(int n, const double *pts, double *res);
void orc_grid_splines(int grid_size, const double *data, double *coeffs);

/* ---- grid lookup ---- */
/* JacobianSphereGrid::InterpolateThetaPhiPlane (Grid.cpp:61-134) */
void orc_interp_plane(int grid_size, const double *data, const double *coeffs, int ri,
                      double theta, double phi, double *out);
/* JacobianSphereGrid::GetSphr (Grid.cpp:487-548) */
void orc_get_sphr(int grid_size, double qmax, const double *data, const double *coeffs, double q,
                  double theta, double phi, double *out);
/* GetCart (Grid.cpp:384-392) */
void orc_get_cart(int grid_size, double qmax, const double *data, const double *coeffs, double qx,
                  double qy, double qz, double *out);
/* 4-plane lookup with the shell-group semantics of MCOAJacobianKernel / AddAmplitudeAtPoint
 * (Backend/GPU/OAJacob.cu:37-116, Backend/GPU/HybridOA.cu:17-159) */
void orc_amp_group(int grid_size, double qmax, const double *data, const double *coeffs, int layer,
                   double theta, double phi, int m, const double *qs, double *out);

/* ---- orientation averaging ----
 * Docking description (hybrid regime, Backend/GPU/HybridOA.cu:17-159):
 *   nrot rotations rot[4*r] = (alpha, beta, gamma [rad], scale), each with ntrans[r]
 *   translations packed in trans[3*k] (nm).  A single untransformed grid is nrot=1,
 *   rot=(0,0,0,1), ntrans=1, trans=(0,0,0). */
typedef struct {
    int grid_size;
    double qmax;
    const double *data;
    const double *coeffs;
    int nrot;
    const double *rot;
    const int *ntrans;
    const double *trans;
} orc_leaf;

/* amplitude of the docked leaves along direction (theta, phi) for the m q-values of one shell group */
void orc_docked_group(int nleaf, const orc_leaf *leaves, int layer, double theta, double phi, int m,
                      const double *qs, double *out);
/* Monte Carlo OA with Philox directions (GPUHybridCalc.cu:164-578, OAJacob.cu:132-383);
 * dirs_used = directions of the longest-running shell group */
int orc_mc_oa(int nleaf, const orc_leaf *leaves, int nq, const double *qs, uint64_t seed,
              long long max_dirs, int chunk, double eps, double *out, long long *dirs_used);
/* VEGAS OA (GPUHybridCalc.cu:581-703, HybridOA.cu:181-501) */
int orc_vegas_oa(int nleaf, const orc_leaf *leaves, int nq, const double *qs, uint64_t seed,
                 int iterations, int points_per_bin, double eps, double *out, int *iters_used);
/* amplitude at a cartesian q for docked leaves, CPU getAmplitude semantics (Amplitude.cpp:3035-3056) */
void orc_docked_cart(int nleaf, const orc_leaf *leaves, double qx, double qy, double qz, double *out);
/* azimuthal average of the 2-D generate path: out[p] = mean over nphi stratified
 * phi of |A(q[p], theta[p], phi)|^2 (phi_k = 2 pi (k + u_k) / nphi) */
void orc_azimuthal(int nleaf, const orc_leaf *leaves, int npts, const double *q, const double *theta, int nphi,
                   uint64_t seed, double *out);
/* adaptive Gauss-Kronrod OA (Amplitude.cpp:1416-1467) */
int orc_gk_oa(int nleaf, const orc_leaf *leaves, int nq, const double *qs, double eps,
              long long max_depth, double *out, long long *evals);
/* seeded single-q MC of JacobianSphereGrid::CalculateIntensity (Grid.cpp:399-484) with the
 * directions supplied by the caller (phi[i], cos(theta)[i]); returns the converged mean */
double orc_grid_intensity_dirs(int grid_size, double qmax, const double *data,
                               const double *coeffs, double q, double eps, long long n,
                               const double *phi, const double *ct, long long *used);

/* interleaved U(a0,b0), U(a1,b1) draws of std::mt19937(seed) + uniform_real_distribution */
void orc_mt19937_pairs(uint32_t seed, long long n, double a0, double b0, double a1, double b1,
                       double *out);

/* ---- manual-symmetry grid-level docking (DefaultModels/Symmetries.cpp:697-716) ---- */
void orc_dock_grid(int out_grid_size, double out_qmax, int nleaf, const orc_leaf *leaves,
                   double scale, double *out_data);

/* ---- direct FP64 Debye sum (PDBAmplitude.cpp:1789-1825 formula) ---- */
void orc_debye(int natoms, const double *xyz, const int *ion, int nion, const double *coef9,
               const double *rad, const int *count, double sol_ed, double c1, int flags,
               double scale, int nq, const double *qs, double *out);

#ifdef __cplusplus
}
#endif
#endif
