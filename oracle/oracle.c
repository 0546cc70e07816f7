/*
 * oracle.c -- FP64 CPU restatement of D+'s scattering hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the FP64 reference the HIP implementation is
 * tested against.  Only the tests, the smoke check and the benchmark's
 * CPU-baseline leg may load it.
 * Every function restates the reference algorithm named in its comment
 * (paths relative to the reference tree) in plain double precision.
 *
 * Pinning (what checks each function independently of this file):
 *   geometry            closed forms 6i(i+1)^2, round trips, reference test_grid.py:19-32
 *   philox              Random123 known-answer vectors
 *   form factors        single-atom golden fixtures (I(q) = f~(q)^2), reviewer tests
 *   amp_atoms / fill    brute-force definition at q=0 (sum f~), Debye identity of |A|^2
 *   splines / lookup    the reference's own Grid.cpp compiled in oracle/_ref; exact
 *                       reproduction of smooth periodic functions; slab.ampj values
 *   mc / vegas / gk     isotropic amplitude (exact), Debye sum for small systems,
 *                       mutual agreement of the three quadratures
 *   dock grid / docked  point-atom subunits docked = Debye sum of the expanded list
 *   boxes               box sums equal the sum over unit voxels (exact decomposition)
 *   azimuthal           two atoms on the x axis: 2 f^2 (1 + J0(2 q d sin theta)) (Bessel closed
 *                       form of the phi average); isotropic amplitude (exact)
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#define M_2PI_D 6.28318530717958647692528676656

/* ------------------------------------------------------------------------- */
/* Grid geometry: Grid.cpp:169-186 (InitializeGrid), 1078-1111 (indices)      */
/* ------------------------------------------------------------------------- */
int orc_actual_size(int grid_size) { return grid_size / 2 + 3; }

long long orc_total_points(int grid_size) {
    long long n = orc_actual_size(grid_size);
    return 1 + 6 * n * (n + 1) * (n + 1);
}

double orc_step_size(int grid_size, double qmax) { return qmax / (double)(grid_size / 2); }

/* number of points before shell qi (the origin is index 0) */
static long long shell_base(long long qi) { return 1 + 6 * (qi - 1) * qi * qi; }

long long orc_index_from_indices(int qi, long long ti, long long pi) {
    if (qi == 0) return 0;
    return shell_base(qi) + ti * 6LL * qi + pi;
}

void orc_indices_from_index(long long index, int *qi, long long *ti, long long *pi) {
    if (index == 0) {
        *qi = 0;
        *ti = 0;
        *pi = 0;
        return;
    }
    long long q = (long long)cbrt((double)index / 6.0);
    if (q < 1) q = 1;
    while (shell_base(q) > index) q--;
    while (shell_base(q + 1) <= index) q++;
    long long rem = index - shell_base(q);
    *qi = (int)q;
    *ti = rem / (6 * q);
    *pi = rem % (6 * q);
}

/* Grid.cpp:795-823: q-vector of a grid point */
void orc_qvector_from_index(int grid_size, double qmax, long long index, double *q3) {
    int qi;
    long long ti, pi;
    orc_indices_from_index(index, &qi, &ti, &pi);
    if (qi == 0) {
        q3[0] = q3[1] = q3[2] = 0.0;
        return;
    }
    double step = orc_step_size(grid_size, qmax);
    double qI = (double)qi * step;
    double tI = M_PI * (double)ti / (double)(3 * qi);
    double pI = M_2PI_D * (double)pi / (double)(6 * qi);
    q3[0] = qI * sin(tI) * cos(pI);
    q3[1] = qI * sin(tI) * sin(pI);
    q3[2] = qI * cos(tI);
}

/* ------------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon et al., SC'11), counter based                        */
/* ------------------------------------------------------------------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* Stratified MC directions of a chunk of C directions (C a multiple of 256):
 * direction d of the chunk belongs to run w = d / 64 of W = C / 64 runs; run w
 * covers patch (w / nv, w % nv) of the nu x nv grid of equal-area patches in
 * (u, v) = ((1 - cos theta) / 2, phi / 2 pi), with nu the largest divisor of W
 * with 2 nu^2 <= W and nv = W / nu.  (The HIP path's sampling design: the
 * reference draws unstratified directions from a clock-seeded MTGP32 stream,
 * GPUHybridCalc.cu:196-198, so only the estimator, not the stream, is shared.) */
void orc_mc_stratify(long long d, int chunk, double *u, double *v) {
    int W = chunk / 64, nu = 1, a;
    if (W < 1 || chunk % 64) return;
    for (a = 1; 2 * a * a <= W; a++)
        if (W % a == 0) nu = a;
    int nv = W / nu;
    long long w = (d % chunk) / 64;
    *u = ((double)(w / nv) + *u) / (double)nu;
    *v = ((double)(w % nv) + *v) / (double)nv;
}

/* two doubles in (0,1) with 53 random bits each */
void orc_uniform_pair(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, double *u,
                      double *v) {
    uint32_t ctr[4] = {c0, c1, c2, c3};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t x[4];
    orc_philox4x32_10(ctr, key, x);
    uint64_t a = ((uint64_t)(x[0] >> 6) << 27) | (uint64_t)(x[1] >> 5);
    uint64_t b = ((uint64_t)(x[2] >> 6) << 27) | (uint64_t)(x[3] >> 5);
    *u = ((double)a + 0.5) * (1.0 / 9007199254740992.0);
    *v = ((double)b + 0.5) * (1.0 / 9007199254740992.0);
}

/* ------------------------------------------------------------------------- */
/* Form factors: Atomic Form Factor.cpp:11-27 (r_m), 91-115 (solvent),        */
/* 160-176 (Cromer-Mann 4 Gaussians + c, s^2 = q^2/(100*16 pi^2))             */
/* ------------------------------------------------------------------------- */
void orc_form_factors(int nion, const double *coef9, const double *rad, const int *count,
                      double sol_ed, double c1, int flags, double q, double *out) {
    double rm = 0.0;
    if (flags & ORC_CALC_DUMMY_SOLVENT) {
        double num = 0.0, den = 0.0;
        for (int i = 0; i < nion; i++) {
            num += rad[i] * rad[i] * rad[i] * (double)count[i];
            den += (double)count[i];
        }
        rm = cbrt(num / den);
    }
    const double sqq = q * q / (100.0 * 157.913670417429737901351855998);
    for (int i = 0; i < nion; i++) {
        const double *c = coef9 + 9 * i;
        double f = 0.0;
        if (flags & ORC_CALC_ATOMIC_FORMFACTORS) {
            f = c[8];
            for (int k = 0; k < 4; k++) f += c[2 * k] * exp(-sqq * c[2 * k + 1]);
        }
        double s = 0.0;
        if (flags & ORC_CALC_DUMMY_SOLVENT) {
            s = 4.1887902047863909846 * rad[i] * rad[i] * rad[i] *
                exp(-(rm * rm) * (q * q) * 0.2067834969664667) * sol_ed;
            s *= c1 * c1 * c1 * exp(-0.6822178052976590 * q * q * rm * rm * (c1 * c1 - 1.0));
        }
        out[i] = f - s;
    }
}

/* ------------------------------------------------------------------------- */
/* Direct amplitudes: PDBAmplitude.cpp:243-317                               */
/* ------------------------------------------------------------------------- */
void orc_amp_atoms(int natoms, const double *xyz, const int *ion, int nion, const double *coef9,
                   const double *rad, const int *count, double sol_ed, double c1, int flags,
                   double scale, int nq, const double *qvec, double *out) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nq; k++) {
        double qx = qvec[3 * k], qy = qvec[3 * k + 1], qz = qvec[3 * k + 2];
        double q = sqrt(qx * qx + qy * qy + qz * qz);
        double ff[256];
        orc_form_factors(nion, coef9, rad, count, sol_ed, c1, flags, q, ff);
        double re = 0.0, im = 0.0;
        for (int a = 0; a < natoms; a++) {
            double ph = qx * xyz[3 * a] + qy * xyz[3 * a + 1] + qz * xyz[3 * a + 2];
            re += ff[ion[a]] * cos(ph);
            im += ff[ion[a]] * sin(ph);
        }
        out[2 * k] = re * scale;
        out[2 * k + 1] = im * scale;
    }
}

/* one solvent box: (Delta rho) prod_a sinc(q_a n_a D/2) n_a D e^{i q.c}
 * (PDBAmplitude.cpp:277-314; PDBAmplitudeCUDA.cu:15-91) */
static void box_term(const int *d, const double *c, double vox, double qx, double qy, double qz,
                     double *re, double *im) {
    double q3[3] = {qx, qy, qz};
    double va = 1.0;
    for (int a = 0; a < 3; a++) {
        double len = (double)d[a] * vox;
        double x = q3[a] * len * 0.5;
        if (q3[a] == 0.0 || d[a] == 0)
            va *= len;
        else
            va *= sin(x) / x * len;
    }
    double ph = qx * c[0] + qy * c[1] + qz * c[2];
    *re = va * cos(ph);
    *im = va * sin(ph);
}

void orc_amp_boxes(int nbox, const int *dims, const double *com, double vox_step, double contrast,
                   double scale, int nq, const double *qvec, double *out) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nq; k++) {
        double sr = 0.0, si = 0.0;
        for (int b = 0; b < nbox; b++) {
            double r, i;
            box_term(dims + 3 * b, com + 3 * b, vox_step, qvec[3 * k], qvec[3 * k + 1],
                     qvec[3 * k + 2], &r, &i);
            sr += r;
            si += i;
        }
        out[2 * k] += sr * contrast * scale;
        out[2 * k + 1] += si * contrast * scale;
    }
}

/* Grid.cpp:760-826 driving PDBAmplitude::calcAmplitude; the form factor is
 * taken at the shell radius qi*step (PDBAmplitudeCUDA.cu:133-186 indexes the
 * per-shell form-factor matrix the same way). */
void orc_grid_fill_pdb(int grid_size, double qmax, int natoms, const double *xyz, const int *ion,
                       int nion, const double *coef9, const double *rad, const int *count,
                       double sol_ed, double c1, int flags, double scale, int nbox,
                       const int *box_dims, const double *box_com, double vox_step,
                       double box_contrast, double *data) {
    const int N = orc_actual_size(grid_size);
    const long long P = orc_total_points(grid_size);
    const double step = orc_step_size(grid_size, qmax);
    double *fft = (double *)malloc(sizeof(double) * (size_t)(N + 1) * (size_t)nion);
    for (int s = 0; s <= N; s++)
        orc_form_factors(nion, coef9, rad, count, sol_ed, c1, flags, s * step, fft + (size_t)s * nion);
#pragma omp parallel for schedule(dynamic, 256)
    for (long long idx = 0; idx < P; idx++) {
        int qi;
        long long ti, pi;
        orc_indices_from_index(idx, &qi, &ti, &pi);
        double q3[3];
        orc_qvector_from_index(grid_size, qmax, idx, q3);
        const double *ff = fft + (size_t)qi * nion;
        double re = 0.0, im = 0.0;
        if (flags & (ORC_CALC_ATOMIC_FORMFACTORS | ORC_CALC_DUMMY_SOLVENT)) {
            for (int a = 0; a < natoms; a++) {
                double ph = q3[0] * xyz[3 * a] + q3[1] * xyz[3 * a + 1] + q3[2] * xyz[3 * a + 2];
                re += ff[ion[a]] * cos(ph);
                im += ff[ion[a]] * sin(ph);
            }
        }
        if (nbox > 0 && box_contrast != 0.0) {
            double br = 0.0, bi = 0.0;
            for (int b = 0; b < nbox; b++) {
                double r, i;
                box_term(box_dims + 3 * b, box_com + 3 * b, vox_step, q3[0], q3[1], q3[2], &r, &i);
                br += r;
                bi += i;
            }
            re += br * box_contrast;
            im += bi * box_contrast;
        }
        data[2 * idx] = re * scale;
        data[2 * idx + 1] = im * scale;
    }
    free(fft);
}

/* ------------------------------------------------------------------------- */
/* Periodic cubic spline: PeriodicSplineSolver.h:8-67.  The reference keeps   */
/* the matrix (diag, rightCol) in single precision and the right-hand side in */
/* double; the elimination below reproduces that arithmetic step by step.     */
/* ------------------------------------------------------------------------- */
void orc_periodic_spline_ref(int n, const double *pts, double *res) {
    float *diag = (float *)malloc(sizeof(float) * n);
    float *rc = (float *)malloc(sizeof(float) * n);
    double *ar = (double *)malloc(sizeof(double) * n);
    double *ai = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < n; i++) {
        diag[i] = 4.0f;
        rc[i] = 0.0f;
    }
    for (int i = 0; i < n; i++) {
        int ip = (i + 1) % n, im = (n + i - 1) % n;
        ar[i] = 3.0 * (pts[2 * ip] - pts[2 * im]);
        ai[i] = 3.0 * (pts[2 * ip + 1] - pts[2 * im + 1]);
    }
    rc[0] = 1.0f;
    rc[n - 2] = 1.0f;
    double vp = 0.25, vc;
    for (int i = 1; i < n - 1; i++) {
        vc = 1.0 / ((double)diag[i] - vp);
        diag[i] = (float)((double)diag[i] - vp);
        rc[i] = (float)((double)rc[i] - vp * (double)rc[i - 1]);
        ar[i] -= ar[i - 1] * vp;
        ai[i] -= ai[i - 1] * vp;
        vp = vc;
    }
    diag[n - 1] = (float)((double)diag[n - 1] - vp * (double)rc[n - 2]);
    ar[n - 1] -= vp * ar[n - 2];
    ai[n - 1] -= vp * ai[n - 2];
    float bottom_left = 1.0f;
    for (int i = 0; i < n - 1; i++) {
        float vcf = bottom_left / diag[i];
        vc = (double)vcf;
        diag[n - 1] = (float)((double)diag[n - 1] - vc * (double)rc[i]);
        ar[n - 1] -= vc * ar[i];
        ai[n - 1] -= vc * ai[i];
        bottom_left = (float)(-vc);
    }
    double fac = (double)(rc[n - 2] / diag[n - 1]);
    ar[n - 2] -= ar[n - 1] * fac;
    ai[n - 2] -= ai[n - 1] * fac;
    for (int i = n - 3; i >= 0; i--) {
        double r = (double)(rc[i] / diag[n - 1]);
        ar[i] -= ar[n - 1] * r;
        ai[i] -= ai[n - 1] * r;
        double u = 1.0 / (double)diag[i + 1];
        ar[i] -= ar[i + 1] * u;
        ai[i] -= ai[i + 1] * u;
    }
    for (int i = 0; i < n; i++) {
        res[2 * i] = ar[i] / (double)diag[i];
        res[2 * i + 1] = ai[i] / (double)diag[i];
    }
    free(diag);
    free(rc);
    free(ar);
    free(ai);
}

/* The same cyclic system (4 on the diagonal, 1 off it and in the corners)
 * solved exactly in double by the Sherman-Morrison reduction of Bartels 1987
 * sec. 3.4; used only to bound the reference's single-precision matrix. */
void orc_periodic_spline_exact(int n, const double *pts, double *res) {
    /* A = T + u v^T with T tridiagonal(1,4,1) modified at the corners */
    double *cp = (double *)malloc(sizeof(double) * n);
    double *b = (double *)malloc(sizeof(double) * n);
    double *y = (double *)malloc(sizeof(double) * 2 * n);
    double *z = (double *)malloc(sizeof(double) * n);
    const double gamma = -4.0;
    for (int i = 0; i < n; i++) b[i] = 4.0;
    b[0] = 4.0 - gamma;
    b[n - 1] = 4.0 - 1.0 / gamma;
    /* solve T x = rhs (Thomas); T has off diagonals 1 */
    for (int part = 0; part < 3; part++) {
        double *rhs = (double *)malloc(sizeof(double) * n);
        for (int i = 0; i < n; i++) {
            if (part < 2) {
                int ip = (i + 1) % n, im = (n + i - 1) % n;
                rhs[i] = 3.0 * (pts[2 * ip + part] - pts[2 * im + part]);
            } else {
                rhs[i] = (i == 0) ? gamma : ((i == n - 1) ? 1.0 : 0.0);
            }
        }
        cp[0] = 1.0 / b[0];
        rhs[0] = rhs[0] / b[0];
        for (int i = 1; i < n; i++) {
            double m = b[i] - cp[i - 1];
            cp[i] = 1.0 / m;
            rhs[i] = (rhs[i] - rhs[i - 1]) / m;
        }
        for (int i = n - 2; i >= 0; i--) rhs[i] -= cp[i] * rhs[i + 1];
        if (part < 2)
            for (int i = 0; i < n; i++) y[2 * i + part] = rhs[i];
        else
            for (int i = 0; i < n; i++) z[i] = rhs[i];
        free(rhs);
    }
    double vz = z[0] + z[n - 1] / gamma;
    for (int part = 0; part < 2; part++) {
        double vy = y[part] + y[2 * (n - 1) + part] / gamma;
        double f = vy / (1.0 + vz);
        for (int i = 0; i < n; i++) res[2 * i + part] = y[2 * i + part] - f * z[i];
    }
    free(cp);
    free(b);
    free(y);
    free(z);
}

/* Grid.cpp:1113-1151: one periodic spline per theta-row of every shell */
void orc_grid_splines(int grid_size, const double *data, double *coeffs) {
    const int N = orc_actual_size(grid_size);
    coeffs[0] = coeffs[1] = 0.0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 1; i <= N; i++) {
        for (long long j = 0; j < 3LL * i + 1; j++) {
            long long pos = 2 * orc_index_from_indices(i, j, 0);
            orc_periodic_spline_ref(6 * i, data + pos, coeffs + pos);
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Lookup                                                                    */
/* ------------------------------------------------------------------------- */
static void hermite(const double *p1, const double *p2, const double *d1, const double *d2,
                    double t, double *out) {
    for (int c = 0; c < 2; c++) {
        out[c] = p1[c] + d1[c] * t + (3.0 * (p2[c] - p1[c]) - 2.0 * d1[c] - d2[c]) * (t * t) +
                 (2.0 * (p1[c] - p2[c]) + d1[c] + d2[c]) * (t * t * t);
    }
}

/* PeriodicSplineSolver.h:227-234 (EvenlySpacedFourPointSpline) */
static void four_point(const double *y1, const double *y2, const double *y3, const double *y4,
                       double *d2, double *d3) {
    for (int c = 0; c < 2; c++) {
        d2[c] = -(7. / 15.) * (y2[c] - y1[c]) + (14. / 15.) * (y3[c] - y1[c]) -
                (4. / 15.) * (y4[c] - y2[c]) + (2. / 15.) * (y4[c] - y3[c]);
        d3[c] = (2. / 15.) * (y2[c] - y1[c]) - (4. / 15.) * (y3[c] - y1[c]) +
                (14. / 15.) * (y4[c] - y2[c]) - (7. / 15.) * (y4[c] - y3[c]);
    }
}

/* Grid.cpp:61-134 with IndicesFromRadians (Grid.cpp:1220-1256), including its
 * handling of theta rows beyond [0, pi] (row index wrapped, phi moved by pi). */
void orc_interp_plane(int grid_size, const double *data, const double *coeffs, int ri,
                      double theta, double phi, double *out) {
    (void)grid_size;
    long long qi = ri - 1;
    long long base = 1 + ((qi + 1) * (6 * qi) * (6 * qi + 3 + 3)) / 6;
    long long phiPoints = 6LL * ri, thePoints = 3LL * ri;
    double edge = M_2PI_D / (double)phiPoints;
    long long tI = (long long)((theta / M_PI) * (double)thePoints);
    long long pI = (long long)((phi / M_2PI_D) * (double)phiPoints);
    double tTh = (theta / edge) - (double)tI;
    double tPh = (phi / edge) - (double)pI;
    if (fabs(tTh) < 1.0e-10) tTh = 0.0;
    if (fabs(tPh) < 1.0e-10) tPh = 0.0;
    if (pI == phiPoints) pI = 0;
    long long pI1 = (pI + 1) % phiPoints;
    double nodes[4][2];
    for (int i = -1; i < 3; i++) {
        long long theInd = (thePoints + 1 + tI + i) % (thePoints + 1);
        long long pos1, pos2;
        if (theInd == tI + i) {
            pos1 = 2 * (base + phiPoints * theInd + pI);
            pos2 = 2 * (base + phiPoints * theInd + pI1);
        } else {
            if (tI + i > thePoints) theInd = (tI + i) - (thePoints + 1);
            pos1 = 2 * (base + phiPoints * theInd + pI);
            pos2 = 2 * (base + phiPoints * theInd + pI1);
            pos1 += ((2 * pI >= phiPoints) ? -phiPoints : phiPoints);
            pos2 += ((2 * pI1 >= phiPoints) ? -phiPoints : phiPoints);
        }
        hermite(data + pos1, data + pos2, coeffs + pos1, coeffs + pos2, tPh, nodes[i + 1]);
    }
    double d1[2], d2[2];
    four_point(nodes[0], nodes[1], nodes[2], nodes[3], d1, d2);
    hermite(nodes[1], nodes[2], d1, d2, tTh, out);
}

/* the reference nudges q/step up by 8 ulps before truncating (Grid.cpp:500-504) */
static double bump8(double x) {
    long long b;
    memcpy(&b, &x, sizeof b);
    b += 8;
    memcpy(&x, &b, sizeof b);
    return x;
}

static void antipode(double th, double ph, double *nth, double *nph) {
    double a = M_PI - th;
    double b = ph + M_PI;
    if (b < 0) b = M_2PI_D + (b - (int)(b / M_2PI_D) * M_2PI_D);
    if (b >= M_2PI_D) b = b - (int)(b / M_2PI_D) * M_2PI_D;
    if (a < 0) a = 0.;
    if (a > M_PI) a = M_PI;
    *nth = a;
    *nph = b;
}

/* Grid.cpp:487-548 */
void orc_get_sphr(int grid_size, double qmax, const double *data, const double *coeffs, double rr,
                  double th, double ph, double *out) {
    const int N = orc_actual_size(grid_size);
    const double step = orc_step_size(grid_size, qmax);
    if (rr == 0.0) {
        out[0] = data[0];
        out[1] = data[1];
        return;
    }
    double frc = bump8(rr / step);
    int fInd = (int)frc;
    double frac = frc - (double)fInd;
    double p1[2], p2[2], p3[2], p4[2];
    if (fInd > 0)
        orc_interp_plane(grid_size, data, coeffs, fInd, th, ph, p2);
    else {
        p2[0] = data[0];
        p2[1] = data[1];
    }
    if (frac < 1.e-9) {
        out[0] = p2[0];
        out[1] = p2[1];
        return;
    }
    orc_interp_plane(grid_size, data, coeffs, fInd + 1, th, ph, p3);
    if (fInd == 0) {
        double nth, nph;
        antipode(th, ph, &nth, &nph);
        orc_interp_plane(grid_size, data, coeffs, fInd + 1, nth, nph, p1);
    } else if (fInd == 1) {
        p1[0] = data[0];
        p1[1] = data[1];
    } else {
        orc_interp_plane(grid_size, data, coeffs, fInd - 1, th, ph, p1);
    }
    if (fInd + 2 < N)
        orc_interp_plane(grid_size, data, coeffs, fInd + 2, th, ph, p4);
    else {
        p4[0] = p3[0];
        p4[1] = p3[1];
    }
    double d1[2], d2[2];
    four_point(p1, p2, p3, p4, d1, d2);
    hermite(p2, p3, d1, d2, frac, out);
}

/* Grid.cpp:384-392 */
void orc_get_cart(int grid_size, double qmax, const double *data, const double *coeffs, double x,
                  double y, double z, double *out) {
    double r = sqrt(x * x + y * y + z * z);
    if (fabs(z) > r) r = fabs(z);
    if (r == 0.0) {
        out[0] = data[0];
        out[1] = data[1];
        return;
    }
    double th = acos(z / r);
    double ph = atan2(y, x);
    if (ph < 0.0) ph += M_2PI_D;
    orc_get_sphr(grid_size, qmax, data, coeffs, r, th, ph, out);
}

/* OAJacob.cu:57-113 / HybridOA.cu:78-133: planes layer-1..layer+2 (plane -1 is
 * shell 1 seen from the antipode, plane 0 the origin), q-Hermite with
 * t = (q - layer*step)/step. */
void orc_amp_group(int grid_size, double qmax, const double *data, const double *coeffs, int layer,
                   double theta, double phi, int m, const double *qs, double *out) {
    const double step = orc_step_size(grid_size, qmax);
    double amps[4][2];
    for (int i = -1; i <= 2; i++) {
        int lqi = i + layer;
        if (lqi == -1) {
            double nth, nph;
            antipode(theta, phi, &nth, &nph);
            orc_interp_plane(grid_size, data, coeffs, 1, nth, nph, amps[i + 1]);
        } else if (lqi == 0) {
            amps[i + 1][0] = data[0];
            amps[i + 1][1] = data[1];
        } else {
            orc_interp_plane(grid_size, data, coeffs, lqi, theta, phi, amps[i + 1]);
        }
    }
    double d1[2], d2[2];
    four_point(amps[0], amps[1], amps[2], amps[3], d1, d2);
    for (int j = 0; j < m; j++) {
        double t = (qs[j] - layer * step) / step;
        hermite(amps[1], amps[2], d1, d2, t, out + 2 * j);
    }
}

/* ------------------------------------------------------------------------- */
/* Docking                                                                   */
/* ------------------------------------------------------------------------- */
/* R^T q with R = Rx(a) Ry(b) Rz(c) (Amplitude.h:557-573; HybridOA.cu:60-63) */
static void rot_t(const double *r, const double *q, double *o) {
    double s1 = sin(r[0]), c1 = cos(r[0]);
    double s2 = sin(r[1]), c2 = cos(r[1]);
    double s3 = sin(r[2]), c3 = cos(r[2]);
    o[0] = (c2 * c3) * q[0] + (c1 * s3 + c3 * s1 * s2) * q[1] + (s1 * s3 - c1 * c3 * s2) * q[2];
    o[1] = -(c2 * s3) * q[0] + (c1 * c3 - s1 * s2 * s3) * q[1] + (c3 * s1 + c1 * s2 * s3) * q[2];
    o[2] = (s2)*q[0] - (c2 * s1) * q[1] + (c1 * c2) * q[2];
}

/* HybridOA.cu:17-159 (AddAmplitudeAtPoint) */
void orc_docked_group(int nleaf, const orc_leaf *leaves, int layer, double theta, double phi, int m,
                      const double *qs, double *out) {
    double qv[3] = {cos(phi) * sin(theta), sin(phi) * sin(theta), cos(theta)};
    for (int j = 0; j < 2 * m; j++) out[j] = 0.0;
    double *tmp = (double *)malloc(sizeof(double) * 2 * m);
    for (int l = 0; l < nleaf; l++) {
        const orc_leaf *L = leaves + l;
        int cum = 0;
        for (int r = 0; r < L->nrot; r++) {
            const double *rt = L->rot + 4 * r;
            double rv[3];
            rot_t(rt, qv, rv);
            if (fabs(rv[2]) > 1.) rv[2] = (rv[2] > 0.) ? 1. : -1.;
            double nth = acos(rv[2]);
            double nph = atan2(rv[1], rv[0]);
            if (nph < 0.0) nph += M_2PI_D;
            orc_amp_group(L->grid_size, L->qmax, L->data, L->coeffs, layer, nth, nph, m, qs, tmp);
            for (int j = 0; j < m; j++) {
                double sc = 0.0, ss = 0.0;
                for (int t = 0; t < L->ntrans[r]; t++) {
                    const double *tr = L->trans + 3 * (cum + t);
                    double qd = qs[j] * (tr[0] * qv[0] + tr[1] * qv[1] + tr[2] * qv[2]);
                    sc += cos(qd);
                    ss += sin(qd);
                }
                double ar = tmp[2 * j], ai = tmp[2 * j + 1];
                out[2 * j] += (ar * sc - ai * ss) * rt[3];
                out[2 * j + 1] += (ai * sc + ar * ss) * rt[3];
            }
            cum += L->ntrans[r];
        }
    }
    free(tmp);
}

/* Amplitude.cpp:3035-3056 (getAmplitude) composed with Symmetries.cpp:697-716 */
void orc_docked_cart(int nleaf, const orc_leaf *leaves, double qx, double qy, double qz,
                     double *out) {
    double q[3] = {qx, qy, qz};
    out[0] = out[1] = 0.0;
    for (int l = 0; l < nleaf; l++) {
        const orc_leaf *L = leaves + l;
        int cum = 0;
        for (int r = 0; r < L->nrot; r++) {
            const double *rt = L->rot + 4 * r;
            double rv[3], a[2];
            rot_t(rt, q, rv);
            orc_get_cart(L->grid_size, L->qmax, L->data, L->coeffs, rv[0], rv[1], rv[2], a);
            double sc = 0.0, ss = 0.0;
            for (int t = 0; t < L->ntrans[r]; t++) {
                const double *tr = L->trans + 3 * (cum + t);
                double qd = q[0] * tr[0] + q[1] * tr[1] + q[2] * tr[2];
                sc += cos(qd);
                ss += sin(qd);
            }
            out[0] += (a[0] * sc - a[1] * ss) * rt[3];
            out[1] += (a[1] * sc + a[0] * ss) * rt[3];
            cum += L->ntrans[r];
        }
    }
}

/* The 2-D generate path's average over phi at fixed (q, theta)
 * (DomainModel::CalculateIntensity with thetaIn >= 0, Amplitude.cpp:1474-1592, and
 * AverageIntensitiesBetweenLayers2D :2825-2897), with one Philox-drawn sample per
 * phi stratum keyed by (seed, sample, point). */
#define ORC_STREAM_AZIMUTH 0x415Au
void orc_azimuthal(int nleaf, const orc_leaf *leaves, int npts, const double *q, const double *theta, int nphi,
                   uint64_t seed, double *out) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int p = 0; p < npts; p++) {
        double sum = 0.0;
        const double st = sin(theta[p]), ct = cos(theta[p]);
        for (int k = 0; k < nphi; k++) {
            double u, v, a[2];
            orc_uniform_pair(seed, (uint32_t)k, (uint32_t)p, 0u, ORC_STREAM_AZIMUTH, &u, &v);
            const double ph = M_2PI_D * ((double)k + u) / (double)nphi;
            orc_docked_cart(nleaf, leaves, q[p] * st * cos(ph), q[p] * st * sin(ph), q[p] * ct, a);
            sum += a[0] * a[0] + a[1] * a[1];
        }
        out[p] = sum / (double)nphi;
    }
}

/* Symmetries.cpp:697-716 evaluated at every node of a new grid (Grid.cpp:760-826) */
void orc_dock_grid(int out_grid_size, double out_qmax, int nleaf, const orc_leaf *leaves,
                   double scale, double *out_data) {
    const long long P = orc_total_points(out_grid_size);
#pragma omp parallel for schedule(dynamic, 256)
    for (long long idx = 0; idx < P; idx++) {
        double q3[3], a[2];
        orc_qvector_from_index(out_grid_size, out_qmax, idx, q3);
        orc_docked_cart(nleaf, leaves, q3[0], q3[1], q3[2], a);
        out_data[2 * idx] = a[0] * scale;
        out_data[2 * idx + 1] = a[1] * scale;
    }
}

/* ------------------------------------------------------------------------- */
/* Orientation averages                                                      */
/* ------------------------------------------------------------------------- */
#define ORC_STREAM_MC 0x4D43u
#define ORC_STREAM_VEGAS 0x5645u

/* shell groups (OAJacob.cu:234-249; GPUHybridCalc.cu:410-420): consecutive q
 * with layer*step < q <= (layer+1)*step share one set of directions */
static int next_group(const double *qs, int nq, int q0, double step, int *layer, int *q1) {
    if (q0 >= nq) return 0;
    while (qs[q0] > (double)(*layer + 1) * step) (*layer)++;
    int e = q0 + 1;
    while (e < nq && !(qs[e] > (double)(*layer + 1) * step)) e++;
    *q1 = e; /* exclusive */
    return 1;
}

static double leaves_step(const orc_leaf *L) { return orc_step_size(L->grid_size, L->qmax); }

/* checkConvergenceOA (GPUHybridCalc.cu:705-720) over the history of running means */
static int converged3(const double *hist, int k, int len, double eps) {
    if (!(k > 2 && eps > 0.0)) return 0;
    for (int i = 0; i < len; i++) {
        double c = hist[(size_t)k * len + i];
        for (int b = 1; b <= 3; b++)
            if (fabs(1.0 - c / hist[(size_t)(k - b) * len + i]) > eps) return 0;
    }
    return 1;
}

int orc_mc_oa(int nleaf, const orc_leaf *leaves, int nq, const double *qs, uint64_t seed,
              long long max_dirs, int chunk, double eps, double *out, long long *dirs_used) {
    const double step = leaves_step(leaves);
    long long nchunks = (max_dirs + chunk - 1) / chunk;
    if (nchunks < 1) nchunks = 1;
    int layer = 0, q0 = 0, q1 = 0;
    long long used_total = 0;
    while (next_group(qs, nq, q0, step, &layer, &q1)) {
        int m = q1 - q0;
        double *hist = (double *)calloc((size_t)nchunks * m, sizeof(double));
        double *sum = (double *)calloc(m, sizeof(double));
        int k;
        for (k = 0; k < nchunks; k++) {
            double *csum = (double *)calloc(m, sizeof(double));
#pragma omp parallel
            {
                double *amp = (double *)malloc(sizeof(double) * 2 * m);
                double *loc = (double *)calloc(m, sizeof(double));
#pragma omp for schedule(static)
                for (long long d = 0; d < chunk; d++) {
                    long long gd = (long long)k * chunk + d;
                    double u, v;
                    orc_uniform_pair(seed, (uint32_t)gd, (uint32_t)(gd >> 32), (uint32_t)layer,
                                     ORC_STREAM_MC, &u, &v);
                    orc_mc_stratify(d, chunk, &u, &v);
                    double th = acos(2.0 * u - 1.0), ph = M_2PI_D * v;
                    orc_docked_group(nleaf, leaves, layer, th, ph, m, qs + q0, amp);
                    for (int j = 0; j < m; j++)
                        loc[j] += amp[2 * j] * amp[2 * j] + amp[2 * j + 1] * amp[2 * j + 1];
                }
#pragma omp critical
                for (int j = 0; j < m; j++) csum[j] += loc[j];
                free(amp);
                free(loc);
            }
            for (int j = 0; j < m; j++) {
                sum[j] += csum[j];
                hist[(size_t)k * m + j] = sum[j] / ((double)(k + 1) * chunk);
            }
            free(csum);
            if ((long long)(k + 1) * chunk > used_total) used_total = (long long)(k + 1) * chunk;
            if (converged3(hist, k, m, eps)) {
                k++;
                break;
            }
        }
        for (int j = 0; j < m; j++) out[q0 + j] = hist[(size_t)(k - 1) * m + j];
        free(hist);
        free(sum);
        q0 = q1;
        layer++;
    }
    if (dirs_used) *dirs_used = used_total;
    return 0;
}

/* VEGAS bin refinement for one axis: ResizeVegasPoints (HybridOA.cu:181-263).
 * New boundaries that the serial re-binning does not reach keep their old value. */
static void vegas_resize(int dim, const double *var, double *bounds) {
    double total = 0.0;
    for (int i = 0; i < dim; i++) total += var[i];
    if (total == 0.) return;
    double *w = (double *)malloc(sizeof(double) * dim);
    double *nb = (double *)malloc(sizeof(double) * dim);
    for (int i = 0; i < dim; i++) {
        double b;
        if (i == 0)
            b = (var[0] + var[1]) / 2.;
        else if (i == dim - 1)
            b = (var[i] + var[i - 1]) / 2.;
        else
            b = (var[i - 1] + var[i] + var[i + 1]) / 3.;
        b = total / b;
        b = pow((b - 1.) / (b * log(b)), 1.5);
        if (b != b) b = 0.;
        w[i] = b;
    }
    double tw = 0.0;
    for (int i = 0; i < dim; i++) tw += w[i];
    for (int i = 0; i < dim; i++) nb[i] = bounds[i + 1];
    double xold, xnew = 0., dw = 0., pts = tw / dim;
    int j = 0;
    for (int i = 0; i < dim; i++) {
        dw += w[i];
        xold = xnew;
        xnew = bounds[i + 1];
        for (; dw > pts; j++) {
            dw -= pts;
            if (j < dim) nb[j] = xnew - dw * (xnew - xold) / w[i];
        }
    }
    for (int i = 0; i < dim - 1; i++) bounds[1 + i] = nb[i];
    free(w);
    free(nb);
}

#define VBINS 32
int orc_vegas_oa(int nleaf, const orc_leaf *leaves, int nq, const double *qs, uint64_t seed,
                 int iterations, int ppb, double eps, double *out, int *iters_used) {
    const double step = leaves_step(leaves);
    const int npts = VBINS * VBINS * ppb;
    int layer = 0, q0 = 0, q1 = 0, used = 0;
    while (next_group(qs, nq, q0, step, &layer, &q1)) {
        int m = q1 - q0;
        double bt[VBINS + 1], bp[VBINS + 1];
        for (int i = 0; i <= VBINS; i++) bt[i] = bp[i] = (double)i / VBINS;
        double *hist = (double *)calloc((size_t)iterations * m, sizeof(double));
        double *I = (double *)malloc(sizeof(double) * (size_t)npts * m);
        int k;
        for (k = 0; k < iterations; k++) {
#pragma omp parallel
            {
                double *amp = (double *)malloc(sizeof(double) * 2 * m);
#pragma omp for schedule(static)
                for (int idx = 0; idx < npts; idx++) {
                    int bin = idx / ppb, ti = bin / VBINS, pj = bin % VBINS;
                    double u, v;
                    orc_uniform_pair(seed, (uint32_t)idx, (uint32_t)k, (uint32_t)layer,
                                     ORC_STREAM_VEGAS, &u, &v);
                    double th = acos(1. - 2. * (u * (bt[ti + 1] - bt[ti]) + bt[ti]));
                    double ph = (v * (bp[pj + 1] - bp[pj]) + bp[pj]) * 2. * M_PI;
                    orc_docked_group(nleaf, leaves, layer, th, ph, m, qs + q0, amp);
                    for (int j = 0; j < m; j++)
                        I[(size_t)idx * m + j] =
                            amp[2 * j] * amp[2 * j] + amp[2 * j + 1] * amp[2 * j + 1];
                }
                free(amp);
            }
            /* HybridVEGASBinReduceToIKernel and HybridVEGASBinReduceKernel */
            double *est = hist + (size_t)k * m;
            double vt[VBINS] = {0}, vp[VBINS] = {0};
            for (int bin = 0; bin < VBINS * VBINS; bin++) {
                int ti = bin / VBINS, pj = bin % VBINS;
                double vol = (bt[ti + 1] - bt[ti]) * (bp[pj + 1] - bp[pj]);
                for (int j = 0; j < m; j++) {
                    double s = 0.;
                    for (int p = 0; p < ppb; p++) s += I[(size_t)(bin * ppb + p) * m + j];
                    est[j] += s * vol / ppb;
                }
                double mean = 0., ssq = 0.;
                for (int p = 0; p < ppb; p++) {
                    double so = 0.;
                    for (int j = 0; j < m; j++) so += I[(size_t)(bin * ppb + p) * m + j];
                    double diff = vol * so - mean;
                    mean += diff / (double)(p + 1);
                    ssq += diff * diff * (double)p / (double)(p + 1);
                }
                vt[ti] += ssq * ppb;
                vp[pj] += ssq * ppb;
            }
            vegas_resize(VBINS, vt, bt);
            vegas_resize(VBINS, vp, bp);
            if (k > 0)
                for (int j = 0; j < m; j++)
                    est[j] = (hist[(size_t)(k - 1) * m + j] * (double)k + est[j]) / (double)(k + 1);
            if (converged3(hist, k, m, eps)) {
                k++;
                break;
            }
        }
        if (k > used) used = k;
        for (int j = 0; j < m; j++) out[q0 + j] = hist[(size_t)(k - 1) * m + j];
        free(hist);
        free(I);
        q0 = q1;
        layer++;
    }
    if (iters_used) *iters_used = used;
    return 0;
}

/* Gauss-Kronrod 15/7 tables (Common/mathfuncs.h:100-125) */
static const double k15x[15] = {
    -0.9914553711208126392068547, -0.9491079123427585245261897, -0.8648644233597690727897128,
    -0.7415311855993944398638648, -0.5860872354676911302941448, -0.4058451513773971669066064,
    -0.2077849550078984676006894, 0.0,
    0.2077849550078984676006894,  0.4058451513773971669066064,  0.5860872354676911302941448,
    0.7415311855993944398638648,  0.8648644233597690727897128,  0.9491079123427585245261897,
    0.9914553711208126392068547};
static const double k15w[15] = {
    0.0229353220105292249637320, 0.0630920926299785532907007, 0.1047900103222501838398763,
    0.1406532597155259187451896, 0.1690047266392679028265834, 0.1903505780647854099132564,
    0.2044329400752988924141620, 0.2094821410847278280129992, 0.2044329400752988924141620,
    0.1903505780647854099132564, 0.1690047266392679028265834, 0.1406532597155259187451896,
    0.1047900103222501838398763, 0.0630920926299785532907007, 0.0229353220105292249637320};
static const double g7w[7] = {0.1294849661688696932706114, 0.2797053914892766679014678,
                              0.3818300505051189449503698, 0.4179591836734693877551020,
                              0.3818300505051189449503698, 0.2797053914892766679014678,
                              0.1294849661688696932706114};

typedef struct {
    int nleaf;
    const orc_leaf *leaves;
    long long evals;
} gk_ctx;

/* Amplitude.cpp:1444-1467 */
static double gk_phi(gk_ctx *c, double q, double st, double ct, double pMin, double pMax, double eps,
                     long long maxDepth, long long minDepth) {
    double kVal = 0.0, gVal = 0.0, ba2 = (pMax - pMin) * 0.5, apb2 = (pMax + pMin) * 0.5;
    for (int i = 0; i < 15; i++) {
        double xi = apb2 + ba2 * k15x[i];
        double a[2];
        orc_docked_cart(c->nleaf, c->leaves, q * st * cos(xi), q * st * sin(xi), q * ct, a);
        c->evals++;
        double I = a[0] * a[0] + a[1] * a[1];
        kVal += I * ba2 * k15w[i];
        if (i % 2 == 1) gVal += I * ba2 * g7w[i / 2];
    }
    if (minDepth < 0 && (maxDepth < 0 || fabs(1.0 - (gVal / kVal)) <= eps || gVal == kVal))
        return kVal;
    return gk_phi(c, q, st, ct, pMin, apb2, eps, maxDepth - 1, minDepth - 1) +
           gk_phi(c, q, st, ct, apb2, pMax, eps, maxDepth - 1, minDepth - 1);
}

/* Amplitude.cpp:1421-1442 */
static double gk_theta(gk_ctx *c, double q, double tMin, double tMax, double eps,
                       long long maxDepth, long long minDepth) {
    double kVal = 0.0, gVal = 0.0, ba2 = (tMax - tMin) * 0.5, apb2 = (tMax + tMin) * 0.5;
    for (int i = 0; i < 15; i++) {
        double sT = sin(apb2 + ba2 * k15x[i]);
        double v = gk_phi(c, q, sT, cos(apb2 + ba2 * k15x[i]), 0.0, M_2PI_D, eps, maxDepth,
                          minDepth);
        kVal += v * k15w[i] * ba2 * sT;
        if (i % 2 == 1) gVal += v * g7w[i / 2] * ba2 * sT;
    }
    if (minDepth < 0 && (maxDepth < 0 || fabs(1.0 - (gVal / kVal)) <= eps || gVal == kVal))
        return kVal;
    double mid = (tMax + tMin) / 2.0;
    return gk_theta(c, q, tMin, mid, eps, maxDepth - 1, minDepth - 1) +
           gk_theta(c, q, mid, tMax, eps, maxDepth - 1, minDepth - 1);
}

/* Amplitude.cpp:1416-1419 (GaussKron2DSphereRecurs) and 1479-1484 (q = 0) */
int orc_gk_oa(int nleaf, const orc_leaf *leaves, int nq, const double *qs, double eps,
              long long max_depth, double *out, long long *evals) {
    long long total = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
    for (int k = 0; k < nq; k++) {
        gk_ctx c = {nleaf, leaves, 0};
        if (qs[k] == 0.0) {
            double a[2];
            orc_docked_cart(nleaf, leaves, 0.0, 0.0, 0.0, a);
            c.evals++;
            out[k] = a[0] * a[0] + a[1] * a[1];
        } else {
            out[k] = gk_theta(&c, qs[k], 0.0, M_PI / 2.0, eps, max_depth - 1, -1) / (2.0 * M_PI);
        }
        total += c.evals;
    }
    if (evals) *evals = total;
    return 0;
}

/* Grid.cpp:399-484 with caller-supplied directions: results[i] = running mean,
 * convergence against i/2, 2i/3 and 3i/4 once i >= iterations/20 */
double orc_grid_intensity_dirs(int grid_size, double qmax, const double *data,
                               const double *coeffs, double q, double eps, long long n,
                               const double *phi, const double *ct, long long *used) {
    if (q == 0.0) {
        if (used) *used = 0;
        return data[0] * data[0] + data[1] * data[1];
    }
    double *res = (double *)malloc(sizeof(double) * (size_t)n);
    long long minIter = n / 20;
    double acc = 0.0;
    for (long long i = 0; i < n; i++) {
        double th = acos(ct[i]);
        double st = sin(th), c = cos(th), sp = sin(phi[i]), cp = cos(phi[i]);
        double a[2];
        orc_get_cart(grid_size, qmax, data, coeffs, q * st * cp, q * st * sp, q * c, a);
        acc += a[0] * a[0] + a[1] * a[1];
        res[i] = acc / (double)(i + 1);
        if (i >= minIter && eps > 0.0) {
            if (fabs(1.0 - res[i] / res[i >> 1]) < eps &&
                fabs(1.0 - res[i] / res[(i << 1) / 3]) < eps &&
                fabs(1.0 - res[i] / res[(3 * i) >> 2]) < eps) {
                double r = res[i];
                free(res);
                if (used) *used = i + 1;
                return r;
            }
        }
    }
    double r = res[n - 1];
    free(res);
    if (used) *used = n;
    return r;
}


/* ------------------------------------------------------------------------- */
/* std::mt19937 (Matsumoto & Nishimura 1998) and libstdc++'s                  */
/* uniform_real_distribution<double> (generate_canonical with two 32-bit      */
/* draws), the stream JacobianSphereGrid::CalculateIntensity uses             */
/* (Grid.cpp:415-425: phi ~ U(phi_min, phi_max), then v2 ~ U(0, 2)).          */
/* ------------------------------------------------------------------------- */
void orc_mt19937_pairs(uint32_t seed, long long n, double a0, double b0, double a1, double b1,
                       double *out) {
    uint32_t mt[624];
    int mti;
    mt[0] = seed;
    for (mti = 1; mti < 624; mti++)
        mt[mti] = 1812433253u * (mt[mti - 1] ^ (mt[mti - 1] >> 30)) + (uint32_t)mti;
    int pos = 624;
    for (long long k = 0; k < 2 * n; k++) {
        uint32_t w[2];
        for (int h = 0; h < 2; h++) {
            if (pos >= 624) {
                for (int i = 0; i < 624; i++) {
                    uint32_t y = (mt[i] & 0x80000000u) | (mt[(i + 1) % 624] & 0x7fffffffu);
                    mt[i] = mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
                }
                pos = 0;
            }
            uint32_t y = mt[pos++];
            y ^= (y >> 11);
            y ^= (y << 7) & 0x9d2c5680u;
            y ^= (y << 15) & 0xefc60000u;
            y ^= (y >> 18);
            w[h] = y;
        }
        double c = ((double)w[0] + (double)w[1] * 4294967296.0) / 18446744073709551616.0;
        if (c >= 1.0) c = nextafter(1.0, 0.0);
        double a = (k & 1) ? a1 : a0, b = (k & 1) ? b1 : b0;
        out[k] = a + (b - a) * c;
    }
}

/* ------------------------------------------------------------------------- */
/* Debye: I(q) = sum_ij f_i f_j sinc(q r_ij)                                  */
/* ------------------------------------------------------------------------- */
void orc_debye(int natoms, const double *xyz, const int *ion, int nion, const double *coef9,
               const double *rad, const int *count, double sol_ed, double c1, int flags,
               double scale, int nq, const double *qs, double *out) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < nq; k++) {
        double ff[256];
        orc_form_factors(nion, coef9, rad, count, sol_ed, c1, flags, qs[k], ff);
        double q = qs[k];
        double s = 0.0;
        for (int i = 0; i < natoms; i++) {
            double fi = ff[ion[i]];
            s += fi * fi;
            double row = 0.0;
            for (int j = i + 1; j < natoms; j++) {
                double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1],
                       dz = xyz[3 * i + 2] - xyz[3 * j + 2];
                double x = q * sqrt(dx * dx + dy * dy + dz * dz);
                row += ff[ion[j]] * (x == 0.0 ? 1.0 : sin(x) / x);
            }
            s += 2.0 * fi * row;
        }
        out[k] = s * scale * scale;
    }
}
