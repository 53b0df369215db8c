/*
 * CPU oracle for the wavelet-compressed sensitivity matrix (method 2 of the
 * reference).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
 * this file.
 *
 * Restates:
 *   - Fwd::compute_G_wavelet   src/Forward_modelling/Fwd.cpp:148-219
 *       row (k*N_obs + i) = field k of point i for the cells in the mesh's
 *       reordered order (column j <-> cell order[j], Mesh::get_reordered_id),
 *       transformed and thresholded by compress_vec
 *   - Fwd::compress_vec        Fwd.cpp:298-371
 *       zero padding to n_dy = the next power of two (n itself when it is
 *       one), forward transform, keep coefficient c when
 *       |c| > thr or ||c| - thr| < 1e-15, thr = relative_threshold * max|c|
 *   - Fwd::G_vec_mul / GT_vec_mul, wavelet branches   Fwd.cpp:80-147
 *   - the transform itself: GSL's gsl_wavelet_transform_forward / _inverse
 *     with gsl_wavelet_daubechies, k = 4 (GSL is an external dependency of
 *     the reference, not vendored, version not pinned).  Its published
 *     algorithm: Mallat's pyramid with periodic boundaries on n = 2^J
 *     samples; one analysis step on the leading n entries maps a to
 *       s[ii] = sum_k h[k] a[(2ii + k) mod n],  d[ii] = sum_k g[k] a[(2ii + k) mod n]
 *     (ii < n/2, non-centred: no offset), stored as a[0..n/2) = s,
 *     a[n/2..n) = d, for n = N, N/2, ..., 2; synthesis runs the transposed
 *     steps for n = 2, 4, ..., N, accumulating h[k] s[ii] + g[k] d[ii] into
 *     position (2ii + k) mod n in ascending ii, then k.  h is the
 *     Daubechies (1988) 4-tap orthonormal low-pass filter
 *     ((1+sqrt3), (3+sqrt3), (3-sqrt3), (1-sqrt3)) / (4 sqrt2) and
 *     g[k] = (-1)^k h[3-k].
 * Built without FMA contraction (oracle/Makefile), every sum in the order
 * written.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

extern int ora_threads;
int ora_prism_field(const double *point, const double *prism, double rho, uint32_t flag,
                    double *out10);

/* (1 + sqrt3)/(4 sqrt2), (3 + sqrt3)/(4 sqrt2), (3 - sqrt3)/(4 sqrt2),
 * (1 - sqrt3)/(4 sqrt2) to 29 digits */
static const double D4H[4] = {0.48296291314453414337487159986, 0.83651630373780790557529378092,
                              0.22414386804201338102597276224, -0.12940952255126038117444941881};
static const double D4G[4] = {-0.12940952255126038117444941881, -0.22414386804201338102597276224,
                              0.83651630373780790557529378092, -0.48296291314453414337487159986};

void ora_d4_filters(double *h, double *g)
{
    for (int k = 0; k < 4; k++) { h[k] = D4H[k]; g[k] = D4G[k]; }
}

static void d4_step_fwd(double *a, int64_t n, double *tmp)
{
    int64_t nh = n >> 1, mask = n - 1;
    for (int64_t ii = 0; ii < nh; ii++) {
        double s = 0.0, d = 0.0;
        for (int k = 0; k < 4; k++) {
            double v = a[(2 * ii + k) & mask];
            s += D4H[k] * v;
            d += D4G[k] * v;
        }
        tmp[ii] = s;
        tmp[ii + nh] = d;
    }
    memcpy(a, tmp, sizeof(double) * (size_t)n);
}

static void d4_step_inv(double *a, int64_t n, double *tmp)
{
    int64_t nh = n >> 1, mask = n - 1;
    for (int64_t j = 0; j < n; j++) tmp[j] = 0.0;
    for (int64_t ii = 0; ii < nh; ii++) {
        double s = a[ii], d = a[ii + nh];
        for (int k = 0; k < 4; k++) tmp[(2 * ii + k) & mask] += D4H[k] * s + D4G[k] * d;
    }
    memcpy(a, tmp, sizeof(double) * (size_t)n);
}

/* In-place full transform of a[0..n), n a power of two >= 2; dir > 0
 * forward, else inverse.  Returns 0, or 1 if n is not a power of two. */
int ora_d4_transform(double *a, int64_t n, int dir)
{
    if (n < 2 || (n & (n - 1))) return 1;
    double *tmp = (double *)malloc(sizeof(double) * (size_t)n);
    if (dir > 0)
        for (int64_t m = n; m >= 2; m >>= 1) d4_step_fwd(a, m, tmp);
    else
        for (int64_t m = 2; m <= n; m <<= 1) d4_step_inv(a, m, tmp);
    free(tmp);
    return 0;
}

/* n_dy of compress_vec: n when log2(n) is an integer, else the next power
 * of two (Fwd.cpp:302-311). */
int64_t ora_wavelet_length(int64_t n)
{
    int64_t p = 1;
    while (p < n) p <<= 1;
    return p;
}

static int comp_bits(uint32_t flag, int *bits)
{
    int n = 0;
    for (int b = 1; b < 10; b++)
        if (flag & (1u << b)) bits[n++] = b;
    return n;
}

/* keep rule of compress_vec (Fwd.cpp:347-352) */
static inline int keep(double c, double thr)
{
    return fabs(c) > thr || fabs(fabs(c) - thr) < 1e-15;
}

/*
 * compute_G_wavelet.  Pass rowptr == NULL to get counts[row] (kept
 * coefficients of each of the n_comp*N_obs rows); then pass the exclusive
 * prefix sums as rowptr (n_rows + 1 entries) to fill ci / v, each row's
 * coefficients in ascending index.  Returns 0, or 1 for degenerate geometry.
 */
int ora_wavelet_rows(const double *prisms, int64_t Nm, const double *pts, int64_t N_obs,
                     uint32_t flag, const int64_t *order, double rel, int64_t n_dy,
                     int64_t *counts, const int64_t *rowptr, int32_t *ci, double *v)
{
    int bits[10];
    int nc = comp_bits(flag, bits);
    int st = 0;
#pragma omp parallel num_threads(ora_threads) reduction(| : st)
    {
        double *row = (double *)malloc(sizeof(double) * (size_t)n_dy * nc);
        double *tmp = (double *)malloc(sizeof(double) * (size_t)n_dy);
#pragma omp for schedule(dynamic, 4)
        for (int64_t i = 0; i < N_obs; i++) {
            double f[10];
            for (int64_t j = 0; j < Nm; j++) {
                st |= ora_prism_field(pts + 3 * i, prisms + 6 * order[j], 1.0, flag, f);
                for (int k = 0; k < nc; k++) row[k * n_dy + j] = f[bits[k]];
            }
            for (int k = 0; k < nc; k++) {
                double *a = row + k * n_dy;
                for (int64_t j = Nm; j < n_dy; j++) a[j] = 0.0;
                for (int64_t m = n_dy; m >= 2; m >>= 1) d4_step_fwd(a, m, tmp);
                double mx = fabs(a[0]);
                for (int64_t j = 1; j < n_dy; j++)
                    if (mx < fabs(a[j])) mx = fabs(a[j]);
                double thr = rel * mx;
                int64_t r = k * N_obs + i;
                if (!rowptr) {
                    int64_t c = 0;
                    for (int64_t j = 0; j < n_dy; j++) c += keep(a[j], thr);
                    counts[r] = c;
                } else {
                    int64_t o = rowptr[r];
                    for (int64_t j = 0; j < n_dy; j++)
                        if (keep(a[j], thr)) { ci[o] = (int32_t)j; v[o] = a[j]; o++; }
                }
            }
        }
        free(row);
        free(tmp);
    }
    return st;
}

/* G_vec_mul, wavelet branch (Fwd.cpp:114-143): y = C W (x in reordered
 * order, zero-padded). */
void ora_wavelet_gemv_n(const int64_t *rowptr, const int32_t *ci, const double *v, int64_t n_rows,
                        int64_t n_dy, const int64_t *order, int64_t Nm, const double *x,
                        double *y)
{
    double *xw = (double *)malloc(sizeof(double) * (size_t)n_dy);
    for (int64_t j = 0; j < n_dy; j++) xw[j] = j < Nm ? x[order[j]] : 0.0;
    ora_d4_transform(xw, n_dy, 1);
#pragma omp parallel for num_threads(ora_threads)
    for (int64_t i = 0; i < n_rows; i++) {
        double s = 0.0;
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++) s += v[k] * xw[ci[k]];
        y[i] = s;
    }
    free(xw);
}

/* GT_vec_mul, wavelet branch (Fwd.cpp:80-113): rows scattered in ascending
 * order, inverse transform, back to the original cell order. */
void ora_wavelet_gemv_t(const int64_t *rowptr, const int32_t *ci, const double *v, int64_t n_rows,
                        int64_t n_dy, const int64_t *order, int64_t Nm, const double *r,
                        double *z)
{
    double *pw = (double *)calloc((size_t)n_dy, sizeof(double));
    for (int64_t i = 0; i < n_rows; i++)
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++) pw[ci[k]] += v[k] * r[i];
    ora_d4_transform(pw, n_dy, -1);
    for (int64_t j = 0; j < Nm; j++) z[order[j]] = pw[j];
    free(pw);
}
