// Device-resident CGLS for GaussNewtonInversion::solve_cg
// (src/Inversion/GaussNewtonInversion.cpp:1417-1698) and the Gauss-Newton
// helpers around it (:856-861 barrier P, :877-911 model update, :913-916
// misfit, :962-1022 IRLS weights).
//
// One CG iteration = 6 launches on the context stream, none of which waits on
// the host:
//   p_update (ends the previous iteration: beta, err, stop test; then
//   p = -r + beta p) -> gemv_n(G, Pp) -> t_rows (+ t^2 partials)
//   -> s_update (alpha; s += a t, dx += a p, w = Wd s_d) -> gemv_t(G, w)
//   -> r_loop (r = A^T s, + r^2 partials)
// and one scalar launch ends the last iteration.  The scalars a kernel needs
// from the previous kernel's block partials (alpha, beta) are summed by every
// block of the consuming kernel in the same fixed order, so each block holds
// the identical value and block 0 records it.  The scalars (r2, t2, alpha,
// beta, err, iterations) live in device memory; when err < tol `done` is set
// and every later kernel of the queued iterations returns at once.  All
// reductions have a fixed shape, so a solve is bit-for-bit repeatable.
//
// Operator association follows the reference's Eigen expressions: a term such
// as sqrt(l) * P * Ws^T * WL_s^T * v is summed as ((sqrt(l)*P_i)*Ws_ji)*WL_j*v_j
// over the row of the transposed operator.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>

#include "engine.hpp"

namespace gra {

namespace {

constexpr int kB = 256;

struct Csr {
    const int64_t *rp;
    const int32_t *ci;
    const double *v;
};

// Everything the row kernels need, passed by value (kernel argument space).
struct CgDev {
    int64_t nd, nm, ld, nrow;
    int nseg;                 // 4, or 7 with the cross-gradient term
    int quirks;
    double sl;                // sqrt(lambda)
    const double *Wd, *P, *dobs, *mk, *mref;
    const double *WL[7];      // s, z, x, y, tz, tx, ty (null = identity)
    const double *WLt_rows[3];// right-diagonals used for t rows 4..6
    Csr A[7];                 // Ws, Wz, Wx, Wy, Tz, Tx, Ty (T scaled by a*S)
    Csr AT[7];                // transposes
    double *s, *t, *r, *p, *Pp, *dx, *w;
    const double *part_n;
    int splits;
    const double *part_t;
    int chunks;
    double *red;
    double *red2;
    CgScalars *sc;
    // single-read fused iteration (see fused_kernels.hip)
    double *rloop;            // loop-operator residual of s0 (vec8 with WL_ty), or null
    const double *zpart;      // [zgroups][ld] G^T (Wd^2 G Pp) partials per row group
    int zgroups;
    const double *t2part;     // [nt2] ||t_d||^2 partials of the row blocks
    int64_t nt2;
    // mixed precision (inner products on a reduced-precision copy of G,
    // FP64 residual refreshes): b kept for the refresh, and its work vectors
    double *bvec;             // [nrow] b, or null
    double *vmk;              // [ld]   mk + P dx  (data rows of the refresh)
    double *vpdx;             // [nm]   P dx       (regularisation rows)
    int w_from_t;             // two-pass recurrence: w = Wd t_d (k_t_rows), not Wd s_d
    // iteration index from sc->it_dev instead of the launch argument (set
    // for iterations replayed from a captured graph); k_r_loop advances it
    int dev_it;
};

__device__ inline double dgv(const double *d, int64_t i) { return d ? d[i] : 1.0; }

// Deterministic block sum (256 threads): xor butterfly per wave, then the
// four wave sums in order.  Result valid in thread 0.
__device__ inline double block_sum(double v)
{
    __shared__ double ws[kB / 64];
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) ws[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        s = ws[0];
        for (int q = 1; q < kB / 64; ++q) s += ws[q];
    }
    return s;
}

// p[0] + p[ld] + ... + p[(n-1) ld], added in that order; the loads are issued
// eight at a time (one lane may sum a hundred row-group partials, and one
// load in flight at a time left the sum latency-bound).
__device__ inline double sum_strided(const double *p, int n, int64_t ld)
{
    double v = p[0];
    int g = 1;
    for (; g + 7 < n; g += 8) {
        double t[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) t[u] = p[(int64_t)(g + u) * ld];
#pragma unroll
        for (int u = 0; u < 8; ++u) v += t[u];
    }
    for (; g < n; ++g) v += p[(int64_t)g * ld];
    return v;
}

// Sum of n partials by one block, fixed order.
__device__ inline double sum_partials(const double *red, int64_t n)
{
    double v = 0.0;
    for (int64_t k = threadIdx.x; k < n; k += kB) v += red[k];
    return block_sum(v);
}

// out = sum_k ((pre*L_i)*A_ik) [*R_k] * x_k over row i of A.
__device__ inline double row_dot(const Csr &A, int64_t i, double pre, const double *L,
                                 const double *R, const double *x)
{
    double li = pre * dgv(L, i);
    double acc = 0.0;
    for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
        int64_t j = A.ci[k];
        double coef = li * A.v[k];
        if (R) coef = coef * R[j];
        acc += coef * x[j];
    }
    return acc;
}

// Segment q's left-multiplier for t rows (quirks: GN.cpp:1573, :1582-1584).
__device__ inline void t_row_factors(const CgDev &d, int q, double &pre, const double *&L)
{
    if (q < 4) {
        pre = d.sl;
        L = (q == 0 && d.quirks) ? nullptr : d.WL[q];
    } else {
        pre = 1.0;
        L = d.WLt_rows[q - 4];
    }
}

__global__ void __launch_bounds__(kB) k_init_scalars(CgScalars *sc)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        memset(sc, 0, sizeof(CgScalars));
    }
}

// b and s = -b (GN.cpp:1459-1472); w = Wd * s_d for the first G^T product.
__global__ void __launch_bounds__(kB) k_b_init(CgDev d)
{
    int64_t g = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (g >= d.nrow) return;
    double b;
    if (g < d.nd) {
        double y = d.part_n[g];
        for (int q = 1; q < d.splits; ++q) y += d.part_n[(int64_t)q * d.nd + g];
        b = dgv(d.Wd, g) * (d.dobs[g] - y);
        d.w[g] = dgv(d.Wd, g) * (-b);
    } else {
        int64_t h = g - d.nd;
        int q = (int)(h / d.nm);
        int64_t i = h - (int64_t)q * d.nm;
        if (q == 0) {
            // -sqrt(l) * WL_s * Ws * (mk - mref)
            double li = -d.sl * dgv(d.WL[0], i);
            double acc = 0.0;
            for (int64_t k = d.A[0].rp[i]; k < d.A[0].rp[i + 1]; ++k) {
                int64_t j = d.A[0].ci[k];
                acc += (li * d.A[0].v[k]) * (d.mk[j] - d.mref[j]);
            }
            b = acc;
        } else if (q < 4) {
            b = row_dot(d.A[q], i, -d.sl, d.WL[q], nullptr, d.mk);
        } else {
            b = row_dot(d.A[q], i, -1.0, d.WL[q], nullptr, d.mk);
        }
    }
    d.s[g] = -b;
    if (d.bvec) d.bvec[g] = b;
}

// FP64 residual refresh of the mixed-precision iteration, part 1: the vectors
// the true s = A_t (P dx) - b needs: mk + P dx for the data rows (one product
// with the FP64 G) and P dx for the regularisation rows.
__global__ void __launch_bounds__(kB) k_refresh_vec(CgDev d)
{
    if (d.sc->done) return;
    int64_t j = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (j >= d.ld) return;
    if (j < d.nm) {
        const double pdx = d.P[j] * d.dx[j];
        d.vpdx[j] = pdx;
        d.vmk[j] = d.mk[j] + pdx;
    } else {
        d.vmk[j] = 0.0;
    }
}

// Part 2, after y = G (mk + P dx) in d.part_n (FP64 G's split count in
// d.splits): s = A_t (P dx) - b row by row -- data rows
// Wd (G (mk + P dx) - dobs), regularisation rows as k_t_rows forms t from
// P dx, minus b -- and w = Wd s_d for the G^T product of the true residual.
// The running s and the iterate dx satisfy s = A_t (P dx) - b in exact
// arithmetic; this replaces the recurrence's s (which carries the
// reduced-precision products) by its FP64 value.
__global__ void __launch_bounds__(kB) k_refresh_s(CgDev d)
{
    if (d.sc->done) return;
    int64_t g = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (g >= d.nrow) return;
    double sv;
    if (g < d.nd) {
        double y = d.part_n[g];
        for (int q = 1; q < d.splits; ++q) y += d.part_n[(int64_t)q * d.nd + g];
        sv = dgv(d.Wd, g) * (y - d.dobs[g]);
        d.w[g] = dgv(d.Wd, g) * sv;
    } else {
        int64_t h = g - d.nd;
        int q = (int)(h / d.nm);
        int64_t i = h - (int64_t)q * d.nm;
        double pre;
        const double *L;
        t_row_factors(d, q, pre, L);
        sv = row_dot(d.A[q], i, pre, L, nullptr, d.vpdx) - d.bvec[g];
    }
    d.s[g] = sv;
}

// r_j = P_j (G^T w)_j + sum_q (pre P_j) * AT_q row j * WL_q * seg_q   (A^T v).
// mode 0: loop residual (vec8 uses WL_ty); mode 1: initial residual plus
// A^T b, whose vec8 uses WL_y under quirks (GN.cpp:1508 vs :1544).
template <int MODE>
__global__ void __launch_bounds__(kB) k_r_update(CgDev d)
{
    if (d.sc->done) return;
    int64_t j = (int64_t)blockIdx.x * kB + threadIdx.x;
    double r2 = 0.0, atb2 = 0.0;
    if (j < d.nm) {
        double gsum = sum_strided(d.part_t + j, d.chunks, d.ld);
        double Pj = d.P[j];
        double r = Pj * gsum;
        double rl = r;
        double atb = 0.0 + Pj * (-gsum);
        const double *seg = d.s + d.nd;
        for (int q = 0; q < d.nseg; ++q) {
            double pre = q < 4 ? d.sl : 1.0;
            const double *R = d.WL[q];
            if (MODE == 1 && q == 6 && d.quirks) R = d.WL[3];   // WL_y (GN.cpp:1508)
            double v = row_dot(d.AT[q], j, pre, d.P, R, seg + (int64_t)q * d.nm);
            r = r + v;
            if (MODE == 1) {
                double vl = (q == 6 && d.quirks)
                                ? row_dot(d.AT[q], j, pre, d.P, d.WL[6], seg + (int64_t)q * d.nm)
                                : v;
                rl = rl + vl;
                atb += -vl;
            }
        }
        d.r[j] = r;
        if (MODE == 1 && d.rloop) d.rloop[j] = rl;
        r2 = r * r;
        atb2 = atb * atb;
    }
    double s1 = block_sum(r2);
    if (threadIdx.x == 0) d.red[blockIdx.x] = s1;
    if (MODE == 1) {
        double s2 = block_sum(atb2);
        if (threadIdx.x == 0) d.red2[blockIdx.x] = s2;
    }
}

// Loop residual r = A^T s with eight lanes per cell j: lane 0 sums the data
// part's chunk partials, lane 1+q forms segment q's transposed row product;
// lane 0 then adds them in the order of the single-lane form
// r = ((P_j g_j + v_0) + v_1) + ..., so results are identical to it while the
// sparse-row latency chains run in parallel.  FUSED: the single-read
// iteration's r <- base + alpha * (P_j z_j + sum_q v_q) over t instead of s,
// z_j the sum of the fused product's row-group partials; in the first
// iteration the base is the loop-operator residual of s0 when it differs from
// the initial one (reference quirk at GN.cpp:1508).
constexpr int kLanesPerCell = 8;

template <bool FUSED, bool FIRST>
__global__ void __launch_bounds__(kB) k_r_loop(CgDev d)
{
    if (d.sc->done) return;
    const int sub = threadIdx.x & (kLanesPerCell - 1);
    const int64_t j = ((int64_t)blockIdx.x * kB + threadIdx.x) / kLanesPerCell;
    const bool ok = j < d.nm;
    double v = 0.0;
    if (ok) {
        if (sub == 0) {
            if (FUSED) v = sum_strided(d.zpart + j, d.zgroups, d.ld);
            else v = sum_strided(d.part_t + j, d.chunks, d.ld);
        } else if (sub <= d.nseg) {
            const int q = sub - 1;
            const double *seg = (FUSED ? d.t : d.s) + d.nd;
            double pre = q < 4 ? d.sl : 1.0;
            v = row_dot(d.AT[q], j, pre, d.P, d.WL[q], seg + (int64_t)q * d.nm);
        }
    }
    // lane group of cell j: lanes sub = 0..7 of the same wave
    const int lane = threadIdx.x & 63, base_lane = lane & ~(kLanesPerCell - 1);
    double vq[kLanesPerCell];
#pragma unroll
    for (int q = 0; q < kLanesPerCell; ++q) vq[q] = __shfl(v, base_lane + q, 64);
    double r2 = 0.0;
    if (ok && sub == 0) {
        double r = d.P[j] * vq[0];
        for (int q = 0; q < d.nseg; ++q) r = r + vq[1 + q];
        if (FUSED) {
            double b = (FIRST && d.rloop) ? d.rloop[j] : d.r[j];
            r = b + d.sc->alpha * r;
        }
        d.r[j] = r;
        r2 = r * r;
    }
    double s = block_sum(r2);
    if (threadIdx.x == 0) d.red[blockIdx.x] = s;
    // the next iteration's kernels read the advanced index (no block of this
    // kernel reads it)
    if (d.dev_it && blockIdx.x == 0 && threadIdx.x == 0) d.sc->it_dev += 1;
}

// After the initial residual: r2, ATb2.
__global__ void __launch_bounds__(kB) k_scalar_init(CgDev d, int64_t nblk)
{
    double r2 = sum_partials(d.red, nblk);
    double atb2 = sum_partials(d.red2, nblk);
    if (threadIdx.x == 0) {
        d.sc->r2[0] = r2;
        d.sc->ATb2 = atb2;
        d.sc->beta = 0.0;
        d.sc->done = 0;
        d.sc->iterations = 0;
    }
}

// End of iteration it (GN.cpp:1674-1691) from the r^2 partials of its
// residual: beta = r2/r2_last, err = r2/ATb2 and the stop test.  Every
// thread gets beta and the stop flag; block 0 records the scalars (no other
// block reads the fields it writes: r2 goes to the slot of iteration it+1).
__device__ inline void cg_end_iteration(const CgDev &d, int64_t nblk, int it, double tol,
                                        double &beta, bool &stop)
{
    __shared__ double beta_sh;
    __shared__ int stop_sh;
    double r2 = sum_partials(d.red, nblk);
    if (threadIdx.x == 0) {
        CgScalars *sc = d.sc;
        double b = r2 / sc->r2[it & 1];
        double err = r2 / sc->ATb2;
        int st = err < tol;
        if (blockIdx.x == 0) {
            sc->r2[(it + 1) & 1] = r2;
            sc->beta = b;
            sc->err = err;
            sc->iterations = it + 1;
            if (st) sc->done = 1;
        }
        beta_sh = b;
        stop_sh = st;
    }
    __syncthreads();
    beta = beta_sh;
    stop = stop_sh != 0;
}

// The stop flag as one value per block: block 0 of the p update may set it
// while other blocks of the same launch read it.
__device__ inline bool cg_done_block(const CgDev &d)
{
    __shared__ int done_sh;
    if (threadIdx.x == 0) done_sh = d.sc->done;
    __syncthreads();
    return done_sh != 0;
}

// Iteration it: end iteration it-1 (FIRST: it = 0, beta = 0 from the initial
// scalars), then p = -r + beta p; Pp = P p   (GN.cpp:1558-1560).  With the
// single-read iteration it also resets the fused product's nfill partials to
// the sentinel (the previous product has finished reading them).
template <bool FIRST>
__global__ void __launch_bounds__(kB) k_p_update(CgDev d, int64_t nblk, int it, double tol,
                                                 uint64_t *fill, int64_t nfill)
{
    for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < nfill;
         i += (int64_t)gridDim.x * kB)
        fill[i] = kFusedSentinel;
    if (cg_done_block(d)) return;
    if (d.dev_it) it = d.sc->it_dev;
    double beta;
    if (FIRST) {
        beta = d.sc->beta;
    } else {
        bool stop;
        cg_end_iteration(d, nblk, it - 1, tol, beta, stop);
        if (stop) return;
    }
    int64_t j = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (j >= d.nm) return;
    double p = -d.r[j] + beta * d.p[j];
    d.p[j] = p;
    d.Pp[j] = d.P[j] * p;
}

// t = A Pp (GN.cpp:1562-1587) with t^2 block partials.
__global__ void __launch_bounds__(kB) k_t_rows(CgDev d)
{
    if (d.sc->done) return;
    int64_t g = (int64_t)blockIdx.x * kB + threadIdx.x;
    double t2 = 0.0;
    if (g < d.nrow) {
        double t;
        if (g < d.nd) {
            double y = d.part_n[g];
            for (int q = 1; q < d.splits; ++q) y += d.part_n[(int64_t)q * d.nd + g];
            t = dgv(d.Wd, g) * y;
        } else {
            int64_t h = g - d.nd;
            int q = (int)(h / d.nm);
            int64_t i = h - (int64_t)q * d.nm;
            double pre;
            const double *L;
            t_row_factors(d, q, pre, L);
            t = row_dot(d.A[q], i, pre, L, nullptr, d.Pp);
        }
        d.t[g] = t;
        if (d.w_from_t && g < d.nd) d.w[g] = dgv(d.Wd, g) * t;
        t2 = t * t;
    }
    double s = block_sum(t2);
    if (threadIdx.x == 0) d.red[blockIdx.x] = s;
}

// alpha = r2 / t2   (GN.cpp:1588-1603) in every thread; block 0 records it.
// t2 sums the t_rows partials (for the single-read iteration those of
// k_t_rows_reg, which fold in the fused product's ||t_d||^2 partials), in a
// fixed order.
template <bool FUSED>
__device__ inline double cg_alpha(const CgDev &d, int64_t nblk, int it)
{
    __shared__ double a_sh;
    // the single-read iteration's partials already hold ||t_d||^2 slices
    // (k_t_rows_reg); both forms sum one partial per block of the t kernel
    (void)FUSED;
    const double t2 = sum_partials(d.red, nblk);
    if (threadIdx.x == 0) {
        double a = d.sc->r2[it & 1] / t2;
        if (blockIdx.x == 0) {
            d.sc->t2 = t2;
            d.sc->alpha = a;
            if (!(t2 != 0.0) || !isfinite(a)) d.sc->breakdown = 1;
        }
        a_sh = a;
    }
    __syncthreads();
    return a_sh;
}

// alpha; s += alpha t; dx += alpha p; w = Wd s_d   (GN.cpp:1588-1629, 1634)
template <bool FUSED>
__global__ void __launch_bounds__(kB) k_s_update(CgDev d, int64_t nblk, int it)
{
    if (d.sc->done) return;
    if (d.dev_it) it = d.sc->it_dev;
    double a = cg_alpha<FUSED>(d, nblk, it);
    int64_t g = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (g < d.nrow) {
        double s = d.s[g] + a * d.t[g];
        d.s[g] = s;
        if (g < d.nd && !d.w_from_t) d.w[g] = dgv(d.Wd, g) * s;
    }
    if (g < d.nm) d.dx[g] = d.dx[g] + a * d.p[g];
}

// End of the last iteration (one block).
__global__ void __launch_bounds__(kB) k_scalar_end(CgDev d, int64_t nblk, int iter, double tol)
{
    if (d.sc->done) return;
    double beta;
    bool stop;
    cg_end_iteration(d, nblk, iter, tol, beta, stop);
}

// ---- single-read fused iteration -------------------------------------------
// t rows of the regularisation blocks (the data rows t_d come from the fused
// kernel) with t^2 block partials.
__global__ void __launch_bounds__(kB) k_t_rows_reg(CgDev d)
{
    if (d.sc->done) return;
    int64_t h = (int64_t)blockIdx.x * kB + threadIdx.x;
    double t2 = 0.0;
    if (h < d.nrow - d.nd) {
        int q = (int)(h / d.nm);
        int64_t i = h - (int64_t)q * d.nm;
        double pre;
        const double *L;
        t_row_factors(d, q, pre, L);
        double t = row_dot(d.A[q], i, pre, L, nullptr, d.Pp);
        d.t[d.nd + h] = t;
        t2 = t * t;
    }
    double s = block_sum(t2);
    // this block's slice of the fused product's ||t_d||^2 partials, so that
    // alpha sums one partial per block instead of every block re-reading all
    // of them (fixed slices and order: repeatable)
    const int64_t per = (d.nt2 + gridDim.x - 1) / gridDim.x;
    const int64_t k0 = (int64_t)blockIdx.x * per;
    int64_t k1 = k0 + per;
    if (k1 > d.nt2) k1 = d.nt2;
    double v = 0.0;
    for (int64_t k = k0 + threadIdx.x; k < k1; k += kB) v += d.t2part[k];
    double sd = block_sum(v);
    if (threadIdx.x == 0) d.red[blockIdx.x] = s + sd;
}

// T*_scaled = (a*S_i) * T0_ij, and for the transpose (a*S_j) * T0_ji.
__global__ void k_scale_T(const int64_t *rp, const int32_t *ci, const double *v0, int64_t n,
                          double a, const double *S, int by_col, double *out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) {
        double si = S ? (by_col ? S[ci[k]] : S[i]) : 1.0;
        out[k] = (a * si) * v0[k];
    }
}

__global__ void k_init_count(int *c) { *c = 0; }

__global__ void k_fill(double *p, int64_t n, double v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---- Gauss-Newton helpers ----

__global__ void k_barrier_P(const double *m, const double *mmin, const double *mmax, int64_t n,
                            double *P)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    P[k] = (mmax[k] - m[k]) * (m[k] - mmin[k]) / (mmax[k] - mmin[k]);
}

// Regular cells in parallel; cells hitting the reference's overflow /
// underflow fallbacks are flagged and resolved in cell order afterwards.
__global__ void k_barrier_main(const double *m, const double *dx, const double *mmin,
                               const double *mmax, int64_t n, double *mnew, int *flags, int *count)
{
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    double e = exp(dx[id]);
    double t2 = mmin[id] * (mmax[id] - m[id]) + mmax[id] * (m[id] - mmin[id]) * e;
    double t3 = (mmax[id] - m[id]) + (m[id] - mmin[id]) * e;
    int f = 0;
    if (isinf(e) || isinf(t2)) f = 1;
    else if (fabs(t3) < 1e-10) f = 2;
    else mnew[id] = t2 / t3;
    flags[id] = f;
    if (f) atomicAdd(count, 1);
}

__global__ void k_barrier_copy(const double *mnew, int64_t n, double *m_trial, const int *count)
{
    if (*count) return;
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) m_trial[id] = mnew[id];
}

__global__ void k_barrier_seq(const double *mnew, const double *mmin, const double *mmax,
                              int64_t n, const int *flags, double *m_trial, const int *count)
{
    if (*count == 0 || threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int64_t id = 0; id < n; ++id) {
        int f = flags[id];
        if (f == 0) {
            m_trial[id] = mnew[id];
            continue;
        }
        double v = f == 1 ? (mmax[id] - 1e-6 + m_trial[id]) / 2.0
                          : (mmin[id] + 1e-6 + m_trial[id]) / 2.0;
        m_trial[id] = v;
        int64_t nb = id == 0 ? (n > 1 ? 1 : 0) : id - 1;
        m_trial[id] = 0.5 * (m_trial[id] + m_trial[nb]);
    }
}

__global__ void __launch_bounds__(kB)
k_misfit_rows(const double *part, int splits, int64_t nd, const double *Wd, const double *dobs,
              double *red)
{
    int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    double v2 = 0.0;
    if (i < nd) {
        double y = part[i];
        for (int q = 1; q < splits; ++q) y += part[(int64_t)q * nd + i];
        double v = dgv(Wd, i) * (y - dobs[i]);
        v2 = v * v;
    }
    double s = block_sum(v2);
    if (threadIdx.x == 0) red[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kB) k_misfit_final(const double *red, int64_t nblk, int64_t nd,
                                                     double *out)
{
    double s = sum_partials(red, nblk);
    if (threadIdx.x == 0) *out = sqrt(s / (double)nd);
}

__global__ void k_irls(const int64_t *rp, const int32_t *ci, const double *v, int64_t n,
                       const double *m, const double *mref, double e, double eps2, double *WL)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = 0.0;
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) {
        int64_t j = ci[k];
        x += v[k] * (mref ? m[j] - mref[j] : m[j]);
    }
    WL[i] = pow(x * x + eps2, e);
}

inline unsigned nblocks(int64_t n) { return (unsigned)((n + kB - 1) / kB); }

// Iteration it's p update; from it = 1 on it also ends iteration it-1 from
// the nblk_r partials of its residual, and fills `nfill` words at `fill` with
// the fused product's sentinel.
inline void launch_p_update(const CgDev &d, int64_t nm, int64_t nblk_r, int it, double tol,
                            hipStream_t st, uint64_t *fill = nullptr, int64_t nfill = 0)
{
    const unsigned grid = nblocks(nm);
    if (it == 0)
        hipLaunchKernelGGL(k_p_update<true>, dim3(grid), dim3(kB), 0, st, d, nblk_r, it, tol, fill,
                           nfill);
    else
        hipLaunchKernelGGL(k_p_update<false>, dim3(grid), dim3(kB), 0, st, d, nblk_r, it, tol, fill,
                           nfill);
}

}  // namespace

// ------------------------------------------------------------------ host

gra_status cgls_solve(Ctx &c, const gra_cg_args &a, const double *d_mk, const double *d_mref,
                      const double *d_dobs, const CgDiag &dg, double *d_dx, double *err,
                      int *iterations, bool have_Gmk, bool defer)
{
    hipStream_t st = c.stream;
    const int64_t nd = c.Nd, nm = c.Nm, ld = c.ld;
    // the last gn_step's affine single-read pass (see Ctx::gn_q_ready)
    const int gm_splits = have_Gmk ? c.gn_gm_splits : 0;
    bool have_q = have_Gmk && c.gn_q_ready && c.gn_q && c.gn_q_groups > 0;
    const void *scratch_in = c.fused_scratch.p;
    c.gn_q_ready = false;
    c.pend_affine = false;
    const bool crg = a.use_cross_gradient != 0;
    if (crg && !c.have_crg_ops)
        return set_error(c, GRA_BAD_ARG, "cross-gradient requested but no T operators uploaded");
    const int nseg = crg ? 7 : 4;
    const int64_t nrow = nd + nseg * nm;
    // Mixed precision: the iteration's products use the reduced-precision
    // copy of the FP64 G (gra_mixed_precision); every `refresh` iterations
    // and after the last one the true residual is recomputed with the FP64
    // G, and only those residuals take the stop test.
    const bool mixed = c.mixed_refresh > 0 && c.Glo.p && c.lo_version == c.g_version && !c.wav &&
                       c.store == GRA_F64;
    const int refresh = mixed ? c.mixed_refresh : 0;
    const gra_dtype st_in = mixed ? c.store_lo : c.store;
    const GemvShape shape = gemv_shape(c);                 // FP64 (context) G
    const GemvShape shape_in = mixed ? default_gemv_shape(c.store_lo, nd, nm) : shape;
    const int splits = std::max(shape.n_splits, shape_in.n_splits);
    const int chunks = std::max(shape.t_chunks, shape_in.t_chunks);
    hipError_t e;
    FusedPlan plan;
    bool fused = a.fused && !c.wav && !c.fused_failed &&
                 fused_plan(st_in, nd, ld, c.device, c.fused_wgs, c.fused_depth, c.fused_math,
                            c.fused_min_blocks, &plan);
    const void *G_in = mixed ? c.Glo.p : c.G.p;
    // products of the iteration (reduced-precision copy when mixed)
    auto prod_n = [&](const double *x, const int *dn) {
        return mixed ? launch_gemv_n(G_in, st_in, nd, ld, shape_in, x, c.part_n.as<double>(), dn, st)
                     : launch_G_n(c, shape, x, c.part_n.as<double>(), dn, st);
    };
    auto prod_t = [&](const double *w, const int *dn) {
        return mixed ? launch_gemv_t(G_in, st_in, nd, ld, shape_in, w, c.part_t.as<double>(), dn, st)
                     : launch_G_t(c, shape, w, c.part_t.as<double>(), dn, st);
    };
    int *fused_err = c.status.as<int>() + 1;
    // Mixed precision with the single-read iteration: each FP64 residual
    // refresh is one affine single-read pass over the FP64 G as well (G x
    // and G^T (Wd^2 (G x - dobs)) from one read), on its own plan; the two
    // plans share the scratch, used one after the other on the stream.
    FusedPlan plan64;
    const bool refresh_single = fused && mixed &&
                                fused_plan(c.store, nd, ld, c.device, c.fused_wgs, c.fused_depth,
                                           c.fused_math, c.fused_min_blocks, &plan64);
    if (fused) {
        size_t need = fused_scratch_bytes(plan, ld);
        if (refresh_single) need = std::max(need, fused_scratch_bytes(plan64, ld));
        if ((e = c.fused_scratch.reserve(need)) != hipSuccess)
            return hip_fail(c, e, "alloc fused scratch");
        if (c.fused_scratch.p != scratch_in) have_q = false;   // reallocated: gn_q is gone
        hipLaunchKernelGGL(k_init_count, dim3(1), dim3(1), 0, st, fused_err);
    }
    const bool need_rloop = (fused || mixed) && crg && a.quirk_mode;
    if (need_rloop && (e = c.rloop.reserve(sizeof(double) * (size_t)ld)) != hipSuccess)
        return hip_fail(c, e, "alloc rloop");

    // work vectors
    size_t need_rows = sizeof(double) * (size_t)nrow;
    if ((e = c.vec_row[0].reserve(need_rows)) != hipSuccess) return hip_fail(c, e, "alloc s");
    if ((e = c.vec_row[1].reserve(need_rows)) != hipSuccess) return hip_fail(c, e, "alloc t");
    size_t vnm = sizeof(double) * (size_t)ld;
    for (int q = 0; q < 3; ++q)
        if ((e = c.vec_nm[q].reserve(vnm)) != hipSuccess) return hip_fail(c, e, "alloc vec");
    if ((e = c.vec_nd[0].reserve(sizeof(double) * (size_t)nd)) != hipSuccess)
        return hip_fail(c, e, "alloc w");
    if (mixed) {
        if ((e = c.vec_row[2].reserve(need_rows)) != hipSuccess) return hip_fail(c, e, "alloc b");
        if ((e = c.vec_nm[4].reserve(vnm)) != hipSuccess ||
            (e = c.vec_nm[5].reserve(vnm)) != hipSuccess)
            return hip_fail(c, e, "alloc refresh vectors");
    }
    if (have_Gmk && c.part_n.bytes < sizeof(double) * (size_t)splits * nd) have_Gmk = false;
    if (!have_Gmk) have_q = false;
    if ((e = c.part_n.reserve(sizeof(double) * (size_t)splits * nd)) != hipSuccess)
        return hip_fail(c, e, "alloc part_n");
    c.partn_holds_gn_model = false;
    if ((e = c.part_t.reserve(sizeof(double) * (size_t)chunks * ld)) != hipSuccess)
        return hip_fail(c, e, "alloc part_t");
    int64_t nblk_row = nblocks(nrow), nblk_nm = nblocks(nm);
    const int64_t nblk_r = nblocks(nm * kLanesPerCell);   // k_r_loop blocks
    int64_t nred = nblk_row > nblk_r ? nblk_row : nblk_r;
    if ((e = c.red.reserve(sizeof(double) * 2 * (size_t)nred)) != hipSuccess)
        return hip_fail(c, e, "alloc red");
    if ((e = c.scal.reserve(sizeof(CgScalars))) != hipSuccess) return hip_fail(c, e, "alloc sc");

    // scaled cross-gradient operators
    double *Tv[3] = {nullptr, nullptr, nullptr}, *TvT[3] = {nullptr, nullptr, nullptr};
    DevCsr *Tops[3] = {&c.Tz0, &c.Tx0, &c.Ty0};
    if (crg) {
        int64_t tot = 0;
        for (int q = 0; q < 3; ++q) tot += 2 * Tops[q]->nnz;
        if ((e = c.vec_nm[3].reserve(sizeof(double) * (size_t)(tot > 0 ? tot : 1))) != hipSuccess)
            return hip_fail(c, e, "alloc T");
        double *base = c.vec_nm[3].as<double>();
        for (int q = 0; q < 3; ++q) {
            Tv[q] = base;
            base += Tops[q]->nnz;
            TvT[q] = base;
            base += Tops[q]->nnz;
            unsigned nb = nblocks(nm);
            hipLaunchKernelGGL(k_scale_T, dim3(nb), dim3(kB), 0, st, Tops[q]->rp.as<int64_t>(),
                               Tops[q]->ci.as<int32_t>(), Tops[q]->v.as<double>(), nm, a.a_crg,
                               dg.S, 0, Tv[q]);
            hipLaunchKernelGGL(k_scale_T, dim3(nb), dim3(kB), 0, st, Tops[q]->rpT.as<int64_t>(),
                               Tops[q]->ciT.as<int32_t>(), Tops[q]->vT.as<double>(), nm, a.a_crg,
                               dg.S, 1, TvT[q]);
        }
    }

    CgDev d;
    memset(&d, 0, sizeof(d));
    d.nd = nd; d.nm = nm; d.ld = ld; d.nrow = nrow; d.nseg = nseg;
    d.quirks = a.quirk_mode ? 1 : 0;
    d.sl = std::sqrt(a.lambda);
    d.Wd = dg.Wd; d.P = dg.P; d.dobs = d_dobs; d.mk = d_mk; d.mref = d_mref;
    d.WL[0] = dg.WLs; d.WL[1] = dg.WLz; d.WL[2] = dg.WLx; d.WL[3] = dg.WLy;
    d.WL[4] = dg.WLtz; d.WL[5] = dg.WLtx; d.WL[6] = dg.WLty;
    if (d.quirks) { d.WLt_rows[0] = dg.WLz; d.WLt_rows[1] = dg.WLx; d.WLt_rows[2] = dg.WLy; }
    else { d.WLt_rows[0] = dg.WLtz; d.WLt_rows[1] = dg.WLtx; d.WLt_rows[2] = dg.WLty; }
    DevCsr *W[4] = {&c.Ws, &c.Wz, &c.Wx, &c.Wy};
    for (int q = 0; q < 4; ++q) {
        d.A[q] = Csr{W[q]->rp.as<int64_t>(), W[q]->ci.as<int32_t>(), W[q]->v.as<double>()};
        d.AT[q] = Csr{W[q]->rpT.as<int64_t>(), W[q]->ciT.as<int32_t>(), W[q]->vT.as<double>()};
    }
    if (crg) {
        for (int q = 0; q < 3; ++q) {
            d.A[4 + q] = Csr{Tops[q]->rp.as<int64_t>(), Tops[q]->ci.as<int32_t>(), Tv[q]};
            d.AT[4 + q] = Csr{Tops[q]->rpT.as<int64_t>(), Tops[q]->ciT.as<int32_t>(), TvT[q]};
        }
    }
    d.s = c.vec_row[0].as<double>();
    d.t = c.vec_row[1].as<double>();
    d.r = c.vec_nm[0].as<double>();
    d.p = c.vec_nm[1].as<double>();
    d.Pp = c.vec_nm[2].as<double>();
    d.dx = d_dx;
    d.w = c.vec_nd[0].as<double>();
    d.part_n = c.part_n.as<double>();
    d.splits = shape.n_splits;
    d.part_t = c.part_t.as<double>();
    d.chunks = shape.t_chunks;
    d.red = c.red.as<double>();
    d.red2 = d.red + nred;
    d.sc = c.scal.as<CgScalars>();
    d.rloop = need_rloop ? c.rloop.as<double>() : nullptr;
    if (fused) {
        d.zgroups = plan.S;
        d.nt2 = (int64_t)plan.S * plan.nrb_group;
    }
    if (mixed) {
        d.bvec = c.vec_row[2].as<double>();
        d.vmk = c.vec_nm[4].as<double>();
        d.vpdx = c.vec_nm[5].as<double>();
    }

    // zero p, Pp (incl. padding up to ld) and dx
    hipLaunchKernelGGL(k_fill, dim3(nblocks(ld)), dim3(kB), 0, st, d.p, ld, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(nblocks(ld)), dim3(kB), 0, st, d.Pp, ld, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(nblocks(nm)), dim3(kB), 0, st, d.dx, nm, 0.0);
    hipLaunchKernelGGL(k_init_scalars, dim3(1), dim3(kB), 0, st, d.sc);

    // b, s, first residual and A^T b (G m_k and, after an affine single-read
    // pass, G^T w as well come from the last gn_step)
    CgDev d0 = d;
    if (have_Gmk && gm_splits > 0) d0.splits = gm_splits;
    if (!have_Gmk) {
        TimedRegion tr(c, kTimeGemvN);
        if ((e = launch_G_n(c, shape, d_mk, c.part_n.as<double>(), nullptr, st)))
            return hip_fail(c, e, "gemv_n");
    }
    hipLaunchKernelGGL(k_b_init, dim3(nblocks(nrow)), dim3(kB), 0, st, d0);
    if (have_q) {
        d0.part_t = const_cast<double *>(c.gn_q);
        d0.chunks = c.gn_q_groups;
    } else {
        TimedRegion tr(c, kTimeGemvT);
        if ((e = launch_G_t(c, shape, d.w, c.part_t.as<double>(), nullptr, st)))
            return hip_fail(c, e, "gemv_t");
    }
    hipLaunchKernelGGL(k_r_update<1>, dim3(nblocks(nm)), dim3(kB), 0, st, d0);
    hipLaunchKernelGGL(k_scalar_init, dim3(1), dim3(kB), 0, st, d, nblk_nm);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(c, e, "cg init");

    const int *done = &d.sc->done;
    // the iteration's kernels read the partial sums with its own launch shape
    CgDev di = d;
    di.splits = shape_in.n_splits;
    di.chunks = shape_in.t_chunks;
    if (mixed && !fused) {
        // Two-pass form of the recurrence: between FP64 refreshes the
        // iteration must stay consistent with its own (reduced-precision)
        // operator, so r <- r + alpha A^T t with G^T (Wd t_d) from the copy,
        // as the fused iteration does; recomputing r = A^T s from an s that a
        // refresh made exact mixes the two operators and diverges.
        di.w_from_t = 1;
        di.zpart = c.part_t.as<double>();
        di.zgroups = shape_in.t_chunks;
    }
    int n_refresh = 0;
    // FP64 residual refresh (mixed precision): s = A_t (P dx) - b and
    // r = A_r^T s with the FP64 G, replacing the recurrence's values; the
    // next p update ends the iteration from this r.
    auto refresh_residual = [&]() -> hipError_t {
        hipLaunchKernelGGL(k_refresh_vec, dim3(nblocks(ld)), dim3(kB), 0, st, d);
        if (refresh_single) {
            // y = G (mk + P dx) into part_n (one split) and G^T (Wd^2 (y -
            // dobs)) = G^T w into the group partials, from one read of G
            CgDev dr = d;
            double *zp = nullptr, *t2p = nullptr;
            {
                TimedRegion tr(c, kTimeRefresh);
                hipError_t er = launch_fused(plan64, c.G.p, nd, ld, d.vmk, dg.Wd, nullptr,
                                             c.fused_scratch.p, fused_err, done,
                                             c.fused_fail_block, &zp, &t2p, st, d_dobs,
                                             c.part_n.as<double>());
                if (er != hipSuccess) return er;
            }
            dr.splits = 1;
            hipLaunchKernelGGL(k_refresh_s, dim3(nblocks(nrow)), dim3(kB), 0, st, dr);
            dr.part_t = zp;
            dr.chunks = plan64.S;
            hipLaunchKernelGGL((k_r_loop<false, false>), dim3((unsigned)nblk_r), dim3(kB), 0, st,
                               dr);
            ++n_refresh;
            return hipGetLastError();
        }
        {
            TimedRegion tr(c, kTimeRefresh);
            hipError_t er = launch_G_n(c, shape, d.vmk, c.part_n.as<double>(), done, st);
            if (er != hipSuccess) return er;
        }
        hipLaunchKernelGGL(k_refresh_s, dim3(nblocks(nrow)), dim3(kB), 0, st, d);
        {
            TimedRegion tr(c, kTimeRefresh);
            hipError_t er = launch_G_t(c, shape, d.w, c.part_t.as<double>(), done, st);
            if (er != hipSuccess) return er;
        }
        hipLaunchKernelGGL((k_r_loop<false, false>), dim3((unsigned)nblk_r), dim3(kB), 0, st, d);
        ++n_refresh;
        return hipGetLastError();
    };
    // stop test: every iteration, or (mixed) only on refreshed residuals
    auto tol_at = [&](int it) { return !mixed || (it > 0 && it % refresh == 0) ? a.tol : 0.0; };
    auto refresh_after = [&](int it) { return mixed && ((it + 1) % refresh == 0 || it == a.maxit - 1); };
    if (!fused) {
        // one iteration of the two-pass form (timed launches when `timed`)
        auto two_pass_iteration = [&](int it, bool timed) -> hipError_t {
            hipError_t er;
            launch_p_update(di, nm, nblk_r, it, tol_at(it), st);
            {
                TimedRegion tr(c, timed ? kTimeGemvN : -1, it);
                if ((er = prod_n(di.Pp, done)) != hipSuccess) return er;
            }
            hipLaunchKernelGGL(k_t_rows, dim3(nblocks(nrow)), dim3(kB), 0, st, di);
            hipLaunchKernelGGL(k_s_update<false>, dim3(nblocks(nrow > nm ? nrow : nm)), dim3(kB), 0,
                               st, di, nblk_row, it);
            {
                TimedRegion tr(c, timed ? kTimeGemvT : -1, it);
                if ((er = prod_t(di.w, done)) != hipSuccess) return er;
            }
            if (!mixed)
                hipLaunchKernelGGL((k_r_loop<false, false>), dim3((unsigned)nblk_r), dim3(kB), 0, st,
                                   di);
            else if (it == 0)
                hipLaunchKernelGGL((k_r_loop<true, true>), dim3((unsigned)nblk_r), dim3(kB), 0, st,
                                   di);
            else
                hipLaunchKernelGGL((k_r_loop<true, false>), dim3((unsigned)nblk_r), dim3(kB), 0, st,
                                   di);
            return hipGetLastError();
        };
        // Captured iterations (plain two-pass form): iteration 0 runs as
        // launched, then a graph of cg_graph iterations is replayed; the
        // kernels take the iteration index from device memory, so every
        // replay continues where the previous one stopped.
        const int gk = (!mixed && c.cg_graph > 0) ? c.cg_graph : 0;
        hipGraph_t graph = nullptr;
        hipGraphExec_t gexec = nullptr;
        int it0 = 0;
        if (gk > 0 && a.maxit > gk) {
            di.dev_it = 1;
            {
                TimedRegion tr_it(c, kTimeCgIter, 0);
                if ((e = two_pass_iteration(0, true)) != hipSuccess) return hip_fail(c, e, "cg iteration");
            }
            it0 = 1;
            const int reps = (a.maxit - 1) / gk;
            bool captured = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess;
            if (captured) {
                hipError_t ec = hipSuccess;
                for (int j = 0; j < gk && ec == hipSuccess; ++j) ec = two_pass_iteration(1 + j, false);
                hipError_t ee = hipStreamEndCapture(st, &graph);
                captured = ec == hipSuccess && ee == hipSuccess && graph &&
                           hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0) == hipSuccess;
            }
            if (captured) {
                for (int r = 0; r < reps; ++r) {
                    TimedRegion tr(c, kTimeCgGraph);
                    if ((e = hipGraphLaunch(gexec, st)) != hipSuccess) break;
                }
                it0 = 1 + reps * gk;
            }
            (void)hipGetLastError();
        }
        for (int it = it0; it < a.maxit && e == hipSuccess; ++it) {
            TimedRegion tr_it(c, kTimeCgIter, it);
            if ((e = two_pass_iteration(it, true)) != hipSuccess) break;
            if (refresh_after(it) && (e = refresh_residual()) != hipSuccess) break;
        }
        if (gexec || graph) {
            // the replays must finish before the graph is destroyed
            hipError_t es = hipStreamSynchronize(st);
            if (gexec) (void)hipGraphExecDestroy(gexec);
            if (graph) (void)hipGraphDestroy(graph);
            if (es != hipSuccess && e == hipSuccess) e = es;
        }
        if (e != hipSuccess) return hip_fail(c, e, "cg iteration");
    } else {
        const int64_t nreg = nrow - nd;
        const int64_t nblk_reg = nblocks(nreg);
        for (int it = 0; it < a.maxit; ++it) {
            TimedRegion tr_it(c, kTimeCgIter, it);
            launch_p_update(di, nm, nblk_r, it, tol_at(it), st,
                            static_cast<uint64_t *>(c.fused_scratch.p), fused_partials(plan));
            {
                TimedRegion tr(c, kTimeFused, it);
                double *zp = nullptr, *t2p = nullptr;
                e = launch_fused(plan, G_in, nd, ld, di.Pp, dg.Wd, di.t, c.fused_scratch.p,
                                 fused_err, done, c.fused_fail_block, &zp, &t2p, st, nullptr,
                                 nullptr, true);
                if (e != hipSuccess) return hip_fail(c, e, "fused product");
                di.zpart = zp;
                di.t2part = t2p;
            }
            hipLaunchKernelGGL(k_t_rows_reg, dim3((unsigned)nblk_reg), dim3(kB), 0, st, di);
            hipLaunchKernelGGL(k_s_update<true>, dim3(nblocks(nrow > nm ? nrow : nm)), dim3(kB), 0,
                               st, di, nblk_reg, it);
            if (it == 0)
                hipLaunchKernelGGL((k_r_loop<true, true>), dim3((unsigned)nblk_r), dim3(kB), 0, st,
                                   di);
            else
                hipLaunchKernelGGL((k_r_loop<true, false>), dim3((unsigned)nblk_r), dim3(kB), 0,
                                   st, di);
            if ((e = hipGetLastError()) != hipSuccess) return hip_fail(c, e, "cg iteration");
            if (refresh_after(it) && (e = refresh_residual()) != hipSuccess)
                return hip_fail(c, e, "residual refresh");
        }
    }
    (void)n_refresh;
    if (a.maxit > 0) {
        hipLaunchKernelGGL(k_scalar_end, dim3(1), dim3(kB), 0, st, d, nblk_r, a.maxit - 1, a.tol);
        if ((e = hipGetLastError()) != hipSuccess) return hip_fail(c, e, "cg end");
    }

    // scalars (and the fused product's status) to pinned host staging
    HostStage *hst = c.host_stage;
    if ((e = hipMemcpyAsync(&hst->sc, d.sc, sizeof(CgScalars), hipMemcpyDeviceToHost, st)) !=
        hipSuccess)
        return hip_fail(c, e, "copy scalars");
    hst->ferr = 0;
    if (fused && (e = hipMemcpyAsync(&hst->ferr, fused_err, sizeof(int), hipMemcpyDeviceToHost,
                                     st)) != hipSuccess)
        return hip_fail(c, e, "copy status");
    c.pend_fused = fused;
    c.pend_maxit = a.maxit;
    c.pend_refresh = refresh;
    if (defer) return GRA_OK;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(c, e, "cg sync");
    bool rerun = false;
    gra_status s = cgls_collect(c, err, iterations, &rerun);
    if (rerun) {
        // A wait of the fused product failed (bounded wait or the test hook):
        // its result is discarded and the same solve reruns with the two-pass
        // iteration from the unchanged inputs; the context stays on the
        // two-pass iteration from now on.
        gra_cg_args b = a;
        b.fused = 0;
        return cgls_solve(c, b, d_mk, d_mref, d_dobs, dg, d_dx, err, iterations, false, false);
    }
    return s;
}

gra_status cgls_collect(Ctx &c, double *err, int *iterations, bool *rerun)
{
    const HostStage *hst = c.host_stage;
    const CgScalars &hs = hst->sc;
    c.last_solve_fused = c.pend_fused;
    *rerun = false;
    if ((c.pend_fused || c.pend_affine) && hst->ferr) {
        c.fused_failed = true;
        c.fused_fail_count += 1;
        *rerun = true;
        return GRA_OK;
    }
    const int maxit = c.pend_maxit, refresh = c.pend_refresh;
    if (maxit > 0 && err) *err = hs.err;
    const int its_done = maxit > 0 ? (hs.done ? hs.iterations : maxit) : 0;
    if (iterations) *iterations = its_done;
    // refreshes that ran: after every refresh-th iteration and the last one;
    // a converged solve stops right after one
    c.last_refreshes = refresh > 0 ? (hs.done ? its_done / refresh : (its_done + refresh - 1) / refresh) : 0;
    if (hs.breakdown)
        return set_error(c, GRA_NUMERIC_BREAKDOWN, "CG breakdown: t2 == 0 or non-finite alpha");
    return GRA_OK;
}

hipError_t launch_barrier_P(const double *m, const double *mmin, const double *mmax, int64_t n,
                            double *P, hipStream_t st)
{
    hipLaunchKernelGGL(k_barrier_P, dim3(nblocks(n)), dim3(kB), 0, st, m, mmin, mmax, n, P);
    return hipGetLastError();
}

hipError_t launch_barrier_update(const double *m, const double *dx, const double *mmin,
                                 const double *mmax, int64_t n, double *m_trial, double *scratch,
                                 int *flags, hipStream_t st)
{
    // flags[0] is the fallback counter, flags[1..n] the per-cell codes.
    int *count = flags;
    hipLaunchKernelGGL(k_init_count, dim3(1), dim3(1), 0, st, count);
    hipLaunchKernelGGL(k_barrier_main, dim3(nblocks(n)), dim3(kB), 0, st, m, dx, mmin, mmax, n,
                       scratch, flags + 1, count);
    hipLaunchKernelGGL(k_barrier_copy, dim3(nblocks(n)), dim3(kB), 0, st, scratch, n, m_trial,
                       count);
    hipLaunchKernelGGL(k_barrier_seq, dim3(1), dim3(64), 0, st, scratch, mmin, mmax, n, flags + 1,
                       m_trial, count);
    return hipGetLastError();
}

hipError_t launch_misfit(const double *part, int splits, int64_t Nd, const double *Wd,
                         const double *dobs, double *red, double *out, hipStream_t st)
{
    int64_t nb = nblocks(Nd);
    hipLaunchKernelGGL(k_misfit_rows, dim3((unsigned)nb), dim3(kB), 0, st, part, splits, Nd, Wd,
                       dobs, red);
    hipLaunchKernelGGL(k_misfit_final, dim3(1), dim3(kB), 0, st, red, nb, Nd, out);
    return hipGetLastError();
}

hipError_t launch_irls(const int64_t *rp, const int32_t *ci, const double *v, int64_t n,
                       const double *m, const double *mref, int p, double eps2, double *WL,
                       hipStream_t st)
{
    double ex = -(2 - p) / 4.0;
    hipLaunchKernelGGL(k_irls, dim3(nblocks(n)), dim3(kB), 0, st, rp, ci, v, n, m, mref, ex, eps2,
                       WL);
    return hipGetLastError();
}

}  // namespace gra
