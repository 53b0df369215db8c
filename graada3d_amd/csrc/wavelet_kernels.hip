// Wavelet-compressed sensitivity matrix (the reference's method 2,
// Fwd::compute_G_wavelet / compress_vec / wavelet branches of G_vec_mul and
// GT_vec_mul, Fwd.cpp:80-219, :298-417).
//
// Every row of G, taken over the cells in the mesh's reordered order and
// zero-padded to n_dy = 2^J >= Nm, is replaced by its periodic Daubechies-4
// wavelet coefficients (Mallat pyramid, non-centred, the algorithm of GSL's
// gsl_wavelet_daubechies k = 4), and only the coefficients with
// |c| >= rel * max|c| of the row (within 1e-15) are kept.  The products are
//   G x   = C (W x_reordered)          forward transform, then CSR rows
//   G^T r = W^T (C^T r), back to cell order
// with C stored twice: row-major (CSR: row start / length, column, value)
// and column-major (CSC from a stable radix sort of the columns, so each
// column's entries keep the row order and every product is deterministic).
//
// Assembly runs in chunks of points: the dense rows of a chunk are assembled
// by the prism kernel (cells in reordered order, leading dimension n_dy),
// transformed in place (one workgroup per row; levels longer than the LDS
// window go through HBM), thresholded and compacted (block scan per row).
// Transforms of vectors inside the products use one workgroup with the whole
// vector in LDS when n_dy fits, with the cell permutation fused into the
// load (forward) or store (inverse).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "engine.hpp"

namespace gra {

namespace {

constexpr int kWave = 64;
constexpr int kBlk = 256;
constexpr int kLdsThreads = 1024;
constexpr int64_t kLdsMaxN = 16384;                     // LDS window cap (128 KiB)
constexpr int kFwdPer = (int)(kLdsMaxN / 2 / kLdsThreads);   // (s, d) pairs per thread

// Daubechies (1988) 4-tap orthonormal filters: h = ((1+sqrt3), (3+sqrt3),
// (3-sqrt3), (1-sqrt3)) / (4 sqrt2); g[k] = (-1)^k h[3-k].
__device__ constexpr double kH0 = 0.48296291314453414337487159986;
__device__ constexpr double kH1 = 0.83651630373780790557529378092;
__device__ constexpr double kH2 = 0.22414386804201338102597276224;
__device__ constexpr double kH3 = -0.12940952255126038117444941881;

__device__ inline void analysis(double v0, double v1, double v2, double v3, double &s, double &d)
{
    s = kH0 * v0 + kH1 * v1 + kH2 * v2 + kH3 * v3;
    d = kH3 * v0 - kH2 * v1 + kH1 * v2 - kH0 * v3;
}

// Synthesis output j of a level of length n from s = a[0..n/2), d = a[n/2..n):
// the two (ii, k) pairs with (2 ii + k) mod n = j, summed in ascending ii.
template <typename F> __device__ inline double synthesis(int64_t j, int64_t n, F at)
{
    const int64_t nh = n >> 1, mask = n - 1;
    const int k0 = (int)(j & 1);
    const int64_t ia = ((j - k0) & mask) >> 1;          // k = k0
    const int64_t ib = ((j - k0 - 2) & mask) >> 1;      // k = k0 + 2
    // h[k], g[k] for k = k0 and k0 + 2
    const double ha = k0 ? kH1 : kH0, ga = k0 ? -kH2 : kH3;
    const double hb = k0 ? kH3 : kH2, gb = k0 ? -kH0 : kH1;
    const double ta = ha * at(ia) + ga * at(ia + nh);
    const double tb = hb * at(ib) + gb * at(ib + nh);
    return ib < ia ? tb + ta : ta + tb;   // n = 2: ia = ib, k0 before k0 + 2
}

__device__ inline bool keep(double c, double thr)
{
    const double a = fabs(c);
    return a > thr || fabs(a - thr) < 1e-15;
}

__device__ inline bool finished(const int *done) { return done && *(volatile const int *)done; }

// ---------------------------------------------------------------- transforms

// Levels n0, n0/2, ..., 2 (FWD) or 2, ..., n0 (inverse) of the leading n0
// entries of each row (blockIdx.x = row), in LDS.  Forward with `order`:
// the input is the cell vector x, loaded as x[order[j]] (0 beyond Nm).
// Inverse with `order`: the result is stored as out[order[j]] (j < Nm) and
// out[Nm..out_len) = 0.  Otherwise rows are read from / written to
// in / out + row * ld.
// `in` and `out` may be the same rows (every read precedes every write).
template <bool FWD>
__global__ void __launch_bounds__(kLdsThreads)
dwt_lds_kernel(const double *in, double *out, int64_t ld, int64_t n0,
               const int32_t *__restrict__ order, int64_t Nm, int64_t out_len, const int *done)
{
    extern __shared__ double lds[];
    if (finished(done)) return;
    const int tid = threadIdx.x;
    const int64_t row = blockIdx.x;
    if (FWD && order) {
        for (int64_t j = tid; j < n0; j += kLdsThreads) lds[j] = j < Nm ? in[order[j]] : 0.0;
    } else {
        const double *a = in + row * ld;
        for (int64_t j = tid; j < n0; j += kLdsThreads) lds[j] = a[j];
    }
    __syncthreads();
    // one thread per (s, d) pair: 16-byte LDS reads of the even-aligned
    // sample pairs (2ii, 2ii+1) and (2ii+2, 2ii+3) forward; the inverse
    // produces outputs 2t and 2t+1 from s[t], d[t], s[t-1], d[t-1]
    typedef double v2 __attribute__((ext_vector_type(2)));
    v2 *lds2 = reinterpret_cast<v2 *>(lds);
    if (FWD) {
        for (int64_t n = n0; n >= 2; n >>= 1) {
            const int64_t nh = n >> 1;
            double s[kFwdPer], d[kFwdPer];
#pragma unroll
            for (int q = 0; q < kFwdPer; ++q) {
                const int64_t ii = tid + (int64_t)q * kLdsThreads;
                if (ii < nh) {
                    const v2 a = lds2[ii], b = lds2[(ii + 1) & (nh - 1)];
                    analysis(a.x, a.y, b.x, b.y, s[q], d[q]);
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < kFwdPer; ++q) {
                const int64_t ii = tid + (int64_t)q * kLdsThreads;
                if (ii < nh) {
                    lds[ii] = s[q];
                    lds[ii + nh] = d[q];
                }
            }
            __syncthreads();
        }
    } else {
        for (int64_t n = 2; n <= n0; n <<= 1) {
            const int64_t nh = n >> 1;
            v2 o[kFwdPer];
#pragma unroll
            for (int q = 0; q < kFwdPer; ++q) {
                const int64_t t = tid + (int64_t)q * kLdsThreads;
                if (t < nh) {
                    const int64_t tb = (t - 1) & (nh - 1);
                    const double sa = lds[t], da = lds[t + nh], sb = lds[tb], db = lds[tb + nh];
                    // output 2t: k = 0 at t, k = 2 at t-1; output 2t+1: k = 1, 3
                    const double ta0 = kH0 * sa + kH3 * da, tb0 = kH2 * sb + kH1 * db;
                    const double ta1 = kH1 * sa + (-kH2) * da, tb1 = kH3 * sb + (-kH0) * db;
                    const bool b_first = tb < t;
                    o[q].x = b_first ? tb0 + ta0 : ta0 + tb0;
                    o[q].y = b_first ? tb1 + ta1 : ta1 + tb1;
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < kFwdPer; ++q) {
                const int64_t t = tid + (int64_t)q * kLdsThreads;
                if (t < nh) lds2[t] = o[q];
            }
            __syncthreads();
        }
    }
    if (!FWD && order) {
        for (int64_t j = tid; j < n0 && j < Nm; j += kLdsThreads) out[order[j]] = lds[j];
        for (int64_t j = Nm + tid; j < out_len; j += kLdsThreads) out[j] = 0.0;
    } else {
        double *o = out + row * ld;
        for (int64_t j = tid; j < n0; j += kLdsThreads) o[j] = lds[j];
    }
}

// One analysis level of length n on every row (grid.y = row): in -> tmp.
__global__ void __launch_bounds__(kBlk)
dwt_level_fwd_kernel(const double *__restrict__ in, double *__restrict__ tmp, int64_t ld, int64_t n,
                     const int *done)
{
    if (finished(done)) return;
    const int64_t nh = n >> 1, mask = n - 1;
    const int64_t ii = (int64_t)blockIdx.x * kBlk + threadIdx.x;
    if (ii >= nh) return;
    const double *a = in + (int64_t)blockIdx.y * ld;
    double *o = tmp + (int64_t)blockIdx.y * ld;
    const int64_t i = 2 * ii;
    double s, d;
    analysis(a[i], a[(i + 1) & mask], a[(i + 2) & mask], a[(i + 3) & mask], s, d);
    o[ii] = s;
    o[ii + nh] = d;
}

// One synthesis level of length n on every row: in -> tmp.
__global__ void __launch_bounds__(kBlk)
dwt_level_inv_kernel(const double *__restrict__ in, double *__restrict__ tmp, int64_t ld, int64_t n,
                     const int *done)
{
    if (finished(done)) return;
    const int64_t j = (int64_t)blockIdx.x * kBlk + threadIdx.x;
    if (j >= n) return;
    const double *a = in + (int64_t)blockIdx.y * ld;
    tmp[(int64_t)blockIdx.y * ld + j] = synthesis(j, n, [&](int64_t p) { return a[p]; });
}

__global__ void __launch_bounds__(kBlk)
copy_rows_kernel(const double *__restrict__ in, double *__restrict__ out, int64_t ld, int64_t n,
                 const int *done)
{
    if (finished(done)) return;
    const int64_t j = (int64_t)blockIdx.x * kBlk + threadIdx.x;
    if (j >= n) return;
    out[(int64_t)blockIdx.y * ld + j] = in[(int64_t)blockIdx.y * ld + j];
}

__global__ void __launch_bounds__(kBlk)
permute_pad_kernel(const double *__restrict__ x, const int32_t *__restrict__ order, int64_t Nm,
                   int64_t n, double *__restrict__ xw, const int *done)
{
    if (finished(done)) return;
    const int64_t j = (int64_t)blockIdx.x * kBlk + threadIdx.x;
    if (j < n) xw[j] = j < Nm ? x[order[j]] : 0.0;
}

__global__ void __launch_bounds__(kBlk)
unpermute_kernel(const double *__restrict__ zw, const int32_t *__restrict__ order, int64_t Nm,
                 int64_t out_len, double *__restrict__ out, const int *done)
{
    if (finished(done)) return;
    const int64_t j = (int64_t)blockIdx.x * kBlk + threadIdx.x;
    if (j < Nm) out[order[j]] = zw[j];
    else if (j < out_len) out[j] = 0.0;
}

// ----------------------------------------------------- threshold + compaction

__device__ inline double block_max(double v, double *sh)
{
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    if (l == 0) sh[w] = v;
    __syncthreads();
    double m = sh[0];
    for (int q = 1; q < kBlk / kWave; ++q) m = fmax(m, sh[q]);
    __syncthreads();
    return m;
}

// Per chunk row: thr = rel * max|a| and the number of kept coefficients.
__global__ void __launch_bounds__(kBlk)
row_stats_kernel(const double *__restrict__ A, int64_t ld, int64_t n, double rel,
                 double *__restrict__ thr, int64_t *__restrict__ cnt)
{
    __shared__ double sh[kBlk / kWave];
    __shared__ int64_t shc[kBlk / kWave];
    const double *a = A + (int64_t)blockIdx.x * ld;
    double m = 0.0;
    for (int64_t j = threadIdx.x; j < n; j += kBlk) m = fmax(m, fabs(a[j]));
    const double t = rel * block_max(m, sh);
    int64_t c = 0;
    for (int64_t j = threadIdx.x; j < n; j += kBlk) c += keep(a[j], t) ? 1 : 0;
    for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off, kWave);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    if (l == 0) shc[w] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t s = 0;
        for (int q = 0; q < kBlk / kWave; ++q) s += shc[q];
        cnt[blockIdx.x] = s;
        thr[blockIdx.x] = t;
    }
}

// Per chunk row: the kept coefficients in ascending index at base[row]
// (column, value, global row id).  Chunk row k*np + i is global row
// k*N_obs + p0 + i.
__global__ void __launch_bounds__(kBlk)
compact_kernel(const double *__restrict__ A, int64_t ld, int64_t n, const double *__restrict__ thr,
               const int64_t *__restrict__ base, int64_t np, int64_t p0, int64_t N_obs,
               int32_t *__restrict__ ci, double *__restrict__ v, int32_t *__restrict__ rowid)
{
    __shared__ int wsum[kBlk / kWave];
    const int64_t r = blockIdx.x;
    const double *a = A + r * ld;
    const double t = thr[r];
    const int32_t grow = (int32_t)((r / np) * N_obs + p0 + (r % np));
    int64_t pos = base[r];
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    for (int64_t j0 = 0; j0 < n; j0 += kBlk) {
        const int64_t j = j0 + threadIdx.x;
        const double c = j < n ? a[j] : 0.0;
        const bool k = j < n && keep(c, t);
        const uint64_t mask = __ballot(k);
        const int before = __popcll(mask & ((l ? (~0ull >> (64 - l)) : 0ull)));
        if (l == 0) wsum[w] = __popcll(mask);
        __syncthreads();
        int off = 0, tot = 0;
        for (int q = 0; q < kBlk / kWave; ++q) {
            if (q < w) off += wsum[q];
            tot += wsum[q];
        }
        if (k) {
            const int64_t o = pos + off + before;
            ci[o] = (int32_t)j;
            v[o] = c;
            rowid[o] = grow;
        }
        pos += tot;
        __syncthreads();
    }
}

// ------------------------------------------------------------------- CSC

__global__ void __launch_bounds__(kBlk)
iota_kernel(int32_t *__restrict__ p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kBlk + threadIdx.x;
    if (i < n) p[i] = (int32_t)i;
}

// cp[c] = first position of column c in the sorted keys (cp[ncols] = nnz).
__global__ void __launch_bounds__(kBlk)
col_ptr_kernel(const int32_t *__restrict__ keys, int64_t nnz, int64_t ncols, int64_t *__restrict__ cp)
{
    const int64_t c = (int64_t)blockIdx.x * kBlk + threadIdx.x;
    if (c > ncols) return;
    int64_t lo = 0, hi = nnz;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (keys[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    cp[c] = lo;
}

__global__ void __launch_bounds__(kBlk)
gather_csc_kernel(const int32_t *__restrict__ perm, const int32_t *__restrict__ rowid,
                  const double *__restrict__ v, int64_t nnz, int32_t *__restrict__ ri,
                  double *__restrict__ vt)
{
    const int64_t q = (int64_t)blockIdx.x * kBlk + threadIdx.x;
    if (q >= nnz) return;
    const int32_t e = perm[q];
    ri[q] = rowid[e];
    vt[q] = v[e];
}

// ------------------------------------------------------------------- SpMV

// Sparse dot products of one wave with a gathered vector: sum over k in
// [b, e) of val[k] * x[idx[k]].  Each lane keeps kSpU entries in flight (the
// index loads, then the value loads and gathers, are issued together); the
// kSpU partial sums are added in a fixed order, then the wave butterfly, so
// the result is repeatable.  Indices past e are clamped (no conditional
// loads) and their products dropped.
constexpr int kSpU = 8;
constexpr int64_t kSegLen = 2048;   // entries per column segment of the CSC product

template <typename X>
__device__ inline double wave_sparse_dot(const int32_t *__restrict__ idx,
                                         const double *__restrict__ val, X x, int64_t b,
                                         int64_t e, int l)
{
    double acc[kSpU];
#pragma unroll
    for (int u = 0; u < kSpU; ++u) acc[u] = 0.0;
    const int64_t last = e - 1;
    for (int64_t k0 = b + l; k0 < e; k0 += (int64_t)kWave * kSpU) {
        int32_t c[kSpU];
        double vv[kSpU], g[kSpU];
#pragma unroll
        for (int u = 0; u < kSpU; ++u) {
            const int64_t k = k0 + (int64_t)kWave * u;
            c[u] = __builtin_nontemporal_load(idx + (k < e ? k : last));
        }
#pragma unroll
        for (int u = 0; u < kSpU; ++u) {
            const int64_t k = k0 + (int64_t)kWave * u;
            vv[u] = __builtin_nontemporal_load(val + (k < e ? k : last));
            g[u] = x(c[u]);
        }
#pragma unroll
        for (int u = 0; u < kSpU; ++u) {
            const int64_t k = k0 + (int64_t)kWave * u;
            if (k < e) acc[u] += vv[u] * g[u];
        }
    }
    double s = acc[0];
#pragma unroll
    for (int u = 1; u < kSpU; ++u) s += acc[u];
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, kWave);
    return s;
}

// The same products with the gathered vector (n entries) staged in LDS: one
// workgroup of kSpWaves waves per `per_block` consecutive outputs; each
// wave takes every kSpWaves-th of them.  Output r sums entries
// [start[r], start[r] + len[r]).  Used when the vector fits the LDS window:
// the gathers then hit LDS instead of scattered L2 lines.
constexpr int kSpWaves = 16;

__global__ void __launch_bounds__(kSpWaves * kWave)
spmv_lds_kernel(const int64_t *__restrict__ start, const int32_t *__restrict__ len,
                const int32_t *__restrict__ idx, const double *__restrict__ val,
                const double *__restrict__ x, int64_t n, int64_t nout, int64_t per_block,
                double *__restrict__ y, const int *done)
{
    extern __shared__ double xs[];
    if (finished(done)) return;
    for (int64_t j = threadIdx.x; j < n; j += kSpWaves * kWave) xs[j] = x[j];
    __syncthreads();
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    const int64_t r0 = (int64_t)blockIdx.x * per_block;
    const int64_t r1 = r0 + per_block < nout ? r0 + per_block : nout;
    auto gx = [&](int32_t c) { return xs[c]; };
    for (int64_t r = r0 + w; r < r1; r += kSpWaves) {
        const int64_t b = start[r];
        const double s = wave_sparse_dot(idx, val, gx, b, b + len[r], l);
        if (l == 0) y[r] = s;
    }
}

// Output r = sum of entries [start[r], start[r] + len[r]) with global
// gathers (one wave per output).
__global__ void __launch_bounds__(kBlk)
spmv_rows_kernel(const int64_t *__restrict__ start, const int32_t *__restrict__ len,
                 const int32_t *__restrict__ idx, const double *__restrict__ val,
                 const double *__restrict__ x, int64_t nout, double *__restrict__ y,
                 const int *done)
{
    if (finished(done)) return;
    const int64_t r = (int64_t)blockIdx.x * (kBlk / kWave) + threadIdx.x / kWave;
    const int l = threadIdx.x % kWave;
    if (r >= nout) return;
    const int64_t b = start[r];
    const double s = wave_sparse_dot(idx, val, [&](int32_t c) { return x[c]; }, b, b + len[r], l);
    if (l == 0) y[r] = s;
}

// z[c] = the column's segment sums in order (columns are split into
// segments of at most kSegLen entries so that the long coarse-scale
// columns spread over many waves).
__global__ void __launch_bounds__(kBlk)
segment_sum_kernel(const int64_t *__restrict__ col_seg, const double *__restrict__ part,
                   int64_t ncols, double *__restrict__ z, const int *done)
{
    if (finished(done)) return;
    const int64_t c = (int64_t)blockIdx.x * kBlk + threadIdx.x;
    if (c >= ncols) return;
    double s = 0.0;
    for (int64_t q = col_seg[c]; q < col_seg[c + 1]; ++q) s += part[q];
    z[c] = s;
}

unsigned blocks_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

// Forward transform of `rows` rows of length n (leading dimension ld) in A,
// with B as the level scratch.
hipError_t dwt_rows_forward(double *A, double *B, int64_t rows, int64_t ld, int64_t n,
                            int64_t nlds, const int *done, hipStream_t st)
{
    int64_t m = n;
    for (; m > nlds; m >>= 1) {
        dim3 g1(blocks_for(m / 2, kBlk), (unsigned)rows), g2(blocks_for(m, kBlk), (unsigned)rows);
        hipLaunchKernelGGL(dwt_level_fwd_kernel, g1, dim3(kBlk), 0, st, A, B, ld, m, done);
        hipLaunchKernelGGL(copy_rows_kernel, g2, dim3(kBlk), 0, st, B, A, ld, m, done);
    }
    if (m >= 2)
        hipLaunchKernelGGL((dwt_lds_kernel<true>), dim3((unsigned)rows), dim3(kLdsThreads),
                           (size_t)m * sizeof(double), st, A, A, ld, m, nullptr, (int64_t)0,
                           (int64_t)0, done);
    return hipGetLastError();
}

}  // namespace

// ------------------------------------------------------------------- host

int64_t wavelet_length(int64_t n)
{
    int64_t p = 1;
    while (p < n) p <<= 1;
    return p;
}

// Largest LDS transform window the device allows (power of two, <= 16384).
static int64_t lds_window(int device)
{
    int bytes = 0;
    if (hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, device) !=
        hipSuccess)
        bytes = 65536;
    int64_t n = 2;
    while (n * 2 <= kLdsMaxN && (n * 2) * (int64_t)sizeof(double) <= bytes) n *= 2;
    if (n * (int64_t)sizeof(double) > 65536) {
        const int b = (int)(n * sizeof(double));
        if (hipFuncSetAttribute((const void *)dwt_lds_kernel<true>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, b) != hipSuccess ||
            hipFuncSetAttribute((const void *)dwt_lds_kernel<false>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, b) != hipSuccess ||
            hipFuncSetAttribute((const void *)spmv_lds_kernel,
                                hipFuncAttributeMaxDynamicSharedMemorySize, b) != hipSuccess) {
            (void)hipGetLastError();
            n = 65536 / sizeof(double);
        }
    }
    return n;
}

// Copies the first `keep` bytes of b into a buffer of at least `n` bytes.
static hipError_t grow_keep(DevBuf &b, size_t n, size_t keep, hipStream_t st)
{
    if (n <= b.bytes && b.p) return hipSuccess;
    size_t want = std::max(n, b.bytes + b.bytes / 2);
    DevBuf nb;
    hipError_t e = nb.reserve(want);
    if (e != hipSuccess && want > n) e = nb.reserve(n);
    if (e != hipSuccess) return e;
    if (keep && (e = hipMemcpyAsync(nb.p, b.p, keep, hipMemcpyDeviceToDevice, st)) != hipSuccess)
        return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    std::swap(b.p, nb.p);
    std::swap(b.bytes, nb.bytes);
    return hipSuccess;
}

gra_status wavelet_assemble(Ctx &c, const double *d_prisms_ro, int64_t Nm, const double *d_pts,
                            int64_t N_obs, uint32_t flag, double rel, int *d_bad)
{
    hipStream_t st = c.stream;
    const int nc = __builtin_popcount(flag & 0x3FEu);
    const int64_t Nd = (int64_t)nc * N_obs;
    const int64_t n = wavelet_length(Nm);
    if (n > INT32_MAX) return set_error(c, GRA_BAD_ARG, "wavelet length exceeds int32 columns");
    const int64_t nlds = lds_window(c.device);
    // chunk of points: about 256 MiB of dense rows per buffer
    int64_t np = (int64_t)((256ull << 20) / ((uint64_t)nc * n * sizeof(double)));
    np = std::max<int64_t>(1, std::min({np, N_obs, (int64_t)(65535 / nc)}));   // grid.y of levels
    const int64_t rows = (int64_t)nc * np;
    DevBuf A, B, thr, cnt, base;
    hipError_t e;
    if ((e = A.reserve(sizeof(double) * (size_t)rows * n)) != hipSuccess ||
        (e = B.reserve(sizeof(double) * (size_t)rows * n)) != hipSuccess ||
        (e = thr.reserve(sizeof(double) * (size_t)rows)) != hipSuccess ||
        (e = cnt.reserve(sizeof(int64_t) * (size_t)rows)) != hipSuccess ||
        (e = base.reserve(sizeof(int64_t) * (size_t)rows)) != hipSuccess)
        return hip_fail(c, e, "allocate wavelet chunk");
    DevBuf rowid;
    std::vector<int64_t> h_rs(Nd), h_cnt(rows), h_base(rows);
    std::vector<int32_t> h_rl(Nd);
    int64_t nnz = 0;
    for (int64_t p0 = 0; p0 < N_obs; p0 += np) {
        const int64_t pn = std::min(np, N_obs - p0);
        const int64_t rr = (int64_t)nc * pn;
        if ((e = hipMemsetAsync(A.p, 0, sizeof(double) * (size_t)rr * n, st)) != hipSuccess)
            return hip_fail(c, e, "clear chunk");
        {
            TimedRegion tr(c, kTimeAssemble);
            if ((e = launch_assemble(d_prisms_ro, Nm, d_pts + 3 * p0, pn, flag, GRA_F64, A.p, n,
                                     d_bad, st)) != hipSuccess)
                return hip_fail(c, e, "assemble chunk");
            if ((e = dwt_rows_forward(A.as<double>(), B.as<double>(), rr, n, n, nlds, nullptr,
                                      st)) != hipSuccess)
                return hip_fail(c, e, "row transforms");
            hipLaunchKernelGGL(row_stats_kernel, dim3((unsigned)rr), dim3(kBlk), 0, st,
                               A.as<double>(), n, n, rel, thr.as<double>(), cnt.as<int64_t>());
        }
        if ((e = hipMemcpyAsync(h_cnt.data(), cnt.p, sizeof(int64_t) * (size_t)rr,
                                hipMemcpyDeviceToHost, st)) != hipSuccess ||
            (e = hipStreamSynchronize(st)) != hipSuccess)
            return hip_fail(c, e, "row counts");
        int64_t tot = 0;
        for (int64_t r = 0; r < rr; ++r) {
            h_base[r] = nnz + tot;
            const int64_t g = (r / pn) * N_obs + p0 + (r % pn);
            h_rs[g] = nnz + tot;
            h_rl[g] = (int32_t)h_cnt[r];
            tot += h_cnt[r];
        }
        if (nnz + tot > INT32_MAX)
            return set_error(c, GRA_BAD_ARG,
                             "wavelet-compressed G exceeds 2^31 coefficients; raise the threshold");
        if ((e = grow_keep(c.wav_ci, sizeof(int32_t) * (size_t)(nnz + tot),
                           sizeof(int32_t) * (size_t)nnz, st)) != hipSuccess ||
            (e = grow_keep(c.wav_v, sizeof(double) * (size_t)(nnz + tot),
                           sizeof(double) * (size_t)nnz, st)) != hipSuccess ||
            (e = grow_keep(rowid, sizeof(int32_t) * (size_t)(nnz + tot),
                           sizeof(int32_t) * (size_t)nnz, st)) != hipSuccess)
            return hip_fail(c, e, "grow compressed rows");
        if ((e = hipMemcpyAsync(base.p, h_base.data(), sizeof(int64_t) * (size_t)rr,
                                hipMemcpyHostToDevice, st)) != hipSuccess)
            return hip_fail(c, e, "row offsets");
        {
            TimedRegion tr(c, kTimeAssemble);
            hipLaunchKernelGGL(compact_kernel, dim3((unsigned)rr), dim3(kBlk), 0, st,
                               A.as<double>(), n, n, thr.as<double>(), base.as<int64_t>(), pn, p0,
                               N_obs, c.wav_ci.as<int32_t>(), c.wav_v.as<double>(),
                               rowid.as<int32_t>());
        }
        if ((e = hipGetLastError()) != hipSuccess) return hip_fail(c, e, "compaction");
        // base is reused by the next chunk: wait before overwriting it
        if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(c, e, "chunk");
        nnz += tot;
    }
    A.release();
    B.release();
    if ((e = c.wav_rs.reserve(sizeof(int64_t) * (size_t)Nd)) != hipSuccess ||
        (e = c.wav_rl.reserve(sizeof(int32_t) * (size_t)Nd)) != hipSuccess ||
        (e = hipMemcpyAsync(c.wav_rs.p, h_rs.data(), sizeof(int64_t) * (size_t)Nd,
                            hipMemcpyHostToDevice, st)) != hipSuccess ||
        (e = hipMemcpyAsync(c.wav_rl.p, h_rl.data(), sizeof(int32_t) * (size_t)Nd,
                            hipMemcpyHostToDevice, st)) != hipSuccess)
        return hip_fail(c, e, "row table");

    // CSC by a stable radix sort of the columns (entries are in storage order,
    // which the sort keeps within each column)
    const int64_t nz = std::max<int64_t>(nnz, 1);
    DevBuf keys_out, perm_in, perm_out, tmp;
    int bits = 1;
    while ((int64_t(1) << bits) < n + 1) ++bits;
    if ((e = keys_out.reserve(sizeof(int32_t) * (size_t)nz)) != hipSuccess ||
        (e = perm_in.reserve(sizeof(int32_t) * (size_t)nz)) != hipSuccess ||
        (e = perm_out.reserve(sizeof(int32_t) * (size_t)nz)) != hipSuccess ||
        (e = c.wav_cp.reserve(sizeof(int64_t) * (size_t)(n + 1))) != hipSuccess ||
        (e = c.wav_ri.reserve(sizeof(int32_t) * (size_t)nz)) != hipSuccess ||
        (e = c.wav_vt.reserve(sizeof(double) * (size_t)nz)) != hipSuccess)
        return hip_fail(c, e, "allocate CSC");
    TimedRegion tr(c, kTimeAssemble);
    if (nnz > 0) {
        hipLaunchKernelGGL(iota_kernel, dim3(blocks_for(nnz, kBlk)), dim3(kBlk), 0, st,
                           perm_in.as<int32_t>(), nnz);
        size_t tb = 0;
        if ((e = hipcub::DeviceRadixSort::SortPairs(nullptr, tb, c.wav_ci.as<int32_t>(),
                                                    keys_out.as<int32_t>(), perm_in.as<int32_t>(),
                                                    perm_out.as<int32_t>(), (int)nnz, 0, bits,
                                                    st)) != hipSuccess ||
            (e = tmp.reserve(tb)) != hipSuccess ||
            (e = hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, c.wav_ci.as<int32_t>(),
                                                    keys_out.as<int32_t>(), perm_in.as<int32_t>(),
                                                    perm_out.as<int32_t>(), (int)nnz, 0, bits,
                                                    st)) != hipSuccess)
            return hip_fail(c, e, "column sort");
        hipLaunchKernelGGL(gather_csc_kernel, dim3(blocks_for(nnz, kBlk)), dim3(kBlk), 0, st,
                           perm_out.as<int32_t>(), rowid.as<int32_t>(), c.wav_v.as<double>(), nnz,
                           c.wav_ri.as<int32_t>(), c.wav_vt.as<double>());
    }
    hipLaunchKernelGGL(col_ptr_kernel, dim3(blocks_for(n + 1, kBlk)), dim3(kBlk), 0, st,
                       keys_out.as<int32_t>(), nnz, n, c.wav_cp.as<int64_t>());
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(c, e, "CSC");
    // column segments of at most kSegLen entries for the G^T product
    {
        std::vector<int64_t> h_cp((size_t)n + 1), h_cs((size_t)n + 1), h_ss;
        std::vector<int32_t> h_sl;
        if ((e = hipMemcpyAsync(h_cp.data(), c.wav_cp.p, sizeof(int64_t) * (size_t)(n + 1),
                                hipMemcpyDeviceToHost, st)) != hipSuccess ||
            (e = hipStreamSynchronize(st)) != hipSuccess)
            return hip_fail(c, e, "column pointers");
        for (int64_t col = 0; col < n; ++col) {
            h_cs[(size_t)col] = (int64_t)h_ss.size();
            for (int64_t b = h_cp[(size_t)col]; b < h_cp[(size_t)col + 1]; b += kSegLen) {
                h_ss.push_back(b);
                h_sl.push_back((int32_t)std::min<int64_t>(kSegLen, h_cp[(size_t)col + 1] - b));
            }
        }
        h_cs[(size_t)n] = (int64_t)h_ss.size();
        const size_t ns = std::max<size_t>(h_ss.size(), 1);
        if ((e = c.wav_seg_start.reserve(sizeof(int64_t) * ns)) != hipSuccess ||
            (e = c.wav_seg_len.reserve(sizeof(int32_t) * ns)) != hipSuccess ||
            (e = c.wav_part.reserve(sizeof(double) * ns)) != hipSuccess ||
            (e = c.wav_col_seg.reserve(sizeof(int64_t) * (size_t)(n + 1))) != hipSuccess ||
            (!h_ss.empty() &&
             ((e = hipMemcpyAsync(c.wav_seg_start.p, h_ss.data(), sizeof(int64_t) * h_ss.size(),
                                  hipMemcpyHostToDevice, st)) != hipSuccess ||
              (e = hipMemcpyAsync(c.wav_seg_len.p, h_sl.data(), sizeof(int32_t) * h_sl.size(),
                                  hipMemcpyHostToDevice, st)) != hipSuccess)) ||
            (e = hipMemcpyAsync(c.wav_col_seg.p, h_cs.data(), sizeof(int64_t) * (size_t)(n + 1),
                                hipMemcpyHostToDevice, st)) != hipSuccess ||
            (e = hipStreamSynchronize(st)) != hipSuccess)
            return hip_fail(c, e, "column segments");
        c.wav_nseg = (int64_t)h_ss.size();
    }
    if ((e = c.wav_x.reserve(sizeof(double) * (size_t)n)) != hipSuccess ||
        (e = c.wav_z.reserve(sizeof(double) * (size_t)n)) != hipSuccess ||
        (e = c.wav_tmp.reserve(sizeof(double) * (size_t)n)) != hipSuccess)
        return hip_fail(c, e, "allocate wavelet vectors");
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(c, e, "wavelet assembly");
    c.wav_ndy = n;
    c.wav_nnz = nnz;
    c.wav_nlds = nlds;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device) != hipSuccess ||
        cus <= 0)
        cus = 256;
    c.wav_cus = cus;
    return GRA_OK;
}

// y = G x on the compressed rows (x: cell order, Nm entries).
static hipError_t wavelet_n(Ctx &c, const double *x, double *y, const int *done, hipStream_t st)
{
    const int64_t n = c.wav_ndy, nlds = c.wav_nlds;
    double *xw = c.wav_x.as<double>(), *tmp = c.wav_tmp.as<double>();
    const int32_t *order = c.wav_order.as<int32_t>();
    if (n <= nlds) {
        if (n >= 2)
            hipLaunchKernelGGL((dwt_lds_kernel<true>), dim3(1), dim3(kLdsThreads),
                               (size_t)n * sizeof(double), st, x, xw, n, n, order, c.Nm,
                               (int64_t)0, done);
        else
            hipLaunchKernelGGL(permute_pad_kernel, dim3(1), dim3(kBlk), 0, st, x, order, c.Nm, n,
                               xw, done);
    } else {
        hipLaunchKernelGGL(permute_pad_kernel, dim3(blocks_for(n, kBlk)), dim3(kBlk), 0, st, x,
                           order, c.Nm, n, xw, done);
        hipError_t e = dwt_rows_forward(xw, tmp, 1, n, n, nlds, done, st);
        if (e != hipSuccess) return e;
    }
    if (n <= nlds) {
        const int64_t per = (c.Nd + c.wav_cus - 1) / c.wav_cus;
        hipLaunchKernelGGL(spmv_lds_kernel, dim3(blocks_for(c.Nd, (int)per)),
                           dim3(kSpWaves * kWave), (size_t)n * sizeof(double), st,
                           c.wav_rs.as<int64_t>(), c.wav_rl.as<int32_t>(), c.wav_ci.as<int32_t>(),
                           c.wav_v.as<double>(), xw, n, c.Nd, per, y, done);
    } else {
        hipLaunchKernelGGL(spmv_rows_kernel, dim3(blocks_for(c.Nd, kBlk / kWave)), dim3(kBlk), 0,
                           st, c.wav_rs.as<int64_t>(), c.wav_rl.as<int32_t>(),
                           c.wav_ci.as<int32_t>(), c.wav_v.as<double>(), xw, c.Nd, y, done);
    }
    return hipGetLastError();
}

// z = G^T w; z holds out_len entries (cell order, zero beyond Nm).
static hipError_t wavelet_t(Ctx &c, const double *w, double *z, int64_t out_len, const int *done,
                            hipStream_t st)
{
    const int64_t n = c.wav_ndy, nlds = c.wav_nlds;
    double *zw = c.wav_z.as<double>(), *tmp = c.wav_tmp.as<double>();
    const int32_t *order = c.wav_order.as<int32_t>();
    const int64_t nseg = c.wav_nseg;
    double *part = c.wav_part.as<double>();
    if (nseg == 0) {
        // no coefficient kept: segment_sum writes zeros
    } else if (c.Nd <= nlds) {
        const int64_t per = (nseg + c.wav_cus - 1) / c.wav_cus;
        hipLaunchKernelGGL(spmv_lds_kernel, dim3(blocks_for(nseg, (int)per)),
                           dim3(kSpWaves * kWave), (size_t)c.Nd * sizeof(double), st,
                           c.wav_seg_start.as<int64_t>(), c.wav_seg_len.as<int32_t>(),
                           c.wav_ri.as<int32_t>(), c.wav_vt.as<double>(), w, c.Nd, nseg, per,
                           part, done);
    } else {
        hipLaunchKernelGGL(spmv_rows_kernel, dim3(blocks_for(nseg, kBlk / kWave)), dim3(kBlk), 0,
                           st, c.wav_seg_start.as<int64_t>(), c.wav_seg_len.as<int32_t>(),
                           c.wav_ri.as<int32_t>(), c.wav_vt.as<double>(), w, nseg, part, done);
    }
    hipLaunchKernelGGL(segment_sum_kernel, dim3(blocks_for(n, kBlk)), dim3(kBlk), 0, st,
                       c.wav_col_seg.as<int64_t>(), part, n, zw, done);
    if (n <= nlds) {
        if (n >= 2)
            hipLaunchKernelGGL((dwt_lds_kernel<false>), dim3(1), dim3(kLdsThreads),
                               (size_t)n * sizeof(double), st, zw, z, n, n, order, c.Nm, out_len,
                               done);
        else
            hipLaunchKernelGGL(unpermute_kernel, dim3(blocks_for(out_len, kBlk)), dim3(kBlk), 0,
                               st, zw, order, c.Nm, out_len, z, done);
    } else {
        hipLaunchKernelGGL((dwt_lds_kernel<false>), dim3(1), dim3(kLdsThreads),
                           (size_t)nlds * sizeof(double), st, zw, zw, n, nlds, nullptr,
                           (int64_t)0, (int64_t)0, done);
        for (int64_t m = nlds * 2; m <= n; m <<= 1) {
            hipLaunchKernelGGL(dwt_level_inv_kernel, dim3(blocks_for(m, kBlk), 1), dim3(kBlk), 0,
                               st, zw, tmp, n, m, done);
            hipLaunchKernelGGL(copy_rows_kernel, dim3(blocks_for(m, kBlk), 1), dim3(kBlk), 0, st,
                               tmp, zw, n, m, done);
        }
        hipLaunchKernelGGL(unpermute_kernel, dim3(blocks_for(std::max(out_len, c.Nm), kBlk)),
                           dim3(kBlk), 0, st, zw, order, c.Nm, out_len, z, done);
    }
    return hipGetLastError();
}

hipError_t launch_G_n(Ctx &c, const GemvShape &s, const double *x, double *part, const int *done,
                      hipStream_t st)
{
    if (c.wav) return wavelet_n(c, x, part, done, st);
    return launch_gemv_n(c.G.p, c.store, c.Nd, c.ld, s, x, part, done, st);
}

hipError_t launch_G_t(Ctx &c, const GemvShape &s, const double *w, double *part, const int *done,
                      hipStream_t st)
{
    if (c.wav) return wavelet_t(c, w, part, c.ld, done, st);
    return launch_gemv_t(c.G.p, c.store, c.Nd, c.ld, s, w, part, done, st);
}

}  // namespace gra
