// Dense G products on the row-major, zero-padded G (leading dimension ld).
//
// Both products are HBM-bound streams over G (2 flops per element), so the
// kernels are shaped for bandwidth: 16-byte (FP64: 2 elements) vector loads
// per lane, several independent row streams per wave for memory-level
// parallelism, and a deterministic two-level reduction (fixed xor butterfly in
// the wave, then the column splits / row chunks summed in index order by the
// consumer), so every product is bit-for-bit repeatable.
//
//   y = G x   (Fwd::G_vec_mul dense branch, Fwd.cpp:145): a wave owns R rows
//             of one column split; x is loaded once per column step and reused
//             for the R rows.
//   z = G^T w (Fwd::GT_vec_mul dense branch, Fwd.cpp:111): a lane owns a
//             pair of adjacent columns and walks a chunk of rows; w[row] is a
//             wave-uniform scalar.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include "engine.hpp"

namespace gra {

namespace {

constexpr int kWave = 64;
constexpr int kBlk = 256;
constexpr int kRowsPerWave = 4;

// Load 2 consecutive elements of G as doubles.
__device__ inline double2 ld2(const double *p) { return *reinterpret_cast<const double2 *>(p); }
__device__ inline double2 ld2(const float *p)
{
    float2 f = *reinterpret_cast<const float2 *>(p);
    return make_double2((double)f.x, (double)f.y);
}
__device__ inline double2 ld2(const __hip_bfloat16 *p)
{
    uint32_t u = *reinterpret_cast<const uint32_t *>(p);
    float lo = __uint_as_float(u << 16), hi = __uint_as_float(u & 0xffff0000u);
    return make_double2((double)lo, (double)hi);
}

template <typename T>
__global__ void __launch_bounds__(kBlk)
gemv_n_kernel(const T *__restrict__ G, int64_t Nd, int64_t ld, int64_t cols_per_split,
              const double *__restrict__ x, double *__restrict__ part, const int *done)
{
    if (done && *done) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wave = (int64_t)blockIdx.x * (kBlk / kWave) + (threadIdx.x >> 6);
    const int64_t row0 = wave * kRowsPerWave;
    const int split = blockIdx.y;
    if (row0 >= Nd) return;
    int64_t c0 = (int64_t)split * cols_per_split;
    int64_t c1 = c0 + cols_per_split;
    if (c1 > ld) c1 = ld;
    const T *rows[kRowsPerWave];
#pragma unroll
    for (int r = 0; r < kRowsPerWave; ++r) {
        int64_t row = row0 + r < Nd ? row0 + r : Nd - 1;   // clamp: duplicate reads, result dropped
        rows[r] = G + row * ld;
    }
    double acc[kRowsPerWave];
#pragma unroll
    for (int r = 0; r < kRowsPerWave; ++r) acc[r] = 0.0;
    int64_t c = c0 + 2 * lane;
    for (; c + 2 * kWave < c1; c += 4 * kWave) {
        double2 xa = *reinterpret_cast<const double2 *>(x + c);
        double2 xb = *reinterpret_cast<const double2 *>(x + c + 2 * kWave);
        double2 ga[kRowsPerWave], gb[kRowsPerWave];
#pragma unroll
        for (int r = 0; r < kRowsPerWave; ++r) {
            ga[r] = ld2(rows[r] + c);
            gb[r] = ld2(rows[r] + c + 2 * kWave);
        }
#pragma unroll
        for (int r = 0; r < kRowsPerWave; ++r) {
            acc[r] = fma(ga[r].x, xa.x, acc[r]);
            acc[r] = fma(ga[r].y, xa.y, acc[r]);
            acc[r] = fma(gb[r].x, xb.x, acc[r]);
            acc[r] = fma(gb[r].y, xb.y, acc[r]);
        }
    }
    for (; c < c1; c += 2 * kWave) {
        double2 xa = *reinterpret_cast<const double2 *>(x + c);
#pragma unroll
        for (int r = 0; r < kRowsPerWave; ++r) {
            double2 g = ld2(rows[r] + c);
            acc[r] = fma(g.x, xa.x, acc[r]);
            acc[r] = fma(g.y, xa.y, acc[r]);
        }
    }
#pragma unroll
    for (int r = 0; r < kRowsPerWave; ++r) {
        double v = acc[r];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
        acc[r] = v;
    }
    if (lane == 0) {
#pragma unroll
        for (int r = 0; r < kRowsPerWave; ++r)
            if (row0 + r < Nd) part[(int64_t)split * Nd + row0 + r] = acc[r];
    }
}

constexpr int kColsPerLane = 2;
constexpr int kColsPerBlock = kBlk * kColsPerLane;

template <typename T>
__global__ void __launch_bounds__(kBlk)
gemv_t_kernel(const T *__restrict__ G, int64_t Nd, int64_t ld, int64_t rows_per_chunk,
              const double *__restrict__ w, double *__restrict__ part, const int *done)
{
    if (done && *done) return;
    const int64_t c = (int64_t)blockIdx.x * kColsPerBlock + kColsPerLane * threadIdx.x;
    const int chunk = blockIdx.y;
    int64_t r0 = (int64_t)chunk * rows_per_chunk;
    int64_t r1 = r0 + rows_per_chunk;
    if (r1 > Nd) r1 = Nd;
    if (c >= ld) return;
    const T *g = G + c;
    double ax = 0.0, ay = 0.0;
    int64_t r = r0;
    for (; r + 4 <= r1; r += 4) {
        double2 g0 = ld2(g + (r + 0) * ld);
        double2 g1 = ld2(g + (r + 1) * ld);
        double2 g2 = ld2(g + (r + 2) * ld);
        double2 g3 = ld2(g + (r + 3) * ld);
        double w0 = w[r], w1 = w[r + 1], w2 = w[r + 2], w3 = w[r + 3];
        ax = fma(g0.x, w0, ax); ay = fma(g0.y, w0, ay);
        ax = fma(g1.x, w1, ax); ay = fma(g1.y, w1, ay);
        ax = fma(g2.x, w2, ax); ay = fma(g2.y, w2, ay);
        ax = fma(g3.x, w3, ax); ay = fma(g3.y, w3, ay);
    }
    for (; r < r1; ++r) {
        double2 g0 = ld2(g + r * ld);
        double w0 = w[r];
        ax = fma(g0.x, w0, ax);
        ay = fma(g0.y, w0, ay);
    }
    *reinterpret_cast<double2 *>(part + (int64_t)chunk * ld + c) = make_double2(ax, ay);
}

__global__ void reduce_parts_kernel(const double *__restrict__ part, int nparts, int64_t stride,
                                    int64_t n, const double *__restrict__ scale,
                                    double *__restrict__ y)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = part[i];
    for (int q = 1; q < nparts; ++q) s += part[(int64_t)q * stride + i];
    y[i] = scale ? scale[i] * s : s;
}

template <typename T>
__global__ void expand_kernel(const T *__restrict__ G, int64_t Nd, int64_t Nm, int64_t ld,
                              double *__restrict__ out, int col_major)
{
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= Nd * Nm) return;
    int64_t i, j;
    if (col_major) { j = idx / Nd; i = idx - j * Nd; }
    else { i = idx / Nm; j = idx - i * Nm; }
    double v;
    if constexpr (sizeof(T) == 2) {
        uint16_t u = reinterpret_cast<const uint16_t *>(G)[i * ld + j];
        v = (double)__uint_as_float((uint32_t)u << 16);
    } else {
        v = (double)G[i * ld + j];
    }
    out[idx] = v;
}

}  // namespace

int gemv_n_splits(int64_t Nd, int64_t Nm)
{
    int64_t waves = (Nd + kRowsPerWave - 1) / kRowsPerWave;
    int64_t want = 8192;
    int64_t s = (want + waves - 1) / waves;
    int64_t max_by_cols = Nm / (8 * kWave) > 0 ? Nm / (8 * kWave) : 1;
    if (s > max_by_cols) s = max_by_cols;
    if (s > 64) s = 64;
    if (s < 1) s = 1;
    return (int)s;
}

int gemv_t_chunks(int64_t Nd, int64_t Nm)
{
    int64_t bx = (Nm + kColsPerBlock - 1) / kColsPerBlock;
    int64_t want = 4096;   // blocks
    int64_t ch = (want + bx - 1) / bx;
    int64_t max_by_rows = Nd / 16 > 0 ? Nd / 16 : 1;
    if (ch > max_by_rows) ch = max_by_rows;
    if (ch > 256) ch = 256;
    if (ch < 1) ch = 1;
    return (int)ch;
}

hipError_t launch_gemv_n(const void *G, gra_dtype dt, int64_t Nd, int64_t Nm, int64_t ld,
                         const double *x, double *part, int splits, const int *done,
                         hipStream_t st)
{
    (void)Nm;
    int64_t waves = (Nd + kRowsPerWave - 1) / kRowsPerWave;
    int64_t blocks = (waves + (kBlk / kWave) - 1) / (kBlk / kWave);
    int64_t cps = (ld + splits - 1) / splits;
    cps = (cps + 2 * kWave - 1) / (2 * kWave) * (2 * kWave);   // keep 16-B lane alignment
    dim3 grid((unsigned)blocks, (unsigned)splits), block(kBlk);
    switch (dt) {
    case GRA_F64:
        hipLaunchKernelGGL(gemv_n_kernel<double>, grid, block, 0, st, (const double *)G, Nd, ld,
                           cps, x, part, done);
        break;
    case GRA_F32:
        hipLaunchKernelGGL(gemv_n_kernel<float>, grid, block, 0, st, (const float *)G, Nd, ld, cps,
                           x, part, done);
        break;
    default:
        hipLaunchKernelGGL(gemv_n_kernel<__hip_bfloat16>, grid, block, 0, st,
                           (const __hip_bfloat16 *)G, Nd, ld, cps, x, part, done);
        break;
    }
    return hipGetLastError();
}

hipError_t launch_gemv_t(const void *G, gra_dtype dt, int64_t Nd, int64_t Nm, int64_t ld,
                         const double *w, double *part, int chunks, const int *done,
                         hipStream_t st)
{
    (void)Nm;
    int64_t bx = (ld + kColsPerBlock - 1) / kColsPerBlock;
    int64_t rpc = (Nd + chunks - 1) / chunks;
    dim3 grid((unsigned)bx, (unsigned)chunks), block(kBlk);
    switch (dt) {
    case GRA_F64:
        hipLaunchKernelGGL(gemv_t_kernel<double>, grid, block, 0, st, (const double *)G, Nd, ld,
                           rpc, w, part, done);
        break;
    case GRA_F32:
        hipLaunchKernelGGL(gemv_t_kernel<float>, grid, block, 0, st, (const float *)G, Nd, ld, rpc,
                           w, part, done);
        break;
    default:
        hipLaunchKernelGGL(gemv_t_kernel<__hip_bfloat16>, grid, block, 0, st,
                           (const __hip_bfloat16 *)G, Nd, ld, rpc, w, part, done);
        break;
    }
    return hipGetLastError();
}

hipError_t launch_reduce_parts(const double *part, int nparts, int64_t stride, int64_t n,
                               const double *scale, double *y, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    dim3 grid((unsigned)((n + 255) / 256)), block(256);
    hipLaunchKernelGGL(reduce_parts_kernel, grid, block, 0, st, part, nparts, stride, n, scale, y);
    return hipGetLastError();
}

hipError_t launch_expand_G(const void *G, gra_dtype dt, int64_t Nd, int64_t Nm, int64_t ld,
                           double *out, int col_major, hipStream_t st)
{
    int64_t n = Nd * Nm;
    if (n <= 0) return hipSuccess;
    dim3 grid((unsigned)((n + 255) / 256)), block(256);
    switch (dt) {
    case GRA_F64:
        hipLaunchKernelGGL(expand_kernel<double>, grid, block, 0, st, (const double *)G, Nd, Nm, ld,
                           out, col_major);
        break;
    case GRA_F32:
        hipLaunchKernelGGL(expand_kernel<float>, grid, block, 0, st, (const float *)G, Nd, Nm, ld,
                           out, col_major);
        break;
    default:
        hipLaunchKernelGGL(expand_kernel<__hip_bfloat16>, grid, block, 0, st,
                           (const __hip_bfloat16 *)G, Nd, Nm, ld, out, col_major);
        break;
    }
    return hipGetLastError();
}

}  // namespace gra
