// Dense G products on the row-major, zero-padded G (leading dimension ld).
//
// Both products are HBM-bound streams over G (2 flops per element), so the
// kernels are shaped for bandwidth: every lane moves 16 bytes of G per load
// whatever the storage type (2 FP64, 4 FP32 or 8 BF16 elements), several
// independent row streams are in flight per wave, G is read with
// non-temporal loads (it is streamed once per product; the vectors stay
// cached), and the reductions are deterministic: a fixed xor butterfly in the
// wave, then the column splits / row chunks summed in index order by the
// consumer, so every product is bit-for-bit repeatable.  Accumulation is FP64
// for every storage type.
//
//   y = G x   (Fwd::G_vec_mul dense branch, Fwd.cpp:145): a wave owns R rows
//             of one column split; each x chunk is loaded once for the R rows.
//             Block-shared form: the four waves of a workgroup share its R
//             rows, so each step reads 4 KB contiguous of every row (the
//             footprint of the G^T w kernel), and the wave sums meet in LDS.
//   z = G^T w (Fwd::GT_vec_mul dense branch, Fwd.cpp:111): a lane owns the
//             16-byte column group it loads and walks a chunk of rows; w[row]
//             is a wave-uniform scalar.
// Launch shapes (R, unroll, splits, chunks, non-temporal) are tunable per
// context (gra_tune_set); the defaults come from a sweep on MI355X.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <type_traits>

#include "engine.hpp"

namespace gra {

namespace {

constexpr int kWave = 64;
constexpr int kBlk = 256;

template <typename T> struct Vec16;
template <> struct Vec16<double> { static constexpr int n = 2; };
template <> struct Vec16<float> { static constexpr int n = 4; };
template <> struct Vec16<__hip_bfloat16> { static constexpr int n = 8; };

typedef double v2f64 __attribute__((ext_vector_type(2)));
typedef float v4f32 __attribute__((ext_vector_type(4)));
typedef uint32_t v4u32 __attribute__((ext_vector_type(4)));

// 16-byte load of G converted to n doubles.
template <typename T, bool NT> struct Load16;

template <bool NT> struct Load16<double, NT> {
    __device__ static inline void run(const double *p, double *o)
    {
        v2f64 v;
        if (NT) v = __builtin_nontemporal_load(reinterpret_cast<const v2f64 *>(p));
        else v = *reinterpret_cast<const v2f64 *>(p);
        o[0] = v.x; o[1] = v.y;
    }
};
template <bool NT> struct Load16<float, NT> {
    __device__ static inline void run(const float *p, double *o)
    {
        v4f32 v;
        if (NT) v = __builtin_nontemporal_load(reinterpret_cast<const v4f32 *>(p));
        else v = *reinterpret_cast<const v4f32 *>(p);
        o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
    }
};
template <bool NT> struct Load16<__hip_bfloat16, NT> {
    __device__ static inline void run(const __hip_bfloat16 *p, double *o)
    {
        v4u32 v;
        if (NT) v = __builtin_nontemporal_load(reinterpret_cast<const v4u32 *>(p));
        else v = *reinterpret_cast<const v4u32 *>(p);
        uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            o[2 * q] = (double)__uint_as_float(w[q] << 16);
            o[2 * q + 1] = (double)__uint_as_float(w[q] & 0xffff0000u);
        }
    }
};

// ------------------------------------------------------------------ G x
template <typename T, int R, int U, bool NT>
__global__ void __launch_bounds__(kBlk)
gemv_n_kernel(const T *__restrict__ G, int64_t Nd, int64_t ld, int64_t cols_per_split,
              const double *__restrict__ x, double *__restrict__ part, const int *done)
{
    if (done && *done) return;
    constexpr int V = Vec16<T>::n;            // elements per lane per load
    constexpr int STEP = V * kWave;           // columns per wave step
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wave = (int64_t)blockIdx.x * (kBlk / kWave) + (threadIdx.x >> 6);
    const int64_t row0 = wave * R;
    const int split = blockIdx.y;
    if (row0 >= Nd) return;
    int64_t c0 = (int64_t)split * cols_per_split;
    int64_t c1 = c0 + cols_per_split;
    if (c1 > ld) c1 = ld;
    const T *rows[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int64_t row = row0 + r < Nd ? row0 + r : Nd - 1;   // clamp: duplicate reads, result dropped
        rows[r] = G + row * ld;
    }
    double acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.0;
    int64_t c = c0 + (int64_t)V * lane;
    for (; c + (U - 1) * STEP < c1; c += U * STEP) {
        double xv[U][V];
        double g[U][R][V];
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int r = 0; r < R; ++r) Load16<T, NT>::run(rows[r] + c + u * STEP, g[u][r]);
#pragma unroll
            for (int q = 0; q < V; q += 2) {
                double2 xx = *reinterpret_cast<const double2 *>(x + c + u * STEP + q);
                xv[u][q] = xx.x;
                xv[u][q + 1] = xx.y;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int q = 0; q < V; ++q) acc[r] = fma(g[u][r][q], xv[u][q], acc[r]);
    }
    for (; c < c1; c += STEP) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            double g[V];
            Load16<T, NT>::run(rows[r] + c, g);
#pragma unroll
            for (int q = 0; q < V; ++q) acc[r] = fma(g[q], x[c + q], acc[r]);
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        double v = acc[r];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
        acc[r] = v;
    }
    if (lane == 0) {
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (row0 + r < Nd) part[(int64_t)split * Nd + row0 + r] = acc[r];
    }
}

// BF16 G x with up to 16 rows per wave: the 16-byte pieces are kept as
// stored (8 BF16 values in 4 registers) and widened at the multiply-add, so
// that a wave can hold 12-16 rows' pieces and load each 64 bytes of x once
// for all of them (with the pieces widened at the load, 8 rows already took
// the registers of occupancy 2).  Per lane and row the products are added in
// the same order as gemv_n_kernel's, so the results are identical for the
// same column splits.
template <int R, int U, bool NT>
__global__ void __launch_bounds__(kBlk)
gemv_nr_kernel(const __hip_bfloat16 *__restrict__ G, int64_t Nd, int64_t ld, int64_t cols_per_split,
               const double *__restrict__ x, double *__restrict__ part, const int *done)
{
    if (done && *done) return;
    constexpr int V = 8;
    constexpr int STEP = V * kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wave = (int64_t)blockIdx.x * (kBlk / kWave) + (threadIdx.x >> 6);
    const int64_t row0 = wave * R;
    const int split = blockIdx.y;
    if (row0 >= Nd) return;
    int64_t c0 = (int64_t)split * cols_per_split;
    int64_t c1 = c0 + cols_per_split;
    if (c1 > ld) c1 = ld;
    const __hip_bfloat16 *rows[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int64_t row = row0 + r < Nd ? row0 + r : Nd - 1;   // clamp: duplicate reads, result dropped
        rows[r] = G + row * ld;
    }
    double acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.0;
    auto load = [&](const __hip_bfloat16 *p) {
        return NT ? __builtin_nontemporal_load(reinterpret_cast<const v4u32 *>(p))
                  : *reinterpret_cast<const v4u32 *>(p);
    };
    auto madd = [&](const v4u32 &g, const double *xv, double &a) {
        const uint32_t w[4] = {g.x, g.y, g.z, g.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            a = fma((double)__uint_as_float(w[q] << 16), xv[2 * q], a);
            a = fma((double)__uint_as_float(w[q] & 0xffff0000u), xv[2 * q + 1], a);
        }
    };
    int64_t c = c0 + (int64_t)V * lane;
    for (; c + (U - 1) * STEP < c1; c += U * STEP) {
        v4u32 g[U][R];
        double xv[U][V];
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int r = 0; r < R; ++r) g[u][r] = load(rows[r] + c + u * STEP);
#pragma unroll
            for (int q = 0; q < V; q += 2) {
                double2 xx = *reinterpret_cast<const double2 *>(x + c + u * STEP + q);
                xv[u][q] = xx.x;
                xv[u][q + 1] = xx.y;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int r = 0; r < R; ++r) madd(g[u][r], xv[u], acc[r]);
    }
    for (; c < c1; c += STEP) {
        double xv[V];
#pragma unroll
        for (int q = 0; q < V; ++q) xv[q] = x[c + q];
#pragma unroll
        for (int r = 0; r < R; ++r) madd(load(rows[r] + c), xv, acc[r]);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        double v = acc[r];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
        acc[r] = v;
    }
    if (lane == 0) {
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (row0 + r < Nd) part[(int64_t)split * Nd + row0 + r] = acc[r];
    }
}

// Block-shared rows: lane t of the workgroup handles columns V*t + k*BSTEP of
// the split, BSTEP = 256*V (4 KB of G per row and step); per row the lane
// sums in column order, the wave by the xor butterfly, then the four wave
// sums in wave order.
template <typename T, int R, int U, bool NT>
__global__ void __launch_bounds__(kBlk)
gemv_nb_kernel(const T *__restrict__ G, int64_t Nd, int64_t ld, int64_t cols_per_split,
               const double *__restrict__ x, double *__restrict__ part, const int *done)
{
    if (done && *done) return;
    constexpr int V = Vec16<T>::n;
    constexpr int BSTEP = V * kBlk;
    constexpr int W = kBlk / kWave;
    __shared__ double red[W][R];
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int64_t row0 = (int64_t)blockIdx.x * R;
    const int split = blockIdx.y;
    int64_t c0 = (int64_t)split * cols_per_split;
    int64_t c1 = c0 + cols_per_split;
    if (c1 > ld) c1 = ld;
    const T *rows[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int64_t row = row0 + r < Nd ? row0 + r : Nd - 1;   // clamp: duplicate reads, result dropped
        rows[r] = G + row * ld;
    }
    double acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.0;
    int64_t c = c0 + (int64_t)V * threadIdx.x;
    for (; c + (U - 1) * BSTEP < c1; c += U * BSTEP) {
        double xv[U][V];
        double g[U][R][V];
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int r = 0; r < R; ++r) Load16<T, NT>::run(rows[r] + c + u * BSTEP, g[u][r]);
#pragma unroll
            for (int q = 0; q < V; q += 2) {
                double2 xx = *reinterpret_cast<const double2 *>(x + c + u * BSTEP + q);
                xv[u][q] = xx.x;
                xv[u][q + 1] = xx.y;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int q = 0; q < V; ++q) acc[r] = fma(g[u][r][q], xv[u][q], acc[r]);
    }
    for (; c < c1; c += BSTEP) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            double g[V];
            Load16<T, NT>::run(rows[r] + c, g);
#pragma unroll
            for (int q = 0; q < V; ++q) acc[r] = fma(g[q], x[c + q], acc[r]);
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        double v = acc[r];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
        if (lane == 0) red[wv][r] = v;
    }
    __syncthreads();
    if (threadIdx.x < R && row0 + threadIdx.x < Nd) {
        double s = red[0][threadIdx.x];
#pragma unroll
        for (int q = 1; q < W; ++q) s += red[q][threadIdx.x];
        part[(int64_t)split * Nd + row0 + threadIdx.x] = s;
    }
}

// ---------------------------------------------------------------- G^T w
// SQ: column sums of squares instead (w unused).
template <typename T, int UR, bool NT, bool SQ = false>
__global__ void __launch_bounds__(kBlk)
gemv_t_kernel(const T *__restrict__ G, int64_t Nd, int64_t ld, int64_t rows_per_chunk,
              const double *__restrict__ w, double *__restrict__ part, const int *done)
{
    if (done && *done) return;
    constexpr int V = Vec16<T>::n;
    const int64_t c = (int64_t)blockIdx.x * (kBlk * V) + (int64_t)V * threadIdx.x;
    const int chunk = blockIdx.y;
    int64_t r0 = (int64_t)chunk * rows_per_chunk;
    int64_t r1 = r0 + rows_per_chunk;
    if (r1 > Nd) r1 = Nd;
    if (c >= ld) return;
    const T *g = G + c;
    double acc[V];
#pragma unroll
    for (int q = 0; q < V; ++q) acc[q] = 0.0;
    int64_t r = r0;
    for (; r + UR <= r1; r += UR) {
        double gv[UR][V];
        double wv[UR];
#pragma unroll
        for (int u = 0; u < UR; ++u) {
            Load16<T, NT>::run(g + (r + u) * ld, gv[u]);
            wv[u] = SQ ? 0.0 : w[r + u];
        }
#pragma unroll
        for (int u = 0; u < UR; ++u)
#pragma unroll
            for (int q = 0; q < V; ++q) acc[q] = fma(gv[u][q], SQ ? gv[u][q] : wv[u], acc[q]);
    }
    for (; r < r1; ++r) {
        double gv[V];
        Load16<T, NT>::run(g + r * ld, gv);
        double w0 = SQ ? 0.0 : w[r];
#pragma unroll
        for (int q = 0; q < V; ++q) acc[q] = fma(gv[q], SQ ? gv[q] : w0, acc[q]);
    }
    double *out = part + (int64_t)chunk * ld + c;
#pragma unroll
    for (int q = 0; q < V; q += 2)
        *reinterpret_cast<double2 *>(out + q) = make_double2(acc[q], acc[q + 1]);
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

template <typename T, int R, int U, bool NT>
void launch_n(const void *G, int64_t Nd, int64_t ld, int splits, const double *x, double *part,
              const int *done, hipStream_t st)
{
    constexpr int STEP = Vec16<T>::n * kWave;
    int64_t waves = (Nd + R - 1) / R;
    int64_t blocks = (waves + (kBlk / kWave) - 1) / (kBlk / kWave);
    int64_t cps = (ld + splits - 1) / splits;
    cps = (cps + STEP - 1) / STEP * STEP;   // keep 16-B lane alignment
    dim3 grid((unsigned)blocks, (unsigned)splits), block(kBlk);
    hipLaunchKernelGGL((gemv_n_kernel<T, R, U, NT>), grid, block, 0, st, (const T *)G, Nd, ld, cps,
                       x, part, done);
}

template <typename T, int R, int U, bool NT>
void launch_nb(const void *G, int64_t Nd, int64_t ld, int splits, const double *x, double *part,
               const int *done, hipStream_t st)
{
    constexpr int BSTEP = Vec16<T>::n * kBlk;
    int64_t blocks = (Nd + R - 1) / R;
    int64_t cps = (ld + splits - 1) / splits;
    cps = (cps + BSTEP - 1) / BSTEP * BSTEP;   // keep 16-B lane alignment
    dim3 grid((unsigned)blocks, (unsigned)splits), block(kBlk);
    hipLaunchKernelGGL((gemv_nb_kernel<T, R, U, NT>), grid, block, 0, st, (const T *)G, Nd, ld,
                       cps, x, part, done);
}

template <typename T, int R, int U>
void launch_n_nt(const void *G, int64_t Nd, int64_t ld, int splits, bool nt, bool blk,
                 const double *x, double *part, const int *done, hipStream_t st)
{
    if (blk) {
        if (nt) launch_nb<T, R, U, true>(G, Nd, ld, splits, x, part, done, st);
        else launch_nb<T, R, U, false>(G, Nd, ld, splits, x, part, done, st);
    } else {
        if (nt) launch_n<T, R, U, true>(G, Nd, ld, splits, x, part, done, st);
        else launch_n<T, R, U, false>(G, Nd, ld, splits, x, part, done, st);
    }
}

template <int R, int U>
void launch_nr(const void *G, int64_t Nd, int64_t ld, int splits, bool nt, const double *x,
               double *part, const int *done, hipStream_t st)
{
    constexpr int STEP = 8 * kWave;
    int64_t waves = (Nd + R - 1) / R;
    int64_t blocks = (waves + (kBlk / kWave) - 1) / (kBlk / kWave);
    int64_t cps = (ld + splits - 1) / splits;
    cps = (cps + STEP - 1) / STEP * STEP;   // keep 16-B lane alignment
    dim3 grid((unsigned)blocks, (unsigned)splits), block(kBlk);
    const __hip_bfloat16 *g = (const __hip_bfloat16 *)G;
    if (nt) hipLaunchKernelGGL((gemv_nr_kernel<R, U, true>), grid, block, 0, st, g, Nd, ld, cps, x, part, done);
    else hipLaunchKernelGGL((gemv_nr_kernel<R, U, false>), grid, block, 0, st, g, Nd, ld, cps, x, part, done);
}

template <typename T>
void launch_n_shape(const void *G, int64_t Nd, int64_t ld, const GemvShape &s, const double *x,
                    double *part, const int *done, hipStream_t st)
{
    const bool nt = (s.nontemporal & 1) != 0, blk = s.n_block != 0;
    const int sp = s.n_splits;
    if constexpr (std::is_same<T, __hip_bfloat16>::value) {
        // 12 or 16 rows per wave: BF16 pieces kept as stored
        if (!blk && s.n_rows >= 12) {
            if (s.n_rows >= 16) launch_nr<16, 1>(G, Nd, ld, sp, nt, x, part, done, st);
            else launch_nr<12, 2>(G, Nd, ld, sp, nt, x, part, done, st);
            return;
        }
    }
    if (s.n_rows == 2) {
        if (s.n_unroll >= 2) launch_n_nt<T, 2, 2>(G, Nd, ld, sp, nt, blk, x, part, done, st);
        else launch_n_nt<T, 2, 1>(G, Nd, ld, sp, nt, blk, x, part, done, st);
    } else if (s.n_rows >= 8) {
        if (s.n_unroll >= 2) launch_n_nt<T, 8, 2>(G, Nd, ld, sp, nt, blk, x, part, done, st);
        else launch_n_nt<T, 8, 1>(G, Nd, ld, sp, nt, blk, x, part, done, st);
    } else {
        if (s.n_unroll >= 2) launch_n_nt<T, 4, 2>(G, Nd, ld, sp, nt, blk, x, part, done, st);
        else launch_n_nt<T, 4, 1>(G, Nd, ld, sp, nt, blk, x, part, done, st);
    }
}

template <typename T>
void launch_t_shape(const void *G, int64_t Nd, int64_t ld, const GemvShape &s, const double *w,
                    double *part, const int *done, hipStream_t st)
{
    constexpr int V = Vec16<T>::n;
    int64_t bx = (ld + kBlk * V - 1) / (kBlk * V);
    int64_t rpc = (Nd + s.t_chunks - 1) / s.t_chunks;
    dim3 grid((unsigned)bx, (unsigned)s.t_chunks), block(kBlk);
    const bool nt = (s.nontemporal & 2) != 0;
#define GRA_T(UR, NTV)                                                                        \
    hipLaunchKernelGGL((gemv_t_kernel<T, UR, NTV>), grid, block, 0, st, (const T *)G, Nd, ld, \
                       rpc, w, part, done)
    if (s.t_unroll >= 8) {
        if (nt) GRA_T(8, true); else GRA_T(8, false);
    } else {
        if (nt) GRA_T(4, true); else GRA_T(4, false);
    }
#undef GRA_T
}

}  // namespace

// Default launch shapes, from a sweep on MI355X (5,000 x 32,000 G in each
// storage type): G x is fastest with 4 rows per wave for FP32 and ~8k waves
// in flight, for FP64 with 4 rows shared by a workgroup, for BF16 with 16
// rows per wave kept as stored (90,000 x 131,072 G: 3.54 ms against 4.01 ms
// with 8 rows widened at the load);
// G^T w with long row chunks, ~256 blocks (one per CU) in total;
// non-temporal loads help both.
GemvShape default_gemv_shape(gra_dtype dt, int64_t Nd, int64_t Nm)
{
    GemvShape s;
    const int V = dt == GRA_F64 ? 2 : (dt == GRA_F32 ? 4 : 8);
    s.n_rows = dt == GRA_F64 ? 2 : (dt == GRA_F32 ? 4 : 16);
    s.n_unroll = dt == GRA_BF16 ? 1 : 2;
    s.t_unroll = 8;
    s.nontemporal = 3;
    int64_t waves = (Nd + s.n_rows - 1) / s.n_rows;
    int64_t sp = (8192 + waves - 1) / waves;
    int64_t max_by_cols = Nm / (4 * V * kWave) > 0 ? Nm / (4 * V * kWave) : 1;
    if (dt == GRA_F64) {
        // FP64: block-shared rows, 4 per workgroup, about 1,250 workgroups
        // (186 us against 190 us for the best wave-owned shape on the
        // 5,000 x 32,000 G, both between G^T w launches as in CG)
        s.n_block = 1;
        s.n_rows = 4;
        s.n_unroll = 1;
        int64_t blocks = (Nd + 3) / 4;
        sp = (1250 + blocks - 1) / blocks;
        max_by_cols = Nm / (2 * V * kBlk) > 0 ? Nm / (2 * V * kBlk) : 1;
    }
    sp = sp > max_by_cols ? max_by_cols : sp;
    s.n_splits = (int)(sp < 1 ? 1 : (sp > 64 ? 64 : sp));
    // G^T w: about 224 workgroups (column blocks x row chunks), never just
    // above a multiple of the 256 CUs -- a launch of 260 runs its last few
    // blocks as a second wave and loses up to 25 % of the bandwidth (swept
    // over Nd 2k-60k, Nm 2.5k-64k on MI355X)
    int64_t bx = (Nm + kBlk * V - 1) / (kBlk * V);
    int64_t ch = 224 / bx > 0 ? 224 / bx : 1;
    if (bx * ch < 160 && bx * (ch + 1) <= 256) ++ch;
    int64_t max_by_rows = Nd / 32 > 0 ? Nd / 32 : 1;
    ch = ch > max_by_rows ? max_by_rows : ch;
    s.t_chunks = (int)(ch < 1 ? 1 : (ch > 256 ? 256 : ch));
    return s;
}

hipError_t launch_gemv_n(const void *G, gra_dtype dt, int64_t Nd, int64_t ld, const GemvShape &s,
                         const double *x, double *part, const int *done, hipStream_t st)
{
    switch (dt) {
    case GRA_F64: launch_n_shape<double>(G, Nd, ld, s, x, part, done, st); break;
    case GRA_F32: launch_n_shape<float>(G, Nd, ld, s, x, part, done, st); break;
    default: launch_n_shape<__hip_bfloat16>(G, Nd, ld, s, x, part, done, st); break;
    }
    return hipGetLastError();
}

hipError_t launch_gemv_t(const void *G, gra_dtype dt, int64_t Nd, int64_t ld, const GemvShape &s,
                         const double *w, double *part, const int *done, hipStream_t st)
{
    switch (dt) {
    case GRA_F64: launch_t_shape<double>(G, Nd, ld, s, w, part, done, st); break;
    case GRA_F32: launch_t_shape<float>(G, Nd, ld, s, w, part, done, st); break;
    default: launch_t_shape<__hip_bfloat16>(G, Nd, ld, s, w, part, done, st); break;
    }
    return hipGetLastError();
}

// G2[row, j] = src[j] >= 0 ? G[row, src[j]] : Gc[row, -src[j] - 1], copying the
// stored bits (storage element of E bytes).
template <typename E>
__global__ void remap_cols_kernel(const E *__restrict__ G, int64_t ld, const E *__restrict__ Gc,
                                  int64_t ldc, const int32_t *__restrict__ src, int64_t Nd,
                                  int64_t Nm2, E *__restrict__ G2, int64_t ld2)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t row = blockIdx.y + (int64_t)gridDim.y * blockIdx.z;
    if (j >= Nm2 || row >= Nd) return;
    const int32_t s = src[j];
    G2[row * ld2 + j] = s >= 0 ? G[row * ld + s] : Gc[row * ldc + (-(int64_t)s - 1)];
}

hipError_t launch_remap_cols(const void *G, int64_t ld, const void *Gc, int64_t ldc,
                             const int32_t *src, int64_t Nd, int64_t Nm2, void *G2, int64_t ld2,
                             int elem_bytes, hipStream_t st)
{
    if (Nd <= 0 || Nm2 <= 0) return hipSuccess;
    const int64_t gy = Nd < 65535 ? Nd : 65535;
    dim3 grid((unsigned)((Nm2 + kBlk - 1) / kBlk), (unsigned)gy, (unsigned)((Nd + gy - 1) / gy));
    switch (elem_bytes) {
    case 8:
        hipLaunchKernelGGL(remap_cols_kernel<uint64_t>, grid, dim3(kBlk), 0, st,
                           (const uint64_t *)G, ld, (const uint64_t *)Gc, ldc, src, Nd, Nm2,
                           (uint64_t *)G2, ld2);
        break;
    case 4:
        hipLaunchKernelGGL(remap_cols_kernel<uint32_t>, grid, dim3(kBlk), 0, st,
                           (const uint32_t *)G, ld, (const uint32_t *)Gc, ldc, src, Nd, Nm2,
                           (uint32_t *)G2, ld2);
        break;
    default:
        hipLaunchKernelGGL(remap_cols_kernel<uint16_t>, grid, dim3(kBlk), 0, st,
                           (const uint16_t *)G, ld, (const uint16_t *)Gc, ldc, src, Nd, Nm2,
                           (uint16_t *)G2, ld2);
        break;
    }
    return hipGetLastError();
}

// Column sums of squares of G into `part` ([t_chunks][ld] partials, summed in
// chunk order by the consumer).
hipError_t launch_col_sumsq(const void *G, gra_dtype dt, int64_t Nd, int64_t ld,
                            const GemvShape &s, double *part, hipStream_t st)
{
    int V = dt == GRA_F64 ? 2 : (dt == GRA_F32 ? 4 : 8);
    int64_t bx = (ld + kBlk * V - 1) / (kBlk * V);
    int64_t rpc = (Nd + s.t_chunks - 1) / s.t_chunks;
    dim3 grid((unsigned)bx, (unsigned)s.t_chunks), block(kBlk);
    switch (dt) {
    case GRA_F64:
        hipLaunchKernelGGL((gemv_t_kernel<double, 8, true, true>), grid, block, 0, st,
                           (const double *)G, Nd, ld, rpc, nullptr, part, nullptr);
        break;
    case GRA_F32:
        hipLaunchKernelGGL((gemv_t_kernel<float, 8, true, true>), grid, block, 0, st,
                           (const float *)G, Nd, ld, rpc, nullptr, part, nullptr);
        break;
    default:
        hipLaunchKernelGGL((gemv_t_kernel<__hip_bfloat16, 8, true, true>), grid, block, 0, st,
                           (const __hip_bfloat16 *)G, Nd, ld, rpc, nullptr, part, nullptr);
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


namespace {
// One conversion per element, rounding as the assembly stores (prism_kernels.hip):
// FP32 = round(double), BF16 = round(round(double) to FP32).
template <typename T> __device__ inline T cvt_store(double v);
template <> __device__ inline float cvt_store<float>(double v) { return (float)v; }
template <> __device__ inline __hip_bfloat16 cvt_store<__hip_bfloat16>(double v)
{
    return __float2bfloat16((float)v);
}
template <typename T>
__global__ void convert_G_kernel(const double *__restrict__ G, int64_t n, T *__restrict__ out)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = cvt_store<T>(__builtin_nontemporal_load(G + i));
}
}  // namespace

hipError_t launch_convert_G(const double *G, int64_t n, gra_dtype dt, void *out, hipStream_t st)
{
    const unsigned grid = 4096;
    if (dt == GRA_F32)
        hipLaunchKernelGGL(convert_G_kernel<float>, dim3(grid), dim3(256), 0, st, G, n,
                           static_cast<float *>(out));
    else
        hipLaunchKernelGGL(convert_G_kernel<__hip_bfloat16>, dim3(grid), dim3(256), 0, st, G, n,
                           static_cast<__hip_bfloat16 *>(out));
    return hipGetLastError();
}

}  // namespace gra
