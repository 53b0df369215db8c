// Multi-right-hand-side products Y = G X (Nd x k) and Z = G^T R (Nm x k) on
// the matrix cores: the batched form of Fwd::G_vec_mul / GT_vec_mul
// (Fwd.cpp:114-147, 80-113) for k vectors at once, so that one pass over G
// serves k requests on the same resident G.
//
// MFMA flavours: FP64 G with v_mfma_f64_16x16x4 (FP64 in, FP64 accumulate);
// FP32 and BF16 G with v_mfma_f32_16x16x4 (the BF16 elements widened to FP32
// exactly, the vectors rounded to FP32), whose FP32 accumulators are added
// into FP64 registers after every 16-32 products, so that FP32 rounding acts
// on short sums only.  Lane layouts (checked on the device): operand A lane l
// = row l % 16, k l / 16; operand B lane l = k l / 16, column l % 16; result
// element i of lane l = row 4 (l / 16) + i (f32) or l / 16 + 4 i (f64),
// column l % 16.
//
// Work split: G x -- a workgroup owns 16 rows of G and up to four 16-wide
// tiles of right-hand sides over one split of the summation range, staging G
// through LDS; G^T r -- a wave owns 16 E columns of G and one tile of
// right-hand sides over one split of the rows, every lane loading 16 bytes
// of one row (coalesced 256-byte row segments).  The splits' FP64 partial
// tiles are summed in split order by a second kernel, so the products are
// bit-for-bit repeatable.
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include "engine.hpp"

namespace gra {

namespace {

constexpr int kWave = 64;
constexpr int kBlk = 256;
constexpr int kWavesPerBlk = kBlk / kWave;

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));
typedef uint32_t v4u __attribute__((ext_vector_type(4)));

// 16 bytes of G as E elements widened to the MFMA input type.
template <typename T> struct Gload;
template <> struct Gload<double> {
    static constexpr int E = 2;
    typedef double In;
    __device__ static inline void run(const double *p, double *o)
    {
        v2d v = __builtin_nontemporal_load(reinterpret_cast<const v2d *>(p));
        o[0] = v.x; o[1] = v.y;
    }
    __device__ static inline void from_lds(const double *p, double *o)
    {
        v2d v = *reinterpret_cast<const v2d *>(p);
        o[0] = v.x; o[1] = v.y;
    }
};
template <> struct Gload<float> {
    static constexpr int E = 4;
    typedef float In;
    __device__ static inline void run(const float *p, float *o)
    {
        v4f v = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(p));
        o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
    }
    __device__ static inline void from_lds(const float *p, float *o)
    {
        v4f v = *reinterpret_cast<const v4f *>(p);
        o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
    }
};
template <> struct Gload<__hip_bfloat16> {
    static constexpr int E = 8;
    typedef float In;
    __device__ static inline void widen(v4u v, float *o)
    {
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            o[2 * q] = __uint_as_float(w[q] << 16);
            o[2 * q + 1] = __uint_as_float(w[q] & 0xffff0000u);
        }
    }
    __device__ static inline void run(const __hip_bfloat16 *p, float *o)
    {
        widen(__builtin_nontemporal_load(reinterpret_cast<const v4u *>(p)), o);
    }
    __device__ static inline void from_lds(const __hip_bfloat16 *p, float *o)
    {
        widen(*reinterpret_cast<const v4u *>(p), o);
    }
};

// One MFMA step and the result layout of the two flavours.
__device__ inline void mfma(double a, double b, v4d &c)
{
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline void mfma(float a, float b, v4f &c)
{
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
template <typename In> __device__ inline int out_row(int lane, int i);
template <> __device__ inline int out_row<double>(int lane, int i) { return lane / 16 + 4 * i; }
template <> __device__ inline int out_row<float>(int lane, int i) { return 4 * (lane / 16) + i; }

template <typename In> struct Acc;
template <> struct Acc<double> { typedef v4d V; };
template <> struct Acc<float> { typedef v4f V; };

// Y tile through LDS: the workgroup's four waves share 16 rows of G; each
// chunk of 256 columns is read with coalesced 16-byte loads (a wave covers
// 1 KB of one row per instruction), held in registers while the previous
// chunk's MFMAs run from LDS, and wave w multiplies columns 64w..64w+63 of
// the chunk.  The four wave sums of each output element meet in LDS and
// are added in wave order.  Rows are padded by 16 bytes in LDS so that the
// 16 rows of an MFMA operand fall on different banks.  NCT tiles of 16
// right-hand sides (tiles blockIdx.z * NCT ..) share each pass over G.
template <typename T, int NCT>
__global__ void __launch_bounds__(kBlk)
gm_lds_kernel(const T *__restrict__ G, int64_t Nd, int64_t ld, int64_t kspl,
              const typename Gload<T>::In *__restrict__ X, int64_t ldx, int ncol,
              double *__restrict__ part)
{
    typedef typename Gload<T>::In In;
    constexpr int E = Gload<T>::E;              // elements per 16 bytes
    constexpr int U = 16 / E;                   // 16-byte steps per wave quarter (64 columns)
    constexpr int kChunk = 256;                 // columns per chunk
    constexpr int CH = kChunk / E;              // 16-byte pieces per row and chunk
    constexpr int PER = 16 * CH / kBlk;         // pieces per thread and chunk
    constexpr int ROWB = kChunk * sizeof(T) + 16;   // padded LDS row, bytes
    __shared__ __attribute__((aligned(16))) unsigned char tile[16 * ROWB];
    __shared__ double red[kWavesPerBlk][16][16];
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int64_t rt = blockIdx.x;
    const int split = blockIdx.y, ct0 = blockIdx.z * NCT;
    const int r = lane & 15, kk = lane >> 4;
    const int64_t k0 = (int64_t)split * kspl;
    const int64_t k1 = k0 + kspl < ld ? k0 + kspl : ld;
    const In *x[NCT];
#pragma unroll
    for (int t = 0; t < NCT; ++t) x[t] = X + (int64_t)((ct0 + t) * 16 + r) * ldx;
    // this thread's pieces: q = threadIdx.x + kBlk * j -> row q / CH, piece q % CH
    // (k1 - k0 is a multiple of 64, not of the chunk: the last chunk of a
    // range may be partial; pieces past k1 are not loaded and wave quarters
    // past k1 skip their MFMAs)
    const T *src[PER];
    int dst[PER], off[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int q = threadIdx.x + kBlk * j, row = q / CH, pc = q % CH;
        const int64_t grow = rt * 16 + row < Nd ? rt * 16 + row : Nd - 1;   // clamp: dropped
        src[j] = G + grow * ld + (int64_t)pc * E;
        dst[j] = row * ROWB + pc * 16;
        off[j] = pc * E;
    }
    v4u buf[PER];
    typename Acc<In>::V acc[NCT];
    double a64[NCT][4];
#pragma unroll
    for (int t = 0; t < NCT; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            acc[t][i] = 0;
            a64[t][i] = 0.0;
        }
    int64_t kc = k0;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        buf[j] = v4u{0u, 0u, 0u, 0u};
        if (kc + off[j] < k1)
            buf[j] = __builtin_nontemporal_load(reinterpret_cast<const v4u *>(src[j] + kc));
    }
    for (; kc < k1; kc += kChunk) {
        __syncthreads();                        // previous chunk's LDS reads done
#pragma unroll
        for (int j = 0; j < PER; ++j) *reinterpret_cast<v4u *>(tile + dst[j]) = buf[j];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            if (kc + kChunk + off[j] < k1)
                buf[j] = __builtin_nontemporal_load(
                    reinterpret_cast<const v4u *>(src[j] + kc + kChunk));
        }
        const int cbase = 64 * wv;              // this wave's 64 columns of the chunk
        if (kc + cbase >= k1) continue;         // uniform per wave
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int col = cbase + u * 4 * E + E * kk;
            In gv[E];
            Gload<T>::from_lds(reinterpret_cast<const T *>(tile + r * ROWB) + col, gv);
#pragma unroll
            for (int t = 0; t < NCT; ++t) {
                In xv[E];
#pragma unroll
                for (int e = 0; e < E; ++e) xv[e] = x[t][kc + col + e];
#pragma unroll
                for (int e = 0; e < E; ++e) mfma(gv[e], xv[e], acc[t]);
                if (sizeof(In) == 4) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        a64[t][i] += (double)acc[t][i];
                        acc[t][i] = 0;
                    }
                }
            }
        }
    }
    if (sizeof(In) == 8) {
#pragma unroll
        for (int t = 0; t < NCT; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) a64[t][i] = (double)acc[t][i];
    }
    const int orow = threadIdx.x / 16, ocol = threadIdx.x % 16;   // 256 threads = 16 x 16
    const int64_t rr = rt * 16 + orow;
#pragma unroll
    for (int t = 0; t < NCT; ++t) {
        __syncthreads();                        // red free again
#pragma unroll
        for (int i = 0; i < 4; ++i) red[wv][out_row<In>(lane, i)][lane & 15] = a64[t][i];
        __syncthreads();
        if (rr < Nd) {
            double s = red[0][orow][ocol];
#pragma unroll
            for (int q = 1; q < kWavesPerBlk; ++q) s += red[q][orow][ocol];
            part[((int64_t)split * ncol + (ct0 + t) * 16 + ocol) * Nd + rr] = s;
        }
    }
}

// Z tile: G columns m0 = mt*16E .. +16E (lane group r takes columns m0+E r..
// +E-1, one MFMA per element phase e), right-hand sides ct*16.., summation
// over G rows [split*rspl, min(+rspl, Nd)).  R: the right-hand sides in MFMA
// input type, column c at R + c*ldr (zero padded to whole tiles of 16).
// part[(split * ncol + c) * ld + column].
template <typename T>
__global__ void __launch_bounds__(kBlk)
gt_kernel(const T *__restrict__ G, int64_t Nd, int64_t Nm, int64_t ld, int64_t rspl,
          const typename Gload<T>::In *__restrict__ R, int64_t ldr, int64_t col_tiles, int ncol,
          double *__restrict__ part)
{
    typedef typename Gload<T>::In In;
    constexpr int E = Gload<T>::E;
    constexpr int U = 4;                      // row steps per iteration (4 rows each)
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t mt = (int64_t)blockIdx.x * kWavesPerBlk + (threadIdx.x / kWave);
    if (mt >= col_tiles) return;
    const int split = blockIdx.y, ct = blockIdx.z;
    const int r = lane & 15, kk = lane >> 4;
    const int64_t m0 = mt * 16 * E;
    const int64_t mcol = m0 + (int64_t)E * r;         // this lane's first G column
    const bool colok = mcol < ld;                      // ld is a multiple of 64 >= E
    const In *rv = R + (int64_t)(ct * 16 + r) * ldr;
    const int64_t r0 = (int64_t)split * rspl;
    const int64_t r1 = r0 + rspl < Nd ? r0 + rspl : Nd;
    typename Acc<In>::V acc[E];
    double a64[E][4];
#pragma unroll
    for (int e = 0; e < E; ++e)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            acc[e][i] = 0;
            a64[e][i] = 0.0;
        }
    for (int64_t rb = r0 + kk; rb < r1; rb += 4 * U) {
        In gv[U][E], b[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t row = rb + 4 * u;
            // the lane's A rows (G columns mcol..) and its B column (right-hand
            // side r) are independent: a lane past the last G column still
            // supplies its right-hand side value
            b[u] = row < r1 ? rv[row] : 0;
            if (colok && row < r1) {
                Gload<T>::run(G + row * ld + mcol, gv[u]);
            } else {
#pragma unroll
                for (int e = 0; e < E; ++e) gv[u][e] = 0;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int e = 0; e < E; ++e) mfma(gv[u][e], b[u], acc[e]);
        if (sizeof(In) == 4) {                // FP32 partial -> FP64 every 16 products
#pragma unroll
            for (int e = 0; e < E; ++e)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    a64[e][i] += (double)acc[e][i];
                    acc[e][i] = 0;
                }
        }
    }
    if (sizeof(In) == 8) {
#pragma unroll
        for (int e = 0; e < E; ++e)
#pragma unroll
            for (int i = 0; i < 4; ++i) a64[e][i] = (double)acc[e][i];
    }
    const int c = ct * 16 + (lane & 15);
    double *out = part + ((int64_t)split * ncol + c) * ld;
#pragma unroll
    for (int e = 0; e < E; ++e)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t m = m0 + (int64_t)E * out_row<In>(lane, i) + e;
            if (m < Nm) out[m] = a64[e][i];
        }
}

// The right-hand sides (host layout: column c at src + c*n, n rows, k
// columns) in MFMA input type, padded with zeros to ldv rows and ncol columns.
template <typename In>
__global__ void pack_rhs_kernel(const double *__restrict__ src, int64_t n, int64_t k,
                                int64_t ldv, int ncol, In *__restrict__ dst)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)ncol * ldv) return;
    const int64_t c = i / ldv, row = i - c * ldv;
    dst[i] = (c < k && row < n) ? (In)src[c * n + row] : (In)0;
}

// out[c*n + i] = sum over splits, in split order, of part[(s*ncol + c)*ldp + i].
__global__ void sum_splits_kernel(const double *__restrict__ part, int splits, int ncol,
                                  int64_t ldp, int64_t n, int64_t k, double *__restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= k * n) return;
    const int64_t c = t / n, i = t - c * n;
    double s = part[c * ldp + i];
    for (int q = 1; q < splits; ++q) s += part[((int64_t)q * ncol + c) * ldp + i];
    out[t] = s;
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

template <typename T>
hipError_t mm_n(const void *G, int64_t Nd, int64_t ld, const void *X, int64_t ldx, int ncol,
                int splits, int64_t kspl, double *part, hipStream_t st)
{
    const int64_t row_tiles = (Nd + 15) / 16;
    const int ctiles = ncol / 16;
    // right-hand-side tiles per pass over G: 4 when they divide evenly, else 2 or 1
    const int nct = ctiles % 4 == 0 ? 4 : (ctiles % 2 == 0 ? 2 : 1);
    dim3 grid((unsigned)row_tiles, (unsigned)splits, (unsigned)(ctiles / nct));
    typedef typename Gload<T>::In In;
#define GRA_GM(N)                                                                              \
    hipLaunchKernelGGL((gm_lds_kernel<T, N>), grid, dim3(kBlk), 0, st, (const T *)G, Nd, ld,   \
                       kspl, (const In *)X, ldx, ncol, part)
    if (nct == 4) GRA_GM(4);
    else if (nct == 2) GRA_GM(2);
    else GRA_GM(1);
#undef GRA_GM
    return hipGetLastError();
}

template <typename T>
hipError_t mm_t(const void *G, int64_t Nd, int64_t Nm, int64_t ld, const void *R, int64_t ldr,
                int ncol, int splits, int64_t rspl, double *part, hipStream_t st)
{
    const int64_t col_tiles = (ld + 16 * Gload<T>::E - 1) / (16 * Gload<T>::E);
    dim3 grid(nblk(col_tiles, kWavesPerBlk), (unsigned)splits, (unsigned)(ncol / 16));
    hipLaunchKernelGGL(gt_kernel<T>, grid, dim3(kBlk), 0, st, (const T *)G, Nd, Nm, ld, rspl,
                       (const typename Gload<T>::In *)R, ldr, col_tiles, ncol, part);
    return hipGetLastError();
}

}  // namespace

int gemm_elems_per_load(gra_dtype dt) { return dt == GRA_F64 ? 2 : (dt == GRA_F32 ? 4 : 8); }

// Splits of the summation range: about 3,000 waves in all.
MatMulPlan plan_mat_mul(gra_dtype dt, bool trans, int64_t Nd, int64_t Nm, int64_t ld, int64_t k)
{
    MatMulPlan p;
    p.ncol = (int)((k + 15) / 16 * 16);
    const int64_t ctiles = p.ncol / 16;
    if (!trans) {
        const int64_t tiles = (Nd + 15) / 16 * ctiles;
        int64_t s = (3000 + tiles - 1) / tiles;
        const int64_t smax = ld / 256 > 0 ? ld / 256 : 1;     // at least 256 columns each
        s = s < 1 ? 1 : (s > smax ? smax : s);
        int64_t kspl = (ld + s - 1) / s;
        p.span = (kspl + 63) / 64 * 64;
        p.splits = (int)((ld + p.span - 1) / p.span);
        p.ldp = Nd;
    } else {
        const int E = gemm_elems_per_load(dt);
        const int64_t tiles = (ld + 16 * E - 1) / (16 * E) * ctiles;
        int64_t s = (3000 + tiles - 1) / tiles;
        const int64_t smax = Nd / 64 > 0 ? Nd / 64 : 1;       // at least 64 rows each
        s = s < 1 ? 1 : (s > smax ? smax : s);
        p.span = (Nd + s - 1) / s;
        p.splits = (int)((Nd + p.span - 1) / p.span);
        p.ldp = ld;
    }
    return p;
}

hipError_t launch_pack_rhs(gra_dtype dt, const double *src, int64_t n, int64_t k, int64_t ldv,
                           int ncol, void *dst, hipStream_t st)
{
    const int64_t tot = (int64_t)ncol * ldv;
    if (dt == GRA_F64)
        hipLaunchKernelGGL(pack_rhs_kernel<double>, dim3(nblk(tot, kBlk)), dim3(kBlk), 0, st, src,
                           n, k, ldv, ncol, (double *)dst);
    else
        hipLaunchKernelGGL(pack_rhs_kernel<float>, dim3(nblk(tot, kBlk)), dim3(kBlk), 0, st, src,
                           n, k, ldv, ncol, (float *)dst);
    return hipGetLastError();
}

hipError_t launch_mat_mul(const void *G, gra_dtype dt, bool trans, int64_t Nd, int64_t Nm,
                          int64_t ld, const void *V, int64_t ldv, const MatMulPlan &p,
                          double *part, hipStream_t st)
{
    if (!trans) {
        switch (dt) {
        case GRA_F64: return mm_n<double>(G, Nd, ld, V, ldv, p.ncol, p.splits, p.span, part, st);
        case GRA_F32: return mm_n<float>(G, Nd, ld, V, ldv, p.ncol, p.splits, p.span, part, st);
        default:
            return mm_n<__hip_bfloat16>(G, Nd, ld, V, ldv, p.ncol, p.splits, p.span, part, st);
        }
    }
    switch (dt) {
    case GRA_F64: return mm_t<double>(G, Nd, Nm, ld, V, ldv, p.ncol, p.splits, p.span, part, st);
    case GRA_F32: return mm_t<float>(G, Nd, Nm, ld, V, ldv, p.ncol, p.splits, p.span, part, st);
    default:
        return mm_t<__hip_bfloat16>(G, Nd, Nm, ld, V, ldv, p.ncol, p.splits, p.span, part, st);
    }
}

hipError_t launch_sum_splits(const double *part, const MatMulPlan &p, int64_t n, int64_t k,
                             double *out, hipStream_t st)
{
    hipLaunchKernelGGL(sum_splits_kernel, dim3(nblk(k * n, kBlk)), dim3(kBlk), 0, st, part,
                       p.splits, p.ncol, p.ldp, n, k, out);
    return hipGetLastError();
}

}  // namespace gra
