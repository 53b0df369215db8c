// Multi-right-hand-side products Y = G X (Nd x k) and Z = G^T R (Nm x k) on
// the matrix cores: the batched form of Fwd::G_vec_mul / GT_vec_mul
// (Fwd.cpp:114-147, 80-113) for k vectors at once, so that one pass over G
// serves k requests on the same resident G.
//
// MFMA flavours per storage type:
// * FP64 G: v_mfma_f64_16x16x4 (FP64 in, FP64 accumulate);
// * FP32 G: v_mfma_f32_16x16x4 with the vectors rounded to FP32;
// * BF16 G: the BF16 matrix cores with the vectors split exactly into three
//   BF16 parts (x32 = hi + mid + lo, 8 significand bits each), so that the
//   products are those of the FP32-rounded vectors: v_mfma_f32_16x16x32_bf16
//   for G X (a lane's 16 bytes of a row are its 8 k values as they lie in
//   memory), v_mfma_f32_16x16x16bf16_1k for G^T R (four G rows per lane,
//   regrouped from row segments by byte permutes).
// FP32 accumulators are added into FP64 registers after at most 32 products
// of the hi part (plus the mid / lo parts' products, 2^-8 and 2^-16 smaller),
// so that FP32 rounding acts on short sums only.
//
// Operand lane layouts (16x16 tiles; l = lane): A lane l = row l % 16,
// k group l / 16; B lane l = k group l / 16, column l % 16; a lane holds 1
// (x4), 4 (x16) or 8 (x32) consecutive k of its group.  Result element i of
// lane l = row 4 (l / 16) + i (f32) or l / 16 + 4 i (f64), column l % 16.
//
// Work split:
// * G X (gmx_kernel): a workgroup of four waves owns 64 rows (16 per wave)
//   and one 16-wide tile of right-hand sides over one split of the columns.
//   G is read one chunk ahead of use in whole 256-byte row segments and
//   turned into the A-operand layout through a small wave-private LDS tile;
//   the chunk's X values (shared by the 64 rows) are staged in LDS,
//   double-buffered, one barrier per chunk.
// * G^T R (gtx_kernel, gtb_kernel): a wave owns a strip of G columns (16 B
//   per lane and row, 256-byte row segments) and one tile of right-hand
//   sides over one split of the rows; the next 16-row step's loads are issued
//   before the current step's MFMAs.  Every lane runs the same number of
//   steps (rows past the split are zero-filled): an MFMA reads the operands
//   of all 64 lanes, so a lane that left a divergent loop would contribute
//   stale registers.
// The splits' FP64 partial tiles are summed in split order by a second
// kernel, so the products are bit-for-bit repeatable.
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <type_traits>

#include "engine.hpp"

namespace gra {

namespace {

constexpr int kWave = 64;
constexpr int kBlk = 256;
constexpr int kWavesPerBlk = kBlk / kWave;

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef uint32_t v4u __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef short v4s __attribute__((ext_vector_type(4)));

// Element q of a 16-byte vector as FP64 (two per vector) or FP32 (four).
__device__ inline double el64(const v4u &v, int q)
{
    return q == 0 ? __hiloint2double((int)v.y, (int)v.x) : __hiloint2double((int)v.w, (int)v.z);
}
__device__ inline float el32(const v4u &v, int q)
{
    return __uint_as_float(q == 0 ? v.x : (q == 1 ? v.y : (q == 2 ? v.z : v.w)));
}

// Per storage type: G elements per 16 bytes (E), the vectors' element type
// and number of parts, G X chunk width (columns; 16 bytes per lane and piece,
// 8 pieces per chunk).
template <typename T> struct MM;
template <> struct MM<double> {
    typedef double In;
    static constexpr int E = 2, NPART = 1, KC = 64;
};
template <> struct MM<float> {
    typedef float In;
    static constexpr int E = 4, NPART = 1, KC = 128;
};
template <> struct MM<__hip_bfloat16> {
    typedef __bf16 In;
    static constexpr int E = 8, NPART = 3, KC = 256;
};

// Y = G X.  Lane (r, kk) of wave w owns row 64 b + 16 w + r in the MFMA
// layout; its piece j of a chunk is the 16 bytes at chunk column cofs(j, kk).
// G is loaded with non-temporal loads of whole 256-byte row segments (load
// instruction j: rows 4 (j & 3) + kk, 16-byte piece 16 (j >> 2) + r, i.e. four
// rows per instruction) and turned into the MFMA layout through a wave-private
// LDS tile, half a chunk at a time (piece p of row r at slot p ^ r: conflict-
// free stores and loads).  Loading the MFMA layout directly touches 16 rows
// per instruction with 64-byte segments, which held the kernel at 6.1 TB/s
// with or without non-temporal loads; the staged form reads 6.5-6.7 TB/s on
// the 90,000 x 131,072 G.  The B operand of the same k values is read from
// the LDS copy of X: rows = right-hand sides, columns = the chunk.
// Xp: NPART parts of ncol rows of ldx elements (part p, right-hand side c at
// Xp + (p * ncol + c) * ldx).  part[(split * ncol + c) * Nd + row].
template <typename T>
__global__ void __launch_bounds__(kBlk)
gmx_kernel(const T *__restrict__ G, int64_t Nd, int64_t ld, int64_t kspl,
           const typename MM<T>::In *__restrict__ Xp, int64_t ldx, int ncol,
           double *__restrict__ part)
{
    typedef typename MM<T>::In In;
    constexpr int E = MM<T>::E, KC = MM<T>::KC, NPART = MM<T>::NPART;
    constexpr int PIECES = KC / (4 * E);                 // 8 (16 bytes each, 512 B of a row per chunk)
    constexpr int XROW = KC * (int)sizeof(In) + 16;      // LDS bytes per right-hand side (padded)
    constexpr int XT = 16 * XROW;                        // per part
    constexpr int XPC = KC * (int)sizeof(In) / 16;       // 16-byte X pieces per right-hand side
    constexpr int XPER = 16 * XPC / kBlk;                // per thread and part
    constexpr int XV = 16 / (int)sizeof(In);             // X elements per 16 bytes
    constexpr bool F64 = sizeof(In) == 8;
    constexpr bool BF = NPART == 3;
    static_assert(PIECES == 8, "512-byte row chunks");
    __shared__ __attribute__((aligned(16))) unsigned char xl[2][NPART][XT];
    __shared__ __attribute__((aligned(16))) v4u gl[kWavesPerBlk][16 * 16];   // 16 rows x half a chunk
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int r = lane & 15, kk = lane >> 4;
    const int64_t rbase = (int64_t)blockIdx.x * 64 + wv * 16;
    const int split = blockIdx.y, ct = blockIdx.z;
    const int64_t k0 = (int64_t)split * kspl;
    const int64_t k1 = k0 + kspl < ld ? k0 + kspl : ld;
    // rows read by load instruction j: 4 (j & 3) + kk (clamped rows are not stored)
    const T *lrow[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int64_t row = rbase + 4 * q + kk;
        lrow[q] = G + (row < Nd ? row : Nd - 1) * ld;
    }
    v4u *tile = gl[wv];

    auto cpiece = [&](int j) { return (j >> 1) * 8 + kk * 2 + (j & 1); };
    // Whole chunks are loaded without masks (a masked load makes the
    // compiler wait for every load in flight, the prefetched chunk's too);
    // a partial last chunk (the end of G) is loaded with masks after them.
    const int64_t nfull = (k1 - k0) / KC;
    v4u g[2][PIECES];
    v4u xs[2][NPART][XPER];
    auto load_g = [&](int64_t kc, v4u (&dst)[PIECES], bool masked) {
#pragma unroll
        for (int j = 0; j < PIECES; ++j) {
            const int64_t c = kc + (int64_t)(16 * (j >> 2) + r) * E;
            const v4u *src = reinterpret_cast<const v4u *>(lrow[j & 3] + c);
            if (!masked) dst[j] = __builtin_nontemporal_load(src);
            else dst[j] = c < k1 ? __builtin_nontemporal_load(src) : v4u{0u, 0u, 0u, 0u};
        }
    };
    // half h of a chunk (pieces 16 h .. 16 h + 15 of every row) from the
    // loaded registers into the MFMA layout: dst[i] = piece cpiece(4 h + i)
    // of row r.  The tile is private to the wave; its stores and loads are
    // ordered at wavefront scope.
    auto stage = [&](const v4u (&src)[PIECES], int h, v4u (&dst)[PIECES / 2]) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = 4 * i + kk;
            tile[row * 16 + (r ^ row)] = src[4 * h + i];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int pc = cpiece(4 * h + i) - 16 * h;
            dst[i] = tile[r * 16 + (pc ^ r)];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    auto load_x = [&](int64_t kc, v4u (&dst)[NPART][XPER], bool masked) {
#pragma unroll
        for (int p = 0; p < NPART; ++p)
#pragma unroll
            for (int u = 0; u < XPER; ++u) {
                const int q = threadIdx.x + kBlk * u, n = q / XPC, pc = q % XPC;
                const int64_t c = kc + (int64_t)pc * XV;
                const In *src = Xp + ((int64_t)p * ncol + ct * 16 + n) * ldx + c;
                if (!masked) dst[p][u] = *reinterpret_cast<const v4u *>(src);
                else dst[p][u] = c < k1 ? *reinterpret_cast<const v4u *>(src) : v4u{0u, 0u, 0u, 0u};
            }
    };
    auto store_x = [&](int b, const v4u (&src)[NPART][XPER]) {
#pragma unroll
        for (int p = 0; p < NPART; ++p)
#pragma unroll
            for (int u = 0; u < XPER; ++u) {
                const int q = threadIdx.x + kBlk * u, n = q / XPC, pc = q % XPC;
                *reinterpret_cast<v4u *>(&xl[b][p][n * XROW + pc * 16]) = src[p][u];
            }
    };

    v4d acc64 = {0.0, 0.0, 0.0, 0.0};
    v4f acc32 = {0.0f, 0.0f, 0.0f, 0.0f}, accml = {0.0f, 0.0f, 0.0f, 0.0f};
    double a64[4] = {0.0, 0.0, 0.0, 0.0};
    auto flush = [&](v4f &acc) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            a64[i] += (double)acc[i];
            acc[i] = 0.0f;
        }
    };
    auto compute = [&](int b, const v4u (&graw)[PIECES]) {
        v4u gh[PIECES / 2];
#pragma unroll
        for (int j = 0; j < PIECES; ++j) {
            if ((j & 3) == 0) stage(graw, j >> 2, gh);
            const v4u gv = gh[j & 3];
            const int off = r * XROW + cpiece(j) * 16;
            if constexpr (BF) {
                const v8bf a = __builtin_bit_cast(v8bf, gv);
                const v8bf xh = *reinterpret_cast<const v8bf *>(&xl[b][0][off]);
                const v8bf xm = *reinterpret_cast<const v8bf *>(&xl[b][1][off]);
                const v8bf xo = *reinterpret_cast<const v8bf *>(&xl[b][2][off]);
                acc32 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, xh, acc32, 0, 0, 0);
                accml = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, xm, accml, 0, 0, 0);
                accml = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, xo, accml, 0, 0, 0);
                flush(acc32);                          // 32 products of the hi part
            } else {
                const v4u xv = *reinterpret_cast<const v4u *>(&xl[b][0][off]);
                if constexpr (F64) {
#pragma unroll
                    for (int e = 0; e < E; ++e)
                        acc64 = __builtin_amdgcn_mfma_f64_16x16x4f64(el64(gv, e), el64(xv, e),
                                                                      acc64, 0, 0, 0);
                } else {
#pragma unroll
                    for (int e = 0; e < E; ++e)
                        acc32 = __builtin_amdgcn_mfma_f32_16x16x4f32(el32(gv, e), el32(xv, e),
                                                                      acc32, 0, 0, 0);
                    if (j & 1) flush(acc32);           // 32 products
                }
            }
        }
        if constexpr (BF) flush(accml);                // 512 products of the mid / lo parts
    };

    // Whole chunks: chunk ch's G in g[ch & 1] and its X in xs[ch & 1] are
    // loaded one chunk ahead (X first, so that storing X to LDS waits only for
    // loads issued before that chunk's G loads); X of chunk ch is in xl[ch & 1].
    auto xchunk = [&](int64_t ch) { return k0 + (ch < nfull ? ch : nfull - 1) * KC; };
    int b = 0;
    if (nfull > 0) {
        load_x(xchunk(0), xs[0], false);
        store_x(0, xs[0]);
        load_g(k0, g[0], false);
        __syncthreads();
        for (int64_t ch = 0; ch < nfull; ch += 2) {
            load_x(xchunk(ch + 1), xs[1], false);
            if (ch + 1 < nfull) load_g(k0 + (ch + 1) * KC, g[1], false);
            compute(b, g[0]);
            if (ch + 1 >= nfull) break;
            store_x(b ^ 1, xs[1]);                     // xl[b ^ 1] was last read before the previous barrier
            __syncthreads();
            b ^= 1;
            load_x(xchunk(ch + 2), xs[0], false);
            if (ch + 2 < nfull) load_g(k0 + (ch + 2) * KC, g[0], false);
            compute(b, g[1]);
            if (ch + 2 >= nfull) break;
            store_x(b ^ 1, xs[0]);
            __syncthreads();
            b ^= 1;
        }
    }
    if (k0 + nfull * KC < k1) {                        // partial last chunk
        const int64_t kc = k0 + nfull * KC;
        load_x(kc, xs[0], true);
        load_g(kc, g[0], true);
        __syncthreads();                               // every wave is done with xl[b ^ 1]
        store_x(b ^ 1, xs[0]);
        __syncthreads();
        compute(b ^ 1, g[0]);
    }
    if constexpr (F64) {
#pragma unroll
        for (int i = 0; i < 4; ++i) a64[i] = acc64[i];
    } else {
        flush(acc32);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int orow = F64 ? (lane >> 4) + 4 * i : 4 * (lane >> 4) + i;
        const int64_t rr = rbase + orow;
        if (rr < Nd) part[((int64_t)split * ncol + ct * 16 + r) * Nd + rr] = a64[i];
    }
}

// Z = G^T R for FP64 / FP32 G.  Wave wt owns columns m0 = wt * MC * 16 E ..
// (G read with non-temporal loads: measured 15 % faster here than plain loads);
// for column tile mc, lane (r, kk) loads 16 B (E columns) of rows kk, kk + 4,
// kk + 8, kk + 12 of a 16-row step and runs one MFMA per element (A = those
// G values, B = R at the lane's row).  Rp: right-hand-side tile t, row i,
// column n at Rp[(t * ldr + i) * 16 + n].  part[(split * ncol + c) * ld + m].
template <typename T, int MC>
__global__ void __launch_bounds__(kBlk)
gtx_kernel(const T *__restrict__ G, int64_t Nd, int64_t Nm, int64_t ld, int64_t rspl,
           const typename MM<T>::In *__restrict__ Rp, int64_t ldr, int64_t wave_tiles, int ncol,
           double *__restrict__ part)
{
    typedef typename MM<T>::In In;
    typedef typename std::conditional<sizeof(In) == 8, v4d, v4f>::type Acc;
    constexpr int E = MM<T>::E;
    constexpr int U = 4;                                 // 4-row groups per step
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wt = (int64_t)blockIdx.x * kWavesPerBlk + threadIdx.x / kWave;
    if (wt >= wave_tiles) return;                       // whole wave
    const int split = blockIdx.y, ct = blockIdx.z;
    const int r = lane & 15, kk = lane >> 4;
    const int64_t m0 = wt * MC * 16 * E;
    const In *rv = Rp + (int64_t)ct * ldr * 16 + r;
    const int64_t r0 = (int64_t)split * rspl;
    const int64_t r1 = r0 + rspl < Nd ? r0 + rspl : Nd;
    Acc acc[MC][E];
    double a64[MC][E][4];
#pragma unroll
    for (int mc = 0; mc < MC; ++mc)
#pragma unroll
        for (int e = 0; e < E; ++e)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                acc[mc][e][i] = 0;
                a64[mc][e][i] = 0.0;
            }
    typedef In V16 __attribute__((ext_vector_type(E)));
    In gv[2][U][MC][E], bv[2][U];
    auto load = [&](int64_t rb, int sl) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t row = rb + kk + 4 * u;
            bv[sl][u] = row < r1 ? rv[row * 16] : (In)0;
#pragma unroll
            for (int mc = 0; mc < MC; ++mc) {
                const int64_t col = m0 + mc * 16 * E + (int64_t)E * r;
                if (col < ld && row < r1) {
                    // loaded as E elements of In: read as four 32-bit words and
                    // reassembled, the compiler re-issued part of the load
                    const V16 v = __builtin_nontemporal_load(reinterpret_cast<const V16 *>(G + row * ld + col));
#pragma unroll
                    for (int e = 0; e < E; ++e) gv[sl][u][mc][e] = v[e];
                } else {
#pragma unroll
                    for (int e = 0; e < E; ++e) gv[sl][u][mc][e] = (In)0;
                }
            }
        }
    };
    auto compute = [&](int sl) {
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int mc = 0; mc < MC; ++mc)
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    if constexpr (sizeof(In) == 8)
                        acc[mc][e] = __builtin_amdgcn_mfma_f64_16x16x4f64(gv[sl][u][mc][e], bv[sl][u],
                                                                          acc[mc][e], 0, 0, 0);
                    else
                        acc[mc][e] = __builtin_amdgcn_mfma_f32_16x16x4f32(gv[sl][u][mc][e], bv[sl][u],
                                                                          acc[mc][e], 0, 0, 0);
                }
        if constexpr (sizeof(In) == 4) {                 // FP32 partial -> FP64 every 16 products
#pragma unroll
            for (int mc = 0; mc < MC; ++mc)
#pragma unroll
                for (int e = 0; e < E; ++e)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        a64[mc][e][i] += (double)acc[mc][e][i];
                        acc[mc][e][i] = 0;
                    }
        }
    };
    load(r0, 0);
    for (int64_t rb = r0; rb < r1; rb += 8 * U) {       // the same trip count in every lane
        load(rb + 4 * U, 1);
        compute(0);
        if (rb + 4 * U >= r1) break;
        load(rb + 8 * U, 0);
        compute(1);
    }
    if constexpr (sizeof(In) == 8) {
#pragma unroll
        for (int mc = 0; mc < MC; ++mc)
#pragma unroll
            for (int e = 0; e < E; ++e)
#pragma unroll
                for (int i = 0; i < 4; ++i) a64[mc][e][i] = acc[mc][e][i];
    }
    const int c = ct * 16 + r;
    double *out = part + ((int64_t)split * ncol + c) * ld;
#pragma unroll
    for (int mc = 0; mc < MC; ++mc)
#pragma unroll
        for (int e = 0; e < E; ++e)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int orow = sizeof(In) == 8 ? (lane >> 4) + 4 * i : 4 * (lane >> 4) + i;
                const int64_t m = m0 + mc * 16 * E + (int64_t)E * orow + e;
                if (m < Nm) out[m] = a64[mc][e][i];
            }
}

// Z = G^T R for BF16 G.  Wave wt owns the 128 columns m0 = 128 wt ..; lane
// (j, kq) loads 16 B (columns m0 + 8j .. + 7) of rows 4kq .. 4kq + 3 of a
// 16-row step, regroups them into eight 4-row column operands (the low /
// high halves of each 32-bit word) and multiplies each with the three parts
// of R.  MFMA row m' of column operand c is G column m0 + 8 m' + c.
// Rb: part p, tile t, rows 4q .. 4q + 3 of right-hand side n at
// Rb[p * pstride + (t * ldr4 + q) * 64 + 4 n]; ldr4 * 4 >= Nd rounded up to 16.
__global__ void __launch_bounds__(kBlk)
gtb_kernel(const __hip_bfloat16 *__restrict__ G, int64_t Nd, int64_t Nm, int64_t ld, int64_t rspl,
           const unsigned short *__restrict__ Rb, int64_t ldr4, int64_t pstride,
           int64_t wave_tiles, int ncol, double *__restrict__ part)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wt = (int64_t)blockIdx.x * kWavesPerBlk + threadIdx.x / kWave;
    if (wt >= wave_tiles) return;
    const int split = blockIdx.y, ct = blockIdx.z;
    const int j = lane & 15, kq = lane >> 4;
    const int64_t m0 = wt * 128;
    const int64_t r0 = (int64_t)split * rspl;           // rspl is a multiple of 16
    const int64_t r1 = r0 + rspl < Nd ? r0 + rspl : Nd;
    const unsigned short *rb0 = Rb + (int64_t)ct * ldr4 * 64 + j * 4;
    v4f acc[8];
    double a64[8][4];
#pragma unroll
    for (int c = 0; c < 8; ++c)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            acc[c][i] = 0.0f;
            a64[c][i] = 0.0;
        }
    v4u g[2][4];
    v4s bp[2][3];
    const int64_t col = m0 + 8 * j;
    // Whole 16-row steps of a wave inside G are loaded without masks (a
    // masked load makes the compiler wait for every load in flight: 4 % on
    // the 90,000 x 131,072 G; the FP64 / FP32 kernel measured no gain, its
    // loads stay masked); the rows past the last whole step with masks.
    // Non-temporal loads: 6.9 TB/s against 6.25 with plain loads on that G.
    auto load = [&](int64_t rb, int sl, bool masked) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t row = rb + 4 * kq + i;
            if (!masked)
                g[sl][i] = __builtin_nontemporal_load(reinterpret_cast<const v4u *>(G + row * ld + col));
            else
                g[sl][i] = (row < r1 && col < ld)
                               ? __builtin_nontemporal_load(reinterpret_cast<const v4u *>(G + row * ld + col))
                               : v4u{0u, 0u, 0u, 0u};
        }
        // (every step starts below r1, so q4 < ldr4; the clamp only guards)
        int64_t q4 = rb / 4 + kq;
        q4 = q4 < ldr4 - 1 ? q4 : ldr4 - 1;
#pragma unroll
        for (int p = 0; p < 3; ++p)
            bp[sl][p] = *reinterpret_cast<const v4s *>(rb0 + p * pstride + q4 * 64);
    };
    int step = 0;
    auto compute = [&](int sl) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t w0 = g[sl][0][q], w1 = g[sl][1][q], w2 = g[sl][2][q], w3 = g[sl][3][q];
            const uint32_t lo01 = (w0 & 0xffffu) | (w1 << 16), lo23 = (w2 & 0xffffu) | (w3 << 16);
            const uint32_t hi01 = (w0 >> 16) | (w1 & 0xffff0000u), hi23 = (w2 >> 16) | (w3 & 0xffff0000u);
            const v4s ae = __builtin_bit_cast(v4s, (unsigned long long)lo01 | ((unsigned long long)lo23 << 32));
            const v4s ao = __builtin_bit_cast(v4s, (unsigned long long)hi01 | ((unsigned long long)hi23 << 32));
            // lo and mid parts first, hi last: the small terms are added
            // before the accumulator grows
#pragma unroll
            for (int p = 2; p >= 0; --p) {
                acc[2 * q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ae, bp[sl][p], acc[2 * q], 0, 0, 0);
                acc[2 * q + 1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ao, bp[sl][p], acc[2 * q + 1], 0, 0, 0);
            }
        }
        if (step & 1) {                                  // 32 products of the hi part
#pragma unroll
            for (int c = 0; c < 8; ++c)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    a64[c][i] += (double)acc[c][i];
                    acc[c][i] = 0.0f;
                }
        }
        ++step;
    };
    const int64_t nfull = m0 + 128 <= ld ? (r1 - r0) / 16 : 0;
    auto step_row = [&](int64_t st) { return r0 + (st < nfull ? st : nfull - 1) * 16; };
    if (nfull > 0) {
        load(r0, 0, false);
        for (int64_t st = 0; st < nfull; st += 2) {   // the same trip count in every lane
            load(step_row(st + 1), 1, false);
            compute(0);
            if (st + 1 >= nfull) break;
            load(step_row(st + 2), 0, false);
            compute(1);
        }
    }
    for (int64_t rb = r0 + nfull * 16; rb < r1; rb += 16) {
        load(rb, 0, true);
        compute(0);
    }
    const int n = ct * 16 + j;
    double *out = part + ((int64_t)split * ncol + n) * ld;
#pragma unroll
    for (int c = 0; c < 8; ++c)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t m = m0 + 8 * (4 * kq + i) + c;
            if (m < Nm) out[m] = a64[c][i] + (double)acc[c][i];
        }
}

// x (FP64) as FP32, and as three BF16 parts whose sum is that FP32 value
// exactly (8 significand bits each; each remainder is exact in FP32).
__device__ inline void split3(double v, __bf16 &h, __bf16 &m, __bf16 &o)
{
    const float f = (float)v;
    h = (__bf16)f;
    const float r1 = f - (float)h;
    m = (__bf16)r1;
    o = (__bf16)(r1 - (float)m);
}

// G X right-hand sides (host layout: column c at src + c*n, n rows, k
// columns) padded with zeros to ncol columns of ldv: FP64 / FP32 rows
// (NPART 1) or three BF16 parts of ncol * ldv each.
template <typename In, int NPART>
__global__ void pack_x_kernel(const double *__restrict__ src, int64_t n, int64_t k, int64_t ldv,
                              int ncol, In *__restrict__ dst)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)ncol * ldv;
    if (i >= per) return;
    const int64_t c = i / ldv, row = i - c * ldv;
    const double v = (c < k && row < n) ? src[c * n + row] : 0.0;
    if constexpr (NPART == 1) {
        dst[i] = (In)v;
    } else {
        __bf16 h, m, o;
        split3(v, h, m, o);
        dst[i] = h;
        dst[per + i] = m;
        dst[2 * per + i] = o;
    }
}

// G^T R right-hand sides for FP64 / FP32 G, row-major per 16-wide tile:
// dst[(t * ldr + i) * 16 + n] = column 16 t + n, row i.
template <typename In>
__global__ void pack_r_kernel(const double *__restrict__ src, int64_t n, int64_t k, int64_t ldr,
                              int ncol, In *__restrict__ dst)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)ncol * ldr) return;
    const int64_t t = i / (ldr * 16), rem = i - t * ldr * 16, row = rem / 16, nn = rem % 16;
    const int64_t c = t * 16 + nn;
    dst[i] = (c < k && row < n) ? (In)src[c * n + row] : (In)0;
}

// G^T R right-hand sides for BF16 G: three BF16 parts, each [t][row / 4][n][row % 4].
__global__ void pack_rb_kernel(const double *__restrict__ src, int64_t n, int64_t k, int64_t ldr4,
                               int ncol, unsigned short *__restrict__ dst)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)ncol * ldr4 * 4;
    if (i >= per) return;
    const int64_t t = i / (ldr4 * 64), rem = i - t * ldr4 * 64;
    const int64_t q4 = rem / 64, w = rem % 64, nn = w / 4, row = q4 * 4 + w % 4;
    const int64_t c = t * 16 + nn;
    __bf16 h, m, o;
    split3((c < k && row < n) ? src[c * n + row] : 0.0, h, m, o);
    dst[i] = __builtin_bit_cast(unsigned short, h);
    dst[per + i] = __builtin_bit_cast(unsigned short, m);
    dst[2 * per + i] = __builtin_bit_cast(unsigned short, o);
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

int elems_per_16b(gra_dtype dt) { return dt == GRA_F64 ? 2 : (dt == GRA_F32 ? 4 : 8); }
int gx_chunk(gra_dtype dt) { return dt == GRA_F64 ? 64 : (dt == GRA_F32 ? 128 : 256); }
// columns per G^T R column tile: 16 lanes x 16 bytes (FP64 / FP32), 128 for BF16
int64_t gt_wave_cols(gra_dtype dt) { return dt == GRA_BF16 ? 128 : 16 * elems_per_16b(dt); }
int64_t gt_waves(gra_dtype dt, int64_t ld, int mc)
{
    const int64_t w = gt_wave_cols(dt) * mc;
    return (ld + w - 1) / w;
}

}  // namespace

// Launch shapes from sweeps on MI355X (16 right-hand sides; 5,000 x 32,000,
// 16,384 x 131,072 and 90,000 x 131,072 G):
// * G X: 6 column splits (chunk-aligned spans) whatever the row count;
// * G^T R: 4 column tiles per wave where that leaves at least 512 waves,
//   then splits of the rows (multiples of 16) for about 1,024 waves;
//   otherwise one tile per wave and about 2,048 waves (BF16: one 128-column
//   strip per wave, about 2,048 waves).
MatMulPlan plan_mat_mul(gra_dtype dt, bool trans, int64_t Nd, int64_t Nm, int64_t ld, int64_t k)
{
    MatMulPlan p;
    (void)Nm;
    p.ncol = (int)((k + 15) / 16 * 16);
    if (!trans) {
        const int64_t kc = gx_chunk(dt);
        int64_t s = 6;
        const int64_t smax = (ld + kc - 1) / kc;          // at least one chunk each
        s = s > smax ? smax : s;
        int64_t span = (ld + s - 1) / s;
        p.span = (span + kc - 1) / kc * kc;
        p.splits = (int)((ld + p.span - 1) / p.span);
        p.ldp = Nd;
        p.pack_rows = ld;
    } else {
        // about 2,048 waves in all (one to four row splits on the 90,000 x
        // 131,072 G: FP64 6.89 TB/s with two against 6.65 with one)
        p.mc = 1;
        const int64_t target = 2048;
        if (dt != GRA_BF16 && (ld + 4 * gt_wave_cols(dt) - 1) / (4 * gt_wave_cols(dt)) >= 512)
            p.mc = 4;
        const int64_t waves = gt_waves(dt, ld, p.mc);
        int64_t s = (target + waves / 2) / waves;
        const int64_t smax = (Nd + 15) / 16;             // at least 16 rows each
        s = s < 1 ? 1 : (s > smax ? smax : s);
        const int64_t span = (Nd + s - 1) / s;
        p.span = (span + 15) / 16 * 16;
        p.splits = (int)((Nd + p.span - 1) / p.span);
        p.ldp = ld;
        p.pack_rows = (Nd + 15) / 16 * 16;
    }
    return p;
}

size_t mat_mul_pack_bytes(gra_dtype dt, const MatMulPlan &p)
{
    const size_t per = (size_t)p.ncol * (size_t)p.pack_rows;
    return dt == GRA_F64 ? 8 * per : (dt == GRA_F32 ? 4 * per : 6 * per);
}

hipError_t launch_pack_rhs(gra_dtype dt, bool trans, const double *src, int64_t n, int64_t k,
                           const MatMulPlan &p, void *dst, hipStream_t st)
{
    const int64_t tot = (int64_t)p.ncol * p.pack_rows;
    const dim3 g(nblk(tot, kBlk)), b(kBlk);
    if (!trans) {
        if (dt == GRA_F64)
            hipLaunchKernelGGL((pack_x_kernel<double, 1>), g, b, 0, st, src, n, k, p.pack_rows,
                               p.ncol, (double *)dst);
        else if (dt == GRA_F32)
            hipLaunchKernelGGL((pack_x_kernel<float, 1>), g, b, 0, st, src, n, k, p.pack_rows,
                               p.ncol, (float *)dst);
        else
            hipLaunchKernelGGL((pack_x_kernel<__bf16, 3>), g, b, 0, st, src, n, k, p.pack_rows,
                               p.ncol, (__bf16 *)dst);
    } else {
        if (dt == GRA_F64)
            hipLaunchKernelGGL(pack_r_kernel<double>, g, b, 0, st, src, n, k, p.pack_rows, p.ncol,
                               (double *)dst);
        else if (dt == GRA_F32)
            hipLaunchKernelGGL(pack_r_kernel<float>, g, b, 0, st, src, n, k, p.pack_rows, p.ncol,
                               (float *)dst);
        else
            hipLaunchKernelGGL(pack_rb_kernel, g, b, 0, st, src, n, k, p.pack_rows / 4, p.ncol,
                               (unsigned short *)dst);
    }
    return hipGetLastError();
}

hipError_t launch_mat_mul(const void *G, gra_dtype dt, bool trans, int64_t Nd, int64_t Nm,
                          int64_t ld, const void *V, const MatMulPlan &p, double *part,
                          hipStream_t st)
{
    const int ctiles = p.ncol / 16;
    if (!trans) {
        const dim3 grid((unsigned)((Nd + 63) / 64), (unsigned)p.splits, (unsigned)ctiles);
        if (dt == GRA_F64)
            hipLaunchKernelGGL(gmx_kernel<double>, grid, dim3(kBlk), 0, st, (const double *)G, Nd,
                               ld, p.span, (const double *)V, p.pack_rows, p.ncol, part);
        else if (dt == GRA_F32)
            hipLaunchKernelGGL(gmx_kernel<float>, grid, dim3(kBlk), 0, st, (const float *)G, Nd, ld,
                               p.span, (const float *)V, p.pack_rows, p.ncol, part);
        else
            hipLaunchKernelGGL(gmx_kernel<__hip_bfloat16>, grid, dim3(kBlk), 0, st,
                               (const __hip_bfloat16 *)G, Nd, ld, p.span, (const __bf16 *)V,
                               p.pack_rows, p.ncol, part);
        return hipGetLastError();
    }
    const int64_t waves = gt_waves(dt, ld, p.mc);
    const dim3 grid(nblk(waves, kWavesPerBlk), (unsigned)p.splits, (unsigned)ctiles);
#define GRA_GTX(T, MC)                                                                           \
    hipLaunchKernelGGL((gtx_kernel<T, MC>), grid, dim3(kBlk), 0, st, (const T *)G, Nd, Nm, ld,  \
                       p.span, (const T *)V, p.pack_rows, waves, p.ncol, part)
    if (dt == GRA_F64) {
        if (p.mc == 4) GRA_GTX(double, 4);
        else GRA_GTX(double, 1);
    } else if (dt == GRA_F32) {
        if (p.mc == 4) GRA_GTX(float, 4);
        else GRA_GTX(float, 1);
    } else
        hipLaunchKernelGGL(gtb_kernel, grid, dim3(kBlk), 0, st, (const __hip_bfloat16 *)G, Nd, Nm,
                           ld, p.span, (const unsigned short *)V, p.pack_rows / 4,
                           (int64_t)p.ncol * p.pack_rows, waves, p.ncol, part);
#undef GRA_GTX
    return hipGetLastError();
}

hipError_t launch_sum_splits(const double *part, const MatMulPlan &p, int64_t n, int64_t k,
                             double *out, hipStream_t st)
{
    hipLaunchKernelGGL(sum_splits_kernel, dim3(nblk(k * n, kBlk)), dim3(kBlk), 0, st, part,
                       p.splits, p.ncol, p.ldp, n, k, out);
    return hipGetLastError();
}

}  // namespace gra
