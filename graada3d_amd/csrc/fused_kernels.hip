// Single-read fused normal-equation product for the CGLS data block:
//
//     y = G x,   t_d = Wd * y,   q = G^T (Wd * t_d) = G^T (Wd^2 * y)
//
// in ONE pass over G.  (The reference's solve_cg reads G twice per iteration,
// GaussNewtonInversion.cpp:1566 and :1637; the fused CG uses the equal
// recurrence r <- r + alpha A^T t instead of r = A^T s.)
//
// Decomposition.  A cooperatively launched grid (every workgroup resident) of
// S row groups x NT column tiles, one 512-thread workgroup per CU.  Workgroup
// (group, tile) owns the TW = 512*V*CW columns of its tile (V = elements per
// 16-byte vector): each thread holds its CW*V entries of x and of the q
// accumulator in registers for the whole pass.  It streams the rows of its
// group in blocks of RB rows (64 KB of G per block and workgroup for every
// storage type), each block loaded once from HBM into a register ring PF
// blocks ahead of use.  Step k:
//
//   phase A   partial dots of block k over the tile (FP64), reduced across
//             the workgroup in a fixed order and posted to global memory;
//   poll      wave 0 collects the NT partials of each row of block k-L, posted
//             by the group's other tiles, sums them in a fixed order (every
//             tile of the group gets bit-identical y), and forms
//             c = Wd^2 y in LDS; the tile (k-L) mod NT also writes t_d and
//             the block's ||t_d||^2 partial;
//   phase B   q += sum over the rows of block k-L of c_row * G[row, tile],
//             from the same registers (L blocks after their phase A).
//
// So a row's partial dots cross workgroups once (one hop, agent scope), and
// nothing else does.  The groups of an XCD-aligned grid lie in one XCD
// (workgroups are dispatched round robin), which keeps a group's traffic in
// one L2.  Publication uses no fences or counters: the partial buffer is
// filled with a signalling-NaN sentinel before the launch, producers store
// with agent-scope atomic stores (device-coherent, single-copy atomic for 8
// bytes) and the polling wave re-reads until no value is the sentinel (a
// computed value is never that bit pattern: arithmetic NaNs are quiet).  The
// poll of block j is issued one step early, after that step's G loads, and
// finished at the start of the next step, so that waiting for it is waiting
// for loads that step needs anyway.
//
// Block k's partials are posted at step k before the workgroup waits on
// block k-L+1 (at step k+1, L >= 2), so with every workgroup resident each
// block's posts happen after waits on strictly earlier blocks only: no
// deadlock.  Every wait is bounded all the same: on time-out the kernel sets
// *err, abandons its remaining waits, runs to the end, and the host discards
// the result and reruns the solve with the two-pass iteration.
//
// Affine pass (dobs given): c = Wd^2 (y - dobs) instead of Wd^2 y, and y
// itself is written (to yout) instead of t_d: one read of G gives both the
// data misfit of x and G^T (Wd^2 (G x - dobs)), the data part of the next
// CG solve's initial residual A^T b.
//
// At the end every thread writes its columns of q for the group (a group's
// columns are owned by exactly one tile, so there is no cross-workgroup sum);
// the consumer adds the S group partials in order, so results are bit-for-bit
// repeatable.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <mutex>
#include <type_traits>

#include "engine.hpp"
#include "fused_plan.hpp"

namespace gra {

namespace {

constexpr int kThreads = 512;
constexpr int kWaves = kThreads / 64;
constexpr int kMaxPoll = 4;            // partials per lane of the polling wave
// Sentinel of unpublished partials (a signalling NaN).
constexpr uint64_t kSentinel = kFusedSentinel;

typedef uint32_t v4u32 __attribute__((ext_vector_type(4)));

template <typename T> struct VecN;
template <> struct VecN<double> { static constexpr int n = 2; };
template <> struct VecN<float> { static constexpr int n = 4; };
template <> struct VecN<__hip_bfloat16> { static constexpr int n = 8; };

// Element q of a 16-byte vector of T, widened to FP64 (q is a compile-time
// constant after unrolling, so the selects fold away; no array is formed,
// which would otherwise be placed in LDS).
template <typename T> __device__ inline double elem(const v4u32 &v, int q);
template <> __device__ inline double elem<double>(const v4u32 &v, int q)
{
    return q == 0 ? __hiloint2double((int)v.y, (int)v.x) : __hiloint2double((int)v.w, (int)v.z);
}
template <> __device__ inline double elem<float>(const v4u32 &v, int q)
{
    const uint32_t w = q == 0 ? v.x : (q == 1 ? v.y : (q == 2 ? v.z : v.w));
    return (double)__uint_as_float(w);
}
template <> __device__ inline double elem<__hip_bfloat16>(const v4u32 &v, int q)
{
    const int h = q >> 1;
    const uint32_t w = h == 0 ? v.x : (h == 1 ? v.y : (h == 2 ? v.z : v.w));
    return (double)__uint_as_float((q & 1) ? (w & 0xffff0000u) : (w << 16));
}

// The same element as FP32 (BF16 and FP32 storage only).
template <typename T> __device__ inline float elemf(const v4u32 &v, int q);
template <> __device__ inline float elemf<double>(const v4u32 &v, int q)
{
    return (float)elem<double>(v, q);
}
template <> __device__ inline float elemf<float>(const v4u32 &v, int q)
{
    return __uint_as_float(q == 0 ? v.x : (q == 1 ? v.y : (q == 2 ? v.z : v.w)));
}
template <> __device__ inline float elemf<__hip_bfloat16>(const v4u32 &v, int q)
{
    const int h = q >> 1;
    const uint32_t w = h == 0 ? v.x : (h == 1 ? v.y : (h == 2 ? v.z : v.w));
    return __uint_as_float((q & 1) ? (w & 0xffff0000u) : (w << 16));
}

// Agent-scope relaxed atomics: device-coherent loads and stores that neither
// read a stale L1 line nor another XCD's stale L2 copy.
template <typename U> __device__ inline void st_agent(U *p, U v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename U> __device__ inline U ld_agent(const U *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline bool is_sentinel(double v) { return __double_as_longlong(v) == (long long)kSentinel; }

struct FusedArgs {
    const void *G;
    int64_t Nd, ld;
    int NT, S;            // column tiles, row groups
    int per_xcd, n_xcd;   // XCD-aligned mapping when per_xcd > 0
    int64_t rows_per_group;
    int64_t nrb;          // row blocks per group
    const double *x;      // [ld] (zero beyond Nm)
    const double *Wd;     // [Nd] (null = 1)
    const double *dobs;   // [Nd]: affine pass (null: the CG product)
    double *yout;         // [Nd]: y = G x of the affine pass
    double *partial;      // [S*nrb][NT][RB], sentinel on entry
    double *tdat;         // [Nd]   Wd * y
    double *t2part;       // [S*nrb]
    double *zpart;        // [S][ld]
    int *err;
    const int *done;
    int64_t wait_ticks;   // bound of every wait, wall-clock ticks
    int64_t fail_block;   // test hook: the poll of this block (of every group) fails; -1 none
};

// Sum over the 64 lanes of RB values per lane; afterwards lane l holds the sum
// of value r(l) = sum over the first log2(RB) halving steps of
// (l & off ? n/2 : 0).  Fixed operation order (the same in every workgroup).
template <int RB> __device__ inline double wave_reduce_rows(double (&v)[RB], int lane, int &row)
{
    int off = 32;
    row = 0;
#pragma unroll
    for (int n = RB; n > 1; n >>= 1, off >>= 1) {
        const bool upper = (lane & off) != 0;
#pragma unroll
        for (int j = 0; j < n / 2; ++j) {
            // both halves cross, each lane keeps its own: a select of values
            // (a select of array elements would turn into an indexed array)
            const double lo = v[j] + __shfl_xor(v[j], off, 64);
            const double hi = v[j + n / 2] + __shfl_xor(v[j + n / 2], off, 64);
            v[j] = upper ? hi : lo;
        }
        if (upper) row += n / 2;
    }
    double s = v[0];
    for (; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
}

template <typename T, int CW, int RB, int L, int PF, bool M32>
__global__ void __launch_bounds__(kThreads, 1)
fused_normal_kernel(FusedArgs a)
{
    if (a.done && *a.done) return;
    constexpr int V = VecN<T>::n;
    constexpr int TW = kThreads * V * CW;
    constexpr int NS = L + PF + 1;     // register ring depth (blocks)
    // steps per unrolled loop body: a multiple of NS (static ring slots) and
    // of 2 (static LDS buffer parities)
    constexpr int U = NS % 2 == 0 ? NS : 2 * NS;
    static_assert(L >= 2 && PF >= 1, "lag of two blocks, prefetch of at least one");
    static_assert(64 % RB == 0 && RB <= 8, "rows per block");
    // the lanes that hold a row sum after wave_reduce_rows: low bits zero
    constexpr int kRowLaneMask = 64 / RB - 1;

    // double-buffered by block parity: one barrier per step orders them
    __shared__ double pa[2][kWaves][RB];   // per-wave partial dots of a block
    __shared__ double cl[2][RB];           // coefficients of a block for phase B

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int lin = blockIdx.x;
    if (a.per_xcd > 0) lin = (int)(blockIdx.x % a.n_xcd) * a.per_xcd + (int)(blockIdx.x / a.n_xcd);
    const int NT = a.NT;
    const int tile = lin % NT, group = lin / NT;
    const int64_t c0 = (int64_t)tile * TW;
    const int64_t g_row0 = (int64_t)group * a.rows_per_group;
    int64_t g_row1 = g_row0 + a.rows_per_group;
    if (g_row1 > a.Nd) g_row1 = a.Nd;
    const int64_t nrb = a.nrb;
    const int64_t rb0 = (int64_t)group * nrb;
    const T *G = static_cast<const T *>(a.G);

    double xr[CW][V], zacc[CW][V];
    float xf[CW][V];                   // M32: x rounded to FP32
    int64_t colc[CW];
#pragma unroll
    for (int cw = 0; cw < CW; ++cw) {
        const int64_t c = c0 + (int64_t)cw * kThreads * V + (int64_t)threadIdx.x * V;
        const bool ok = c < a.ld;
        // columns past ld read a valid column (x is zero there, q is not stored)
        colc[cw] = ok ? c : a.ld - V;
#pragma unroll
        for (int q = 0; q < V; ++q) {
            xr[cw][q] = ok ? a.x[c + q] : 0.0;
            xf[cw][q] = (float)xr[cw][q];
            zacc[cw][q] = 0.0;
        }
    }

    v4u32 ring[NS][RB][CW];
    // Loads are unconditional, so the compiler counts them in flight and waits
    // only for the block it is about to use.  Rows past the matrix read its
    // last row (their coefficient is zero, their partial is discarded).
    auto load_block = [&](int64_t k, v4u32 (&slot)[RB][CW]) {
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            int64_t row = g_row0 + k * RB + i;
            if (row > a.Nd - 1) row = a.Nd - 1;
            const T *gr = G + row * a.ld;
#pragma unroll
            for (int cw = 0; cw < CW; ++cw)
                slot[i][cw] = __builtin_nontemporal_load(reinterpret_cast<const v4u32 *>(gr + colc[cw]));
        }
    };

    // M32 (BF16 or FP32 storage): a thread's CW*V products of a row are formed
    // and summed in FP32 (the stored values widen exactly; x and the
    // coefficients are rounded to FP32), and each such sum is added in FP64,
    // so a row's dot and a column's q accumulate in FP64 across threads and
    // blocks.  Fewer FP64 instructions per element; rounding of the order of
    // FP32's, far below BF16's.
    auto phase_a = [&](v4u32 (&slot)[RB][CW], int par) {
        double s[RB];
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            if (M32) {
                float f = 0.0f;
#pragma unroll
                for (int cw = 0; cw < CW; ++cw)
#pragma unroll
                    for (int q = 0; q < V; ++q) f = fmaf(elemf<T>(slot[i][cw], q), xf[cw][q], f);
                s[i] = (double)f;
            } else {
                s[i] = 0.0;
#pragma unroll
                for (int cw = 0; cw < CW; ++cw)
#pragma unroll
                    for (int q = 0; q < V; ++q) s[i] = fma(elem<T>(slot[i][cw], q), xr[cw][q], s[i]);
            }
        }
        int row;
        const double v = wave_reduce_rows<RB>(s, lane, row);
        if ((lane & kRowLaneMask) == 0) pa[par][wave][row] = v;
    };

    auto phase_b = [&](v4u32 (&slot)[RB][CW], int par) {
        if (M32) {
            float zf[CW][V];
#pragma unroll
            for (int cw = 0; cw < CW; ++cw)
#pragma unroll
                for (int q = 0; q < V; ++q) zf[cw][q] = 0.0f;
#pragma unroll
            for (int i = 0; i < RB; ++i) {
                const float cr = (float)cl[par][i];
#pragma unroll
                for (int cw = 0; cw < CW; ++cw)
#pragma unroll
                    for (int q = 0; q < V; ++q) zf[cw][q] = fmaf(elemf<T>(slot[i][cw], q), cr, zf[cw][q]);
            }
#pragma unroll
            for (int cw = 0; cw < CW; ++cw)
#pragma unroll
                for (int q = 0; q < V; ++q) zacc[cw][q] += (double)zf[cw][q];
        } else {
#pragma unroll
            for (int i = 0; i < RB; ++i) {
                const double cr = cl[par][i];
#pragma unroll
                for (int cw = 0; cw < CW; ++cw)
#pragma unroll
                    for (int q = 0; q < V; ++q) zacc[cw][q] = fma(elem<T>(slot[i][cw], q), cr, zacc[cw][q]);
            }
        }
    };

    // ---- wave 0: posts and polls
    const int np = NT * RB;                // partials per block
    bool failed = false;                   // sticky: no more waits after a failure
    auto post = [&](int64_t k, int par) {
        if (lane < RB) {
            double s = pa[par][0][lane];
#pragma unroll
            for (int w = 1; w < kWaves; ++w) s += pa[par][w][lane];
            st_agent(a.partial + ((rb0 + k) * NT + tile) * RB + lane, s);
        }
    };
    // The poll of block j and its rows' Wd are loaded one step early (vector
    // loads complete in order: a load issued and waited for within one step
    // would make the wave wait for that step's whole prefetched block).
    auto issue_poll = [&](int64_t j, double (&v)[kMaxPoll], double &wd, double &dv) {
        const double *pb = a.partial + (rb0 + j) * NT * RB;
#pragma unroll
        for (int u = 0; u < kMaxPoll; ++u) {
            const int idx = lane + 64 * u;
            v[u] = idx < np ? ld_agent(pb + idx) : 0.0;
        }
        const int64_t row = g_row0 + j * RB + lane;
        const bool inr = lane < RB && row < g_row1;
        wd = (inr && a.Wd) ? a.Wd[row] : 1.0;
        dv = (inr && a.dobs) ? a.dobs[row] : 0.0;
    };
    auto pending = [&](const double (&v)[kMaxPoll]) {
        bool pend = false;
#pragma unroll
        for (int u = 0; u < kMaxPoll; ++u) pend = pend || (lane + 64 * u < np && is_sentinel(v[u]));
        return __any(pend);
    };
    auto finish_poll = [&](int64_t j, double (&v)[kMaxPoll], double wd, double dv, int par) {
        const double *pb = a.partial + (rb0 + j) * NT * RB;
        if (j == a.fail_block && !failed) {
            failed = true;
            if (lane == 0) atomicOr(a.err, 4);
        }
        // the first check is outside the re-poll loop, so that the common case
        // waits only for the early poll loads, not for the prefetched block
        if (!failed && pending(v)) {
            int64_t t0 = 0;
            unsigned spins = 0;
            do {
                if ((++spins & 31u) == 0u) {
                    const int64_t t = wall_clock64();
                    if (t0 == 0) t0 = t;
                    const bool other = ld_agent(a.err) != 0;
                    if (other || t - t0 > a.wait_ticks) {
                        if (!other && lane == 0) atomicOr(a.err, 2);
                        failed = true;
                        break;
                    }
                }
                __builtin_amdgcn_s_sleep(1);
#pragma unroll
                for (int u = 0; u < kMaxPoll; ++u) {
                    const int idx = lane + 64 * u;
                    if (idx < np && is_sentinel(v[u])) v[u] = ld_agent(pb + idx);
                }
            } while (pending(v));
        }
        // lane l sums the tiles l/RB + (64/RB) u of row l % RB, then the lanes
        // of a row combine with a fixed butterfly
        double y = 0.0;
#pragma unroll
        for (int u = 0; u < kMaxPoll; ++u)
            if (lane + 64 * u < np) y += v[u];
#pragma unroll
        for (int off = RB; off < 64; off <<= 1) y += __shfl_xor(y, off, 64);
        const int64_t row = g_row0 + j * RB + lane;
        const bool valid = lane < RB && row < g_row1;
        const bool writer = tile == (int)(j % NT);
        double t2 = 0.0;
        if (lane < RB) {
            double c = 0.0;
            if (valid) {
                if (a.dobs) {
                    // affine pass: q = G^T (Wd^2 (y - dobs)), y kept
                    const double td = wd * (y - dv);
                    c = wd * td;
                    t2 = td * td;
                    if (writer) a.yout[row] = y;
                } else {
                    const double td = wd * y;
                    c = wd * td;
                    t2 = td * td;
                    if (writer) a.tdat[row] = td;     // read by later kernels only
                }
            }
            cl[par][lane] = c;
        }
#pragma unroll
        for (int off = 1; off < RB; off <<= 1) t2 += __shfl_xor(t2, off, 64);
        if (writer && lane == 0) a.t2part[rb0 + j] = t2;
    };

    double pv[kMaxPoll], pwd = 1.0, pdv = 0.0;
#pragma unroll
    for (int u = 0; u < kMaxPoll; ++u) pv[u] = 0.0;

    // Step k, with one barrier:
    //   wave 0: finish the poll of block k-L (issued last step) -> cl[(k-L)&1]
    //   all:    G loads of block k+PF; phase A(k) -> pa[k&1]
    //   barrier
    //   wave 0: post block k from pa[k&1]; issue the poll of block k-L+1
    //   all:    phase B(k-L) with cl[(k-L)&1]
    // Waiting for the poll at the start of a step waits only for loads issued
    // before this step's G loads (vector loads complete in order), i.e. for
    // the block phase A needs next.  pa[k&1] is next written at step k+2,
    // after the barrier of step k+1, which wave 0 passes after posting block
    // k; cl[(k-L)&1] is next written at step k+2 before that step's barrier,
    // after every wave has passed the barrier of step k+1, i.e. finished
    // phase B(k-L).
    const int64_t last = nrb - 1;
#pragma unroll
    for (int p = 0; p < PF; ++p) load_block(p < last ? p : last, ring[p]);
    // block k lives in ring slot k % NS; kk is a multiple of U, so k % NS
    // and k & 1 are s % NS and s & 1
    for (int64_t kk = 0; kk < nrb + L; kk += U) {
#pragma unroll
        for (int s = 0; s < U; ++s) {
            const int64_t k = kk + s;
            if (k >= nrb + L) break;
            const int64_t j = k - L;
            const int pj = (s - L + 2 * U) & 1, pk = s & 1;
            if (wave == 0 && j >= 0 && j < nrb) finish_poll(j, pv, pwd, pdv, pj);
            // slot (s+PF)%NS held block k+PF-NS = k-L-1, consumed at step k-1
            load_block(k + PF < last ? k + PF : last, ring[(s + PF) % NS]);
            if (k < nrb) phase_a(ring[s % NS], pk);
            __syncthreads();
            if (wave == 0) {
                if (k < nrb) post(k, pk);
                const int64_t jp = k - L + 1;
                if (jp >= 0 && jp < nrb) issue_poll(jp, pv, pwd, pdv);
            }
            if (j >= 0 && j < nrb) phase_b(ring[(s - L + U) % NS], pj);
        }
    }
    // q of this tile for this row group
#pragma unroll
    for (int cw = 0; cw < CW; ++cw) {
        const int64_t c = c0 + (int64_t)cw * kThreads * V + (int64_t)threadIdx.x * V;
#pragma unroll
        for (int q = 0; q < V; ++q)
            if (c + q < a.ld) a.zpart[(int64_t)group * a.ld + c + q] = zacc[cw][q];
    }
}

__global__ void fill_u64(uint64_t *p, int64_t n, uint64_t v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// (lag L, prefetch PF) variants; the ring holds L + PF + 1 blocks.
constexpr int kNumDepths = 3;
constexpr int kDepthL[kNumDepths] = {2, 2, 3};     // prefetch 1, 2, 2

template <typename T, int CW, int RB> struct FusedCfg {
    static constexpr int V = VecN<T>::n;
    static constexpr int TW = kThreads * V * CW;
    template <bool M32> static const void *pick(int depth_index)
    {
        switch (depth_index) {
        case 1: return (const void *)fused_normal_kernel<T, CW, RB, 2, 2, M32>;
        case 2: return (const void *)fused_normal_kernel<T, CW, RB, 3, 2, M32>;
        default: return (const void *)fused_normal_kernel<T, CW, RB, 2, 1, M32>;
        }
    }
    static const void *kernel(int depth_index, bool m32)
    {
        if constexpr (!std::is_same<T, double>::value) {
            if (m32) return pick<true>(depth_index);
        }
        return pick<false>(depth_index);
    }
};

// Per storage type: 16-byte vectors per row and thread, rows per block
// (64 KB of G per block and workgroup; 4,096-column tiles for FP64 / FP32,
// 8,192 for BF16: two vectors of 4 rows measured 3 % faster than one of 8 on
// the 90,000 x 131,072 G -- half the partials to poll and a shorter
// cross-lane reduction of the row dots).
using CfgF64 = FusedCfg<double, 4, 2>;
using CfgF32 = FusedCfg<float, 2, 4>;
using CfgBf16 = FusedCfg<__hip_bfloat16, 2, 4>;

}  // namespace

bool fused_plan(gra_dtype dt, int64_t Nd, int64_t ld, int device, int target_wgs, int depth,
                int math, int min_blocks, FusedPlan *out)
{
    int cus = 0, coop = 0, clock_khz = 0, n_xcd = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return false;
    if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device) != hipSuccess ||
        !coop)
        return false;
    if (hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess ||
        clock_khz <= 0)
        return false;
    if (hipDeviceGetAttribute(&n_xcd, hipDeviceAttributeNumberOfXccs, device) != hipSuccess ||
        n_xcd < 1)
        n_xcd = 1;
    const int di = depth >= 1 && depth <= kNumDepths ? depth - 1 : 0;
    const void *kern;
    int TW, RB;
    // math: 0 = the storage type's default (FP64 products for FP64 and FP32
    // storage, FP32 products for BF16), 1 = FP64 products, 2 = FP32 products
    const bool m32 = math == 2 || (math == 0 && dt == GRA_BF16);
    switch (dt) {
    case GRA_F64: kern = CfgF64::kernel(di, false); TW = CfgF64::TW; RB = 2; break;
    case GRA_F32: kern = CfgF32::kernel(di, m32); TW = CfgF32::TW; RB = 4; break;
    default: kern = CfgBf16::kernel(di, m32); TW = CfgBf16::TW; RB = 4; break;
    }
    // co-resident workgroups per CU from the runtime's occupancy figure for
    // this kernel's registers and LDS (the cooperative launch re-checks it)
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kThreads, 0) != hipSuccess)
        return false;
    FusedShape s;
    if (!fused_plan_shape(Nd, ld, TW, RB, cus, n_xcd, per_cu, target_wgs, kMaxPoll, &s,
                          min_blocks))
        return false;
    FusedPlan p;
    p.NT = s.NT;
    p.S = s.S;
    p.per_xcd = s.per_xcd;
    p.n_xcd = n_xcd;
    p.rows_per_group = s.rows_per_group;
    p.nrb_group = s.nrb;
    p.RB = RB;
    p.threads = kThreads;
    p.grid = s.grid;
    p.L = kDepthL[di];
    p.kernel = kern;
    p.wait_ticks = (uint64_t)clock_khz * 2000ull;   // 2 s
    *out = p;
    return true;
}

int64_t fused_partials(const FusedPlan &p) { return (int64_t)p.S * p.nrb_group * p.NT * p.RB; }

size_t fused_scratch_bytes(const FusedPlan &p, int64_t ld)
{
    int64_t nrb = (int64_t)p.S * p.nrb_group;
    return sizeof(double) * (size_t)(fused_partials(p) + nrb) +   // partial, t2part
           sizeof(double) * (size_t)p.S * (size_t)ld;               // zpart
}

namespace {
constexpr int kMaxDevices = 64;
// The last cooperative grid launched on each device (event recorded after
// it, and the stream it was queued on); created on first use, kept until the
// process ends.
struct CoopOrder {
    hipEvent_t last = nullptr;
    hipStream_t stream = nullptr;
};
std::mutex coop_mu;
CoopOrder coop_order[kMaxDevices];
}  // namespace

hipError_t launch_fused(const FusedPlan &p, const void *G, int64_t Nd, int64_t ld,
                        const double *x, const double *Wd, double *tdat, void *scratch, int *err,
                        const int *done, int64_t fail_block, double **zpart_out,
                        double **t2part_out, hipStream_t st, const double *dobs, double *yout,
                        bool partials_filled)
{
    const int64_t nrb = (int64_t)p.S * p.nrb_group;
    char *b = static_cast<char *>(scratch);
    FusedArgs a;
    a.G = G;
    a.Nd = Nd;
    a.ld = ld;
    a.NT = p.NT;
    a.S = p.S;
    a.per_xcd = p.per_xcd;
    a.n_xcd = p.n_xcd;
    a.rows_per_group = p.rows_per_group;
    a.nrb = p.nrb_group;
    a.x = x;
    a.Wd = Wd;
    a.dobs = dobs;
    a.yout = yout;
    a.partial = reinterpret_cast<double *>(b);
    b += sizeof(double) * (size_t)(nrb * p.NT * p.RB);
    a.t2part = reinterpret_cast<double *>(b);
    b += sizeof(double) * (size_t)nrb;
    a.zpart = reinterpret_cast<double *>(b);
    a.tdat = tdat;
    a.err = err;
    a.done = done;
    a.wait_ticks = (int64_t)p.wait_ticks;
    a.fail_block = fail_block;
    const int64_t npart = fused_partials(p);
    if (!partials_filled)
        hipLaunchKernelGGL(fill_u64, dim3((unsigned)((npart + 255) / 256)), dim3(256), 0, st,
                           reinterpret_cast<uint64_t *>(a.partial), npart, kSentinel);
    if (zpart_out) *zpart_out = a.zpart;
    if (t2part_out) *t2part_out = a.t2part;
    // One cooperative grid at a time per device, in launch order, whatever
    // stream it is queued on: each grid is planned to hold every CU, so two
    // in flight at once could keep each other's workgroups from becoming
    // resident until their bounded waits fail.
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(coop_mu);
    CoopOrder &o = coop_order[dev];
    if (!o.last) {
        // ordering only, waited on by streams of this device: no system-scope
        // release needed at the record
        if ((e = hipEventCreateWithFlags(&o.last, hipEventDisableTiming | hipEventDisableSystemFence)) !=
            hipSuccess)
            return e;
    } else if (o.stream != st) {
        if ((e = hipStreamWaitEvent(st, o.last, 0)) != hipSuccess) return e;
    }
    void *args[] = {&a};
    e = hipLaunchCooperativeKernel(p.kernel, dim3((unsigned)p.grid), dim3(p.threads), args, 0, st);
    if (e != hipSuccess) return e;
    if ((e = hipEventRecord(o.last, st)) != hipSuccess) return e;
    o.stream = st;
    return hipSuccess;
}

}  // namespace gra
