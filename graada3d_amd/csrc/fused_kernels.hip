// Single-read fused normal-equation product for the CGLS data block:
//
//     y = G x,   t_d = Wd * y,   q = G^T (Wd * t_d) = G^T (Wd^2 * y)
//
// in ONE pass over G.  (The reference's solve_cg reads G twice per iteration,
// GaussNewtonInversion.cpp:1566 and :1637; the fused CG uses the equal
// recurrence r <- r + alpha A^T t instead of r = A^T s.)
//
// Decomposition.  A persistent, cooperatively launched grid (every workgroup
// co-resident) of NT x S workgroups.  Workgroup (tile, group) owns a column
// tile of TW = 64*V*CW columns (V = elements per 16-byte vector) and streams the
// rows of its row group in row blocks of RB = 4*RW rows.  A workgroup has five
// waves with two roles:
//
//  * waves 0-3 (compute): each owns RW rows of every block.  A block is loaded
//    from HBM once, PF blocks ahead, into a register ring; "phase A" forms the
//    partial dots of its rows over the tile and posts them to LDS; "phase B",
//    L blocks later, takes the block's coefficients c = Wd^2 y from LDS and
//    accumulates q_tile += sum_rows c * G[row, tile] from the same registers.
//    These waves synchronise only through LDS, so their outstanding G loads
//    are never drained by a memory fence.
//  * wave 4 (sync), step s: hands block s-L's coefficients to the compute
//    waves (loaded one step earlier), stores block s's partial dots to global
//    memory, and -- in the group's tile (s-D) mod NT -- reduces block s-D: the
//    NT partials of each row in tile order, t_d = Wd y, c = Wd^2 y and a
//    partial of ||t_d||^2.
//
// Cross-workgroup data (partials, coefficients) is published without fences
// or counters: both buffers are filled with a signalling-NaN sentinel before
// the launch, producers only store (agent-scope atomic stores: device-coherent,
// single-copy atomic for 8 bytes), and consumers poll the values themselves
// until none is the sentinel.  A computed value is never that bit pattern
// (arithmetic NaNs are quiet).  Every wait points to a strictly earlier step
// of some workgroup (1 <= D < L), so with all workgroups co-resident the grid
// cannot deadlock.
//
// At the end the four compute waves' q partials are summed in wave order and
// the workgroup writes its tile of q for its row group (the consumer sums the
// groups in order), so results are bit-for-bit repeatable.
//
// Ring-slot reuse is safe by construction: LDS slots are reused after DS = L+1
// blocks and register slots after L+1+PF blocks; each reuse is preceded by a
// wait that implies the previous occupant was consumed (see the comments at
// the waits).  Every wait is bounded: on time-out the kernel sets *err, every
// wave abandons its remaining waits, the grid runs to completion, and the host
// reports the failure.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include "engine.hpp"

namespace gra {

namespace {

constexpr int kCompute = 4;
constexpr int kThreads = 64 * (kCompute + 2);   // + a sync wave and a reduce wave
// Sentinel of unpublished partials and coefficients (a signalling NaN).
constexpr uint64_t kSentinel = 0x7FF4DEADBEEF0001ull;

typedef uint32_t v4u32 __attribute__((ext_vector_type(4)));

template <typename T> struct VecN;
template <> struct VecN<double> { static constexpr int n = 2; };
template <> struct VecN<float> { static constexpr int n = 4; };
template <> struct VecN<__hip_bfloat16> { static constexpr int n = 8; };

template <typename T> __device__ inline void unpack(const v4u32 &v, double *o);
template <> __device__ inline void unpack<double>(const v4u32 &v, double *o)
{
    o[0] = __hiloint2double((int)v.y, (int)v.x);
    o[1] = __hiloint2double((int)v.w, (int)v.z);
}
template <> __device__ inline void unpack<float>(const v4u32 &v, double *o)
{
    o[0] = (double)__uint_as_float(v.x);
    o[1] = (double)__uint_as_float(v.y);
    o[2] = (double)__uint_as_float(v.z);
    o[3] = (double)__uint_as_float(v.w);
}
template <> __device__ inline void unpack<__hip_bfloat16>(const v4u32 &v, double *o)
{
    uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        o[2 * q] = (double)__uint_as_float(w[q] << 16);
        o[2 * q + 1] = (double)__uint_as_float(w[q] & 0xffff0000u);
    }
}

// Compiler barrier: keeps LDS accesses in program order (the hardware
// executes one wave's LDS instructions in order).
__device__ inline void compiler_fence() { asm volatile("" ::: "memory"); }

__device__ inline unsigned lds_load(const unsigned *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ inline void lds_store(unsigned *p, unsigned v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Agent-scope relaxed atomics: device-coherent loads and stores (sc1) that do
// not hit another XCD's stale L2 copy.
template <typename U> __device__ inline void st_agent(U *p, U v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename U> __device__ inline U ld_agent(const U *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline bool is_sentinel(double v) { return __double_as_longlong(v) == (long long)kSentinel; }

// Bounded-wait bookkeeping of one wave.  `failed` is sticky: after the first
// failure the wave skips all further waits and runs to the end.
struct Waiter {
    const int *err;
    uint64_t limit;          // wall-clock ticks (about 2 s)
    bool failed = false;
    unsigned spins = 0;
    uint64_t t0 = 0;

    __device__ void begin() { spins = 0; t0 = 0; }
    // One poll step; returns false when the wait must be abandoned.
    __device__ bool keep_waiting(int code, int *err_w)
    {
        if ((++spins & 255u) == 0u) {
            uint64_t t = wall_clock64();
            if (t0 == 0) t0 = t;
            bool other = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
            if (other || t - t0 > limit) {
                if (!other && (threadIdx.x & 63) == 0) atomicOr(err_w, code);
                failed = true;
                return false;
            }
        }
        // back off: many waves may poll one line (every tile of a row group
        // waits for the same coefficients), so the polls must not starve
        // the store they wait for
        if (spins < 16) __builtin_amdgcn_s_sleep(1);
        else if (spins < 64) __builtin_amdgcn_s_sleep(8);
        else __builtin_amdgcn_s_sleep(32);
        return true;
    }
};

struct FusedArgs {
    const void *G;
    int64_t Nd, ld;
    int NT, S;            // column tiles, row groups
    int64_t rows_per_group;
    int64_t nrb_group;    // row blocks per group
    const double *x;      // [ld]
    const double *Wd;     // [Nd] (null = 1)
    double *partial;      // [S*nrb_group][NT][RB], sentinel on entry
    double *coef;         // [Nd]   Wd^2 * y, sentinel on entry
    double *tdat;         // [Nd]   Wd * y
    double *t2part;       // [S*nrb_group]
    double *zpart;        // [S][ld]
    int *err;
    const int *done;
    uint64_t wait_ticks;  // bound of every wait, wall-clock ticks
};

template <typename T, int CW, int RW, int L, int PF>
__global__ void __launch_bounds__(kThreads)
fused_normal_kernel(FusedArgs a)
{
    if (a.done && *a.done) return;
    constexpr int V = VecN<T>::n;
    constexpr int TW = 64 * V * CW;
    constexpr int RB = kCompute * RW;
    constexpr int DS = L + 1;          // LDS ring depth (blocks)
    constexpr int DR = L + 1 + PF;     // register ring depth (blocks)
    static_assert(RB <= 64, "one sync-wave lane per row of a block");

    __shared__ double p_lds[DS][RB];
    __shared__ double c_lds[DS][RB];
    __shared__ unsigned pcount[DS];    // 4 * (uses so far) once all compute waves posted
    __shared__ unsigned cready[DS];    // j + 1 once block j's coefficients are in c_lds
    __shared__ double z_lds[kCompute][TW];

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tile = blockIdx.x % a.NT, group = blockIdx.x / a.NT;
    const int64_t c0 = (int64_t)tile * TW;
    const int64_t g_row0 = (int64_t)group * a.rows_per_group;
    int64_t g_row1 = g_row0 + a.rows_per_group;
    if (g_row1 > a.Nd) g_row1 = a.Nd;
    const int64_t nrb = a.nrb_group;
    const int64_t rb_base = (int64_t)group * nrb;

    if (threadIdx.x < DS) {
        pcount[threadIdx.x] = 0u;
        cready[threadIdx.x] = 0u;
    }
    __syncthreads();

    if (wave < kCompute) {
        // ---------------------------------------------------- compute waves
        const T *G = static_cast<const T *>(a.G);
        Waiter wt{a.err, a.wait_ticks};
        bool col_ok[CW];
        double xr[CW][V];
        double zacc[CW][V];
#pragma unroll
        for (int cw = 0; cw < CW; ++cw) {
            int64_t c = c0 + (int64_t)cw * 64 * V + (int64_t)lane * V;
            col_ok[cw] = c < a.ld;
#pragma unroll
            for (int q = 0; q < V; ++q) {
                xr[cw][q] = col_ok[cw] ? a.x[c + q] : 0.0;
                zacc[cw][q] = 0.0;
            }
        }
        v4u32 ring[DR][RW][CW];

        // Loads are unconditional (no per-lane branch), so the compiler can
        // count them in flight and wait only for the block it is about to use.
        // Rows past the group read a valid row (their coefficient is zero, their
        // partial dot is discarded); columns past ld read a valid column (x is
        // zero there and their q entries are not stored).
        int64_t colc[CW];
#pragma unroll
        for (int cw = 0; cw < CW; ++cw) {
            int64_t c = c0 + (int64_t)cw * 64 * V + (int64_t)lane * V;
            colc[cw] = c < a.ld ? c : a.ld - V;
        }
        auto load_block = [&](int64_t k, v4u32 (&slot)[RW][CW]) {
#pragma unroll
            for (int i = 0; i < RW; ++i) {
                int64_t row = g_row0 + k * RB + wave * RW + i;
                if (row > a.Nd - 1) row = a.Nd - 1;
                const T *gr = G + row * a.ld;
#pragma unroll
                for (int cw = 0; cw < CW; ++cw)
                    slot[i][cw] = __builtin_nontemporal_load(
                        reinterpret_cast<const v4u32 *>(gr + colc[cw]));
            }
        };

        auto phase_a = [&](int64_t k, v4u32 (&slot)[RW][CW]) {
            double pd[RW];
#pragma unroll
            for (int i = 0; i < RW; ++i) {
                double s = 0.0;
#pragma unroll
                for (int cw = 0; cw < CW; ++cw) {
                    double g[V];
                    unpack<T>(slot[i][cw], g);
#pragma unroll
                    for (int q = 0; q < V; ++q) s = fma(g[q], xr[cw][q], s);
                }
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
                pd[i] = s;
            }
            const int sl = (int)(k % DS);
            // The slot's previous block (k - DS) was copied out by the sync
            // wave before it fetched block k-1-L, which this wave waited for.
            if (lane == 0) {
#pragma unroll
                for (int i = 0; i < RW; ++i) p_lds[sl][wave * RW + i] = pd[i];
            }
            compiler_fence();
            if (lane == 0) __hip_atomic_fetch_add(&pcount[sl], 1u, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_WORKGROUP);
        };

        auto phase_b = [&](int64_t j, v4u32 (&slot)[RW][CW]) {
            const int sl = (int)(j % DS);
            wt.begin();
            while (!wt.failed && lds_load(&cready[sl]) < (unsigned)(j + 1))
                if (!wt.keep_waiting(2, a.err)) break;
            compiler_fence();
#pragma unroll
            for (int i = 0; i < RW; ++i) {
                double cr = c_lds[sl][wave * RW + i];
#pragma unroll
                for (int cw = 0; cw < CW; ++cw) {
                    double g[V];
                    unpack<T>(slot[i][cw], g);
#pragma unroll
                    for (int q = 0; q < V; ++q) zacc[cw][q] = fma(g[q], cr, zacc[cw][q]);
                }
            }
        };

        // Every step issues exactly one block of loads (indexes past the last
        // block re-read the last block), so the number of loads in flight
        // behind each ring slot is a compile-time constant.
        const int64_t last = nrb - 1;
        // prologue: blocks 0..PF-1
#pragma unroll
        for (int p = 0; p < PF; ++p) load_block(p < last ? p : last, ring[p]);
        // steady state; block k lives in register slot k % DR
        for (int64_t kk = 0; kk < nrb + L; kk += DR) {
#pragma unroll
            for (int s = 0; s < DR; ++s) {
                const int64_t k = kk + s;
                // slot (s+PF)%DR held block k+PF-DR = k-L-1, consumed at step k-1
                load_block(k + PF < last ? k + PF : last, ring[(s + PF) % DR]);
                if (k < nrb) phase_a(k, ring[s]);
                if (k >= L && k - L < nrb) phase_b(k - L, ring[(s + DR - L) % DR]);
            }
        }
#pragma unroll
        for (int cw = 0; cw < CW; ++cw)
#pragma unroll
            for (int q = 0; q < V; ++q) z_lds[wave][cw * 64 * V + lane * V + q] = zacc[cw][q];
    } else {
        // ------------------------------------------- sync wave and reduce wave
        Waiter wt{a.err, a.wait_ticks};
        const int NT = a.NT;
        const bool row_lane = lane < RB;
        // coefficient prefetch: the load for the block fetched at step s is
        // issued at step s - PQ into register slot s % PQ
        constexpr int PQ = L > 3 ? 2 : 1;
        double pre[PQ];
#pragma unroll
        for (int u = 0; u < PQ; ++u) pre[u] = 0.0;
        auto coef_row = [&](int64_t f, int64_t &row) {
            row = g_row0 + f * RB + lane;
            return row_lane && row < g_row1;
        };
        auto issue_prefetch = [&](int64_t f, double &slot) {
            int64_t row;
            slot = coef_row(f, row) ? ld_agent(a.coef + row) : 0.0;
        };
        auto fetch = [&](int64_t f, double prefetched) {
            int64_t row;
            const bool mine = coef_row(f, row);
            double v = prefetched;
            wt.begin();
            while (!wt.failed && !__all(!(mine && is_sentinel(v)))) {
                if (!wt.keep_waiting(8, a.err)) break;
                v = mine ? ld_agent(a.coef + row) : 0.0;
            }
            const int sl = (int)(f % DS);
            if (row_lane) c_lds[sl][lane] = v;
            compiler_fence();
            // c_lds[sl] previously held block f-DS: every compute wave has
            // posted block f+L-1 (the post of the previous step), hence
            // finished phase B of block f-2 >= f-DS.
            if (lane == 0) lds_store(&cready[sl], (unsigned)(f + 1));
        };
        auto post = [&](int64_t j) {
            const int sl = (int)(j % DS);
            const unsigned want = (unsigned)(kCompute * (j / DS + 1));
            wt.begin();
            while (!wt.failed && lds_load(&pcount[sl]) < want)
                if (!wt.keep_waiting(4, a.err)) break;
            compiler_fence();
            // the slot's previous block (j - DS) was copied at step j - DS
            const double v = row_lane ? p_lds[sl][lane] : 0.0;
            if (row_lane) st_agent(a.partial + ((rb_base + j) * NT + tile) * RB + lane, v);
        };
        auto reduce = [&](int64_t r) {
            // lane l: row l % RB, tiles l / RB + m * (64 / RB), in tile order;
            // then the lanes of a row combine with a fixed butterfly
            constexpr int PER = 64 / RB;
            constexpr int BATCH = 16;
            const int rr = lane % RB, t0 = lane / RB;
            const int64_t rb = rb_base + r;
            const double *pb = a.partial + rb * NT * RB;
            double y = 0.0;
            for (int tb = t0; tb - t0 < NT; tb += PER * BATCH) {
                double v[BATCH];
                bool need[BATCH];
#pragma unroll
                for (int q = 0; q < BATCH; ++q) {
                    const int tt = tb + q * PER;
                    need[q] = tt < NT;
                    v[q] = ld_agent(pb + (int64_t)(need[q] ? tt : 0) * RB + rr);
                }
                wt.begin();
                for (;;) {
                    bool ok = true;
#pragma unroll
                    for (int q = 0; q < BATCH; ++q) ok = ok && !(need[q] && is_sentinel(v[q]));
                    if (wt.failed || __all(ok)) break;
                    if (!wt.keep_waiting(16, a.err)) break;
#pragma unroll
                    for (int q = 0; q < BATCH; ++q)
                        if (need[q] && is_sentinel(v[q]))
                            v[q] = ld_agent(pb + (int64_t)(tb + q * PER) * RB + rr);
                }
#pragma unroll
                for (int q = 0; q < BATCH; ++q)
                    if (need[q]) y += v[q];
            }
#pragma unroll
            for (int off = RB; off < 64; off <<= 1) y += __shfl_xor(y, off, 64);
            const int64_t row = g_row0 + r * RB + rr;
            double t2 = 0.0;
            if (row_lane && row < g_row1) {
                const double wd = a.Wd ? a.Wd[row] : 1.0;
                const double td = wd * y;
                a.tdat[row] = td;               // read by later kernels only
                st_agent(a.coef + row, wd * td);
                t2 = td * td;
            }
#pragma unroll
            for (int off = 1; off < RB; off <<= 1) t2 += __shfl_xor(t2, off, 64);
            if (lane == 0) a.t2part[rb] = t2;
        };
        if (wave == kCompute) {
            // sync wave: step s hands block s-L's coefficients to the compute
            // waves and posts block s's partial dots
            for (int64_t s0 = 0; s0 < nrb + L; s0 += PQ) {
#pragma unroll
                for (int u = 0; u < PQ; ++u) {
                    const int64_t s = s0 + u;
                    if (s >= nrb + L) break;
                    const int64_t f = s - L;
                    if (f >= 0 && f < nrb) fetch(f, pre[u]);   // issued at step s - PQ
                    if (s < nrb) post(s);
                    const int64_t fn = s + PQ - L;              // fetched at step s + PQ
                    if (fn >= 0 && fn < nrb) issue_prefetch(fn, pre[u]);
                }
            }
        } else {
            // reduce wave: the blocks r with r mod NT == tile, in order; each
            // waits only for the posts of block r, so reductions never hold
            // up this workgroup's posts
            for (int64_t r = tile; r < nrb; r += NT) reduce(r);
        }
    }
    __syncthreads();
    // q tile of this row group: the compute waves' partials in wave order
    for (int j = threadIdx.x; j < TW; j += kThreads) {
        int64_t c = c0 + j;
        if (c >= a.ld) continue;
        double z = z_lds[0][j];
        for (int w = 1; w < kCompute; ++w) z += z_lds[w][j];
        a.zpart[(int64_t)group * a.ld + c] = z;
    }
}

__global__ void fill_u64(uint64_t *p, int64_t n, uint64_t v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Lag depths L (blocks between posting a block's partial dots and applying
// its coefficients) with their prefetch depths PF: the register ring holds
// L+1+PF blocks.
constexpr int kNumDepths = 7;
constexpr int kDepthL[kNumDepths] = {3, 4, 5, 6, 6, 6, 8};   // prefetch 2 2 1 1 2 3 2

template <typename T, int CW, int RW> struct FusedCfg {
    static constexpr int V = VecN<T>::n;
    static constexpr int TW = 64 * V * CW;
    static constexpr int RB = kCompute * RW;
    static const void *kernel(int depth_index)
    {
        switch (depth_index) {
        case 0: return (const void *)fused_normal_kernel<T, CW, RW, 3, 2>;
        case 1: return (const void *)fused_normal_kernel<T, CW, RW, 4, 2>;
        case 2: return (const void *)fused_normal_kernel<T, CW, RW, 5, 1>;
        case 3: return (const void *)fused_normal_kernel<T, CW, RW, 6, 1>;
        case 4: return (const void *)fused_normal_kernel<T, CW, RW, 6, 2>;
        case 5: return (const void *)fused_normal_kernel<T, CW, RW, 6, 3>;
        default: return (const void *)fused_normal_kernel<T, CW, RW, 8, 2>;
        }
    }
};

template <class C> FusedPlan make_plan(int64_t Nd, int64_t ld, int target_wgs, int depth_index)
{
    FusedPlan p;
    p.NT = (int)((ld + C::TW - 1) / C::TW);
    int S = target_wgs / p.NT;             // NT * S <= target_wgs when NT <= target_wgs
    int64_t max_groups = (Nd + C::RB - 1) / C::RB;
    if (S > max_groups) S = (int)max_groups;
    if (S < 1) S = 1;
    p.S = S;
    p.rows_per_group = ((Nd + S - 1) / S + C::RB - 1) / C::RB * C::RB;
    p.nrb_group = p.rows_per_group / C::RB;
    p.RB = C::RB;
    p.threads = kThreads;
    p.kernel = C::kernel(depth_index);
    return p;
}

// Per storage type: columns per lane (in 16-byte vectors), rows per compute
// wave.
using CfgF64 = FusedCfg<double, 2, 2>;
using CfgF32 = FusedCfg<float, 1, 2>;
using CfgBf16 = FusedCfg<__hip_bfloat16, 1, 2>;

}  // namespace

static bool plan_variant(gra_dtype dt, int64_t Nd, int64_t ld, int cus, int target_wgs, int di,
                         FusedPlan *out)
{
    const void *kern;
    switch (dt) {
    case GRA_F64: kern = CfgF64::kernel(di); break;
    case GRA_F32: kern = CfgF32::kernel(di); break;
    default: kern = CfgBf16::kernel(di); break;
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kThreads, 0) != hipSuccess)
        return false;
    // Two six-wave workgroups share a CU only when their waves balance over
    // the four SIMDs, so a grid that fills every reported slot may not be
    // fully resident, and every row block waits for all tiles of its group.
    // A group of at most half the reported slots fills the device (on the
    // 5,000 x 32,000 G: FP64 2 groups of 125 tiles at one workgroup per CU,
    // BF16 16 groups of 32; stopping cus/32 workgroups short of the limit
    // halves the groups there and runs 1.7x slower).  A wider group stays
    // cus/32 workgroups below the limit (a 512-tile group at the limit
    // exceeded the bounded wait) and, if that does not fit, the two-pass
    // iteration is used.
    const int full = per_cu * cus;
    auto plan_for = [&](int lim) {
        int target = target_wgs > 0 ? target_wgs : 2 * cus;
        if (target > lim) target = lim;
        switch (dt) {
        case GRA_F64: return make_plan<CfgF64>(Nd, ld, target, di);
        case GRA_F32: return make_plan<CfgF32>(Nd, ld, target, di);
        default: return make_plan<CfgBf16>(Nd, ld, target, di);
        }
    };
    FusedPlan p = plan_for(full);
    int limit = full;
    if (2 * p.NT > full) {
        limit = full - cus / 32;
        p = plan_for(limit);
    }
    if ((int64_t)p.NT * p.S > (int64_t)limit) return false;
    p.L = kDepthL[di];
    *out = p;
    return true;
}

bool fused_plan(gra_dtype dt, int64_t Nd, int64_t ld, int device, int target_wgs, int depth,
                FusedPlan *out)
{
    int cus = 0, coop = 0, clock_khz = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return false;
    if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device) != hipSuccess ||
        !coop)
        return false;
    if (hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess ||
        clock_khz <= 0)
        return false;
    // depth: 0 = the storage type's default (fastest in a sweep on the
    // 5,000 x 32,000 G), else an index + 1 into the (L, PF) table; if the grid
    // cannot be co-resident (wide G, many column tiles) the leaner (5, 1)
    // variant is tried before giving up
    const int def = dt == GRA_F64 ? 6 : (dt == GRA_F32 ? 2 : 4);
    const int di = depth >= 1 && depth <= kNumDepths ? depth - 1 : def;
    FusedPlan p;
    if (!plan_variant(dt, Nd, ld, cus, target_wgs, di, &p) &&
        !(depth == 0 && plan_variant(dt, Nd, ld, cus, target_wgs, 2, &p)))
        return false;
    p.wait_ticks = (uint64_t)clock_khz * 2000ull;   // 2 s
    *out = p;
    return true;
}

size_t fused_scratch_bytes(const FusedPlan &p, int64_t ld)
{
    int64_t nrb = (int64_t)p.S * p.nrb_group;
    return sizeof(double) * (size_t)(nrb * p.NT * p.RB + nrb) +   // partial, t2part
           sizeof(double) * (size_t)p.S * (size_t)ld;               // zpart
}

hipError_t launch_fused(const FusedPlan &p, const void *G, int64_t Nd, int64_t ld,
                        const double *x, const double *Wd, double *coef, double *tdat,
                        void *scratch, int *err, const int *done, double **zpart_out,
                        double **t2part_out, hipStream_t st)
{
    const int64_t nrb = (int64_t)p.S * p.nrb_group;
    char *b = static_cast<char *>(scratch);
    FusedArgs a;
    a.G = G;
    a.Nd = Nd;
    a.ld = ld;
    a.NT = p.NT;
    a.S = p.S;
    a.rows_per_group = p.rows_per_group;
    a.nrb_group = p.nrb_group;
    a.x = x;
    a.Wd = Wd;
    a.partial = reinterpret_cast<double *>(b);
    b += sizeof(double) * (size_t)(nrb * p.NT * p.RB);
    a.t2part = reinterpret_cast<double *>(b);
    b += sizeof(double) * (size_t)nrb;
    a.zpart = reinterpret_cast<double *>(b);
    a.coef = coef;
    a.tdat = tdat;
    a.err = err;
    a.done = done;
    a.wait_ticks = p.wait_ticks;
    const int64_t npart = nrb * p.NT * p.RB;
    hipLaunchKernelGGL(fill_u64, dim3((unsigned)((npart + 255) / 256)), dim3(256), 0, st,
                       reinterpret_cast<uint64_t *>(a.partial), npart, kSentinel);
    hipLaunchKernelGGL(fill_u64, dim3((unsigned)((Nd + 255) / 256)), dim3(256), 0, st,
                       reinterpret_cast<uint64_t *>(coef), Nd, kSentinel);
    void *args[] = {&a};
    hipError_t e = hipLaunchCooperativeKernel(p.kernel, dim3((unsigned)(p.NT * p.S)),
                                              dim3(p.threads), args, 0, st);
    if (zpart_out) *zpart_out = a.zpart;
    if (t2part_out) *t2part_out = a.t2part;
    return e;
}

}  // namespace gra
