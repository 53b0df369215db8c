// FP64 prism kernels: dense G assembly (replaces Fwd::compute_G,
// src/Forward_modelling/Fwd.cpp:47-79) and the matrix-free forward sum
// (replaces compute.cpp:246-260).  Built with -ffp-contract=off so every
// expression of prism_math.hpp rounds as the reference's does.
//
// Work mapping: one thread per (point, prism); a 256-thread block covers 256
// consecutive prisms of one point, so G stores are contiguous along a row of
// the row-major G.  The field bits are a template parameter for the common
// component sets, so unused faces compile away and register pressure drops.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include "engine.hpp"
#include "prism_math.hpp"

namespace gra {

namespace {

template <typename T> __device__ inline T cvt(double v);
template <> __device__ inline double cvt<double>(double v) { return v; }
template <> __device__ inline float cvt<float>(double v) { return (float)v; }
template <> __device__ inline __hip_bfloat16 cvt<__hip_bfloat16>(double v)
{
    return __float2bfloat16((float)v);
}

constexpr int kBlock = 256;

// FLAG == 0 means "runtime flag" (generic instantiation).
template <uint32_t FLAG, typename T>
__global__ void __launch_bounds__(kBlock)
assemble_kernel(const double *__restrict__ prisms, int64_t Nm, const double *__restrict__ pts,
                int64_t N_obs, uint32_t rflag, T *__restrict__ G, int64_t ld,
                int64_t blocks_per_row, int *__restrict__ bad)
{
    const uint32_t flag = FLAG ? FLAG : rflag;
    int64_t b = blockIdx.x;
    int64_t i = b / blocks_per_row;
    int64_t j = (b - i * blocks_per_row) * kBlock + threadIdx.x;
    if (i >= N_obs || j >= Nm) return;
    double pt[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    double pr[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) pr[q] = prisms[6 * j + q];
    double out[10];
    int st;
    if constexpr (FLAG != 0u) st = prism_fields_t<FLAG>(pt, pr, 1.0, out);
    else st = prism_fields(pt, pr, 1.0, flag, out);
    if (st) atomicOr(bad, 1);
    int k = 0;
#pragma unroll
    for (int bit = 1; bit < 10; ++bit) {
        if (flag & (1u << bit)) {
            G[((int64_t)k * N_obs + i) * ld + j] = cvt<T>(out[bit]);
            ++k;
        }
    }
}

template <uint32_t FLAG>
__global__ void __launch_bounds__(kBlock)
forward_kernel(const double *__restrict__ prisms, const double *__restrict__ rho, int64_t Nm,
               const double *__restrict__ pts, int64_t N_obs, uint32_t rflag,
               double *__restrict__ out, int *__restrict__ bad)
{
    const uint32_t flag = FLAG ? FLAG : rflag;
    __shared__ double red[kBlock / 64][10];
    int64_t i = blockIdx.x;
    if (i >= N_obs) return;
    double pt[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    double acc[10];
#pragma unroll
    for (int q = 0; q < 10; ++q) acc[q] = 0.0;
    int st = 0;
    for (int64_t j = threadIdx.x; j < Nm; j += kBlock) {
        double pr[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) pr[q] = prisms[6 * j + q];
        double f[10];
        if constexpr (FLAG != 0u) st |= prism_fields_t<FLAG>(pt, pr, rho[j], f);
        else st |= prism_fields(pt, pr, rho[j], flag, f);
#pragma unroll
        for (int bit = 1; bit < 10; ++bit)
            if (flag & (1u << bit)) acc[bit] += f[bit];
    }
    if (st) atomicOr(bad, 1);
    // Deterministic reduction: fixed xor-butterfly inside each wave, then the
    // waves' results summed in wave order by thread 0.
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int bit = 1; bit < 10; ++bit) {
        if (!(flag & (1u << bit))) continue;
        double v = acc[bit];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) red[wave][bit] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int k = 0;
        int nc = __popc(flag & kAllFields);
        for (int bit = 1; bit < 10; ++bit) {
            if (!(flag & (1u << bit))) continue;
            double s = red[0][bit];
            for (int w = 1; w < kBlock / 64; ++w) s += red[w][bit];
            out[i * nc + k] = s;
            ++k;
        }
    }
}

// Many G assemblies of uneven sizes in one launch: block b belongs to the
// last item whose first block is <= b (items sorted by block0); inside an
// item the mapping is assemble_kernel's, with the item's runtime flag.
template <typename T>
__global__ void __launch_bounds__(kBlock)
assemble_batch_kernel(const BatchItem *__restrict__ items, int n_items, int *__restrict__ bad)
{
    __shared__ int s_item;
    const int64_t b = blockIdx.x;
    if (threadIdx.x == 0) {
        int lo = 0, hi = n_items - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (items[mid].block0 <= b) lo = mid;
            else hi = mid - 1;
        }
        s_item = lo;
    }
    __syncthreads();
    const BatchItem it = items[s_item];
    const int64_t lb = b - it.block0;
    const int64_t i = lb / it.bpr;
    const int64_t j = (lb - i * it.bpr) * kBlock + threadIdx.x;
    if (i >= it.N_obs || j >= it.Nm) return;
    double pt[3] = {it.pts[3 * i], it.pts[3 * i + 1], it.pts[3 * i + 2]};
    double pr[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) pr[q] = it.prisms[6 * j + q];
    double out[10];
    if (prism_fields(pt, pr, 1.0, it.flag, out)) atomicOr(bad + s_item, 1);
    T *G = static_cast<T *>(it.G);
    int k = 0;
#pragma unroll
    for (int bit = 1; bit < 10; ++bit) {
        if (it.flag & (1u << bit)) {
            G[((int64_t)k * it.N_obs + i) * it.ld + j] = cvt<T>(out[bit]);
            ++k;
        }
    }
}

template <typename T>
hipError_t assemble_dispatch(const double *prisms, int64_t Nm, const double *pts, int64_t N_obs,
                             uint32_t flag, T *G, int64_t ld, int *bad, hipStream_t st)
{
    int64_t bpr = (Nm + kBlock - 1) / kBlock;
    int64_t nb = bpr * N_obs;
    if (nb <= 0) return hipSuccess;
    dim3 grid((unsigned)nb), block(kBlock);
    switch (flag) {
    case kGz | kTzz:
        hipLaunchKernelGGL((assemble_kernel<kGz | kTzz, T>), grid, block, 0, st, prisms, Nm, pts,
                           N_obs, flag, G, ld, bpr, bad);
        break;
    case kGz:
        hipLaunchKernelGGL((assemble_kernel<kGz, T>), grid, block, 0, st, prisms, Nm, pts, N_obs,
                           flag, G, ld, bpr, bad);
        break;
    case kTzz:
        hipLaunchKernelGGL((assemble_kernel<kTzz, T>), grid, block, 0, st, prisms, Nm, pts, N_obs,
                           flag, G, ld, bpr, bad);
        break;
    case kAllFields:
        hipLaunchKernelGGL((assemble_kernel<kAllFields, T>), grid, block, 0, st, prisms, Nm, pts,
                           N_obs, flag, G, ld, bpr, bad);
        break;
    default:
        hipLaunchKernelGGL((assemble_kernel<0u, T>), grid, block, 0, st, prisms, Nm, pts, N_obs,
                           flag, G, ld, bpr, bad);
        break;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_assemble(const double *d_prisms, int64_t Nm, const double *d_pts, int64_t N_obs,
                           uint32_t flag, gra_dtype store, void *G, int64_t ld, int *d_bad,
                           hipStream_t st)
{
    switch (store) {
    case GRA_F64:
        return assemble_dispatch<double>(d_prisms, Nm, d_pts, N_obs, flag, (double *)G, ld, d_bad,
                                         st);
    case GRA_F32:
        return assemble_dispatch<float>(d_prisms, Nm, d_pts, N_obs, flag, (float *)G, ld, d_bad,
                                        st);
    default:
        return assemble_dispatch<__hip_bfloat16>(d_prisms, Nm, d_pts, N_obs, flag,
                                                 (__hip_bfloat16 *)G, ld, d_bad, st);
    }
}

hipError_t launch_assemble_batch(const BatchItem *d_items, int n_items, int64_t total_blocks,
                                 gra_dtype store, int *d_bad, hipStream_t st)
{
    if (total_blocks <= 0) return hipSuccess;
    dim3 grid((unsigned)total_blocks), block(kBlock);
    switch (store) {
    case GRA_F64:
        hipLaunchKernelGGL(assemble_batch_kernel<double>, grid, block, 0, st, d_items, n_items,
                           d_bad);
        break;
    case GRA_F32:
        hipLaunchKernelGGL(assemble_batch_kernel<float>, grid, block, 0, st, d_items, n_items,
                           d_bad);
        break;
    default:
        hipLaunchKernelGGL(assemble_batch_kernel<__hip_bfloat16>, grid, block, 0, st, d_items,
                           n_items, d_bad);
        break;
    }
    return hipGetLastError();
}

hipError_t launch_forward(const double *d_prisms, const double *d_rho, int64_t Nm,
                          const double *d_pts, int64_t N_obs, uint32_t flag, double *d_out,
                          int *d_bad, hipStream_t st)
{
    if (N_obs <= 0) return hipSuccess;
    dim3 grid((unsigned)N_obs), block(kBlock);
    switch (flag) {
    case kGz | kTzz:
        hipLaunchKernelGGL((forward_kernel<kGz | kTzz>), grid, block, 0, st, d_prisms, d_rho, Nm,
                           d_pts, N_obs, flag, d_out, d_bad);
        break;
    case kAllFields:
        hipLaunchKernelGGL((forward_kernel<kAllFields>), grid, block, 0, st, d_prisms, d_rho, Nm,
                           d_pts, N_obs, flag, d_out, d_bad);
        break;
    default:
        hipLaunchKernelGGL((forward_kernel<0u>), grid, block, 0, st, d_prisms, d_rho, Nm, d_pts,
                           N_obs, flag, d_out, d_bad);
        break;
    }
    return hipGetLastError();
}

}  // namespace gra
