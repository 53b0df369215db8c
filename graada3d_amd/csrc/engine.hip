// C-ABI entry points of libgraeng (declared in include/graeng.h).
//
// Each entry point validates its arguments on the host, takes the context
// lock, enqueues device work on the context's own stream and, for calls with
// host outputs, synchronises that stream before returning (the reference's
// Fwd / solve_cg calls are synchronous).  Errors come back as gra_status with
// a message in gra_last_error(); nothing throws across the boundary.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "engine.hpp"

using namespace gra;

namespace gra {

gra_status set_error(Ctx &c, gra_status s, const std::string &msg)
{
    c.last_error = msg;
    return s;
}

gra_status hip_fail(Ctx &c, hipError_t e, const char *what)
{
    c.last_error = std::string(what) + ": " + hipGetErrorString(e);
    if (e == hipErrorOutOfMemory) return GRA_OOM;
    return GRA_HIP_ERROR;
}

TimedRegion::TimedRegion(Ctx &ctx, int klass, int iter) : c(ctx), cls(klass)
{
    if (klass < 0 || c.timing <= 0 || iter % c.timing != 0) return;   // -1: untimed
    while (c.ev_used + 2 > c.ev_pool.size()) {
        hipEvent_t ev;
        // timing only: read after a stream synchronisation, so the record
        // needs no system-scope cache writeback (which would slow the kernels
        // around it)
        if (hipEventCreateWithFlags(&ev, hipEventDisableSystemFence) != hipSuccess) return;
        c.ev_pool.push_back(ev);
    }
    a = c.ev_pool[c.ev_used++];
    b = c.ev_pool[c.ev_used++];
    (void)hipEventRecord(a, c.stream);
}

TimedRegion::~TimedRegion()
{
    if (!a) return;
    (void)hipEventRecord(b, c.stream);
    c.ev_marks.push_back({cls, {a, b}});
}

namespace {

// Host CSR -> device CSR plus its transpose (columns ascending per row).
void transpose_csr(const gra_csr *h, std::vector<int64_t> &rp, std::vector<int32_t> &ci,
                   std::vector<double> &v)
{
    int64_t n = h->n, nnz = h->nnz;
    rp.assign((size_t)n + 1, 0);
    ci.resize((size_t)nnz);
    v.resize((size_t)nnz);
    for (int64_t k = 0; k < nnz; ++k) rp[(size_t)h->col_idx[k] + 1]++;
    for (int64_t i = 0; i < n; ++i) rp[(size_t)i + 1] += rp[(size_t)i];
    std::vector<int64_t> pos(rp.begin(), rp.end() - 1);
    for (int64_t i = 0; i < n; ++i) {
        for (int64_t k = h->row_ptr[i]; k < h->row_ptr[i + 1]; ++k) {
            int64_t j = h->col_idx[k];
            int64_t q = pos[(size_t)j]++;
            ci[(size_t)q] = (int32_t)i;
            v[(size_t)q] = h->val[k];
        }
    }
}

bool csr_valid(const gra_csr *h, int64_t n)
{
    if (!h || h->n != n || h->nnz < 0 || !h->row_ptr) return false;
    if (h->nnz > 0 && (!h->col_idx || !h->val)) return false;
    if (h->row_ptr[0] != 0 || h->row_ptr[n] != h->nnz) return false;
    for (int64_t i = 0; i < n; ++i) {
        if (h->row_ptr[i + 1] < h->row_ptr[i]) return false;
        for (int64_t k = h->row_ptr[i]; k < h->row_ptr[i + 1]; ++k)
            if (h->col_idx[k] < 0 || h->col_idx[k] >= n) return false;
    }
    return true;
}

template <typename T>
hipError_t put(DevBuf &b, const T *h, size_t n, hipStream_t st)
{
    hipError_t e = b.reserve(sizeof(T) * (n ? n : 1));
    if (e != hipSuccess) return e;
    if (n) e = hipMemcpyAsync(b.p, h, sizeof(T) * n, hipMemcpyHostToDevice, st);
    return e;
}

// Uploads an Nm-vector into a buffer of ld entries with zero padding.
hipError_t put_padded(DevBuf &b, const double *h, int64_t n, int64_t ld, hipStream_t st)
{
    hipError_t e = b.reserve(sizeof(double) * (size_t)ld);
    if (e != hipSuccess) return e;
    if (ld > n) {
        e = hipMemsetAsync(b.as<double>() + n, 0, sizeof(double) * (size_t)(ld - n), st);
        if (e != hipSuccess) return e;
    }
    return hipMemcpyAsync(b.p, h, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, st);
}

bool prisms_valid(const double *pr, int64_t n)
{
    for (int64_t j = 0; j < n; ++j) {
        const double *p = pr + 6 * j;
        for (int q = 0; q < 6; ++q)
            if (!std::isfinite(p[q])) return false;
        if (!(p[0] < p[1] && p[2] < p[3] && p[4] < p[5])) return false;
    }
    return true;
}

double max_point_z(const double *pts, int64_t n)
{
    double z = pts[2];
    for (int64_t i = 1; i < n; ++i) z = std::max(z, pts[3 * i + 2]);
    return z;
}

bool points_valid(const double *pt, int64_t n)
{
    for (int64_t i = 0; i < 3 * n; ++i)
        if (!std::isfinite(pt[i])) return false;
    return true;
}

int ncomp(uint32_t flag) { return __builtin_popcount(flag & 0x3FEu); }

bool increasing(const double *v, int64_t n)
{
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(v[i]) || (i && !(v[i] > v[i - 1]))) return false;
    return true;
}

size_t dtype_size(gra_dtype d) { return d == GRA_F64 ? 8 : (d == GRA_F32 ? 4 : 2); }

// A G (dense or wavelet-compressed) is assembled.
bool have_G(const Ctx &c) { return c.Nd > 0 && (c.wav || c.G.p); }

}  // namespace

GemvShape gemv_shape(const Ctx &c)
{
    GemvShape s = default_gemv_shape(c.store, c.Nd, c.Nm);
    const GemvShape &o = c.tune_override;
    if (o.n_rows) s.n_rows = o.n_rows;
    if (o.n_unroll) s.n_unroll = o.n_unroll;
    if (o.n_splits) s.n_splits = o.n_splits;
    if (o.t_chunks) s.t_chunks = o.t_chunks;
    if (o.t_unroll) s.t_unroll = o.t_unroll;
    if (o.nontemporal >= 0) s.nontemporal = o.nontemporal;
    if (o.n_block >= 0) s.n_block = o.n_block;
    if (c.wav) s.n_splits = s.t_chunks = 1;   // the compressed products write whole sums
    return s;
}

gra_status upload_csr(Ctx &c, DevCsr &d, const gra_csr *h)
{
    std::vector<int64_t> rpT;
    std::vector<int32_t> ciT;
    std::vector<double> vT;
    transpose_csr(h, rpT, ciT, vT);
    hipError_t e;
    if ((e = put(d.rp, h->row_ptr, (size_t)h->n + 1, c.stream)) != hipSuccess ||
        (e = put(d.ci, h->col_idx, (size_t)h->nnz, c.stream)) != hipSuccess ||
        (e = put(d.v, h->val, (size_t)h->nnz, c.stream)) != hipSuccess ||
        (e = put(d.rpT, rpT.data(), rpT.size(), c.stream)) != hipSuccess ||
        (e = put(d.ciT, ciT.data(), ciT.size(), c.stream)) != hipSuccess ||
        (e = put(d.vT, vT.data(), vT.size(), c.stream)) != hipSuccess)
        return hip_fail(c, e, "upload CSR");
    // host vectors die at return: finish the copies first
    if ((e = hipStreamSynchronize(c.stream)) != hipSuccess) return hip_fail(c, e, "upload CSR");
    d.n = h->n;
    d.nnz = h->nnz;
    d.present = true;
    return GRA_OK;
}

}  // namespace gra

#define CTX_GUARD(ctx)                                  \
    if (!(ctx)) return GRA_BAD_ARG;                     \
    Ctx &c = (ctx)->c;                                  \
    std::lock_guard<std::mutex> lock_(c.mu);            \
    c.last_error.clear();                               \
    if (hipSetDevice(c.device) != hipSuccess)           \
        return set_error(c, GRA_HIP_ERROR, "hipSetDevice failed");

#define HIPCHK(expr, what)                                  \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return hip_fail(c, e_, what); \
    } while (0)

extern "C" {

const char *gra_status_string(int s)
{
    switch (s) {
    case GRA_OK: return "ok";
    case GRA_BAD_ARG: return "bad argument";
    case GRA_OOM: return "out of memory";
    case GRA_HIP_ERROR: return "GPU runtime error";
    case GRA_NUMERIC_BREAKDOWN: return "numeric breakdown";
    case GRA_DEGENERATE_GEOMETRY: return "degenerate geometry";
    case GRA_NO_DEVICE: return "no GPU device";
    default: return "unknown status";
    }
}

// ------------------------------------------- allocation-free G x (scratch)
// gra_G_vec_mul_noalloc stages x and keeps the column-split partial sums of
// G x in this array instead of context buffers.  Each GPU holds its own copy
// (32 MiB) from the module load until the process ends; host code reaches it
// only through the device address hipGetSymbolAddress returns for the
// current device, resolved once per device when a context is created.
namespace gra {
namespace {
constexpr int64_t kScratchDoubles = int64_t(4) << 20;     // 32 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex scratch_mu;
double *scratch_dev[64] = {};      // per device ordinal, null until resolved

// With `device` current: the device address of its `scratch`.
hipError_t scratch_address(int device, double **out)
{
    if (device < 0 || device >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        scratch_dev[device] = static_cast<double *>(p);
    }
    *out = scratch_dev[device];
    return hipSuccess;
}

// out[i] = x[i] for i < n, 0 up to ld (the G x kernels read ld entries).
__global__ void k_stage_padded(const double *__restrict__ x, int64_t n, int64_t ld,
                               double *__restrict__ out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ld) out[i] = i < n ? x[i] : 0.0;
}
}  // namespace
}  // namespace gra

const char *gra_last_error(const gra_ctx *ctx) { return ctx ? ctx->c.last_error.c_str() : ""; }

gra_status gra_ctx_create(int device, gra_ctx **out)
{
    if (!out) return GRA_BAD_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return GRA_NO_DEVICE;
    if (device < 0 || device >= n) return GRA_BAD_ARG;
    if (hipSetDevice(device) != hipSuccess) return GRA_HIP_ERROR;
    gra_ctx *ctx = new (std::nothrow) gra_ctx();
    if (!ctx) return GRA_OOM;
    ctx->c.device = device;
    if (hipStreamCreateWithFlags(&ctx->c.own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return GRA_HIP_ERROR;
    }
    ctx->c.stream = ctx->c.own_stream;
    if (ctx->c.status.reserve(sizeof(int) * 4) != hipSuccess) {
        (void)hipStreamDestroy(ctx->c.own_stream);
        delete ctx;
        return GRA_OOM;
    }
    // the allocation-free entry point's scratch on this device (see above)
    if (scratch_address(device, &ctx->c.noalloc_scratch) != hipSuccess) {
        (void)hipStreamDestroy(ctx->c.own_stream);
        delete ctx;
        return GRA_HIP_ERROR;
    }
    // error messages of short length then need no allocation
    ctx->c.last_error.reserve(256);
    // per-solve results come back through pinned memory, so that result
    // copies queued on the stream do not block the host
    if (hipHostMalloc((void **)&ctx->c.host_stage, sizeof(HostStage), hipHostMallocDefault) !=
        hipSuccess) {
        (void)hipStreamDestroy(ctx->c.own_stream);
        delete ctx;
        return GRA_OOM;
    }
    *out = ctx;
    return GRA_OK;
}

void gra_ctx_destroy(gra_ctx *ctx)
{
    if (!ctx) return;
    {
        Ctx &c = ctx->c;
        (void)hipSetDevice(c.device);
        (void)hipStreamSynchronize(c.stream);
        for (hipEvent_t ev : c.ev_pool) (void)hipEventDestroy(ev);
        c.ev_pool.clear();
        if (c.batch_event) (void)hipEventDestroy(c.batch_event);
        if (c.host_stage) (void)hipHostFree(c.host_stage);
        (void)hipStreamDestroy(c.own_stream);
    }
    delete ctx;
}

gra_status gra_ctx_set_stream(gra_ctx *ctx, void *stream)
{
    CTX_GUARD(ctx);
    HIPCHK(hipStreamSynchronize(c.stream), "sync before a stream change");
    c.stream = stream ? static_cast<hipStream_t>(stream) : c.own_stream;
    return GRA_OK;
}

void *gra_ctx_stream(const gra_ctx *ctx) { return ctx ? ctx->c.stream : nullptr; }

gra_status gra_sync(gra_ctx *ctx)
{
    CTX_GUARD(ctx);
    HIPCHK(hipStreamSynchronize(c.stream), "sync");
    return GRA_OK;
}

gra_status gra_assemble(gra_ctx *ctx, const double *prisms, int64_t Nm, const double *pts,
                        int64_t N_obs, uint32_t field_flag, gra_dtype store)
{
    CTX_GUARD(ctx);
    if (!prisms || !pts || Nm <= 0 || N_obs <= 0)
        return set_error(c, GRA_BAD_ARG, "empty mesh or point set");
    if ((field_flag & 1u) || (field_flag & ~0x3FFu) || !(field_flag & 0x3FEu))
        return set_error(c, GRA_BAD_ARG, "field_flag must select components among bits 1..9");
    if (store != GRA_F64 && store != GRA_F32 && store != GRA_BF16)
        return set_error(c, GRA_BAD_ARG, "unknown storage type");
    if (Nm > INT32_MAX) return set_error(c, GRA_BAD_ARG, "Nm exceeds int32 column indices");
    if (!prisms_valid(prisms, Nm))
        return set_error(c, GRA_BAD_ARG, "prism bounds must be finite with x0<x1, y0<y1, z0<z1");
    if (!points_valid(pts, N_obs)) return set_error(c, GRA_BAD_ARG, "non-finite point");
    const int nc = ncomp(field_flag);
    const int64_t Nd = (int64_t)nc * N_obs;
    const int64_t ld = (Nm + 63) / 64 * 64;
    c.Nd = c.Nm = c.N_obs = c.ld = 0;          // G is invalid until the kernel has run
    c.gn_ready = false;
    c.wav = false;
    c.partn_holds_gn_model = false;
    HIPCHK(c.G.reserve(dtype_size(store) * (size_t)Nd * (size_t)ld), "allocate G");  // reuses
    if (ld > Nm)
        HIPCHK(hipMemsetAsync(c.G.p, 0, dtype_size(store) * (size_t)Nd * (size_t)ld, c.stream),
               "clear G padding");
    DevBuf dpr;
    DevBuf &dpt = c.pts_dev;
    HIPCHK(put(dpr, prisms, (size_t)(6 * Nm), c.stream), "upload prisms");
    HIPCHK(put(dpt, pts, (size_t)(3 * N_obs), c.stream), "upload points");
    int *bad = c.status.as<int>();
    HIPCHK(hipMemsetAsync(bad, 0, sizeof(int), c.stream), "clear status");
    {
        TimedRegion tr(c, kTimeAssemble);
        HIPCHK(launch_assemble(dpr.as<double>(), Nm, dpt.as<double>(), N_obs, field_flag, store,
                               c.G.p, ld, bad, c.stream),
               "assemble kernel");
    }
    int hbad = 0;
    HIPCHK(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, c.stream), "status");
    HIPCHK(hipStreamSynchronize(c.stream), "assemble");
    c.Nd = Nd; c.Nm = Nm; c.N_obs = N_obs; c.ld = ld; c.flag = field_flag; c.store = store;
    c.g_corners = false;
    ++c.g_version;
    if (hbad)
        return set_error(c, GRA_DEGENERATE_GEOMETRY,
                         "an observation point lies on a prism edge (s0*s1 <= 0, Integral.cpp:77)");
    return GRA_OK;
}

gra_status gra_assemble_rectilinear(gra_ctx *ctx, const double *xs, int64_t nx, const double *ys,
                                   int64_t ny, const double *zs, int64_t nz, const double *pts,
                                   int64_t N_obs, uint32_t field_flag, gra_dtype store)
{
    CTX_GUARD(ctx);
    if (!xs || !ys || !zs || !pts || nx <= 0 || ny <= 0 || nz <= 0 || N_obs <= 0)
        return set_error(c, GRA_BAD_ARG, "empty mesh or point set");
    if ((field_flag & 1u) || (field_flag & ~0x3FFu) || !(field_flag & 0x3FEu))
        return set_error(c, GRA_BAD_ARG, "field_flag must select components among bits 1..9");
    if (store != GRA_F64 && store != GRA_F32 && store != GRA_BF16)
        return set_error(c, GRA_BAD_ARG, "unknown storage type");
    if (!increasing(xs, nx + 1) || !increasing(ys, ny + 1) || !increasing(zs, nz + 1))
        return set_error(c, GRA_BAD_ARG, "node coordinates must be finite and strictly increasing");
    if (nx * ny > (int64_t)INT32_MAX / nz)
        return set_error(c, GRA_BAD_ARG, "Nm exceeds int32 column indices");
    if (!points_valid(pts, N_obs)) return set_error(c, GRA_BAD_ARG, "non-finite point");
    for (int64_t i = 0; i < N_obs; ++i)
        if (!(pts[3 * i + 2] < zs[0]))
            return set_error(c, GRA_BAD_ARG,
                             "points must lie strictly above the mesh (z < zs[0], z down)");
    const int nc = ncomp(field_flag);
    const int tile = c.node_tile > 0 ? c.node_tile : node_assemble_tile(nc, nz);
    if (tile < 1)
        return set_error(c, GRA_BAD_ARG, "too many z levels for the corner-sum assembly");
    if (8.0 * nc * (tile + 1) * (tile + 1) * (double)(nz + 1) > 61440.0)
        return set_error(c, GRA_BAD_ARG, "corner-sum tile too large for LDS");
    const int64_t Nm = nx * ny * nz;
    const int64_t Nd = (int64_t)nc * N_obs;
    const int64_t ld = (Nm + 63) / 64 * 64;
    c.Nd = c.Nm = c.N_obs = c.ld = 0;          // G is invalid until the kernel has run
    c.gn_ready = false;
    c.wav = false;
    c.partn_holds_gn_model = false;
    HIPCHK(c.G.reserve(dtype_size(store) * (size_t)Nd * (size_t)ld), "allocate G");
    if (ld > Nm)
        HIPCHK(hipMemsetAsync(c.G.p, 0, dtype_size(store) * (size_t)Nd * (size_t)ld, c.stream),
               "clear G padding");
    std::vector<double> axes((size_t)(nx + ny + nz + 3));
    std::copy(xs, xs + nx + 1, axes.begin());
    std::copy(ys, ys + ny + 1, axes.begin() + nx + 1);
    std::copy(zs, zs + nz + 1, axes.begin() + nx + ny + 2);
    HIPCHK(put(c.rect_axes, axes.data(), axes.size(), c.stream), "upload axes");
    HIPCHK(put(c.pts_dev, pts, (size_t)(3 * N_obs), c.stream), "upload points");
    const double *da = c.rect_axes.as<double>();
    {
        TimedRegion tr(c, kTimeAssemble);
        HIPCHK(launch_node_assemble(da, da + nx + 1, da + nx + ny + 2, nx, ny, nz,
                                    c.pts_dev.as<double>(), N_obs, field_flag, store, c.G.p, ld,
                                    tile, c.stream),
               "corner-sum assembly");
    }
    HIPCHK(hipStreamSynchronize(c.stream), "assemble");
    c.Nd = Nd; c.Nm = Nm; c.N_obs = N_obs; c.ld = ld; c.flag = field_flag; c.store = store;
    c.g_corners = true;
    c.pts_zmax = max_point_z(pts, N_obs);
    ++c.g_version;
    return GRA_OK;
}

gra_status gra_assemble_corners(gra_ctx *ctx, const double *prisms, int64_t Nm, const double *pts,
                                int64_t N_obs, uint32_t field_flag, gra_dtype store)
{
    CTX_GUARD(ctx);
    if (!prisms || !pts || Nm <= 0 || N_obs <= 0)
        return set_error(c, GRA_BAD_ARG, "empty mesh or point set");
    if ((field_flag & 1u) || (field_flag & ~0x3FFu) || !(field_flag & 0x3FEu))
        return set_error(c, GRA_BAD_ARG, "field_flag must select components among bits 1..9");
    if (store != GRA_F64 && store != GRA_F32 && store != GRA_BF16)
        return set_error(c, GRA_BAD_ARG, "unknown storage type");
    if (Nm > INT32_MAX) return set_error(c, GRA_BAD_ARG, "Nm exceeds int32 column indices");
    if (!prisms_valid(prisms, Nm))
        return set_error(c, GRA_BAD_ARG, "prism bounds must be finite with x0<x1, y0<y1, z0<z1");
    if (!points_valid(pts, N_obs)) return set_error(c, GRA_BAD_ARG, "non-finite point");
    double top = prisms[4];
    for (int64_t j = 1; j < Nm; ++j) top = std::min(top, prisms[6 * j + 4]);
    for (int64_t i = 0; i < N_obs; ++i)
        if (!(pts[3 * i + 2] < top))
            return set_error(c, GRA_BAD_ARG,
                             "points must lie strictly above every prism (z < min z0, z down)");
    const int nc = ncomp(field_flag);
    const int64_t Nd = (int64_t)nc * N_obs;
    const int64_t ld = (Nm + 63) / 64 * 64;
    c.Nd = c.Nm = c.N_obs = c.ld = 0;          // G is invalid until the kernel has run
    c.gn_ready = false;
    c.wav = false;
    c.partn_holds_gn_model = false;
    HIPCHK(c.G.reserve(dtype_size(store) * (size_t)Nd * (size_t)ld), "allocate G");
    if (ld > Nm)
        HIPCHK(hipMemsetAsync(c.G.p, 0, dtype_size(store) * (size_t)Nd * (size_t)ld, c.stream),
               "clear G padding");
    DevBuf dpr;
    HIPCHK(put(dpr, prisms, (size_t)(6 * Nm), c.stream), "upload prisms");
    HIPCHK(put(c.pts_dev, pts, (size_t)(3 * N_obs), c.stream), "upload points");
    {
        TimedRegion tr(c, kTimeAssemble);
        HIPCHK(launch_cell_corner(dpr.as<double>(), Nm, c.pts_dev.as<double>(), N_obs, field_flag,
                                  store, c.G.p, ld, c.stream),
               "per-cell corner assembly");
    }
    HIPCHK(hipStreamSynchronize(c.stream), "assemble");
    c.Nd = Nd; c.Nm = Nm; c.N_obs = N_obs; c.ld = ld; c.flag = field_flag; c.store = store;
    c.g_corners = true;
    c.pts_zmax = max_point_z(pts, N_obs);
    ++c.g_version;
    return GRA_OK;
}

gra_status gra_get_dims(const gra_ctx *ctx, int64_t *Nd, int64_t *Nm, int64_t *N_obs,
                        uint32_t *field_flag, int *store)
{
    if (!ctx) return GRA_BAD_ARG;
    const Ctx &c = ctx->c;
    if (Nd) *Nd = c.Nd;
    if (Nm) *Nm = c.Nm;
    if (N_obs) *N_obs = c.N_obs;
    if (field_flag) *field_flag = c.flag;
    if (store) *store = (int)c.store;
    return GRA_OK;
}

gra_status gra_assemble_wavelet(gra_ctx *ctx, const double *prisms, int64_t Nm, const double *pts,
                                int64_t N_obs, uint32_t field_flag, const int64_t *order,
                                double relative_threshold)
{
    CTX_GUARD(ctx);
    if (!prisms || !pts || !order || Nm <= 0 || N_obs <= 0)
        return set_error(c, GRA_BAD_ARG, "empty mesh, point set or order");
    if ((field_flag & 1u) || (field_flag & ~0x3FFu) || !(field_flag & 0x3FEu))
        return set_error(c, GRA_BAD_ARG, "field_flag must select components among bits 1..9");
    if (Nm > INT32_MAX / 2) return set_error(c, GRA_BAD_ARG, "Nm exceeds int32 column indices");
    if (!(relative_threshold >= 0.0) || !std::isfinite(relative_threshold))
        return set_error(c, GRA_BAD_ARG, "relative threshold must be finite and >= 0");
    if (!prisms_valid(prisms, Nm))
        return set_error(c, GRA_BAD_ARG, "prism bounds must be finite with x0<x1, y0<y1, z0<z1");
    if (!points_valid(pts, N_obs)) return set_error(c, GRA_BAD_ARG, "non-finite point");
    const int64_t Nd = (int64_t)ncomp(field_flag) * N_obs;
    if (Nd > INT32_MAX) return set_error(c, GRA_BAD_ARG, "rows exceed int32 row indices");
    std::vector<char> seen((size_t)Nm, 0);
    std::vector<int32_t> ord((size_t)Nm);
    std::vector<double> pro((size_t)(6 * Nm));
    for (int64_t j = 0; j < Nm; ++j) {
        const int64_t o = order[j];
        if (o < 0 || o >= Nm || seen[(size_t)o])
            return set_error(c, GRA_BAD_ARG, "order must be a permutation of 0..Nm-1");
        seen[(size_t)o] = 1;
        ord[(size_t)j] = (int32_t)o;
        for (int q = 0; q < 6; ++q) pro[(size_t)(6 * j + q)] = prisms[6 * o + q];
    }
    c.Nd = c.Nm = c.N_obs = c.ld = 0;
    c.gn_ready = false;
    c.partn_holds_gn_model = false;
    c.wav = false;
    c.G.release();
    DevBuf dpr;
    HIPCHK(put(dpr, pro.data(), pro.size(), c.stream), "upload prisms");
    HIPCHK(put(c.pts_dev, pts, (size_t)(3 * N_obs), c.stream), "upload points");
    HIPCHK(put(c.wav_order, ord.data(), ord.size(), c.stream), "upload order");
    int *bad = c.status.as<int>();
    HIPCHK(hipMemsetAsync(bad, 0, sizeof(int), c.stream), "clear status");
    gra_status st = wavelet_assemble(c, dpr.as<double>(), Nm, c.pts_dev.as<double>(), N_obs,
                                     field_flag, relative_threshold, bad);
    if (st != GRA_OK) return st;
    int hbad = 0;
    HIPCHK(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, c.stream), "status");
    HIPCHK(hipStreamSynchronize(c.stream), "wavelet assembly");
    c.Nd = Nd; c.Nm = Nm; c.N_obs = N_obs; c.ld = (Nm + 63) / 64 * 64; c.flag = field_flag;
    c.store = GRA_F64;
    c.wav = true;
    c.g_corners = false;
    ++c.g_version;
    if (hbad)
        return set_error(c, GRA_DEGENERATE_GEOMETRY,
                         "an observation point lies on a prism edge (s0*s1 <= 0, Integral.cpp:77)");
    return GRA_OK;
}

gra_status gra_wavelet_info(gra_ctx *ctx, int64_t *n_dy, int64_t *nnz)
{
    CTX_GUARD(ctx);
    if (!c.wav || c.Nd == 0) return set_error(c, GRA_BAD_ARG, "no wavelet-compressed G");
    if (n_dy) *n_dy = c.wav_ndy;
    if (nnz) *nnz = c.wav_nnz;
    return GRA_OK;
}

gra_status gra_copy_G(gra_ctx *ctx, double *out, int col_major)
{
    CTX_GUARD(ctx);
    if (!out) return set_error(c, GRA_BAD_ARG, "null output");
    if (c.wav) return set_error(c, GRA_BAD_ARG, "not available for the wavelet-compressed G");
    if (!c.G.p || c.Nd == 0) return set_error(c, GRA_BAD_ARG, "no G assembled");
    DevBuf tmp;
    HIPCHK(tmp.reserve(sizeof(double) * (size_t)c.Nd * (size_t)c.Nm), "allocate copy");
    HIPCHK(launch_expand_G(c.G.p, c.store, c.Nd, c.Nm, c.ld, tmp.as<double>(), col_major, c.stream),
           "expand G");
    HIPCHK(hipMemcpyAsync(out, tmp.p, sizeof(double) * (size_t)c.Nd * (size_t)c.Nm,
                          hipMemcpyDeviceToHost, c.stream),
           "copy G");
    HIPCHK(hipStreamSynchronize(c.stream), "copy G");
    return GRA_OK;
}

static gra_status gemv_host(gra_ctx *ctx, int trans, const double *vec, double *product)
{
    CTX_GUARD(ctx);
    if (!vec || !product) return set_error(c, GRA_BAD_ARG, "null vector");
    if (!have_G(c)) return set_error(c, GRA_BAD_ARG, "no G assembled");
    const int64_t nin = trans ? c.Nd : c.Nm, nout = trans ? c.Nm : c.Nd;
    DevBuf &x = c.vec_nd[1], &y = c.vec_nd[2];
    if (trans) HIPCHK(put(x, vec, (size_t)nin, c.stream), "upload vec");
    else HIPCHK(put_padded(x, vec, c.Nm, c.ld, c.stream), "upload vec");
    HIPCHK(y.reserve(sizeof(double) * (size_t)nout), "alloc out");
    const GemvShape shape = gemv_shape(c);
    c.partn_holds_gn_model = false;
    if (!trans) {
        HIPCHK(c.part_n.reserve(sizeof(double) * (size_t)shape.n_splits * c.Nd), "alloc part");
        {
            TimedRegion tr(c, kTimeGemvN);
            HIPCHK(launch_G_n(c, shape, x.as<double>(), c.part_n.as<double>(), nullptr, c.stream),
                   "G x");
        }
        HIPCHK(launch_reduce_parts(c.part_n.as<double>(), shape.n_splits, c.Nd, c.Nd, nullptr,
                                   y.as<double>(), c.stream),
               "reduce");
    } else {
        HIPCHK(c.part_t.reserve(sizeof(double) * (size_t)shape.t_chunks * c.ld), "alloc part");
        {
            TimedRegion tr(c, kTimeGemvT);
            HIPCHK(launch_G_t(c, shape, x.as<double>(), c.part_t.as<double>(), nullptr, c.stream),
                   "G^T w");
        }
        HIPCHK(launch_reduce_parts(c.part_t.as<double>(), shape.t_chunks, c.ld, c.Nm, nullptr,
                                   y.as<double>(), c.stream),
               "reduce");
    }
    HIPCHK(hipMemcpyAsync(product, y.p, sizeof(double) * (size_t)nout, hipMemcpyDeviceToHost,
                          c.stream),
           "download");
    HIPCHK(hipStreamSynchronize(c.stream), "gemv");
    return GRA_OK;
}

static gra_status matmul_host(gra_ctx *ctx, int trans, const double *in, int64_t k, double *out)
{
    CTX_GUARD(ctx);
    if (!in || !out) return set_error(c, GRA_BAD_ARG, "null matrix");
    if (k < 1 || k > 1024) return set_error(c, GRA_BAD_ARG, "right-hand sides: 1..1024");
    if (c.wav) return set_error(c, GRA_BAD_ARG, "not available for the wavelet-compressed G");
    if (!c.G.p || c.Nd == 0) return set_error(c, GRA_BAD_ARG, "no G assembled");
    const int64_t nin = trans ? c.Nd : c.Nm, nout = trans ? c.Nm : c.Nd;
    const MatMulPlan p = plan_mat_mul(c.store, trans != 0, c.Nd, c.Nm, c.ld, k);
    HIPCHK(put(c.mm_in, in, (size_t)(nin * k), c.stream), "upload right-hand sides");
    HIPCHK(c.mm_pack.reserve(mat_mul_pack_bytes(c.store, p)), "alloc packed");
    HIPCHK(c.mm_part.reserve(sizeof(double) * (size_t)p.splits * p.ncol * (size_t)p.ldp),
           "alloc partials");
    HIPCHK(c.mm_out.reserve(sizeof(double) * (size_t)(nout * k)), "alloc out");
    HIPCHK(launch_pack_rhs(c.store, trans != 0, c.mm_in.as<double>(), nin, k, p, c.mm_pack.p,
                           c.stream),
           "pack");
    {
        TimedRegion tr(c, kTimeMatMul);
        HIPCHK(launch_mat_mul(c.G.p, c.store, trans != 0, c.Nd, c.Nm, c.ld, c.mm_pack.p, p,
                              c.mm_part.as<double>(), c.stream),
               "matrix-core product");
    }
    HIPCHK(launch_sum_splits(c.mm_part.as<double>(), p, nout, k, c.mm_out.as<double>(), c.stream),
           "sum splits");
    HIPCHK(hipMemcpyAsync(out, c.mm_out.p, sizeof(double) * (size_t)(nout * k),
                          hipMemcpyDeviceToHost, c.stream),
           "download");
    HIPCHK(hipStreamSynchronize(c.stream), "mat_mul");
    return GRA_OK;
}

namespace {
// gra_assemble_batch (face/edge values) and gra_assemble_batch_corners
// (per-prism corner sums) over one item table.
gra_status assemble_batch(gra_ctx *const *ctxs, int64_t n, const gra_batch_item *items,
                          gra_dtype store, bool corners)
{
    if (!ctxs || !items || n <= 0 || n > INT32_MAX) return GRA_BAD_ARG;
    std::vector<gra_ctx *> order(ctxs, ctxs + n);
    for (gra_ctx *x : order)
        if (!x) return GRA_BAD_ARG;
    std::sort(order.begin(), order.end());
    if (std::adjacent_find(order.begin(), order.end()) != order.end()) return GRA_BAD_ARG;
    std::vector<std::unique_lock<std::mutex>> locks;
    locks.reserve(order.size());
    for (gra_ctx *x : order) locks.emplace_back(x->c.mu);   // address order: no lock cycles
    Ctx &c = ctxs[0]->c;
    c.last_error.clear();
    for (int64_t b = 0; b < n; ++b)
        if (ctxs[b]->c.device != c.device)
            return set_error(c, GRA_BAD_ARG, "batched contexts must share one device");
    if (hipSetDevice(c.device) != hipSuccess) return set_error(c, GRA_HIP_ERROR, "hipSetDevice");
    if (store != GRA_F64 && store != GRA_F32 && store != GRA_BF16)
        return set_error(c, GRA_BAD_ARG, "unknown storage type");
    int64_t tot_prisms = 0;
    for (int64_t b = 0; b < n; ++b) {
        const gra_batch_item &it = items[b];
        const uint32_t f = it.field_flag;
        if (!it.prisms || !it.pts || it.Nm <= 0 || it.N_obs <= 0 || it.Nm > INT32_MAX ||
            (f & 1u) || (f & ~0x3FFu) || !(f & 0x3FEu) || !prisms_valid(it.prisms, it.Nm) ||
            !points_valid(it.pts, it.N_obs))
            return set_error(c, GRA_BAD_ARG,
                             "batch item " + std::to_string(b) + ": invalid mesh, points or flag");
        if (corners) {
            double top = it.prisms[4];
            for (int64_t j = 1; j < it.Nm; ++j) top = std::min(top, it.prisms[6 * j + 4]);
            for (int64_t i = 0; i < it.N_obs; ++i)
                if (!(it.pts[3 * i + 2] < top))
                    return set_error(c, GRA_BAD_ARG,
                                     "batch item " + std::to_string(b) +
                                         ": points must lie strictly above every prism");
        }
        tot_prisms += it.Nm;
    }
    std::vector<double> prisms((size_t)(6 * tot_prisms));
    std::vector<BatchItem> tab((size_t)n);
    int64_t off = 0, blocks = 0;
    for (int64_t b = 0; b < n; ++b) {
        const gra_batch_item &it = items[b];
        Ctx &cb = ctxs[b]->c;
        const int64_t Nd = (int64_t)ncomp(it.field_flag) * it.N_obs, ld = (it.Nm + 63) / 64 * 64;
        cb.Nd = cb.Nm = cb.N_obs = cb.ld = 0;
        cb.gn_ready = false;
        cb.wav = false;
        cb.have_crg_ops = false;
        cb.partn_holds_gn_model = false;
        HIPCHK(cb.G.reserve(dtype_size(store) * (size_t)Nd * (size_t)ld), "allocate G");
        if (ld > it.Nm)
            HIPCHK(hipMemsetAsync(cb.G.p, 0, dtype_size(store) * (size_t)Nd * (size_t)ld, c.stream),
                   "clear G padding");
        HIPCHK(put(cb.pts_dev, it.pts, (size_t)(3 * it.N_obs), c.stream), "upload points");
        std::copy(it.prisms, it.prisms + 6 * it.Nm, prisms.begin() + 6 * off);
        BatchItem &t = tab[(size_t)b];
        t.prisms = nullptr;     // set after the arena upload
        t.pts = cb.pts_dev.as<double>();
        t.G = cb.G.p;
        t.Nm = it.Nm;
        t.N_obs = it.N_obs;
        t.ld = ld;
        t.bpr = (it.Nm + 255) / 256;
        t.block0 = blocks;
        t.flag = it.field_flag;
        blocks += t.bpr * it.N_obs;
        off += it.Nm;
    }
    if (blocks > (int64_t)INT32_MAX) return set_error(c, GRA_BAD_ARG, "batch too large for one launch");
    HIPCHK(put(c.batch_prisms, prisms.data(), prisms.size(), c.stream), "upload prisms");
    off = 0;
    for (int64_t b = 0; b < n; ++b) {
        tab[(size_t)b].prisms = c.batch_prisms.as<double>() + 6 * off;
        off += items[b].Nm;
    }
    HIPCHK(c.batch_items.reserve(sizeof(BatchItem) * (size_t)n), "alloc item table");
    HIPCHK(hipMemcpyAsync(c.batch_items.p, tab.data(), sizeof(BatchItem) * (size_t)n,
                          hipMemcpyHostToDevice, c.stream),
           "upload item table");
    HIPCHK(c.batch_bad.reserve(sizeof(int) * (size_t)n), "alloc status");
    HIPCHK(hipMemsetAsync(c.batch_bad.p, 0, sizeof(int) * (size_t)n, c.stream), "clear status");
    {
        TimedRegion tr(c, kTimeAssemble);
        if (corners)
            HIPCHK(launch_cell_corner_batch(c.batch_items.as<BatchItem>(), (int)n, blocks, store,
                                            c.stream),
                   "batched corner-sum assembly");
        else
            HIPCHK(launch_assemble_batch(c.batch_items.as<BatchItem>(), (int)n, blocks, store,
                                         c.batch_bad.as<int>(), c.stream),
                   "batched assembly");
    }
    std::vector<int> bad((size_t)n, 0);
    HIPCHK(hipMemcpyAsync(bad.data(), c.batch_bad.p, sizeof(int) * (size_t)n,
                          hipMemcpyDeviceToHost, c.stream),
           "status");
    HIPCHK(hipStreamSynchronize(c.stream), "batched assembly");
    for (int64_t b = 0; b < n; ++b) {
        Ctx &cb = ctxs[b]->c;
        cb.Nd = (int64_t)ncomp(items[b].field_flag) * items[b].N_obs;
        cb.Nm = items[b].Nm;
        cb.N_obs = items[b].N_obs;
        cb.ld = tab[(size_t)b].ld;
        cb.flag = items[b].field_flag;
        cb.store = store;
        cb.g_corners = corners;
        cb.pts_zmax = max_point_z(items[b].pts, items[b].N_obs);
        ++cb.g_version;
    }
    for (int64_t b = 0; b < n; ++b)
        if (bad[(size_t)b])
            return set_error(c, GRA_DEGENERATE_GEOMETRY,
                             "batch item " + std::to_string(b) +
                                 ": an observation point lies on a prism edge (Integral.cpp:77)");
    return GRA_OK;
}
}  // namespace

gra_status gra_assemble_batch(gra_ctx *const *ctxs, int64_t n, const gra_batch_item *items,
                              gra_dtype store)
{
    return assemble_batch(ctxs, n, items, store, false);
}

gra_status gra_assemble_batch_corners(gra_ctx *const *ctxs, int64_t n,
                                      const gra_batch_item *items, gra_dtype store)
{
    return assemble_batch(ctxs, n, items, store, true);
}

gra_status gra_expand_cols(gra_ctx *ctx, const int64_t *split_cols, int64_t n_split,
                           const double *child_prisms)
{
    CTX_GUARD(ctx);
    if (c.wav) return set_error(c, GRA_BAD_ARG, "not available for the wavelet-compressed G");
    if (!c.G.p || c.Nd == 0) return set_error(c, GRA_BAD_ARG, "no G assembled");
    if (!split_cols || !child_prisms || n_split <= 0)
        return set_error(c, GRA_BAD_ARG, "empty split list");
    for (int64_t s = 0; s < n_split; ++s)
        if (split_cols[s] < 0 || split_cols[s] >= c.Nm || (s && split_cols[s] <= split_cols[s - 1]))
            return set_error(c, GRA_BAD_ARG,
                             "split columns must be strictly increasing column indices < Nm");
    const int64_t nchild = 8 * n_split;
    if (!prisms_valid(child_prisms, nchild))
        return set_error(c, GRA_BAD_ARG, "child prism bounds must be finite with x0<x1, y0<y1, z0<z1");
    // the children of a corner-sum G are corner sums too (the two forms
    // differ where the face/edge one loses digits, so mixing them would
    // break the children-sum-to-parent identity there)
    if (c.g_corners)
        for (int64_t j = 0; j < nchild; ++j)
            if (!(c.pts_zmax < child_prisms[6 * j + 4]))
                return set_error(c, GRA_BAD_ARG,
                                 "points must lie strictly above every child prism (corner-sum G)");
    const int64_t Nm2 = c.Nm + 7 * n_split;
    if (Nm2 > INT32_MAX) return set_error(c, GRA_BAD_ARG, "Nm exceeds int32 column indices");
    const int64_t ld2 = (Nm2 + 63) / 64 * 64, ldc = (nchild + 63) / 64 * 64;
    // new column j <- old column src[j] (>= 0) or child -src[j]-1, children of a
    // parent replacing it in place (AdaptiveInversion.cpp:58-104)
    std::vector<int32_t> src((size_t)Nm2);
    int64_t out = 0, s = 0;
    for (int64_t col = 0; col < c.Nm; ++col) {
        if (s < n_split && split_cols[s] == col) {
            for (int j = 0; j < 8; ++j) src[(size_t)out++] = -(int32_t)(8 * s + j) - 1;
            ++s;
        } else {
            src[(size_t)out++] = (int32_t)col;
        }
    }
    const size_t esz = dtype_size(c.store);
    DevBuf dch, dsrc, Gc, G2;
    HIPCHK(put(dch, child_prisms, (size_t)(6 * nchild), c.stream), "upload child prisms");
    HIPCHK(put(dsrc, src.data(), src.size(), c.stream), "upload column map");
    HIPCHK(Gc.reserve(esz * (size_t)c.Nd * (size_t)ldc), "allocate child columns");
    HIPCHK(G2.reserve(esz * (size_t)c.Nd * (size_t)ld2), "allocate expanded G");
    if (ld2 > Nm2) HIPCHK(hipMemsetAsync(G2.p, 0, G2.bytes, c.stream), "clear G padding");
    int *bad = c.status.as<int>();
    HIPCHK(hipMemsetAsync(bad, 0, sizeof(int), c.stream), "clear status");
    {
        TimedRegion tr(c, kTimeAssemble);
        if (c.g_corners)
            HIPCHK(launch_cell_corner(dch.as<double>(), nchild, c.pts_dev.as<double>(), c.N_obs,
                                      c.flag, c.store, Gc.p, ldc, c.stream),
                   "child columns by corner sums");
        else
            HIPCHK(launch_assemble(dch.as<double>(), nchild, c.pts_dev.as<double>(), c.N_obs,
                                   c.flag, c.store, Gc.p, ldc, bad, c.stream),
                   "assemble child columns");
    }
    HIPCHK(launch_remap_cols(c.G.p, c.ld, Gc.p, ldc, dsrc.as<int32_t>(), c.Nd, Nm2, G2.p, ld2,
                             (int)esz, c.stream),
           "remap columns");
    int hbad = 0;
    HIPCHK(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, c.stream), "status");
    HIPCHK(hipStreamSynchronize(c.stream), "expand");
    std::swap(c.G.p, G2.p);
    std::swap(c.G.bytes, G2.bytes);
    ++c.g_version;
    c.Nm = Nm2;
    c.ld = ld2;
    c.gn_ready = false;
    c.have_crg_ops = false;
    c.partn_holds_gn_model = false;
    if (hbad)
        return set_error(c, GRA_DEGENERATE_GEOMETRY,
                         "an observation point lies on a child prism edge (Integral.cpp:77)");
    return GRA_OK;
}

gra_status gra_col_norms(gra_ctx *ctx, double *norms)
{
    CTX_GUARD(ctx);
    if (!norms) return set_error(c, GRA_BAD_ARG, "null output");
    if (c.wav) return set_error(c, GRA_BAD_ARG, "not available for the wavelet-compressed G");
    if (!c.G.p || c.Nd == 0) return set_error(c, GRA_BAD_ARG, "no G assembled");
    DevBuf &y = c.vec_nd[2];
    HIPCHK(y.reserve(sizeof(double) * (size_t)c.Nm), "alloc out");
    const GemvShape shape = gemv_shape(c);
    HIPCHK(c.part_t.reserve(sizeof(double) * (size_t)shape.t_chunks * c.ld), "alloc part");
    HIPCHK(launch_col_sumsq(c.G.p, c.store, c.Nd, c.ld, shape, c.part_t.as<double>(), c.stream),
           "column sums of squares");
    HIPCHK(launch_reduce_parts(c.part_t.as<double>(), shape.t_chunks, c.ld, c.Nm, nullptr,
                               y.as<double>(), c.stream),
           "reduce");
    HIPCHK(hipMemcpyAsync(norms, y.p, sizeof(double) * (size_t)c.Nm, hipMemcpyDeviceToHost,
                          c.stream),
           "download");
    HIPCHK(hipStreamSynchronize(c.stream), "col_norms");
    for (int64_t j = 0; j < c.Nm; ++j) norms[j] = std::sqrt(norms[j]);
    return GRA_OK;
}

gra_status gra_G_vec_mul(gra_ctx *ctx, const double *vec, double *product)
{
    return gemv_host(ctx, 0, vec, product);
}

gra_status gra_G_vec_mul_noalloc(gra_ctx *ctx, const double *d_x, double *d_y)
{
    CTX_GUARD(ctx);
    if (!d_x || !d_y) return set_error(c, GRA_BAD_ARG, "null vector");
    if (c.wav || !c.G.p || c.Nd == 0) return set_error(c, GRA_BAD_ARG, "no dense G assembled");
    const GemvShape shape = gemv_shape(c);
    // staged x (ld) + the split partial sums (n_splits x Nd)
    const int64_t need = c.ld + (int64_t)shape.n_splits * c.Nd;
    if (need > kScratchDoubles)
        return set_error(c, GRA_BAD_ARG, "G too large for the allocation-free product (needs "
                                         "ld + splits*Nd <= 4194304 doubles of scratch)");
    double *xs = c.noalloc_scratch, *part = c.noalloc_scratch + c.ld;
    hipLaunchKernelGGL(k_stage_padded, dim3((unsigned)((c.ld + 255) / 256)), dim3(256), 0,
                       c.stream, d_x, c.Nm, c.ld, xs);
    HIPCHK(hipGetLastError(), "stage x");
    HIPCHK(launch_gemv_n(c.G.p, c.store, c.Nd, c.ld, shape, xs, part, nullptr, c.stream), "G x");
    HIPCHK(launch_reduce_parts(part, shape.n_splits, c.Nd, c.Nd, nullptr, d_y, c.stream),
           "reduce");
    return GRA_OK;
}

gra_status gra_GT_vec_mul(gra_ctx *ctx, const double *vec, double *product)
{
    return gemv_host(ctx, 1, vec, product);
}

gra_status gra_G_mat_mul(gra_ctx *ctx, const double *X, int64_t k, double *Y)
{
    return matmul_host(ctx, 0, X, k, Y);
}

gra_status gra_GT_mat_mul(gra_ctx *ctx, const double *R, int64_t k, double *Z)
{
    return matmul_host(ctx, 1, R, k, Z);
}

gra_status gra_forward(gra_ctx *ctx, const double *prisms, const double *rho, int64_t Nm,
                       const double *pts, int64_t N_obs, uint32_t field_flag, double *out)
{
    CTX_GUARD(ctx);
    if (!prisms || !rho || !pts || !out || Nm <= 0 || N_obs <= 0)
        return set_error(c, GRA_BAD_ARG, "empty input");
    if ((field_flag & 1u) || (field_flag & ~0x3FFu) || !(field_flag & 0x3FEu))
        return set_error(c, GRA_BAD_ARG, "field_flag must select components among bits 1..9");
    if (!prisms_valid(prisms, Nm))
        return set_error(c, GRA_BAD_ARG, "prism bounds must be finite with x0<x1, y0<y1, z0<z1");
    if (!points_valid(pts, N_obs)) return set_error(c, GRA_BAD_ARG, "non-finite point");
    const int nc = ncomp(field_flag);
    DevBuf &dpr = c.fwd_prisms, &drho = c.fwd_rho, &dpt = c.fwd_pts, &dout = c.fwd_out;
    HIPCHK(put(dpr, prisms, (size_t)(6 * Nm), c.stream), "upload prisms");
    HIPCHK(put(drho, rho, (size_t)Nm, c.stream), "upload rho");
    HIPCHK(put(dpt, pts, (size_t)(3 * N_obs), c.stream), "upload points");
    HIPCHK(dout.reserve(sizeof(double) * (size_t)N_obs * nc), "alloc out");
    int *bad = c.status.as<int>();
    HIPCHK(hipMemsetAsync(bad, 0, sizeof(int), c.stream), "clear status");
    HIPCHK(launch_forward(dpr.as<double>(), drho.as<double>(), Nm, dpt.as<double>(), N_obs,
                          field_flag, dout.as<double>(), bad, c.stream),
           "forward kernel");
    int hbad = 0;
    HIPCHK(hipMemcpyAsync(out, dout.p, sizeof(double) * (size_t)N_obs * nc, hipMemcpyDeviceToHost,
                          c.stream),
           "download");
    HIPCHK(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, c.stream), "status");
    HIPCHK(hipStreamSynchronize(c.stream), "forward");
    if (hbad)
        return set_error(c, GRA_DEGENERATE_GEOMETRY,
                         "an observation point lies on a prism edge (s0*s1 <= 0, Integral.cpp:77)");
    return GRA_OK;
}

gra_status gra_forward_rectilinear(gra_ctx *ctx, const double *xs, int64_t nx, const double *ys,
                                   int64_t ny, const double *zs, int64_t nz, const double *rho,
                                   const double *pts, int64_t N_obs, uint32_t field_flag,
                                   double *out)
{
    CTX_GUARD(ctx);
    if (!xs || !ys || !zs || !rho || !pts || !out || nx <= 0 || ny <= 0 || nz <= 0 || N_obs <= 0)
        return set_error(c, GRA_BAD_ARG, "empty input");
    if ((field_flag & 1u) || (field_flag & ~0x3FFu) || !(field_flag & 0x3FEu))
        return set_error(c, GRA_BAD_ARG, "field_flag must select components among bits 1..9");
    if (!increasing(xs, nx + 1) || !increasing(ys, ny + 1) || !increasing(zs, nz + 1))
        return set_error(c, GRA_BAD_ARG, "node coordinates must be finite and strictly increasing");
    if ((nx + 1) * (ny + 1) > (int64_t)INT32_MAX / (nz + 1))
        return set_error(c, GRA_BAD_ARG, "mesh too large (more than 2^31 nodes)");
    if (!points_valid(pts, N_obs)) return set_error(c, GRA_BAD_ARG, "non-finite point");
    for (int64_t i = 0; i < N_obs; ++i)
        if (!(pts[3 * i + 2] < zs[0]))
            return set_error(c, GRA_BAD_ARG,
                             "points must lie strictly above the mesh (z < zs[0], z down)");
    const int nc = ncomp(field_flag);
    const int64_t Nm = nx * ny * nz, NN = (nx + 1) * (ny + 1) * (nz + 1);
    std::vector<double> axes((size_t)(nx + ny + nz + 3));
    std::copy(xs, xs + nx + 1, axes.begin());
    std::copy(ys, ys + ny + 1, axes.begin() + nx + 1);
    std::copy(zs, zs + nz + 1, axes.begin() + nx + ny + 2);
    HIPCHK(put(c.fwd_axes, axes.data(), axes.size(), c.stream), "upload axes");
    HIPCHK(put(c.fwd_rho, rho, (size_t)Nm, c.stream), "upload rho");
    HIPCHK(put(c.fwd_pts, pts, (size_t)(3 * N_obs), c.stream), "upload points");
    HIPCHK(c.fwd_w.reserve(sizeof(double) * (size_t)NN), "alloc node weights");
    HIPCHK(c.fwd_out.reserve(sizeof(double) * (size_t)N_obs * nc), "alloc out");
    const double *da = c.fwd_axes.as<double>();
    HIPCHK(launch_node_forward(da, da + nx + 1, da + nx + ny + 2, nx, ny, nz,
                               c.fwd_rho.as<double>(), c.fwd_w.as<double>(),
                               c.fwd_pts.as<double>(), N_obs, field_flag, c.fwd_out.as<double>(),
                               c.stream),
           "node forward");
    HIPCHK(hipMemcpyAsync(out, c.fwd_out.p, sizeof(double) * (size_t)N_obs * nc,
                          hipMemcpyDeviceToHost, c.stream),
           "download");
    HIPCHK(hipStreamSynchronize(c.stream), "forward_rectilinear");
    return GRA_OK;
}

gra_status gra_forward_rectilinear_batch(gra_ctx *ctx, int64_t n, const gra_fwd_rect_item *items)
{
    CTX_GUARD(ctx);
    if (!items || n <= 0 || n > 4096) return set_error(c, GRA_BAD_ARG, "items: 1..4096");
    int64_t n_axes = 0, n_rho = 0, n_pts = 0, n_w = 0, n_out = 0, blocks = 0;
    for (int64_t b = 0; b < n; ++b) {
        const gra_fwd_rect_item &it = items[b];
        const std::string tag = "item " + std::to_string(b) + ": ";
        if (!it.xs || !it.ys || !it.zs || !it.rho || !it.pts || !it.out || it.nx <= 0 ||
            it.ny <= 0 || it.nz <= 0 || it.N_obs <= 0)
            return set_error(c, GRA_BAD_ARG, tag + "empty input");
        const uint32_t f = it.field_flag;
        if ((f & 1u) || (f & ~0x3FFu) || !(f & 0x3FEu))
            return set_error(c, GRA_BAD_ARG, tag + "field_flag must select components among bits 1..9");
        if (!increasing(it.xs, it.nx + 1) || !increasing(it.ys, it.ny + 1) ||
            !increasing(it.zs, it.nz + 1))
            return set_error(c, GRA_BAD_ARG, tag + "node coordinates must be finite and strictly increasing");
        if ((it.nx + 1) * (it.ny + 1) > (int64_t)INT32_MAX / (it.nz + 1))
            return set_error(c, GRA_BAD_ARG, tag + "mesh too large (more than 2^31 nodes)");
        if (!points_valid(it.pts, it.N_obs)) return set_error(c, GRA_BAD_ARG, tag + "non-finite point");
        for (int64_t i = 0; i < it.N_obs; ++i)
            if (!(it.pts[3 * i + 2] < it.zs[0]))
                return set_error(c, GRA_BAD_ARG,
                                 tag + "points must lie strictly above the mesh (z < zs[0], z down)");
        n_axes += it.nx + it.ny + it.nz + 3;
        n_rho += it.nx * it.ny * it.nz;
        n_pts += 3 * it.N_obs;
        n_w += (it.nx + 1) * (it.ny + 1) * (it.nz + 1);
        n_out += it.N_obs * ncomp(f);
        blocks += it.N_obs;
    }
    if (blocks > (int64_t)INT32_MAX) return set_error(c, GRA_BAD_ARG, "too many points for one launch");
    // one host arena per array kind, uploaded once
    std::vector<double> axes((size_t)n_axes), rho((size_t)n_rho), pts((size_t)n_pts);
    std::vector<FwdItem> tab((size_t)n);
    int64_t oa = 0, orh = 0, op = 0, ow = 0, oo = 0, ob = 0;
    for (int64_t b = 0; b < n; ++b) {
        const gra_fwd_rect_item &it = items[b];
        double *a = axes.data() + oa;
        std::copy(it.xs, it.xs + it.nx + 1, a);
        std::copy(it.ys, it.ys + it.ny + 1, a + it.nx + 1);
        std::copy(it.zs, it.zs + it.nz + 1, a + it.nx + it.ny + 2);
        std::copy(it.rho, it.rho + it.nx * it.ny * it.nz, rho.data() + orh);
        std::copy(it.pts, it.pts + 3 * it.N_obs, pts.data() + op);
        FwdItem &t = tab[(size_t)b];
        // offsets now, device addresses once the arenas exist
        t.xs = reinterpret_cast<const double *>(oa);
        t.rho = reinterpret_cast<const double *>(orh);
        t.pts = reinterpret_cast<const double *>(op);
        t.w = reinterpret_cast<double *>(ow);
        t.out = reinterpret_cast<double *>(oo);
        t.nx = it.nx; t.ny = it.ny; t.nz = it.nz; t.N_obs = it.N_obs; t.block0 = ob;
        t.flag = it.field_flag;
        oa += it.nx + it.ny + it.nz + 3;
        orh += it.nx * it.ny * it.nz;
        op += 3 * it.N_obs;
        ow += (it.nx + 1) * (it.ny + 1) * (it.nz + 1);
        oo += it.N_obs * ncomp(it.field_flag);
        ob += it.N_obs;
    }
    HIPCHK(put(c.fwd_axes, axes.data(), axes.size(), c.stream), "upload axes");
    HIPCHK(put(c.fwd_rho, rho.data(), rho.size(), c.stream), "upload rho");
    HIPCHK(put(c.fwd_pts, pts.data(), pts.size(), c.stream), "upload points");
    HIPCHK(c.fwd_w.reserve(sizeof(double) * (size_t)n_w), "alloc node weights");
    HIPCHK(c.fwd_out.reserve(sizeof(double) * (size_t)n_out), "alloc out");
    for (FwdItem &t : tab) {
        const int64_t a = reinterpret_cast<int64_t>(t.xs);
        t.xs = c.fwd_axes.as<double>() + a;
        t.ys = t.xs + t.nx + 1;
        t.zs = t.ys + t.ny + 1;
        t.rho = c.fwd_rho.as<double>() + reinterpret_cast<int64_t>(t.rho);
        t.pts = c.fwd_pts.as<double>() + reinterpret_cast<int64_t>(t.pts);
        t.w = c.fwd_w.as<double>() + reinterpret_cast<int64_t>(t.w);
        t.out = c.fwd_out.as<double>() + reinterpret_cast<int64_t>(t.out);
    }
    HIPCHK(c.fwd_items.reserve(sizeof(FwdItem) * (size_t)n), "alloc item table");
    HIPCHK(hipMemcpyAsync(c.fwd_items.p, tab.data(), sizeof(FwdItem) * (size_t)n,
                          hipMemcpyHostToDevice, c.stream),
           "upload item table");
    HIPCHK(launch_node_forward_batch(tab.data(), c.fwd_items.as<FwdItem>(), (int)n, blocks, c.stream),
           "batched node forward");
    std::vector<double> out((size_t)n_out);
    HIPCHK(hipMemcpyAsync(out.data(), c.fwd_out.p, sizeof(double) * (size_t)n_out,
                          hipMemcpyDeviceToHost, c.stream),
           "download");
    HIPCHK(hipStreamSynchronize(c.stream), "forward_rectilinear_batch");
    oo = 0;
    for (int64_t b = 0; b < n; ++b) {
        const int64_t m = items[b].N_obs * ncomp(items[b].field_flag);
        std::copy(out.begin() + oo, out.begin() + oo + m, items[b].out);
        oo += m;
    }
    return GRA_OK;
}

gra_status gra_set_regularization(gra_ctx *ctx, const gra_csr *Ws, const gra_csr *Wx,
                                  const gra_csr *Wy, const gra_csr *Wz, const gra_csr *Tx0,
                                  const gra_csr *Ty0, const gra_csr *Tz0)
{
    CTX_GUARD(ctx);
    if (c.Nm == 0) return set_error(c, GRA_BAD_ARG, "assemble G first (Nm unknown)");
    const int64_t n = c.Nm;
    if (!csr_valid(Ws, n) || !csr_valid(Wx, n) || !csr_valid(Wy, n) || !csr_valid(Wz, n))
        return set_error(c, GRA_BAD_ARG, "W operators must be valid Nm x Nm CSR matrices");
    const bool crg = Tx0 || Ty0 || Tz0;
    if (crg && (!csr_valid(Tx0, n) || !csr_valid(Ty0, n) || !csr_valid(Tz0, n)))
        return set_error(c, GRA_BAD_ARG, "T operators must be valid Nm x Nm CSR matrices");
    gra_status s;
    if ((s = upload_csr(c, c.Ws, Ws)) || (s = upload_csr(c, c.Wx, Wx)) ||
        (s = upload_csr(c, c.Wy, Wy)) || (s = upload_csr(c, c.Wz, Wz)))
        return s;
    c.have_crg_ops = false;
    if (crg) {
        if ((s = upload_csr(c, c.Tx0, Tx0)) || (s = upload_csr(c, c.Ty0, Ty0)) ||
            (s = upload_csr(c, c.Tz0, Tz0)))
            return s;
        c.have_crg_ops = true;
    }
    return GRA_OK;
}

gra_status gra_solve_cg(gra_ctx *ctx, const gra_cg_args *a, double *dx, double *err,
                        int *iterations)
{
    CTX_GUARD(ctx);
    if (!a || !dx || !a->mk || !a->mref || !a->Wd || !a->dobs || !a->P)
        return set_error(c, GRA_BAD_ARG, "missing solve_cg argument");
    if (!have_G(c)) return set_error(c, GRA_BAD_ARG, "no G assembled");
    if (!c.Ws.present || c.Ws.n != c.Nm)
        return set_error(c, GRA_BAD_ARG, "regularisation operators not uploaded for this Nm");
    if (a->maxit < 0 || !(a->lambda >= 0)) return set_error(c, GRA_BAD_ARG, "maxit/lambda");
    const int64_t nm = c.Nm, nd = c.Nd, ld = c.ld;
    // input arena: mk(ld) mref P dx WL*7 S (ld each) | Wd dobs (nd each)
    DevBuf &in = c.host_io;
    const size_t nvec = 12;
    HIPCHK(in.reserve(sizeof(double) * ((size_t)ld * nvec + 2 * (size_t)nd)), "alloc inputs");
    double *base = in.as<double>();
    double *d_mk = base, *d_mref = base + ld, *d_P = base + 2 * ld, *d_dx = base + 3 * ld;
    double *d_WL = base + 4 * ld, *d_S = base + 11 * ld;
    double *d_Wd = base + 12 * ld, *d_dobs = d_Wd + nd;
    HIPCHK(hipMemsetAsync(d_mk, 0, sizeof(double) * (size_t)ld, c.stream), "clear");
    HIPCHK(hipMemcpyAsync(d_mk, a->mk, sizeof(double) * nm, hipMemcpyHostToDevice, c.stream), "mk");
    HIPCHK(hipMemcpyAsync(d_mref, a->mref, sizeof(double) * nm, hipMemcpyHostToDevice, c.stream),
           "mref");
    HIPCHK(hipMemcpyAsync(d_P, a->P, sizeof(double) * nm, hipMemcpyHostToDevice, c.stream), "P");
    HIPCHK(hipMemcpyAsync(d_Wd, a->Wd, sizeof(double) * nd, hipMemcpyHostToDevice, c.stream), "Wd");
    HIPCHK(hipMemcpyAsync(d_dobs, a->dobs, sizeof(double) * nd, hipMemcpyHostToDevice, c.stream),
           "dobs");
    const double *hWL[7] = {a->WL_s, a->WL_z, a->WL_x, a->WL_y, a->WL_tz, a->WL_tx, a->WL_ty};
    const double *dWL[7];
    for (int q = 0; q < 7; ++q) {
        if (hWL[q]) {
            HIPCHK(hipMemcpyAsync(d_WL + q * ld, hWL[q], sizeof(double) * nm,
                                  hipMemcpyHostToDevice, c.stream),
                   "WL");
            dWL[q] = d_WL + q * ld;
        } else {
            dWL[q] = nullptr;
        }
    }
    const double *dS = nullptr;
    if (a->S_crg) {
        HIPCHK(hipMemcpyAsync(d_S, a->S_crg, sizeof(double) * nm, hipMemcpyHostToDevice, c.stream),
               "S");
        dS = d_S;
    }
    CgDiag dg{d_Wd, d_P, dWL[0], dWL[2], dWL[3], dWL[1], dWL[5], dWL[6], dWL[4], dS};
    gra_status s = cgls_solve(c, *a, d_mk, d_mref, d_dobs, dg, d_dx, err, iterations);
    if (s != GRA_OK && s != GRA_NUMERIC_BREAKDOWN) return s;
    HIPCHK(hipMemcpyAsync(dx, d_dx, sizeof(double) * nm, hipMemcpyDeviceToHost, c.stream), "dx");
    HIPCHK(hipStreamSynchronize(c.stream), "solve_cg");
    return s;
}

gra_status gra_mixed_precision(gra_ctx *ctx, gra_dtype inner, int refresh_every)
{
    CTX_GUARD(ctx);
    if (inner == GRA_F64) {
        c.Glo.release();
        c.mixed_refresh = 0;
        c.lo_version = ~0ull;
        c.store_lo = GRA_F64;
        return GRA_OK;
    }
    if (inner != GRA_F32 && inner != GRA_BF16)
        return set_error(c, GRA_BAD_ARG, "inner precision: GRA_F64 (off), GRA_F32 or GRA_BF16");
    if (refresh_every < 1) return set_error(c, GRA_BAD_ARG, "refresh_every must be >= 1");
    if (!c.G.p || c.Nd == 0 || c.wav || c.store != GRA_F64)
        return set_error(c, GRA_BAD_ARG, "mixed precision needs a dense FP64 G assembled");
    const int64_t n = c.Nd * c.ld;
    HIPCHK(c.Glo.reserve(dtype_size(inner) * (size_t)n), "allocate reduced-precision G");
    HIPCHK(launch_convert_G(c.G.as<double>(), n, inner, c.Glo.p, c.stream), "convert G");
    HIPCHK(hipStreamSynchronize(c.stream), "convert G");
    c.store_lo = inner;
    c.lo_version = c.g_version;
    c.mixed_refresh = refresh_every;
    return GRA_OK;
}

gra_status gra_barrier_P(gra_ctx *ctx, const double *m, const double *m_min, const double *m_max,
                         double *P)
{
    CTX_GUARD(ctx);
    if (!m || !m_min || !m_max || !P || c.Nm == 0) return set_error(c, GRA_BAD_ARG, "args");
    const int64_t n = c.Nm;
    DevBuf &b = c.vec_nm[8];
    HIPCHK(b.reserve(sizeof(double) * 4 * (size_t)n), "alloc");
    double *dm = b.as<double>(), *dlo = dm + n, *dhi = dm + 2 * n, *dP = dm + 3 * n;
    HIPCHK(hipMemcpyAsync(dm, m, sizeof(double) * n, hipMemcpyHostToDevice, c.stream), "m");
    HIPCHK(hipMemcpyAsync(dlo, m_min, sizeof(double) * n, hipMemcpyHostToDevice, c.stream), "min");
    HIPCHK(hipMemcpyAsync(dhi, m_max, sizeof(double) * n, hipMemcpyHostToDevice, c.stream), "max");
    HIPCHK(launch_barrier_P(dm, dlo, dhi, n, dP, c.stream), "barrier_P");
    HIPCHK(hipMemcpyAsync(P, dP, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream), "P");
    HIPCHK(hipStreamSynchronize(c.stream), "barrier_P");
    return GRA_OK;
}

gra_status gra_barrier_update(gra_ctx *ctx, const double *m, const double *dx, const double *m_min,
                              const double *m_max, double *m_trial)
{
    CTX_GUARD(ctx);
    if (!m || !dx || !m_min || !m_max || !m_trial || c.Nm == 0)
        return set_error(c, GRA_BAD_ARG, "args");
    const int64_t n = c.Nm;
    DevBuf &b = c.vec_nm[9];
    HIPCHK(b.reserve(sizeof(double) * 6 * (size_t)n + sizeof(int) * (size_t)(n + 1)), "alloc");
    double *dm = b.as<double>(), *ddx = dm + n, *dlo = dm + 2 * n, *dhi = dm + 3 * n,
           *dmt = dm + 4 * n, *scr = dm + 5 * n;
    int *flags = reinterpret_cast<int *>(dm + 6 * n);
    HIPCHK(hipMemcpyAsync(dm, m, sizeof(double) * n, hipMemcpyHostToDevice, c.stream), "m");
    HIPCHK(hipMemcpyAsync(ddx, dx, sizeof(double) * n, hipMemcpyHostToDevice, c.stream), "dx");
    HIPCHK(hipMemcpyAsync(dlo, m_min, sizeof(double) * n, hipMemcpyHostToDevice, c.stream), "min");
    HIPCHK(hipMemcpyAsync(dhi, m_max, sizeof(double) * n, hipMemcpyHostToDevice, c.stream), "max");
    HIPCHK(hipMemcpyAsync(dmt, m_trial, sizeof(double) * n, hipMemcpyHostToDevice, c.stream), "mt");
    HIPCHK(launch_barrier_update(dm, ddx, dlo, dhi, n, dmt, scr, flags, c.stream), "update");
    HIPCHK(hipMemcpyAsync(m_trial, dmt, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream), "mt");
    HIPCHK(hipStreamSynchronize(c.stream), "barrier_update");
    return GRA_OK;
}

gra_status gra_misfit(gra_ctx *ctx, const double *m, const double *Wd, const double *dobs,
                      double *misfit)
{
    CTX_GUARD(ctx);
    if (!m || !Wd || !dobs || !misfit) return set_error(c, GRA_BAD_ARG, "args");
    if (!have_G(c)) return set_error(c, GRA_BAD_ARG, "no G assembled");
    const int64_t nd = c.Nd;
    DevBuf &x = c.vec_nm[10], &v = c.vec_nd[3];
    HIPCHK(put_padded(x, m, c.Nm, c.ld, c.stream), "m");
    HIPCHK(v.reserve(sizeof(double) * 2 * (size_t)nd), "alloc");
    HIPCHK(hipMemcpyAsync(v.p, Wd, sizeof(double) * nd, hipMemcpyHostToDevice, c.stream), "Wd");
    HIPCHK(hipMemcpyAsync(v.as<double>() + nd, dobs, sizeof(double) * nd, hipMemcpyHostToDevice,
                          c.stream),
           "dobs");
    const GemvShape shape = gemv_shape(c);
    c.partn_holds_gn_model = false;
    HIPCHK(c.part_n.reserve(sizeof(double) * (size_t)shape.n_splits * nd), "alloc");
    HIPCHK(c.red.reserve(sizeof(double) * (size_t)((nd + 255) / 256 + 1)), "alloc");
    HIPCHK(c.vec_nm[11].reserve(sizeof(double)), "alloc");
    HIPCHK(launch_G_n(c, shape, x.as<double>(), c.part_n.as<double>(), nullptr, c.stream), "G x");
    HIPCHK(launch_misfit(c.part_n.as<double>(), shape.n_splits, nd, v.as<double>(), v.as<double>() + nd,
                         c.red.as<double>(), c.vec_nm[11].as<double>(), c.stream),
           "misfit");
    HIPCHK(hipMemcpyAsync(misfit, c.vec_nm[11].p, sizeof(double), hipMemcpyDeviceToHost, c.stream),
           "misfit");
    HIPCHK(hipStreamSynchronize(c.stream), "misfit");
    return GRA_OK;
}

gra_status gra_irls_weights(gra_ctx *ctx, const gra_csr *W, const double *m, int p, double eps2,
                            double *WL)
{
    CTX_GUARD(ctx);
    if (!W || !m || !WL || W->n <= 0 || !csr_valid(W, W->n))
        return set_error(c, GRA_BAD_ARG, "args");
    const int64_t n = W->n;
    DevBuf rp, ci, v, dm, dout;
    HIPCHK(put(rp, W->row_ptr, (size_t)n + 1, c.stream), "rp");
    HIPCHK(put(ci, W->col_idx, (size_t)W->nnz, c.stream), "ci");
    HIPCHK(put(v, W->val, (size_t)W->nnz, c.stream), "v");
    HIPCHK(put(dm, m, (size_t)n, c.stream), "m");
    HIPCHK(dout.reserve(sizeof(double) * (size_t)n), "alloc");
    HIPCHK(launch_irls(rp.as<int64_t>(), ci.as<int32_t>(), v.as<double>(), n, dm.as<double>(),
                       nullptr, p, eps2, dout.as<double>(), c.stream),
           "irls");
    HIPCHK(hipMemcpyAsync(WL, dout.p, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, c.stream),
           "WL");
    HIPCHK(hipStreamSynchronize(c.stream), "irls");
    return GRA_OK;
}

gra_status gra_gn_setup_problem(gra_ctx *ctx, const gra_gn_setup *s)
{
    CTX_GUARD(ctx);
    if (!s || !s->dobs || !s->Wd || !s->m_min || !s->m_max || !s->mref || !s->m_ini)
        return set_error(c, GRA_BAD_ARG, "missing setup array");
    if (!have_G(c)) return set_error(c, GRA_BAD_ARG, "no G assembled");
    const int64_t nd = c.Nd, nm = c.Nm, ld = c.ld;
    HIPCHK(put(c.gn_dobs, s->dobs, (size_t)nd, c.stream), "dobs");
    HIPCHK(put(c.gn_Wd, s->Wd, (size_t)nd, c.stream), "Wd");
    HIPCHK(put(c.gn_mmin, s->m_min, (size_t)nm, c.stream), "m_min");
    HIPCHK(put(c.gn_mmax, s->m_max, (size_t)nm, c.stream), "m_max");
    HIPCHK(put(c.gn_mref, s->mref, (size_t)nm, c.stream), "mref");
    // The reference's start model rule (GaussNewtonInversion.cpp:769-780):
    // values at or beyond a bound move to the middle of the interval.
    std::vector<double> m0(s->m_ini, s->m_ini + nm);
    for (int64_t i = 0; i < nm; ++i) {
        double lo = s->m_min[i], hi = s->m_max[i];
        if (m0[i] > hi || std::fabs(m0[i] - hi) < 1e-7) m0[i] = 0.5 * (hi + lo);
        if (m0[i] < lo || std::fabs(m0[i] - lo) < 1e-7) m0[i] = 0.5 * (hi + lo);
    }
    HIPCHK(put_padded(c.gn_m, m0.data(), nm, ld, c.stream), "m");
    HIPCHK(c.gn_mtrial.reserve(sizeof(double) * (size_t)ld), "alloc");
    HIPCHK(c.gn_P.reserve(sizeof(double) * (size_t)nm), "alloc");
    HIPCHK(c.gn_dx.reserve(sizeof(double) * (size_t)nm), "alloc");
    HIPCHK(c.gn_one.reserve(sizeof(double) * 2 * (size_t)nm + sizeof(int) * (size_t)(nm + 1) + 64),
           "alloc");
    c.gn_lp = s->lp;
    c.gn_eps2 = s->eps2;
    c.gn_fused = s->fused;
    c.gn_crg = s->use_cross_gradient != 0;
    c.gn_a_crg = s->a_crg;
    c.gn_have_S = false;
    if (c.gn_crg) {
        if (!c.have_crg_ops || c.Tx0.n != nm)
            return set_error(c, GRA_BAD_ARG, "cross-gradient requested but no T operators for this Nm");
        if (s->S_crg) {
            HIPCHK(put(c.gn_S, s->S_crg, (size_t)nm, c.stream), "S_crg");
            c.gn_have_S = true;
        }
    }
    c.partn_holds_gn_model = false;
    c.gn_have_wl = false;
    if (s->lp != 2) HIPCHK(c.gn_WL.reserve(sizeof(double) * 4 * (size_t)nm), "alloc WL");
    HIPCHK(hipStreamSynchronize(c.stream), "gn setup");
    c.gn_ready = true;
    return GRA_OK;
}

namespace gra {
namespace {
// After a two-pass solve the misfit product is an affine single-read pass only
// on a G of at least this many bytes: there one pass less per step saves
// milliseconds (2 % of the steps on a 94 GB G), while on a 1.28 GB G the
// cooperative pass (200 us, with launch gaps of 10-15 us either side) and the
// two ordinary products (185 + 178 us) came out even.
constexpr double kAffineMinBytes = 4e9;

// One Gauss-Newton step queued on the context stream (no host
// synchronisation): its results land in c.host_stage.  When the step may use
// a single-read (cooperative) pass -- the fused iteration or the affine
// misfit pass -- the model and the IRLS weights are first copied aside so
// that a failed wait can be undone (gn_collect).
gra_status gn_enqueue(Ctx &c, double lambda, double cg_tol, int cg_maxit)
{
    if (!c.gn_ready) return set_error(c, GRA_BAD_ARG, "gra_gn_setup_problem not called");
    if (!c.Ws.present || c.Ws.n != c.Nm)
        return set_error(c, GRA_BAD_ARG, "regularisation operators not uploaded for this Nm");
    if (cg_maxit < 0 || !(lambda >= 0)) return set_error(c, GRA_BAD_ARG, "maxit/lambda");
    // the step's misfit product may be an affine single-read pass (below)
    const bool big_G = (double)c.Nd * (double)c.ld * (double)dtype_size(c.store) >= kAffineMinBytes;
    c.gn_backup = !c.fused_failed && (c.gn_fused || (big_G && !c.wav));
    if (c.gn_backup) {
        const size_t bm = sizeof(double) * (size_t)c.ld;
        HIPCHK(c.gn_m_prev.reserve(bm), "alloc model backup");
        HIPCHK(hipMemcpyAsync(c.gn_m_prev.p, c.gn_m.p, bm, hipMemcpyDeviceToDevice, c.stream),
               "model backup");
        if (c.gn_have_wl) {
            const size_t bw = sizeof(double) * 4 * (size_t)c.Nm;
            HIPCHK(c.gn_WL_prev.reserve(bw), "alloc weight backup");
            HIPCHK(hipMemcpyAsync(c.gn_WL_prev.p, c.gn_WL.p, bw, hipMemcpyDeviceToDevice, c.stream),
                   "weight backup");
        }
        c.gn_backup_wl = c.gn_have_wl;
        c.gn_backup_partn = c.partn_holds_gn_model;
    }
    const int64_t nm = c.Nm, nd = c.Nd, ld = c.ld;
    double *m = c.gn_m.as<double>(), *mt = c.gn_mtrial.as<double>();
    double *scr = c.gn_one.as<double>();
    double *dmis = scr + nm;
    int *flags = reinterpret_cast<int *>(scr + nm + 1);
    // P from the current model; m_trial starts from m (GN.cpp:825, :856-861)
    HIPCHK(launch_barrier_P(m, c.gn_mmin.as<double>(), c.gn_mmax.as<double>(), nm,
                            c.gn_P.as<double>(), c.stream),
           "barrier_P");
    HIPCHK(hipMemcpyAsync(mt, m, sizeof(double) * (size_t)ld, hipMemcpyDeviceToDevice, c.stream),
           "m_trial");
    gra_cg_args a;
    memset(&a, 0, sizeof(a));
    a.tol = cg_tol;
    a.maxit = cg_maxit;
    a.lambda = lambda;
    a.quirk_mode = 1;
    a.fused = c.gn_fused;
    a.use_cross_gradient = c.gn_crg ? 1 : 0;
    a.a_crg = c.gn_a_crg;
    CgDiag dg{c.gn_Wd.as<double>(), c.gn_P.as<double>(), nullptr, nullptr, nullptr, nullptr,
              nullptr, nullptr, nullptr, c.gn_have_S ? c.gn_S.as<double>() : nullptr};
    if (c.gn_have_wl) {
        double *wl = c.gn_WL.as<double>();
        dg.WLs = wl;
        dg.WLx = wl + nm;
        dg.WLy = wl + 2 * nm;
        dg.WLz = wl + 3 * nm;
    }
    gra_status s = cgls_solve(c, a, m, c.gn_mref.as<double>(), c.gn_dobs.as<double>(), dg,
                              c.gn_dx.as<double>(), nullptr, nullptr, c.partn_holds_gn_model,
                              true);
    if (s != GRA_OK) return s;
    HIPCHK(launch_barrier_update(m, c.gn_dx.as<double>(), c.gn_mmin.as<double>(),
                                 c.gn_mmax.as<double>(), nm, mt, scr, flags, c.stream),
           "barrier_update");
    const GemvShape shape = gemv_shape(c);
    HIPCHK(c.part_n.reserve(sizeof(double) * (size_t)shape.n_splits * nd), "alloc");
    // The misfit product is an affine single-read pass where a grid fits:
    // G m_trial (into part_n, one split) and, from the same read of G,
    // G^T (Wd^2 (G m_trial - dobs)) for the next solve's initial residual --
    // one pass over G instead of two per step, after single-read and
    // two-pass solves alike (launch_fused orders the cooperative grids of a
    // device, also across the streams of gra_gn_step_batch).  Its bounded
    // waits report through the solve's status flag, and a failed wait redoes
    // the step (gn_collect).  (With
    // the mixed-precision copy the pass reads the FP64 G, as the misfit and
    // the initial residual of that mode do.)
    FusedPlan plan;
    const bool two_pass_step = !(c.gn_fused && c.pend_fused);
    bool affine = c.gn_backup && !c.wav && (!two_pass_step || big_G) &&
                  fused_plan(c.store, nd, ld, c.device, c.fused_wgs, c.fused_depth, c.fused_math,
                             c.fused_min_blocks, &plan);
    if (affine && c.fused_scratch.bytes < fused_scratch_bytes(plan, ld)) {
        // a two-pass solve reserved none: the next solve reads gn_q from here
        if (!two_pass_step || c.fused_scratch.reserve(fused_scratch_bytes(plan, ld)) != hipSuccess)
            affine = false;
    }
    c.pend_affine = affine;
    if (affine) {
        int *fused_err = c.status.as<int>() + 1;
        if (two_pass_step)     // the solve did not clear the flag
            HIPCHK(hipMemsetAsync(fused_err, 0, sizeof(int), c.stream), "clear status");
        double *zp = nullptr, *t2p = nullptr;
        {
            TimedRegion tr(c, kTimeAffine);
            HIPCHK(launch_fused(plan, c.G.p, nd, ld, mt, c.gn_Wd.as<double>(), nullptr,
                                c.fused_scratch.p, fused_err, nullptr, c.fused_fail_block, &zp,
                                &t2p, c.stream, c.gn_dobs.as<double>(), c.part_n.as<double>()),
                   "affine single-read pass");
        }
        HIPCHK(launch_misfit(c.part_n.as<double>(), 1, nd, c.gn_Wd.as<double>(),
                             c.gn_dobs.as<double>(), c.red.as<double>(), dmis, c.stream),
               "misfit");
        HIPCHK(hipMemcpyAsync(&c.host_stage->ferr, fused_err, sizeof(int), hipMemcpyDeviceToHost,
                              c.stream),
               "copy status");
        c.gn_gm_splits = 1;
        c.gn_q = zp;
        c.gn_q_groups = plan.S;
        c.gn_q_ready = true;
    } else {
        HIPCHK(launch_G_n(c, shape, mt, c.part_n.as<double>(), nullptr, c.stream), "G x");
        HIPCHK(launch_misfit(c.part_n.as<double>(), shape.n_splits, nd, c.gn_Wd.as<double>(),
                             c.gn_dobs.as<double>(), c.red.as<double>(), dmis, c.stream),
               "misfit");
        c.gn_gm_splits = 0;
        c.gn_q_ready = false;
    }
    // part_n now holds G * m_trial, which becomes the next step's G * m_k
    c.partn_holds_gn_model = true;
    // accept the trial model
    HIPCHK(hipMemcpyAsync(m, mt, sizeof(double) * (size_t)nm, hipMemcpyDeviceToDevice, c.stream),
           "accept");
    // IRLS weights for the next step from the accepted model (GN.cpp:962-1022)
    if (c.gn_lp != 2) {
        double *wl = c.gn_WL.as<double>();
        DevCsr *W[4] = {&c.Ws, &c.Wx, &c.Wy, &c.Wz};
        for (int q = 0; q < 4; ++q)
            HIPCHK(launch_irls(W[q]->rp.as<int64_t>(), W[q]->ci.as<int32_t>(), W[q]->v.as<double>(),
                               nm, m, q == 0 ? c.gn_mref.as<double>() : nullptr, c.gn_lp, c.gn_eps2,
                               wl + q * nm, c.stream),
                   "irls");
        c.gn_have_wl = true;
    }
    HIPCHK(hipMemcpyAsync(&c.host_stage->misfit, dmis, sizeof(double), hipMemcpyDeviceToHost,
                          c.stream),
           "misfit");
    c.gn_args = {lambda, cg_tol, cg_maxit};
    return GRA_OK;
}

// After the stream has been synchronised: the step's results; a failed fused
// wait restores the model and weights and redoes the step two-pass (the
// context stays two-pass from then on).
gra_status gn_collect(Ctx &c, double *misfit, double *cg_err, int *cg_iterations)
{
    double err = NAN;
    int its = 0;
    bool rerun = false;
    gra_status s = cgls_collect(c, &err, &its, &rerun);
    if (rerun) {
        const size_t bm = sizeof(double) * (size_t)c.ld;
        HIPCHK(hipMemcpyAsync(c.gn_m.p, c.gn_m_prev.p, bm, hipMemcpyDeviceToDevice, c.stream),
               "model restore");
        if (c.gn_backup_wl)
            HIPCHK(hipMemcpyAsync(c.gn_WL.p, c.gn_WL_prev.p, sizeof(double) * 4 * (size_t)c.Nm,
                                  hipMemcpyDeviceToDevice, c.stream),
                   "weight restore");
        c.gn_have_wl = c.gn_backup_wl;
        // part_n was overwritten by the misfit product of the failed step
        c.partn_holds_gn_model = false;
        c.gn_q_ready = false;
        (void)c.gn_backup_partn;
        if ((s = gn_enqueue(c, c.gn_args.lambda, c.gn_args.tol, c.gn_args.maxit)) != GRA_OK)
            return s;
        HIPCHK(hipStreamSynchronize(c.stream), "gn step");
        return gn_collect(c, misfit, cg_err, cg_iterations);
    }
    if (s != GRA_OK) return s;
    if (misfit) *misfit = c.host_stage->misfit;
    if (cg_err) *cg_err = err;
    if (cg_iterations) *cg_iterations = its;
    return GRA_OK;
}
}  // namespace
}  // namespace gra

gra_status gra_gn_step(gra_ctx *ctx, double lambda, double cg_tol, int cg_maxit, double *misfit,
                       double *cg_err, int *cg_iterations)
{
    CTX_GUARD(ctx);
    gra_status s = gn_enqueue(c, lambda, cg_tol, cg_maxit);
    if (s != GRA_OK) return s;
    HIPCHK(hipStreamSynchronize(c.stream), "gn step");
    return gn_collect(c, misfit, cg_err, cg_iterations);
}

gra_status gra_gn_step_batch(gra_ctx *const *ctxs, int64_t n, const double *lambdas,
                             double cg_tol, int cg_maxit, double *misfits, double *cg_errs,
                             int *cg_iterations)
{
    if (!ctxs || n <= 0 || !lambdas) return GRA_BAD_ARG;
    for (int64_t b = 0; b < n; ++b)
        if (!ctxs[b]) return GRA_BAD_ARG;
    // Every instance's step is queued on its own stream before any is waited
    // for.  Steps whose CG may use the fused product (a cooperatively
    // launched grid that needs the whole device) are chained by events, so
    // that no two such grids are in flight together; two-pass steps run
    // concurrently with everything.
    hipEvent_t prev_fused = nullptr;
    gra_status first = GRA_OK;
    int64_t queued = 0;
    for (int64_t b = 0; b < n && first == GRA_OK; ++b) {
        Ctx &c = ctxs[b]->c;
        std::lock_guard<std::mutex> lock(c.mu);
        c.last_error.clear();
        if (hipSetDevice(c.device) != hipSuccess) {
            first = set_error(c, GRA_HIP_ERROR, "hipSetDevice failed");
            break;
        }
        const bool coop = c.gn_fused && !c.fused_failed;
        if (coop && prev_fused && hipStreamWaitEvent(c.stream, prev_fused, 0) != hipSuccess) {
            first = set_error(c, GRA_HIP_ERROR, "stream wait");
            break;
        }
        gra_status s = gn_enqueue(c, lambdas[b], cg_tol, cg_maxit);
        if (s != GRA_OK) {
            first = s;
            break;
        }
        if (coop) {
            if (!c.batch_event &&
                hipEventCreateWithFlags(&c.batch_event, hipEventDisableTiming) != hipSuccess) {
                first = set_error(c, GRA_HIP_ERROR, "event");
                break;
            }
            if (hipEventRecord(c.batch_event, c.stream) != hipSuccess) {
                first = set_error(c, GRA_HIP_ERROR, "event record");
                break;
            }
            prev_fused = c.batch_event;
        }
        ++queued;
    }
    for (int64_t b = 0; b < queued; ++b) {
        Ctx &c = ctxs[b]->c;
        std::lock_guard<std::mutex> lock(c.mu);
        if (hipSetDevice(c.device) != hipSuccess || hipStreamSynchronize(c.stream) != hipSuccess) {
            if (first == GRA_OK) first = set_error(c, GRA_HIP_ERROR, "batch sync");
            continue;
        }
        if (first != GRA_OK) continue;
        gra_status s = gn_collect(c, misfits ? misfits + b : nullptr,
                                  cg_errs ? cg_errs + b : nullptr,
                                  cg_iterations ? cg_iterations + b : nullptr);
        if (s != GRA_OK) first = s;
    }
    return first;
}

gra_status gra_gn_get_model(gra_ctx *ctx, double *m)
{
    CTX_GUARD(ctx);
    if (!c.gn_ready || !m) return set_error(c, GRA_BAD_ARG, "no resident model");
    HIPCHK(hipMemcpyAsync(m, c.gn_m.p, sizeof(double) * (size_t)c.Nm, hipMemcpyDeviceToHost,
                          c.stream),
           "model");
    HIPCHK(hipStreamSynchronize(c.stream), "model");
    return GRA_OK;
}

gra_status gra_tune_set(gra_ctx *ctx, int param, int64_t value)
{
    CTX_GUARD(ctx);
    GemvShape &o = c.tune_override;
    c.partn_holds_gn_model = false;   // the split layout of the partials may change
    switch (param) {
    case GRA_TUNE_N_ROWS:
        if (value != 0 && value != 2 && value != 4 && value != 8)
            return set_error(c, GRA_BAD_ARG, "rows per wave: 0, 2, 4 or 8");
        o.n_rows = (int)value;
        break;
    case GRA_TUNE_N_UNROLL:
        if (value < 0 || value > 2) return set_error(c, GRA_BAD_ARG, "unroll: 0, 1 or 2");
        o.n_unroll = (int)value;
        break;
    case GRA_TUNE_N_SPLITS:
        if (value < 0 || value > 256) return set_error(c, GRA_BAD_ARG, "splits: 0..256");
        o.n_splits = (int)value;
        break;
    case GRA_TUNE_T_CHUNKS:
        if (value < 0 || value > 1024) return set_error(c, GRA_BAD_ARG, "chunks: 0..1024");
        o.t_chunks = (int)value;
        break;
    case GRA_TUNE_T_UNROLL:
        if (value != 0 && value != 4 && value != 8)
            return set_error(c, GRA_BAD_ARG, "row unroll: 0, 4 or 8");
        o.t_unroll = (int)value;
        break;
    case GRA_TUNE_NONTEMPORAL:
        if (value < -1 || value > 3) return set_error(c, GRA_BAD_ARG, "nontemporal: -1..3");
        o.nontemporal = (int)value;
        break;
    case GRA_TUNE_FUSED_WGS:
        if (value < 0 || value > 65536) return set_error(c, GRA_BAD_ARG, "fused workgroups: 0..65536");
        c.fused_wgs = (int)value;
        break;
    case GRA_TUNE_FUSED_DEPTH:
        if (value < 0 || value > 3)
            return set_error(c, GRA_BAD_ARG, "fused lag variant: 0 (default) or 1..3");
        c.fused_depth = (int)value;
        break;
    case GRA_TUNE_LAST_REFRESHES:
        return set_error(c, GRA_BAD_ARG, "last refreshes is read-only");
    case GRA_TUNE_FUSED_MIN_BLOCKS:
        if (value < 0 || value > 1000000)
            return set_error(c, GRA_BAD_ARG, "fused minimum row blocks: 0..1000000");
        c.fused_min_blocks = (int)value;
        break;
    case GRA_TUNE_CG_GRAPH:
        if (value < 0 || value > 1000)
            return set_error(c, GRA_BAD_ARG, "CG iterations per graph: 0 (off) .. 1000");
        c.cg_graph = (int)value;
        break;
    case GRA_TUNE_FUSED_MATH:
        if (value < 0 || value > 2) return set_error(c, GRA_BAD_ARG, "fused math: 0, 1 or 2");
        c.fused_math = (int)value;
        break;
    case GRA_TUNE_FUSED_FAIL:
        if (value < -1) return set_error(c, GRA_BAD_ARG, "fused failing row block: -1 or >= 0");
        c.fused_fail_block = value;
        break;
    case GRA_TUNE_FUSED_FALLBACKS:
        if (value != 0) return set_error(c, GRA_BAD_ARG, "fused fallbacks: only 0 (re-enable)");
        c.fused_failed = false;
        c.fused_fail_count = 0;
        break;
    case GRA_TUNE_N_BLOCK:
        if (value < -1 || value > 1) return set_error(c, GRA_BAD_ARG, "block-shared rows: -1..1");
        o.n_block = (int)value;
        break;
    case GRA_TUNE_NODE_TILE:
        if (value < 0 || value > 32) return set_error(c, GRA_BAD_ARG, "corner-sum tile: 0..32");
        c.node_tile = (int)value;
        break;

    default:
        return set_error(c, GRA_BAD_ARG, "unknown tuning parameter");
    }
    return GRA_OK;
}

// Average duration (ms) of G x and of G^T w over `reps` alternating launches
// issued back to back on the context stream, as the CG iteration issues them.
static hipError_t time_pair(Ctx &c, const GemvShape &s, int reps, const double *x,
                            const double *w, double *ms_n, double *ms_t)
{
    std::vector<hipEvent_t> ev(3 * (size_t)reps, nullptr);
    hipError_t e = hipSuccess;
    for (auto &v : ev)
        if ((e = hipEventCreateWithFlags(&v, hipEventDisableSystemFence)) != hipSuccess) break;
    if (e == hipSuccess) {
        // one untimed pair first
        e = launch_gemv_n(c.G.p, c.store, c.Nd, c.ld, s, x, c.part_n.as<double>(), nullptr, c.stream);
        if (e == hipSuccess)
            e = launch_gemv_t(c.G.p, c.store, c.Nd, c.ld, s, w, c.part_t.as<double>(), nullptr,
                              c.stream);
        for (int r = 0; r < reps && e == hipSuccess; ++r) {
            if ((e = hipEventRecord(ev[3 * r], c.stream)) != hipSuccess) break;
            e = launch_gemv_n(c.G.p, c.store, c.Nd, c.ld, s, x, c.part_n.as<double>(), nullptr,
                              c.stream);
            if (e == hipSuccess) e = hipEventRecord(ev[3 * r + 1], c.stream);
            if (e == hipSuccess)
                e = launch_gemv_t(c.G.p, c.store, c.Nd, c.ld, s, w, c.part_t.as<double>(), nullptr,
                                  c.stream);
            if (e == hipSuccess) e = hipEventRecord(ev[3 * r + 2], c.stream);
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
    double tn = 0.0, tt = 0.0;
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
        float a = 0.f, b = 0.f;
        if ((e = hipEventElapsedTime(&a, ev[3 * r], ev[3 * r + 1])) != hipSuccess) break;
        if ((e = hipEventElapsedTime(&b, ev[3 * r + 1], ev[3 * r + 2])) != hipSuccess) break;
        tn += a;
        tt += b;
    }
    for (auto &v : ev)
        if (v) (void)hipEventDestroy(v);
    *ms_n = tn / reps;
    *ms_t = tt / reps;
    return e;
}

gra_status gra_autotune(gra_ctx *ctx, int reps, double *ms_n, double *ms_t)
{
    CTX_GUARD(ctx);
    if (c.wav) return set_error(c, GRA_BAD_ARG, "not available for the wavelet-compressed G");
    if (!c.G.p || c.Nd == 0) return set_error(c, GRA_BAD_ARG, "no G assembled");
    const double gbytes = (double)c.Nd * (double)c.Nm * (c.store == GRA_F64 ? 8 : c.store == GRA_F32 ? 4 : 2);
    if (reps <= 0) reps = gbytes < 8e9 ? 6 : 2;
    c.partn_holds_gn_model = false;
    constexpr int kMaxSplits = 16, kMaxChunks = 32;
    HIPCHK(c.part_n.reserve(sizeof(double) * (size_t)kMaxSplits * c.Nd), "alloc part_n");
    HIPCHK(c.part_t.reserve(sizeof(double) * (size_t)kMaxChunks * c.ld), "alloc part_t");
    DevBuf &x = c.vec_nd[1], &w = c.vec_nd[2];
    HIPCHK(x.reserve(sizeof(double) * (size_t)c.ld), "alloc x");
    HIPCHK(w.reserve(sizeof(double) * (size_t)(c.Nd > c.ld ? c.Nd : c.ld)), "alloc w");
    HIPCHK(hipMemsetAsync(x.p, 0, x.bytes, c.stream), "clear x");
    HIPCHK(hipMemsetAsync(w.p, 0, w.bytes, c.stream), "clear w");
    const double *xp = x.as<double>(), *wp = w.as<double>();

    GemvShape best = gemv_shape(c);
    double bn = 0.0, bt = 0.0;
    HIPCHK(time_pair(c, best, reps, xp, wp, &bn, &bt), "time default");
    // G^T w: row chunks x rows per iteration
    const int64_t max_chunks = c.Nd / 32 > 0 ? c.Nd / 32 : 1;
    for (int ch : {1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 16, 24, 32}) {
        if (ch > max_chunks) break;
        for (int un : {4, 8}) {
            GemvShape s = best;
            s.t_chunks = ch;
            s.t_unroll = un;
            double tn, tt;
            HIPCHK(time_pair(c, s, reps, xp, wp, &tn, &tt), "time G^T w");
            if (tt < bt) { best = s; bt = tt; bn = tn; }
        }
    }
    // G x: wave-owned / block-shared rows x rows x column steps x column splits
    const int V = c.store == GRA_F64 ? 2 : (c.store == GRA_F32 ? 4 : 8);
    for (int blk : {0, 1}) {
        const int64_t span = blk ? 2 * V * 256 : 4 * V * 64;   // columns a split keeps at least
        const int64_t max_splits = c.Nm / span > 0 ? c.Nm / span : 1;
        for (int rows : {2, 4, 8, 12, 16}) {
            if (c.store == GRA_F64 && rows == 8 && !blk) continue;
            if (rows > 8 && (c.store != GRA_BF16 || blk)) continue;   // BF16, wave-owned only
            for (int un : {1, 2})
                for (int sp : {1, 2, 4, 8, 16}) {
                    if (sp > max_splits) continue;
                    GemvShape s = best;
                    s.n_block = blk;
                    s.n_rows = rows;
                    s.n_unroll = un;
                    s.n_splits = sp;
                    double tn, tt;
                    HIPCHK(time_pair(c, s, reps, xp, wp, &tn, &tt), "time G x");
                    if (tn < bn) { best = s; bn = tn; bt = tt; }
                }
        }
    }
    // non-temporal loads: judged on the pair (they interact across kernels)
    for (int nt = 0; nt <= 3; ++nt) {
        if (nt == best.nontemporal) continue;
        GemvShape s = best;
        s.nontemporal = nt;
        double tn, tt;
        HIPCHK(time_pair(c, s, reps, xp, wp, &tn, &tt), "time non-temporal");
        if (tn + tt < bn + bt) { best = s; bn = tn; bt = tt; }
    }
    c.tune_override = best;
    if (ms_n) *ms_n = bn;
    if (ms_t) *ms_t = bt;
    return GRA_OK;
}

gra_status gra_tune_get(gra_ctx *ctx, int param, int64_t *value)
{
    CTX_GUARD(ctx);
    if (!value) return set_error(c, GRA_BAD_ARG, "null output");
    GemvShape s = gemv_shape(c);
    switch (param) {
    case GRA_TUNE_N_ROWS: *value = s.n_rows; break;
    case GRA_TUNE_N_UNROLL: *value = s.n_unroll; break;
    case GRA_TUNE_N_SPLITS: *value = s.n_splits; break;
    case GRA_TUNE_T_CHUNKS: *value = s.t_chunks; break;
    case GRA_TUNE_T_UNROLL: *value = s.t_unroll; break;
    case GRA_TUNE_NONTEMPORAL: *value = s.nontemporal; break;
    case GRA_TUNE_FUSED_WGS: *value = c.fused_wgs; break;
    case GRA_TUNE_FUSED_DEPTH: *value = c.fused_depth; break;
    case GRA_TUNE_FUSED_FAIL: *value = c.fused_fail_block; break;
    case GRA_TUNE_FUSED_MATH: *value = c.fused_math; break;
    case GRA_TUNE_LAST_REFRESHES: *value = c.last_refreshes; break;
    case GRA_TUNE_CG_GRAPH: *value = c.cg_graph; break;
    case GRA_TUNE_FUSED_MIN_BLOCKS: *value = c.fused_min_blocks; break;
    case GRA_TUNE_FUSED_FALLBACKS: *value = c.fused_fail_count; break;
    case GRA_TUNE_N_BLOCK: *value = s.n_block; break;
    case GRA_TUNE_NODE_TILE: *value = c.node_tile; break;
    default: return set_error(c, GRA_BAD_ARG, "unknown tuning parameter");
    }
    return GRA_OK;
}

gra_status gra_timing_enable(gra_ctx *ctx, int on)
{
    CTX_GUARD(ctx);
    if (on < 0) return set_error(c, GRA_BAD_ARG, "timing: negative stride");
    c.timing = on;
    return GRA_OK;
}

gra_status gra_timing_reset(gra_ctx *ctx)
{
    CTX_GUARD(ctx);
    HIPCHK(hipStreamSynchronize(c.stream), "sync");
    c.ev_marks.clear();
    c.ev_used = 0;
    return GRA_OK;
}

gra_status gra_timing_get(gra_ctx *ctx, int which, double *avg_ms, int64_t *count)
{
    CTX_GUARD(ctx);
    if (which < 0 || which >= kTimeClasses) return set_error(c, GRA_BAD_ARG, "timing class");
    HIPCHK(hipStreamSynchronize(c.stream), "sync");
    double tot = 0.0;
    int64_t n = 0;
    for (auto &mk : c.ev_marks) {
        if (mk.first != which) continue;
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, mk.second.first, mk.second.second), "elapsed");
        tot += ms;
        ++n;
    }
    if (avg_ms) *avg_ms = n ? tot / (double)n : 0.0;
    if (count) *count = n;
    return GRA_OK;
}

}  // extern "C"
