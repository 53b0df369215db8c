// Internal declarations shared by the engine's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "graeng.h"

namespace gra {

// ------------------------------------------------------------ device buffer
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    // Grows to at least `n` bytes; contents are not preserved.
    hipError_t reserve(size_t n)
    {
        if (n <= bytes && p) return hipSuccess;
        release();
        hipError_t e = hipMalloc(&p, n ? n : 256);
        if (e == hipSuccess) bytes = n;
        return e;
    }
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

struct DevCsr {
    int64_t n = 0, nnz = 0;
    DevBuf rp, ci, v;          // A
    DevBuf rpT, ciT, vT;       // A^T (columns ascending in each row)
    bool present = false;
};

// Device-resident scalar state of one CGLS solve.  Every CG kernel reads
// `done` first and returns when it is set, so a captured or pre-queued
// sequence of iterations drains as no-ops after convergence.
struct CgScalars {
    double r2[2];     // ||r||^2 of the residual entering iteration `it`: r2[it & 1]
    double t2;
    double alpha;
    double beta;
    double ATb2;
    double err;
    int done;
    int iterations;
    int breakdown;
    int it_dev;       // iteration index for graph-replayed iterations (dev_it)
};

// Launch shape of the two G products (see linalg_kernels.hip).
struct GemvShape {
    int n_rows = 4;        // rows per wave, G x (2, 4, 8)
    int n_unroll = 2;      // column steps per loop iteration (1, 2)
    int n_splits = 1;      // column splits (partial sums summed by the consumer)
    int t_chunks = 1;      // row chunks of G^T w
    int t_unroll = 8;      // rows per loop iteration (4, 8)
    int nontemporal = 3;   // non-temporal loads of G: bit 0 for G x, bit 1 for G^T w
    int n_block = 0;       // G x: 1 = the waves of a block share its n_rows rows
};

// Kernel timing classes (gra_timing_get).
enum TimingClass {
    kTimeGemvN = 0, kTimeGemvT = 1, kTimeAssemble = 2, kTimeCgIter = 3, kTimeFused = 4,
    kTimeMatMul = 5, kTimeRefresh = 6, kTimeCgGraph = 7, kTimeAffine = 8, kTimeClasses = 9
};

// Launch plan of the single-read fused product (fused_kernels.hip).
struct FusedPlan {
    int NT = 0, S = 0;            // column tiles, row groups (grid = NT * S)
    int per_xcd = 0, n_xcd = 1;   // XCD-aligned workgroup mapping when per_xcd > 0
    int grid = 0;
    int64_t rows_per_group = 0;
    int64_t nrb_group = 0;        // row blocks per group
    int RB = 0;                   // rows per block
    int threads = 0;
    uint64_t wait_ticks = 0;      // bound of the kernel's waits (wall-clock ticks)
    int L = 0;                    // coefficient lag (blocks)
    const void *kernel = nullptr;
};

// Pinned host staging of a context's per-solve results (read after a sync).
struct HostStage {
    CgScalars sc;
    int ferr;
    double misfit;
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;      // where the context's work is queued
    hipStream_t own_stream = nullptr;  // created with the context
    std::mutex mu;
    std::string last_error;

    // G (row-major, leading dimension ld >= Nm, padding columns are zero)
    DevBuf G;
    int64_t Nd = 0, Nm = 0, N_obs = 0, ld = 0;
    uint32_t flag = 0;
    gra_dtype store = GRA_F64;
    uint64_t g_version = 0;       // bumped whenever G changes
    // G's values come from corner sums (gra_assemble_corners / _rectilinear /
    // _batch_corners), not the face/edge form; pts_zmax: the deepest point
    // (all points lie strictly above every prism of such a G)
    bool g_corners = false;
    double pts_zmax = 0.0;

    // mixed precision (gra_mixed_precision): a reduced-precision copy of the
    // FP64 G for the CG iteration's products, FP64 residual refreshes every
    // mixed_refresh iterations
    DevBuf Glo;
    gra_dtype store_lo = GRA_F64;
    uint64_t lo_version = ~0ull;  // g_version the copy was made from
    int mixed_refresh = 0;        // 0: off
    int last_refreshes = 0;       // FP64 refreshes of the last solve

    // regularisation operators
    DevCsr Ws, Wx, Wy, Wz, Tx0, Ty0, Tz0;
    bool have_crg_ops = false;

    // scratch
    DevBuf part_n, part_t;     // GEMV partial sums
    DevBuf fused_scratch, rloop;
    DevBuf vec_nd[4], vec_nm[12], vec_row[3];
    DevBuf fwd_prisms, fwd_rho, fwd_pts, fwd_out;   // gra_forward inputs/outputs (grow-only)
    DevBuf fwd_axes, fwd_w;                          // gra_forward_rectilinear
    DevBuf fwd_items;                                // gra_forward_rectilinear_batch
    DevBuf rect_axes;                                // gra_assemble_rectilinear
    DevBuf mm_in, mm_pack, mm_part, mm_out;         // gra_G_mat_mul / gra_GT_mat_mul
    int node_tile = 0;                               // its tile side (0: by LDS budget)
    DevBuf pts_dev;                                  // the assembled G's points (expansion)
    DevBuf batch_prisms, batch_items, batch_bad;     // gra_assemble_batch (first context)
    DevBuf red;                // block partial sums for reductions
    DevBuf scal;               // CgScalars
    DevBuf status;             // int flags written by kernels
    double *noalloc_scratch = nullptr;   // this device's `scratch` (gra_G_vec_mul_noalloc)
    HostStage *host_stage = nullptr;     // pinned (hipHostMalloc at context creation)
    // the solve / Gauss-Newton step whose results host_stage will hold
    bool pend_fused = false;
    bool pend_affine = false;            // the queued step's misfit is an affine single-read pass
    int pend_maxit = 0, pend_refresh = 0;
    bool gn_backup = false;              // gn_m_prev / gn_WL_prev hold the step's inputs
    bool gn_backup_wl = false, gn_backup_partn = false;
    DevBuf gn_m_prev, gn_WL_prev;
    struct { double lambda, tol; int maxit; } gn_args{0.0, 0.0, 0};   // the queued step
    hipEvent_t batch_event = nullptr;    // gra_gn_step_batch ordering of fused steps
    DevBuf host_io;            // unused placeholder

    // wavelet-compressed G (gra_assemble_wavelet): replaces the dense G while
    // `wav` is set.  Rows: start / length into (column, value); columns:
    // pointers into (row, value); n_dy-long work vectors.
    bool wav = false;
    int64_t wav_ndy = 0, wav_nnz = 0, wav_nlds = 0, wav_cus = 256;
    DevBuf wav_order;                    // int32, Nm: reordered position -> cell
    DevBuf wav_rs, wav_rl, wav_ci, wav_v;
    DevBuf wav_cp, wav_ri, wav_vt;
    DevBuf wav_seg_start, wav_seg_len, wav_col_seg, wav_part;   // CSC column segments
    int64_t wav_nseg = 0;
    DevBuf wav_x, wav_z, wav_tmp;

    // resident Gauss-Newton state
    bool gn_ready = false;
    bool gn_crg = false, gn_have_S = false;   // resident loop with the cross-gradient term
    double gn_a_crg = 0.0;
    DevBuf gn_S;
    DevBuf gn_dobs, gn_Wd, gn_mmin, gn_mmax, gn_mref, gn_m, gn_mtrial, gn_P, gn_dx, gn_one;
    DevBuf gn_WL;              // 4 x Nm: WL_s, WL_x, WL_y, WL_z (IRLS)
    // part_n holds the column-split partials of G * gn_m (left by the last
    // gn_step's misfit product); the next step reuses them for G * m_k.
    bool partn_holds_gn_model = false;
    // With the single-read iteration, the gn_step's misfit product is an
    // affine single-read pass: part_n then holds G * gn_m in one split
    // (gn_gm_splits) and gn_q the row-group partials (gn_q_groups of ld) of
    // G^T (Wd^2 (G gn_m - dobs)), which the next solve uses for its initial
    // residual instead of a G^T product; valid with partn_holds_gn_model.
    int gn_gm_splits = 0;
    bool gn_q_ready = false;
    const double *gn_q = nullptr;
    int gn_q_groups = 0;
    bool last_solve_fused = false;   // the last CG solve used the single-read product
    int gn_fused = 0;
    int gn_lp = 2;
    double gn_eps2 = 0.0;
    bool gn_have_wl = false;

    // launch-shape overrides (0 = automatic; nontemporal -1 = automatic)
    GemvShape tune_override{0, 0, 0, 0, 0, -1, -1};
    int fused_wgs = 0;            // 0: one per CU
    int fused_depth = 0;          // 0: default (lag, prefetch) variant
    int fused_math = 0;           // 0: default per storage type, 1: FP64, 2: FP32 products
    int cg_graph = 0;             // two-pass CG iterations per captured hipGraph (0: none)
    int fused_min_blocks = 0;     // fused only with at least this many row blocks per group
    int64_t fused_fail_block = -1;   // test hook: force a failed wait at this row block
    bool fused_failed = false;    // a fused solve failed on this context: two-pass from now on
    int fused_fail_count = 0;     // fused solves rerun with the two-pass iteration

    // timing
    int timing = 0;            // 0 off; n: every launch, CG-loop launches every n-th iteration
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_marks;
};

// Records a timed region on the context stream when timing is enabled.
struct TimedRegion {
    Ctx &c;
    int cls;
    hipEvent_t a = nullptr, b = nullptr;
    // iter: the CG iteration of a launch inside the CG loop (timed when
    // iter % timing == 0), 0 elsewhere
    TimedRegion(Ctx &ctx, int klass, int iter = 0);
    ~TimedRegion();
};

// ------------------------------------------------------------ launchers
// All launchers enqueue on `st` and return the launch error.

// G assembly (prism_kernels.hip): G[(k*N_obs+i)*ld + j] in `store` type.
hipError_t launch_assemble(const double *d_prisms, int64_t Nm, const double *d_pts, int64_t N_obs,
                           uint32_t flag, gra_dtype store, void *G, int64_t ld, int *d_bad,
                           hipStream_t st);
// Matrix-free forward: out[i*nc+k].
// One item of a batched assembly (device-side table, sorted by block0).
struct BatchItem {
    const double *prisms, *pts;   // device
    void *G;                      // device, Nd x ld of the item's storage type
    int64_t Nm, N_obs, ld;
    int64_t bpr;                  // blocks per point row: ceil(Nm / 256)
    int64_t block0;               // first block of the item in the launch
    uint32_t flag;
};
hipError_t launch_assemble_batch(const BatchItem *d_items, int n_items, int64_t total_blocks,
                                 gra_dtype store, int *d_bad, hipStream_t st);
// Per-prism corner sums over the same item table (rectilinear_kernels.hip).
hipError_t launch_cell_corner_batch(const BatchItem *d_items, int n_items, int64_t total_blocks,
                                    gra_dtype store, hipStream_t st);

// One request of a batched node-sum forward (device pointers; items sorted
// by block0, the first point's workgroup in the launch).
struct FwdItem {
    const double *xs, *ys, *zs, *rho, *pts;
    double *w, *out;
    int64_t nx, ny, nz, N_obs, block0;
    uint32_t flag;
};
hipError_t launch_node_forward_batch(const FwdItem *h_items, const FwdItem *d_items, int n_items,
                                     int64_t total_blocks, hipStream_t st);
// Node-sum forward on a rectilinear mesh (rectilinear_kernels.hip); w holds
// (nx+1)(ny+1)(nz+1) node weights.
// Multi-right-hand-side products on the matrix cores (gemm_kernels.hip).
// ncol: right-hand sides padded to tiles of 16; span: summation columns
// (G x) or rows (G^T r) per split; ldp: leading dimension of the partials.
struct MatMulPlan {
    int ncol = 16;          // right-hand sides rounded up to whole 16-wide tiles
    int splits = 1;         // splits of the summation range
    int64_t span = 0;       // its length per split
    int64_t ldp = 0;        // partial-sum rows per right-hand side
    int64_t pack_rows = 0;  // rows of the packed right-hand sides
    int mc = 1;             // G^T R: column tiles per wave
};
MatMulPlan plan_mat_mul(gra_dtype dt, bool trans, int64_t Nd, int64_t Nm, int64_t ld, int64_t k);
// Bytes of the packed right-hand sides of a plan.
size_t mat_mul_pack_bytes(gra_dtype dt, const MatMulPlan &p);
// Right-hand sides (host layout, n rows x k columns) in the layout and
// element type the product kernels read: FP64 for FP64 G, FP32 for FP32 G,
// three BF16 parts for BF16 G.
hipError_t launch_pack_rhs(gra_dtype dt, bool trans, const double *src, int64_t n, int64_t k,
                           const MatMulPlan &p, void *dst, hipStream_t st);
hipError_t launch_mat_mul(const void *G, gra_dtype dt, bool trans, int64_t Nd, int64_t Nm,
                          int64_t ld, const void *V, const MatMulPlan &p, double *part,
                          hipStream_t st);
hipError_t launch_sum_splits(const double *part, const MatMulPlan &p, int64_t n, int64_t k,
                             double *out, hipStream_t st);

// G of a rectilinear mesh by corner sums (rectilinear_kernels.hip): tile =
// cell columns per tile side; its corner values take
// 8 * ncomp * (tile+1)^2 * (nz+1) bytes of LDS.
hipError_t launch_cell_corner(const double *prisms, int64_t Nm, const double *pts, int64_t N_obs,
                              uint32_t flag, gra_dtype store, void *G, int64_t ld, hipStream_t st);
hipError_t launch_node_assemble(const double *xs, const double *ys, const double *zs, int64_t nx,
                                int64_t ny, int64_t nz, const double *pts, int64_t N_obs,
                                uint32_t flag, gra_dtype store, void *G, int64_t ld, int tile,
                                hipStream_t st);
int node_assemble_tile(int ncomp, int64_t nz);
hipError_t launch_node_forward(const double *xs, const double *ys, const double *zs, int64_t nx,
                               int64_t ny, int64_t nz, const double *rho, double *w,
                               const double *pts, int64_t N_obs, uint32_t flag, double *out,
                               hipStream_t st);
hipError_t launch_forward(const double *d_prisms, const double *d_rho, int64_t Nm,
                          const double *d_pts, int64_t N_obs, uint32_t flag, double *d_out,
                          int *d_bad, hipStream_t st);

// GEMV (linalg_kernels.hip).  gemv_n writes part[s*Nd + i] for s < n_splits;
// x must hold ld entries (zero beyond Nm).  gemv_t writes part[c*ld + j] for
// c < t_chunks.
GemvShape default_gemv_shape(gra_dtype dt, int64_t Nd, int64_t Nm);
GemvShape gemv_shape(const Ctx &c);   // defaults of the context's G + overrides
hipError_t launch_gemv_n(const void *G, gra_dtype dt, int64_t Nd, int64_t ld, const GemvShape &s,
                         const double *x, double *part, const int *done, hipStream_t st);
hipError_t launch_gemv_t(const void *G, gra_dtype dt, int64_t Nd, int64_t ld, const GemvShape &s,
                         const double *w, double *part, const int *done, hipStream_t st);
// FP64 G -> FP32 / BF16 copy with the assembly's rounding (BF16 via FP32).
hipError_t launch_convert_G(const double *G, int64_t n, gra_dtype dt, void *out, hipStream_t st);
hipError_t launch_remap_cols(const void *G, int64_t ld, const void *Gc, int64_t ldc,
                             const int32_t *src, int64_t Nd, int64_t Nm2, void *G2, int64_t ld2,
                             int elem_bytes, hipStream_t st);
hipError_t launch_col_sumsq(const void *G, gra_dtype dt, int64_t Nd, int64_t ld,
                            const GemvShape &s, double *part, hipStream_t st);
// y[i] = scale_i * sum_s part[s*stride + i]  (scale may be null => 1)
hipError_t launch_reduce_parts(const double *part, int nparts, int64_t stride, int64_t n,
                               const double *scale, double *y, hipStream_t st);

// Products of the context's G, dense or wavelet-compressed (wavelet_kernels.hip).
// The wavelet form writes y (Nd) / z (ld entries, zero beyond Nm) directly
// (one split / chunk).
hipError_t launch_G_n(Ctx &c, const GemvShape &s, const double *x, double *part, const int *done,
                      hipStream_t st);
hipError_t launch_G_t(Ctx &c, const GemvShape &s, const double *w, double *part, const int *done,
                      hipStream_t st);
int64_t wavelet_length(int64_t n);
// compute_G_wavelet on the device: prisms already in reordered order,
// c.wav_order uploaded; fills the wav_* buffers.
gra_status wavelet_assemble(Ctx &c, const double *d_prisms_ro, int64_t Nm, const double *d_pts,
                            int64_t N_obs, uint32_t flag, double rel, int *d_bad);

// Copy G to a float64 host-layout device buffer (row- or column-major).
hipError_t launch_expand_G(const void *G, gra_dtype dt, int64_t Nd, int64_t Nm, int64_t ld,
                           double *out, int col_major, hipStream_t st);

// Fused product (fused_kernels.hip).  fused_plan returns false when no grid
// can be co-resident on `device` for this G; target_wgs <= 0 means one
// workgroup per CU, larger values are capped at the co-resident limit.
// math: 0 default for the storage type, 1 FP64 products, 2 FP32 products.
bool fused_plan(gra_dtype dt, int64_t Nd, int64_t ld, int device, int target_wgs, int depth,
                int math, int min_blocks, FusedPlan *out);
size_t fused_scratch_bytes(const FusedPlan &p, int64_t ld);
// The fused product's cross-workgroup partials: the first fused_partials(p)
// 64-bit words of its scratch, which must hold kFusedSentinel (a signalling
// NaN) when a launch starts.
constexpr uint64_t kFusedSentinel = 0x7FF4DEADBEEF0001ull;
int64_t fused_partials(const FusedPlan &p);
// fail_block >= 0 makes the wait on that row block fail (test hook).
// dobs / yout: the affine pass (q = G^T (Wd^2 (G x - dobs)), y = G x to yout).
// partials_filled: the caller has queued the sentinel fill itself (the CG
// iteration does it inside its p update, one launch fewer per iteration);
// otherwise launch_fused queues a fill kernel before the grid.
hipError_t launch_fused(const FusedPlan &p, const void *G, int64_t Nd, int64_t ld,
                        const double *x, const double *Wd, double *tdat, void *scratch, int *err,
                        const int *done, int64_t fail_block, double **zpart_out,
                        double **t2part_out, hipStream_t st, const double *dobs = nullptr,
                        double *yout = nullptr, bool partials_filled = false);

// CGLS (cgls.hip)
struct CgDiag {
    const double *Wd, *P;
    const double *WLs, *WLx, *WLy, *WLz, *WLtx, *WLty, *WLtz;  // may be null (identity)
    const double *S;                                            // may be null
};
// have_Gmk: c.part_n already holds the partials of G * d_mk (same launch
// shape), so the initial product is not recomputed.  a.fused selects the
// single-read iteration (falls back to the two-pass one when the fused grid
// cannot be co-resident).
// defer: enqueue only (results to c.host_stage); cgls_collect after a sync
// reads them (rerun: the fused product failed, the solve must be redone
// with the two-pass iteration).
gra_status cgls_solve(Ctx &c, const gra_cg_args &a, const double *d_mk, const double *d_mref,
                      const double *d_dobs, const CgDiag &dg, double *d_dx, double *err,
                      int *iterations, bool have_Gmk = false, bool defer = false);
gra_status cgls_collect(Ctx &c, double *err, int *iterations, bool *rerun);
hipError_t launch_barrier_P(const double *m, const double *mmin, const double *mmax, int64_t n,
                            double *P, hipStream_t st);
hipError_t launch_barrier_update(const double *m, const double *dx, const double *mmin,
                                 const double *mmax, int64_t n, double *m_trial, double *scratch,
                                 int *flags, hipStream_t st);
hipError_t launch_misfit(const double *part, int splits, int64_t Nd, const double *Wd,
                         const double *dobs, double *red, double *out, hipStream_t st);
// WL_i = ((W (m - mref))_i^2 + eps2)^(-(2-p)/4); mref may be null.
hipError_t launch_irls(const int64_t *rp, const int32_t *ci, const double *v, int64_t n,
                       const double *m, const double *mref, int p, double eps2, double *WL,
                       hipStream_t st);

// Host helpers
gra_status set_error(Ctx &c, gra_status s, const std::string &msg);
gra_status hip_fail(Ctx &c, hipError_t e, const char *what);
gra_status upload_csr(Ctx &c, DevCsr &d, const gra_csr *h);

}  // namespace gra

struct gra_ctx {
    gra::Ctx c;
};
