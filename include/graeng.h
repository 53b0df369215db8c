/*
 * graeng — MI355X-native engine for GraAda3D's prism-kernel / CG hot path.
 *
 * Plain C ABI: plain pointers and sizes, no C++ or GPU-runtime types.  Every
 * entry point returns a gra_status; no exception crosses the boundary.  All
 * host pointers are borrowed for the duration of the call only.  Calls on one
 * context are serialised by the context's own lock; separate contexts may be
 * used from separate threads concurrently.
 *
 * Units and conventions follow the reference (src/Forward_modelling/gs.h):
 * lengths in m, z positive DOWN, density in kg/m^3, G0 = 6.673e-11,
 * g-components in mGal (1e5 SI), gradient components in Eotvos (1e9 SI).
 *
 * Prisms are float64 (Nm x 6) rows {x0, x1, y0, y1, z0, z1} with x0<x1,
 * y0<y1, z0<z1 (RectPrism, src/Forward_modelling/RectPrism.h:41-43); points
 * are float64 (N_obs x 3) rows {x, y, z}.
 */
#ifndef GRAENG_H
#define GRAENG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gra_status {
    GRA_OK = 0,
    GRA_BAD_ARG = 1,               /* invalid argument / state                     */
    GRA_OOM = 2,                   /* device or host allocation failed            */
    GRA_HIP_ERROR = 3,             /* GPU runtime error (message: gra_last_error)  */
    GRA_NUMERIC_BREAKDOWN = 4,     /* CG step with t2 == 0 or a non-finite scalar  */
    GRA_DEGENERATE_GEOMETRY = 5,   /* point on an edge extension (Integral.cpp:77) */
    GRA_NO_DEVICE = 6              /* no GPU visible                               */
} gra_status;

typedef enum gra_dtype { GRA_F64 = 0, GRA_F32 = 1, GRA_BF16 = 2 } gra_dtype;

/* Field bits, identical to GS::ComputationOptions (gs.h:46-57).  Bit 0 (the
 * potential V) is not part of this engine and is rejected. */
#define GRA_GZ  0x0002u
#define GRA_GX  0x0004u
#define GRA_GY  0x0008u
#define GRA_TZZ 0x0010u
#define GRA_TXZ 0x0020u
#define GRA_TYZ 0x0040u
#define GRA_TXX 0x0080u
#define GRA_TXY 0x0100u
#define GRA_TYY 0x0200u

typedef struct gra_ctx gra_ctx;

/* Square CSR matrix (n x n): row_ptr[n+1] (int64), col_idx[nnz] (int32),
 * val[nnz].  The reference's SMatrix is Eigen::SparseMatrix<double,RowMajor>
 * (gs.h:29), i.e. the same storage. */
typedef struct gra_csr {
    int64_t n;
    int64_t nnz;
    const int64_t *row_ptr;
    const int32_t *col_idx;
    const double *val;
} gra_csr;

/* ---------------------------------------------------------------- context */
const char *gra_status_string(int status);
/* Message of the last failing call on this context ("" if none). */
const char *gra_last_error(const gra_ctx *ctx);
/* Creates a context on GPU `device` with its own (default-priority) stream. */
gra_status gra_ctx_create(int device, gra_ctx **out);
void gra_ctx_destroy(gra_ctx *ctx);
/* Waits for all work queued on the context's stream. */
gra_status gra_sync(gra_ctx *ctx);
/*
 * No counterpart in the reference (its products run on the host).  For
 * callers that run several contexts' work from one thread of control on one
 * stream, such as a server whose worker streams borrow the context of a
 * resident problem: the context's later calls queue their work on `stream`
 * (a hipStream_t of the context's device, owned by the caller) instead of
 * its own; NULL returns it to its own stream.  The call first waits for the
 * work queued on the stream in use.  Calls that return results wait for the
 * stream they used; gra_G_vec_mul_noalloc is ordered on it only.  The caller
 * keeps `stream` alive while the context uses it.
 */
gra_status gra_ctx_set_stream(gra_ctx *ctx, void *stream);
/* The stream the context's work is queued on (a hipStream_t). */
void *gra_ctx_stream(const gra_ctx *ctx);

/* ---------------------------------------------------------- forward model */
/*
 * Replaces Fwd::compute_G (src/Forward_modelling/Fwd.cpp:47-79).
 * Assembles the dense sensitivity matrix G on the device (it stays there):
 * row k*N_obs + i, column j holds component k (ascending field bit) of the
 * unit-density field of prism j at point i, evaluated in FP64 with the
 * reference's face/edge formulation (GravFormula.cpp:75-272,
 * Integral.cpp:59-85, 540-633, 764-823) and stored in `store` precision.
 * Any previous G of the context is released first.
 */
gra_status gra_assemble(gra_ctx *ctx, const double *prisms, int64_t Nm, const double *pts,
                        int64_t N_obs, uint32_t field_flag, gra_dtype store);

/*
 * Fwd::compute_G (Fwd.cpp:47-79) for a mesh built by
 * Mesh::generate_regular_mesh (Mesh.cpp:828-846) or any other tensor-product
 * mesh: cells are the products of the node coordinates xs[nx+1], ys[ny+1],
 * zs[nz+1] (strictly increasing, z down), column (i*ny + j)*nz + k is cell
 * (i, j, k), the row layout and storage are gra_assemble's.  Each entry is
 * the signed sum of the closed-form corner function over the cell's 8
 * corners (the quantity the reference's face/edge integrals evaluate), with
 * the corner values of a tile of cells computed once per point and shared
 * through LDS.  Agrees with gra_assemble to ~1e-13 of each component block's
 * largest entry; the relative accuracy of a single entry falls with its
 * distance (corner-sum cancellation).  Points must lie strictly above the
 * mesh (z < zs[0]), else GRA_BAD_ARG; a mesh too deep for the LDS tile
 * (more than ~1,900 levels for one component, ~210 for all nine) is
 * GRA_BAD_ARG as well.
 */
gra_status gra_assemble_rectilinear(gra_ctx *ctx, const double *xs, int64_t nx, const double *ys,
                                   int64_t ny, const double *zs, int64_t nz, const double *pts,
                                   int64_t N_obs, uint32_t field_flag, gra_dtype store);

/*
 * Fwd::compute_G (Fwd.cpp:47-79) by corner sums for any prisms (octree and
 * other non-tensor meshes): same arguments, G layout and storage rounding as
 * gra_assemble, each entry c_k * sum over the prism's 8 corners of
 * s * F_k(corner - point), evaluated per (point, prism) with no sharing
 * between prisms, in the difference order of gra_assemble_rectilinear (so a
 * tensor-product mesh gives the same G both ways).  Agrees with gra_assemble
 * to ~1e-12 of each component block's largest entry; the relative accuracy
 * of a single entry falls with its distance (corner-sum cancellation).
 * Points must lie strictly above every prism (z < min z0), else GRA_BAD_ARG.
 */
gra_status gra_assemble_corners(gra_ctx *ctx, const double *prisms, int64_t Nm, const double *pts,
                                int64_t N_obs, uint32_t field_flag, gra_dtype store);

/*
 * Batched Fwd::compute_G over uneven instances (no counterpart in the
 * reference, which assembles one G per inversion): ctxs[b] (distinct, one
 * device) receives the G of items[b] -- its own mesh, points and component
 * set -- exactly as gra_assemble would compute it, but all n instances are
 * evaluated by one kernel launch (on ctxs[0]'s stream) whose blocks are
 * distributed over the instances in proportion to their sizes.  Errors are
 * reported on ctxs[0] (gra_last_error names the instance).
 */
typedef struct {
    const double *prisms;  /* Nm x 6 */
    int64_t Nm;
    const double *pts;     /* N_obs x 3 */
    int64_t N_obs;
    uint32_t field_flag;
} gra_batch_item;
gra_status gra_assemble_batch(gra_ctx *const *ctxs, int64_t n, const gra_batch_item *items,
                              gra_dtype store);

/*
 * gra_assemble_batch with the values of gra_assemble_corners (per-prism
 * corner sums; no counterpart in the reference): the same items, contexts
 * and launch distribution; every item's points must lie strictly above
 * every prism of that item, else GRA_BAD_ARG naming the item.
 */
gra_status gra_assemble_batch_corners(gra_ctx *const *ctxs, int64_t n,
                                      const gra_batch_item *items, gra_dtype store);

/*
 * Replaces Fwd::compute_G_wavelet (src/Forward_modelling/Fwd.cpp:148-219,
 * compress_vec :298-371; the reference's method 2).  order[j] is the cell
 * at reordered position j (Mesh::get_reordered_id, a permutation of
 * 0..Nm-1).  Each row of G over the cells order[0..Nm), zero-padded to
 * n_dy = the next power of two >= Nm, is transformed by the periodic
 * Daubechies-4 wavelet transform (GSL gsl_wavelet_daubechies, k = 4) and
 * only its coefficients c with |c| > t or ||c| - t| < 1e-15,
 * t = relative_threshold * max|c| of the row, are kept on the device.
 * Until the next gra_assemble, gra_G_vec_mul, gra_GT_vec_mul, gra_misfit,
 * gra_solve_cg and the resident Gauss-Newton loop use this compressed G
 * (Fwd.cpp:80-147 wavelet branches); gra_copy_G, gra_col_norms,
 * gra_expand_cols (re-assemble instead, as AdaptiveInversion::refine_mesh
 * does) and gra_autotune return GRA_BAD_ARG, and the single-read CG
 * iteration falls back to the two-pass one.  Any dense G is released.
 */
gra_status gra_assemble_wavelet(gra_ctx *ctx, const double *prisms, int64_t Nm, const double *pts,
                                int64_t N_obs, uint32_t field_flag, const int64_t *order,
                                double relative_threshold);
/* n_dy and the number of kept coefficients of the compressed G
 * (the reference prints both, Fwd.cpp:210-218). */
gra_status gra_wavelet_info(gra_ctx *ctx, int64_t *n_dy, int64_t *nnz);

/* Shape of the context's G: Nd = n_comp * N_obs rows, Nm columns. */
gra_status gra_get_dims(const gra_ctx *ctx, int64_t *Nd, int64_t *Nm, int64_t *N_obs,
                        uint32_t *field_flag, int *store);

/*
 * Replaces Fwd::get_G (Fwd.h:46): copies G to host as float64, column-major
 * (Eigen::MatrixXd layout, leading dimension Nd) when col_major != 0, else
 * row-major.  `out` holds Nd*Nm doubles.
 */
gra_status gra_copy_G(gra_ctx *ctx, double *out, int col_major);

/* Replaces Fwd::G_vec_mul (Fwd.cpp:114-147): product[Nd] = G * vec[Nm]. */
gra_status gra_G_vec_mul(gra_ctx *ctx, const double *vec, double *product);
/* Replaces Fwd::GT_vec_mul (Fwd.cpp:80-113): product[Nm] = G^T * vec[Nd]. */
gra_status gra_GT_vec_mul(gra_ctx *ctx, const double *vec, double *product);
/*
 * No counterpart in the reference: G x for callers that cannot allocate
 * device memory (code under hipGraph stream capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same kernels and launch shape as
 * gra_G_vec_mul, but device pointers in and out (d_x: Nm doubles, d_y: Nd
 * doubles) and no allocation of any kind: x is staged and the partial sums
 * kept in a fixed 32 MiB per-GPU scratch array inside the library, so G must
 * satisfy ld + splits * Nd <= 4,194,304 (ld = Nm rounded up to 64; splits as
 * gra_tune_get(GRA_TUNE_N_SPLITS) reports), else GRA_BAD_ARG.  Dense G only.
 * Asynchronous on the context's stream (gra_sync waits for d_y).  Every call
 * on one GPU uses that GPU's single scratch array: calls that overlap on one
 * GPU (contexts on different streams, or host threads) race and must be
 * serialised by the caller; calls on one context are ordered by its stream.
 */
gra_status gra_G_vec_mul_noalloc(gra_ctx *ctx, const double *d_x, double *d_y);

/*
 * k right-hand sides at once on the matrix cores: Y = G X (X: Nm x k, Y:
 * Nd x k) and Z = G^T R (R: Nd x k, Z: Nm x k), all column-major (column j
 * at base + j * rows).  The batched form of Fwd::G_vec_mul / GT_vec_mul
 * (Fwd.cpp:114-147, 80-113), which the reference calls one vector at a time;
 * no counterpart for k > 1.  FP64-stored G: FP64 MFMA (FP64 throughout);
 * FP32 / BF16-stored G: FP32 MFMA with the vectors rounded to FP32 and the
 * FP32 partial sums added into FP64 every 16-32 products.  1 <= k <= 1024;
 * not for the wavelet-compressed G.  Bit-for-bit repeatable.
 */
gra_status gra_G_mat_mul(gra_ctx *ctx, const double *X, int64_t k, double *Y);
gra_status gra_GT_mat_mul(gra_ctx *ctx, const double *R, int64_t k, double *Z);

/*
 * Replaces AdaptiveInversion::expand_G (AdaptiveInversion.cpp:23-106): the
 * cells at the n_split old columns split_cols[] (strictly increasing) are
 * refined, and each parent's column is replaced in place by the columns of
 * its 8 children, child_prisms[(8*s + j)*6 .. +5] (x0,x1,y0,y1,z0,z1) for the
 * s-th split cell in the order of Cell::child_cells (Mesh.cpp:1581-1598:
 * j = 4i + 2j' + k over x, y, z halves), so the later columns shift by 7 per
 * earlier split -- the column order of the refined leaf list
 * (Mesh.cpp:1934-1941).  Only the child columns are computed (on the GPU, in
 * the context's storage type); Nm grows by 7 * n_split.  Regularisation
 * operators and the resident Gauss-Newton state refer to the old Nm and must
 * be set again.
 */
gra_status gra_expand_cols(gra_ctx *ctx, const int64_t *split_cols, int64_t n_split,
                           const double *child_prisms);

/*
 * Column 2-norms of the stored G: norms[j] = ||G(:, j)||_2, j < Nm.  Replaces
 * the G.col(i).norm() reads of InversionBase::set_depth_weighting(beta, 1)
 * (InversionBase.cpp:519-526), whose weight is sqrt(norm / Nd).
 */
gra_status gra_col_norms(gra_ctx *ctx, double *norms);

/*
 * Replaces the matrix-free forward loop of compute.cpp:246-260:
 * out[i*n_comp + k] = sum_j field_k(point i, prism j, rho_j), no G stored.
 */
gra_status gra_forward(gra_ctx *ctx, const double *prisms, const double *rho, int64_t Nm,
                       const double *pts, int64_t N_obs, uint32_t field_flag, double *out);

/*
 * The same forward for a rectilinear mesh: cells are the tensor product of
 * the node coordinates xs[nx+1], ys[ny+1], zs[nz+1] (strictly increasing; z
 * down), rho[(i*ny + j)*nz + k] the density of cell (i, j, k) -- the cell
 * order of Mesh::generate_regular_mesh (Mesh.cpp:828-846).  The field is
 * summed over shared mesh nodes (one corner-function evaluation per node
 * instead of the per-prism face/edge integrals of GravFormula.cpp), which
 * agrees with gra_forward to ~1e-11 of the largest field value but loses
 * relative accuracy far outside the mesh (corner-sum cancellation; ~1e-6
 * relative at 5 mesh widths).  Points must lie strictly above the mesh
 * (z < zs[0]), else GRA_BAD_ARG.  No counterpart in the reference, whose
 * compute tool takes an explicit prism list.
 */
gra_status gra_forward_rectilinear(gra_ctx *ctx, const double *xs, int64_t nx, const double *ys,
                                   int64_t ny, const double *zs, int64_t nz, const double *rho,
                                   const double *pts, int64_t N_obs, uint32_t field_flag,
                                   double *out);

/*
 * Several gra_forward_rectilinear requests in one launch (no counterpart in
 * the reference): the node weights of each item, then one workgroup per
 * observation point of every item.  Each item's output is bit-identical to
 * its own gra_forward_rectilinear call; the same argument rules per item.
 */
typedef struct gra_fwd_rect_item {
    const double *xs; int64_t nx;
    const double *ys; int64_t ny;
    const double *zs; int64_t nz;
    const double *rho;       /* [nx*ny*nz] */
    const double *pts;       /* [N_obs*3]  */
    int64_t N_obs;
    uint32_t field_flag;
    double *out;             /* [N_obs*n_comp] */
} gra_fwd_rect_item;
gra_status gra_forward_rectilinear_batch(gra_ctx *ctx, int64_t n, const gra_fwd_rect_item *items);

/* ------------------------------------------------------------- inversion */
/*
 * Uploads the regularisation operators used by gra_solve_cg: Ws, Wx, Wy, Wz
 * (W* = a*S*V*D*Z, GaussNewtonInversion.cpp:711-718) and, for the
 * cross-gradient term, Tx0, Ty0, Tz0 (InversionBase::set_Tmatrix,
 * InversionBase.cpp:468-495); pass NULL for the T's when the cross-gradient
 * constraint is off.  All are Nm x Nm with Nm equal to G's column count.
 */
gra_status gra_set_regularization(gra_ctx *ctx, const gra_csr *Ws, const gra_csr *Wx,
                                  const gra_csr *Wy, const gra_csr *Wz, const gra_csr *Tx0,
                                  const gra_csr *Ty0, const gra_csr *Tz0);

/* Diagonal operators are passed as vectors; NULL WL_* / S_crg = identity. */
typedef struct gra_cg_args {
    double tol;
    int maxit;
    double lambda;
    const double *mk;      /* [Nm] current model                 */
    const double *mref;    /* [Nm] reference model (m0)          */
    const double *Wd;      /* [Nd] data weights (diagonal)       */
    const double *dobs;    /* [Nd] observed data                 */
    const double *P;       /* [Nm] barrier Jacobian (diagonal)   */
    const double *WL_s, *WL_x, *WL_y, *WL_z, *WL_tx, *WL_ty, *WL_tz;  /* [Nm] */
    int use_cross_gradient;
    double a_crg;
    const double *S_crg;   /* [Nm] */
    /* 1: reproduce the reference operator exactly, including its asymmetric
     *    weighting (t_s omits WL_s, GN.cpp:1573; cross-gradient rows of t use
     *    WL_z/WL_x/WL_y, :1582-1584; initial residual's Ty term uses WL_y,
     *    :1508).  0: consistent operator (A^T of the same A).            */
    int quirk_mode;
    /* 0: the reference's iteration, two passes over G (r = A^T s).
     * 1: single-read iteration: one fused pass computes G Pp and
     *    G^T Wd^2 G Pp together and r is updated as r + alpha A^T t (equal in
     *    exact arithmetic, different rounding).  Falls back to 0 when the
     *    fused grid cannot be co-resident on the device; when one of the
     *    fused product's bounded waits fails, the solve is rerun with 0
     *    from the same inputs (GRA_TUNE_FUSED_FALLBACKS counts these).  */
    int fused;
} gra_cg_args;

/*
 * Replaces GaussNewtonInversion::solve_cg
 * (src/Inversion/GaussNewtonInversion.h:70-93, .cpp:1417-1698): CGLS on
 * [Wd G P; sqrt(l) WL_s Ws P; sqrt(l) WL_z Wz P; sqrt(l) WL_x Wx P;
 *  sqrt(l) WL_y Wy P; (WL_t* a S T* P)] dx ~= b, in FP64 on the device with
 * the reference's two passes over G per iteration.  Writes dx[Nm], *err
 * (= ||r||^2/||A^T b||^2, untouched when maxit == 0) and *iterations.
 */
gra_status gra_solve_cg(gra_ctx *ctx, const gra_cg_args *args, double *dx, double *err,
                        int *iterations);

/*
 * Mixed-precision CG (no counterpart in the reference, whose solve_cg runs
 * in FP64: GaussNewtonInversion.cpp:1556-1691).  Keeps a copy of the
 * context's dense FP64 G in `inner` precision (GRA_F32 or GRA_BF16, rounded
 * as gra_assemble stores them) beside it; from then on every gra_solve_cg
 * and gra_gn_step on this context runs the CG iteration's G products on the
 * copy (fused or two-pass as requested) and, every `refresh_every`
 * iterations and after the last one, replaces the iteration's residuals by
 * their FP64 values -- s = A (P dx) - b and r = A^T s with the FP64 G -- so
 * that the solve converges to the FP64 least-squares solution.  The
 * tolerance is tested on the refreshed residuals only, and the returned err
 * is the FP64 one.  Products outside the iteration (b, the initial
 * residual, misfits) use the FP64 G.  GRA_F64 switches it off and frees the
 * copy; a new G (assembly, expansion) invalidates it (solves then run in
 * FP64 until it is made again).
 */
gra_status gra_mixed_precision(gra_ctx *ctx, gra_dtype inner, int refresh_every);

/* P_ii = (M-m)(m-m_min)/(M-m_min)  (GaussNewtonInversion.cpp:856-861). */
gra_status gra_barrier_P(gra_ctx *ctx, const double *m, const double *m_min,
                         const double *m_max, double *P);

/* Log-barrier model update with the reference's overflow fallbacks
 * (GaussNewtonInversion.cpp:877-911); m_trial is read and written. */
gra_status gra_barrier_update(gra_ctx *ctx, const double *m, const double *dx,
                              const double *m_min, const double *m_max, double *m_trial);

/* misfit = sqrt(||Wd (G m - dobs)||^2 / Nd)  (GaussNewtonInversion.cpp:913-916). */
gra_status gra_misfit(gra_ctx *ctx, const double *m, const double *Wd, const double *dobs,
                      double *misfit);

/* IRLS weights WL_ii = ((W m)_i^2 + eps2)^(-(2-p)/4) for a CSR W
 * (GaussNewtonInversion.cpp:962-1022); p is the integer Lp exponent. */
gra_status gra_irls_weights(gra_ctx *ctx, const gra_csr *W, const double *m, int p,
                            double eps2, double *WL);

/* ----------------------------------------- resident Gauss-Newton driver */
/*
 * A Gauss-Newton loop whose model, weights and CG state stay in device
 * memory (the per-outer-iteration body of invert_with_own_CG,
 * GaussNewtonInversion.cpp:815-1048, with the lambda schedule chosen by the
 * caller and every trial accepted).  gra_gn_setup uploads the problem (the
 * regularisation operators come from gra_set_regularization);
 * gra_gn_step runs: P from m; dx = solve_cg(...) (reference operator,
 * quirk_mode 1); m_trial = barrier update; misfit(m_trial); m = m_trial; and,
 * for an Lp norm (lp != 2), the IRLS weights WL_s = ((Ws(m-mref))^2 +
 * eps2)^(-(2-lp)/4), WL_x/y/z likewise from Wx/Wy/Wz m, used by the next step
 * (GaussNewtonInversion.cpp:962-1022 with the trial model accepted).  WL_* start
 * as the identity.
 */
typedef struct gra_gn_setup {
    const double *dobs;    /* [Nd] */
    const double *Wd;      /* [Nd] */
    const double *m_min;   /* [Nm] */
    const double *m_max;   /* [Nm] */
    const double *mref;    /* [Nm] */
    const double *m_ini;   /* [Nm] */
    int lp;                /* Lp exponent of the model norms (2 = L2, no IRLS) */
    double eps2;           /* IRLS epsilon^2 (GaussNewtonInversion.h:57)      */
    int fused;             /* CG iteration as gra_cg_args.fused               */
    int use_cross_gradient;/* 1: the cross-gradient rows with the T operators of
                              gra_set_regularization (GN.cpp:1477-1484)       */
    double a_crg;          /* cross-gradient weight a (GaussNewtonInversion.h) */
    const double *S_crg;   /* [Nm] cross-gradient mask (update_S_crg), or null */
} gra_gn_setup;

gra_status gra_gn_setup_problem(gra_ctx *ctx, const gra_gn_setup *setup);
gra_status gra_gn_step(gra_ctx *ctx, double lambda, double cg_tol, int cg_maxit,
                       double *misfit, double *cg_err, int *cg_iterations);
/*
 * One gra_gn_step on each of n contexts (independent inversions of uneven
 * sizes on one device, e.g. the lambda trials or instances of a batch;
 * lambdas[b] for context b): every step is queued on its context's stream
 * before any is waited for, so the instances' kernels overlap and no host
 * synchronisation separates them; steps whose CG uses the single-read
 * product run one after another (each needs the whole device).  Each
 * context's results are bit-identical to its own gra_gn_step.  Outputs per
 * context (may be null).  Returns the first failure (its context's
 * gra_last_error says why).  No counterpart in the reference (its
 * instances run one after another).
 */
gra_status gra_gn_step_batch(gra_ctx *const *ctxs, int64_t n, const double *lambdas,
                             double cg_tol, int cg_maxit, double *misfits, double *cg_errs,
                             int *cg_iterations);
gra_status gra_gn_get_model(gra_ctx *ctx, double *m);

/* --------------------------------------------------------------- tuning */
/*
 * Launch shapes of the G products (no counterpart in the reference, whose
 * Eigen GEMV has none): used by the autotuner sweep.  0 restores the
 * automatic choice for the context's current G (nontemporal: -1).  Values
 * outside the listed sets are rejected.  Results do not depend on these
 * settings except through the (deterministic) summation order of splits and
 * chunks.
 */
#define GRA_TUNE_N_ROWS 0       /* rows per wave of G*x: 2, 4, 8           */
#define GRA_TUNE_N_UNROLL 1     /* column steps per iteration: 1, 2        */
#define GRA_TUNE_N_SPLITS 2     /* column splits of G*x: 1..256            */
#define GRA_TUNE_T_CHUNKS 3     /* row chunks of G^T*w: 1..1024            */
#define GRA_TUNE_T_UNROLL 4     /* rows per iteration of G^T*w: 4, 8       */
#define GRA_TUNE_NONTEMPORAL 5  /* non-temporal loads of G, mask: 1 G*x, 2 G^T*w */
#define GRA_TUNE_FUSED_WGS 6    /* workgroups of the fused CG product: 1..65536,
                                   capped at what can be co-resident      */
#define GRA_TUNE_FUSED_DEPTH 7  /* fused: (lag L, prefetch PF) variant 1..3 =
                                   (2,1) (2,2) (3,2); 0 = (2,1)            */
#define GRA_TUNE_N_BLOCK 8      /* G*x rows owned by a wave (0) or shared by the
                                   4 waves of a workgroup (1); -1 = auto */
#define GRA_TUNE_NODE_TILE 9    /* gra_assemble_rectilinear tile side in cell
                                   columns, 1..32; 0 = auto (LDS budget) */
#define GRA_TUNE_FUSED_FAIL 10  /* test hook: the fused product's wait on this
                                   row block fails (-1 = none), forcing the
                                   two-pass rerun of the solve            */
#define GRA_TUNE_FUSED_MATH 12  /* fused product arithmetic: 0 = default (FP32
                                   products with FP64 sums for BF16 storage,
                                   FP64 otherwise), 1 = FP64, 2 = FP32
                                   products (BF16 / FP32 storage)         */
#define GRA_TUNE_LAST_REFRESHES 13 /* get: FP64 residual refreshes of the last
                                   mixed-precision solve                    */
#define GRA_TUNE_CG_GRAPH 14    /* two-pass CG iterations captured per hipGraph
                                   (after the first iteration, the rest are
                                   graph replays); 0 = every kernel launched
                                   on its own.  Not used by the fused or the
                                   mixed-precision iteration.             */
#define GRA_TUNE_FUSED_MIN_BLOCKS 15 /* fused CG only where each row group of
                                   the single-read grid holds at least this
                                   many row blocks (else two-pass); 0 = any */
#define GRA_TUNE_FUSED_FALLBACKS 11 /* get: fused solves rerun two-pass on this
                                   context (it then stays two-pass); set 0:
                                   clear and allow the fused iteration again */
gra_status gra_tune_set(gra_ctx *ctx, int param, int64_t value);
gra_status gra_tune_get(gra_ctx *ctx, int param, int64_t *value);
/*
 * Autotuner for the context's current G (no counterpart in the reference):
 * times `reps` (<= 0: automatic) alternating G x / G^T w launches issued back
 * to back, as the CG iteration issues them, for candidate shapes (coordinate
 * descent: G^T w row chunks and unroll, then G x rows / unroll / splits, then
 * the non-temporal mask, judged on the pair) and keeps the fastest as the
 * context's overrides (gra_tune_get reads them; gra_tune_set(..., 0) resets).
 * Writes the winners' average launch times in ms.  Overwrites the product
 * scratch buffers, so a Gauss-Newton step after it recomputes G m_k.
 */
gra_status gra_autotune(gra_ctx *ctx, int reps, double *ms_n, double *ms_t);

/* ---------------------------------------------------------- measurement */
/*
 * Average device time (ms, from GPU events on the context's stream) of the
 * launches of kernel class `which` since the last gra_timing_reset:
 * 0 = GEMV G*x, 1 = GEMV G^T*x, 2 = G assembly, 3 = one full CG iteration,
 * 4 = fused single-read product (G Pp and G^T Wd^2 G Pp), 5 = the
 * matrix-core multi-right-hand-side product (gra_G_mat_mul / gra_GT_mat_mul),
 * 6 = the FP64 G products of the mixed-precision residual refreshes,
 * 7 = one replay of the captured CG iterations (GRA_TUNE_CG_GRAPH),
 * 8 = the affine single-read pass of a Gauss-Newton step (G m_trial for the
 * misfit and G^T Wd^2 (G m_trial - dobs) for the next solve, one read of G).
 * Timing events are recorded only while enabled: `on` = 0 disables; n >= 1
 * times every launch of these classes, except inside the CG loop, where only
 * iterations 0, n, 2n, ... are timed (each event is a queue packet of its
 * own and slows the launches around it, so a stride keeps most iterations
 * free of them).
 */
gra_status gra_timing_enable(gra_ctx *ctx, int on);
gra_status gra_timing_reset(gra_ctx *ctx);
gra_status gra_timing_get(gra_ctx *ctx, int which, double *avg_ms, int64_t *count);

#ifdef __cplusplus
}
#endif

#endif /* GRAENG_H */
