/*
 * CPU oracle for the Gauss-Newton / CGLS inner solver and its helpers.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
 * this file; tests/, the smoke check and bench.py's CPU-baseline leg use it
 * as the checker.
 *
 * Restates, single-threaded and with sequential left-to-right sums:
 *   - GaussNewtonInversion::solve_cg   src/Inversion/GaussNewtonInversion.cpp:1417-1698
 *     (CGLS on the stacked system, including the reference's operator
 *     asymmetries when quirks != 0:
 *       (i)   t_s = sqrt(lambda) Ws Pp omits WL_s            (:1573)
 *       (ii)  cross-gradient t rows use WL_z/WL_x/WL_y       (:1582-1584)
 *       (iii) the initial residual's Ty term uses WL_y       (:1508)
 *     with quirks == 0 the operator is the consistent one: WL_s in t_s,
 *     WL_t* in the cross-gradient rows and in the initial residual.)
 *   - barrier parameter P and model update    GaussNewtonInversion.cpp:856-861, 877-911
 *   - data misfit                             GaussNewtonInversion.cpp:913-916
 *   - IRLS weights                            GaussNewtonInversion.cpp:962-1022
 *   - dense products G*v, G^T*v               Fwd.cpp:111, 145 (G row-major here)
 *
 * Sparse operators are CSR (row pointers int64, column indices int32, values
 * double).  Diagonal operators (Wd, P, WL_*, S_crg) are vectors; a NULL
 * WL_* or S_crg means the identity.  Products such as
 * "sqrt(lambda) * P * Ws^T * WL_s^T * v" are formed entry by entry as
 * ((sqrt(lambda)*P_i)*Ws_ji)*WL_j, the association order Eigen's chained
 * sparse products use; the transposes are taken from the caller-supplied
 * transposed CSR (columns ascending within each row).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int64_t n;               /* rows (== cols, Nm) */
    const int64_t *rp;       /* n+1 */
    const int32_t *ci;
    const double *v;
} ora_csr;

/* y = G x, G row-major nd x nm (Fwd::G_vec_mul, Fwd.cpp:145) */
extern int ora_threads;   /* prism_oracle.c */

void ora_gemv_n(const double *G, int64_t nd, int64_t nm, const double *x, double *y)
{
#pragma omp parallel for num_threads(ora_threads) schedule(static)
    for (int64_t i = 0; i < nd; i++) {
        const double *row = G + i * nm;
        double acc = 0.0;
        for (int64_t j = 0; j < nm; j++) acc += row[j] * x[j];
        y[i] = acc;
    }
}

/* y = G^T x (Fwd::GT_vec_mul, Fwd.cpp:111) */
void ora_gemv_t(const double *G, int64_t nd, int64_t nm, const double *x, double *y)
{
    /* column blocks per thread; each column accumulates rows in order */
    const int64_t nb = (nm + 511) / 512;
#pragma omp parallel for num_threads(ora_threads) schedule(static)
    for (int64_t b = 0; b < nb; b++) {
        const int64_t j0 = b * 512, j1 = j0 + 512 < nm ? j0 + 512 : nm;
        for (int64_t j = j0; j < j1; j++) y[j] = 0.0;
        for (int64_t i = 0; i < nd; i++) {
            const double *row = G + i * nm;
            double xi = x[i];
            for (int64_t j = j0; j < j1; j++) y[j] += row[j] * xi;
        }
    }
}

static inline double dg(const double *d, int64_t i) { return d ? d[i] : 1.0; }

/* out_i = sum_j ((pre * L_i) * A_ij) * R_j * x_j  with L (row diag) and R
 * (col diag) optional.  Used for both A and A^T (pass the transposed CSR). */
static void spmv_scaled(const ora_csr *A, double pre, const double *L, const double *R,
                        const double *x, double *out)
{
    for (int64_t i = 0; i < A->n; i++) {
        double acc = 0.0;
        double li = pre * dg(L, i);
        for (int64_t k = A->rp[i]; k < A->rp[i + 1]; k++) {
            int64_t j = A->ci[k];
            double coef = li * A->v[k];
            if (R) coef = coef * R[j];
            acc += coef * x[j];
        }
        out[i] = acc;
    }
}

/* Tz = a_crg * S_crg * Tz0 folded into the value array (GN.cpp:1445-1450) */
static double *crg_scale(const ora_csr *T0, double a_crg, const double *S)
{
    int64_t nnz = T0->rp[T0->n];
    double *v = (double *)malloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
    for (int64_t i = 0; i < T0->n; i++)
        for (int64_t k = T0->rp[i]; k < T0->rp[i + 1]; k++)
            v[k] = (a_crg * dg(S, i)) * T0->v[k];
    return v;
}

typedef struct {
    const double *G; int64_t nd, nm;
    double sl;                          /* sqrt(lambda) */
    const double *Wd, *P;
    ora_csr Ws, Wx, Wy, Wz, WsT, WxT, WyT, WzT;
    int crg;
    ora_csr Tz, Tx, Ty, TzT, TxT, TyT;  /* already scaled by a_crg*S_crg */
    const double *WLs, *WLx, *WLy, *WLz, *WLtx, *WLty, *WLtz;
    double *tmp_nd, *tmp_nm, *tmp_nm2;
} cg_ops;

/* r = A^T v (GaussNewtonInversion.cpp:1484-1520 and :1636-1673).  `wl_ty`
 * is the right-diagonal used for the Ty row block. */
static void apply_AT(cg_ops *o, const double *v, double *r, const double *wl_ty)
{
    int64_t nd = o->nd, nm = o->nm;
    for (int64_t i = 0; i < nd; i++) o->tmp_nd[i] = dg(o->Wd, i) * v[i];
    ora_gemv_t(o->G, nd, nm, o->tmp_nd, o->tmp_nm);
    double *acc = o->tmp_nm2;
    for (int64_t k = 0; k < nm; k++) r[k] = o->P[k] * o->tmp_nm[k];          /* vec1 */
    spmv_scaled(&o->WsT, o->sl, o->P, o->WLs, v + nd, acc);                     /* vec2 */
    for (int64_t k = 0; k < nm; k++) r[k] = r[k] + acc[k];
    spmv_scaled(&o->WzT, o->sl, o->P, o->WLz, v + nd + nm, acc);                /* vec3 */
    for (int64_t k = 0; k < nm; k++) r[k] = r[k] + acc[k];
    spmv_scaled(&o->WxT, o->sl, o->P, o->WLx, v + nd + 2 * nm, acc);            /* vec4 */
    for (int64_t k = 0; k < nm; k++) r[k] = r[k] + acc[k];
    spmv_scaled(&o->WyT, o->sl, o->P, o->WLy, v + nd + 3 * nm, acc);            /* vec5 */
    for (int64_t k = 0; k < nm; k++) r[k] = r[k] + acc[k];
    if (o->crg) {
        spmv_scaled(&o->TzT, 1.0, o->P, o->WLtz, v + nd + 4 * nm, acc);        /* vec6 */
        for (int64_t k = 0; k < nm; k++) r[k] = r[k] + acc[k];
        spmv_scaled(&o->TxT, 1.0, o->P, o->WLtx, v + nd + 5 * nm, acc);        /* vec7 */
        for (int64_t k = 0; k < nm; k++) r[k] = r[k] + acc[k];
        spmv_scaled(&o->TyT, 1.0, o->P, wl_ty, v + nd + 6 * nm, acc);          /* vec8 */
        for (int64_t k = 0; k < nm; k++) r[k] = r[k] + acc[k];
    }
}

static double sumsq(const double *v, int64_t n)
{
    double s = 0.0;
    for (int64_t k = 0; k < n; k++) s += v[k] * v[k];
    return s;
}

/*
 * solve_cg.  Arguments mirror GaussNewtonInversion.h:70-93; the G matrix is
 * passed explicitly (row-major, nd x nm), *T0 and the transposes of every
 * CSR operator are supplied by the caller.  Returns dx in `dx` (length nm),
 * writes *err and *iterations as the reference does (err is left untouched
 * when maxit == 0, as in the reference).
 *
 * recurrence = 0 is the reference's iteration (r = A^T s after every update of
 * s, GN.cpp:1637-1672).  recurrence = 1 updates r <- r + alpha A^T t instead,
 * equal in exact arithmetic because s <- s + alpha t and A^T is linear; in the
 * first iteration the base is A^T s0 with the loop's weights (WL_ty in the
 * last cross-gradient term), since the initial residual of the reference uses
 * WL_y there (GN.cpp:1508).  This is the single-read iteration's algorithm.
 */
int ora_solve_cg(double tol, int maxit, double *err, int *iterations, double lambda,
                 const double *mk, const double *mref, const double *Wd, const double *dobs,
                 const double *P,
                 const ora_csr *Ws, const ora_csr *Wx, const ora_csr *Wy, const ora_csr *Wz,
                 const ora_csr *WsT, const ora_csr *WxT, const ora_csr *WyT, const ora_csr *WzT,
                 int use_crg, double a_crg, const double *S_crg,
                 const ora_csr *Tx0, const ora_csr *Ty0, const ora_csr *Tz0,
                 const ora_csr *Tx0T, const ora_csr *Ty0T, const ora_csr *Tz0T,
                 const double *WLs, const double *WLx, const double *WLy, const double *WLz,
                 const double *WLtx, const double *WLty, const double *WLtz,
                 int quirks, int recurrence, const double *G, int64_t nd, int64_t nm,
                 double *dx)
{
    cg_ops o;
    memset(&o, 0, sizeof(o));
    o.G = G; o.nd = nd; o.nm = nm; o.sl = sqrt(lambda); o.Wd = Wd; o.P = P;
    o.Ws = *Ws; o.Wx = *Wx; o.Wy = *Wy; o.Wz = *Wz;
    o.WsT = *WsT; o.WxT = *WxT; o.WyT = *WyT; o.WzT = *WzT;
    o.WLs = WLs; o.WLx = WLx; o.WLy = WLy; o.WLz = WLz;
    o.WLtx = WLtx; o.WLty = WLty; o.WLtz = WLtz;
    o.crg = use_crg;
    double *tv[6] = {0};
    if (use_crg) {
        /* Transposes of S*T0 have the row scaling on their columns. */
        tv[0] = crg_scale(Tz0, a_crg, S_crg); tv[1] = crg_scale(Tx0, a_crg, S_crg);
        tv[2] = crg_scale(Ty0, a_crg, S_crg);
        o.Tz = *Tz0; o.Tz.v = tv[0]; o.Tx = *Tx0; o.Tx.v = tv[1]; o.Ty = *Ty0; o.Ty.v = tv[2];
        const ora_csr *TT[3] = {Tz0T, Tx0T, Ty0T};
        for (int q = 0; q < 3; q++) {
            int64_t nnz = TT[q]->rp[TT[q]->n];
            tv[3 + q] = (double *)malloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
            for (int64_t i = 0; i < TT[q]->n; i++)
                for (int64_t k = TT[q]->rp[i]; k < TT[q]->rp[i + 1]; k++)
                    tv[3 + q][k] = (a_crg * dg(S_crg, TT[q]->ci[k])) * TT[q]->v[k];
        }
        o.TzT = *Tz0T; o.TzT.v = tv[3]; o.TxT = *Tx0T; o.TxT.v = tv[4];
        o.TyT = *Ty0T; o.TyT.v = tv[5];
    }
    int64_t nrow = use_crg ? nd + 7 * nm : nd + 4 * nm;
    double *s = (double *)calloc((size_t)nrow, sizeof(double));
    double *b = (double *)calloc((size_t)nrow, sizeof(double));
    double *t = (double *)calloc((size_t)nrow, sizeof(double));
    double *r = (double *)calloc((size_t)nm, sizeof(double));
    double *p = (double *)calloc((size_t)nm, sizeof(double));
    double *Pp = (double *)calloc((size_t)nm, sizeof(double));
    double *ATb = (double *)calloc((size_t)nm, sizeof(double));
    double *w = (double *)calloc((size_t)nm, sizeof(double));
    o.tmp_nd = (double *)calloc((size_t)nd, sizeof(double));
    o.tmp_nm = (double *)calloc((size_t)nm, sizeof(double));
    o.tmp_nm2 = (double *)calloc((size_t)nm, sizeof(double));
    double *gv = (double *)calloc((size_t)nd, sizeof(double));
    for (int64_t k = 0; k < nm; k++) dx[k] = 0.0;

    /* b (GN.cpp:1459-1471) */
    ora_gemv_n(G, nd, nm, mk, gv);
    for (int64_t i = 0; i < nd; i++) b[i] = dg(Wd, i) * (dobs[i] - gv[i]);
    for (int64_t k = 0; k < nm; k++) w[k] = mk[k] - mref[k];
    spmv_scaled(Ws, -o.sl, WLs, NULL, w, b + nd);
    spmv_scaled(Wz, -o.sl, WLz, NULL, mk, b + nd + nm);
    spmv_scaled(Wx, -o.sl, WLx, NULL, mk, b + nd + 2 * nm);
    spmv_scaled(Wy, -o.sl, WLy, NULL, mk, b + nd + 3 * nm);
    if (use_crg) {
        spmv_scaled(&o.Tz, -1.0, WLtz, NULL, mk, b + nd + 4 * nm);
        spmv_scaled(&o.Tx, -1.0, WLtx, NULL, mk, b + nd + 5 * nm);
        spmv_scaled(&o.Ty, -1.0, WLty, NULL, mk, b + nd + 6 * nm);
    }
    for (int64_t k = 0; k < nrow; k++) s[k] = -b[k];

    apply_AT(&o, s, r, quirks ? WLy : WLty);
    /* ATb (GN.cpp:1531-1555), accumulated term by term */
    {
        double *one = (double *)calloc((size_t)nm, sizeof(double));
        for (int64_t i = 0; i < nd; i++) o.tmp_nd[i] = dg(Wd, i) * b[i];
        ora_gemv_t(G, nd, nm, o.tmp_nd, o.tmp_nm);
        for (int64_t k = 0; k < nm; k++) ATb[k] = 0.0 + P[k] * o.tmp_nm[k];
        const ora_csr *AT[4] = {WsT, WzT, WxT, WyT};
        const double *WL[4] = {WLs, WLz, WLx, WLy};
        for (int q = 0; q < 4; q++) {
            spmv_scaled(AT[q], o.sl, P, WL[q], b + nd + q * nm, one);
            for (int64_t k = 0; k < nm; k++) ATb[k] += one[k];
        }
        if (use_crg) {
            const ora_csr *TT[3] = {&o.TzT, &o.TxT, &o.TyT};
            const double *WT[3] = {WLtz, WLtx, WLty};
            for (int q = 0; q < 3; q++) {
                spmv_scaled(TT[q], 1.0, P, WT[q], b + nd + (4 + q) * nm, one);
                for (int64_t k = 0; k < nm; k++) ATb[k] += one[k];
            }
        }
        free(one);
    }
    double ATb2 = sumsq(ATb, nm);
    double *rloop = NULL, *q = NULL;
    if (recurrence) {
        rloop = (double *)calloc((size_t)nm, sizeof(double));
        q = (double *)calloc((size_t)nm, sizeof(double));
        if (quirks && use_crg) apply_AT(&o, s, rloop, WLty);
        else memcpy(rloop, r, sizeof(double) * (size_t)nm);
    }
    double beta = 0.0, r2 = 0.0, t2, rlast2, alpha;
    int i;
    for (i = 0; i < maxit; i++) {
        for (int64_t k = 0; k < nm; k++) p[k] = -r[k] + beta * p[k];
        for (int64_t k = 0; k < nm; k++) Pp[k] = P[k] * p[k];
        ora_gemv_n(G, nd, nm, Pp, gv);
        for (int64_t q = 0; q < nd; q++) t[q] = dg(Wd, q) * gv[q];
        spmv_scaled(Ws, o.sl, quirks ? NULL : WLs, NULL, Pp, t + nd);
        spmv_scaled(Wz, o.sl, WLz, NULL, Pp, t + nd + nm);
        spmv_scaled(Wx, o.sl, WLx, NULL, Pp, t + nd + 2 * nm);
        spmv_scaled(Wy, o.sl, WLy, NULL, Pp, t + nd + 3 * nm);
        if (use_crg) {
            spmv_scaled(&o.Tz, 1.0, quirks ? WLz : WLtz, NULL, Pp, t + nd + 4 * nm);
            spmv_scaled(&o.Tx, 1.0, quirks ? WLx : WLtx, NULL, Pp, t + nd + 5 * nm);
            spmv_scaled(&o.Ty, 1.0, quirks ? WLy : WLty, NULL, Pp, t + nd + 6 * nm);
        }
        r2 = sumsq(r, nm);
        t2 = sumsq(t, nrow);
        alpha = r2 / t2;
        for (int64_t k = 0; k < nrow; k++) s[k] += alpha * t[k];
        for (int64_t k = 0; k < nm; k++) dx[k] += alpha * p[k];
        if (recurrence) {
            apply_AT(&o, t, q, WLty);
            const double *base = i == 0 ? rloop : r;
            for (int64_t k = 0; k < nm; k++) r[k] = base[k] + alpha * q[k];
        } else {
            apply_AT(&o, s, r, WLty);
        }
        rlast2 = r2;
        r2 = sumsq(r, nm);
        beta = r2 / rlast2;
        *err = r2 / ATb2;
        if (*err < tol) {
            *iterations = i + 1;
            break;
        }
    }
    if (i == maxit) *iterations = maxit;

    free(s); free(b); free(t); free(r); free(p); free(Pp); free(ATb); free(w);
    free(o.tmp_nd); free(o.tmp_nm); free(o.tmp_nm2); free(gv); free(rloop); free(q);
    for (int q = 0; q < 6; q++) free(tv[q]);
    return 0;
}

/* P_ii = (M - m)(m - m_min)/(M - m_min)   (GaussNewtonInversion.cpp:856-861) */
void ora_barrier_P(const double *m, const double *mmin, const double *mmax, int64_t nm, double *P)
{
    for (int64_t k = 0; k < nm; k++)
        P[k] = (mmax[k] - m[k]) * (m[k] - mmin[k]) / (mmax[k] - mmin[k]);
}

/*
 * Log-barrier model update with the reference's overflow / underflow
 * fallbacks, sequential in cell order (GaussNewtonInversion.cpp:877-911).
 * m_trial holds the previous trial model on entry and the new one on exit.
 */
void ora_barrier_update(const double *m, const double *dx, const double *mmin,
                        const double *mmax, int64_t nm, double *m_trial)
{
    for (int64_t id = 0; id < nm; id++) {
        double e = exp(dx[id]);
        double t2 = mmin[id] * (mmax[id] - m[id]) + mmax[id] * (m[id] - mmin[id]) * e;
        double t3 = (mmax[id] - m[id]) + (m[id] - mmin[id]) * e;
        if (isinf(e) || isinf(t2)) {
            m_trial[id] = (mmax[id] - 1e-6 + m_trial[id]) / 2.0;
            if (id == 0) m_trial[id] = 0.5 * (m_trial[id] + m_trial[id + 1]);
            else m_trial[id] = 0.5 * (m_trial[id] + m_trial[id - 1]);
        } else if (fabs(t3) < 1e-10) {
            m_trial[id] = (mmin[id] + 1e-6 + m_trial[id]) / 2.0;
            if (id == 0) m_trial[id] = 0.5 * (m_trial[id] + m_trial[id + 1]);
            else m_trial[id] = 0.5 * (m_trial[id] + m_trial[id - 1]);
        } else {
            m_trial[id] = t2 / t3;
        }
    }
}

/* misfit = sqrt(||Wd (G m - dobs)||^2 / Nd)   (GaussNewtonInversion.cpp:913-916) */
double ora_misfit(const double *G, int64_t nd, int64_t nm, const double *m, const double *Wd,
                  const double *dobs)
{
    double *y = (double *)malloc(sizeof(double) * (size_t)nd);
    ora_gemv_n(G, nd, nm, m, y);
    double s = 0.0;
    for (int64_t i = 0; i < nd; i++) {
        double v = dg(Wd, i) * (y[i] - dobs[i]);
        s += v * v;
    }
    free(y);
    return sqrt(s / (double)nd);
}

/* WL_ii = (x_i^2 + eps2)^(-(2-p)/4), x = W m   (GaussNewtonInversion.cpp:962-1022).
 * p is the integer Lp exponent (the reference stores Lp as int). */
void ora_irls_weights(const ora_csr *W, const double *m, int p, double eps2, double *WL)
{
    double e = -(2 - p) / 4.0;
    for (int64_t i = 0; i < W->n; i++) {
        double x = 0.0;
        for (int64_t k = W->rp[i]; k < W->rp[i + 1]; k++) x += W->v[k] * m[W->ci[k]];
        WL[i] = pow(x * x + eps2, e);
    }
}
