/*
 * CPU oracle for the rectangular-prism gravity / gravity-gradient kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (graada3d_amd/, the
 * libgraeng C-ABI) links, loads or calls this file.  It is the checker that
 * the HIP kernels are compared against, plus the CPU baseline leg of bench.py.
 *
 * It restates, in plain C99, the face/edge (divergence-theorem) formulation
 * that the reference evaluates:
 *   - Point algebra            src/Forward_modelling/Point.cpp:158-243
 *   - Integral::Line_R_1       src/Forward_modelling/Integral.cpp:59-85
 *   - Integral::getBeta        src/Forward_modelling/Integral.cpp:540-570
 *   - Integral::Surface_R_1    src/Forward_modelling/Integral.cpp:593-633
 *   - Integral::Surface_R_3    src/Forward_modelling/Integral.cpp:764-793
 *   - Integral::Surface_rR_3   src/Forward_modelling/Integral.cpp:795-823
 *   - GravFormula::{gz,gx,gy,Txx,Txy,Txz,Tyy,Tyz,Tzz}
 *                              src/Forward_modelling/GravFormula.cpp:75-272
 *   - Fwd::compute_G layout    src/Forward_modelling/Fwd.cpp:47-79
 *   - compute.cpp forward sum  src/compute.cpp:246-260
 * Every floating-point expression keeps the reference's association order
 * (and is built without FMA contraction), so with the same libm the results
 * are bit-identical to the reference build; the tests check exactly that
 * against oracle/_ref where it has been built.
 *
 * Degenerate geometry: the reference asserts s0*s1 > 0 when an observation
 * point lies on the extension of an edge through a vertex
 * (Integral.cpp:70-78).  Here that condition sets *status = 1 instead of
 * aborting and the value follows the reference's formula.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stddef.h>

#define ORA_TOL 1.000e-12               /* gs.h:19 */
#define ORA_G0 6.673e-11                /* gs.h:20 */
#define ORA_SI2EOTVOS 1000000000.0      /* gs.h:24 */
#define ORA_SI2MGAL 100000.0            /* gs.h:25 */

typedef struct { double x, y, z; } v3;

static inline v3 mk(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 scl(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double norm(v3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
static inline v3 unit(v3 a) { double l = norm(a); return mk(a.x / l, a.y / l, a.z / l); }
static inline double dist(v3 a, v3 b)
{
    double dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
    return sqrt(dx * dx + dy * dy + dz * dz);
}
/* Point.cpp:158-169 (cross) */
static inline v3 cross(v3 a, v3 b)
{
    return mk(a.y * b.z - b.y * a.z, b.x * a.z - a.x * b.z, a.x * b.y - b.x * a.y);
}
/* Point.cpp:171-182 (unitCross) */
static inline v3 ucross(v3 a, v3 b) { return unit(cross(a, b)); }

/* Integral::Line_R_1, Integral.cpp:59-85: integral of 1/R along v0->v1. */
static double line_r_1(v3 ob, v3 v0, v3 v1, int *status)
{
    v3 t = unit(sub(v1, v0));
    double s0 = dot(sub(v0, ob), t);
    double s1 = dot(sub(v1, ob), t);
    double R0 = dist(v0, ob);
    double R1 = dist(v1, ob);
    if (fabs(s0 + R0) < ORA_TOL) {
        if (!(s0 * s1 > 0)) *status = 1;
        return fabs(log(s1 / s0));
    }
    return log((R1 + s1) / (R0 + s0));
}

/* Integral::getBeta, Integral.cpp:540-570: signed angle sum of a face. */
static double get_beta(v3 ob, const v3 *f)
{
    v3 n = ucross(sub(f[1], f[0]), sub(f[2], f[1]));
    double beta = 0.0;
    for (int i = 0; i < 4; i++) {
        int in = (i + 1) % 4;
        v3 t = unit(sub(f[in], f[i]));
        double s0 = dot(sub(f[i], ob), t);
        double s1 = dot(sub(f[in], ob), t);
        v3 m = cross(t, n);
        double mv = dot(m, sub(f[i], ob));
        double mp = dot(m, sub(f[in], ob));
        double tt;
        if (fabs(mp) < ORA_TOL) tt = 0.;
        else if (mp > 0) tt = 1.;
        else tt = -1.;
        if (fabs(mv) < ORA_TOL) beta += 0.0;
        else beta = beta + tt * (atan(s1 / fabs(mv)) - atan(s0 / fabs(mv)));
    }
    return beta;
}

/* Integral::Surface_R_1, Integral.cpp:593-633: integral of 1/R over a face. */
static double surface_r_1(v3 ob, const v3 *f)
{
    double beta = get_beta(ob, f);
    v3 n = ucross(sub(f[1], f[0]), sub(f[3], f[0]));
    double h = dot(sub(ob, f[0]), n);
    double ah = fabs(h);
    double res = 0.0;
    for (int j = 0; j < 4; j++) {
        int nx = (j + 1) % 4;
        v3 t = unit(sub(f[nx], f[j]));
        v3 m = ucross(t, n);
        double mv = dot(sub(f[j], ob), m);
        double bj = 0.0;
        if (fabs(mv) < ORA_TOL) {
            bj = 0.0;
        } else {
            double s0 = dot(t, sub(f[j], ob));
            double s1 = dot(t, sub(f[nx], ob));
            double R0 = dist(ob, f[j]);
            double R1 = dist(ob, f[nx]);
            bj = ah * (atan((ah * s1) / (mv * R1)) - atan((ah * s0) / (mv * R0)))
                 + mv * log((s1 + R1) / (s0 + R0));
        }
        res += bj;
    }
    return res - beta * ah;
}

/* Integral::Surface_R_3, Integral.cpp:764-793: integral of 1/R^3 (|h|>=TOL). */
static double surface_r_3(v3 ob, const v3 *f)
{
    v3 n = ucross(sub(f[1], f[0]), sub(f[2], f[1]));
    double h = dot(sub(ob, f[0]), n);
    double ah = fabs(h);
    double res = 0.0;
    for (int i = 0; i < 4; i++) {
        int nx = (i + 1) % 4;
        v3 e = unit(sub(f[nx], f[i]));
        v3 m = ucross(sub(f[nx], f[i]), n);
        /* Point::projection_line(p1=f[nx], p2=f[i]), Point.cpp:232-237 */
        v3 tl = unit(sub(f[i], f[nx]));
        double pr = dot(sub(ob, f[nx]), tl);
        v3 sc = scl(tl, pr);
        v3 prj = mk(f[nx].x + sc.x, f[nx].y + sc.y, f[nx].z + sc.z);
        double L0 = norm(sub(prj, ob));
        double s1 = dot(sub(f[nx], ob), e);
        double s0 = dot(sub(f[i], ob), e);
        double R1 = dist(ob, f[nx]);
        double R0 = dist(ob, f[i]);
        double m0 = dot(sub(f[i], ob), m);
        res += atan((m0 * s1) / (L0 * L0 + ah * R1)) - atan((m0 * s0) / (L0 * L0 + ah * R0));
    }
    return res / ah;
}

/* Integral::Surface_rR_3, Integral.cpp:795-823: vector integral of r/R^3. */
static v3 surface_rr_3(v3 ob, const v3 *f, int *status)
{
    v3 n = ucross(sub(f[1], f[0]), sub(f[2], f[1]));
    double h = dot(sub(ob, f[0]), n);
    v3 res = mk(0, 0, 0);
    for (int i = 0; i < 4; i++) {
        int nx = (i + 1) % 4;
        v3 m = ucross(sub(f[nx], f[i]), n);
        res = sub(res, scl(m, line_r_1(ob, f[i], f[nx], status)));
    }
    if (!(fabs(h) < ORA_TOL)) {
        double sr3 = surface_r_3(ob, f);
        res = sub(res, scl(scl(n, h), sr3));
    }
    return res;
}

/* Field bits (gs.h:46-57): 0 V, 1 gz, 2 gx, 3 gy, 4 Tzz, 5 Txz, 6 Tyz,
 * 7 Txx, 8 Txy, 9 Tyy.  Writes out[bit] for every requested bit 1..9;
 * bit 0 (potential) is not part of this path and is rejected by callers. */
static void prism_fields(v3 ob, const double *pr, double rho, uint32_t flag,
                         double *out, int *status)
{
    double x0 = pr[0], x1 = pr[1], y0 = pr[2], y1 = pr[3], z0 = pr[4], z1 = pr[5];
    v3 p0 = mk(x0, y0, z0), p1 = mk(x1, y0, z0), p2 = mk(x1, y1, z0), p3 = mk(x0, y1, z0);
    v3 p4 = mk(x0, y0, z1), p5 = mk(x1, y0, z1), p6 = mk(x1, y1, z1), p7 = mk(x0, y1, z1);
    /* face vertex orders of GravFormula.cpp:87-88, 110-111, 133-134 */
    v3 fx0[4] = {p0, p4, p7, p3}, fx1[4] = {p1, p2, p6, p5};
    v3 fy0[4] = {p0, p1, p5, p4}, fy1[4] = {p2, p3, p7, p6};
    v3 fz0[4] = {p3, p2, p1, p0}, fz1[4] = {p4, p5, p6, p7};
    double cg = -ORA_SI2MGAL * ORA_G0 * rho;
    double ct = -ORA_SI2EOTVOS * ORA_G0 * rho;
    if (flag & (1u << 1)) out[1] = cg * (-surface_r_1(ob, fz0) + surface_r_1(ob, fz1));
    if (flag & (1u << 2)) out[2] = cg * (-surface_r_1(ob, fx0) + surface_r_1(ob, fx1));
    if (flag & (1u << 3)) out[3] = cg * (-surface_r_1(ob, fy0) + surface_r_1(ob, fy1));
    if (flag & (1u << 4)) {
        v3 a = surface_rr_3(ob, fz0, status), b = surface_rr_3(ob, fz1, status);
        out[4] = ct * (-a.z + b.z);
    }
    if (flag & ((1u << 5) | (1u << 7) | (1u << 8))) {
        v3 a = surface_rr_3(ob, fx0, status), b = surface_rr_3(ob, fx1, status);
        if (flag & (1u << 5)) out[5] = ct * (-a.z + b.z);
        if (flag & (1u << 7)) out[7] = ct * (-a.x + b.x);
        if (flag & (1u << 8)) out[8] = ct * (-a.y + b.y);
    }
    if (flag & ((1u << 6) | (1u << 9))) {
        v3 a = surface_rr_3(ob, fy0, status), b = surface_rr_3(ob, fy1, status);
        if (flag & (1u << 6)) out[6] = ct * (-a.z + b.z);
        if (flag & (1u << 9)) out[9] = ct * (-a.y + b.y);
    }
}

static int comp_bits(uint32_t flag, int *bits)
{
    int n = 0;
    for (int b = 1; b < 10; b++)
        if (flag & (1u << b)) bits[n++] = b;
    return n;
}

/*
 * One (point, prism) evaluation.  prism = {x0,x1,y0,y1,z0,z1}; out[10] is
 * indexed by field bit.  Returns 0, or 1 for the degenerate-geometry case.
 */
int ora_prism_field(const double *point, const double *prism, double rho,
                    uint32_t flag, double *out10)
{
    int st = 0;
    for (int k = 0; k < 10; k++) out10[k] = 0.0;
    prism_fields(mk(point[0], point[1], point[2]), prism, rho, flag, out10, &st);
    return st;
}

/*
 * Fwd::compute_G (Fwd.cpp:47-79) for rows [row0, row0+nrows) of the points:
 * G is written ROW-MAJOR, G[(k*N_obs + i) * Nm + j] for component k (ascending
 * field bit), point i, prism j; only point rows row0..row0+nrows-1 are filled
 * (the caller's buffer holds n_comp*N_obs rows).  Returns 0 or 1 (degenerate).
 */
/*
 * Threads used by the loops below (points for assembly / forward, rows of
 * G x, column blocks of G^T x).  Each thread computes whole output entries in
 * the serial order, so results are identical for every thread count; 1 (the
 * default) reproduces the reference's single-threaded Eigen GEMV timing.
 */
int ora_threads = 1;
void ora_set_threads(int n) { ora_threads = n > 0 ? n : 1; }
int ora_max_threads(void) { return omp_get_max_threads(); }

int ora_assemble(const double *prisms, int64_t Nm, const double *pts, int64_t N_obs,
                 uint32_t flag, int64_t row0, int64_t nrows, double *G)
{
    int bits[10];
    int nc = comp_bits(flag, bits);
    int st = 0;
    int64_t i1 = row0 + nrows < N_obs ? row0 + nrows : N_obs;
#pragma omp parallel for num_threads(ora_threads) schedule(dynamic, 4) reduction(| : st)
    for (int64_t i = row0; i < i1; i++) {
        double f[10];
        int sti = 0;
        v3 ob = mk(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
        for (int64_t j = 0; j < Nm; j++) {
            prism_fields(ob, prisms + 6 * j, 1.0, flag, f, &sti);
            for (int k = 0; k < nc; k++)
                G[((int64_t)k * N_obs + i) * Nm + j] = f[bits[k]];
        }
        st |= sti;
    }
    return st;
}

/*
 * compute.cpp:246-260 matrix-free forward: out[i*n_comp + k] is the sum over
 * prisms j (in order) of field_k(point i, prism j, rho_j).
 */
int ora_forward(const double *prisms, const double *rho, int64_t Nm, const double *pts,
                int64_t N_obs, uint32_t flag, double *out)
{
    int bits[10];
    int nc = comp_bits(flag, bits);
    int st = 0;
#pragma omp parallel for num_threads(ora_threads) schedule(dynamic, 4) reduction(| : st)
    for (int64_t i = 0; i < N_obs; i++) {
        double f[10];
        int sti = 0;
        v3 ob = mk(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
        for (int k = 0; k < nc; k++) out[i * nc + k] = 0.0;
        for (int64_t j = 0; j < Nm; j++) {
            prism_fields(ob, prisms + 6 * j, rho[j], flag, f, &sti);
            for (int k = 0; k < nc; k++) out[i * nc + k] += f[bits[k]];
        }
        st |= sti;
    }
    return st;
}
