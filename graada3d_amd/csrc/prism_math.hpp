// Per-(point, prism) field evaluation in FP64, face/edge formulation.
//
// This is the reference's closed form (GravFormula.cpp:75-272 over
// Integral::Surface_R_1 / getBeta / Surface_rR_3 / Surface_R_3 / Line_R_1,
// Integral.cpp:59-85, 540-633, 764-823), specialised to axis-aligned prisms.
//
// Why the specialisation is exact.  For a RectPrism every edge vector has one
// non-zero coordinate, so the reference's unit tangents t, unit face normals n
// and in-plane edge normals m (unitCross / getUnit, Point.cpp:158-194) are
// signed unit axis vectors *exactly* (sqrt(fl(d*d)) == |d| in binary IEEE
// arithmetic, and products of +-1 with +-0 are exact).  A dot product with such
// a vector therefore returns one signed, rounded coordinate difference, e.g.
// (f - a).t == +-fl(f_x - a_x), bit for bit.  Every quantity below is formed
// from the six differences D = fl(bound - a) and eight vertex distances
// R = sqrt((dx^2 + dy^2) + dz^2) in the reference's operation order, so with
// the same libm the results equal the reference's bit for bit (checked on the
// host against the oracle); on the GPU only atan/log differ (device libm).
//
// Per face the edge table lists, in the reference's vertex order: the
// coordinate differences giving s0, s1 (along the edge, signed by the
// traversal direction), mv (along the outward in-plane edge normal m), the
// vertex distances R0, R1 and, for the 1/R^3 term, the perpendicular distance
// L0 of Point::projection_line (Point.cpp:232-237) including its rounding.
#pragma once

#ifndef GRA_HD
#define GRA_HD __host__ __device__
#endif

// GRA_ATAN_PAIRS = 1 evaluates every difference atan(x) - atan(y) of the
// formulas below as one atan (see atan_diff): half the atans and one
// division per pair instead of two.  The GPU build uses it; the host build
// keeps the reference's operation order (0) unless asked otherwise.
#ifndef GRA_ATAN_PAIRS
#if defined(__HIP_DEVICE_COMPILE__)
#define GRA_ATAN_PAIRS 1
#else
#define GRA_ATAN_PAIRS 0
#endif
#endif

#include <math.h>
#include <stdint.h>

// GRA_TABLE_LOG = 1 takes the edge logs from table_log (fastlog.hpp, about a
// quarter of the library log's instructions, within ~1 ulp); the GPU build
// uses it, the host build keeps the library log unless asked otherwise.
#ifndef GRA_TABLE_LOG
#if defined(__HIP_DEVICE_COMPILE__)
#define GRA_TABLE_LOG 1
#else
#define GRA_TABLE_LOG 0
#endif
#endif
#if GRA_TABLE_LOG
#include "fastlog.hpp"
#endif
// GRA_FAST_ATAN = 1 takes the atans of atan_diff from atan_ratio
// (fastatan.hpp: one division, a degree-20 polynomial, within ~1 ulp); the
// GPU build uses it, the host build keeps the library atan unless asked.
#ifndef GRA_FAST_ATAN
#if defined(__HIP_DEVICE_COMPILE__)
#define GRA_FAST_ATAN 1
#else
#define GRA_FAST_ATAN 0
#endif
#endif
#if GRA_FAST_ATAN
#include "fastatan.hpp"
#endif

namespace gra {

constexpr double kTol = 1.000e-12;            // gs.h:19
constexpr double kG0 = 6.673e-11;             // gs.h:20
constexpr double kSI2Eotvos = 1000000000.0;   // gs.h:24
constexpr double kSI2mGal = 100000.0;         // gs.h:25

// Field bits of gs.h:46-57.
enum : uint32_t {
    kGz = 1u << 1, kGx = 1u << 2, kGy = 1u << 3, kTzz = 1u << 4, kTxz = 1u << 5,
    kTyz = 1u << 6, kTxx = 1u << 7, kTxy = 1u << 8, kTyy = 1u << 9,
    kAllFields = 0x3FEu
};

// One edge of a face as seen from the observation point.
struct Edge {
    double s0, s1;   // t.(v0 - a), t.(v1 - a)
    double mv;       // m.(v0 - a) == m.(v1 - a)
    double R0, R1;   // |v0 - a|, |v1 - a|
    int ax;          // axis of the edge (0 x, 1 y, 2 z) -- for L0
    double c_next;   // coordinate of v1 along ax (projection base point)
    double pa, pb;   // fl(v1 - a) on the two axes other than ax, in xyz order
    int max_;        // axis of m (for the vector integral)
    double msign;    // sign of m along max_
};

struct Geo {
    double a[3];               // observation point
    double D[3][2];            // D[ax][lo/hi] = fl(bound - a[ax])
    double R[2][2][2];         // R[ix][iy][iz]
};

GRA_HD inline void geo_init(Geo &g, const double *pt, const double *pr)
{
    g.a[0] = pt[0]; g.a[1] = pt[1]; g.a[2] = pt[2];
    for (int ax = 0; ax < 3; ++ax) {
        g.D[ax][0] = pr[2 * ax] - g.a[ax];
        g.D[ax][1] = pr[2 * ax + 1] - g.a[ax];
    }
    double sx[2] = {g.D[0][0] * g.D[0][0], g.D[0][1] * g.D[0][1]};
    double sy[2] = {g.D[1][0] * g.D[1][0], g.D[1][1] * g.D[1][1]};
    double sz[2] = {g.D[2][0] * g.D[2][0], g.D[2][1] * g.D[2][1]};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) g.R[i][j][k] = sqrt((sx[i] + sy[j]) + sz[k]);
}

// Face descriptor: 4 vertices as (ix,iy,iz) bound indices in the reference's
// order, the face's normal axis and its outward sign.
struct FaceDef {
    int8_t v[4][3];
    int8_t nax;
    int8_t nsign;
};

// GravFormula.cpp vertex orders: p0=(0,0,0) p1=(1,0,0) p2=(1,1,0) p3=(0,1,0)
// p4=(0,0,1) p5=(1,0,1) p6=(1,1,1) p7=(0,1,1).
//   z faces {p3,p2,p1,p0} / {p4,p5,p6,p7}   (GravFormula.cpp:133-134)
//   x faces {p0,p4,p7,p3} / {p1,p2,p6,p5}   (GravFormula.cpp:87-88)
//   y faces {p0,p1,p5,p4} / {p2,p3,p7,p6}   (GravFormula.cpp:110-111)
// The table is constexpr and every use below has a compile-time face index
// and unrolled edge loops, so all vertex / axis indices fold to constants
// and the per-thread arrays stay in registers.
enum FaceId { kZ0 = 0, kZ1, kX0, kX1, kY0, kY1 };

constexpr FaceDef kFaces[6] = {
    {{{0, 1, 0}, {1, 1, 0}, {1, 0, 0}, {0, 0, 0}}, 2, -1},
    {{{0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}}, 2, +1},
    {{{0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0}}, 0, -1},
    {{{1, 0, 0}, {1, 1, 0}, {1, 1, 1}, {1, 0, 1}}, 0, +1},
    {{{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}}, 1, -1},
    {{{1, 1, 0}, {0, 1, 0}, {0, 1, 1}, {1, 1, 1}}, 1, +1},
};

// Builds edge J of face F (v0 = vertex J, v1 = vertex J+1).
template <int F, int J>
GRA_HD inline Edge make_edge(const Geo &g, const double *pr)
{
    constexpr FaceDef fd = kFaces[F];
    constexpr int8_t u[3] = {fd.v[J][0], fd.v[J][1], fd.v[J][2]};
    constexpr int8_t w[3] = {fd.v[(J + 1) & 3][0], fd.v[(J + 1) & 3][1], fd.v[(J + 1) & 3][2]};
    constexpr int ax = (u[0] != w[0]) ? 0 : ((u[1] != w[1]) ? 1 : 2);
    constexpr double dir = (w[ax] > u[ax]) ? 1.0 : -1.0;   // t = dir * e_ax
    Edge e;
    e.ax = ax;
    e.s0 = dir * g.D[ax][u[ax]];
    e.s1 = dir * g.D[ax][w[ax]];
    e.R0 = g.R[u[0]][u[1]][u[2]];
    e.R1 = g.R[w[0]][w[1]][w[2]];
    // In-plane outward normal m = t x n lies on the remaining axis; the edge
    // sits on the bound of that axis it shares with both vertices.
    constexpr int max_ = 3 - ax - fd.nax;
    constexpr int side = u[max_];
    e.max_ = max_;
    e.msign = side ? 1.0 : -1.0;
    e.mv = e.msign * g.D[max_][side];
    e.c_next = pr[2 * ax + w[ax]];
    constexpr int o1 = (ax == 0) ? 1 : 0, o2 = (ax == 2) ? 1 : 2;
    e.pa = g.D[o1][w[o1]];
    e.pb = g.D[o2][w[o2]];
    return e;
}

// Perpendicular distance of Point::projection_line(v1, v0) from a, with the
// rounding of prj = v1 + ((a - v1).tl) tl along the edge axis.
GRA_HD inline double edge_L0(const Geo &g, const Edge &e)
{
    double d = g.a[e.ax] - e.c_next;
    double prj = e.c_next + d;
    double r = prj - g.a[e.ax];
    double q[3];
    int o1 = (e.ax == 0) ? 1 : 0, o2 = (e.ax == 2) ? 1 : 2;
    q[e.ax] = r;
    q[o1] = e.pa;
    q[o2] = e.pb;
    return sqrt((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]);
}

// h = (a - f0).n for face F (n = nsign * e_nax).
template <int F> GRA_HD inline double face_h(const Geo &g)
{
    constexpr FaceDef fd = kFaces[F];
    return -(double(fd.nsign) * g.D[fd.nax][fd.v[0][fd.nax]]);
}

template <int F> GRA_HD inline void face_edges(const Geo &g, const double *pr, Edge *E)
{
    E[0] = make_edge<F, 0>(g, pr);
    E[1] = make_edge<F, 1>(g, pr);
    E[2] = make_edge<F, 2>(g, pr);
    E[3] = make_edge<F, 3>(g, pr);
}

// atan(x) - atan(y) for x - y = num / q, 1 + x y = den / q with q > 0:
// atan(num / den), plus pi sgn(num) when den < 0 (the difference leaves the
// principal range).  den = 0 gives atan(+-inf) = +-pi/2 as the exact value.
GRA_HD inline double atan_diff(double num, double den)
{
#if GRA_FAST_ATAN
    double a = atan_ratio(num, den);
#else
    double a = atan(num / den);
#endif
    if (den < 0.0) a += (num < 0.0) ? -3.14159265358979323846 : 3.14159265358979323846;
    return a;
}

// log((s1 + R1)/(s0 + R0)) of an edge (shared by Surface_R_1 and Line_R_1).
GRA_HD inline double edge_log(const Edge &e)
{
#if GRA_TABLE_LOG
    const double v = (e.s1 + e.R1) / (e.s0 + e.R0);
    // the table log for positive finite ratios; 0, inf and NaN keep the
    // library's value (the reference's special cases rely on them), on a
    // branch of its own so that the library log is not evaluated beside
    // the table log for every edge
    double lv = fast_log(v);
    if (__builtin_expect(!(v > 0.0 && v < INFINITY), 0)) lv = log(v);
    return lv;
#else
    return log((e.s1 + e.R1) / (e.s0 + e.R0));
#endif
}

// Integral::Surface_R_1 (Integral.cpp:593-633) including getBeta (:540-570).
GRA_HD inline double surface_r_1(const Edge *E, const double *lg, double h)
{
    double ah = fabs(h);
    double beta = 0.0;
    for (int j = 0; j < 4; ++j) {
        double am = fabs(E[j].mv);
        if (am < kTol) {
            beta += 0.0;
        } else {
            double tt = (E[j].mv > 0) ? 1. : -1.;
#if GRA_ATAN_PAIRS
            // x = s1/am, y = s0/am
            beta = beta + tt * atan_diff((E[j].s1 - E[j].s0) * am, am * am + E[j].s0 * E[j].s1);
#else
            beta = beta + tt * (atan(E[j].s1 / am) - atan(E[j].s0 / am));
#endif
        }
    }
    double res = 0.0;
    for (int j = 0; j < 4; ++j) {
        double bj = 0.0;
        if (!(fabs(E[j].mv) < kTol)) {
#if GRA_ATAN_PAIRS
            // x = ah s1 / (mv R1), y = ah s0 / (mv R0)
            const double mv = E[j].mv;
            bj = ah * atan_diff(ah * mv * (E[j].s1 * E[j].R0 - E[j].s0 * E[j].R1),
                                mv * mv * (E[j].R0 * E[j].R1) + (ah * ah) * (E[j].s0 * E[j].s1)) +
                 mv * lg[j];
#else
            bj = ah * (atan((ah * E[j].s1) / (E[j].mv * E[j].R1)) -
                       atan((ah * E[j].s0) / (E[j].mv * E[j].R0))) +
                 E[j].mv * lg[j];
#endif
        }
        res += bj;
    }
    return res - beta * ah;
}

// Integral::Surface_R_3 (Integral.cpp:764-793), |h| >= TOL.
GRA_HD inline double surface_r_3(const Geo &g, const Edge *E, double h)
{
    double ah = fabs(h);
    double res = 0.0;
    for (int i = 0; i < 4; ++i) {
        double L0 = edge_L0(g, E[i]);
        double LL = L0 * L0;
        double m0 = E[i].mv;
#if GRA_ATAN_PAIRS
        // x = m0 s1 / A1, y = m0 s0 / A0, A = LL + ah R > 0
        const double A1 = LL + ah * E[i].R1, A0 = LL + ah * E[i].R0;
        res += atan_diff(m0 * (E[i].s1 * A0 - E[i].s0 * A1), A0 * A1 + (m0 * m0) * (E[i].s0 * E[i].s1));
#else
        res += atan((m0 * E[i].s1) / (LL + ah * E[i].R1)) -
               atan((m0 * E[i].s0) / (LL + ah * E[i].R0));
#endif
    }
    return res / ah;
}

// Integral::Line_R_1 (Integral.cpp:59-85) given the shared log; sets *bad for
// the asserted case s0*s1 <= 0 (observation point on the edge itself).
GRA_HD inline double line_r_1(const Edge &e, double lg, int *bad)
{
    if (fabs(e.s0 + e.R0) < kTol) {
        if (!(e.s0 * e.s1 > 0)) *bad = 1;
        return fabs(log(e.s1 / e.s0));
    }
    return lg;
}

// Integral::Surface_rR_3 (Integral.cpp:795-823): vector integral of r/R^3.
// The edge terms land on the axis of each edge's normal m; the 1/R^3 term on
// the face-normal axis.  A non-finite edge term poisons every component, as
// the reference's 0 * inf products do.
GRA_HD inline void surface_rr_3(const Geo &g, const Edge *E, const double *lg, double h,
                                int nax, double nsign, double *out, int *bad)
{
    double r[3] = {0.0, 0.0, 0.0};
    bool poison = false;
    for (int i = 0; i < 4; ++i) {
        double L = line_r_1(E[i], lg[i], bad);
        if (!isfinite(L)) poison = true;
        r[E[i].max_] = r[E[i].max_] - E[i].msign * L;
    }
    if (!(fabs(h) < kTol)) {
        double sr3 = surface_r_3(g, E, h);
        r[nax] = r[nax] - (nsign * h) * sr3;
    }
    if (poison) {
        double nan = NAN;
        r[0] = r[1] = r[2] = nan;
    }
    out[0] = r[0]; out[1] = r[1]; out[2] = r[2];
}

// Only the face-normal component of Surface_rR_3 (what Tzz/Txx/Tyy use from
// a face normal to their own axis): the edge terms do not reach it.
GRA_HD inline double surface_rr_3_normal(const Geo &g, const Edge *E, double h, double nsign,
                                         int *bad)
{
    bool poison = false;
    for (int i = 0; i < 4; ++i) {
        if (fabs(E[i].s0 + E[i].R0) < kTol) {
            if (!(E[i].s0 * E[i].s1 > 0)) *bad = 1;
            if (!isfinite(log(E[i].s1 / E[i].s0))) poison = true;
        }
    }
    double r = 0.0;
    if (!(fabs(h) < kTol)) r = r - (nsign * h) * surface_r_3(g, E, h);
    return poison ? (double)NAN : r;
}

// One face normal to z: Surface_R_1 (gz) and the normal component of
// Surface_rR_3 (Tzz).
template <int F, uint32_t FLAG>
GRA_HD inline void z_face(const Geo &g, const double *pr, double &vR1, double &vT, int *bad)
{
    Edge E[4];
    face_edges<F>(g, pr, E);
    const double h = face_h<F>(g);
    if (FLAG & kGz) {
        double lg[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) lg[j] = fabs(E[j].mv) < kTol ? 0.0 : edge_log(E[j]);
        vR1 = surface_r_1(E, lg, h);
    }
    if (FLAG & kTzz) vT = surface_rr_3_normal(g, E, h, double(kFaces[F].nsign), bad);
}

// One face normal to x or y: Surface_R_1 (gx / gy) and the full vector
// Surface_rR_3 (or only its normal component when that is all FLAG needs).
template <int F, bool R1, bool FULL, bool NORMAL>
GRA_HD inline void side_face(const Geo &g, const double *pr, double &vR1, double *vT, int *bad)
{
    Edge E[4];
    double lg[4];
    face_edges<F>(g, pr, E);
#pragma unroll
    for (int j = 0; j < 4; ++j) lg[j] = edge_log(E[j]);
    const double h = face_h<F>(g);
    constexpr int nax = kFaces[F].nax;
    if (R1) vR1 = surface_r_1(E, lg, h);
    if (FULL) {
        surface_rr_3(g, E, lg, h, nax, double(kFaces[F].nsign), vT, bad);
    } else if (NORMAL) {
        vT[nax] = surface_rr_3_normal(g, E, h, double(kFaces[F].nsign), bad);
    }
}

// Face groups.  Each evaluates the two faces normal to one axis for a
// compile-time component subset and writes only those outputs; `bad` is set
// only by the evaluations the subset needs, as in the reference.
template <bool GZ, bool TZZ>
GRA_HD inline void z_group(const Geo &g, const double *pr, double cg, double ct, double *out,
                           int *bad)
{
    constexpr uint32_t F = (GZ ? kGz : 0u) | (TZZ ? kTzz : 0u);
    double r0 = 0, r1 = 0, t0 = 0, t1 = 0;
    z_face<kZ0, F>(g, pr, r0, t0, bad);
    z_face<kZ1, F>(g, pr, r1, t1, bad);
    if (GZ) out[1] = cg * (-r0 + r1);
    if (TZZ) out[4] = ct * (-t0 + t1);
}

// x faces: R1 -> gx; FULL -> Txy and Txz (and Txx); NORMAL -> Txx only.
template <bool R1, bool FULL, bool NORMAL>
GRA_HD inline void x_group(const Geo &g, const double *pr, double cg, double ct, uint32_t flag,
                           double *out, int *bad)
{
    double r0 = 0, r1 = 0, t0[3] = {0, 0, 0}, t1[3] = {0, 0, 0};
    side_face<kX0, R1, FULL, NORMAL>(g, pr, r0, t0, bad);
    side_face<kX1, R1, FULL, NORMAL>(g, pr, r1, t1, bad);
    if (R1) out[2] = cg * (-r0 + r1);
    if (flag & kTxz) out[5] = ct * (-t0[2] + t1[2]);
    if (flag & kTxx) out[7] = ct * (-t0[0] + t1[0]);
    if (flag & kTxy) out[8] = ct * (-t0[1] + t1[1]);
}

// y faces: R1 -> gy; FULL -> Tyz (and Tyy); NORMAL -> Tyy only.
template <bool R1, bool FULL, bool NORMAL>
GRA_HD inline void y_group(const Geo &g, const double *pr, double cg, double ct, uint32_t flag,
                           double *out, int *bad)
{
    double r0 = 0, r1 = 0, t0[3] = {0, 0, 0}, t1[3] = {0, 0, 0};
    side_face<kY0, R1, FULL, NORMAL>(g, pr, r0, t0, bad);
    side_face<kY1, R1, FULL, NORMAL>(g, pr, r1, t1, bad);
    if (R1) out[3] = cg * (-r0 + r1);
    if (flag & kTyz) out[6] = ct * (-t0[2] + t1[2]);
    if (flag & kTyy) out[9] = ct * (-t0[1] + t1[1]);
}

// Components of a compile-time FLAG (field bits 1..9) of one (point, prism)
// pair with density rho into out[bit].  Returns 1 when the reference would
// have asserted (degenerate geometry), else 0.
template <uint32_t FLAG>
GRA_HD inline int prism_fields_t(const double *pt, const double *pr, double rho, double *out)
{
    Geo g;
    geo_init(g, pt, pr);
    int bad = 0;
    const double cg = -kSI2mGal * kG0 * rho;
    const double ct = -kSI2Eotvos * kG0 * rho;
    if (FLAG & (kGz | kTzz))
        z_group<(FLAG & kGz) != 0, (FLAG & kTzz) != 0>(g, pr, cg, ct, out, &bad);
    if (FLAG & (kGx | kTxx | kTxy | kTxz))
        x_group<(FLAG & kGx) != 0, (FLAG & (kTxy | kTxz)) != 0, (FLAG & kTxx) != 0>(
            g, pr, cg, ct, FLAG, out, &bad);
    if (FLAG & (kGy | kTyy | kTyz))
        y_group<(FLAG & kGy) != 0, (FLAG & kTyz) != 0, (FLAG & kTyy) != 0>(g, pr, cg, ct, FLAG,
                                                                            out, &bad);
    return bad;
}

// Runtime-flag entry: dispatches each face group to its instantiation.
GRA_HD inline int prism_fields(const double *pt, const double *pr, double rho, uint32_t flag,
                               double *out)
{
    Geo g;
    geo_init(g, pt, pr);
    int bad = 0;
    const double cg = -kSI2mGal * kG0 * rho;
    const double ct = -kSI2Eotvos * kG0 * rho;
    switch (flag & (kGz | kTzz)) {
    case kGz: z_group<true, false>(g, pr, cg, ct, out, &bad); break;
    case kTzz: z_group<false, true>(g, pr, cg, ct, out, &bad); break;
    case kGz | kTzz: z_group<true, true>(g, pr, cg, ct, out, &bad); break;
    default: break;
    }
    {
        const int sel = ((flag & kGx) ? 4 : 0) | ((flag & (kTxy | kTxz)) ? 2 : 0) |
                        ((flag & kTxx) ? 1 : 0);
        switch (sel) {
        case 1: x_group<false, false, true>(g, pr, cg, ct, flag, out, &bad); break;
        case 2: x_group<false, true, false>(g, pr, cg, ct, flag, out, &bad); break;
        case 3: x_group<false, true, true>(g, pr, cg, ct, flag, out, &bad); break;
        case 4: x_group<true, false, false>(g, pr, cg, ct, flag, out, &bad); break;
        case 5: x_group<true, false, true>(g, pr, cg, ct, flag, out, &bad); break;
        case 6: x_group<true, true, false>(g, pr, cg, ct, flag, out, &bad); break;
        case 7: x_group<true, true, true>(g, pr, cg, ct, flag, out, &bad); break;
        default: break;
        }
    }
    {
        const int sel = ((flag & kGy) ? 4 : 0) | ((flag & kTyz) ? 2 : 0) | ((flag & kTyy) ? 1 : 0);
        switch (sel) {
        case 1: y_group<false, false, true>(g, pr, cg, ct, flag, out, &bad); break;
        case 2: y_group<false, true, false>(g, pr, cg, ct, flag, out, &bad); break;
        case 3: y_group<false, true, true>(g, pr, cg, ct, flag, out, &bad); break;
        case 4: y_group<true, false, false>(g, pr, cg, ct, flag, out, &bad); break;
        case 5: y_group<true, false, true>(g, pr, cg, ct, flag, out, &bad); break;
        case 6: y_group<true, true, false>(g, pr, cg, ct, flag, out, &bad); break;
        case 7: y_group<true, true, true>(g, pr, cg, ct, flag, out, &bad); break;
        default: break;
        }
    }
    return bad;
}

}  // namespace gra
