// C entry points over the reference's own prism-field sources, compiled from
// src/Forward_modelling/{Point,RectPrism,Integral,GravFormula}.cpp where they
// lie in the reference tree (see oracle/Makefile, target `ref`).
//
// TEST INFRASTRUCTURE ONLY: the product never loads this library.  It is used
// to check the C restatement (prism_oracle.c) bit for bit, to write the gx/gy
// golden vectors, and as the "reference" CPU baseline for G assembly.
//
// The reference's field dispatcher is
//   GravFormula::field_caused_by_single_prism (GravFormula.cpp:2-36),
// its dense assembly loop is Fwd::compute_G (Fwd.cpp:47-79) and its forward
// loop is compute.cpp:246-260; Fwd.cpp itself cannot be compiled here (it
// needs GSL headers the image lacks), so the two loops below restate those
// loops around the reference's own per-entry kernel.
#include <bitset>
#include <cstdint>
#include <vector>

#include "GravFormula.h"

extern "C" {

int ref_prism_field(const double *pt, const double *prism, double rho, uint32_t flag,
                    double *out10)
{
    GravFormula gra;
    std::vector<double> field;
    RectPrism rect(prism[0], prism[2], prism[4], prism[1], prism[3], prism[5]);
    Point p(pt[0], pt[1], pt[2]);
    gra.field_caused_by_single_prism(p, rect, rho, field, std::bitset<10>(flag));
    for (int k = 0; k < 10; k++) out10[k] = field[k];
    return 0;
}

// Row-major G[(k*N_obs + i)*Nm + j] for point rows [row0, row0+nrows).
int ref_assemble(const double *prisms, int64_t Nm, const double *pts, int64_t N_obs,
                 uint32_t flag, int64_t row0, int64_t nrows, double *G)
{
    std::vector<int> bits;
    for (int b = 0; b < 10; b++)
        if (flag & (1u << b)) bits.push_back(b);
    std::bitset<10> bs(flag);
    for (int64_t i = row0; i < row0 + nrows && i < N_obs; i++) {
        Point p(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
        for (int64_t j = 0; j < Nm; j++) {
            const double *pr = prisms + 6 * j;
            RectPrism rect(pr[0], pr[2], pr[4], pr[1], pr[3], pr[5]);
            GravFormula gra;
            std::vector<double> field;
            gra.field_caused_by_single_prism(p, rect, 1.0, field, bs);
            for (size_t k = 0; k < bits.size(); k++)
                G[((int64_t)k * N_obs + i) * Nm + j] = field[bits[k]];
        }
    }
    return 0;
}

int ref_forward(const double *prisms, const double *rho, int64_t Nm, const double *pts,
                int64_t N_obs, uint32_t flag, double *out)
{
    std::vector<int> bits;
    for (int b = 0; b < 10; b++)
        if (flag & (1u << b)) bits.push_back(b);
    std::bitset<10> bs(flag);
    int nc = (int)bits.size();
    for (int64_t i = 0; i < N_obs; i++) {
        Point p(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
        for (int k = 0; k < nc; k++) out[i * nc + k] = 0.0;
        for (int64_t j = 0; j < Nm; j++) {
            const double *pr = prisms + 6 * j;
            RectPrism rect(pr[0], pr[2], pr[4], pr[1], pr[3], pr[5]);
            GravFormula gra;
            std::vector<double> field;
            gra.field_caused_by_single_prism(p, rect, rho[j], field, bs);
            for (int k = 0; k < nc; k++) out[i * nc + k] += field[bits[k]];
        }
    }
    return 0;
}

}  // extern "C"
