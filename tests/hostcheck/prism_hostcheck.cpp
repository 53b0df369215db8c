// Test support only: compiles the engine's per-(point, prism) math
// (graada3d_amd/csrc/prism_math.hpp) for the host with g++, so tests can
// compare it bit for bit with the CPU oracle under the same libm.  On the GPU
// the same header is compiled for gfx950; only atan/log come from a
// different library there.
#define GRA_HD
#include <cstdint>

#include "prism_math.hpp"
#include "fastlog.hpp"
#include "fastatan.hpp"

extern "C" double hc_atan_ratio(double n, double d) { return gra::atan_ratio(n, d); }

extern "C" double hc_table_log(double v) { return gra::fast_log(v); }

extern "C" int hc_prism_field(const double *pt, const double *pr, double rho, uint32_t flag,
                              double *out10)
{
    for (int k = 0; k < 10; ++k) out10[k] = 0.0;
    // the compile-time instantiations the GPU kernels use, else the runtime path
    switch (flag) {
    case gra::kGz | gra::kTzz: return gra::prism_fields_t<gra::kGz | gra::kTzz>(pt, pr, rho, out10);
    case gra::kGz: return gra::prism_fields_t<gra::kGz>(pt, pr, rho, out10);
    case gra::kTzz: return gra::prism_fields_t<gra::kTzz>(pt, pr, rho, out10);
    case gra::kAllFields: return gra::prism_fields_t<gra::kAllFields>(pt, pr, rho, out10);
    default: return gra::prism_fields(pt, pr, rho, flag, out10);
    }
}

// The fused product's launch planner (graada3d_amd/csrc/fused_plan.hpp).
// out: NT, S, per_xcd, rows_per_group, nrb, grid.
#include "fused_plan.hpp"

extern "C" int hc_fused_plan(int64_t Nd, int64_t ld, int TW, int RB, int cus, int n_xcd,
                             int per_cu, int target, int max_poll, int min_blocks, int64_t *out6)
{
    gra::FusedShape s;
    if (!gra::fused_plan_shape(Nd, ld, TW, RB, cus, n_xcd, per_cu, target, max_poll, &s,
                               min_blocks))
        return 0;
    out6[0] = s.NT;
    out6[1] = s.S;
    out6[2] = s.per_xcd;
    out6[3] = s.rows_per_group;
    out6[4] = s.nrb;
    out6[5] = s.grid;
    return 1;
}
