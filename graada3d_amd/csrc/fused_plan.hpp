// Launch shape of the single-read fused product (fused_kernels.hip): a pure
// function of the problem and device figures, so that the host can test it
// without a GPU (tests/hostcheck/fused_plan_check.cpp).
#pragma once

#include <cstdint>

namespace gra {

struct FusedShape {
    int NT = 0;                  // column tiles per row group
    int S = 0;                   // row groups (grid = NT * S workgroups)
    int per_xcd = 0;             // workgroups per XCD when the grid is XCD-aligned, else 0
    int64_t rows_per_group = 0;  // multiple of RB
    int64_t nrb = 0;             // row blocks per group
    int grid = 0;
};

// Nd x ld G; a workgroup covers TW columns and RB rows per step.  `cus`
// compute units in `n_xcd` XCDs, `per_cu` co-resident workgroups per CU (the
// runtime's occupancy figure for the kernel), `target` workgroups requested
// (<= 0: one per CU).  The kernel's polling wave reads the NT * RB partials
// of a block with at most `max_poll` values per lane, so NT * RB <=
// 64 * max_poll.  Returns false when no shape fits; the caller then runs the
// two-pass iteration.
//
// The grid never exceeds per_cu * cus, so every workgroup is resident at once
// (the kernel's waits need that); the cooperative launch checks it again.
// When the grid is a multiple of n_xcd and each XCD's share is a multiple of
// NT, workgroup b is mapped to the share of XCD b mod n_xcd (workgroups are
// dispatched to the XCDs round robin), so a row group's partials stay in one
// XCD; the protocol is correct for any mapping.
// min_blocks > 0: refuse a shape whose groups hold fewer row blocks (the
// pipeline's fill and drain, L + PF steps, would dominate; the two-pass
// iteration is faster there).
inline bool fused_plan_shape(int64_t Nd, int64_t ld, int TW, int RB, int cus, int n_xcd,
                             int per_cu, int target, int max_poll, FusedShape *out,
                             int min_blocks = 0)
{
    if (Nd < 1 || ld < 1 || TW < 1 || RB < 1 || cus < 1 || per_cu < 1) return false;
    if (n_xcd < 1) n_xcd = 1;
    FusedShape s;
    s.NT = (int)((ld + TW - 1) / TW);
    if ((int64_t)s.NT * RB > 64 * (int64_t)max_poll) return false;
    const int64_t cap = (int64_t)per_cu * cus;
    int64_t want = target > 0 ? target : cus;
    if (want > cap) want = cap;
    if (want < s.NT) return false;
    const int64_t s_max = (Nd + RB - 1) / RB;     // at least one row block per group
    int64_t S = want / s.NT;
    int per_xcd = 0;
    if (n_xcd > 1 && want % n_xcd == 0) {
        int64_t gpx = (want / n_xcd) / s.NT;      // groups per XCD
        if (gpx * n_xcd > s_max) gpx = s_max / n_xcd;
        if (gpx >= 1) {
            S = gpx * n_xcd;
            per_xcd = (int)(gpx * s.NT);
        }
    }
    if (per_xcd == 0 && S > s_max) S = s_max;
    if (S < 1) return false;
    const int64_t rows = (Nd + S - 1) / S;
    s.rows_per_group = (rows + RB - 1) / RB * RB;
    // rounding the rows up to whole blocks may leave trailing groups empty:
    // drop them unless the grid is XCD-aligned (there an empty group just
    // writes zero partials)
    if (per_xcd == 0) S = (Nd + s.rows_per_group - 1) / s.rows_per_group;
    s.S = (int)S;
    s.per_xcd = per_xcd;
    s.nrb = s.rows_per_group / RB;
    if (min_blocks > 0 && s.nrb < min_blocks) return false;
    s.grid = s.NT * s.S;
    *out = s;
    return true;
}

}  // namespace gra
