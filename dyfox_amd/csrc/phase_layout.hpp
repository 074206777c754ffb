// How a Laue grid maps onto the march kernels, shared by the phase-table
// producers and the march itself so both agree on the table layout.
//
// Fast path (exact one-pixel shift): one wavefront per (row, s_dev); lane
// l holds columns [l*C, l*C + C) in registers.  The e^{iH} table of a row
// and step is stored lane-interleaved, element (i, lane) at i*64 + lane,
// so each of the C loads per step is one contiguous, coalesced 64-wide
// access.  Columns x >= Nx hold e^{iH} = 0, which keeps the padding
// columns from coupling back into the grid.
#pragma once

#include "common.hpp"

namespace dyf {

// Column counts per lane that have compiled fast kernels, per precision
// family (register budget: FP64 state + phase prefetch ~12 VGPRs/column,
// two-float state ~10 VGPRs/column).
constexpr int kFastCols[] = {2, 4, 6, 8, 10, 12, 13, 14, 16, 17, 18};
constexpr int kNumFastCols = sizeof(kFastCols) / sizeof(kFastCols[0]);

struct PhaseLayout {
    int path;       // DYF_PATH_FAST or DYF_PATH_GENERAL
    int C;          // columns per lane (fast path)
    int iy_lo, iy_hi, nrows;
    int nsteps;     // Nz - 1
    int fan_lo, fan_hi;   // columns where the incident beam can be non-zero
    size_t bytes;   // phase buffer size for (mode, prec)
};

// Columns [lo, hi] where the march's incident profile (the top-hat
// |(x - Nx/2) dx| < sheet_hw, or the Gaussian of sigma -sheet_hw) can be
// non-zero, evaluated with the device's own expression.  A Gaussian is
// exactly zero once (x/sigma)^2 / 2 > 745 (exp underflows); 800 is used.
// A degenerate profile (sheet_hw = 0, or no column lit) gives the whole
// row.
inline void beam_support(const dyf_laue_args *a, int *lo, int *hi) {
    int l = a->Nx, h = -1;
    const double hw = a->sheet_hw;
    for (int x = 0; x < a->Nx; ++x) {
        const double xp = (double)(x - a->Nx / 2) * a->dx;
        const bool lit = hw > 0.0 ? std::fabs(xp) < hw
                                  : xp * xp < 1600.0 * hw * hw;
        if (lit) l = x < l ? x : l, h = x > h ? x : h;
    }
    if (hw == 0.0 || l > h) l = 0, h = a->Nx - 1;
    *lo = l, *hi = h;
}

// The march's fields start non-zero only on columns [fan_lo, fan_hi] and
// spread one column per step each way (X moves left, Y right), so before
// step k they vanish outside [fan_lo - k, fan_hi + k] and the step reads
// e^{iH} only on [fan_lo - k - 1, fan_hi + k + 1]; outside, it multiplies
// exact zeros.  Lane `lane` (columns [lane C, lane C + C)) needs its
// e^{iH} from this step on (one column of margin each way); the table
// producers skip the entries before it and the marches do not load them.
__host__ __device__ inline int lane_first_step(int lane, int C, int fan_lo,
                                               int fan_hi) {
    const int x0 = lane * C, x1 = x0 + C - 1;
    const int kl = fan_lo - 2 - x1;   // lanes left of the fan
    const int kr = x0 - fan_hi - 2;   // lanes right of it
    const int k = kl > kr ? kl : kr;
    return k > 0 ? k : 0;
}

inline int fast_cols_for(int Nx) {
    const int need = (Nx + kWave - 1) / kWave;
    for (int j = 0; j < kNumFastCols; ++j)
        if (kFastCols[j] >= need) return kFastCols[j];
    return 0;
}

inline int check_laue_args(const dyf_laue_args *a) {
    if (!a) return fail(DYF_EINVAL, "laue args are NULL");
    if (a->Nz < 2 || a->Nx < 1 || a->Ny < 1)
        return fail(DYF_EINVAL, "bad grid Nz=%d Nx=%d Ny=%d", a->Nz, a->Nx,
                    a->Ny);
    if (!(a->ds > 0.0) || !(a->dx > 0.0))
        return fail(DYF_EINVAL, "ds and dx must be positive");
    return DYF_OK;
}

inline int phase_layout(const dyf_laue_args *a, int mode, int prec,
                        PhaseLayout *L) {
    int rc = check_laue_args(a);
    if (rc) return rc;
    if (prec != DYF_FP64 && prec != DYF_FP32 && prec != DYF_MIXED)
        return fail(DYF_EKEY, "unknown precision %d", prec);
    if (mode != DYF_MODE_GRID && mode != DYF_MODE_ANALYTIC &&
        mode != DYF_MODE_PRISTINE)
        return fail(DYF_EKEY, "unknown march mode %d", mode);
    const RowBand band = laue_rows(a);
    L->iy_lo = band.iy_lo;
    L->iy_hi = band.iy_hi;
    L->nrows = band.iy_hi - band.iy_lo;
    L->nsteps = a->Nz - 1;
    L->C = fast_cols_for(a->Nx);
    beam_support(a, &L->fan_lo, &L->fan_hi);
    // The interaction-picture march divides by the diagonal propagator;
    // keep it well conditioned: |sigma_h| ds small (always true on exact
    // grids at sub-micron pitch).
    // The grid-mode march addresses one row's table with 32-bit byte
    // offsets (a buffer resource per row): nsteps * C * 64 entries must
    // stay below 2^31 bytes (Nz up to ~116,000 at C = 18 in FP64).
    const double sh = std::sqrt(a->shr * a->shr + a->shi * a->shi);
    const size_t elem = prec == DYF_FP64 ? 16 : 8;
    const bool row_fits =
        (size_t)L->nsteps * (size_t)(L->C > 0 ? L->C : 1) * kWave * elem <
        ((size_t)1 << 31);
    const bool fast =
        a->shift_px == 1.0 && L->C > 0 && sh * a->ds < 0.25 && row_fits;
    L->path = fast ? DYF_PATH_FAST : DYF_PATH_GENERAL;
    L->bytes = 0;
    if (fast && mode == DYF_MODE_GRID)
        L->bytes = (size_t)L->nrows * L->nsteps * L->C * kWave * elem;
    else if (fast && mode == DYF_MODE_ANALYTIC)
        L->bytes = (size_t)L->nrows * L->nsteps * elem;
    return DYF_OK;
}

}  // namespace dyf
