// Launch helpers shared between translation units of libdyfox (not part
// of the C ABI).
#pragma once

#include "common.hpp"
#include "phase_layout.hpp"

namespace dyf {

// Builds the device-wide table of log_tab once per device (a kernel and a
// synchronisation on first use); call before enqueueing any kernel that
// evaluates segments, outside stream capture.
int ensure_log_table();

// e^{iH} table of one Laue instance from segments already on the device
// (the batched path; graph-capturable: kernels only; ensure_log_table
// must have run on the device).
int launch_phase_segments_dev(const dyf_laue_args *a, const PhaseLayout &L,
                              int prec, const SegParams *segs_dev, int nseg,
                              double a2, void *phase, hipStream_t stream);

// Interaction-picture coefficients of n_e angles into `coef`.
int launch_laue_coef(const dyf_laue_args *a, const double *s_dev, int n_e,
                     void *coef, hipStream_t stream);
size_t laue_coef_bytes(int n_e, int nsteps);

// Zero the rows of `out` outside the marched band.
int launch_zero_rows(const dyf_laue_args *a, const PhaseLayout &L, int prec,
                     int n_e, void *out, hipStream_t stream);

// Device descriptor of one instance of the persistent march.
struct InstDesc {
    const void *phase;
    const void *coef;
    void *out;
    int n_e, nrows, iy_lo, nsteps, Nx, Ny, C, pad_;
    int fan_lo, fan_hi;   // columns the incident beam lights
    int pair_K, pair_g0;  // pair march (C <= pair_max_c): steps with two angles per
                          // wavefront and their lane window (K = 0: none)
    double dx, sheet_hw;
    long long item_base, n_items;
};

// The pair march's lane window [g0, g0 + 32) for a layout and its step K:
// the largest K before which every lane outside the window is still
// before its first step in the beam's fan.
void laue_pair_window(const PhaseLayout &L, int *K, int *g0);

// The largest column count whose persistent march takes (row, pair of
// angles) items: beyond it the pair march spills registers inside its
// step loops (FP64 at C = 17, every precision at C = 18).
constexpr int pair_max_c(int prec) { return prec == DYF_FP64 ? 16 : 17; }

// Persistent march over `n_inst` instances that share column count C: every wave takes (row, angle)
// items (C <= pair_max_c(prec): (row, pair of angles) items, marched as
// fast_row_pair does)
// items from 8 XCD-local queues (stealing when its own is empty) until
// all are drained.  counters: 8 unsigned ints, zeroed by this call.
int launch_laue_persistent(const InstDesc *inst_dev, int n_inst,
                           long long total_items, int prec, int C,
                           unsigned *counters, int waves_per_block,
                           int blocks, hipStream_t stream);

// ---- The Bragg preconditioner's transforms (fft.hip) -------------------
constexpr int kFftMaxPasses = 16;

// How rows of Nx points are transformed: at length N = Nx (conv = 0,
// every prime factor of Nx <= 13) or as a linear convolution at a
// 13-smooth N >= 2 Nx - 1 (conv = 1); radix[0..nst) are the passes.
struct FftPlan {
    int N, conv, nst;
    int radix[kFftMaxPasses];
};
int fft_plan_for(int Nx, FftPlan *p);

// Complex elements per system of the preconditioner symbol: the
// 1 / (T22s Nx) row (Nx), then for convolution plans the kernel's
// spectrum (N).
size_t fft_symbol_elems(const FftPlan &p, int Nx);

// For convolution plans: the kernel spectrum of every system from its
// 1 / (T22s Nx) row (symbol layout above; nothing to do otherwise).
int fft_prepare_spectrum(const FftPlan &p, int Nx, cx<double> *symbol,
                         int nsys, hipStream_t stream);

// w = ifft(fft(v) / T22s) on nsys * rows_per_sys rows of Nx points (row r
// belongs to system r / rows_per_sys); v and w may be the same buffer.
int fft_precondition(const FftPlan &p, int Nx, const cx<double> *symbol,
                     int nsys, int rows_per_sys, const cx<double> *v,
                     cx<double> *w, hipStream_t stream);

}  // namespace dyf
