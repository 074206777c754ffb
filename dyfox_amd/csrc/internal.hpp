// Launch helpers shared between translation units of libdyfox (not part
// of the C ABI).
#pragma once

#include "common.hpp"
#include "phase_layout.hpp"

namespace dyf {

// e^{iH} table of one Laue instance from segments already on the device
// (the batched path; graph-capturable: kernels only).
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
    double dx, sheet_hw;
    long long item_base, n_items;
};

// Persistent march over `n_inst` instances that share column count C: every wave takes (row, angle)
// items from 8 XCD-local queues (stealing when its own is empty) until
// all are drained.  counters: 8 unsigned ints, zeroed by this call.
int launch_laue_persistent(const InstDesc *inst_dev, int n_inst,
                           long long total_items, int prec, int C,
                           unsigned *counters, int waves_per_block,
                           int blocks, hipStream_t stream);

}  // namespace dyf
