// Volterra displacement phase H = 2 pi g.u of straight dislocations
// (isotropic elasticity, reference dislocations.py:113-211), evaluated in
// FP64 on gfx950.  Three producers:
//   * the H grid (ny, Nz, Nx) of compute_H_grid / H_plane;
//   * the lane-interleaved e^{iH} table the fast Laue march streams;
//   * e^{iH(y, z_k)} per row and step for a line along x_lab.
// Segments live in constant memory (uniform across every wave, so every
// read is a scalar load); calls on one device are serialised by stream
// order, which is what makes one symbol per library enough.
#include <mutex>

#include "internal.hpp"

namespace dyf {

constexpr int kMaxSeg = 224;
constexpr int kMaxRows = 2048;
__constant__ SegParams c_segs[kMaxSeg];
__constant__ double c_yvals[kMaxRows];

// The log_tab table, built once per device by fill_log_table and copied
// into LDS by every workgroup that evaluates segments (a copy of 2 KB
// instead of 128 library logs per workgroup).
__device__ LogEntry g_logtab[kLogTab];

__global__ void log_table_init_kernel() { fill_log_table(g_logtab); }

__device__ __forceinline__ void load_log_table(LogEntry *tab) {
    for (int j = threadIdx.x; j < kLogTab; j += blockDim.x) tab[j] = g_logtab[j];
}

int ensure_log_table() {
    static std::mutex mu;
    static bool done[64] = {};
    int dev = 0;
    DYF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(DYF_EINVAL, "device %d", dev);
    std::lock_guard<std::mutex> lock(mu);
    if (done[dev]) return DYF_OK;
    // synchronous, so every stream sees the table from the first use on
    log_table_init_kernel<<<1, kLogTab, 0, nullptr>>>();
    DYF_LAUNCHED("log_table_init_kernel");
    DYF_HIP(hipStreamSynchronize(nullptr));
    done[dev] = true;
    return DYF_OK;
}

static int upload_segments(const double *segs, int nseg, const double *g,
                           double nu, int *kept, hipStream_t stream) {
    if (nseg < 0 || (nseg > 0 && !segs))
        return fail(DYF_EINVAL, "bad segment list (n=%d)", nseg);
    if (nseg > kMaxSeg)
        return fail(DYF_EINVAL, "%d segments exceed the limit of %d", nseg,
                    kMaxSeg);
    if (!g) return fail(DYF_EINVAL, "g_vec is NULL");
    if (int rc = ensure_log_table()) return rc;
    SegParams host[kMaxSeg];
    *kept = prepare_segments(segs, nseg, g, nu, host);
    // Synchronous: the host array is gone once this call returns, and a
    // later call must not overwrite the symbol before this one's kernels
    // ran (the copy is ordered after all prior work on the device).
    if (*kept > 0)
        DYF_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_segs), host,
                                  sizeof(SegParams) * (*kept), 0,
                                  hipMemcpyHostToDevice));
    (void)stream;
    return DYF_OK;
}

// ---- H grid: one thread per voxel, x fastest (coalesced stores) -------
template <typename OutT>
__global__ __launch_bounds__(256) void volterra_grid_kernel(
    OutT *__restrict__ H, int nseg,
    int Nx, int Nz, double dx, double dz, double a2) {
    const int ix = blockIdx.x * 256 + threadIdx.x;
    const int k = blockIdx.y;
    const int j = blockIdx.z;
    if (ix >= Nx) return;
    const double x = ((double)ix - (double)Nx / 2.0) * dx;
    const double z = (double)k * dz;
    const double y = c_yvals[j];
    OutT h;
    if constexpr (sizeof(OutT) == 8)
        h = volterra_h(c_segs, nseg, x, y, z, a2);
    else
        h = volterra_h_f32acc(c_segs, nseg, x, y, z, a2);
    H[((size_t)j * Nz + k) * Nx + ix] = h;
}

// ---- fused e^{iH} table for the fast march ---------------------------
// Thread t of a (row, step block) covers column x = t (0 <= x < C * 64,
// lane x / C, slot x % C) for kPhaseZ consecutive steps and writes that
// column's element of each of those (row, step) blocks.  A wavefront
// covers 64 consecutive columns, so the wavefronts wholly outside the
// beam's fan skip their evaluation as a unit.  The columns past Nx (the
// lane padding, which the last lane reads) are written as zeros.  Entries
// of a lane before its first step (lane_first_step) are left unwritten
// unless zero_pre_fan is set: every march except the persistent FP32-state
// march at C = 18 skips those loads (its fields there are exact zeros), and
// at a table of one segment the zero stores were a third of the kernel's
// time (200 um worked example, FP64: 0.82 -> 0.57 ms without them).
// The grid is flat over (row, step block, t): a tile's C * 64 threads are
// whole wavefronts, so no wavefront is idle.
constexpr int kPhaseZ = 4;

static dim3 phase_table_grid(const PhaseLayout &L) {
    const size_t nzb = (L.nsteps + kPhaseZ - 1) / kPhaseZ;
    const size_t n = (size_t)L.nrows * nzb * L.C * kWave;
    return dim3((unsigned)((n + 255) / 256));
}

template <typename T>
__device__ __forceinline__ void phase_table_points(
    cx<T> *__restrict__ E, const SegParams *__restrict__ segs, int nseg,
    int C, int Nx, int Ny, int nsteps, int nrows, int iy_lo, int fan_lo,
    int fan_hi, double dx, double dz, double a2, int zero_pre_fan) {
    __shared__ LogEntry tab[kLogTab];
    load_log_table(tab);
    __syncthreads();
    const int per = C * kWave, nzb = (nsteps + kPhaseZ - 1) / kPhaseZ;
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t tile = g / per;
    if (tile >= (size_t)nrows * nzb) return;
    const int x = (int)(g - tile * per);
    const int r = (int)(tile / nzb);
    const int k0 = (int)(tile - (size_t)r * nzb) * kPhaseZ;
    const int lane = x / C, i = x - lane * C;
    const int nk = min(kPhaseZ, nsteps - k0);
    const size_t stride = (size_t)C * kWave;
    cx<T> *out = E + ((size_t)r * nsteps + k0) * stride + (size_t)i * kWave + lane;
    // padding, or every step of this block before the lane's first step
    if (x >= Nx || k0 + nk - 1 < lane_first_step(lane, C, fan_lo, fan_hi)) {
        if (x < Nx && !zero_pre_fan) return;
#pragma unroll
        for (int j = 0; j < kPhaseZ; ++j)
            if (j < nk) out[j * stride] = cx<T>{T(0), T(0)};
        return;
    }
    const double xp = ((double)x - (double)Nx / 2.0) * dx;
    // y as compute_H_grid places it (dislocations.py:146): (iy - Ny/2) dx
    // with a real Ny/2, so an odd Ny puts rows half a pixel off the integer
    const double yp = ((double)(iy_lo + r) - (double)Ny * 0.5) * dx;
    double z[kPhaseZ], Pr[kPhaseZ], Pi[kPhaseZ], R[kPhaseZ];
#pragma unroll
    for (int j = 0; j < kPhaseZ; ++j) z[j] = (double)(k0 + j) * dz;
    volterra_phase_z<kPhaseZ>(segs, nseg, xp, yp, z, a2, tab, Pr, Pi, R);
#pragma unroll
    for (int j = 0; j < kPhaseZ; ++j) {
        if (j >= nk) break;
        if constexpr (sizeof(T) == 8)
            out[j * stride] = phase_parts_f64(Pr[j], Pi[j], R[j]);
        else
            out[j * stride] = phase_parts_f32(Pr[j], Pi[j], R[j]);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void phase_from_segments_kernel(
    cx<T> *__restrict__ E, int nseg, int C, int Nx, int Ny, int nsteps,
    int nrows, int iy_lo, int fan_lo, int fan_hi, double dx, double dz,
    double a2) {
    // read only by the stand-alone marches, which skip pre-fan entries
    phase_table_points<T>(E, c_segs, nseg, C, Nx, Ny, nsteps, nrows, iy_lo,
                          fan_lo, fan_hi, dx, dz, a2, 0);
}

// Same table with the segments read from device memory (uniform loads),
// so a batch of instances can be enqueued - and graph-captured - without
// touching the constant symbol.
template <typename T>
__global__ __launch_bounds__(256) void phase_segments_dev_kernel(
    cx<T> *__restrict__ E, const SegParams *__restrict__ segs, int nseg,
    int C, int Nx, int Ny, int nsteps, int nrows, int iy_lo, int fan_lo,
    int fan_hi, double dx, double dz, double a2, int zero_pre_fan) {
    phase_table_points<T>(E, segs, nseg, C, Nx, Ny, nsteps, nrows, iy_lo,
                          fan_lo, fan_hi, dx, dz, a2, zero_pre_fan);
}

int launch_phase_segments_dev(const dyf_laue_args *a, const PhaseLayout &L,
                              int prec, const SegParams *segs, int nseg,
                              double a2, void *phase, hipStream_t stream) {
    if (L.nrows == 0) return DYF_OK;
    const dim3 grid = phase_table_grid(L);
    // the persistent FP32-state march at C = 18 reads its slices from step
    // 0 on (selecting the offset spills there), so that table keeps zeros
    // before each lane's first step; every other march skips those loads
    const int zero_pre_fan = prec != DYF_FP64 && L.C >= 18;
    if (prec == DYF_FP64)
        phase_segments_dev_kernel<double><<<grid, 256, 0, stream>>>(
            (cx<double> *)phase, segs, nseg, L.C, a->Nx, a->Ny, L.nsteps,
            L.nrows, L.iy_lo, L.fan_lo, L.fan_hi, a->dx, a->dz, a2,
            zero_pre_fan);
    else
        phase_segments_dev_kernel<float><<<grid, 256, 0, stream>>>(
            (cx<float> *)phase, segs, nseg, L.C, a->Nx, a->Ny, L.nsteps,
            L.nrows, L.iy_lo, L.fan_lo, L.fan_hi, a->dx, a->dz, a2,
            zero_pre_fan);
    DYF_LAUNCHED("phase_segments_dev_kernel");
    return DYF_OK;
}

// ---- e^{iH(y, z_k)} for one line along x_lab (the analytic kernel's
// in-loop phase, laue/kernels.py:216-229), one value per (row, step) ---
template <typename T>
__global__ __launch_bounds__(256) void phase_analytic_kernel(
    cx<T> *__restrict__ E, dyf_line_params lp, int Ny, int nsteps,
    int nrows, int iy_lo, double dx, double dz) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= nrows * nsteps) return;
    const int r = idx / nsteps, k = idx % nsteps;
    // the closed-form line's row coordinate uses the integer Ny/2 of the
    // reference's analytic kernel (laue/kernels.py:172); the host takes
    // this path only for even Ny, where both conventions agree
    const double dy = (double)(iy_lo + r - Ny / 2) * dx - lp.y0;
    const double dzv = (double)k * dz - lp.z0;
    const double d1 = dy * lp.e1y + dzv * lp.e1z;
    const double d2 = dy * lp.e2y + dzv * lp.e2z;
    const double r2 = d1 * d1 + d2 * d2;
    const double r2c = r2 > lp.a2 ? r2 : lp.a2;
    double H = lp.c_atan * atan2(d2, d1);
    if (lp.b_perp > 0.0) {
        const double onu = 1.0 - lp.nu, c1 = lp.b_perp / kTwoPi;
        const double ux = c1 * d1 * d2 / (2.0 * onu * r2c);
        const double uy = -c1 * ((1.0 - 2.0 * lp.nu) / (4.0 * onu) * log(r2c) +
                                 (d1 * d1 - d2 * d2) / (4.0 * onu * r2c));
        H += kTwoPi * (lp.ge1 * ux + lp.ge2 * uy);
    }
    if constexpr (sizeof(T) == 8)
        E[idx] = phase_f64(H);
    else
        E[idx] = phase_f32(H);
}

// ---- e^{iH} table from a stored H grid (the H_gpu= path) --------------
template <typename T, typename HT>
__global__ __launch_bounds__(256) void phase_from_h_kernel(
    cx<T> *__restrict__ E, const HT *__restrict__ H, int C, int Nx, int Nz,
    int nsteps, int iy_lo) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int k = blockIdx.y;
    const int r = blockIdx.z;
    if (t >= C * kWave) return;
    const int i = t / kWave, lane = t % kWave;
    const int x = lane * C + i;
    cx<T> e{T(0), T(0)};
    if (x < Nx) {
        const double h = (double)H[((size_t)(iy_lo + r) * Nz + k) * Nx + x];
        if constexpr (sizeof(T) == 8)
            e = phase_f64(h);
        else
            e = phase_f32(h);
    }
    E[((size_t)r * nsteps + k) * (size_t)(C * kWave) + t] = e;
}

// ---- e^{iH} on Bragg chunk planes, (Nz, ny, Nx) ----------------------
__global__ __launch_bounds__(256) void bragg_phase_kernel(
    cx<double> *__restrict__ E, int nseg, int Nx, int ny, double dx,
    double dz, double a2) {
    __shared__ LogEntry tab[kLogTab];
    load_log_table(tab);
    __syncthreads();
    const int ix = blockIdx.x * 256 + threadIdx.x;
    const int k = blockIdx.y, j = blockIdx.z;
    if (ix >= Nx) return;
    const double x = ((double)ix - (double)Nx / 2.0) * dx;
    double Pr, Pi, R;
    volterra_phase(c_segs, nseg, x, c_yvals[j], (double)k * dz, a2, tab, Pr,
                   Pi, R);
    E[((size_t)k * ny + j) * Nx + ix] = phase_parts_f64(Pr, Pi, R);
}

}  // namespace dyf

using namespace dyf;

extern "C" int dyf_bragg_phase(const double *segs, int nseg, const double *g,
                               double nu, double a_core, const dyf_volume *vol,
                               const double *y_values, int ny, void *E,
                               void *stream_) {
    hipStream_t stream = (hipStream_t)stream_;
    if (!vol || vol->Nx <= 0 || vol->Nz <= 0 || vol->Nz > 65535)
        return fail(DYF_EINVAL, "bad volume");
    if (ny < 1 || ny > kMaxRows || !y_values || !E)
        return fail(DYF_EINVAL, "bad planes (ny=%d, limit %d)", ny, kMaxRows);
    int kept = 0;
    int rc = upload_segments(segs, nseg, g, nu, &kept, stream);
    if (rc) return rc;
    DYF_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_yvals), y_values,
                              sizeof(double) * ny, 0, hipMemcpyHostToDevice));
    dim3 grid((vol->Nx + 255) / 256, vol->Nz, ny);
    bragg_phase_kernel<<<grid, 256, 0, stream>>>(
        (cx<double> *)E, kept, vol->Nx, ny, vol->dx, vol->dz,
        a_core * a_core);
    DYF_LAUNCHED("bragg_phase_kernel");
    return DYF_OK;
}

extern "C" int dyf_volterra_h(const double *segs, int nseg, const double *g,
                              double nu, double a_core, const dyf_volume *vol,
                              const double *y_values, int ny, int out_prec,
                              void *H, void *stream_) {
    hipStream_t stream = (hipStream_t)stream_;
    if (!vol || vol->Nx <= 0 || vol->Nz <= 0)
        return fail(DYF_EINVAL, "bad volume");
    if (ny < 0 || (ny > 0 && (!y_values || !H)))
        return fail(DYF_EINVAL, "bad rows (ny=%d)", ny);
    if (out_prec != DYF_FP64 && out_prec != DYF_FP32)
        return fail(DYF_EKEY, "H precision must be DYF_FP64 or DYF_FP32");
    if (ny == 0) return DYF_OK;
    int kept = 0;
    int rc = upload_segments(segs, nseg, g, nu, &kept, stream);
    if (rc) return rc;
    if (vol->Nz > 65535)
        return fail(DYF_EINVAL, "grid too large for one launch");
    const double a2 = a_core * a_core;
    // rows go through constant memory in blocks of kMaxRows
    for (int j0 = 0; j0 < ny; j0 += kMaxRows) {
        const int nb = std::min(kMaxRows, ny - j0);
        DYF_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_yvals), y_values + j0,
                                  sizeof(double) * nb, 0,
                                  hipMemcpyHostToDevice));
        dim3 grid((vol->Nx + 255) / 256, vol->Nz, nb);
        const size_t off = (size_t)j0 * vol->Nz * vol->Nx;
        if (out_prec == DYF_FP64)
            volterra_grid_kernel<double><<<grid, 256, 0, stream>>>(
                (double *)H + off, kept, vol->Nx, vol->Nz, vol->dx, vol->dz, a2);
        else
            volterra_grid_kernel<float><<<grid, 256, 0, stream>>>(
                (float *)H + off, kept, vol->Nx, vol->Nz, vol->dx, vol->dz, a2);
        DYF_LAUNCHED("volterra_grid_kernel");
    }
    return DYF_OK;
}

extern "C" int dyf_laue_phase_from_segments(const dyf_laue_args *a, int prec,
                                            const double *segs, int nseg,
                                            const double *g, double nu,
                                            double a_core, void *phase,
                                            void *stream_) {
    hipStream_t stream = (hipStream_t)stream_;
    PhaseLayout L;
    int rc = phase_layout(a, DYF_MODE_GRID, prec, &L);
    if (rc) return rc;
    if (L.path != DYF_PATH_FAST)
        return fail(DYF_EINVAL, "grid takes the general path; build an H grid");
    if (L.nrows == 0) return DYF_OK;
    if (!phase) return fail(DYF_EINVAL, "phase buffer is NULL");
    int kept = 0;
    rc = upload_segments(segs, nseg, g, nu, &kept, stream);
    if (rc) return rc;
    const dim3 grid = phase_table_grid(L);
    const double a2 = a_core * a_core;
    if (prec == DYF_FP64)
        phase_from_segments_kernel<double><<<grid, 256, 0, stream>>>(
            (cx<double> *)phase, kept, L.C, a->Nx, a->Ny, L.nsteps, L.nrows,
            L.iy_lo, L.fan_lo, L.fan_hi, a->dx, a->dz, a2);
    else
        phase_from_segments_kernel<float><<<grid, 256, 0, stream>>>(
            (cx<float> *)phase, kept, L.C, a->Nx, a->Ny, L.nsteps, L.nrows,
            L.iy_lo, L.fan_lo, L.fan_hi, a->dx, a->dz, a2);
    DYF_LAUNCHED("phase_from_segments_kernel");
    return DYF_OK;
}

extern "C" int dyf_laue_phase_analytic(const dyf_laue_args *a, int prec,
                                       const dyf_line_params *line,
                                       void *phase, void *stream_) {
    hipStream_t stream = (hipStream_t)stream_;
    PhaseLayout L;
    int rc = phase_layout(a, DYF_MODE_ANALYTIC, prec, &L);
    if (rc) return rc;
    if (!line) return fail(DYF_EINVAL, "line parameters are NULL");
    if (L.path != DYF_PATH_FAST)
        return fail(DYF_EINVAL, "grid takes the general path");
    const int n = L.nrows * L.nsteps;
    if (n == 0) return DYF_OK;
    if (!phase) return fail(DYF_EINVAL, "phase buffer is NULL");
    if (prec == DYF_FP64)
        phase_analytic_kernel<double><<<(n + 255) / 256, 256, 0, stream>>>(
            (cx<double> *)phase, *line, a->Ny, L.nsteps, L.nrows, L.iy_lo,
            a->dx, a->dz);
    else
        phase_analytic_kernel<float><<<(n + 255) / 256, 256, 0, stream>>>(
            (cx<float> *)phase, *line, a->Ny, L.nsteps, L.nrows, L.iy_lo,
            a->dx, a->dz);
    DYF_LAUNCHED("phase_analytic_kernel");
    return DYF_OK;
}

extern "C" int dyf_laue_phase_from_h(const dyf_laue_args *a, int prec,
                                     const void *H, int H_prec, void *phase,
                                     void *stream_) {
    hipStream_t stream = (hipStream_t)stream_;
    PhaseLayout L;
    int rc = phase_layout(a, DYF_MODE_GRID, prec, &L);
    if (rc) return rc;
    if (L.path != DYF_PATH_FAST)
        return fail(DYF_EINVAL, "grid takes the general path; march H directly");
    if (H_prec != DYF_FP64 && H_prec != DYF_FP32)
        return fail(DYF_EKEY, "H precision must be DYF_FP64 or DYF_FP32");
    if (L.nrows == 0) return DYF_OK;
    if (!H || !phase) return fail(DYF_EINVAL, "NULL H or phase buffer");
    dim3 grid((L.C * kWave + 255) / 256, L.nsteps, L.nrows);
#define DYF_PFH(T, HT)                                                      \
    phase_from_h_kernel<T, HT><<<grid, 256, 0, stream>>>(                  \
        (cx<T> *)phase, (const HT *)H, L.C, a->Nx, a->Nz, L.nsteps, L.iy_lo)
    if (prec == DYF_FP64 && H_prec == DYF_FP64) DYF_PFH(double, double);
    else if (prec == DYF_FP64) DYF_PFH(double, float);
    else if (H_prec == DYF_FP64) DYF_PFH(float, double);
    else DYF_PFH(float, float);
#undef DYF_PFH
    DYF_LAUNCHED("phase_from_h_kernel");
    return DYF_OK;
}
