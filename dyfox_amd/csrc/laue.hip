// Laue (transmission) depth march on gfx950.
//
// The reference advances, per (s_dev, row), two amplitude rows by
//     D0'[x] = M11 D0[x+1] + M12[x] Dg[x-1]
//     Dg'[x] = M21[x] D0[x+1] + M22 Dg[x-1]
// with M11 = ep (cg - i b), M22 = ep (cg + i b), M12 = ep sg sigma_hbar
// e^{+iH}, M21 = ep sg sigma_h e^{-iH} (laue/kernels.py:90-139).
//
// Fast path.  M11 and M22 are the same at every column, so the march is
// carried in the interaction picture X_k = D0_k / M11^k, Y_k = Dg_k / M22^k:
//     X'[x] = X[x+1] + c_k e^{+iH}[x] Y[x-1]
//     Y'[x] = Y[x-1] + b_k e^{-iH}[x] X[x+1]
//     c_k = sg sigma_hbar/(cg - i b) e^{+2ik phi},
//     b_k = sg sigma_h  /(cg + i b) e^{-2ik phi},  phi = atan2(b, cg)
// and Dg = M22^(Nz-1) Y at the end.  That is 16 real FMAs per
// pixel-step instead of 24, and every step is a small correction to the
// state (|c_k|, |b_k| <= |sigma| ds), which is what lets the mixed mode
// carry the state as an FP32 value plus an FP32 compensation term.
//
// One wavefront owns one (row, s_dev); lane l keeps columns
// [l*C, l*C + C) of both fields in VGPRs, and the one-pixel transport is a
// register rename inside the lane plus one cross-lane move per field per
// step.  e^{iH} streams from the lane-interleaved table (coalesced), the
// next step's slice prefetched while the current one is used.
//
// General path (any shift, any Nx): the reference's direct form, FP64, one
// workgroup per (row, s_dev) with the rows double-buffered in LDS.
#include <mutex>

#include "internal.hpp"

namespace dyf {

struct CoefD {
    double cr, ci, br, bi;
};
struct CoefF {
    float cr, ci, br, bi;
};

// The per-step coefficients are uniform per wave and read-only: read
// through the constant address space they become scalar loads.  As
// vector loads (the compiler cannot rule out that they alias the output)
// they were issued after the prefetched e^{iH} slice, so waiting for them
// waited for the whole slice too (the vector-memory counter drains in
// issue order).
template <typename T>
using kptr = const T __attribute__((address_space(4))) *;

// Per-call workspace: [CoefD n_e*nsteps][CoefF n_e*nsteps][qN 2*n_e doubles]
static size_t coef_bytes(int n_e, int nsteps) {
    const size_t a = (size_t)n_e * nsteps * sizeof(CoefD);
    const size_t b = (size_t)n_e * nsteps * sizeof(CoefF);
    return a + b + (size_t)n_e * 2 * sizeof(double);
}

struct CoefPtrs {
    CoefD *d;
    CoefF *f;
    double *qN;
};
__host__ __device__ inline CoefPtrs coef_ptrs(void *ws, int n_e, int nsteps) {
    char *p = (char *)ws;
    CoefPtrs c;
    c.d = (CoefD *)p;
    c.f = (CoefF *)(p + (size_t)n_e * nsteps * sizeof(CoefD));
    c.qN = (double *)(p + (size_t)n_e * nsteps * (sizeof(CoefD) + sizeof(CoefF)));
    return c;
}

// One thread per (s_dev, step): the interaction-picture coefficients.
__global__ __launch_bounds__(256) void laue_coef_kernel(
    const double *__restrict__ s_dev, int n_e, int nsteps, double ds,
    double s0r, double s0i, double shr, double shi, double sbr, double sbi,
    CoefD *__restrict__ cd, CoefF *__restrict__ cf, double *__restrict__ qN) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= n_e * nsteps) return;
    const int ie = idx / nsteps, k = idx % nsteps;
    const double alpha = -2.0 * kPi * s_dev[ie];
    const double gam = sqrt(shr * shr + shi * shi + alpha * alpha * 0.25);
    double cg, sg;
    if (gam > 1e-30) {
        cg = cos(gam * ds);
        sg = sin(gam * ds) / gam;
    } else {
        cg = 1.0;
        sg = ds;
    }
    const double b = alpha * 0.5 * sg;
    const double rho2 = cg * cg + b * b;
    const double phi = atan2(b, cg);
    // c0 = sg sbar (cg + i b) / rho2 ; b0 = sg sh (cg - i b) / rho2
    const double c0r = sg * (sbr * cg - sbi * b) / rho2;
    const double c0i = sg * (sbi * cg + sbr * b) / rho2;
    const double b0r = sg * (shr * cg + shi * b) / rho2;
    const double b0i = sg * (shi * cg - shr * b) / rho2;
    double sn, cs;
    sincos(2.0 * (double)k * phi, &sn, &cs);
    CoefD c;
    c.cr = c0r * cs - c0i * sn;
    c.ci = c0r * sn + c0i * cs;
    c.br = b0r * cs + b0i * sn;
    c.bi = b0i * cs - b0r * sn;
    cd[idx] = c;
    cf[idx] = CoefF{(float)c.cr, (float)c.ci, (float)c.br, (float)c.bi};
    if (k == 0) {
        const double N = (double)nsteps;
        const double mag = exp(N * ds * s0r + 0.5 * N * log(rho2));
        const double ang = N * ds * (s0i + alpha * 0.5) + N * phi;
        double s2, c2;
        sincos(ang, &s2, &c2);
        qN[2 * ie] = mag * c2;
        qN[2 * ie + 1] = mag * s2;
    }
}

// ---- cross-lane transport --------------------------------------------
// value of lane+1 (0 on the last lane: nothing flows in past x = Nx)
__device__ __forceinline__ float from_right(float v, int lane) {
    const int w = __builtin_amdgcn_ds_bpermute((lane + 1) << 2,
                                               __float_as_int(v));
    return lane == kWave - 1 ? 0.0f : __int_as_float(w);
}
// value of lane-1 (0 on lane 0: nothing flows in before x = 0)
__device__ __forceinline__ float from_left(float v, int lane) {
    const int w = __builtin_amdgcn_ds_bpermute((lane - 1) << 2,
                                               __float_as_int(v));
    return lane == 0 ? 0.0f : __int_as_float(w);
}
__device__ __forceinline__ double from_right(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_bpermute((lane + 1) << 2, (int)b);
    const int hi = __builtin_amdgcn_ds_bpermute((lane + 1) << 2, (int)(b >> 32));
    const double r = __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
    return lane == kWave - 1 ? 0.0 : r;
}
__device__ __forceinline__ double from_left(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_bpermute((lane - 1) << 2, (int)b);
    const int hi = __builtin_amdgcn_ds_bpermute((lane - 1) << 2, (int)(b >> 32));
    const double r = __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
    return lane == 0 ? 0.0 : r;
}

__device__ __forceinline__ double beam_profile(int x, int Nx, double dx,
                                               double sheet_hw) {
    const double xp = (double)(x - Nx / 2) * dx;
    return sheet_hw > 0.0 ? (fabs(xp) < sheet_hw ? 1.0 : 0.0)
                          : exp(-xp * xp / (2.0 * sheet_hw * sheet_hw));
}

struct FastArgs {
    const void *phase;   // GRID: lane-interleaved table; ANALYTIC: (nrows, nsteps)
    const CoefD *cd;
    const CoefF *cf;
    const double *qN;
    void *out;           // (n_e, Ny, Nx) complex
    int n_e, nrows, iy_lo, nsteps, Nx, Ny, per_xcd, total;
    int fan_lo, fan_hi;  // columns the incident beam lights (PhaseLayout)
    double dx, sheet_hw;
};

// Block b runs on XCD b % 8; give each XCD a contiguous range of (row,
// s_dev) pairs so that every s_dev of a row reads its e^{iH} rows through
// the same L2.
__device__ __forceinline__ int logical_block(int per_xcd) {
    const int b = blockIdx.x;
    return (b & 7) * per_xcd + (b >> 3);
}

// State element: plain (FP64 or FP32) or two-float compensated (MIXED).
template <int PREC>
struct StateT;
template <>
struct StateT<DYF_FP64> {
    using T = double;
    static constexpr bool comp = false;
};
template <>
struct StateT<DYF_FP32> {
    using T = float;
    static constexpr bool comp = false;
};
template <>
struct StateT<DYF_MIXED> {
    using T = float;
    static constexpr bool comp = true;
};

template <typename T>
__device__ __forceinline__ T fmaT(T a, T b, T c) {
    if constexpr (sizeof(T) == 8)
        return __builtin_fma(a, b, c);
    else
        return __builtin_fmaf(a, b, c);
}

// March one (row r, angle ie) with the calling wavefront (64 lanes):
// the analytic mode (one e^{iH} per row and step, folded into the step's
// coefficients) and the pristine mode.  Grid-mode marches, which stream a
// per-column e^{iH} table, are fast_row_inplace and fast_row_unrolled.
template <int C, int MODE, int PREC>
__device__ __forceinline__ void fast_row(const FastArgs &a, int r, int ie) {
    static_assert(MODE != DYF_MODE_GRID, "grid mode has its own marches");
    using S = StateT<PREC>;
    using T = typename S::T;
    constexpr bool COMP = S::comp;
    const int lane = threadIdx.x & (kWave - 1);
    const int iy = a.iy_lo + r;

    // fields: hi parts (and compensations for MIXED)
    T Xr[C], Xi[C], Yr[C], Yi[C];
    T Xlr[COMP ? C : 1], Xli[COMP ? C : 1], Ylr[COMP ? C : 1], Yli[COMP ? C : 1];
    T msk[C];
#pragma unroll
    for (int i = 0; i < C; ++i) {
        const int x = lane * C + i;
        Xr[i] = x < a.Nx ? (T)beam_profile(x, a.Nx, a.dx, a.sheet_hw) : T(0);
        Xi[i] = Yr[i] = Yi[i] = T(0);
        if constexpr (COMP) Xlr[i] = Xli[i] = Ylr[i] = Yli[i] = T(0);
        msk[i] = x < a.Nx ? T(1) : T(0);
    }

    const cx<T> *Erow = nullptr;
    if constexpr (MODE == DYF_MODE_ANALYTIC)
        Erow = (const cx<T> *)a.phase + (size_t)r * a.nsteps;

    for (int k = 0; k < a.nsteps; ++k) {
        T cr, ci, br, bi;
        const size_t ci_ = (size_t)__builtin_amdgcn_readfirstlane(ie) * a.nsteps + k;
        if constexpr (sizeof(T) == 8) {
            kptr<CoefD> c = (kptr<CoefD>)a.cd + ci_;
            cr = c->cr, ci = c->ci, br = c->br, bi = c->bi;
        } else {
            kptr<CoefF> c = (kptr<CoefF>)a.cf + ci_;
            cr = c->cr, ci = c->ci, br = c->br, bi = c->bi;
        }
        if constexpr (MODE == DYF_MODE_ANALYTIC) {
            // fold the row-uniform e^{iH_k} into the coefficients
            const cx<T> e = Erow[k];
            const T ncr = cr * e.r - ci * e.i, nci = cr * e.i + ci * e.r;
            const T nbr = br * e.r + bi * e.i, nbi = bi * e.r - br * e.i;
            cr = ncr, ci = nci, br = nbr, bi = nbi;
        }
        // inflow across lane boundaries (old values)
        const T xin_r = from_right(Xr[0], lane), xin_i = from_right(Xi[0], lane);
        const T yin_r = from_left(Yr[C - 1], lane), yin_i = from_left(Yi[C - 1], lane);
        T xinl_r = T(0), xinl_i = T(0), yinl_r = T(0), yinl_i = T(0);
        if constexpr (COMP) {
            xinl_r = from_right(Xlr[0], lane), xinl_i = from_right(Xli[0], lane);
            yinl_r = from_left(Ylr[C - 1], lane), yinl_i = from_left(Yli[C - 1], lane);
        }
        T nXr[C], nXi[C], nYr[C], nYi[C];
        T nXlr[COMP ? C : 1], nXli[COMP ? C : 1], nYlr[COMP ? C : 1], nYli[COMP ? C : 1];
#pragma unroll
        for (int i = 0; i < C; ++i) {
            // sources: X from x+1, Y from x-1
            const T xr = i + 1 < C ? Xr[i + 1] : xin_r;
            const T xi = i + 1 < C ? Xi[i + 1] : xin_i;
            const T yr = i > 0 ? Yr[i - 1] : yin_r;
            const T yi = i > 0 ? Yi[i - 1] : yin_i;
            // the row-uniform e^{iH} is in the coefficients; the mask
            // keeps the padding columns from coupling back
            const T tr = msk[i] * yr, ti = msk[i] * yi;
            const T ur = msk[i] * xr, ui = msk[i] * xi;
            if constexpr (!COMP) {
                nXr[i] = fmaT(cr, tr, fmaT(-ci, ti, xr));
                nXi[i] = fmaT(cr, ti, fmaT(ci, tr, xi));
                nYr[i] = fmaT(br, ur, fmaT(-bi, ui, yr));
                nYi[i] = fmaT(br, ui, fmaT(bi, ur, yi));
            } else {
                const T xlr = i + 1 < C ? Xlr[i + 1] : xinl_r;
                const T xli = i + 1 < C ? Xli[i + 1] : xinl_i;
                const T ylr = i > 0 ? Ylr[i - 1] : yinl_r;
                const T yli = i > 0 ? Yli[i - 1] : yinl_i;
                // increment plus the carried compensation, then an exact
                // (Fast2Sum) split of hi + increment
                const T dXr = fmaT(cr, tr, fmaT(-ci, ti, xlr));
                const T dXi = fmaT(cr, ti, fmaT(ci, tr, xli));
                const T dYr = fmaT(br, ur, fmaT(-bi, ui, ylr));
                const T dYi = fmaT(br, ui, fmaT(bi, ur, yli));
                const T sXr = xr + dXr, sXi = xi + dXi;
                const T sYr = yr + dYr, sYi = yi + dYi;
                nXlr[i] = dXr - (sXr - xr);
                nXli[i] = dXi - (sXi - xi);
                nYlr[i] = dYr - (sYr - yr);
                nYli[i] = dYi - (sYi - yi);
                nXr[i] = sXr, nXi[i] = sXi, nYr[i] = sYr, nYi[i] = sYi;
            }
        }
#pragma unroll
        for (int i = 0; i < C; ++i) {
            Xr[i] = nXr[i], Xi[i] = nXi[i], Yr[i] = nYr[i], Yi[i] = nYi[i];
            if constexpr (COMP)
                Xlr[i] = nXlr[i], Xli[i] = nXli[i], Ylr[i] = nYlr[i],
                Yli[i] = nYli[i];
        }
    }

    const double qr = a.qN[2 * ie], qi = a.qN[2 * ie + 1];
    const size_t row = ((size_t)ie * a.Ny + iy) * a.Nx;
#pragma unroll
    for (int i = 0; i < C; ++i) {
        const int x = lane * C + i;
        if (x >= a.Nx) continue;
        double yr = (double)Yr[i], yi = (double)Yi[i];
        if constexpr (COMP) yr += (double)Ylr[i], yi += (double)Yli[i];
        const double o_r = qr * yr - qi * yi, o_i = qr * yi + qi * yr;
        if constexpr (PREC == DYF_FP64)
            ((cx<double> *)a.out)[row + x] = cx<double>{o_r, o_i};
        else
            ((cx<float> *)a.out)[row + x] = cx<float>{(float)o_r, (float)o_i};
    }
}

// The grid-mode march with FP32 state (stand-alone, ROLL) and the
// persistent batched march with FP32 state: the step loop is unrolled by
// C.  The shift moves every column one register over per step, so after
// C steps the register assignment is back where it started and the
// allocator needs no copies for the loop-carried state (C = 17 mixed: 489
// VALU instructions per step instead of 587).  Without ROLL each step's
// e^{iH} slice is loaded at the start of its step (the persistent kernel,
// capped at two waves per SIMD: 256-instance ensemble, mixed 4.80 s ->
// 4.10 s when this replaced the rolled loop there).
// The persistent kernel reads its table, coefficient and output pointers
// from descriptors in device memory, so the compiler cannot tell they are
// global and would access them as flat memory, whose loads also count
// against the LDS counter every step's lane exchange waits on.  These
// casts restate the global address space.
template <typename T>
using gptr = const T __attribute__((address_space(1))) *;
template <typename T>
using gptr_w = T __attribute__((address_space(1))) *;

// Buffer-resource word 3 for raw (unswizzled, stride 0) buffers on gfx9
// parts: 32-bit data format, no index stride.
constexpr int kBufferWord3 = 0x00020000;
// A lane's byte offset for the slices it does not need (before its first
// step in the beam's fan): past every row's range (rows stay below 2^31
// bytes), so the buffer load returns zeros without reading memory.
constexpr int kOutOfRange = 0x7FFFFFF0;

template <typename T>
__device__ __forceinline__ cx<T> load_cx(__amdgpu_buffer_rsrc_t rs, int voff,
                                         int soff) {
    if constexpr (sizeof(T) == 8) {
        using u4 = unsigned __attribute__((ext_vector_type(4)));
        const u4 v = __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, 0);
        return __builtin_bit_cast(cx<T>, v);
    } else {
        using u2 = unsigned __attribute__((ext_vector_type(2)));
        const u2 v = __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 0);
        return __builtin_bit_cast(cx<T>, v);
    }
}

template <int C, int PREC, bool ROLL = false>
__device__ __forceinline__ void fast_row_unrolled(const FastArgs &a, int r,
                                                  int ie) {
    using S = StateT<PREC>;
    using T = typename S::T;
    constexpr bool COMP = S::comp;
    constexpr int kSlice = C * kWave;                   // entries per step
    const int lane = threadIdx.x & (kWave - 1);
    const int iy = a.iy_lo + r;

    // fields: hi parts (and compensations for MIXED)
    T Xr[C], Xi[C], Yr[C], Yi[C];
    T Xlr[COMP ? C : 1], Xli[COMP ? C : 1], Ylr[COMP ? C : 1], Yli[COMP ? C : 1];
#pragma unroll
    for (int i = 0; i < C; ++i) {
        const int x = lane * C + i;
        Xr[i] = x < a.Nx ? (T)beam_profile(x, a.Nx, a.dx, a.sheet_hw) : T(0);
        Xi[i] = Yr[i] = Yi[i] = T(0);
        if constexpr (COMP) Xlr[i] = Xli[i] = Ylr[i] = Yli[i] = T(0);
    }
    // the row's table through a buffer resource, the lane's byte offset
    // the only vector operand (as in fast_row_inplace)
    constexpr int kEl = (int)sizeof(cx<T>);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)((const cx<T> *)a.phase + (size_t)r * a.nsteps * kSlice), 0,
        a.nsteps * kSlice * kEl, kBufferWord3);
    const int loff = lane * kEl;
    // Slices before the lane's first step in the beam's fan are not read
    // (the fields there are exact zeros), except at C = 18 without ROLL,
    // where selecting the offset spills registers inside the step loop
    // (the table holds zeros there, so reading them is harmless).
    const int kfirst = lane_first_step(lane, C, a.fan_lo, a.fan_hi);
    auto slice_at = [&](int voff, int k, int i) {
        return load_cx<T>(rs, voff, (k * C + i) * kWave * kEl);
    };
    auto slice = [&](int k, int i) {
        return slice_at(k >= kfirst ? loff : kOutOfRange, k, i);
    };

    // ROLL: the e^{iH} slice is prefetched on a rolling basis, as in
    // fast_row (column i's value for step k+1 is loaded as soon as column
    // i of step k has used it); otherwise each step loads its slice first.
    cx<T> Ec[C];
    if constexpr (ROLL) {
#pragma unroll
        for (int i = 0; i < C; ++i) {
            Ec[i] = slice(0, i);
            if constexpr (sizeof(T) == 4) __builtin_amdgcn_sched_barrier(0);
        }
    }

    auto step = [&](int k) {
        const int kn = ROLL ? (k + 1 < a.nsteps ? k + 1 : k) : k;
        if constexpr (!ROLL && C >= 18) {
#pragma unroll
            for (int i = 0; i < C; ++i) Ec[i] = slice_at(loff, k, i);
        } else if constexpr (!ROLL) {
            // the offset chosen once per step (the empty asm keeps the
            // compiler from re-deriving it for each load)
            int voff = k >= kfirst ? loff : kOutOfRange;
            asm volatile("" : "+v"(voff));
#pragma unroll
            for (int i = 0; i < C; ++i) Ec[i] = slice_at(voff, k, i);
        }
        T cr, ci, br, bi;
        const size_t ci_ = (size_t)__builtin_amdgcn_readfirstlane(ie) * a.nsteps + k;
        if constexpr (sizeof(T) == 8) {
            kptr<CoefD> c = (kptr<CoefD>)a.cd + ci_;
            cr = c->cr, ci = c->ci, br = c->br, bi = c->bi;
        } else {
            kptr<CoefF> c = (kptr<CoefF>)a.cf + ci_;
            cr = c->cr, ci = c->ci, br = c->br, bi = c->bi;
        }
        // inflow across lane boundaries (old values)
        const T xin_r = from_right(Xr[0], lane), xin_i = from_right(Xi[0], lane);
        const T yin_r = from_left(Yr[C - 1], lane), yin_i = from_left(Yi[C - 1], lane);
        T xinl_r = T(0), xinl_i = T(0), yinl_r = T(0), yinl_i = T(0);
        if constexpr (COMP) {
            xinl_r = from_right(Xlr[0], lane), xinl_i = from_right(Xli[0], lane);
            yinl_r = from_left(Ylr[C - 1], lane), yinl_i = from_left(Yli[C - 1], lane);
        }
        T nXr[C], nXi[C], nYr[C], nYi[C];
        T nXlr[COMP ? C : 1], nXli[COMP ? C : 1], nYlr[COMP ? C : 1], nYli[COMP ? C : 1];
#pragma unroll
        for (int i = 0; i < C; ++i) {
            // sources: X from x+1, Y from x-1
            const T xr = i + 1 < C ? Xr[i + 1] : xin_r;
            const T xi = i + 1 < C ? Xi[i + 1] : xin_i;
            const T yr = i > 0 ? Yr[i - 1] : yin_r;
            const T yi = i > 0 ? Yi[i - 1] : yin_i;
            const T er = Ec[i].r, ei = Ec[i].i;
            const T tr = fmaT(er, yr, -ei * yi);    // t = e^{iH} y
            const T ti = fmaT(er, yi, ei * yr);
            const T ur = fmaT(er, xr, ei * xi);     // u = e^{-iH} x
            const T ui = fmaT(er, xi, -ei * xr);
            if constexpr (ROLL) {
                asm volatile("" ::: "memory");
                Ec[i] = slice(kn, i);
                if constexpr (sizeof(T) == 4) __builtin_amdgcn_sched_barrier(0);
            }
            if constexpr (!COMP) {
                nXr[i] = fmaT(cr, tr, fmaT(-ci, ti, xr));
                nXi[i] = fmaT(cr, ti, fmaT(ci, tr, xi));
                nYr[i] = fmaT(br, ur, fmaT(-bi, ui, yr));
                nYi[i] = fmaT(br, ui, fmaT(bi, ur, yi));
            } else {
                const T xlr = i + 1 < C ? Xlr[i + 1] : xinl_r;
                const T xli = i + 1 < C ? Xli[i + 1] : xinl_i;
                const T ylr = i > 0 ? Ylr[i - 1] : yinl_r;
                const T yli = i > 0 ? Yli[i - 1] : yinl_i;
                // increment plus the carried compensation, then an exact
                // (Fast2Sum) split of hi + increment
                const T dXr = fmaT(cr, tr, fmaT(-ci, ti, xlr));
                const T dXi = fmaT(cr, ti, fmaT(ci, tr, xli));
                const T dYr = fmaT(br, ur, fmaT(-bi, ui, ylr));
                const T dYi = fmaT(br, ui, fmaT(bi, ur, yli));
                const T sXr = xr + dXr, sXi = xi + dXi;
                const T sYr = yr + dYr, sYi = yi + dYi;
                nXlr[i] = dXr - (sXr - xr);
                nXli[i] = dXi - (sXi - xi);
                nYlr[i] = dYr - (sYr - yr);
                nYli[i] = dYi - (sYi - yi);
                nXr[i] = sXr, nXi[i] = sXi, nYr[i] = sYr, nYi[i] = sYi;
            }
        }
#pragma unroll
        for (int i = 0; i < C; ++i) {
            Xr[i] = nXr[i], Xi[i] = nXi[i], Yr[i] = nYr[i], Yi[i] = nYi[i];
            if constexpr (COMP)
                Xlr[i] = nXlr[i], Xli[i] = nXli[i], Ylr[i] = nYlr[i],
                Yli[i] = nYli[i];
        }
    };
    int k = 0;
    for (; k + C <= a.nsteps; k += C) {
#pragma unroll
        for (int kk = 0; kk < C; ++kk) {
            step(k + kk);
            if constexpr (ROLL) __builtin_amdgcn_sched_barrier(0);
        }
    }
    for (; k < a.nsteps; ++k) step(k);

    gptr<double> qN = (gptr<double>)a.qN;
    const double qr = qN[2 * ie], qi = qN[2 * ie + 1];
    const size_t row = ((size_t)ie * a.Ny + iy) * a.Nx;
#pragma unroll
    for (int i = 0; i < C; ++i) {
        const int x = lane * C + i;
        if (x >= a.Nx) continue;
        double yr = (double)Yr[i], yi = (double)Yi[i];
        if constexpr (COMP) yr += (double)Ylr[i], yi += (double)Yli[i];
        const double o_r = qr * yr - qi * yi, o_i = qr * yi + qi * yr;
        if constexpr (PREC == DYF_FP64) {
            gptr_w<cx<double>> o = (gptr_w<cx<double>>)a.out + row + x;
            o->r = o_r, o->i = o_i;
        } else {
            gptr_w<cx<float>> o = (gptr_w<cx<float>>)a.out + row + x;
            o->r = (float)o_r, o->i = (float)o_i;
        }
    }
}

// fast_row for the grid mode with the state updated in place, in two
// passes per step, so the rolled loop needs no register copies.
//
// Column i of a step reads X[i+1] and Y[i-1] only.  A pass in ascending
// column order can overwrite X[i] (its last reader, column i-1, has run)
// but not Y[i] (column i+1 still reads it); descending order is the other
// way round.  So one pass updates one field in place and leaves, in the
// register that held column i's e^{iH}, the product the other field needs
// (u = e^{-iH} X[i+1] or t = e^{iH} Y[i-1]); the second pass runs in the
// opposite order, updates the other field in place from it, and loads the
// next step's e^{iH} of column i into that register.  Consecutive steps
// swap the pass order (X ascending then Y descending; Y descending then X
// ascending), so every column's next e^{iH} is loaded a whole step before
// its use.  Each value is formed by the same operations as in fast_row,
// so the results are bit-identical.
template <int C, int PREC, int U>
__device__ __forceinline__ void fast_row_inplace(const FastArgs &a, int r,
                                                 int ie) {
    using S = StateT<PREC>;
    using T = typename S::T;
    constexpr bool COMP = S::comp;
    constexpr int kSlice = C * kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int iy = a.iy_lo + r;

    T Xr[C], Xi[C], Yr[C], Yi[C];
    T Xlr[COMP ? C : 1], Xli[COMP ? C : 1], Ylr[COMP ? C : 1], Yli[COMP ? C : 1];
#pragma unroll
    for (int i = 0; i < C; ++i) {
        const int x = lane * C + i;
        Xr[i] = x < a.Nx ? (T)beam_profile(x, a.Nx, a.dx, a.sheet_hw) : T(0);
        Xi[i] = Yr[i] = Yi[i] = T(0);
        if constexpr (COMP) Xlr[i] = Xli[i] = Ylr[i] = Yli[i] = T(0);
    }
    // The row's table is read through a buffer resource (scalar registers)
    // with the lane's byte offset as the only vector operand: the column
    // and step offsets are scalar, so no 64-bit addresses are held or
    // advanced in vector registers.  The row spans nsteps * C * 64 entries
    // (8.5 MB at Nx 1056, Nz 490), well inside the 32-bit offset range.
    constexpr int kEl = (int)sizeof(cx<T>);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)((const cx<T> *)a.phase + (size_t)r * a.nsteps * kSlice), 0,
        a.nsteps * kSlice * kEl, kBufferWord3);
    const int loff = lane * kEl;
    // slices before the lane's first step in the beam's fan are not read
    // (the fields there are exact zeros)
    const int kfirst = lane_first_step(lane, C, a.fan_lo, a.fan_hi);
    auto slice = [&](int k, int i) {
        return load_cx<T>(rs, k >= kfirst ? loff : kOutOfRange,
                          (k * C + i) * kWave * kEl);
    };
    // column i's e^{iH}, or between the two passes the product the second
    // pass needs
    T Sr[C], Si[C];
#pragma unroll
    for (int i = 0; i < C; ++i) {
        const cx<T> e = slice(0, i);
        Sr[i] = e.r, Si[i] = e.i;
    }

    auto coefs = [&](int k, T &cr, T &ci, T &br, T &bi) {
        const size_t at = (size_t)__builtin_amdgcn_readfirstlane(ie) * a.nsteps + k;
        if constexpr (sizeof(T) == 8) {
            kptr<CoefD> c = (kptr<CoefD>)a.cd + at;
            cr = c->cr, ci = c->ci, br = c->br, bi = c->bi;
        } else {
            kptr<CoefF> c = (kptr<CoefF>)a.cf + at;
            cr = c->cr, ci = c->ci, br = c->br, bi = c->bi;
        }
    };
    // X'[i] = X[i+1] + c t (and its compensation), in place
    auto new_x = [&](int i, T xr, T xi, T xlr, T xli, T tr, T ti, T cr, T ci) {
        if constexpr (!COMP) {
            Xr[i] = fmaT(cr, tr, fmaT(-ci, ti, xr));
            Xi[i] = fmaT(cr, ti, fmaT(ci, tr, xi));
        } else {
            const T dr = fmaT(cr, tr, fmaT(-ci, ti, xlr));
            const T di = fmaT(cr, ti, fmaT(ci, tr, xli));
            const T sr = xr + dr, si = xi + di;
            Xlr[i] = dr - (sr - xr), Xli[i] = di - (si - xi);
            Xr[i] = sr, Xi[i] = si;
        }
    };
    // Y'[i] = Y[i-1] + b u (and its compensation), in place
    auto new_y = [&](int i, T yr, T yi, T ylr, T yli, T ur, T ui, T br, T bi) {
        if constexpr (!COMP) {
            Yr[i] = fmaT(br, ur, fmaT(-bi, ui, yr));
            Yi[i] = fmaT(br, ui, fmaT(bi, ur, yi));
        } else {
            const T dr = fmaT(br, ur, fmaT(-bi, ui, ylr));
            const T di = fmaT(br, ui, fmaT(bi, ur, yli));
            const T sr = yr + dr, si = yi + di;
            Ylr[i] = dr - (sr - yr), Yli[i] = di - (si - yi);
            Yr[i] = sr, Yi[i] = si;
        }
    };

    // one step; XFIRST: X pass ascending, then Y pass descending
    auto step = [&](int k, auto xfirst) {
        constexpr bool XFIRST = decltype(xfirst)::value;
        T cr, ci, br, bi;
        coefs(k, cr, ci, br, bi);
        const int kn = k + 1 < a.nsteps ? k + 1 : k;
        // inflow across lane boundaries (old values)
        const T xin_r = from_right(Xr[0], lane), xin_i = from_right(Xi[0], lane);
        const T yin_r = from_left(Yr[C - 1], lane), yin_i = from_left(Yi[C - 1], lane);
        T xinl_r = T(0), xinl_i = T(0), yinl_r = T(0), yinl_i = T(0);
        if constexpr (COMP) {
            xinl_r = from_right(Xlr[0], lane), xinl_i = from_right(Xli[0], lane);
            yinl_r = from_left(Ylr[C - 1], lane), yinl_i = from_left(Yli[C - 1], lane);
        }
        // sources of column i: X from x+1, Y from x-1 (old values while
        // the pass order keeps them unwritten)
        auto xs = [&](int i, T &r_, T &i_, T &lr, T &li) {
            r_ = i + 1 < C ? Xr[i + 1] : xin_r, i_ = i + 1 < C ? Xi[i + 1] : xin_i;
            if constexpr (COMP)
                lr = i + 1 < C ? Xlr[i + 1] : xinl_r, li = i + 1 < C ? Xli[i + 1] : xinl_i;
        };
        auto ys = [&](int i, T &r_, T &i_, T &lr, T &li) {
            r_ = i > 0 ? Yr[i - 1] : yin_r, i_ = i > 0 ? Yi[i - 1] : yin_i;
            if constexpr (COMP)
                lr = i > 0 ? Ylr[i - 1] : yinl_r, li = i > 0 ? Yli[i - 1] : yinl_i;
        };
        auto first = [&](int i) {
            T xr, xi, xlr = T(0), xli = T(0), yr, yi, ylr = T(0), yli = T(0);
            xs(i, xr, xi, xlr, xli);
            ys(i, yr, yi, ylr, yli);
            const T er = Sr[i], ei = Si[i];
            const T tr = fmaT(er, yr, -ei * yi);    // t = e^{iH} y
            const T ti = fmaT(er, yi, ei * yr);
            const T ur = fmaT(er, xr, ei * xi);     // u = e^{-iH} x
            const T ui = fmaT(er, xi, -ei * xr);
            if constexpr (XFIRST) {
                new_x(i, xr, xi, xlr, xli, tr, ti, cr, ci);
                Sr[i] = ur, Si[i] = ui;
            } else {
                new_y(i, yr, yi, ylr, yli, ur, ui, br, bi);
                Sr[i] = tr, Si[i] = ti;
            }
        };
        auto second = [&](int i) {
            if constexpr (XFIRST) {
                T yr, yi, ylr = T(0), yli = T(0);
                ys(i, yr, yi, ylr, yli);
                new_y(i, yr, yi, ylr, yli, Sr[i], Si[i], br, bi);
            } else {
                T xr, xi, xlr = T(0), xli = T(0);
                xs(i, xr, xi, xlr, xli);
                new_x(i, xr, xi, xlr, xli, Sr[i], Si[i], cr, ci);
            }
            asm volatile("" ::: "memory");
            const cx<T> e = slice(kn, i);
            Sr[i] = e.r, Si[i] = e.i;
        };
        if constexpr (XFIRST) {
#pragma unroll
            for (int i = 0; i < C; ++i) first(i);
#pragma unroll
            for (int i = C - 1; i >= 0; --i) second(i);
        } else {
#pragma unroll
            for (int i = C - 1; i >= 0; --i) first(i);
#pragma unroll
            for (int i = 0; i < C; ++i) second(i);
        }
    };
    // U steps per iteration (U / 2 of each pass order)
    using XF = std::integral_constant<bool, true>;
    using YF = std::integral_constant<bool, false>;
    int k = 0;
    if constexpr (U > 2) {
        for (; k + U <= a.nsteps; k += U) {
#pragma unroll
            for (int kk = 0; kk < U; kk += 2) {
                step(k + kk, XF{});
                step(k + kk + 1, YF{});
            }
        }
    }
    for (; k + 2 <= a.nsteps; k += 2) {
        step(k, XF{});
        step(k + 1, YF{});
    }
    if (k < a.nsteps) step(k, XF{});

    gptr<double> qN = (gptr<double>)a.qN;
    const double qr = qN[2 * ie], qi = qN[2 * ie + 1];
    const size_t row = ((size_t)ie * a.Ny + iy) * a.Nx;
#pragma unroll
    for (int i = 0; i < C; ++i) {
        const int x = lane * C + i;
        if (x >= a.Nx) continue;
        double yr = (double)Yr[i], yi = (double)Yi[i];
        if constexpr (COMP) yr += (double)Ylr[i], yi += (double)Yli[i];
        const double o_r = qr * yr - qi * yi, o_i = qr * yi + qi * yr;
        if constexpr (PREC == DYF_FP64) {
            gptr_w<cx<double>> o = (gptr_w<cx<double>>)a.out + row + x;
            o->r = o_r, o->i = o_i;
        } else {
            gptr_w<cx<float>> o = (gptr_w<cx<float>>)a.out + row + x;
            o->r = (float)o_r, o->i = (float)o_i;
        }
    }
}

// Steps per loop iteration of the in-place march: four where that fits
// two waves per SIMD without spilling inside the step loop, else two.
template <int C>
constexpr int inplace_unroll() {
    return C == 13 || C == 18 ? 2 : 4;
}

template <int C, int PREC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
void laue_fast_inplace_kernel(FastArgs a) {
    const int L = logical_block(a.per_xcd);
    if (L >= a.total) return;
    const int r = L / a.n_e, ie = L - r * a.n_e;
    fast_row_inplace<C, PREC, inplace_unroll<C>()>(a, r, ie);
}

// ---- two angles per wavefront while the beam's fan is narrow ------------
// Before step K every non-zero field value, and every e^{iH} a step
// reads, lies on the 32 lanes [g0, g0 + 32) (the fan argument of
// lane_first_step; the host picks g0 and the largest such K).  Until then
// one wavefront marches two angles of the same row, one per half: lane l
// stands for lane g0 + (l & 31) of the one-angle layout, both halves read
// the same slices, each half takes its own angle's coefficients, and a
// half's edge lanes take zero inflow (the lanes beyond them hold exact
// zeros).  At step K the second angle's state goes to LDS, the first
// angle's moves to its one-angle lanes, and each angle finishes on the
// full wavefront in turn.  Every value is formed by the same operations
// as in fast_row_inplace (results equal up to the sign of zeros outside
// the fan); a row costs 2 nsteps - K wavefront-steps instead of 2 nsteps.
struct PairGeom {
    int gl;              // lane of the one-angle layout this lane stands for
    int kfirst;          // first step the lane needs its e^{iH}
    bool ledge, redge;   // no inflow from the left / right neighbour
};

template <typename T>
__device__ __forceinline__ T pair_take(T v, int src) {
    if constexpr (sizeof(T) == 8) {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_ds_bpermute(src << 2, (int)b);
        const int hi = __builtin_amdgcn_ds_bpermute(src << 2, (int)(b >> 32));
        return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
    } else {
        return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v)));
    }
}
// value of the right / left neighbour lane, 0 at an edge
template <typename T>
__device__ __forceinline__ T pair_right(T v, int lane, bool edge) {
    const T r = pair_take(v, lane + 1);
    return edge ? T(0) : r;
}
template <typename T>
__device__ __forceinline__ T pair_left(T v, int lane, bool edge) {
    const T r = pair_take(v, lane - 1);
    return edge ? T(0) : r;
}

// ANY_K: K may be 0 (the persistent march); the stand-alone kernel is
// launched with K > 0 only, and compiling without the K = 0 paths keeps
// its first phase free of register spills.
template <int C, int PREC, int U, bool ANY_K>
__device__ __forceinline__ void fast_row_pair(const FastArgs &a, int r,
                                              int ie0, int ie1, int K,
                                              int g0, char *lds_raw) {
    using S = StateT<PREC>;
    using T = typename S::T;
    constexpr bool COMP = S::comp;
    constexpr int NF = COMP ? 8 : 4;          // fields held per column
    constexpr int kSlice = C * kWave;
    constexpr int kEl = (int)sizeof(cx<T>);
    const int lane = threadIdx.x & (kWave - 1);
    const int half = lane >> 5, hl = lane & 31;
    const int iy = a.iy_lo + r;
    T *lds = (T *)lds_raw;

    // state: field f of column i is F[f][i]: X re, X im, Y re, Y im, then
    // (mixed) their compensations
    T F[NF][C], Sr[C], Si[C];
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)((const cx<T> *)a.phase + (size_t)r * a.nsteps * kSlice), 0,
        a.nsteps * kSlice * kEl, kBufferWord3);
    auto load_slice = [&](const PairGeom &g, int k) {
        const int voff = k >= g.kfirst ? g.gl * kEl : kOutOfRange;
#pragma unroll
        for (int i = 0; i < C; ++i) {
            const cx<T> e = load_cx<T>(rs, voff, (k * C + i) * kWave * kEl);
            Sr[i] = e.r, Si[i] = e.i;
        }
    };
    auto scalar_coefs = [&](int ie, int k, T &cr, T &ci, T &br, T &bi) {
        const size_t at = (size_t)ie * a.nsteps + k;
        if constexpr (sizeof(T) == 8) {
            kptr<CoefD> c = (kptr<CoefD>)a.cd + at;
            cr = c->cr, ci = c->ci, br = c->br, bi = c->bi;
        } else {
            kptr<CoefF> c = (kptr<CoefF>)a.cf + at;
            cr = c->cr, ci = c->ci, br = c->br, bi = c->bi;
        }
    };
    // X'[i] = X[i+1] + c t, Y'[i] = Y[i-1] + b u, in place (with the
    // compensated split of fast_row_inplace in mixed precision)
    auto new_x = [&](int i, T xr, T xi, T xlr, T xli, T tr, T ti, T cr, T ci) {
        if constexpr (!COMP) {
            F[0][i] = fmaT(cr, tr, fmaT(-ci, ti, xr));
            F[1][i] = fmaT(cr, ti, fmaT(ci, tr, xi));
        } else {
            const T dr = fmaT(cr, tr, fmaT(-ci, ti, xlr));
            const T di = fmaT(cr, ti, fmaT(ci, tr, xli));
            const T sr = xr + dr, si = xi + di;
            F[4][i] = dr - (sr - xr), F[5][i] = di - (si - xi);
            F[0][i] = sr, F[1][i] = si;
        }
    };
    auto new_y = [&](int i, T yr, T yi, T ylr, T yli, T ur, T ui, T br, T bi) {
        if constexpr (!COMP) {
            F[2][i] = fmaT(br, ur, fmaT(-bi, ui, yr));
            F[3][i] = fmaT(br, ui, fmaT(bi, ur, yi));
        } else {
            const T dr = fmaT(br, ur, fmaT(-bi, ui, ylr));
            const T di = fmaT(br, ui, fmaT(bi, ur, yli));
            const T sr = yr + dr, si = yi + di;
            F[6][i] = dr - (sr - yr), F[7][i] = di - (si - yi);
            F[2][i] = sr, F[3][i] = si;
        }
    };

    // steps [kb, ke) on geometry g; coefs(k, cr, ci, br, bi)
    auto march = [&](int kb, int ke, const PairGeom &g, auto coefs) {
        auto step = [&](int k, auto xfirst) {
            constexpr bool XFIRST = decltype(xfirst)::value;
            T cr, ci, br, bi;
            coefs(k, cr, ci, br, bi);
            const int kn = k + 1 < ke ? k + 1 : k;
            const int voff = kn >= g.kfirst ? g.gl * kEl : kOutOfRange;
            // inflow across lane boundaries (old values)
            T xin[COMP ? 4 : 2], yin[COMP ? 4 : 2];
            xin[0] = pair_right(F[0][0], lane, g.redge);
            xin[1] = pair_right(F[1][0], lane, g.redge);
            yin[0] = pair_left(F[2][C - 1], lane, g.ledge);
            yin[1] = pair_left(F[3][C - 1], lane, g.ledge);
            if constexpr (COMP) {
                xin[2] = pair_right(F[4][0], lane, g.redge);
                xin[3] = pair_right(F[5][0], lane, g.redge);
                yin[2] = pair_left(F[6][C - 1], lane, g.ledge);
                yin[3] = pair_left(F[7][C - 1], lane, g.ledge);
            }
            // sources of column i: X from x+1, Y from x-1
            auto xs = [&](int i, T &vr, T &vi, T &lr, T &li) {
                vr = i + 1 < C ? F[0][i + 1] : xin[0];
                vi = i + 1 < C ? F[1][i + 1] : xin[1];
                if constexpr (COMP) {
                    lr = i + 1 < C ? F[4][i + 1] : xin[2];
                    li = i + 1 < C ? F[5][i + 1] : xin[3];
                }
            };
            auto ys = [&](int i, T &vr, T &vi, T &lr, T &li) {
                vr = i > 0 ? F[2][i - 1] : yin[0];
                vi = i > 0 ? F[3][i - 1] : yin[1];
                if constexpr (COMP) {
                    lr = i > 0 ? F[6][i - 1] : yin[2];
                    li = i > 0 ? F[7][i - 1] : yin[3];
                }
            };
            auto first = [&](int i) {
                T xr, xi, xlr = T(0), xli = T(0), yr, yi, ylr = T(0), yli = T(0);
                xs(i, xr, xi, xlr, xli);
                ys(i, yr, yi, ylr, yli);
                const T er = Sr[i], ei = Si[i];
                const T tr = fmaT(er, yr, -ei * yi);    // t = e^{iH} y
                const T ti = fmaT(er, yi, ei * yr);
                const T ur = fmaT(er, xr, ei * xi);     // u = e^{-iH} x
                const T ui = fmaT(er, xi, -ei * xr);
                if constexpr (XFIRST) {
                    new_x(i, xr, xi, xlr, xli, tr, ti, cr, ci);
                    Sr[i] = ur, Si[i] = ui;
                } else {
                    new_y(i, yr, yi, ylr, yli, ur, ui, br, bi);
                    Sr[i] = tr, Si[i] = ti;
                }
            };
            auto second = [&](int i) {
                if constexpr (XFIRST) {
                    T yr, yi, ylr = T(0), yli = T(0);
                    ys(i, yr, yi, ylr, yli);
                    new_y(i, yr, yi, ylr, yli, Sr[i], Si[i], br, bi);
                } else {
                    T xr, xi, xlr = T(0), xli = T(0);
                    xs(i, xr, xi, xlr, xli);
                    new_x(i, xr, xi, xlr, xli, Sr[i], Si[i], cr, ci);
                }
                asm volatile("" ::: "memory");
                const cx<T> e = load_cx<T>(rs, voff, (kn * C + i) * kWave * kEl);
                Sr[i] = e.r, Si[i] = e.i;
            };
            if constexpr (XFIRST) {
#pragma unroll
                for (int i = 0; i < C; ++i) first(i);
#pragma unroll
                for (int i = C - 1; i >= 0; --i) second(i);
            } else {
#pragma unroll
                for (int i = C - 1; i >= 0; --i) first(i);
#pragma unroll
                for (int i = 0; i < C; ++i) second(i);
            }
        };
        using XF = std::integral_constant<bool, true>;
        using YF = std::integral_constant<bool, false>;
        int k = kb;
        if constexpr (U > 2) {
            for (; k + U <= ke; k += U) {
#pragma unroll
                for (int kk = 0; kk < U; kk += 2) {
                    step(k + kk, XF{});
                    step(k + kk + 1, YF{});
                }
            }
        }
        for (; k + 2 <= ke; k += 2) {
            step(k, XF{});
            step(k + 1, YF{});
        }
        if (k < ke) step(k, XF{});
    };
    auto write_out = [&](int ie) {
        gptr<double> qN = (gptr<double>)a.qN;
        const double qr = qN[2 * ie], qi = qN[2 * ie + 1];
        const size_t row = ((size_t)ie * a.Ny + iy) * a.Nx;
#pragma unroll
        for (int i = 0; i < C; ++i) {
            const int x = lane * C + i;
            if (x >= a.Nx) continue;
            double yr = (double)F[2][i], yi = (double)F[3][i];
            if constexpr (COMP) yr += (double)F[6][i], yi += (double)F[7][i];
            const double o_r = qr * yr - qi * yi, o_i = qr * yi + qi * yr;
            if constexpr (PREC == DYF_FP64) {
                gptr_w<cx<double>> o = (gptr_w<cx<double>>)a.out + row + x;
                o->r = o_r, o->i = o_i;
            } else {
                gptr_w<cx<float>> o = (gptr_w<cx<float>>)a.out + row + x;
                o->r = (float)o_r, o->i = (float)o_i;
            }
        }
    };

    // the incident beam on geometry g
    auto init_state = [&](const PairGeom &g) {
#pragma unroll
        for (int i = 0; i < C; ++i) {
            const int x = g.gl * C + i;
            F[0][i] = x < a.Nx ? (T)beam_profile(x, a.Nx, a.dx, a.sheet_hw) : T(0);
#pragma unroll
            for (int f = 1; f < NF; ++f) F[f][i] = T(0);
        }
    };
    const int s0 = __builtin_amdgcn_readfirstlane(ie0);
    const int s1 = __builtin_amdgcn_readfirstlane(ie1);
    PairGeom g;
    if (!ANY_K || K > 0) {
        // phase 1: both angles, one per half
        g.gl = g0 + hl;
        g.kfirst = lane_first_step(g.gl, C, a.fan_lo, a.fan_hi);
        g.ledge = hl == 0;
        g.redge = hl == 31;
        init_state(g);
        load_slice(g, 0);
        march(0, K, g, [&](int k, T &cr, T &ci, T &br, T &bi) {
            T c0r, c0i, b0r, b0i, c1r, c1i, b1r, b1i;
            scalar_coefs(s0, k, c0r, c0i, b0r, b0i);
            scalar_coefs(s1, k, c1r, c1i, b1r, b1i);
            cr = half ? c1r : c0r, ci = half ? c1i : c0i;
            br = half ? b1r : b0r, bi = half ? b1i : b0i;
        });

        // the second angle's state to LDS ([field][column][32 lanes])
        if (half) {
#pragma unroll
            for (int f = 0; f < NF; ++f)
#pragma unroll
                for (int i = 0; i < C; ++i) lds[(f * C + i) * 32 + hl] = F[f][i];
        }
        // the first angle's state to its one-angle lanes
        const int j = lane - g0;
        const bool inwin = j >= 0 && j < 32;
        const int src = inwin ? j : 0;
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int i = 0; i < C; ++i) {
                const T v = pair_take(F[f][i], src);
                F[f][i] = inwin ? v : T(0);
            }
        __syncthreads();
    }

    // phase 2: each angle on the full wavefront (K = 0: the whole march,
    // one angle after the other)
    g.gl = lane;
    g.kfirst = lane_first_step(lane, C, a.fan_lo, a.fan_hi);
    g.ledge = lane == 0;
    g.redge = lane == kWave - 1;
    if constexpr (ANY_K)
        if (K == 0) init_state(g);
    // (no slice past the row's last step: the buffer's range check covers
    // the lane offset, not the scalar one)
    if (K < a.nsteps) load_slice(g, K);
    march(K, a.nsteps, g, [&](int k, T &cr, T &ci, T &br, T &bi) {
        scalar_coefs(s0, k, cr, ci, br, bi);
    });
    write_out(ie0);
    if (ie1 == ie0) return;
    if (!ANY_K || K > 0) {
        const int j = lane - g0;
        const bool inwin = j >= 0 && j < 32;
        const int src = inwin ? j : 0;
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int i = 0; i < C; ++i)
                F[f][i] = inwin ? lds[(f * C + i) * 32 + src] : T(0);
    } else {
        init_state(g);
    }
    if (K < a.nsteps) load_slice(g, K);
    march(K, a.nsteps, g, [&](int k, T &cr, T &ci, T &br, T &bi) {
        scalar_coefs(s1, k, cr, ci, br, bi);
    });
    write_out(ie1);
}

// One wavefront per (row, pair of angles); LDS holds the second angle's
// state between the phases (4 fields x C x 32 lanes in FP64, 8 in FP32
// for the mixed state: 1 KB per column either way).  ie1 = ie0 for the
// last pair of an odd count (its second half marches a copy, not
// written).
template <int C, int PREC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
void laue_fast_pair_kernel(FastArgs a, int K, int g0) {
    extern __shared__ double pair_lds[];
    const int L = logical_block(a.per_xcd);
    if (L >= a.total) return;
    const int npair = (a.n_e + 1) / 2;
    const int r = L / npair, p = L - r * npair;
    const int ie0 = 2 * p, ie1 = 2 * p + 1 < a.n_e ? 2 * p + 1 : 2 * p;
    // two steps per loop iteration: four spill inside the loops at C = 17
    fast_row_pair<C, PREC, 2, false>(a, r, ie0, ie1, K, g0, (char *)pair_lds);
}

template <int C, int MODE, int PREC>
__global__ __launch_bounds__(64) void laue_fast_kernel(FastArgs a) {
    const int L = logical_block(a.per_xcd);
    if (L >= a.total) return;
    const int r = L / a.n_e, ie = L - r * a.n_e;
    fast_row<C, MODE, PREC>(a, r, ie);
}

// The grid-mode march with FP32 state (FP32 and MIXED): the step loop
// unrolled by C with the rolling prefetch.  Each step moves every column
// one register over, so after C steps the register assignment is back
// where it started and the loop-carried state needs no copies (C = 17
// mixed: 497 VALU instructions per step instead of 553).  With FP64 state
// the register allocator does not find that assignment within two waves
// per SIMD and spills, so FP64 keeps the rolled fast_row.
template <int C, int PREC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
void laue_fast_unrolled_kernel(FastArgs a) {
    const int L = logical_block(a.per_xcd);
    if (L >= a.total) return;
    const int r = L / a.n_e, ie = L - r * a.n_e;
    fast_row_unrolled<C, PREC, true>(a, r, ie);
}

// ---- general path: direct form, any shift, LDS rows ------------------
struct GeneralArgs {
    const void *H;          // GRID: (Ny, Nz, Nx) double or float
    int H_f32;
    dyf_line_params line;   // ANALYTIC
    const double *s_dev;
    void *out;
    int out_f32;
    int n_e, iy_lo, nrows, Nz, Nx, Ny;
    double s0r, s0i, shr, shi, sbr, sbi, dz, ds, dx, shift, sheet_hw;
};

__device__ __forceinline__ void interp_pair(const double *re, const double *im,
                                            double src, int Nx, double *vr,
                                            double *vi) {
    const int i0 = (int)floor(src);
    const double f = src - (double)i0;
    if (i0 >= 0 && i0 + 1 < Nx) {
        *vr = (1.0 - f) * re[i0] + f * re[i0 + 1];
        *vi = (1.0 - f) * im[i0] + f * im[i0 + 1];
    } else if (i0 >= 0 && i0 < Nx) {
        *vr = (1.0 - f) * re[i0];
        *vi = (1.0 - f) * im[i0];
    } else {
        *vr = 0.0;
        *vi = 0.0;
    }
}

template <int MODE>
__global__ __launch_bounds__(256) void laue_general_kernel(GeneralArgs a) {
    extern __shared__ double lds[];
    const int r = blockIdx.x / a.n_e, ie = blockIdx.x % a.n_e;
    const int iy = a.iy_lo + r;
    const int Nx = a.Nx, P = Nx + 1;
    double *buf[2] = {lds, lds + 4 * P};   // [D0r, D0i, Dgr, Dgi] x 2
    for (int x = threadIdx.x; x < Nx; x += blockDim.x) {
        buf[0][x] = beam_profile(x, Nx, a.dx, a.sheet_hw);
        buf[0][P + x] = buf[0][2 * P + x] = buf[0][3 * P + x] = 0.0;
    }
    __syncthreads();
    const double alpha = -2.0 * kPi * a.s_dev[ie];
    const double gam = sqrt(a.shr * a.shr + a.shi * a.shi + alpha * alpha * 0.25);
    double cg, sg;
    if (gam > 1e-30) {
        cg = cos(gam * a.ds);
        sg = sin(gam * a.ds) / gam;
    } else {
        cg = 1.0;
        sg = a.ds;
    }
    const double ex = exp(a.s0r * a.ds), ph = (a.s0i + alpha * 0.5) * a.ds;
    const double epr = ex * cos(ph), epi = ex * sin(ph);
    const double b = alpha * 0.5 * sg;
    const double M11r = epr * cg + epi * b, M11i = -epr * b + epi * cg;
    const double M22r = epr * cg - epi * b, M22i = epr * b + epi * cg;
    const double y_pix = (double)(iy - a.Ny / 2) * a.dx;
    int cur = 0;
    for (int k = 0; k < a.Nz - 1; ++k) {
        double cHu = 1.0, sHu = 0.0;   // row-uniform phase (analytic)
        if constexpr (MODE == DYF_MODE_ANALYTIC) {
            const dyf_line_params &lp = a.line;
            const double dy = y_pix - lp.y0, dzv = (double)k * a.dz - lp.z0;
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
            sincos(H, &sHu, &cHu);
        }
        const double *s = buf[cur];
        double *d = buf[1 - cur];
        for (int x = threadIdx.x; x < Nx; x += blockDim.x) {
            double cH = cHu, sH = sHu;
            if constexpr (MODE == DYF_MODE_GRID) {
                const size_t at = ((size_t)iy * a.Nz + k) * Nx + x;
                const double h = a.H_f32 ? (double)((const float *)a.H)[at]
                                         : ((const double *)a.H)[at];
                sincos(h, &sH, &cH);
            }
            const double P12r = (a.sbr * cH - a.sbi * sH) * sg;
            const double P12i = (a.sbi * cH + a.sbr * sH) * sg;
            const double P21r = (a.shr * cH + a.shi * sH) * sg;
            const double P21i = (a.shi * cH - a.shr * sH) * sg;
            const double M12r = epr * P12r - epi * P12i, M12i = epr * P12i + epi * P12r;
            const double M21r = epr * P21r - epi * P21i, M21i = epr * P21i + epi * P21r;
            double d0r, d0i, dgr, dgi;
            interp_pair(s, s + P, (double)x + a.shift, Nx, &d0r, &d0i);
            interp_pair(s + 2 * P, s + 3 * P, (double)x - a.shift, Nx, &dgr, &dgi);
            d[x] = M11r * d0r - M11i * d0i + M12r * dgr - M12i * dgi;
            d[P + x] = M11r * d0i + M11i * d0r + M12r * dgi + M12i * dgr;
            d[2 * P + x] = M21r * d0r - M21i * d0i + M22r * dgr - M22i * dgi;
            d[3 * P + x] = M21r * d0i + M21i * d0r + M22r * dgi + M22i * dgr;
        }
        __syncthreads();
        cur = 1 - cur;
    }
    const double *s = buf[cur];
    const size_t row = ((size_t)ie * a.Ny + iy) * Nx;
    for (int x = threadIdx.x; x < Nx; x += blockDim.x) {
        if (a.out_f32)
            ((cx<float> *)a.out)[row + x] =
                cx<float>{(float)s[2 * P + x], (float)s[3 * P + x]};
        else
            ((cx<double> *)a.out)[row + x] = cx<double>{s[2 * P + x], s[3 * P + x]};
    }
}

// Zeroes the rows outside the field of view (a fill kernel rather than a
// memset, so the whole solve can be captured in a graph).
__global__ __launch_bounds__(256) void zero_rows_kernel(void *out, int elem,
                                                        int n_e, int Ny, int Nx,
                                                        int iy_lo, int iy_hi) {
    const size_t per = (size_t)(Ny - (iy_hi - iy_lo)) * Nx;
    const size_t n = per * n_e;
    for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < n;
         t += (size_t)gridDim.x * 256) {
        const int ie = (int)(t / per);
        const size_t rem = t % per;
        int iy = (int)(rem / Nx);
        const int x = (int)(rem % Nx);
        if (iy >= iy_lo) iy += iy_hi - iy_lo;
        const size_t at = ((size_t)ie * Ny + iy) * Nx + x;
        if (elem == 16)
            ((cx<double> *)out)[at] = cx<double>{0.0, 0.0};
        else
            ((cx<float> *)out)[at] = cx<float>{0.0f, 0.0f};
    }
}

// ---- persistent march over many instances ----------------------------
__global__ void zero_counters_kernel(unsigned *c) {
    if (threadIdx.x < 8) c[threadIdx.x] = 0u;
}

// One persistent kernel per column count C (register allocation is per
// kernel, so each bucket keeps the occupancy of its standalone kernel).
template <int C, int PREC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void laue_persistent_kernel(
    const InstDesc *__restrict__ inst, int n_inst, long long total,
    long long per_queue, unsigned *__restrict__ counters) {
    extern __shared__ double pers_lds[];   // pair items: 1 KB x C per wave
    const int lane = threadIdx.x & (kWave - 1);
    const int q0 = blockIdx.x & 7;   // block b runs on XCD b % 8
    for (int qq = 0; qq < 8; ++qq) {
        const int q = (q0 + qq) & 7;
        const long long lo = (long long)q * per_queue;
        const long long hi = lo + per_queue < total ? lo + per_queue : total;
        for (;;) {
            unsigned t = 0;
            if (lane == 0) t = atomicAdd(&counters[q], 1u);
            t = __builtin_amdgcn_readfirstlane(t);
            const long long item = lo + (long long)t;
            if (item >= hi) break;
            int a = 0, b = n_inst - 1;
            while (a < b) {
                const int m = (a + b + 1) >> 1;
                if (inst[m].item_base <= item) a = m; else b = m - 1;
            }
            const InstDesc &d = inst[a];
            const long long local = item - d.item_base;
            FastArgs fa;
            fa.phase = d.phase;
            const CoefPtrs cp = coef_ptrs((void *)d.coef, d.n_e, d.nsteps);
            fa.cd = cp.d, fa.cf = cp.f, fa.qN = cp.qN;
            fa.out = d.out;
            fa.n_e = d.n_e, fa.nrows = d.nrows, fa.iy_lo = d.iy_lo;
            fa.nsteps = d.nsteps, fa.Nx = d.Nx, fa.Ny = d.Ny;
            fa.per_xcd = 0, fa.total = 0;
            fa.fan_lo = d.fan_lo, fa.fan_hi = d.fan_hi;
            fa.dx = d.dx, fa.sheet_hw = d.sheet_hw;
            if constexpr (C <= pair_max_c(PREC)) {
                // (row, pair of angles): two angles per wavefront for the
                // first pair_K steps, then one after the other
                const int npair = (d.n_e + 1) / 2;
                const int r = (int)(local / npair), p = (int)(local % npair);
                const int ie0 = 2 * p, ie1 = 2 * p + 1 < d.n_e ? 2 * p + 1 : 2 * p;
                char *lds = (char *)pers_lds + (size_t)(threadIdx.x / kWave) * 1024 * C;
                fast_row_pair<C, PREC, 2, true>(fa, r, ie0, ie1, d.pair_K, d.pair_g0, lds);
            } else {
                const int r = (int)(local / d.n_e), ie = (int)(local % d.n_e);
                // FP64: the in-place march (no register copies, slices
                // skipped before a lane's first step in the fan); FP32
                // state: the step loop unrolled by C
                if constexpr (PREC == DYF_FP64)
                    fast_row_inplace<C, PREC, inplace_unroll<C>()>(fa, r, ie);
                else
                    fast_row_unrolled<C, PREC>(fa, r, ie);
            }
        }
    }
}

using PersistFn = void (*)(const InstDesc *, int, long long, long long,
                           unsigned *);

template <int PREC, int... Cs>
static PersistFn persist_fn(int C) {
    PersistFn f = nullptr;
    ((C == Cs ? (f = laue_persistent_kernel<Cs, PREC>, 0) : 0), ...);
    return f;
}

size_t laue_coef_bytes(int n_e, int nsteps) { return coef_bytes(n_e, nsteps); }

int launch_laue_coef(const dyf_laue_args *a, const double *s_dev, int n_e,
                     void *coef, hipStream_t stream) {
    const int nsteps = a->Nz - 1;
    const CoefPtrs cp = coef_ptrs(coef, n_e, nsteps);
    const int n = n_e * nsteps;
    laue_coef_kernel<<<(n + 255) / 256, 256, 0, stream>>>(
        s_dev, n_e, nsteps, a->ds, a->s0r, a->s0i, a->shr, a->shi, a->sbr,
        a->sbi, cp.d, cp.f, cp.qN);
    DYF_LAUNCHED("laue_coef_kernel");
    return DYF_OK;
}

int launch_zero_rows(const dyf_laue_args *a, const PhaseLayout &L, int prec,
                     int n_e, void *out, hipStream_t stream) {
    if (L.nrows >= a->Ny || n_e == 0) return DYF_OK;
    const int elem = prec == DYF_FP64 ? 16 : 8;
    const size_t n = (size_t)(a->Ny - L.nrows) * a->Nx * n_e;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
    zero_rows_kernel<<<blocks, 256, 0, stream>>>(out, elem, n_e, a->Ny, a->Nx,
                                                  L.iy_lo, L.iy_hi);
    DYF_LAUNCHED("zero_rows_kernel");
    return DYF_OK;
}

int launch_laue_persistent(const InstDesc *inst, int n_inst, long long total,
                           int prec, int C, unsigned *counters, int wpb,
                           int blocks, hipStream_t stream) {
    if (n_inst < 1 || total < 1) return DYF_OK;
    if (wpb < 1 || wpb > 4 || blocks < 1)
        return fail(DYF_EINVAL, "bad persistent launch shape (%d waves, %d "
                    "blocks)", wpb, blocks);
    PersistFn fn = nullptr;
    if (prec == DYF_FP64)
        fn = persist_fn<DYF_FP64, 2, 4, 6, 8, 10, 12, 13, 14, 16, 17, 18>(C);
    else if (prec == DYF_FP32)
        fn = persist_fn<DYF_FP32, 2, 4, 6, 8, 10, 12, 13, 14, 16, 17, 18>(C);
    else
        fn = persist_fn<DYF_MIXED, 2, 4, 6, 8, 10, 12, 13, 14, 16, 17, 18>(C);
    if (!fn) return fail(DYF_EKEY, "no persistent kernel for C=%d", C);
    // the pair march keeps one angle's state in LDS: 1 KB per column per wave
    const size_t lds = C <= pair_max_c(prec) ? (size_t)wpb * 1024 * C : 0;
    if (lds > 65536)
        DYF_HIP(hipFuncSetAttribute((const void *)fn,
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
    zero_counters_kernel<<<1, 64, 0, stream>>>(counters);
    DYF_LAUNCHED("zero_counters_kernel");
    const long long per_queue = (total + 7) / 8;
    if (per_queue > 0xffffffffLL) return fail(DYF_EINVAL, "too many items");
    hipLaunchKernelGGL(fn, dim3(blocks), dim3(kWave * wpb), lds, stream, inst,
                       n_inst, total, per_queue, counters);
    DYF_LAUNCHED("laue_persistent_kernel");
    return DYF_OK;
}

// ---- fast-kernel dispatch table ---------------------------------------
using FastFn = void (*)(FastArgs);

template <int MODE, int PREC, int... Cs>
struct FastTable {
    static FastFn get(int C) {
        FastFn f = nullptr;
        if constexpr (MODE == DYF_MODE_GRID && PREC == DYF_FP64)
            ((C == Cs ? (f = laue_fast_inplace_kernel<Cs, PREC>, 0) : 0), ...);
        else if constexpr (MODE == DYF_MODE_GRID)
            ((C == Cs ? (f = laue_fast_unrolled_kernel<Cs, PREC>, 0) : 0), ...);
        else
            ((C == Cs ? (f = laue_fast_kernel<Cs, MODE, PREC>, 0) : 0), ...);
        return f;
    }
};

template <int MODE, int PREC>
static FastFn fast_fn(int C) {
    return FastTable<MODE, PREC, 2, 4, 6, 8, 10, 12, 13, 14, 16, 17, 18>::get(C);
}

using PairFn = void (*)(FastArgs, int, int);

template <int PREC, int... Cs>
static PairFn pair_fn(int C) {
    PairFn f = nullptr;
    ((C == Cs ? (f = laue_fast_pair_kernel<Cs, PREC>, 0) : 0), ...);
    return f;
}

// The pair march's window [g0, g0 + 32) and its step K: the largest K with
// every lane outside the window still before its first step in the fan.
static void pair_window(const PhaseLayout &L, int *K, int *g0) {
    int bestK = 0, bestg = 0;
    for (int g = 0; g <= kWave / 2; ++g) {
        int k = L.nsteps;
        for (int l = 0; l < kWave; ++l)
            if (l < g || l >= g + kWave / 2)
                k = std::min(k, lane_first_step(l, L.C, L.fan_lo, L.fan_hi));
        if (k > bestK) bestK = k, bestg = g;
    }
    *K = bestK, *g0 = bestg;
}

// Whether a fast-path march pairs angles, and its (K, g0) if so (K = 0:
// one angle per wavefront).  Grid mode, two angles or more, while the fan
// fits half a wavefront for at least an eighth of the steps, and when
// there are at least 6 pairs per CU (1.5 pair wavefronts per SIMD):
// halving the wavefronts of a smaller launch costs more than the pairing
// saves.  Measured on the 273 um rocking-scan grid (256 rows, C = 17),
// march with / without pairs: 8 angles (4 pairs per CU) FP64 1.01 / 0.83
// ms, 12 angles (6 per CU) 1.15 / 1.17 ms (mixed 0.88 / 1.10), 16 angles
// 1.39 / 1.48 (mixed 0.99 / 1.24), 24 angles 1.85 / 2.15; the 2-angle
// worked example (481 pairs) 0.55 / 0.46 ms.  Not at C = 18, where the
// pair march spills registers.
static int pair_plan(const PhaseLayout &L, int mode, int n_e, int *K,
                     int *g0) {
    *K = 0, *g0 = 0;
    if (L.path != DYF_PATH_FAST || mode != DYF_MODE_GRID || n_e < 2 ||
        L.C >= 18 || L.nrows == 0)
        return DYF_OK;
    int dev = 0, cus = 256;
    DYF_HIP(hipGetDevice(&dev));
    DYF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if ((long long)L.nrows * ((n_e + 1) / 2) < 6LL * cus) return DYF_OK;
    int k = 0, g = 0;
    pair_window(L, &k, &g);
    if (k >= 1 && 8 * k >= L.nsteps) *K = k, *g0 = g;
    return DYF_OK;
}

void laue_pair_window(const PhaseLayout &L, int *K, int *g0) {
    pair_window(L, K, g0);
}

static FastFn pick_fast(int mode, int prec, int C) {
#define DYF_PICK(M, P) \
    if (mode == M && prec == P) return fast_fn<M, P>(C);
    DYF_PICK(DYF_MODE_GRID, DYF_FP64)
    DYF_PICK(DYF_MODE_GRID, DYF_FP32)
    DYF_PICK(DYF_MODE_GRID, DYF_MIXED)
    DYF_PICK(DYF_MODE_ANALYTIC, DYF_FP64)
    DYF_PICK(DYF_MODE_ANALYTIC, DYF_FP32)
    DYF_PICK(DYF_MODE_ANALYTIC, DYF_MIXED)
    DYF_PICK(DYF_MODE_PRISTINE, DYF_FP64)
    DYF_PICK(DYF_MODE_PRISTINE, DYF_FP32)
    DYF_PICK(DYF_MODE_PRISTINE, DYF_MIXED)
#undef DYF_PICK
    return nullptr;
}

}  // namespace dyf

using namespace dyf;

extern "C" int dyf_laue_plan(const dyf_laue_args *a, int prec, int *path,
                             int *cols, int *iy_lo, int *iy_hi) {
    PhaseLayout L;
    int rc = phase_layout(a, DYF_MODE_GRID, prec, &L);
    if (rc) return rc;
    if (path) *path = L.path;
    if (cols) *cols = L.path == DYF_PATH_FAST ? L.C : 0;
    if (iy_lo) *iy_lo = L.iy_lo;
    if (iy_hi) *iy_hi = L.iy_hi;
    return DYF_OK;
}

extern "C" int dyf_laue_phase_bytes(const dyf_laue_args *a, int mode, int prec,
                                    size_t *bytes) {
    PhaseLayout L;
    int rc = phase_layout(a, mode, prec, &L);
    if (rc) return rc;
    if (!bytes) return fail(DYF_EINVAL, "bytes is NULL");
    *bytes = L.bytes;
    return DYF_OK;
}

extern "C" int dyf_laue_coef_bytes(const dyf_laue_args *a, int n_e,
                                   size_t *bytes) {
    int rc = check_laue_args(a);
    if (rc) return rc;
    if (n_e < 0 || !bytes) return fail(DYF_EINVAL, "bad n_e or bytes");
    *bytes = coef_bytes(n_e, a->Nz - 1);
    return DYF_OK;
}

extern "C" int dyf_laue_march_fraction(const dyf_laue_args *a, int mode,
                                       int prec, int n_e, double *fraction) {
    PhaseLayout L;
    int rc = phase_layout(a, mode, prec, &L);
    if (rc) return rc;
    if (n_e < 0 || !fraction) return fail(DYF_EINVAL, "bad n_e or fraction");
    *fraction = 1.0;
    int K = 0, g0 = 0;
    if ((rc = pair_plan(L, mode, n_e, &K, &g0))) return rc;
    if (K == 0 || n_e == 0) return DYF_OK;
    // wavefront-steps of the paired launch over those of one angle each
    const double pairs = (double)(n_e / 2), ns = (double)L.nsteps;
    const double odd = (n_e % 2) ? ns : 0.0;
    *fraction = (pairs * (2.0 * ns - K) + odd) / ((double)n_e * ns);
    return DYF_OK;
}

extern "C" int dyf_laue_march(const dyf_laue_args *a, int mode, int prec,
                              const void *src, int H_prec,
                              const dyf_line_params *line,
                              const double *s_dev, int n_e, void *coef,
                              void *out, void *stream_) {
    hipStream_t stream = (hipStream_t)stream_;
    PhaseLayout L;
    int rc = phase_layout(a, mode, prec, &L);
    if (rc) return rc;
    if (n_e < 0) return fail(DYF_EINVAL, "n_e < 0");
    if (n_e == 0) return DYF_OK;
    if (!out || !s_dev) return fail(DYF_EINVAL, "NULL output or s_dev");
    // every argument is checked before the first launch, so an error
    // leaves nothing enqueued (the reference raises before launching:
    // laue/kernels.py:365-386)
    const size_t lds = (size_t)2 * 4 * (a->Nx + 1) * sizeof(double);
    if (L.path == DYF_PATH_FAST && L.nrows > 0) {
        if (mode != DYF_MODE_PRISTINE && !src)
            return fail(DYF_EINVAL, "phase table is NULL");
        if (!coef) return fail(DYF_EINVAL, "coefficient workspace is NULL");
        if (!pick_fast(mode, prec, L.C))
            return fail(DYF_EKEY, "no fast kernel for C=%d", L.C);
    } else if (L.nrows > 0) {
        if (mode == DYF_MODE_GRID && !src)
            return fail(DYF_EINVAL, "H grid is NULL");
        if (mode == DYF_MODE_GRID && H_prec != DYF_FP64 && H_prec != DYF_FP32)
            return fail(DYF_EKEY, "H precision must be DYF_FP64 or DYF_FP32");
        if (mode == DYF_MODE_ANALYTIC && !line)
            return fail(DYF_EINVAL, "line parameters are NULL");
        if (lds > 160 * 1024)
            return fail(DYF_ENOMEM,
                        "the general march needs %zu KB of LDS per workgroup "
                        "but the device allows 160 KB; reduce the x-grid",
                        lds / 1024);
    }
    // the march's form is settled before the first launch too
    int K = 0, g0 = 0;
    if (int prc = pair_plan(L, mode, n_e, &K, &g0)) return prc;
    const int elem = prec == DYF_FP64 ? 16 : 8;
    const int nrows = L.nrows;
    if (nrows < a->Ny) {
        const size_t n = (size_t)(a->Ny - nrows) * a->Nx * n_e;
        const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
        zero_rows_kernel<<<blocks, 256, 0, stream>>>(out, elem, n_e, a->Ny,
                                                      a->Nx, L.iy_lo, L.iy_hi);
        DYF_LAUNCHED("zero_rows_kernel");
    }
    if (nrows == 0) return DYF_OK;
    if (L.path == DYF_PATH_FAST) {
        const CoefPtrs cp = coef_ptrs(coef, n_e, L.nsteps);
        const int ncoef = n_e * L.nsteps;
        laue_coef_kernel<<<(ncoef + 255) / 256, 256, 0, stream>>>(
            s_dev, n_e, L.nsteps, a->ds, a->s0r, a->s0i, a->shr, a->shi,
            a->sbr, a->sbi, cp.d, cp.f, cp.qN);
        DYF_LAUNCHED("laue_coef_kernel");
        FastArgs fa;
        fa.phase = src;
        fa.cd = cp.d;
        fa.cf = cp.f;
        fa.qN = cp.qN;
        fa.out = out;
        fa.n_e = n_e;
        fa.nrows = nrows;
        fa.iy_lo = L.iy_lo;
        fa.nsteps = L.nsteps;
        fa.Nx = a->Nx;
        fa.Ny = a->Ny;
        fa.total = nrows * n_e;
        fa.per_xcd = (fa.total + 7) / 8;
        fa.fan_lo = L.fan_lo;
        fa.fan_hi = L.fan_hi;
        fa.dx = a->dx;
        fa.sheet_hw = a->sheet_hw;
        if (K > 0) {
            PairFn pf = nullptr;
            if (prec == DYF_FP64)
                pf = pair_fn<DYF_FP64, 2, 4, 6, 8, 10, 12, 13, 14, 16, 17>(L.C);
            else if (prec == DYF_FP32)
                pf = pair_fn<DYF_FP32, 2, 4, 6, 8, 10, 12, 13, 14, 16, 17>(L.C);
            else
                pf = pair_fn<DYF_MIXED, 2, 4, 6, 8, 10, 12, 13, 14, 16, 17>(L.C);
            if (!pf) return fail(DYF_EKEY, "no pair march for C=%d", L.C);
            fa.total = nrows * ((n_e + 1) / 2);
            fa.per_xcd = (fa.total + 7) / 8;
            const size_t lds = (size_t)L.C * 1024;   // 4 doubles or 8 floats per lane
            hipLaunchKernelGGL(pf, dim3(8 * fa.per_xcd), dim3(kWave), lds,
                               stream, fa, K, g0);
            DYF_LAUNCHED("laue_fast_pair_kernel");
            return DYF_OK;
        }
        FastFn fn = pick_fast(mode, prec, L.C);
        hipLaunchKernelGGL(fn, dim3(8 * fa.per_xcd), dim3(kWave), 0, stream, fa);
        DYF_LAUNCHED("laue_fast_kernel");
        return DYF_OK;
    }
    // general path
    GeneralArgs ga;
    ga.H = src;
    ga.H_f32 = H_prec == DYF_FP32;
    if (line) ga.line = *line;
    ga.s_dev = s_dev;
    ga.out = out;
    ga.out_f32 = prec != DYF_FP64;
    ga.n_e = n_e;
    ga.iy_lo = L.iy_lo;
    ga.nrows = nrows;
    ga.Nz = a->Nz;
    ga.Nx = a->Nx;
    ga.Ny = a->Ny;
    ga.s0r = a->s0r, ga.s0i = a->s0i, ga.shr = a->shr, ga.shi = a->shi;
    ga.sbr = a->sbr, ga.sbi = a->sbi;
    ga.dz = a->dz, ga.ds = a->ds, ga.dx = a->dx, ga.shift = a->shift_px;
    ga.sheet_hw = a->sheet_hw;
    dim3 grid(nrows * n_e);
    if (mode == DYF_MODE_GRID) {
        DYF_HIP(hipFuncSetAttribute((const void *)laue_general_kernel<DYF_MODE_GRID>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
        hipLaunchKernelGGL(laue_general_kernel<DYF_MODE_GRID>, grid, dim3(256),
                           lds, stream, ga);
    } else if (mode == DYF_MODE_ANALYTIC) {
        DYF_HIP(hipFuncSetAttribute(
            (const void *)laue_general_kernel<DYF_MODE_ANALYTIC>,
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(laue_general_kernel<DYF_MODE_ANALYTIC>, grid,
                           dim3(256), lds, stream, ga);
    } else {
        DYF_HIP(hipFuncSetAttribute(
            (const void *)laue_general_kernel<DYF_MODE_PRISTINE>,
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(laue_general_kernel<DYF_MODE_PRISTINE>, grid,
                           dim3(256), lds, stream, ga);
    }
    DYF_LAUNCHED("laue_general_kernel");
    return DYF_OK;
}

// ---- The march without a caller workspace ---------------------------------
//
// dyf_laue_march takes its per-(angle, step) coefficient table from a
// workspace the caller allocates.  Callers that cannot allocate (code under
// stream capture, tight serving loops) use dyf_laue_march_noalloc, which
// keeps that table in `scratch`, a file-scope device array each GPU holds
// for the life of the process.  Host code never passes `scratch` itself to a
// kernel (on the host the name is only a stand-in for the device symbol);
// the device address of each GPU's copy is looked up once per device with
// hipGetSymbolAddress and kept in scratch_dev[].  The table is written in
// full by laue_coef_kernel before the march reads it, so nothing needs
// clearing between calls; calls that overlap on one GPU (different streams)
// share the array and race.

namespace dyf {

constexpr size_t kNoAllocScratchBytes = size_t(4) << 20;   // 4 MiB
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex scratch_mu;
static void *scratch_dev[kMaxDevices];

// The device address of `scratch` on the current device (looked up on the
// first call made with that device current).
static int scratch_address(void **addr) {
    int dev = -1;
    DYF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return fail(DYF_EINVAL, "device %d outside the %d the library tracks",
                    dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return fail(DYF_EHIP, "hipGetSymbolAddress(scratch) on device "
                        "%d: %s", dev, hipGetErrorString(e));
        }
        scratch_dev[dev] = p;
    }
    *addr = scratch_dev[dev];
    return DYF_OK;
}

}  // namespace dyf

extern "C" size_t dyf_laue_march_noalloc_max_bytes(void) {
    return dyf::kNoAllocScratchBytes;
}

extern "C" int dyf_laue_march_noalloc(const dyf_laue_args *a, int mode,
                                      int prec, const void *src, int H_prec,
                                      const dyf_line_params *line,
                                      const double *s_dev, int n_e,
                                      void *out, void *stream) {
    int rc = check_laue_args(a);
    if (rc) return rc;
    if (n_e < 0) return fail(DYF_EINVAL, "n_e < 0");
    const size_t need = coef_bytes(n_e, a->Nz - 1);
    if (need > kNoAllocScratchBytes)
        return fail(DYF_EINVAL,
                    "%d angles x %d steps need %zu bytes of coefficient "
                    "scratch, above this entry point's %zu; march fewer "
                    "angles per call or use dyf_laue_march with a workspace",
                    n_e, a->Nz - 1, need, kNoAllocScratchBytes);
    void *ws = nullptr;
    rc = scratch_address(&ws);
    if (rc) return rc;
    return dyf_laue_march(a, mode, prec, src, H_prec, line, s_dev, n_e, ws,
                          out, stream);
}
