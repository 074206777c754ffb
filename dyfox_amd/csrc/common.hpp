// Shared helpers of libdyfox: status codes, the thread-local error string,
// HIP error checks and small complex-arithmetic helpers for device code.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../include/dyfox.h"

namespace dyf {

// Sets the thread-local message returned by dyf_last_error and returns `code`.
int fail(int code, const char *fmt, ...);

// Maps a HIP status to DYF_EHIP (or DYF_ENOMEM) with the call site named.
int hip_fail(hipError_t e, const char *what);

#define DYF_HIP(call)                                      \
    do {                                                   \
        hipError_t e_ = (call);                            \
        if (e_ != hipSuccess) return dyf::hip_fail(e_, #call); \
    } while (0)

// Checks the launch status of the kernel just enqueued.
#define DYF_LAUNCHED(name)                                 \
    do {                                                   \
        hipError_t e_ = hipGetLastError();                 \
        if (e_ != hipSuccess) return dyf::hip_fail(e_, name); \
    } while (0)

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 6.28318530717958647692;
constexpr int kWave = 64;

// Complex pair in registers; plain struct so it stays in VGPRs.
template <typename T>
struct cx {
    T r, i;
};

// One Volterra segment reduced to what a voxel evaluation needs (FP64).
// Built on the host from (ra, rb, b) by prepare_segments.
struct SegParams {
    double cx, cy, cz;      // line point (segment midpoint)
    double xx, xy, xz;      // unit line direction xi
    double e1x, e1y, e1z;   // unit edge direction
    double e2x, e2y, e2z;   // xi x e1
    double c_atan;          // g.e1 b_perp + g.xi b_screw  (= g.b)
    double k_xy;            // 2 pi g.e1 c1 / (2(1-nu))        * d1 d2 / r2c
    double k_dd;            // -2 pi g.e2 c1 / (4(1-nu))       * (d1^2-d2^2)/r2c
    double k_log;           // -2 pi g.e2 c1 (1-2nu)/(4(1-nu)) * ln r2c
    double edge;            // 1 when b_perp > 1e-20
    double wind;            // g.b rounded, when it is an integer in
                            // [-8, 8] (within 1e-9); NaN otherwise
    double f1x, f1y, f1z;   // e1 - (e1.xi) xi: d1 = f1.(r - c) equals
    double f2x, f2y, f2z;   // e1.(r - c projected off xi); same for e2
};

// Reduces raw (ra, rb, b) segments to SegParams; segments shorter than
// 1e-20 m are dropped, as the reference does.  Returns the count kept.
int prepare_segments(const double *segs, int nseg, const double *g_vec,
                     double nu, SegParams *out);

// Workgroup barrier for LDS traffic only: waits for this wave's LDS
// operations, then s_barrier.  __syncthreads() also fences global memory,
// which drains the vector-memory counter and so waits for every global
// load still in flight - including loads issued early to prefetch the
// next step - on every step.  Only for loops whose cross-thread data goes
// through LDS (global memory read-only, or written and read back by the
// same thread).
__device__ __forceinline__ void lds_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// H = 2 pi g.u at (x, y, z), segments summed in list order (FP64).
__device__ __forceinline__ double volterra_h(const SegParams *__restrict__ s,
                                             int nseg, double x, double y,
                                             double z, double a2) {
    double H = 0.0;
    for (int n = 0; n < nseg; ++n) {
        const SegParams &p = s[n];
        const double X = x - p.cx, Y = y - p.cy, Z = z - p.cz;
        const double eta = X * p.xx + Y * p.xy + Z * p.xz;
        const double Px = X - eta * p.xx, Py = Y - eta * p.xy,
                     Pz = Z - eta * p.xz;
        const double d1 = Px * p.e1x + Py * p.e1y + Pz * p.e1z;
        const double d2 = Px * p.e2x + Py * p.e2y + Pz * p.e2z;
        double Hc = p.c_atan * atan2(d2, d1);
        if (p.edge != 0.0) {
            const double r2c = fmax(d1 * d1 + d2 * d2, a2);
            const double inv = 1.0 / r2c;
            Hc += (p.k_xy * (d1 * d2) + p.k_dd * (d1 * d1 - d2 * d2)) * inv +
                  p.k_log * log(r2c);
        }
        H += Hc;
    }
    return H;
}

// Natural logarithm of a positive, normal x, to about one ulp:
// x = 2^e m with m in [sqrt(1/2), sqrt(2)), and log m = 2 atanh(s),
// s = (m - 1)/(m + 1), |s| <= 0.1716, summed as the odd series
// 2 (s + s^3/3 + ... + s^19/19); the first dropped term is below 2^-55
// of the sum.  m - 1 is exact (Sterbenz), 1/(m + 1) takes one Newton step
// from v_rcp_f64 and s one residual correction.
__device__ __forceinline__ double log_pos(double x) {
    int e;
    double m = frexp(x, &e);                  // m in [1/2, 1)
    const bool lo = m < 0.70710678118654752440;
    m = lo ? m + m : m;
    e = lo ? e - 1 : e;
    const double f = m - 1.0, d = m + 1.0;
    double r = __builtin_amdgcn_rcp(d);
    r = fma(r, fma(-d, r, 1.0), r);
    double sv = f * r;
    sv = fma(r, fma(-sv, d, f), sv);
    const double u = sv * sv;
    double q = 1.0 / 19.0;
    q = fma(q, u, 1.0 / 17.0);
    q = fma(q, u, 1.0 / 15.0);
    q = fma(q, u, 1.0 / 13.0);
    q = fma(q, u, 1.0 / 11.0);
    q = fma(q, u, 1.0 / 9.0);
    q = fma(q, u, 1.0 / 7.0);
    q = fma(q, u, 1.0 / 5.0);
    q = fma(q, u, 1.0 / 3.0);
    const double two_s = sv + sv;
    const double logm = fma(two_s * u, q, two_s);
    // ln 2 = kLn2Hi + kLn2Lo to 3e-31: kLn2Hi is ln 2 rounded to 42
    // bits, so e * kLn2Hi is exact for every double exponent e.
    constexpr double kLn2Hi = 0x1.62e42fefa3800p-1;
    constexpr double kLn2Lo = 0x1.ef35793c76730p-45;
    const double de = (double)e;
    return fma(de, kLn2Hi, fma(de, kLn2Lo, logm));
}

// e^{iH} at (x, y, z) as P e^{iR}: P = prod (w/|w|)^(g.b) over the
// segments whose g.b is an integer (w = d1 + i d2, so (w/|w|)^n =
// e^{i n atan2(d2, d1)} without the atan2), R = the remaining terms (the
// atan2 winding of non-integer g.b and the edge terms).  Equal to
// e^{i volterra_h} up to rounding; P is a unit complex number.
// One reciprocal of r^2 = d1^2 + d2^2 serves both: with
// (c, s) = ((d1^2 - d2^2), 2 d1 d2) / r^2 = (w/|w|)^2, the edge terms are
// k_xy s/2 + k_dd c, scaled by r^2/a^2 inside the core floor a, and even
// windings are powers of (c, s) (odd ones take one more factor w/|w|).
// A voxel on the line (r^2 = 0) gets the factor 1 (atan2(0, 0) = 0) and
// no k_xy/k_dd term.  The core floor is taken as at least the smallest
// normal double, so a zero core radius leaves a voxel on the line finite
// (log of 2^-1022) where the unfloored formula has 0/0.
__device__ __forceinline__ void volterra_phase(const SegParams *__restrict__ s,
                                               int nseg, double x, double y,
                                               double z, double a2,
                                               double &Pr, double &Pi,
                                               double &R) {
    Pr = 1.0, Pi = 0.0, R = 0.0;
    a2 = fmax(a2, 0x1p-1022);
    const double inv_a2 = 1.0 / a2;
    for (int n = 0; n < nseg; ++n) {
        const SegParams &p = s[n];
        const double X = x - p.cx, Y = y - p.cy, Z = z - p.cz;
        const double d1 = X * p.f1x + Y * p.f1y + Z * p.f1z;
        const double d2 = X * p.f2x + Y * p.f2y + Z * p.f2z;
        const double r2 = fma(d1, d1, d2 * d2);
        const bool on = r2 >= 0x1p-1022;
        double q = __builtin_amdgcn_rcp(r2);
        q = fma(q, fma(-r2, q, 1.0), q);
        q = fma(q, fma(-r2, q, 1.0), q);
        const double c2 = on ? fma(d1, d1, -d2 * d2) * q : 1.0;
        const double sh = on ? d1 * d2 * q : 0.0;        // s / 2
        if (p.wind == p.wind) {
            int m = (int)p.wind;
            if (m != 0) {
                const double sg = m < 0 ? -1.0 : 1.0;
                m = m < 0 ? -m : m;
                const double br = c2, bi = sg * (sh + sh);
                double wr, wi;
                if (m & 1) {
                    const double rs = rsqrt(r2);
                    wr = on ? d1 * rs : 1.0;
                    wi = on ? sg * d2 * rs : 0.0;
                } else {
                    wr = br, wi = bi;
                    m -= 2;
                }
                for (; m >= 2; m -= 2) {
                    const double tr = wr * br - wi * bi;
                    wi = wr * bi + wi * br;
                    wr = tr;
                }
                const double tr = Pr * wr - Pi * wi;
                Pi = Pr * wi + Pi * wr;
                Pr = tr;
            }
        } else {
            R += p.c_atan * atan2(d2, d1);
        }
        if (p.edge != 0.0) {
            const double t = fma(p.k_xy, sh, p.k_dd * c2);
            const double core = fmin(r2 * inv_a2, 1.0);
            R += fma(t, core, p.k_log * log_pos(fmax(r2, a2)));
        }
    }
}

// FP32 segment accumulation (the reference's FP32 H grid: each segment's
// FP64 value is rounded, then summed in float).
__device__ __forceinline__ float volterra_h_f32acc(
    const SegParams *__restrict__ s, int nseg, double x, double y, double z,
    double a2) {
    float H = 0.0f;
    for (int n = 0; n < nseg; ++n) H += (float)volterra_h(s + n, 1, x, y, z, a2);
    return H;
}

// e^{iH} rounded to float after an FP64 reduction of H into [-pi, pi].
__device__ __forceinline__ cx<float> phase_f32(double H) {
    const double n = rint(H * (1.0 / kTwoPi));
    const float r = (float)fma(-n, kTwoPi, H);
    float s, c;
    sincosf(r, &s, &c);
    return {c, s};
}

__device__ __forceinline__ cx<double> phase_f64(double H) {
    double s, c;
    sincos(H, &s, &c);
    return {c, s};
}

// P e^{iR} (the parts of volterra_phase) in FP64, and in FP32 after an
// FP64 reduction of R into [-pi, pi].
__device__ __forceinline__ cx<double> phase_parts_f64(double Pr, double Pi,
                                                      double R) {
    double s, c;
    sincos(R, &s, &c);
    return {Pr * c - Pi * s, Pr * s + Pi * c};
}

__device__ __forceinline__ cx<float> phase_parts_f32(double Pr, double Pi,
                                                     double R) {
    const double n = rint(R * (1.0 / kTwoPi));
    const float r = (float)fma(-n, kTwoPi, R);
    float s, c;
    sincosf(r, &s, &c);
    const float pr = (float)Pr, pi = (float)Pi;
    return {pr * c - pi * s, pr * s + pi * c};
}

// Deterministic, segment-independent grid description for kernels.
struct RowBand {
    int iy_lo, iy_hi;
};

// Marched-row band of a Laue grid: |(iy - Ny/2) dx| < fov_hy in FP64.
RowBand laue_rows(const dyf_laue_args *a);

}  // namespace dyf
