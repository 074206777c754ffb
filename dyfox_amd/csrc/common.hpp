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
    int32_t nwind;          // g.b rounded, when it is an integer in
                            // [-8, 8] (within 1e-9); 0 otherwise
    int32_t frac;           // 1 when g.b is not such an integer (its
                            // winding is taken with atan2)
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

// d = a b + c in one v_fma_f64 with c in a scalar register.  Written
// out for polynomial steps: the compiler otherwise selects the two-address
// v_fmac_f64 and copies the (loop-invariant) coefficient into the
// destination before every use.
__device__ __forceinline__ double fma_sc(double a, double b, double c) {
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
    return d;
}

// Table for log_tab: entry j covers mantissas m in [1 + j/128,
// 1 + (j+1)/128) with c = 1/(1 + (j + 1/2)/128) (rounded) and
// nl = -log(c) of that rounded c.  Built per workgroup in LDS by
// fill_log_table (one library log per entry).
constexpr int kLogTab = 128;
struct alignas(16) LogEntry {
    double c, nl;
};

__device__ __forceinline__ void fill_log_table(LogEntry *tab) {
    for (int j = threadIdx.x; j < kLogTab; j += blockDim.x) {
        const double c = 1.0 / (1.0 + ((double)j + 0.5) / kLogTab);
        tab[j] = LogEntry{c, -log(c)};
    }
}

// Natural logarithm of a positive, normal x, to about one ulp:
// x = 2^e m, m in [1, 2); with the table entry (c, nl) of m's top seven
// fraction bits, r = m c - 1 (one fma; |r| <= 2^-8) and
// log x = e ln 2 + nl + log1p(r), log1p(r) = r - r^2/2 + ... - r^6/6 +
// r^7/7 (the first dropped term is below 2e-18 of r).
__device__ __forceinline__ double log_tab(double x,
                                          const LogEntry *__restrict__ tab) {
    const unsigned long long bits = __double_as_longlong(x);
    const unsigned hi = (unsigned)(bits >> 32);
    const int e = (int)(hi >> 20) - 1023;
    const LogEntry t = tab[(hi >> 13) & (kLogTab - 1)];
    const double m = __longlong_as_double(
        (bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    const double r = fma(m, t.c, -1.0);
    double q = fma(r, 1.0 / 7.0, -1.0 / 6.0);
    q = fma_sc(q, r, 1.0 / 5.0);
    q = fma_sc(q, r, -1.0 / 4.0);
    q = fma_sc(q, r, 1.0 / 3.0);
    q = fma_sc(q, r, -1.0 / 2.0);
    const double lm = fma(r * r, q, r) + t.nl;
    // ln 2 = kLn2Hi + kLn2Lo to 3e-31: kLn2Hi is ln 2 rounded to 42
    // bits, so e * kLn2Hi is exact for every double exponent e.
    constexpr double kLn2Hi = 0x1.62e42fefa3800p-1;
    constexpr double kLn2Lo = 0x1.ef35793c76730p-45;
    const double de = (double)e;
    return fma(de, kLn2Hi, fma(de, kLn2Lo, lm));
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
//
// NP points share (x, y) and differ in z: a segment's parameters, its
// (uniform) branches and the x, y parts of d1, d2 are taken once for all
// NP.  `tab` is the LDS table of log_tab.
template <int NP>
__device__ __forceinline__ void volterra_phase_z(
    const SegParams *__restrict__ s, int nseg, double x, double y,
    const double (&z)[NP], double a2, const LogEntry *__restrict__ tab,
    double (&Pr)[NP], double (&Pi)[NP], double (&R)[NP]) {
    constexpr double kTiny = 0x1p-1022;
#pragma unroll
    for (int j = 0; j < NP; ++j) Pr[j] = 1.0, Pi[j] = 0.0, R[j] = 0.0;
    a2 = fmax(a2, kTiny);
    const double inv_a2 = 1.0 / a2;
    for (int n = 0; n < nseg; ++n) {
        const SegParams &p = s[n];
        const double X = x - p.cx, Y = y - p.cy;
        const double b1 = fma(Y, p.f1y, X * p.f1x);
        const double b2 = fma(Y, p.f2y, X * p.f2x);
        double d1[NP], d2[NP], r2[NP], c2[NP], sh[NP];
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            const double Z = z[j] - p.cz;
            d1[j] = fma(Z, p.f1z, b1);
            d2[j] = fma(Z, p.f2z, b2);
            r2[j] = fma(d1[j], d1[j], d2[j] * d2[j]);
            const double rc = fmax(r2[j], kTiny);
            double q = __builtin_amdgcn_rcp(rc);
            q = fma(q, fma(-rc, q, 1.0), q);
            q = fma(q, fma(-rc, q, 1.0), q);
            c2[j] = fma(d1[j], d1[j], -d2[j] * d2[j]) * q;    // 0 on the line
            sh[j] = d1[j] * d2[j] * q;                       // s / 2
        }
        const int mw = p.nwind;
        if (mw != 0) {
            const double sg = mw < 0 ? -1.0 : 1.0;
            const int m = mw < 0 ? -mw : mw;
#pragma unroll
            for (int j = 0; j < NP; ++j) {
                const bool on = r2[j] >= kTiny;
                const double br = on ? c2[j] : 1.0, bi = sg * (sh[j] + sh[j]);
                double wr, wi;
                int k;
                if (m & 1) {
                    const double rs = rsqrt(r2[j]);
                    wr = on ? d1[j] * rs : 1.0;
                    wi = on ? sg * d2[j] * rs : 0.0;
                    k = m - 1;
                } else {
                    wr = br, wi = bi;
                    k = m - 2;
                }
                for (; k >= 2; k -= 2) {
                    const double tr = wr * br - wi * bi;
                    wi = wr * bi + wi * br;
                    wr = tr;
                }
                const double tr = Pr[j] * wr - Pi[j] * wi;
                Pi[j] = Pr[j] * wi + Pi[j] * wr;
                Pr[j] = tr;
            }
        } else if (p.frac) {
#pragma unroll
            for (int j = 0; j < NP; ++j) R[j] += p.c_atan * atan2(d2[j], d1[j]);
        }
        if (p.edge != 0.0) {
#pragma unroll
            for (int j = 0; j < NP; ++j) {
                const double t = fma(p.k_xy, sh[j], p.k_dd * c2[j]);
                const double core = fmin(r2[j] * inv_a2, 1.0);
                R[j] = fma(p.k_log, log_tab(fmax(r2[j], a2), tab), R[j]);
                R[j] = fma(t, core, R[j]);
            }
        }
    }
}

__device__ __forceinline__ void volterra_phase(const SegParams *__restrict__ s,
                                               int nseg, double x, double y,
                                               double z, double a2,
                                               const LogEntry *__restrict__ tab,
                                               double &Pr, double &Pi,
                                               double &R) {
    const double zz[1] = {z};
    double pr[1], pi[1], rr[1];
    volterra_phase_z<1>(s, nseg, x, y, zz, a2, tab, pr, pi, rr);
    Pr = pr[0], Pi = pi[0], R = rr[0];
}

// The library's sincos out of line, for the rare argument the reduction
// below does not take: inlined, the compiler runs its Payne-Hanek branch
// unconditionally on every lane.
__device__ __attribute__((noinline)) inline cx<double> sincos_lib(double x) {
    double s, c;
    sincos(x, &s, &c);
    return {c, s};
}

// sin and cos of x to about one ulp: x = n pi/2 + r with n the nearest
// integer to 2x/pi and pi/2 = kPio2Hi + kPio2Lo (a double-double, to
// 1e-33 relative; with fma the two subtractions leave r good to an ulp
// for any |x| < 2^30), then Taylor series on |r| <= pi/4 (sin through
// r^17/17!, cos through r^16/16!: the first dropped terms are below
// 1e-19) and the quadrant n mod 4.  Larger |x|, infinities and NaN take
// the library routine.
__device__ __forceinline__ void sincos_red(double x, double &sn, double &cs) {
    if (!(fabs(x) < 0x1p30)) {
        const cx<double> e = sincos_lib(x);
        sn = e.i, cs = e.r;
        return;
    }
    constexpr double k2OverPi = 0x1.45f306dc9c883p-1;
    constexpr double kPio2Hi = 0x1.921fb54442d18p+0;
    constexpr double kPio2Lo = 0x1.1a62633145c07p-54;
    const double n = rint(x * k2OverPi);
    double r = fma(-n, kPio2Hi, x);
    r = fma(-n, kPio2Lo, r);
    const double u = r * r;
    double ps = fma(u, 1.0 / 355687428096000.0,                // 1/17!
                    -1.0 / 1307674368000.0);                    // 1/15!
    ps = fma_sc(ps, u, 1.0 / 6227020800.0);                     // 1/13!
    ps = fma_sc(ps, u, -1.0 / 39916800.0);                      // 1/11!
    ps = fma_sc(ps, u, 1.0 / 362880.0);                         // 1/9!
    ps = fma_sc(ps, u, -1.0 / 5040.0);                          // 1/7!
    ps = fma_sc(ps, u, 1.0 / 120.0);                            // 1/5!
    ps = fma_sc(ps, u, -1.0 / 6.0);                             // 1/3!
    const double sr = fma(r * u, ps, r);
    double pc = fma(u, 1.0 / 20922789888000.0,                 // 1/16!
                    -1.0 / 87178291200.0);                      // 1/14!
    pc = fma_sc(pc, u, 1.0 / 479001600.0);                      // 1/12!
    pc = fma_sc(pc, u, -1.0 / 3628800.0);                       // 1/10!
    pc = fma_sc(pc, u, 1.0 / 40320.0);                          // 1/8!
    pc = fma_sc(pc, u, -1.0 / 720.0);                           // 1/6!
    pc = fma_sc(pc, u, 1.0 / 24.0);                             // 1/4!
    const double cr = fma(u * u, pc, fma(-0.5, u, 1.0));
    const int qd = (int)(long long)n;
    const bool swap = qd & 1;
    const double a = swap ? cr : sr, b = swap ? sr : cr;
    sn = (qd & 2) ? -a : a;
    cs = ((qd + 1) & 2) ? -b : b;
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
    sincos_red(R, s, c);
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
