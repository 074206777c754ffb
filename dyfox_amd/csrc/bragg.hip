// Bragg (reflection) shooting solve on gfx950: the down-march as a batched
// matvec, the perfect-crystal Fourier preconditioner, and restarted GMRES
// whose Arnoldi basis, orthogonalisation and Givens rotations live on the
// device (reference bragg/solver.py:28-60, 182-379; SciPy 1.15 gmres).
//
// A batch is `nsys` independent systems (one s_dev each) over the same
// chunk of `nyc` y-planes, so they share the e^{iH} table E (Nz, nyc, Nx).
// Vectors are (nsys, nyc, Nx) complex128, system-major.
//
// GMRES follows SciPy's control flow exactly - left preconditioning,
// inner stop at presid <= ptol, true residual b - A x after every cycle,
// the adaptive ptol update, maxiter counted in cycles - so a system
// converges under the same rule and reports the same quantities.  The
// orthogonalisation is classical Gram-Schmidt applied twice (CGS2): two
// batched reductions instead of SciPy's col+1 sequential MGS sweeps,
// equally stable, and the only change is in rounding.
#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "internal.hpp"

namespace dyf {

using cd = cx<double>;

__device__ __forceinline__ cd cmul(cd a, cd b) {
    return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
__device__ __forceinline__ cd cconjmul(cd a, cd b) {   // conj(a) * b
    return {a.r * b.r + a.i * b.i, a.r * b.i - a.i * b.r};
}
__device__ __forceinline__ cd cadd(cd a, cd b) { return {a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ cd cscale(cd a, double s) { return {a.r * s, a.i * s}; }
__device__ __forceinline__ cd cexp(double re, double im) {
    double s, c;
    sincos(im, &s, &c);
    const double m = exp(re);
    return {m * c, m * s};
}
__device__ __forceinline__ cd cdiv(cd a, cd b) {
    const double d = b.r * b.r + b.i * b.i;
    return {(a.r * b.r + a.i * b.i) / d, (a.i * b.r - a.r * b.i) / d};
}

struct BraggK {
    double s0r, s0i, shr, shi, sbr, sbi, ds;
    int Nz, Nx, nyc, nsys;
    int tpl;             // y-planes per depth slice of the e^{iH} table
    int sper;            // systems per table (planes of it when each
                         // system is one plane, else 1)
    long long tstride;   // elements between tables (0: one shared table)
};

// First element of system s's e^{iH} table at depth 0, plane 0; depth
// slices are tpl * Nx apart.
__host__ __device__ inline long long table_offset(const BraggK &k, int s) {
    return (long long)(s / k.sper) * k.tstride +
           (long long)(s % k.sper) * k.nyc * k.Nx;
}

static BraggK make_k(const dyf_bragg_args *a, int nsys) {
    BraggK k;
    k.s0r = a->s0r, k.s0i = a->s0i, k.shr = a->shr, k.shi = a->shi;
    k.sbr = a->sbr, k.sbi = a->sbi, k.ds = a->ds;
    k.Nz = a->Nz, k.Nx = a->Nx, k.nyc = a->nyc, k.nsys = nsys;
    k.sper = a->plane_systems > 0 ? a->plane_systems : 1;
    k.tpl = a->plane_systems > 0 ? a->plane_systems : a->nyc;
    k.tstride = a->per_system_phase ? (long long)a->Nz * k.tpl * a->Nx : 0;
    return k;
}

// ---- the down-march: one workgroup per (system, plane) ----------------
// The reference's step (per source column x, periodic in x)
//   p = D0[x] + ds sbar e^{+iH} Dg[x]  -> D0'[x+1] = e0 p
//   q = Dg[x] - ds sh  e^{-iH} D0[x]   -> Dg'[x-1] = q / eg
// is marched in the interaction picture X_k = D0_k / e0^k,
// Y_k = Dg_k eg^k, where the diagonal factors drop out of the step:
//   X'[x+1] = X[x] + c_k e^{+iH}[x] Y[x],  c_k =  ds sbar (e0 eg)^-k
//   Y'[x-1] = Y[x] + b_k e^{-iH}[x] X[x],  b_k = -ds sh   (e0 eg)^k
// and Dg = eg^-(Nz-1) Y at the end (16 FP64 operations per column-step
// instead of 28).  c_k, b_k are formed per system with exp/sincos of
// k times the exponent (no running products) into LDS.
constexpr int kMarchThreads = 256;
constexpr int kMaxPerThread = 8;          // Nx <= 2048

// LDS of one march workgroup: [X | Y] x 2 buffers of Nx + 1 entries, then
// c_k and b_k for the Nz - 1 steps.
static size_t march_lds_bytes(const BraggK &k) {
    return ((size_t)4 * (k.Nx + 1) + (size_t)2 * std::max(k.Nz - 1, 1)) *
           sizeof(cd);
}

// M = ceil(Nx / 256) columns per thread.  The loop is branch-free: loads
// use a clamped index and out-of-range columns write to a spare LDS slot
// (index Nx), so the compiler can keep the next step's e^{iH} prefetch in
// flight across the step (with guarded loads it waited for every load
// before each column).
template <int M>
__global__ __launch_bounds__(kMarchThreads) void bragg_march_kernel(
    BraggK k, const double *__restrict__ s_dev, const cd *__restrict__ E,
    const cd *__restrict__ inc, const cd *__restrict__ vin, size_t vin_stride,
    cd *__restrict__ out, size_t out_stride, double out_scale,
    const int *__restrict__ active) {
    extern __shared__ cd lds_b[];
    const int s = blockIdx.x / k.nyc, j = blockIdx.x % k.nyc;
    if (active && !active[s]) return;
    E += table_offset(k, s);
    const int Nx = k.Nx, tid = threadIdx.x, P = Nx + 1;
    const int nst = k.Nz - 1;
    // LDS: [X | Y] x 2 buffers of P = Nx + 1 entries, addressed by
    // integer offsets into lds_b (selecting among pointers kept in an
    // array loses the LDS address space and turns every access into a
    // flat one, which then waits for outstanding global loads too), then
    // c_k at 4P and b_k at 4P + nst
    for (int x = tid; x < Nx; x += kMarchThreads) {
        lds_b[x] = inc ? inc[x] : cd{0.0, 0.0};
        lds_b[P + x] = vin ? vin[(size_t)s * vin_stride + (size_t)j * Nx + x]
                           : cd{0.0, 0.0};
    }
    const double alpha = -2.0 * kPi * s_dev[s];
    // e0 eg = exp(lr + i li) per step
    const double lr = 2.0 * k.s0r * k.ds, li = (2.0 * k.s0i + alpha) * k.ds;
    const int oc = 4 * P, ob = 4 * P + nst;
    for (int t = tid; t < nst; t += kMarchThreads) {
        const cd f = cexp(-(double)t * lr, -(double)t * li);   // (e0 eg)^-t
        const cd g = cexp((double)t * lr, (double)t * li);     // (e0 eg)^t
        lds_b[oc + t] = cmul(cd{k.ds * k.sbr, k.ds * k.sbi}, f);
        lds_b[ob + t] = cmul(cd{-k.ds * k.shr, -k.ds * k.shi}, g);
    }
    // per column m: the source column (clamped for loads and LDS reads)
    // and the destinations of the shifted fields (slot Nx when unused)
    int xs[M], to0[M], tog[M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int x = tid + m * kMarchThreads;
        const bool ok = x < Nx;
        xs[m] = ok ? x : Nx - 1;
        to0[m] = ok ? (x + 1 < Nx ? x + 1 : 0) : Nx;
        tog[m] = ok ? (x > 0 ? x - 1 : Nx - 1) : Nx;
    }
    __syncthreads();
    int cur = 0;
    const size_t plane = (size_t)k.tpl * Nx;
    const cd *Ej = E + (size_t)j * Nx;
    cd en[M];
#pragma unroll
    for (int m = 0; m < M; ++m) en[m] = Ej[xs[m]];
    for (int step = 0; step < nst; ++step) {
        cd ec[M];
#pragma unroll
        for (int m = 0; m < M; ++m) ec[m] = en[m];
        const int sn = step + 1 < nst ? step + 1 : step;
        const cd *En = Ej + (size_t)sn * plane;
#pragma unroll
        for (int m = 0; m < M; ++m) en[m] = En[xs[m]];
        const cd ck = lds_b[oc + step], bk = lds_b[ob + step];
        const int o0 = cur * 2 * P, n0 = (1 - cur) * 2 * P;
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const cd X = lds_b[o0 + xs[m]], Y = lds_b[o0 + P + xs[m]];
            const cd e = ec[m];
            const cd u = cmul(ck, e);              // c_k e^{+iH}
            const cd w = cmul(bk, cd{e.r, -e.i});  // b_k e^{-iH}
            lds_b[n0 + to0[m]] = cd{fma(u.r, Y.r, fma(-u.i, Y.i, X.r)),
                                    fma(u.r, Y.i, fma(u.i, Y.r, X.i))};
            lds_b[n0 + P + tog[m]] = cd{fma(w.r, X.r, fma(-w.i, X.i, Y.r)),
                                        fma(w.r, X.i, fma(w.i, X.r, Y.i))};
        }
        lds_barrier();
        cur = 1 - cur;
    }
    // Dg = eg^-(Nz-1) Y, times the caller's scale
    const double egr = k.s0r * k.ds, egi = (k.s0i + alpha) * k.ds;
    const cd fin = cscale(cexp(-(double)nst * egr, -(double)nst * egi),
                          out_scale);
    const int og = cur * 2 * P + P;
    for (int x = tid; x < Nx; x += kMarchThreads)
        out[(size_t)s * out_stride + (size_t)j * Nx + x] =
            cmul(lds_b[og + x], fin);
}

// ---- the same march, one wavefront per (system, plane) ----------------
// For Nx = 64-or-fewer lanes times C with C | Nx: lane l keeps columns
// [lC, lC + C) of X and Y in registers (Lu = Nx / C lanes; the rest idle).
// A step is C column updates in registers, a rename by one register, and
// one ds_bpermute per field for the column crossing a lane boundary; the
// periodic wrap is the same bpermute with lane 0 reading lane Lu - 1 (X)
// and lane Lu - 1 reading lane 0 (Y), because Nx - 1 is the last register
// of lane Lu - 1.  No LDS buffers and no barrier per step (the
// workgroup-per-plane kernel above spends most of a step on its LDS
// round trip and barrier).  c_k, b_k are formed as there, into LDS.
__device__ __forceinline__ cd wave_fetch(cd v, int src4) {
    const long long r = __double_as_longlong(v.r), i = __double_as_longlong(v.i);
    const int rl = __builtin_amdgcn_ds_bpermute(src4, (int)r);
    const int rh = __builtin_amdgcn_ds_bpermute(src4, (int)(r >> 32));
    const int il = __builtin_amdgcn_ds_bpermute(src4, (int)i);
    const int ih = __builtin_amdgcn_ds_bpermute(src4, (int)(i >> 32));
    return {__longlong_as_double(((long long)rh << 32) | (unsigned)rl),
            __longlong_as_double(((long long)ih << 32) | (unsigned)il)};
}

template <int C, int D = (C <= 10 ? 3 : 2)>
__global__ __launch_bounds__(kWave, (C <= 11 ? 2 : 1)) void bragg_march_wave_kernel(
    BraggK k, const double *__restrict__ s_dev, const cd *__restrict__ E,
    const cd *__restrict__ inc, const cd *__restrict__ vin, size_t vin_stride,
    cd *__restrict__ out, size_t out_stride, double out_scale,
    const int *__restrict__ active) {
    extern __shared__ cd coef[];                 // c_k at 0, b_k at nst
    const int s = blockIdx.x / k.nyc, j = blockIdx.x % k.nyc;
    if (active && !active[s]) return;
    E += table_offset(k, s);
    const int Nx = k.Nx, lane = threadIdx.x, Lu = Nx / C;
    const int nst = k.Nz - 1;
    const double alpha = -2.0 * kPi * s_dev[s];
    const double lr = 2.0 * k.s0r * k.ds, li = (2.0 * k.s0i + alpha) * k.ds;
    for (int t = lane; t < nst; t += kWave) {
        const cd f = cexp(-(double)t * lr, -(double)t * li);   // (e0 eg)^-t
        const cd g = cexp((double)t * lr, (double)t * li);     // (e0 eg)^t
        coef[t] = cmul(cd{k.ds * k.sbr, k.ds * k.sbi}, f);
        coef[nst + t] = cmul(cd{-k.ds * k.shr, -k.ds * k.shi}, g);
    }
    const bool real = lane < Lu;
    const int x0 = (real ? lane : Lu - 1) * C;   // idle lanes read in bounds
    const int srcX = (real ? (lane == 0 ? Lu - 1 : lane - 1) : lane) << 2;
    const int srcY = (real ? (lane == Lu - 1 ? 0 : lane + 1) : lane) << 2;
    cd X[C], Y[C];
    const cd *vj = vin ? vin + (size_t)s * vin_stride + (size_t)j * Nx : nullptr;
#pragma unroll
    for (int i = 0; i < C; ++i) {
        X[i] = inc ? inc[x0 + i] : cd{0.0, 0.0};
        Y[i] = vj ? vj[x0 + i] : cd{0.0, 0.0};
    }
    const size_t plane = (size_t)k.tpl * Nx;
    const cd *Ej = E + (size_t)j * Nx + x0;
    // e^{iH} slices run D steps ahead in D register sets with fixed roles
    // (the loop is unrolled by D), so a slice's load has D - 1 steps of
    // work to land in and no register copy waits for it.
    cd eb[D][C];
    auto load = [&](cd (&e)[C], int t) {
        const cd *Et = Ej + (size_t)(t < nst ? t : nst - 1) * plane;
#pragma unroll
        for (int i = 0; i < C; ++i) e[i] = Et[i];
    };
    auto advance = [&](const cd (&ec)[C], int step) {
        const cd ck = coef[step], bk = coef[nst + step];
        cd xn[C], yn[C];
#pragma unroll
        for (int i = 0; i < C; ++i) {
            const cd e = ec[i];
            const cd u = cmul(ck, e);              // c_k e^{+iH}
            const cd w = cmul(bk, cd{e.r, -e.i});  // b_k e^{-iH}
            xn[i] = cd{fma(u.r, Y[i].r, fma(-u.i, Y[i].i, X[i].r)),
                       fma(u.r, Y[i].i, fma(u.i, Y[i].r, X[i].i))};
            yn[i] = cd{fma(w.r, X[i].r, fma(-w.i, X[i].i, Y[i].r)),
                       fma(w.r, X[i].i, fma(w.i, X[i].r, Y[i].i))};
        }
        X[0] = wave_fetch(xn[C - 1], srcX);
#pragma unroll
        for (int i = 1; i < C; ++i) X[i] = xn[i - 1];
        Y[C - 1] = wave_fetch(yn[0], srcY);
#pragma unroll
        for (int i = 0; i < C - 1; ++i) Y[i] = yn[i + 1];
    };
#pragma unroll
    for (int d = 0; d < D; ++d) load(eb[d], d);
    __syncthreads();
    int step = 0;
    for (; step + D <= nst; step += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            advance(eb[d], step + d);
            load(eb[d], step + d + D);
        }
    }
#pragma unroll
    for (int d = 0; d < D; ++d)
        if (step + d < nst) advance(eb[d], step + d);
    const double egr = k.s0r * k.ds, egi = (k.s0i + alpha) * k.ds;
    const cd fin = cscale(cexp(-(double)nst * egr, -(double)nst * egi),
                          out_scale);
    if (real) {
        cd *o = out + (size_t)s * out_stride + (size_t)j * Nx + x0;
#pragma unroll
        for (int i = 0; i < C; ++i) o[i] = cmul(Y[i], fin);
    }
}

// The same march with each row split over the two wavefronts of a
// workgroup, for launches with too few rows to load every SIMD evenly
// (one wavefront per row leaves some SIMDs with two rows and others with
// one).  Wavefront w owns columns [w Nx/2, (w+1) Nx/2) in Lo lanes and
// marches Hc = Hl C further columns on each side in Hl otherwise idle
// lanes.  Column x after t steps depends only on columns x - t .. x + t,
// so the halos absorb the wrong values that enter at the wavefront's
// outer lanes for up to Hc steps; every P <= Hc steps the halos are
// refreshed from the other wavefront's owned edge columns through LDS
// (one barrier, buffers alternating between refreshes).  The owned
// columns see exactly the arithmetic of the one-wavefront kernel.
template <int C, int D = 3>
__global__ __launch_bounds__(2 * kWave, (C <= 6 ? 3 : 2)) void bragg_march_halo_kernel(
    BraggK k, const double *__restrict__ s_dev, const cd *__restrict__ E,
    const cd *__restrict__ inc, const cd *__restrict__ vin, size_t vin_stride,
    cd *__restrict__ out, size_t out_stride, double out_scale,
    const int *__restrict__ active) {
    extern __shared__ cd coef[];   // c_k, b_k, then [2][wave][edge][X,Y][Hc]
    const int s = blockIdx.x / k.nyc, j = blockIdx.x % k.nyc;
    if (active && !active[s]) return;
    E += table_offset(k, s);
    const int Nx = k.Nx, half = Nx / 2, lane = threadIdx.x % kWave;
    const int wv = threadIdx.x / kWave;
    const int Lo = half / C, Hl = (kWave - Lo) / 2, Hc = Hl * C;
    const int P = Hc / D * D;                 // steps between refreshes
    const int nst = k.Nz - 1;
    const double alpha = -2.0 * kPi * s_dev[s];
    const double lr = 2.0 * k.s0r * k.ds, li = (2.0 * k.s0i + alpha) * k.ds;
    for (int t = threadIdx.x; t < nst; t += 2 * kWave) {
        const cd f = cexp(-(double)t * lr, -(double)t * li);   // (e0 eg)^-t
        const cd g = cexp((double)t * lr, (double)t * li);     // (e0 eg)^t
        coef[t] = cmul(cd{k.ds * k.sbr, k.ds * k.sbi}, f);
        coef[nst + t] = cmul(cd{-k.ds * k.shr, -k.ds * k.shi}, g);
    }
    cd *edge = coef + 2 * nst;
    const int rel = lane - Hl;                // owned lane index if in [0, Lo)
    const int x0 = ((wv * half + rel * C) % Nx + Nx) % Nx;
    const int srcX = (lane == 0 ? kWave - 1 : lane - 1) << 2;
    const int srcY = (lane == kWave - 1 ? 0 : lane + 1) << 2;
    cd X[C], Y[C];
    const cd *vj = vin ? vin + (size_t)s * vin_stride + (size_t)j * Nx : nullptr;
#pragma unroll
    for (int i = 0; i < C; ++i) {
        X[i] = inc ? inc[x0 + i] : cd{0.0, 0.0};
        Y[i] = vj ? vj[x0 + i] : cd{0.0, 0.0};
    }
    const size_t plane = (size_t)k.tpl * Nx;
    const cd *Ej = E + (size_t)j * Nx + x0;
    cd eb[D][C];
    auto load = [&](cd (&e)[C], int t) {
        const cd *Et = Ej + (size_t)(t < nst ? t : nst - 1) * plane;
#pragma unroll
        for (int i = 0; i < C; ++i) e[i] = Et[i];
    };
    auto advance = [&](const cd (&ec)[C], int step) {
        const cd ck = coef[step], bk = coef[nst + step];
        cd xn[C], yn[C];
#pragma unroll
        for (int i = 0; i < C; ++i) {
            const cd e = ec[i];
            const cd u = cmul(ck, e);              // c_k e^{+iH}
            const cd w = cmul(bk, cd{e.r, -e.i});  // b_k e^{-iH}
            xn[i] = cd{fma(u.r, Y[i].r, fma(-u.i, Y[i].i, X[i].r)),
                       fma(u.r, Y[i].i, fma(u.i, Y[i].r, X[i].i))};
            yn[i] = cd{fma(w.r, X[i].r, fma(-w.i, X[i].i, Y[i].r)),
                       fma(w.r, X[i].i, fma(w.i, X[i].r, Y[i].i))};
        }
        X[0] = wave_fetch(xn[C - 1], srcX);
#pragma unroll
        for (int i = 1; i < C; ++i) X[i] = xn[i - 1];
        Y[C - 1] = wave_fetch(yn[0], srcY);
#pragma unroll
        for (int i = 0; i < C - 1; ++i) Y[i] = yn[i + 1];
    };
    // edge[((par * 2 + wave) * 2 + side) * 2 + field][Hc]
    auto slot = [&](int par, int w, int side, int field) {
        return edge + (size_t)(((par * 2 + w) * 2 + side) * 2 + field) * Hc;
    };
    int par = 0;
    auto refresh = [&]() {
        if (rel >= 0 && rel < Hl) {           // first Hc owned columns
#pragma unroll
            for (int i = 0; i < C; ++i)
                slot(par, wv, 0, 0)[rel * C + i] = X[i],
                slot(par, wv, 0, 1)[rel * C + i] = Y[i];
        }
        if (rel >= Lo - Hl && rel < Lo) {     // last Hc owned columns
            const int r = rel - (Lo - Hl);
#pragma unroll
            for (int i = 0; i < C; ++i)
                slot(par, wv, 1, 0)[r * C + i] = X[i],
                slot(par, wv, 1, 1)[r * C + i] = Y[i];
        }
        __syncthreads();
        if (lane < Hl) {                      // left halo: other's last
#pragma unroll
            for (int i = 0; i < C; ++i)
                X[i] = slot(par, 1 - wv, 1, 0)[lane * C + i],
                Y[i] = slot(par, 1 - wv, 1, 1)[lane * C + i];
        } else if (rel >= Lo && rel < Lo + Hl) {   // right halo: other's first
            const int r = rel - Lo;
#pragma unroll
            for (int i = 0; i < C; ++i)
                X[i] = slot(par, 1 - wv, 0, 0)[r * C + i],
                Y[i] = slot(par, 1 - wv, 0, 1)[r * C + i];
        }
        par ^= 1;
    };
#pragma unroll
    for (int d = 0; d < D; ++d) load(eb[d], d);
    __syncthreads();
    int step = 0, since = 0;
    for (; step + D <= nst; step += D) {
        if (since == P) refresh(), since = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            advance(eb[d], step + d);
            load(eb[d], step + d + D);
        }
        since += D;
    }
    if (step < nst && since == P) refresh();
#pragma unroll
    for (int d = 0; d < D; ++d)
        if (step + d < nst) advance(eb[d], step + d);
    const double egr = k.s0r * k.ds, egi = (k.s0i + alpha) * k.ds;
    const cd fin = cscale(cexp(-(double)nst * egr, -(double)nst * egi),
                          out_scale);
    if (rel >= 0 && rel < Lo) {
        cd *o = out + (size_t)s * out_stride + (size_t)j * Nx + x0;
#pragma unroll
        for (int i = 0; i < C; ++i) o[i] = cmul(Y[i], fin);
    }
}

using MarchFn = void (*)(BraggK, const double *, const cd *, const cd *,
                         const cd *, size_t, cd *, size_t, double,
                         const int *);

constexpr int kWaveMaxC = 16;

template <int... Cs>
static MarchFn wave_fn(int C) {
    MarchFn f = nullptr;
    ((C == Cs ? (f = bragg_march_wave_kernel<Cs>, 0) : 0), ...);
    return f;
}

// Columns per lane of the wavefront march for Nx: the smallest C with
// C | Nx and Nx / C <= 64, if one is at most kWaveMaxC; else 0.
static int wave_columns(int Nx) {
    for (int c = (Nx + kWave - 1) / kWave; c <= kWaveMaxC; ++c)
        if (c >= 1 && Nx % c == 0) return c;
    return 0;
}

template <int... Cs>
static MarchFn halo_fn(int C) {
    MarchFn f = nullptr;
    ((C == Cs ? (f = bragg_march_halo_kernel<Cs>, 0) : 0), ...);
    return f;
}

// Columns per lane of the two-wavefront march: the smallest C <= 16 with
// C | Nx/2 that leaves at least one spare lane on each side of the Nx/2/C
// owned lanes, a halo Hc = C (64 - Nx/2/C)/2 of at least 3 columns (one
// unrolled group of steps between refreshes) and no wider than the owned
// part; 0 if none (or Nx is odd).
static int halo_columns(int Nx) {
    if (Nx % 2) return 0;
    const int half = Nx / 2;
    for (int c = 1; c <= kWaveMaxC; ++c) {
        if (half % c) continue;
        const int Lo = half / c, Hl = (kWave - Lo) / 2;
        if (Hl >= 1 && Hl * c >= 3 && Hl <= Lo) return c;
    }
    return 0;
}

// Below this many rows (system-planes) a march launch splits each row over
// two wavefronts: one wavefront per row would load the SIMDs unevenly.
constexpr long long kHaloRows = 4096;

template <int... Ms>
static MarchFn march_fn(int M) {
    MarchFn f = nullptr;
    ((M == Ms ? (f = bragg_march_kernel<Ms>, 0) : 0), ...);
    return f;
}

static int launch_march(const BraggK &k, const double *s_dev, const cd *E,
                        const cd *inc, const cd *vin, size_t vin_stride,
                        cd *out, size_t out_stride, double out_scale,
                        const int *active, hipStream_t stream) {
    const int C = wave_columns(k.Nx);
    const size_t coef = (size_t)2 * std::max(k.Nz - 1, 1) * sizeof(cd);
    const int Ch = halo_columns(k.Nx);
    if (Ch && (long long)k.nsys * k.nyc < kHaloRows) {
        const int Hc = (kWave - k.Nx / 2 / Ch) / 2 * Ch;
        const size_t lds = coef + (size_t)16 * Hc * sizeof(cd);
        if (lds <= 64 * 1024) {
            MarchFn hf = halo_fn<1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14,
                                 15, 16>(Ch);
            hipLaunchKernelGGL(hf, dim3(k.nsys * k.nyc), dim3(2 * kWave), lds,
                               stream, k, s_dev, E, inc, vin, vin_stride, out,
                               out_stride, out_scale, active);
            DYF_LAUNCHED("bragg_march_halo_kernel");
            return DYF_OK;
        }
    }
    if (C && coef <= 64 * 1024) {
        MarchFn wf = wave_fn<1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15,
                             16>(C);
        hipLaunchKernelGGL(wf, dim3(k.nsys * k.nyc), dim3(kWave), coef,
                           stream, k, s_dev, E, inc, vin, vin_stride, out,
                           out_stride, out_scale, active);
        DYF_LAUNCHED("bragg_march_wave_kernel");
        return DYF_OK;
    }
    const size_t lds = march_lds_bytes(k);
    const int M = (k.Nx + kMarchThreads - 1) / kMarchThreads;
    MarchFn fn = march_fn<1, 2, 3, 4, 5, 6, 7, 8>(M);
    if (!fn) return fail(DYF_EINVAL, "Nx=%d exceeds %d", k.Nx,
                         kMarchThreads * kMaxPerThread);
    static std::once_flag once;
    std::call_once(once, [] {
        for (int m = 1; m <= kMaxPerThread; ++m)
            (void)hipFuncSetAttribute(
                (const void *)march_fn<1, 2, 3, 4, 5, 6, 7, 8>(m),
                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipGetLastError();
    });
    hipLaunchKernelGGL(fn, dim3(k.nsys * k.nyc), dim3(kMarchThreads), lds,
                       stream, k, s_dev, E, inc, vin, vin_stride, out,
                       out_stride, out_scale, active);
    DYF_LAUNCHED("bragg_march_kernel");
    return DYF_OK;
}

// ---- perfect-crystal symbol and the preconditioner scale --------------
// One block per system: T22(m) of the unstrained chain over Nz-1 steps
// (bragg/solver.py:38-60), floored at 1e-8 max|T22| by a real value
// (163-165), stored as 1 / (T22s Nx) so the inverse FFT needs no extra pass.
__global__ __launch_bounds__(256) void bragg_symbol_kernel(
    BraggK k, const double *__restrict__ s_dev, cd *__restrict__ pinv) {
    __shared__ double red[256];
    const int s = blockIdx.x, Nx = k.Nx;
    const double alpha = -2.0 * kPi * s_dev[s];
    const cd e0 = cexp(k.s0r * k.ds, k.s0i * k.ds);
    const cd eg = cexp(k.s0r * k.ds, (k.s0i + alpha) * k.ds);
    const cd sh = {k.shr, k.shi}, sb = {k.sbr, k.sbi};
    cd *T = pinv + (size_t)s * Nx;
    double mx = 0.0;
    for (int idx = threadIdx.x; idx < Nx; idx += 256) {
        const int m = idx <= (Nx - 1) / 2 ? idx : idx - Nx;
        double sn, cs;
        sincos(-2.0 * kPi * (double)m / (double)Nx, &sn, &cs);
        const cd p = {cs, sn}, pc = {cs, -sn};
        const cd a11 = cmul(e0, p);
        const cd a12 = cmul(cscale(cmul(e0, p), k.ds), sb);
        const cd a21 = cdiv(cmul(cscale(cd{-pc.r, -pc.i}, k.ds), sh), eg);
        const cd a22 = cdiv(pc, eg);
        // T = A^(Nz-1) by repeated squaring (the step matrix A is the
        // same at every depth): about 2 log2(Nz) 2x2 products instead of
        // Nz - 1 in a dependent chain; the rounding differs from the
        // step-by-step product in the last bits only
        cd T11 = {1, 0}, T12 = {0, 0}, T21 = {0, 0}, T22 = {1, 0};
        cd p11 = a11, p12 = a12, p21 = a21, p22 = a22;
        for (int e = k.Nz - 1; e > 0; e >>= 1) {
            if (e & 1) {
                const cd n11 = cadd(cmul(p11, T11), cmul(p12, T21));
                const cd n12 = cadd(cmul(p11, T12), cmul(p12, T22));
                const cd n21 = cadd(cmul(p21, T11), cmul(p22, T21));
                const cd n22 = cadd(cmul(p21, T12), cmul(p22, T22));
                T11 = n11, T12 = n12, T21 = n21, T22 = n22;
            }
            if (e > 1) {
                const cd q11 = cadd(cmul(p11, p11), cmul(p12, p21));
                const cd q12 = cadd(cmul(p11, p12), cmul(p12, p22));
                const cd q21 = cadd(cmul(p21, p11), cmul(p22, p21));
                const cd q22 = cadd(cmul(p21, p12), cmul(p22, p22));
                p11 = q11, p12 = q12, p21 = q21, p22 = q22;
            }
        }
        T[idx] = T22;
        mx = fmax(mx, sqrt(T22.r * T22.r + T22.i * T22.i));
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    const double floor_v = 1e-8 * red[0];
    for (int idx = threadIdx.x; idx < Nx; idx += 256) {
        cd t = T[idx];
        if (sqrt(t.r * t.r + t.i * t.i) < floor_v) t = cd{floor_v, 0.0};
        T[idx] = cdiv(cd{1.0 / (double)Nx, 0.0}, t);
    }
}

// w = M v: the fused transform - scale - inverse transform of fft.hip on
// every row (nsys systems of nyc planes), with the symbol written by
// bragg_symbol_kernel (and, for convolution plans, its kernel spectrum).
static int apply_precon(const BraggK &k, const cd *symbol, cd *v, cd *w,
                        hipStream_t stream) {
    FftPlan p;
    int rc = fft_plan_for(k.Nx, &p);
    if (rc) return rc;
    return fft_precondition(p, k.Nx, symbol, k.nsys, k.nyc, v, w, stream);
}

// The symbol of every system: 1 / (T22s Nx) per mode, then the
// convolution plan's kernel spectrum where the plan has one.
static int make_symbol(const BraggK &k, const double *s_dev, cd *symbol,
                       hipStream_t stream) {
    FftPlan p;
    int rc = fft_plan_for(k.Nx, &p);
    if (rc) return rc;
    bragg_symbol_kernel<<<k.nsys, 256, 0, stream>>>(k, s_dev, symbol);
    DYF_LAUNCHED("bragg_symbol_kernel");
    return fft_prepare_spectrum(p, k.Nx, symbol, k.nsys, stream);
}

// Complex elements per system of the symbol (an upper bound when Nx has no
// plan: the solve then fails before using the workspace).
static size_t symbol_elems(int Nx) {
    FftPlan p;
    if (fft_plan_for(Nx, &p) != DYF_OK) return (size_t)Nx + 4096;
    return fft_symbol_elems(p, Nx);
}

// ---- batched reductions ----------------------------------------------
// part[(s*kc + q)*P + p] = sum over chunk p of conj(A_sq[i]) * B_s[i]
constexpr int kRedThreads = 1024;

__global__ __launch_bounds__(kRedThreads) void batched_dot_kernel(
    const cd *__restrict__ A, size_t a_sys, size_t a_vec, int kc,
    const cd *__restrict__ Bv, size_t b_sys, size_t N, int P,
    cd *__restrict__ part, const int *__restrict__ active,
    cd *__restrict__ out, int ostride) {
    __shared__ double rr[kRedThreads], ri[kRedThreads];
    const int sq = blockIdx.x, p = blockIdx.y;
    const int s = sq / kc, q = sq % kc;
    if (active && !active[s]) return;
    const cd *a = A + s * a_sys + q * a_vec;
    const cd *b = Bv + s * b_sys;
    const size_t chunk = (N + P - 1) / P;
    const size_t lo = (size_t)p * chunk, hi = std::min(N, lo + chunk);
    double sr = 0.0, si = 0.0;
#pragma unroll 4
    for (size_t i = lo + threadIdx.x; i < hi; i += kRedThreads) {
        const cd x = a[i], y = b[i];
        sr += x.r * y.r + x.i * y.i;
        si += x.r * y.i - x.i * y.r;
    }
    rr[threadIdx.x] = sr;
    ri[threadIdx.x] = si;
    __syncthreads();
    for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            rr[threadIdx.x] += rr[threadIdx.x + w];
            ri[threadIdx.x] += ri[threadIdx.x + w];
        }
        __syncthreads();
    }
    // one part (N <= 32768): the block's sum is the dot product itself
    if (threadIdx.x == 0) {
        if (P == 1)
            out[(size_t)s * ostride + q] = cd{rr[0], ri[0]};
        else
            part[(size_t)sq * P + p] = cd{rr[0], ri[0]};
    }
}

// out[s*ostride + q] = sum_p part[...] (fixed order: deterministic)
__global__ void sum_parts_kernel(const cd *__restrict__ part, int P, int n,
                                 cd *__restrict__ out, int ostride,
                                 const int *active, int kc) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const int s = t / kc, q = t % kc;
    if (active && !active[s]) return;
    double r = 0.0, i = 0.0;
    for (int p = 0; p < P; ++p)
        r += part[(size_t)t * P + p].r, i += part[(size_t)t * P + p].i;
    out[(size_t)s * ostride + q] = cd{r, i};
}

struct SysState;
__device__ int sys_cols(const SysState *st, int s);

// y_s[i] += sign * sum_q c_sq * V_sq[i], q < kc (or the system's own
// column count col+1 when `st` is given)
__global__ __launch_bounds__(256) void combine_kernel(
    const cd *__restrict__ V, size_t v_sys, size_t v_vec, int kc,
    const SysState *__restrict__ st, const cd *__restrict__ coef,
    size_t c_sys, double sign, cd *__restrict__ y, size_t y_sys, size_t N,
    const int *__restrict__ active) {
    const int s = blockIdx.y;
    if (active && !active[s]) return;
    if (st) kc = sys_cols(st, s);
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    const cd *c = coef + s * c_sys;
    const cd *v = V + s * v_sys + i;
    double ar = 0.0, ai = 0.0;
    for (int q = 0; q < kc; ++q) {
        const cd a = c[q], b = v[(size_t)q * v_vec];
        ar += a.r * b.r - a.i * b.i;
        ai += a.r * b.i + a.i * b.r;
    }
    cd &o = y[s * y_sys + i];
    o.r += sign * ar;
    o.i += sign * ai;
}

// dst_s = src_s * scale_s  (scale given as 1/norm; skipped if zero)
__global__ __launch_bounds__(256) void scale_copy_kernel(
    const cd *__restrict__ src, size_t s_sys, cd *__restrict__ dst,
    size_t d_sys, const double *__restrict__ inv, size_t N,
    const int *__restrict__ active) {
    const int s = blockIdx.y;
    if (active && !active[s]) return;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    dst[s * d_sys + i] = cscale(src[s * s_sys + i], inv[s]);
}

// r_s = b_s - a_s
__global__ __launch_bounds__(256) void sub_kernel(const cd *__restrict__ b,
                                                  const cd *__restrict__ a,
                                                  cd *__restrict__ r, size_t N,
                                                  const int *__restrict__ active) {
    const int s = blockIdx.y;
    if (active && !active[s]) return;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    const size_t at = (size_t)s * N + i;
    r[at] = cd{b[at].r - a[at].r, b[at].i - a[at].i};
}

// ---- per-system small dense work (one thread per system) --------------
struct SysState {
    double bnrm2, atol, ptol, pmf, presid, rnorm, h0;
    int col, breakdown, inner_done, done, iters, code, pad_;
};

__device__ int sys_cols(const SysState *st, int s) { return st[s].col + 1; }

// The words the host steers the loop by: per column "some system still
// iterates", then per cycle "some system continues" (R + 1 words).  The
// host reads `host` (page-locked host memory); `gate` is a device copy
// that lets only the first system to raise a word write it across the
// bus (with 1312 one-plane systems, one write per column instead of
// 1312).  Both are cleared at the start of a cycle.
struct Flags {
    int *gate, *host;
};

__device__ __forceinline__ void raise_flag(Flags f, int i) {
    if (atomicExch(f.gate + i, 1) == 0) *(volatile int *)(f.host + i) = 1;
}

// LAPACK zlartg convention: [c s; -conj(s) c] [f; g] = [r; 0], c real.
__device__ void lartg(cd f, cd g, double *c, cd *s, cd *r) {
    const double fa = hypot(f.r, f.i), ga = hypot(g.r, g.i);
    if (ga == 0.0) {
        *c = 1.0, *s = cd{0, 0}, *r = f;
    } else if (fa == 0.0) {
        *c = 0.0, *s = cd{g.r / ga, -g.i / ga}, *r = cd{ga, 0.0};
    } else {
        const double d = hypot(fa, ga);
        const cd ph = {f.r / fa, f.i / fa};
        *c = fa / d;
        *s = cd{(ph.r * g.r + ph.i * g.i) / d, (ph.i * g.r - ph.r * g.i) / d};
        *r = cd{ph.r * d, ph.i * d};
    }
}

// After CGS2 of column `col`: h[col][0..col] = h1 + h2, h[col][col+1] =
// ||w||, breakdown test, past and new Givens rotations, S update, presid.
// inv[s] receives 1/||w|| (0 on breakdown) for the basis store.
// Column `col` of every active system: Hessenberg column from the two
// CGS2 passes, breakdown test against h0 = ||w|| before them (nrm0),
// the stored Givens rotations and a new one, the residual estimate, and
// SciPy's inner-loop exit (presid <= ptol, breakdown or the last column),
// which clears the system's active flag; `still` is set while any system
// continues, so the host reads one int per column.  nw = ||w||^2 after
// CGS2, n0 = ||w||^2 before it.  Returns whether system s continues.
// h1c, h2c: this column's projections from the two CGS2 passes (r <=
// col); gp: the stored rotations (2 col entries); hc: a working column of
// col + 2 entries, H's own column or a faster copy written to it at the
// end; Sc = S[col].  (The fused step passes LDS copies, so the serial
// rotation loop does not wait on global memory at every entry.)
__device__ int arnoldi_finish_sys(int s, int restart, int col, cd *H, cd *G,
                                  cd *S, const cd *h1c, const cd *h2c,
                                  const cd *gp, cd *hc, cd Sc, double nw,
                                  double n0, SysState *st, double *inv,
                                  int *active, double eps, Flags fl) {
    SysState &q = st[s];
    q.col = col;
    q.h0 = sqrt(n0);
    cd *h = H + (size_t)s * restart * (restart + 1) + (size_t)col * (restart + 1);
    cd *g = G + (size_t)s * restart * 2;
    cd *Sv = S + (size_t)s * (restart + 1);
    for (int r = 0; r <= col; ++r) {
        const cd a = h1c[r];
        const cd b = h2c[r];
        hc[r] = cd{a.r + b.r, a.i + b.i};
    }
    const double h1 = sqrt(nw);
    hc[col + 1] = cd{h1, 0.0};
    int breakdown = 0;
    if (h1 <= eps * q.h0) {
        hc[col + 1] = cd{0.0, 0.0};
        breakdown = 1;
        inv[s] = 0.0;
    } else {
        inv[s] = 1.0 / h1;
    }
    for (int r = 0; r < col; ++r) {
        const double c = gp[2 * r].r;
        const cd sn = gp[2 * r + 1];
        const cd n0 = hc[r], n1 = hc[r + 1];
        hc[r] = cd{c * n0.r + (sn.r * n1.r - sn.i * n1.i),
                   c * n0.i + (sn.r * n1.i + sn.i * n1.r)};
        // -conj(s) n0 + c n1
        hc[r + 1] = cd{-(sn.r * n0.r + sn.i * n0.i) + c * n1.r,
                       -(sn.r * n0.i - sn.i * n0.r) + c * n1.i};
    }
    double c;
    cd sn, mag;
    lartg(hc[col], hc[col + 1], &c, &sn, &mag);
    g[2 * col] = cd{c, 0.0};
    g[2 * col + 1] = sn;
    hc[col] = mag;
    hc[col + 1] = cd{0.0, 0.0};
    if (hc != h)
        for (int r = 0; r <= col + 1; ++r) h[r] = hc[r];
    const cd tmp = {-(sn.r * Sc.r + sn.i * Sc.i), -(sn.r * Sc.i - sn.i * Sc.r)};
    Sv[col] = cd{c * Sc.r, c * Sc.i};
    Sv[col + 1] = tmp;
    q.presid = hypot(tmp.r, tmp.i);
    q.breakdown = breakdown;
    q.iters += 1;
    if (q.presid <= q.ptol || breakdown || col == restart - 1) {
        q.inner_done = 1;
        active[s] = 0;
        return 0;
    }
    raise_flag(fl, col);
    return 1;
}

// One Arnoldi step after w = M A v_col for systems of N <= kCgsMaxN
// unknowns, one workgroup per system with w in registers: CGS2 (||w||^2,
// then twice h = V^H w over the kc = col + 1 basis vectors and w -= V h,
// then ||w||^2), the Hessenberg/Givens update of arnoldi_finish_sys on
// one lane, and v_{col+1} = w / ||w|| for a system that continues (what
// the dot, combine, finish and scale kernels do in nine launches).
// Reductions: a butterfly within each wavefront, then the wavefront sums
// in a fixed order.  The workgroup has NW = min(16, ceil(N / 64))
// wavefronts, so most lanes hold one element.
constexpr int kCgsMaxThreads = 1024;
constexpr int kCgsMaxN = 2048;

// Cross-lane moves of a double through DPP (two 32-bit halves), and one
// lane's double read into a uniform value.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi =
        __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double lane_f64(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum over the 64 lanes of a fully active wavefront, in every lane: DPP
// butterflies within each row of 16 lanes (swap pairs, swap pair-pairs,
// rotate by 4, rotate by 8), then the four row sums in a fixed order.
// (ds_bpermute butterflies cost an LDS round trip per stage.)
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_f64<0xB1>(v);    // quad_perm [1, 0, 3, 2]
    v += dpp_f64<0x4E>(v);    // quad_perm [2, 3, 0, 1]
    v += dpp_f64<0x124>(v);   // row_ror:4
    v += dpp_f64<0x128>(v);   // row_ror:8
    return (lane_f64(v, 0) + lane_f64(v, 16)) + (lane_f64(v, 32) + lane_f64(v, 48));
}

// P = elements of w per lane (N <= P x block size).  Every lane loads at
// a clamped in-range index, so the loads of a basis vector carry no
// branch and two vectors' loads are in flight together; lanes past N hold
// w = 0 and keep it.
template <int P>
__global__ __launch_bounds__(kCgsMaxThreads) void arnoldi_step_kernel(
    cd *__restrict__ V, size_t v_sys, size_t N, int restart, int col,
    const cd *__restrict__ w, cd *__restrict__ h1, cd *__restrict__ h2,
    cd *H, cd *G, cd *S, SysState *st, double *inv, int *active, double eps,
    Flags fl) {
    // LDS (doubles): [NW waves][2 kc] wave sums, h [2 kc], then the first
    // pass's h and the working column [2 (kc + 1)], then the stored
    // rotations [4 col]
    extern __shared__ double red[];
    const int s = blockIdx.x, tid = threadIdx.x, T = blockDim.x;
    if (!active[s]) return;
    const int wv = tid >> 6, lane = tid & 63, NW = T >> 6;
    const int kc = col + 1, hstride = restart + 1;
    const int hb = 2 * NW * kc;               // h after the wave sums
    cd *h1l = (cd *)(red + hb + 2 * kc);
    cd *gl = h1l + (kc + 1);
    for (int t = tid; t < 2 * col; t += T)
        gl[t] = G[(size_t)s * restart * 2 + t];
    const cd Sc0 = tid == 0 ? S[(size_t)s * hstride + col] : cd{0.0, 0.0};
    const cd *ws = w + (size_t)s * N;
    cd *Vs = V + (size_t)s * v_sys;
    cd x[P];
    size_t idx[P];
    bool ok[P];
    double nn = 0.0;
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const size_t i = tid + (size_t)m * T;
        ok[m] = i < N;
        idx[m] = ok[m] ? i : N - 1;
        const cd v = ws[idx[m]];
        x[m] = ok[m] ? v : cd{0.0, 0.0};
        nn += x[m].r * x[m].r + x[m].i * x[m].i;
    }
    // block sum of one value through red[0..NW)
    auto block_sum = [&](double v) {
        v = wave_sum(v);
        __syncthreads();
        if (lane == 0) red[wv] = v;
        __syncthreads();
        double t = red[0];
        for (int u = 1; u < NW; ++u) t += red[u];
        return t;
    };
    const double n0 = block_sum(nn);
    for (int pass = 0; pass < 2; ++pass) {
        cd *hp = (pass == 0 ? h1 : h2) + (size_t)s * hstride;
        __syncthreads();
        for (int q = 0; q < kc; q += 2) {
            const bool two = q + 1 < kc;
            const cd *v0 = Vs + (size_t)q * N;
            const cd *v1 = Vs + (size_t)(two ? q + 1 : q) * N;
            cd a[P], b[P];
#pragma unroll
            for (int m = 0; m < P; ++m) a[m] = v0[idx[m]], b[m] = v1[idx[m]];
            double sr = 0.0, si = 0.0, tr = 0.0, ti = 0.0;
#pragma unroll
            for (int m = 0; m < P; ++m) {
                sr += a[m].r * x[m].r + a[m].i * x[m].i;
                si += a[m].r * x[m].i - a[m].i * x[m].r;
                tr += b[m].r * x[m].r + b[m].i * x[m].i;
                ti += b[m].r * x[m].i - b[m].i * x[m].r;
            }
            sr = wave_sum(sr), si = wave_sum(si);
            tr = wave_sum(tr), ti = wave_sum(ti);
            if (lane == 0) {
                red[wv * 2 * kc + 2 * q] = sr, red[wv * 2 * kc + 2 * q + 1] = si;
                if (two)
                    red[wv * 2 * kc + 2 * q + 2] = tr,
                    red[wv * 2 * kc + 2 * q + 3] = ti;
            }
        }
        __syncthreads();
        if (tid < kc) {
            double r = red[2 * tid], i = red[2 * tid + 1];
            for (int u = 1; u < NW; ++u)
                r += red[u * 2 * kc + 2 * tid], i += red[u * 2 * kc + 2 * tid + 1];
            hp[tid] = cd{r, i};
            red[hb + 2 * tid] = r, red[hb + 2 * tid + 1] = i;
            if (pass == 0) h1l[tid] = cd{r, i};
        }
        __syncthreads();
        double ar[P], ai[P];
#pragma unroll
        for (int m = 0; m < P; ++m) ar[m] = 0.0, ai[m] = 0.0;
        for (int q = 0; q < kc; q += 2) {
            const bool two = q + 1 < kc;
            const double c0r = red[hb + 2 * q], c0i = red[hb + 2 * q + 1];
            const double c1r = two ? red[hb + 2 * q + 2] : 0.0;
            const double c1i = two ? red[hb + 2 * q + 3] : 0.0;
            const cd *v0 = Vs + (size_t)q * N;
            const cd *v1 = Vs + (size_t)(two ? q + 1 : q) * N;
            cd a[P], b[P];
#pragma unroll
            for (int m = 0; m < P; ++m) a[m] = v0[idx[m]], b[m] = v1[idx[m]];
#pragma unroll
            for (int m = 0; m < P; ++m) {
                ar[m] += c0r * a[m].r - c0i * a[m].i;
                ai[m] += c0r * a[m].i + c0i * a[m].r;
                ar[m] += c1r * b[m].r - c1i * b[m].i;
                ai[m] += c1r * b[m].i + c1i * b[m].r;
            }
        }
#pragma unroll
        for (int m = 0; m < P; ++m) {
            x[m].r = ok[m] ? x[m].r - ar[m] : 0.0;
            x[m].i = ok[m] ? x[m].i - ai[m] : 0.0;
        }
    }
    nn = 0.0;
#pragma unroll
    for (int m = 0; m < P; ++m)
        nn += x[m].r * x[m].r + x[m].i * x[m].i;
    nn = block_sum(nn);
    // red[0..NW) may still be read by other lanes: broadcast through red[NW]
    if (tid == 0)
        red[NW] = arnoldi_finish_sys(s, restart, col, H, G, S, h1l,
                                     (const cd *)(red + hb), gl, h1l, Sc0,
                                     nn, n0, st, inv, active, eps, fl)
                      ? inv[s]
                      : 0.0;
    __syncthreads();
    const double sc = red[NW];
    if (sc == 0.0) return;
    cd *vn = Vs + (size_t)kc * N;
#pragma unroll
    for (int m = 0; m < P; ++m)
        if (ok[m]) vn[idx[m]] = cscale(x[m], sc);
}

using step_fn = void (*)(cd *, size_t, size_t, int, int, const cd *, cd *,
                         cd *, cd *, cd *, cd *, SysState *, double *, int *,
                         double, Flags);

// Workgroup size of the Arnoldi step for N unknowns (N <= kCgsMaxN):
// about kCgsPer elements per lane, at most 8
constexpr int kCgsPer = 4;
static int step_threads(size_t N) {
    const size_t waves = std::max<size_t>(
        (N + 64 * kCgsPer - 1) / (64 * kCgsPer), (N + 511) / 512);
    return (int)std::min<size_t>(kCgsMaxThreads, waves * 64);
}

static step_fn step_kernel(size_t N) {
    static const step_fn tab[8] = {
        arnoldi_step_kernel<1>, arnoldi_step_kernel<2>, arnoldi_step_kernel<3>,
        arnoldi_step_kernel<4>, arnoldi_step_kernel<5>, arnoldi_step_kernel<6>,
        arnoldi_step_kernel<7>, arnoldi_step_kernel<8>};
    const size_t T = step_threads(N);
    return tab[(N + T - 1) / T - 1];
}

// ---- Arnoldi step for large systems: CGS2 over slices -----------------
// Systems too large for one workgroup (the joint chunk: N = planes x Nx,
// 16,640 on a 32-plane, 520-column chunk) take one workgroup per (slice of
// kSplitLen elements, system), each holding its slice of w in registers.
// A pass that needs a whole-system sum writes one partial per slice; the
// next kernel sums that system's partials in slice order (deterministic,
// and every workgroup of the system gets the same value) before it uses
// them:
//   cgs_dots_kernel               h1 = V^H w and ||w||^2   -> partials A
//   cgs_update_dots_kernel        w -= V h1, h2 = V^H w    -> partials B
//   cgs_update_norm_kernel        w -= V h2, ||w||^2       -> partials C
//   arnoldi_finish_scale_kernel   the norms, v_{col+1} = w / ||w||, and in
//                                 slice 0 the Hessenberg/Givens update
// four launches where the batched dot, combine, finish and scale kernels
// took nine, and w crosses memory five times instead of ten.  The same CGS2 as arnoldi_step_kernel; only the order
// of the floating-point sums differs.
constexpr int kSplitThreads = 256;
constexpr int kSplitWaves = kSplitThreads / kWave;
constexpr int kSplitPer = 4;
constexpr int kSplitLen = kSplitThreads * kSplitPer;

static int split_slices(size_t N) {
    return (int)((N + kSplitLen - 1) / kSplitLen);
}

// This lane's elements of slice `b`: clamped in-range indices (loads carry
// no branch), and x = 0 past N.
__device__ __forceinline__ void split_load(const cd *__restrict__ src,
                                           size_t N, int b, int tid,
                                           size_t (&idx)[kSplitPer],
                                           bool (&ok)[kSplitPer],
                                           cd (&x)[kSplitPer]) {
#pragma unroll
    for (int m = 0; m < kSplitPer; ++m) {
        const size_t i = (size_t)b * kSplitLen + tid + (size_t)m * kSplitThreads;
        ok[m] = i < N;
        idx[m] = ok[m] ? i : N - 1;
        const cd v = src[idx[m]];
        x[m] = ok[m] ? v : cd{0.0, 0.0};
    }
}

// red[wv * rs + 2q], red[wv * rs + 2q + 1] = this wavefront's share of
// conj(V_q) . x for q < kc (x is zero past N); two vectors' loads in flight
__device__ __forceinline__ void split_dots(const cd *__restrict__ Vs, size_t N,
                                           int kc,
                                           const size_t (&idx)[kSplitPer],
                                           const cd (&x)[kSplitPer],
                                           double *red, int rs, int wv,
                                           int lane) {
    for (int q = 0; q < kc; q += 2) {
        const bool two = q + 1 < kc;
        const cd *v0 = Vs + (size_t)q * N;
        const cd *v1 = Vs + (size_t)(two ? q + 1 : q) * N;
        cd a[kSplitPer], b[kSplitPer];
#pragma unroll
        for (int m = 0; m < kSplitPer; ++m) a[m] = v0[idx[m]], b[m] = v1[idx[m]];
        double sr = 0.0, si = 0.0, tr = 0.0, ti = 0.0;
#pragma unroll
        for (int m = 0; m < kSplitPer; ++m) {
            sr += a[m].r * x[m].r + a[m].i * x[m].i;
            si += a[m].r * x[m].i - a[m].i * x[m].r;
            tr += b[m].r * x[m].r + b[m].i * x[m].i;
            ti += b[m].r * x[m].i - b[m].i * x[m].r;
        }
        sr = wave_sum(sr), si = wave_sum(si);
        tr = wave_sum(tr), ti = wave_sum(ti);
        if (lane == 0) {
            red[wv * rs + 2 * q] = sr, red[wv * rs + 2 * q + 1] = si;
            if (two) red[wv * rs + 2 * q + 2] = tr, red[wv * rs + 2 * q + 3] = ti;
        }
    }
}

// x -= sum_q coef_q V_q over q < kc (coef: 2 kc doubles, re/im)
__device__ __forceinline__ void split_update(const cd *__restrict__ Vs,
                                             size_t N, int kc,
                                             const double *coef,
                                             const size_t (&idx)[kSplitPer],
                                             const bool (&ok)[kSplitPer],
                                             cd (&x)[kSplitPer]) {
    double ar[kSplitPer], ai[kSplitPer];
#pragma unroll
    for (int m = 0; m < kSplitPer; ++m) ar[m] = 0.0, ai[m] = 0.0;
    for (int q = 0; q < kc; q += 2) {
        const bool two = q + 1 < kc;
        const double c0r = coef[2 * q], c0i = coef[2 * q + 1];
        const double c1r = two ? coef[2 * q + 2] : 0.0;
        const double c1i = two ? coef[2 * q + 3] : 0.0;
        const cd *v0 = Vs + (size_t)q * N;
        const cd *v1 = Vs + (size_t)(two ? q + 1 : q) * N;
        cd a[kSplitPer], b[kSplitPer];
#pragma unroll
        for (int m = 0; m < kSplitPer; ++m) a[m] = v0[idx[m]], b[m] = v1[idx[m]];
#pragma unroll
        for (int m = 0; m < kSplitPer; ++m) {
            ar[m] += c0r * a[m].r - c0i * a[m].i;
            ai[m] += c0r * a[m].i + c0i * a[m].r;
            ar[m] += c1r * b[m].r - c1i * b[m].i;
            ai[m] += c1r * b[m].i + c1i * b[m].r;
        }
    }
#pragma unroll
    for (int m = 0; m < kSplitPer; ++m) {
        x[m].r = ok[m] ? x[m].r - ar[m] : 0.0;
        x[m].i = ok[m] ? x[m].i - ai[m] : 0.0;
    }
}

// coef[2q], coef[2q+1] = sum over the B slices of part[q * B + b], q < kc,
// in slice order; slice 0 also stores the sums in h when given (this
// system's first-pass projections, read by the finish)
__device__ __forceinline__ void split_sum_parts(const cd *__restrict__ part,
                                                int B, int kc, double *coef,
                                                cd *h, int b, int tid) {
    for (int q = tid; q < kc; q += kSplitThreads) {
        const cd *p = part + (size_t)q * B;
        double r = 0.0, i = 0.0;
        for (int u = 0; u < B; ++u) r += p[u].r, i += p[u].i;
        coef[2 * q] = r, coef[2 * q + 1] = i;
        if (h && b == 0) h[q] = cd{r, i};
    }
}

// Partials A: pA[(s (R+2) + q) B + b] = slice b's conj(V_q) . w (q < kc),
// and at q = R + 1 its ||w||^2.
__global__ __launch_bounds__(kSplitThreads) void cgs_dots_kernel(
    const cd *__restrict__ V, size_t v_sys, size_t N, int R, int kc,
    const cd *__restrict__ w, cd *__restrict__ pA, int B,
    const int *__restrict__ active) {
    extern __shared__ double red[];            // [waves][2 (kc + 1)]
    const int b = blockIdx.x, s = blockIdx.y;
    if (!active[s]) return;
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int rs = 2 * (kc + 1);
    size_t idx[kSplitPer];
    bool ok[kSplitPer];
    cd x[kSplitPer];
    split_load(w + (size_t)s * N, N, b, tid, idx, ok, x);
    double nn = 0.0;
#pragma unroll
    for (int m = 0; m < kSplitPer; ++m) nn += x[m].r * x[m].r + x[m].i * x[m].i;
    nn = wave_sum(nn);
    if (lane == 0) red[wv * rs + 2 * kc] = nn, red[wv * rs + 2 * kc + 1] = 0.0;
    split_dots(V + (size_t)s * v_sys, N, kc, idx, x, red, rs, wv, lane);
    __syncthreads();
    for (int j = tid; j < rs; j += kSplitThreads) {
        double t = red[j];
        for (int u = 1; u < kSplitWaves; ++u) t += red[u * rs + j];
        const int q = j >> 1 < kc ? j >> 1 : R + 1;
        ((double *)&pA[((size_t)s * (R + 2) + q) * B + b])[j & 1] = t;
    }
}

// Sums partials A into h1 (its column stored in h1 by slice 0), w -= V h1
// on the slice, then partials B: pB[(s (R+1) + q) B + b] = conj(V_q) . w.
__global__ __launch_bounds__(kSplitThreads) void cgs_update_dots_kernel(
    const cd *__restrict__ V, size_t v_sys, size_t N, int R, int kc,
    cd *__restrict__ w, const cd *__restrict__ pA, cd *__restrict__ pB,
    cd *__restrict__ h1, int B, const int *__restrict__ active) {
    extern __shared__ double red[];            // [waves][2 kc], coef [2 kc]
    const int b = blockIdx.x, s = blockIdx.y;
    if (!active[s]) return;
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int rs = 2 * kc;
    double *coef = red + kSplitWaves * rs;
    split_sum_parts(pA + (size_t)s * (R + 2) * B, B, kc, coef,
                    h1 + (size_t)s * (R + 1), b, tid);
    size_t idx[kSplitPer];
    bool ok[kSplitPer];
    cd x[kSplitPer];
    cd *ws = w + (size_t)s * N;
    split_load(ws, N, b, tid, idx, ok, x);
    __syncthreads();
    const cd *Vs = V + (size_t)s * v_sys;
    split_update(Vs, N, kc, coef, idx, ok, x);
#pragma unroll
    for (int m = 0; m < kSplitPer; ++m)
        if (ok[m]) ws[idx[m]] = x[m];
    split_dots(Vs, N, kc, idx, x, red, rs, wv, lane);
    __syncthreads();
    for (int j = tid; j < rs; j += kSplitThreads) {
        double t = red[j];
        for (int u = 1; u < kSplitWaves; ++u) t += red[u * rs + j];
        ((double *)&pB[((size_t)s * (R + 1) + (j >> 1)) * B + b])[j & 1] = t;
    }
}

// Sums partials B into h2 (stored by slice 0 for the finish), w -= V h2
// on the slice, and the slice's ||w||^2 into partials C.
__global__ __launch_bounds__(kSplitThreads) void cgs_update_norm_kernel(
    const cd *__restrict__ V, size_t v_sys, size_t N, int R, int kc,
    cd *__restrict__ w, const cd *__restrict__ pB, cd *__restrict__ h2,
    cd *__restrict__ pC, int B, const int *__restrict__ active) {
    extern __shared__ double red[];            // h2 [2 kc], wave sums
    const int b = blockIdx.x, s = blockIdx.y;
    if (!active[s]) return;
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    double *coef = red;
    double *wsum = red + 2 * kc;
    split_sum_parts(pB + (size_t)s * (R + 1) * B, B, kc, coef,
                    h2 + (size_t)s * (R + 1), b, tid);
    size_t idx[kSplitPer];
    bool ok[kSplitPer];
    cd x[kSplitPer];
    cd *ws = w + (size_t)s * N;
    split_load(ws, N, b, tid, idx, ok, x);
    __syncthreads();
    split_update(V + (size_t)s * v_sys, N, kc, coef, idx, ok, x);
    double nn = 0.0;
#pragma unroll
    for (int m = 0; m < kSplitPer; ++m) {
        if (ok[m]) ws[idx[m]] = x[m];
        nn += x[m].r * x[m].r + x[m].i * x[m].i;
    }
    nn = wave_sum(nn);
    if (lane == 0) wsum[wv] = nn;
    __syncthreads();
    if (tid == 0) {
        double t = wsum[0];
        for (int u = 1; u < kSplitWaves; ++u) t += wsum[u];
        pC[(size_t)s * B + b] = cd{t, 0.0};
    }
}

// ||w||^2 before (partials A) and after (partials C) CGS2, summed in slice
// order by every workgroup of the system, then v_{col+1} = w / ||w|| on
// the slice (zero on breakdown, as arnoldi_finish_sys sets 1/||w||).
// Slice 0 also runs the Hessenberg/Givens update of arnoldi_finish_sys on
// LDS copies of the column and the stored rotations.  When that update
// stops the system, the other slices may or may not see the cleared
// active flag; a stopped system never reads v_{col+1}, so either is fine.
__global__ __launch_bounds__(kSplitThreads) void arnoldi_finish_scale_kernel(
    cd *__restrict__ V, size_t v_sys, size_t N, int R, int col,
    const cd *__restrict__ w, const cd *__restrict__ pA,
    const cd *__restrict__ pC, const cd *__restrict__ h1,
    const cd *__restrict__ h2, int B, cd *H, cd *G, cd *S, SysState *st,
    double *inv, int *active, double eps, Flags fl) {
    // LDS (doubles): partials [2 B], then slice 0's h1 / working column
    // [2 (kc + 1)], h2 [2 kc] and rotations [4 col]
    extern __shared__ double red[];
    const int b = blockIdx.x, s = blockIdx.y;
    if (!active[s]) return;
    const int tid = threadIdx.x, kc = col + 1;
    double *pn = red;
    cd *h1l = (cd *)(red + 2 * B);
    cd *h2l = h1l + (kc + 1);
    cd *gl = h2l + kc;
    const cd *pn0 = pA + ((size_t)s * (R + 2) + R + 1) * B;
    for (int u = tid; u < B; u += kSplitThreads)
        pn[u] = pn0[u].r, pn[B + u] = pC[(size_t)s * B + u].r;
    if (b == 0) {
        for (int t = tid; t < kc; t += kSplitThreads)
            h1l[t] = h1[(size_t)s * (R + 1) + t],
            h2l[t] = h2[(size_t)s * (R + 1) + t];
        for (int t = tid; t < 2 * col; t += kSplitThreads)
            gl[t] = G[(size_t)s * R * 2 + t];
    }
    __syncthreads();
    double n0 = 0.0, nw = 0.0;
    for (int u = 0; u < B; ++u) n0 += pn[u], nw += pn[B + u];
    // 1/||w|| exactly as arnoldi_finish_sys forms it
    const double hn = sqrt(nw);
    const double sc = hn <= eps * sqrt(n0) ? 0.0 : 1.0 / hn;
    if (b == 0 && tid == 0)
        arnoldi_finish_sys(s, R, col, H, G, S, h1l, h2l, gl, h1l,
                           S[(size_t)s * (R + 1) + col], nw, n0, st, inv,
                           active, eps, fl);
    const cd *ws = w + (size_t)s * N;
    cd *vn = V + (size_t)s * v_sys + (size_t)kc * N;
#pragma unroll
    for (int m = 0; m < kSplitPer; ++m) {
        const size_t i = (size_t)b * kSplitLen + tid + (size_t)m * kSplitThreads;
        if (i < N) vn[i] = cscale(ws[i], sc);
    }
}

// Start of a restart cycle: inv_s = 1/||M r_s||, S_s = [||M r_s||, 0, ...]
// (the first column of the Arnoldi basis is M r_s * inv_s).
__global__ void cycle_start_kernel(int nsys, int R, const cd *nrm,
                                   double *inv, cd *S, Flags fl) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i <= (size_t)R) fl.gate[i] = 0, fl.host[i] = 0;   // R columns, `any`
    if (i >= (size_t)nsys * (R + 1)) return;
    const int s = (int)(i / (R + 1)), q = (int)(i % (R + 1));
    const double t = sqrt(nrm[s].r);
    S[i] = q == 0 ? cd{t, 0.0} : cd{0.0, 0.0};
    if (q == 0) inv[s] = t > 0.0 ? 1.0 / t : 0.0;
}

// SciPy's gmres set-up for every system from ||b||^2, ||M b||^2 and
// ||r0||^2, and the first cycle's active flags.
__global__ void gmres_setup_kernel(int nsys, const cd *nb, const cd *nmb,
                                   const cd *nr, double tol, int maxiter,
                                   SysState *st, int *active) {
    const int s = blockIdx.x * 64 + threadIdx.x;
    if (s >= nsys) return;
    SysState q{};
    q.bnrm2 = sqrt(nb[s].r);
    q.atol = tol * q.bnrm2;
    q.pmf = 1.0;
    q.code = maxiter;
    if (q.bnrm2 == 0.0) q.done = 1, q.code = 0;   // SciPy returns b (= 0)
    const double Mb = sqrt(nmb[s].r);
    q.ptol = q.bnrm2 > 0.0 ? Mb * fmin(1.0, q.atol / q.bnrm2) : 0.0;
    q.rnorm = sqrt(nr[s].r);
    if (!q.done && q.rnorm < q.atol) q.done = 1, q.code = 0;
    st[s] = q;
    active[s] = !q.done;
}

// End of a restart cycle for every system not yet done: SciPy's
// convergence test on the true residual ||r||^2 (nr), the breakdown exit
// and the update of the inner tolerance; then the next cycle's flags.
// `any` (cleared at the cycle's start) is set while a system continues.
__global__ void gmres_cycle_end_kernel(int nsys, const cd *nr, double eps,
                                       SysState *st, int *active, Flags fl,
                                       int any) {
    const int s = blockIdx.x * 64 + threadIdx.x;
    if (s >= nsys) return;
    SysState &q = st[s];
    if (!q.done) {
        q.rnorm = sqrt(nr[s].r);
        if (q.rnorm <= q.atol) {
            q.done = 1, q.code = 0;
        } else if (q.breakdown) {
            q.done = 1;
        } else {
            q.pmf = q.presid <= q.ptol ? fmax(eps, 0.25 * q.pmf)
                                       : fmin(1.0, 1.5 * q.pmf);
            q.ptol = q.presid * fmin(q.pmf, q.atol / q.rnorm);
        }
    }
    q.col = 0, q.breakdown = 0, q.inner_done = 0;
    active[s] = !q.done;
    if (!q.done) raise_flag(fl, any);
}

// After the inner loop (which clears a system's flag as its Arnoldi
// process ends): every system not yet done takes the update again.
__global__ void restore_active_kernel(int nsys, const SysState *st,
                                      int *active) {
    const int s = blockIdx.x * 64 + threadIdx.x;
    if (s < nsys) active[s] = !st[s].done;
}

// Upper-triangular solve of the Hessenberg least squares (SciPy's loop),
// y stored per system; ncol = col + 1 columns.
__global__ void lsq_kernel(int nsys, int restart, const cd *H, cd *S, cd *Y,
                           const SysState *st, const int *active) {
    const int s = blockIdx.x * 64 + threadIdx.x;
    if (s >= nsys || !active[s]) return;
    const int col = st[s].col;
    const cd *h = H + (size_t)s * restart * (restart + 1);
    cd *Sv = S + (size_t)s * (restart + 1);
    cd *y = Y + (size_t)s * restart;
    auto hh = [&](int c, int r) { return h[(size_t)c * (restart + 1) + r]; };
    const cd dcc = hh(col, col);
    if (dcc.r == 0.0 && dcc.i == 0.0) Sv[col] = cd{0.0, 0.0};
    for (int r = 0; r <= col; ++r) y[r] = Sv[r];
    for (int kk = col; kk > 0; --kk) {
        if (y[kk].r != 0.0 || y[kk].i != 0.0) {
            const cd d = hh(kk, kk);
            const double dn = d.r * d.r + d.i * d.i;
            y[kk] = cd{(y[kk].r * d.r + y[kk].i * d.i) / dn,
                       (y[kk].i * d.r - y[kk].r * d.i) / dn};
            const cd t = y[kk];
            for (int r = 0; r < kk; ++r) {
                const cd a = hh(kk, r);
                y[r].r -= t.r * a.r - t.i * a.i;
                y[r].i -= t.r * a.i + t.i * a.r;
            }
        }
    }
    if (y[0].r != 0.0 || y[0].i != 0.0) {
        const cd d = hh(0, 0);
        const double dn = d.r * d.r + d.i * d.i;
        y[0] = cd{(y[0].r * d.r + y[0].i * d.i) / dn,
                  (y[0].i * d.r - y[0].r * d.i) / dn};
    }
    for (int r = col + 1; r < restart; ++r) y[r] = cd{0.0, 0.0};
}

__global__ void exit_wave_kernel(const cd *__restrict__ v,
                                 const cd *__restrict__ E, BraggK k, size_t n,
                                 cd *__restrict__ out) {
    const size_t plane = (size_t)k.nyc * k.Nx;
    for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < n;
         t += (size_t)gridDim.x * 256) {
        const size_t s = t / plane, pix = t % plane;
        out[t] = cconjmul(E[table_offset(k, (int)s) + pix], v[t]);
    }
}

// ---- workspace layout --------------------------------------------------
struct Work {
    cd *V, *b, *r, *w, *av, *H, *G, *S, *Y, *h1, *h2, *nrm, *nrm0, *pinv,
        *part, *pA, *pB, *pC;
    double *inv;
    SysState *st;
    int *active, *gate;
    Flags flags() const { return Flags{gate, still}; }
    int *still;   // page-locked host words, set by dyf_bragg_solve
    size_t bytes;
};

static size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// parts of a batched dot product: one 1024-lane workgroup per 32768
// elements (four loads of each vector in flight per lane)
static int reduce_parts(size_t N) {
    const int P = (int)std::min<size_t>((N + 32767) / 32768, 64);
    return std::max(P, 1);
}

static Work carve(void *base, int nsys, size_t N, int restart, int Nx) {
    Work w{};
    char *p = (char *)base;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char *q = p ? p + off : nullptr;
        off += align256(bytes);
        return q;
    };
    const size_t R = restart;
    const int P = reduce_parts(N);
    w.V = (cd *)take(sizeof(cd) * nsys * (R + 1) * N);
    w.b = (cd *)take(sizeof(cd) * nsys * N);
    w.r = (cd *)take(sizeof(cd) * nsys * N);
    w.w = (cd *)take(sizeof(cd) * nsys * N);
    w.av = (cd *)take(sizeof(cd) * nsys * N);
    w.H = (cd *)take(sizeof(cd) * nsys * R * (R + 1));
    w.G = (cd *)take(sizeof(cd) * nsys * R * 2);
    w.S = (cd *)take(sizeof(cd) * nsys * (R + 1));
    w.Y = (cd *)take(sizeof(cd) * nsys * R);
    w.h1 = (cd *)take(sizeof(cd) * nsys * (R + 1));
    w.h2 = (cd *)take(sizeof(cd) * nsys * (R + 1));
    w.nrm = (cd *)take(sizeof(cd) * nsys);
    w.nrm0 = (cd *)take(sizeof(cd) * nsys);
    w.pinv = (cd *)take(sizeof(cd) * nsys * symbol_elems(Nx));
    w.part = (cd *)take(sizeof(cd) * nsys * (R + 2) * P);
    // partials of the sliced Arnoldi step (systems too large for one
    // workgroup); sized for every N, since the path is chosen per solve
    const size_t B = split_slices(N);
    w.pA = (cd *)take(sizeof(cd) * nsys * (R + 2) * B);
    w.pB = (cd *)take(sizeof(cd) * nsys * (R + 1) * B);
    w.pC = (cd *)take(sizeof(cd) * nsys * B);
    w.inv = (double *)take(sizeof(double) * nsys);
    w.st = (SysState *)take(sizeof(SysState) * nsys);
    w.active = (int *)take(sizeof(int) * nsys);
    w.gate = (int *)take(sizeof(int) * (R + 1));   // per column, `any`
    w.bytes = off;
    return w;
}

// dot products of kc vectors A_q (per system) against B, summed into out
static int dots(const cd *A, size_t a_sys, size_t a_vec, int kc, const cd *Bv,
                size_t b_sys, size_t N, cd *out, int ostride, cd *part,
                int nsys, const int *active, hipStream_t stream) {
    const int P = reduce_parts(N);
    batched_dot_kernel<<<dim3(nsys * kc, P), kRedThreads, 0, stream>>>(
        A, a_sys, a_vec, kc, Bv, b_sys, N, P, part, active, out, ostride);
    DYF_LAUNCHED("batched_dot_kernel");
    if (P == 1) return DYF_OK;            // written to out directly
    const int n = nsys * kc;
    sum_parts_kernel<<<(n + 255) / 256, 256, 0, stream>>>(part, P, n, out,
                                                          ostride, active, kc);
    DYF_LAUNCHED("sum_parts_kernel");
    return DYF_OK;
}

// ---- fixed-point (trapezoid round-trip) solve, bragg/solver.py:63-132 -
// One workgroup per (system, plane) runs every iteration.  Each thread
// owns columns x = tid + 256 m of both fields (kept in global memory: the
// other sweep needs every depth); only the periodic one-column shift
// crosses threads, through a double-buffered LDS row, so one barrier per
// depth step suffices.  The next step's e^{iH} and frozen-field values
// are loaded before the barrier of the current step.
constexpr int kFpThreads = 256;

__device__ __forceinline__ double block_max(double v, double *red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    const int w = threadIdx.x >> 6;
    __syncthreads();                       // red[] free from the last use
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double m = red[0];
    for (int i = 1; i < kFpThreads / 64; ++i) m = fmax(m, red[i]);
    return m;
}

__global__ __launch_bounds__(kFpThreads) void bragg_fixed_point_kernel(
    BraggK k, const double *__restrict__ s_dev, const cd *__restrict__ E,
    const cd *__restrict__ inc, int n_iter, double tol, cd *__restrict__ D0,
    cd *__restrict__ Dg, cd *__restrict__ out, dyf_fixed_info *__restrict__ info) {
    extern __shared__ cd lds_fp[];
    __shared__ double red[kFpThreads / 64];
    const int sys = blockIdx.x / k.nyc, j = blockIdx.x % k.nyc;
    const int Nx = k.Nx, Nz = k.Nz, tid = threadIdx.x;
    const cd *Es = E + table_offset(k, sys) + (size_t)j * Nx;
    const size_t zs = (size_t)k.tpl * Nx;             // e^{iH} depth stride
    cd *d0 = D0 + (size_t)blockIdx.x * Nz * Nx;
    cd *dg = Dg + (size_t)blockIdx.x * Nz * Nx;
    const double alpha = -2.0 * kPi * s_dev[sys];
    const cd e0 = cexp(k.s0r * k.ds, k.s0i * k.ds);
    const cd eg = cexp(k.s0r * k.ds, (k.s0i + alpha) * k.ds);
    const cd sb = {k.sbr, k.sbi}, sh = {k.shr, k.shi};
    const double hds = 0.5 * k.ds;

    cd pv[kMaxPerThread];
#pragma unroll
    for (int m = 0; m < kMaxPerThread; ++m) {
        const int x = tid + m * kFpThreads;
        if (x >= Nx) continue;
        d0[x] = inc ? inc[x] : cd{1.0, 0.0};
        for (int z = 0; z < Nz; ++z) dg[(size_t)z * Nx + x] = cd{0.0, 0.0};
        pv[m] = cd{0.0, 0.0};
    }
    double res = __builtin_nan("");
    int iters = n_iter, conv = 0, par = 0;
    for (int it = 0; it < n_iter; ++it) {
        // forward: D0[z+1] = e0 roll(D0[z] + hds Cb[z] Dg[z]) + hds Cb[z+1] Dg[z+1]
        cd cur[kMaxPerThread], en[kMaxPerThread], gn[kMaxPerThread];
#pragma unroll
        for (int m = 0; m < kMaxPerThread; ++m) {
            const int x = tid + m * kFpThreads;
            if (x >= Nx) continue;
            cur[m] = d0[x];
            en[m] = Es[x];
            gn[m] = dg[x];
        }
        for (int z = 0; z < Nz - 1; ++z) {
            cd *buf = lds_fp + par * Nx;
            par ^= 1;
#pragma unroll
            for (int m = 0; m < kMaxPerThread; ++m) {
                const int x = tid + m * kFpThreads;
                if (x >= Nx) continue;
                const cd cbg = cmul(cmul(sb, en[m]), gn[m]);
                buf[x] = cmul(e0, cadd(cur[m], cscale(cbg, hds)));
                en[m] = Es[(size_t)(z + 1) * zs + x];
                gn[m] = dg[(size_t)(z + 1) * Nx + x];
            }
            lds_barrier();
#pragma unroll
            for (int m = 0; m < kMaxPerThread; ++m) {
                const int x = tid + m * kFpThreads;
                if (x >= Nx) continue;
                const cd cbg = cmul(cmul(sb, en[m]), gn[m]);
                cur[m] = cadd(buf[x > 0 ? x - 1 : Nx - 1], cscale(cbg, hds));
                d0[(size_t)(z + 1) * Nx + x] = cur[m];
            }
        }
        // backward: Dg[z] = eg roll(Dg[z+1] + hds Ch[z+1] D0[z+1]) + hds Ch[z] D0[z]
        cd nx[kMaxPerThread];
#pragma unroll
        for (int m = 0; m < kMaxPerThread; ++m) {
            const int x = tid + m * kFpThreads;
            if (x >= Nx) continue;
            nx[m] = cd{0.0, 0.0};
            dg[(size_t)(Nz - 1) * Nx + x] = nx[m];
            en[m] = Es[(size_t)(Nz - 1) * zs + x];
            gn[m] = cur[m];                          // D0[Nz-1]
        }
        for (int z = Nz - 2; z >= 0; --z) {
            cd *buf = lds_fp + par * Nx;
            par ^= 1;
#pragma unroll
            for (int m = 0; m < kMaxPerThread; ++m) {
                const int x = tid + m * kFpThreads;
                if (x >= Nx) continue;
                const cd chp = cmul(cconjmul(en[m], sh), gn[m]);
                buf[x] = cmul(eg, cadd(nx[m], cscale(chp, hds)));
                en[m] = Es[(size_t)z * zs + x];
                gn[m] = d0[(size_t)z * Nx + x];
            }
            lds_barrier();
#pragma unroll
            for (int m = 0; m < kMaxPerThread; ++m) {
                const int x = tid + m * kFpThreads;
                if (x >= Nx) continue;
                const cd chp = cmul(cconjmul(en[m], sh), gn[m]);
                nx[m] = cadd(buf[x > 0 ? x - 1 : Nx - 1], cscale(chp, hds));
                dg[(size_t)z * Nx + x] = nx[m];
            }
        }
        // convergence on the surface value Dg[0]
        double dmax = 0.0, amax = 0.0;
#pragma unroll
        for (int m = 0; m < kMaxPerThread; ++m) {
            const int x = tid + m * kFpThreads;
            if (x >= Nx) continue;
            const double dr = nx[m].r - pv[m].r, di = nx[m].i - pv[m].i;
            dmax = fmax(dmax, sqrt(dr * dr + di * di));
            amax = fmax(amax, sqrt(nx[m].r * nx[m].r + nx[m].i * nx[m].i));
            pv[m] = nx[m];
        }
        dmax = block_max(dmax, red);
        amax = block_max(amax, red);
        if (it > 0) {
            res = dmax / (amax + 1e-300);
            if (res < tol) {
                conv = 1, iters = it + 1;
                break;
            }
        }
    }
#pragma unroll
    for (int m = 0; m < kMaxPerThread; ++m) {
        const int x = tid + m * kFpThreads;
        if (x >= Nx) continue;
        out[(size_t)blockIdx.x * Nx + x] = cconjmul(Es[x], pv[m]);
    }
    if (tid == 0) info[blockIdx.x] = dyf_fixed_info{res, iters, conv, };
}

// Wavefront form of the same solve, for Nx = Lu C with C | Nx, Lu <= 64
// (the march's wavefront layout): lane l keeps columns [lC, lC + C) of the
// field being marched in registers, the one-column roll is a register
// rename plus one ds_bpermute (lane 0 reads lane Lu - 1), and there is no
// barrier.  The stored fields are lane-interleaved per depth row
// (element (z, i, lane) at (z C + i) 64 + lane), so every row access is
// one contiguous 1 KB wave access per register.  The convergence test
// reads the previous sweep's Dg(0) from memory just before it is
// overwritten, so it keeps no copy in registers.
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

template <int C>
__global__ __launch_bounds__(kWave) void bragg_fixed_point_wave_kernel(
    BraggK k, const double *__restrict__ s_dev, const cd *__restrict__ E,
    const cd *__restrict__ inc, int n_iter, double tol, cd *__restrict__ D0,
    cd *__restrict__ Dg, cd *__restrict__ out, dyf_fixed_info *__restrict__ info) {
    const int sys = blockIdx.x / k.nyc, j = blockIdx.x % k.nyc;
    const int Nx = k.Nx, Nz = k.Nz, lane = threadIdx.x, Lu = Nx / C;
    const bool real = lane < Lu;
    const int x0 = (real ? lane : Lu - 1) * C;   // idle lanes read in bounds
    const int src = (real ? (lane == 0 ? Lu - 1 : lane - 1) : lane) << 2;
    const cd *Es = E + table_offset(k, sys) + (size_t)j * Nx + x0;
    const size_t zs = (size_t)k.tpl * Nx;             // e^{iH} depth stride
    const size_t row = (size_t)C * kWave;
    cd *d0 = D0 + (size_t)blockIdx.x * Nz * row + lane;
    cd *dg = Dg + (size_t)blockIdx.x * Nz * row + lane;
    const double alpha = -2.0 * kPi * s_dev[sys];
    const cd e0 = cexp(k.s0r * k.ds, k.s0i * k.ds);
    const cd eg = cexp(k.s0r * k.ds, (k.s0i + alpha) * k.ds);
    const cd sb = {k.sbr, k.sbi}, sh = {k.shr, k.shi};
    const double hds = 0.5 * k.ds;

#pragma unroll
    for (int i = 0; i < C; ++i)
        d0[i * kWave] = inc ? inc[x0 + i] : cd{1.0, 0.0};
    for (int z = 0; z < Nz; ++z)
#pragma unroll
        for (int i = 0; i < C; ++i) dg[(size_t)z * row + i * kWave] = cd{0.0, 0.0};
    double res = __builtin_nan("");
    int iters = n_iter, conv = 0;
    // The e^{iH} and frozen-field values of depth z+1 are loaded during
    // step z-1 (index clamped at the last row, so the loop has no branch)
    // and used in step z: a whole step covers their latency.
    cd cur[C], cb[C], en[C], fn[C];
    for (int it = 0; it < n_iter; ++it) {
        // forward: D0[z+1] = e0 roll(D0[z] + hds Cb[z] Dg[z]) + hds Cb[z+1] Dg[z+1]
#pragma unroll
        for (int i = 0; i < C; ++i) {
            cur[i] = d0[i * kWave];
            cb[i] = cscale(cmul(cmul(sb, Es[i]), dg[i * kWave]), hds);
            en[i] = Es[zs + i];
            fn[i] = dg[row + i * kWave];
        }
        for (int z = 0; z < Nz - 1; ++z) {
            const int zp = z + 2 < Nz ? z + 2 : Nz - 1;
            const cd *En = Es + (size_t)zp * zs;
            const cd *Gn = dg + (size_t)zp * row;
#pragma unroll
            for (int i = 0; i < C; ++i) {
                cur[i] = cmul(e0, cadd(cur[i], cb[i]));
                cb[i] = cscale(cmul(cmul(sb, en[i]), fn[i]), hds);
            }
            // keep the next loads below the last use of en, fn (hoisted
            // above it they need a second register set and spill)
            asm volatile("" ::: "memory");
#pragma unroll
            for (int i = 0; i < C; ++i) {
                en[i] = En[i];
                fn[i] = Gn[i * kWave];
            }
            const cd w = wave_fetch(cur[C - 1], src);
#pragma unroll
            for (int i = C - 1; i > 0; --i) cur[i] = cadd(cur[i - 1], cb[i]);
            cur[0] = cadd(w, cb[0]);
            cd *Dn = d0 + (size_t)(z + 1) * row;
#pragma unroll
            for (int i = 0; i < C; ++i) Dn[i * kWave] = cur[i];
        }
        // backward: Dg[z] = eg roll(Dg[z+1] + hds Ch[z+1] D0[z+1]) + hds Ch[z] D0[z]
        const cd *Ez = Es + (size_t)(Nz - 1) * zs;
#pragma unroll
        for (int i = 0; i < C; ++i) {
            dg[(size_t)(Nz - 1) * row + i * kWave] = cd{0.0, 0.0};
            cb[i] = cscale(cmul(cconjmul(Ez[i], sh), cur[i]), hds);
            cur[i] = cd{0.0, 0.0};
            en[i] = Es[(size_t)(Nz - 2) * zs + i];
            fn[i] = d0[(size_t)(Nz - 2) * row + i * kWave];
        }
        // the step z (z = 0 last, separately: its convergence test would
        // otherwise be if-converted into every step)
        auto back_step = [&](int z) {
            const int zp = z > 0 ? z - 1 : 0;
            const cd *En = Es + (size_t)zp * zs;
            const cd *Fn = d0 + (size_t)zp * row;
#pragma unroll
            for (int i = 0; i < C; ++i) {
                cur[i] = cmul(eg, cadd(cur[i], cb[i]));
                cb[i] = cscale(cmul(cconjmul(en[i], sh), fn[i]), hds);
            }
            asm volatile("" ::: "memory");
#pragma unroll
            for (int i = 0; i < C; ++i) {
                en[i] = En[i];
                fn[i] = Fn[i * kWave];
            }
            const cd w = wave_fetch(cur[C - 1], src);
#pragma unroll
            for (int i = C - 1; i > 0; --i) cur[i] = cadd(cur[i - 1], cb[i]);
            cur[0] = cadd(w, cb[0]);
        };
        for (int z = Nz - 2; z >= 1; --z) {
            back_step(z);
            cd *Dn = dg + (size_t)z * row;
#pragma unroll
            for (int i = 0; i < C; ++i) Dn[i * kWave] = cur[i];
        }
        back_step(0);
        // convergence on the surface value Dg[0] against the last sweep's
        double dmax = 0.0, amax = 0.0;
#pragma unroll
        for (int i = 0; i < C; ++i) {
            const cd pv = dg[i * kWave];
            const double dr = cur[i].r - pv.r, di = cur[i].i - pv.i;
            dmax = fmax(dmax, sqrt(dr * dr + di * di));
            amax = fmax(amax, sqrt(cur[i].r * cur[i].r + cur[i].i * cur[i].i));
            dg[i * kWave] = cur[i];
        }
        dmax = wave_max(real ? dmax : 0.0);
        amax = wave_max(real ? amax : 0.0);
        if (it > 0) {
            res = dmax / (amax + 1e-300);
            if (res < tol) {
                conv = 1, iters = it + 1;
                break;
            }
        }
    }
    if (real) {
#pragma unroll
        for (int i = 0; i < C; ++i)
            out[(size_t)blockIdx.x * Nx + x0 + i] = cconjmul(Es[i], cur[i]);
    }
    if (lane == 0) info[blockIdx.x] = dyf_fixed_info{res, iters, conv, };
}

// The same solve with each row split over W wavefronts of a workgroup,
// for launches with few rows (one wavefront per row leaves most SIMDs idle
// and every depth step waiting on memory).  Wavefront w owns columns
// [w Nx/W, (w+1) Nx/W) on its last Lo lanes and marches the Hc = Hl C
// columns to their left in its first Hl lanes.  Both sweeps roll the
// marched field one column towards +x per depth step, so wrong values
// enter only at lane 0 and move right one column per step: the owned
// columns stay exact for Hc steps, and every P <= Hc steps the halo takes
// the left neighbour's current values of those columns through LDS (one
// barrier, buffers alternating).  The coupling terms come from the stored
// fields at each lane's own columns, exact in the halo too.  Fields are
// stored (row, z, x); the e^{iH} and field values of the next D depths are
// in flight in D register sets with fixed roles (steps unrolled by D).
// Stores carry no branch: halo lanes (and the z = 0 row of the backward
// sweep, written after the convergence test) store to a per-lane slot of
// `trash`.  A branch around them made the compiler wait for the loads
// of later depths too (stores count in the same counter).  The sweeps
// read columns other wavefronts of the workgroup stored; the barrier
// between sweeps orders those global stores before the loads (the
// wavefronts of a workgroup share one CU and its vector L1).
template <int C, int W, int D = (4 / C > 2 ? 4 / C : 2)>
__global__ __launch_bounds__(kWave * W) void bragg_fixed_point_halo_kernel(
    BraggK k, const double *__restrict__ s_dev, const cd *__restrict__ E,
    const cd *__restrict__ inc, int n_iter, double tol, cd *__restrict__ D0,
    cd *__restrict__ Dg, cd *__restrict__ out, dyf_fixed_info *__restrict__ info,
    cd *__restrict__ trash) {
    __shared__ cd xch[2][W][kWave / 2 * C];      // [parity][wave][Hc]
    __shared__ double red[2][W];
    const int sys = blockIdx.x / k.nyc, j = blockIdx.x % k.nyc;
    const int Nx = k.Nx, Nz = k.Nz, lane = threadIdx.x % kWave;
    const int wv = threadIdx.x / kWave, wl = wv == 0 ? W - 1 : wv - 1;
    const int Nw = Nx / W, Lo = Nw / C, Hl = kWave - Lo, Hc = Hl * C;
    const int P = Hc / D * D;                  // steps between refreshes
    const int rel = lane - Hl;
    const bool own = rel >= 0;
    const int x0 = ((wv * Nw + rel * C) % Nx + Nx) % Nx;
    const int src = (lane == 0 ? kWave - 1 : lane - 1) << 2;
    const cd *Es = E + table_offset(k, sys) + (size_t)j * Nx + x0;
    const size_t zs = (size_t)k.tpl * Nx;             // e^{iH} depth stride
    cd *d0 = D0 + (size_t)blockIdx.x * Nz * Nx + x0;
    cd *dg = Dg + (size_t)blockIdx.x * Nz * Nx + x0;
    // this lane's slot of `trash` (halo lanes are lanes < Hl <= 32)
    cd *tr = trash + ((size_t)blockIdx.x * W + wv) * (kWave / 2 * C) +
             (own ? 0 : lane * C);
    const double alpha = -2.0 * kPi * s_dev[sys];
    const cd e0 = cexp(k.s0r * k.ds, k.s0i * k.ds);
    const cd eg = cexp(k.s0r * k.ds, (k.s0i + alpha) * k.ds);
    const cd sb = {k.sbr, k.sbi}, sh = {k.shr, k.shi};
    const double hds = 0.5 * k.ds;

    if (own) {
#pragma unroll
        for (int i = 0; i < C; ++i) d0[i] = inc ? inc[x0 + i] : cd{1.0, 0.0};
        for (int z = 0; z < Nz; ++z)
#pragma unroll
            for (int i = 0; i < C; ++i) dg[(size_t)z * Nx + i] = cd{0.0, 0.0};
    }
    __syncthreads();
    double res = __builtin_nan("");
    int iters = n_iter, conv = 0, par = 0;
    cd cur[C], cb[C], en[D][C], fn[D][C];
    auto refresh = [&]() {
        if (lane >= kWave - Hl) {
#pragma unroll
            for (int i = 0; i < C; ++i)
                xch[par][wv][(lane - (kWave - Hl)) * C + i] = cur[i];
        }
        __syncthreads();
        if (lane < Hl) {
#pragma unroll
            for (int i = 0; i < C; ++i) cur[i] = xch[par][wl][lane * C + i];
        }
        par ^= 1;
    };
    // forward step z with the set holding depth z + 1; reloads it with
    // depth z + 1 + D
    auto fwd = [&](int z, cd (&e)[C], cd (&f)[C]) {
#pragma unroll
        for (int i = 0; i < C; ++i) {
            cur[i] = cmul(e0, cadd(cur[i], cb[i]));
            cb[i] = cscale(cmul(cmul(sb, e[i]), f[i]), hds);
        }
        asm volatile("" ::: "memory");
        const int zp = z + 1 + D < Nz ? z + 1 + D : Nz - 1;
#pragma unroll
        for (int i = 0; i < C; ++i) {
            e[i] = Es[(size_t)zp * zs + i];
            f[i] = dg[(size_t)zp * Nx + i];
        }
        const cd w = wave_fetch(cur[C - 1], src);
#pragma unroll
        for (int i = C - 1; i > 0; --i) cur[i] = cadd(cur[i - 1], cb[i]);
        cur[0] = cadd(w, cb[0]);
        cd *dst = own ? d0 + (size_t)(z + 1) * Nx : tr;
#pragma unroll
        for (int i = 0; i < C; ++i) dst[i] = cur[i];
    };
    // backward step z with the set holding depth z; reloads it with
    // depth z - D
    auto bwd = [&](int z, cd (&e)[C], cd (&f)[C]) {
#pragma unroll
        for (int i = 0; i < C; ++i) {
            cur[i] = cmul(eg, cadd(cur[i], cb[i]));
            cb[i] = cscale(cmul(cconjmul(e[i], sh), f[i]), hds);
        }
        asm volatile("" ::: "memory");
        const int zp = z - D > 0 ? z - D : 0;
#pragma unroll
        for (int i = 0; i < C; ++i) {
            e[i] = Es[(size_t)zp * zs + i];
            f[i] = d0[(size_t)zp * Nx + i];
        }
        const cd w = wave_fetch(cur[C - 1], src);
#pragma unroll
        for (int i = C - 1; i > 0; --i) cur[i] = cadd(cur[i - 1], cb[i]);
        cur[0] = cadd(w, cb[0]);
        cd *dst = own && z >= 1 ? dg + (size_t)z * Nx : tr;
#pragma unroll
        for (int i = 0; i < C; ++i) dst[i] = cur[i];
    };
    const int ns = Nz - 1;                     // steps per sweep
    for (int it = 0; it < n_iter; ++it) {
        // forward: D0[z+1] = e0 roll(D0[z] + hds Cb[z] Dg[z]) + hds Cb[z+1] Dg[z+1]
#pragma unroll
        for (int i = 0; i < C; ++i) {
            cur[i] = d0[i];
            cb[i] = cscale(cmul(cmul(sb, Es[i]), dg[i]), hds);
        }
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int zq = 1 + d < Nz ? 1 + d : Nz - 1;
#pragma unroll
            for (int i = 0; i < C; ++i) {
                en[d][i] = Es[(size_t)zq * zs + i];
                fn[d][i] = dg[(size_t)zq * Nx + i];
            }
        }
        // the first group is peeled: the loop is then entered with its
        // loads and stores in flight in the order of its back-edge, and
        // each set waits only for its own loads (stores count too)
        int t = 0, since = 0;
        if (D <= ns) {
#pragma unroll
            for (int d = 0; d < D; ++d) fwd(d, en[d], fn[d]);
            t = D, since = D;
        }
        for (; t + D <= ns; t += D) {
            if (since == P) refresh(), since = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) fwd(t + d, en[d], fn[d]);
            since += D;
        }
        if (t < ns && since == P) refresh();
#pragma unroll
        for (int d = 0; d < D; ++d)
            if (t + d < ns) fwd(t + d, en[d], fn[d]);
        __syncthreads();                       // D0 of every column stored
        // backward: Dg[z] = eg roll(Dg[z+1] + hds Ch[z+1] D0[z+1]) + hds Ch[z] D0[z]
        if (own) {
#pragma unroll
            for (int i = 0; i < C; ++i)
                dg[(size_t)(Nz - 1) * Nx + i] = cd{0.0, 0.0};
        }
#pragma unroll
        for (int i = 0; i < C; ++i) {
            const cd dz = d0[(size_t)(Nz - 1) * Nx + i];
            cb[i] = cscale(cmul(cconjmul(Es[(size_t)(Nz - 1) * zs + i], sh), dz),
                           hds);
            cur[i] = cd{0.0, 0.0};
        }
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int zq = Nz - 2 - d > 0 ? Nz - 2 - d : 0;
#pragma unroll
            for (int i = 0; i < C; ++i) {
                en[d][i] = Es[(size_t)zq * zs + i];
                fn[d][i] = d0[(size_t)zq * Nx + i];
            }
        }
        t = 0, since = 0;
        if (D <= ns) {
#pragma unroll
            for (int d = 0; d < D; ++d) bwd(Nz - 2 - d, en[d], fn[d]);
            t = D, since = D;
        }
        for (; t + D <= ns; t += D) {
            if (since == P) refresh(), since = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) bwd(Nz - 2 - (t + d), en[d], fn[d]);
            since += D;
        }
        if (t < ns && since == P) refresh();
#pragma unroll
        for (int d = 0; d < D; ++d)
            if (t + d < ns) bwd(Nz - 2 - (t + d), en[d], fn[d]);
        // convergence on the surface value Dg[0] against the last sweep's
        double dmax = 0.0, amax = 0.0;
        if (own) {
#pragma unroll
            for (int i = 0; i < C; ++i) {
                const cd pv = dg[i];
                const double dr = cur[i].r - pv.r, di = cur[i].i - pv.i;
                dmax = fmax(dmax, sqrt(dr * dr + di * di));
                amax = fmax(amax,
                            sqrt(cur[i].r * cur[i].r + cur[i].i * cur[i].i));
                dg[i] = cur[i];
            }
        }
        dmax = wave_max(dmax);
        amax = wave_max(amax);
        if (lane == 0) red[0][wv] = dmax, red[1][wv] = amax;
        __syncthreads();                       // also: Dg of every column stored
#pragma unroll
        for (int w = 0; w < W; ++w)
            dmax = fmax(dmax, red[0][w]), amax = fmax(amax, red[1][w]);
        __syncthreads();                       // red[] read before its reuse
        if (it > 0) {
            res = dmax / (amax + 1e-300);
            if (res < tol) {
                conv = 1, iters = it + 1;
                break;
            }
        }
    }
    if (own) {
#pragma unroll
        for (int i = 0; i < C; ++i)
            out[(size_t)blockIdx.x * Nx + x0 + i] = cconjmul(Es[i], cur[i]);
    }
    if (threadIdx.x == 0) info[blockIdx.x] = dyf_fixed_info{res, iters, conv, };
}

// Rows per launch below which the fixed-point solve splits rows over
// wavefronts: at most this many wavefronts in all.
constexpr long long kFpHaloWaves = 2048;

// Wavefronts per row and columns per lane of the fixed-point halo kernel:
// the largest W in {16, 8, 4, 2} with W | Nx and rows W <= kFpHaloWaves, and
// the smallest C <= 8 (<= 4 at W = 16) with C | Nx/W, Lo = Nx/(W C) <= 63 owned lanes,
// Hl = 64 - Lo <= Lo halo lanes and a halo of at least D(C) columns;
// {0, 0} if none.
static void fixed_halo_shape(int Nx, long long rows, int *W_, int *C_) {
    *W_ = 0, *C_ = 0;
    for (int W = 16; W >= 2; W /= 2) {
        if (Nx % W || rows * W > kFpHaloWaves) continue;
        const int Nw = Nx / W;
        for (int c = 1; c <= (W == 16 ? 4 : 8); ++c) {   // as instantiated
            if (Nw % c) continue;
            const int Lo = Nw / c, Hl = kWave - Lo;
            const int D = 4 / c > 2 ? 4 / c : 2;
            if (Lo <= kWave - 1 && Hl <= Lo && Hl * c >= D) {
                *W_ = W, *C_ = c;
                return;
            }
        }
    }
}

template <int W, int... Cs>
static auto fixed_halo_fn(int C) {
    decltype(&bragg_fixed_point_halo_kernel<1, 2>) f = nullptr;
    ((C == Cs ? (f = bragg_fixed_point_halo_kernel<Cs, W>, 0) : 0), ...);
    return f;
}

template <int... Cs>
static auto fixed_wave_fn(int C) {
    decltype(&bragg_fixed_point_wave_kernel<1>) f = nullptr;
    ((C == Cs ? (f = bragg_fixed_point_wave_kernel<Cs>, 0) : 0), ...);
    return f;
}

}  // namespace dyf

using namespace dyf;

static int check_bragg(const dyf_bragg_args *a, int nsys) {
    if (!a || a->Nz < 2 || a->Nx < 2 || a->nyc < 1 || nsys < 1)
        return fail(DYF_EINVAL, "bad Bragg arguments");
    if (a->Nx > kMarchThreads * kMaxPerThread)
        return fail(DYF_EINVAL, "Nx=%d exceeds %d", a->Nx,
                    kMarchThreads * kMaxPerThread);
    if (!(a->ds > 0.0)) return fail(DYF_EINVAL, "ds must be positive");
    if (a->plane_systems < 0 ||
        (a->plane_systems > 0 &&
         (a->nyc != 1 || nsys % a->plane_systems != 0)))
        return fail(DYF_EINVAL,
                    "plane_systems=%d needs nyc=1 and nsys (%d) a multiple",
                    a->plane_systems, nsys);
    return DYF_OK;
}

extern "C" int dyf_bragg_march(const dyf_bragg_args *a, const double *s_dev,
                               int nsys, const void *E, const void *incident,
                               const void *Dg0, void *out, void *stream) {
    int rc = check_bragg(a, nsys);
    if (rc) return rc;
    if (!s_dev || !E || !out) return fail(DYF_EINVAL, "NULL s_dev, E or out");
    const BraggK k = make_k(a, nsys);
    const size_t N = (size_t)a->nyc * a->Nx;
    return launch_march(k, s_dev, (const cd *)E, (const cd *)incident,
                        (const cd *)Dg0, N, (cd *)out, N, 1.0, nullptr,
                        (hipStream_t)stream);
}

extern "C" int dyf_bragg_precondition(const dyf_bragg_args *a,
                                      const double *s_dev, int nsys,
                                      const void *v, void *w, void *pinv_ws,
                                      void *stream_) {
    int rc = check_bragg(a, nsys);
    if (rc) return rc;
    hipStream_t stream = (hipStream_t)stream_;
    const BraggK k = make_k(a, nsys);
    if ((rc = make_symbol(k, s_dev, (cd *)pinv_ws, stream))) return rc;
    return apply_precon(k, (const cd *)pinv_ws, (cd *)v, (cd *)w, stream);
}

extern "C" int dyf_bragg_symbol_bytes(const dyf_bragg_args *a, int nsys,
                                      size_t *bytes) {
    int rc = check_bragg(a, nsys);
    if (rc) return rc;
    if (!bytes) return fail(DYF_EINVAL, "bytes is NULL");
    FftPlan p;
    if ((rc = fft_plan_for(a->Nx, &p))) return rc;
    *bytes = sizeof(cd) * (size_t)nsys * fft_symbol_elems(p, a->Nx);
    return DYF_OK;
}

extern "C" int dyf_bragg_workspace_bytes(const dyf_bragg_args *a, int nsys,
                                         int restart, size_t *bytes) {
    int rc = check_bragg(a, nsys);
    if (rc) return rc;
    if (restart < 1 || !bytes) return fail(DYF_EINVAL, "bad restart/bytes");
    const size_t N = (size_t)a->nyc * a->Nx;
    *bytes = carve(nullptr, nsys, N, restart, a->Nx).bytes;
    return DYF_OK;
}

extern "C" int dyf_bragg_exit_wave(const dyf_bragg_args *a, int nsys,
                                   const void *E, const void *v, void *out,
                                   void *stream) {
    int rc = check_bragg(a, nsys);
    if (rc) return rc;
    const size_t n = (size_t)nsys * a->nyc * a->Nx;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 8192);
    exit_wave_kernel<<<blocks, 256, 0, (hipStream_t)stream>>>(
        (const cd *)v, (const cd *)E, make_k(a, nsys), n, (cd *)out);
    DYF_LAUNCHED("exit_wave_kernel");
    return DYF_OK;
}

// Entries per stored depth row of the fixed-point fields: the wavefront
// kernel's lane-interleaved row (64 C) when Nx takes it, else Nx.
static size_t fixed_row(int Nx) {
    const int C = wave_columns(Nx);
    return C ? (size_t)C * kWave : (size_t)Nx;
}

extern "C" int dyf_bragg_fixed_point_bytes(const dyf_bragg_args *a, int nsys,
                                           size_t *bytes) {
    int rc = check_bragg(a, nsys);
    if (rc) return rc;
    if (!bytes) return fail(DYF_EINVAL, "bytes is NULL");
    const size_t n = (size_t)nsys * a->nyc;
    int Wh = 0, Ch = 0;
    fixed_halo_shape(a->Nx, (long long)n, &Wh, &Ch);
    *bytes = 2 * n * a->Nz * fixed_row(a->Nx) * sizeof(cd) +
             n * sizeof(dyf_fixed_info) +
             (Wh ? 256 + n * Wh * (kWave / 2) * Ch * sizeof(cd) : 0);
    return DYF_OK;
}

extern "C" int dyf_bragg_fixed_point(const dyf_bragg_args *a,
                                     const double *s_dev, int nsys,
                                     const void *E, const void *incident,
                                     const dyf_fixed_opts *opt, void *work,
                                     size_t work_bytes, void *out,
                                     dyf_fixed_info *info, void *stream_) {
    int rc = check_bragg(a, nsys);
    if (rc) return rc;
    if (!opt || opt->n_iter < 1 || !(opt->tol >= 0.0))
        return fail(DYF_EINVAL, "bad fixed-point options");
    if (!s_dev || !E || !work || !out || !info)
        return fail(DYF_EINVAL, "NULL argument to dyf_bragg_fixed_point");
    size_t need = 0;
    dyf_bragg_fixed_point_bytes(a, nsys, &need);
    if (work_bytes < need)
        return fail(DYF_EINVAL, "workspace of %zu bytes, need %zu", work_bytes,
                    need);
    hipStream_t stream = (hipStream_t)stream_;
    const BraggK k = make_k(a, nsys);
    const size_t n = (size_t)nsys * a->nyc,
                 field = n * a->Nz * fixed_row(a->Nx);
    cd *D0 = (cd *)work, *Dg = D0 + field;
    dyf_fixed_info *idev = (dyf_fixed_info *)(Dg + field);
    const int C = wave_columns(a->Nx);
    int Wh = 0, Ch = 0;
    fixed_halo_shape(a->Nx, (long long)n, &Wh, &Ch);
    if (Wh) {
        auto hf = Wh == 16  ? fixed_halo_fn<16, 1, 2, 3, 4>(Ch)
                  : Wh == 8 ? fixed_halo_fn<8, 1, 2, 3, 4, 5, 6, 7, 8>(Ch)
                  : Wh == 4 ? fixed_halo_fn<4, 1, 2, 3, 4, 5, 6, 7, 8>(Ch)
                            : fixed_halo_fn<2, 1, 2, 3, 4, 5, 6, 7, 8>(Ch);
        if (!hf) return fail(DYF_EINVAL, "no fixed-point kernel for W=%d C=%d",
                             Wh, Ch);
        cd *trash = (cd *)(((uintptr_t)(idev + n) + 255) & ~(uintptr_t)255);
        hipLaunchKernelGGL(hf, dim3((unsigned)n), dim3(kWave * Wh), 0, stream,
                           k, s_dev, (const cd *)E, (const cd *)incident,
                           opt->n_iter, opt->tol, D0, Dg, (cd *)out, idev,
                           trash);
        DYF_LAUNCHED("bragg_fixed_point_halo_kernel");
    } else if (C) {
        auto wf = fixed_wave_fn<1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14,
                                15, 16>(C);
        hipLaunchKernelGGL(wf, dim3((unsigned)n), dim3(kWave), 0, stream, k,
                           s_dev, (const cd *)E, (const cd *)incident,
                           opt->n_iter, opt->tol, D0, Dg, (cd *)out, idev);
        DYF_LAUNCHED("bragg_fixed_point_wave_kernel");
    } else {
        // two rows of at most 2048 columns: within the default 64 KB of LDS
        const size_t lds = (size_t)2 * a->Nx * sizeof(cd);
        hipLaunchKernelGGL(bragg_fixed_point_kernel, dim3((unsigned)n),
                           dim3(kFpThreads), lds, stream, k, s_dev,
                           (const cd *)E, (const cd *)incident, opt->n_iter,
                           opt->tol, D0, Dg, (cd *)out, idev);
        DYF_LAUNCHED("bragg_fixed_point_kernel");
    }
    DYF_HIP(hipMemcpyAsync(info, idev, n * sizeof(dyf_fixed_info),
                           hipMemcpyDeviceToHost, stream));
    DYF_HIP(hipStreamSynchronize(stream));
    return DYF_OK;
}

extern "C" int dyf_bragg_solve(const dyf_bragg_args *a, const double *s_dev,
                               int nsys, const void *E_, const void *incident,
                               const dyf_gmres_opts *opt, const void *x0,
                               void *x_, void *work, size_t work_bytes,
                               dyf_gmres_info *info, void *stream_) {
    int rc = check_bragg(a, nsys);
    if (rc) return rc;
    if (!opt || opt->restart < 1 || opt->maxiter < 1 || !(opt->tol > 0.0))
        return fail(DYF_EINVAL, "bad GMRES options");
    if (!s_dev || !E_ || !incident || !x_ || !work || !info)
        return fail(DYF_EINVAL, "NULL argument to dyf_bragg_solve");
    hipStream_t stream = (hipStream_t)stream_;
    const BraggK k = make_k(a, nsys);
    const size_t N = (size_t)a->nyc * a->Nx;
    const int R = std::min<size_t>((size_t)opt->restart, N);
    Work W = carve(work, nsys, N, R, a->Nx);
    if (W.bytes > work_bytes)
        return fail(DYF_EINVAL, "workspace of %zu bytes, need %zu", work_bytes,
                    W.bytes);
    const cd *E = (const cd *)E_;
    cd *x = (cd *)x_;
    const double eps = 2.220446049250313e-16;
    const size_t Vsys = (size_t)(R + 1) * N;
    const dim3 vgrid((unsigned)((N + 255) / 256), nsys);

    // The per-system GMRES state lives on the device (set-up, inner-loop
    // and end-of-cycle kernels).  The flags the host steers by - per
    // column "some system still iterates" and per cycle "some system
    // continues" - are written by those kernels straight into page-locked
    // host memory (R + 1 words); the host reads a column's word after an
    // event behind it (pipelined one step late, below), a cycle's after a
    // stream synchronisation, and the states once at the end.
    static thread_local int *still_h = nullptr;
    static thread_local int still_cap = 0;
    static thread_local hipEvent_t still_ev[2] = {nullptr, nullptr};
    static thread_local SysState *st_h = nullptr;
    static thread_local int st_cap = 0;
    if (!still_ev[0])
        for (int e = 0; e < 2; ++e)
            DYF_HIP(hipEventCreateWithFlags(&still_ev[e],
                                            hipEventDisableTiming));
    if (still_cap < R + 1) {
        if (still_h) DYF_HIP(hipHostFree(still_h));
        still_h = nullptr, still_cap = 0;
        DYF_HIP(hipHostMalloc((void **)&still_h, sizeof(int) * (R + 1),
                              hipHostMallocDefault));
        still_cap = R + 1;
    }
    if (st_cap < nsys) {
        if (st_h) DYF_HIP(hipHostFree(st_h));
        st_h = nullptr, st_cap = 0;
        DYF_HIP(hipHostMalloc((void **)&st_h, sizeof(SysState) * nsys,
                              hipHostMallocDefault));
        st_cap = nsys;
    }
    W.still = still_h;
    volatile const int *flag = still_h;
    auto norm_into = [&](const cd *v, cd *out, const int *active) -> int {
        return dots(v, N, 0, 1, v, N, N, out, 1, W.part, nsys, active,
                    stream);
    };

    // preconditioner symbol for every system; b = -march(incident, 0);
    // ||b||, ||M b|| and ||r|| (r = b - A x0, or b) into the set-up
    if ((rc = make_symbol(k, s_dev, W.pinv, stream))) return rc;
    rc = launch_march(k, s_dev, E, (const cd *)incident, nullptr, 0, W.b, N,
                      -1.0, nullptr, stream);
    if (rc) return rc;
    if (x0) {
        DYF_HIP(hipMemcpyAsync(x, x0, sizeof(cd) * nsys * N,
                               hipMemcpyDeviceToDevice, stream));
    } else {
        DYF_HIP(hipMemsetAsync(x, 0, sizeof(cd) * nsys * N, stream));
    }
    DYF_HIP(hipMemsetAsync(W.H, 0, sizeof(cd) * nsys * R * (R + 1), stream));
    if ((rc = norm_into(W.b, W.nrm, nullptr))) return rc;
    if ((rc = apply_precon(k, W.pinv, W.b, W.w, stream))) return rc;
    if ((rc = norm_into(W.w, W.nrm0, nullptr))) return rc;
    if (x0) {
        rc = launch_march(k, s_dev, E, nullptr, x, N, W.av, N, 1.0, nullptr,
                          stream);
        if (rc) return rc;
        sub_kernel<<<vgrid, 256, 0, stream>>>(W.b, W.av, W.r, N, nullptr);
        DYF_LAUNCHED("sub_kernel");
        if ((rc = norm_into(W.r, W.h2, nullptr))) return rc;
    } else {
        DYF_HIP(hipMemcpyAsync(W.r, W.b, sizeof(cd) * nsys * N,
                               hipMemcpyDeviceToDevice, stream));
    }
    const unsigned tgrid = (unsigned)((nsys + 63) / 64);
    gmres_setup_kernel<<<tgrid, 64, 0, stream>>>(
        nsys, W.nrm, W.nrm0, x0 ? W.h2 : W.nrm, opt->tol, opt->maxiter, W.st,
        W.active);
    DYF_LAUNCHED("gmres_setup_kernel");

    // The inner loop reads each step's "still iterating" word one step
    // late: the copy of step c lands in a page-locked slot with an event,
    // and the host waits on step c-1's event only after enqueueing step
    // c, so the device never idles on the host.  A step enqueued after
    // every system stopped is masked out by the active flags (only the
    // preconditioner's FFTs run on it).  The first cycle starts without a
    // look at the set-up: if no system needs it, it is masked out the
    // same way and its end-of-cycle word ends the loop.
    const unsigned sgrid = (unsigned)(((size_t)nsys * (R + 1) + 255) / 256);
    const bool fused_step =
        N <= (size_t)kCgsMaxN && R < step_threads(N) &&
        sizeof(double) * (2 * (step_threads(N) / 64) + 8) * (size_t)R <=
            64 * 1024;
    // the sliced step's LDS at the last column within 64 KB: the wave
    // sums of cgs_update_dots_kernel, the partials, column and rotations
    // of arnoldi_finish_scale_kernel
    if (!fused_step &&
        sizeof(double) * std::max<size_t>((kSplitWaves + 1) * 2 * (size_t)R,
                                          2 * (size_t)split_slices(N) +
                                              8 * (size_t)R + 2) >
            64 * 1024)
        return fail(DYF_EINVAL, "restart %d is too large for the Arnoldi "
                    "step of a system of %zu unknowns", R, N);

    for (int cycle = 0; cycle < opt->maxiter; ++cycle) {
        // v0 = M r / ||M r||, S = [||M r||, 0, ...]
        if ((rc = apply_precon(k, W.pinv, W.r, W.w, stream))) return rc;
        if ((rc = norm_into(W.w, W.nrm, W.active))) return rc;
        cycle_start_kernel<<<sgrid, 256, 0, stream>>>(nsys, R, W.nrm, W.inv,
                                                      W.S, W.flags());
        DYF_LAUNCHED("cycle_start_kernel");
        scale_copy_kernel<<<vgrid, 256, 0, stream>>>(W.w, N, W.V, Vsys, W.inv,
                                                     N, W.active);
        DYF_LAUNCHED("scale_copy_kernel");

        for (int col = 0; col < R; ++col) {
            // w = M A v_col
            rc = launch_march(k, s_dev, E, nullptr, W.V + (size_t)col * N,
                              Vsys, W.av, N, 1.0, W.active, stream);
            if (rc) return rc;
            rc = apply_precon(k, W.pinv, W.av, W.w, stream);
            if (rc) return rc;
            // h0 = ||w||; CGS2 against v_0..v_col
            // one workgroup per system while h fits its lanes and its
            // wave sums fit the default 64 KB of LDS at the last column
            if (fused_step) {
                const int T = step_threads(N);
                const size_t lds =
                    sizeof(double) * (2 * (T / 64) + 8) * (size_t)(col + 1);
                step_kernel(N)<<<nsys, T, lds, stream>>>(
                    W.V, Vsys, N, R, col, W.w, W.h1, W.h2, W.H, W.G, W.S,
                    W.st, W.inv, W.active, eps, W.flags());
                DYF_LAUNCHED("arnoldi_step_kernel");
            } else {
                // CGS2 over slices of kSplitLen elements (three passes
                // over the basis), then the finish on the partial sums
                const int kc = col + 1, B = split_slices(N);
                const dim3 sg((unsigned)B, (unsigned)nsys);
                cgs_dots_kernel<<<sg, kSplitThreads,
                                  sizeof(double) * kSplitWaves * 2 * (kc + 1),
                                  stream>>>(W.V, Vsys, N, R, kc, W.w, W.pA, B,
                                            W.active);
                DYF_LAUNCHED("cgs_dots_kernel");
                cgs_update_dots_kernel<<<sg, kSplitThreads,
                                         sizeof(double) * (kSplitWaves + 1) *
                                             2 * kc,
                                         stream>>>(W.V, Vsys, N, R, kc, W.w,
                                                   W.pA, W.pB, W.h1, B,
                                                   W.active);
                DYF_LAUNCHED("cgs_update_dots_kernel");
                cgs_update_norm_kernel<<<sg, kSplitThreads,
                                         sizeof(double) * (2 * kc + kSplitWaves),
                                         stream>>>(W.V, Vsys, N, R, kc, W.w,
                                                   W.pB, W.h2, W.pC, B,
                                                   W.active);
                DYF_LAUNCHED("cgs_update_norm_kernel");
                arnoldi_finish_scale_kernel<<<
                    sg, kSplitThreads,
                    sizeof(double) * (2 * B + 4 * kc + 2 + 4 * col),
                    stream>>>(W.V, Vsys, N, R, col, W.w, W.pA, W.pC, W.h1,
                              W.h2, B, W.H, W.G, W.S, W.st, W.inv, W.active,
                              eps, W.flags());
                DYF_LAUNCHED("arnoldi_finish_scale_kernel");
            }
            DYF_HIP(hipEventRecord(still_ev[col & 1], stream));
            if (col > 0) {
                DYF_HIP(hipEventSynchronize(still_ev[(col - 1) & 1]));
                if (!flag[col - 1]) break;
            }
        }
        // solve the small least-squares problems, update x and form the
        // true residual for the systems not yet done; then one round trip
        restore_active_kernel<<<(nsys + 63) / 64, 64, 0, stream>>>(
            nsys, W.st, W.active);
        DYF_LAUNCHED("restore_active_kernel");
        lsq_kernel<<<(nsys + 63) / 64, 64, 0, stream>>>(nsys, R, W.H, W.S,
                                                        W.Y, W.st, W.active);
        DYF_LAUNCHED("lsq_kernel");
        combine_kernel<<<vgrid, 256, 0, stream>>>(W.V, Vsys, N, R, W.st, W.Y,
                                                  R, 1.0, x, N, N, W.active);
        DYF_LAUNCHED("combine_kernel");
        rc = launch_march(k, s_dev, E, nullptr, x, N, W.av, N, 1.0, W.active,
                          stream);
        if (rc) return rc;
        sub_kernel<<<vgrid, 256, 0, stream>>>(W.b, W.av, W.r, N, W.active);
        DYF_LAUNCHED("sub_kernel");
        if ((rc = norm_into(W.r, W.nrm, W.active))) return rc;
        gmres_cycle_end_kernel<<<tgrid, 64, 0, stream>>>(
            nsys, W.nrm, eps, W.st, W.active, W.flags(), R);
        DYF_LAUNCHED("gmres_cycle_end_kernel");
        // reset the Hessenberg for the next cycle
        DYF_HIP(hipMemsetAsync(W.H, 0, sizeof(cd) * nsys * R * (R + 1), stream));
        DYF_HIP(hipStreamSynchronize(stream));
        if (!flag[R]) break;
    }
    DYF_HIP(hipMemcpyAsync(st_h, W.st, sizeof(SysState) * nsys,
                           hipMemcpyDeviceToHost, stream));
    DYF_HIP(hipStreamSynchronize(stream));
    for (int s = 0; s < nsys; ++s) {
        const SysState &q = st_h[s];
        info[s].residual = q.bnrm2 > 0.0 ? q.rnorm / q.bnrm2 : 0.0;
        info[s].iterations = q.iters;
        info[s].code = q.rnorm <= q.atol ? 0 : opt->maxiter;
    }
    return DYF_OK;
}
