// The Bragg preconditioner w = ifft(fft(v) / T22s) along x, per row, as one
// fused kernel (bragg/solver.py:167-168, 272-275, 297-300).
//
// A workgroup of 512 threads takes R rows into LDS, transforms them,
// multiplies by the row's system spectrum, transforms back and stores:
// one read and one write of each row in HBM, one launch.
//
// Transforms are mixed-radix Stockham passes (radices 8, 4, 2, 3, 5, 7, 11,
// 13) in place in LDS.  A pass reads every butterfly's inputs into
// registers, applies the twiddles and the radix-r DFT there, and after a
// barrier writes the outputs to their autosorted places:
//     inputs   j + q N/r                           q = 0 .. r-1
//     twiddle  exp(-+2 pi i q (j mod Ns) / (Ns r))  (from one table of
//              exp(-2 pi i t / N), t < N, built per workgroup in LDS)
//     outputs  (j div Ns) Ns r + (j mod Ns) + q Ns,   Ns = product of the
//              radices of the earlier passes
// (the Stockham iteration as in Govindaraju et al., "High performance
// discrete Fourier transforms on graphics processors", SC 2008).
//
// Row lengths Nx whose prime factors are all <= 13 are transformed at
// length Nx (direct plan: the spectrum is 1 / (T22s Nx)).  Other Nx use
// the circular convolution the preconditioner is: w = v (*) k with
// k = ifft(1 / T22s), computed exactly as a linear convolution at a
// 13-smooth length N >= 2 Nx - 1 (convolution plan):
//     kk[d + Nx - 1] = k[d mod Nx]   for -(Nx-1) <= d <= Nx-1,  0 beyond
//     w[n] = ifft_N(fft_N(v zero-padded) fft_N(kk))[n + Nx - 1]
// since every index n - j + Nx - 1 (0 <= j, n < Nx) lies in [0, 2Nx - 2]
// no wrap occurs.  fft_N(kk) / N is formed once per system.
#include "internal.hpp"

namespace dyf {

using cd = cx<double>;

namespace {

__device__ __forceinline__ cd c_add(cd a, cd b) { return {a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ cd c_sub(cd a, cd b) { return {a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ cd c_mul(cd a, cd b) {
    return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
// a * (-i) for the forward transform, a * (+i) for the inverse
template <bool INV>
__device__ __forceinline__ cd c_rot(cd a) {
    return INV ? cd{-a.i, a.r} : cd{a.i, -a.r};
}

constexpr int kThreads = 512;

// LDS pointers: the passes are separate functions, and a generic pointer
// there would become flat accesses (slower, and waiting on the global
// memory counter too)
typedef __attribute__((address_space(3))) double lds_d;

__device__ __forceinline__ cd ld(const lds_d *p, int i) {
    return {p[2 * i], p[2 * i + 1]};
}
__device__ __forceinline__ void st(lds_d *p, int i, cd v) {
    p[2 * i] = v.r;
    p[2 * i + 1] = v.i;
}

// Row data in LDS is padded by one element every 2^kPadShift: the passes
// write with strides of RAD * 16 bytes, which would otherwise put many
// lanes of a wave on the same LDS banks
#ifndef FFT_PAD_SHIFT
#define FFT_PAD_SHIFT 4
#endif
constexpr int kPadShift = FFT_PAD_SHIFT;
__device__ __forceinline__ int pad(int i) { return i + (i >> kPadShift); }
constexpr int kMaxLen = 4096;          // R * N <= kMaxLen per workgroup

// cos and sin of 2 pi m / RAD for m = 1 .. (RAD-1)/2 (odd radices),
// decimal expansions of the exact values rounded to double.  As
// compile-time literals they live in scalar registers.
template <int RAD> struct Roots;
template <> struct Roots<3> {
    static constexpr double c[1] = {-0.5};
    static constexpr double s[1] = {0.8660254037844386};
};
template <> struct Roots<5> {
    static constexpr double c[2] = {0.30901699437494745, -0.8090169943749475};
    static constexpr double s[2] = {0.9510565162951535, 0.5877852522924731};
};
template <> struct Roots<7> {
    static constexpr double c[3] = {0.6234898018587335, -0.2225209339563144,
                                    -0.9009688679024191};
    static constexpr double s[3] = {0.7818314824680298, 0.9749279121818236,
                                    0.4338837391175581};
};
template <> struct Roots<11> {
    static constexpr double c[5] = {0.8412535328311812, 0.41541501300188644,
                                    -0.14231483827328514, -0.6548607339452851,
                                    -0.9594929736144974};
    static constexpr double s[5] = {0.5406408174555976, 0.9096319953545183,
                                    0.9898214418809327, 0.7557495743542583,
                                    0.28173255684142967};
};
template <> struct Roots<13> {
    static constexpr double c[6] = {0.8854560256532099, 0.5680647467311558,
                                    0.12053668025532305, -0.3546048870425356,
                                    -0.7485107481711011, -0.970941817426052};
    static constexpr double s[6] = {0.46472317204376856, 0.8229838658936564,
                                    0.992708874098054, 0.9350162426854148,
                                    0.6631226582407952, 0.23931566428755777};
};
constexpr double kSqrtHalf = 0.7071067811865476;

template <bool INV>
__device__ __forceinline__ void dft4(cd &a0, cd &a1, cd &a2, cd &a3) {
    const cd t0 = c_add(a0, a2), t1 = c_sub(a0, a2);
    const cd t2 = c_add(a1, a3), t3 = c_rot<INV>(c_sub(a1, a3));
    a0 = c_add(t0, t2);
    a2 = c_sub(t0, t2);
    a1 = c_add(t1, t3);
    a3 = c_sub(t1, t3);
}

// DFT of length RAD in registers: forward exp(-2 pi i q k / RAD), inverse
// the conjugate.
template <int RAD, bool INV>
__device__ __forceinline__ void dft(cd (&a)[RAD]) {
    if constexpr (RAD == 2) {
        const cd t = a[1];
        a[1] = c_sub(a[0], t);
        a[0] = c_add(a[0], t);
    } else if constexpr (RAD == 4) {
        dft4<INV>(a[0], a[1], a[2], a[3]);
    } else if constexpr (RAD == 8) {
        // length-4 DFTs of the even and odd samples, then one radix-2
        // combination with the eighth roots of unity
        cd e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
        cd o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
        dft4<INV>(e0, e1, e2, e3);
        dft4<INV>(o0, o1, o2, o3);
        const double h = kSqrtHalf;
        // o1 w8, o2 w8^2 = o2 (-+i), o3 w8^3
        const cd w1 = INV ? cd{h * (o1.r - o1.i), h * (o1.r + o1.i)}
                          : cd{h * (o1.r + o1.i), h * (o1.i - o1.r)};
        const cd w2 = c_rot<INV>(o2);
        const cd w3 = INV ? cd{-h * (o3.r + o3.i), h * (o3.r - o3.i)}
                          : cd{h * (o3.i - o3.r), -h * (o3.r + o3.i)};
        a[0] = c_add(e0, o0), a[4] = c_sub(e0, o0);
        a[1] = c_add(e1, w1), a[5] = c_sub(e1, w1);
        a[2] = c_add(e2, w2), a[6] = c_sub(e2, w2);
        a[3] = c_add(e3, w3), a[7] = c_sub(e3, w3);
    } else {
        // odd prime: with s_q = a_q + a_{RAD-q}, d_q = a_q - a_{RAD-q},
        //   y_k      = a_0 + sum_q s_q cos(2 pi qk/RAD) -+ i d_q sin(...)
        //   y_{RAD-k} = the same with the sine term's sign flipped
        constexpr int h = (RAD - 1) / 2;
#pragma unroll
        for (int q = 1; q <= h; ++q) {
            const cd sq = c_add(a[q], a[RAD - q]);
            const cd dq = c_sub(a[q], a[RAD - q]);
            a[q] = sq;
            a[RAD - q] = dq;
        }
        const cd a0 = a[0];
        cd re[h], im[h];
#pragma unroll
        for (int k = 1; k <= h; ++k) {
            re[k - 1] = a0;
            im[k - 1] = {0.0, 0.0};
        }
        cd y0 = a0;
#pragma unroll
        for (int q = 1; q <= h; ++q) {
            const cd sq = a[q], dq = a[RAD - q];
            y0 = c_add(y0, sq);
#pragma unroll
            for (int k = 1; k <= h; ++k) {
                const int m = (q * k) % RAD;
                const bool lo = m <= h;
                const double c = Roots<RAD>::c[(lo ? m : RAD - m) - 1];
                const double sn = lo ? Roots<RAD>::s[m - 1]
                                     : -Roots<RAD>::s[RAD - m - 1];
                re[k - 1].r = fma(sq.r, c, re[k - 1].r);
                re[k - 1].i = fma(sq.i, c, re[k - 1].i);
                im[k - 1].r = fma(dq.r, sn, im[k - 1].r);
                im[k - 1].i = fma(dq.i, sn, im[k - 1].i);
            }
        }
        a[0] = y0;
        // forward: a_q w^{qk} + a_{RAD-q} w^{-qk} = s_q cos - i d_q sin
#pragma unroll
        for (int k = 1; k <= h; ++k) {
            const cd r = re[k - 1], m = im[k - 1];
            const cd yk = INV ? cd{r.r - m.i, r.i + m.r}
                              : cd{r.r + m.i, r.i - m.r};
            const cd yn = INV ? cd{r.r + m.i, r.i - m.r}
                              : cd{r.r - m.i, r.i + m.r};
            a[k] = yk;
            a[RAD - k] = yn;
        }
    }
}

// q = a / b, r = a - q b for 0 <= a < 2^24 and a wave-uniform b >= 1,
// through a single-precision reciprocal and one correction each way
__device__ __forceinline__ void divmod(int a, int b, float rb, int &q, int &r) {
    q = (int)((float)a * rb);
    r = a - q * b;
    if (r < 0) { --q; r += b; }
    if (r >= b) { ++q; r -= b; }
}

// One Stockham pass of radix RAD over the R rows of length N in `buf`.
template <int RAD, bool INV>
__device__ __noinline__ void pass(lds_d *__restrict__ buf,
                                  const lds_d *__restrict__ tw, int N, int R,
                                  int Ns) {
    constexpr int T = (kMaxLen / RAD + kThreads - 1) / kThreads;
    const int nb = N / RAD;
    const int total = R * nb;
    const int tid = threadIdx.x;
    const int tw_step = N / (Ns * RAD);
    const float rnb = __builtin_amdgcn_rcpf((float)nb);
    const float rns = __builtin_amdgcn_rcpf((float)Ns);
    cd v[T][RAD];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int b = tid + t * kThreads;
        if (b < total) {
            int r, j, jq, jm;
            divmod(b, nb, rnb, r, j);
            divmod(j, Ns, rns, jq, jm);
            const int base = r * N + j;
#pragma unroll
            for (int q = 0; q < RAD; ++q) v[t][q] = ld(buf, pad(base + q * nb));
            // twiddles w^q, w = exp(-+2 pi i jm / (Ns RAD)), by products
            // of w (q - 1 roundings, q < 16)
            if (jm != 0) {
                cd w = ld(tw, tw_step * jm);
                if (INV) w.i = -w.i;
                cd wq = w;
#pragma unroll
                for (int q = 1; q < RAD; ++q) {
                    v[t][q] = c_mul(v[t][q], wq);
                    if (q + 1 < RAD) wq = c_mul(wq, w);
                }
            }
            dft<RAD, INV>(v[t]);
        }
    }
    lds_barrier();
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int b = tid + t * kThreads;
        if (b < total) {
            int r, j, jq, jm;
            divmod(b, nb, rnb, r, j);
            divmod(j, Ns, rns, jq, jm);
            const int dst = r * N + jq * Ns * RAD + jm;
#pragma unroll
            for (int q = 0; q < RAD; ++q) st(buf, pad(dst + q * Ns), v[t][q]);
        }
    }
    lds_barrier();
}

template <bool INV>
__device__ void transform(lds_d *buf, const lds_d *tw, const FftPlan &p,
                          int R) {
    int Ns = 1;
    for (int s = 0; s < p.nst; ++s) {
        const int rad = p.radix[s];
        switch (rad) {
        case 8: pass<8, INV>(buf, tw, p.N, R, Ns); break;
        case 4: pass<4, INV>(buf, tw, p.N, R, Ns); break;
        case 2: pass<2, INV>(buf, tw, p.N, R, Ns); break;
        case 3: pass<3, INV>(buf, tw, p.N, R, Ns); break;
        case 5: pass<5, INV>(buf, tw, p.N, R, Ns); break;
        case 7: pass<7, INV>(buf, tw, p.N, R, Ns); break;
        case 11: pass<11, INV>(buf, tw, p.N, R, Ns); break;
        default: pass<13, INV>(buf, tw, p.N, R, Ns); break;
        }
        Ns *= rad;
    }
}

// What one launch does to each of its rows.
struct FftJob {
    FftPlan p;
    int R;              // rows per workgroup
    int nrows;          // rows in the launch
    int len_in;         // elements per row in `in` (zero-padded up to N)
    int len_out;        // elements per row written to `out`
    int out_off;        // first transform index written
    int rows_per_sys;   // rows sharing one spectrum row
    int round_trip;     // 1: forward, times spectrum, inverse; 0: forward
                        // times `scale`
    double scale;
};

// At most 128 VGPRs (four waves per SIMD): two workgroups per CU, so one
// workgroup's barriers and memory waits overlap the other's passes
// (measured on the 20 um Bragg rows, 1312 x 520: 28.4 -> 23.0 us; R = 3
// rows per workgroup was the best of 1 to 6)
__global__ __launch_bounds__(kThreads, 4) void fft_rows_kernel(
    FftJob job, const cd *__restrict__ in, cd *__restrict__ out,
    const cd *__restrict__ spec) {
    extern __shared__ cd lds[];
    const FftPlan &p = job.p;
    const int N = p.N, R = job.R, tid = threadIdx.x;
    cd *tw = lds;                 // exp(-2 pi i t / N), t < N
    cd *buf = lds + N;            // R rows of N
    const int row0 = blockIdx.x * R;
    for (int t = tid; t < N; t += kThreads) {
        double s, c;
        sincospi(-2.0 * (double)t / (double)N, &s, &c);
        tw[t] = {c, s};
    }
    for (int idx = tid; idx < R * N; idx += kThreads) {
        const int r = idx / N, n = idx - r * N;
        const int row = row0 + r;
        cd v = {0.0, 0.0};
        if (row < job.nrows && n < job.len_in)
            v = in[(size_t)row * job.len_in + n];
        buf[pad(idx)] = v;
    }
    lds_barrier();
    lds_d *lbuf = (lds_d *)(double *)buf;
    const lds_d *ltw = (const lds_d *)(const double *)tw;
    transform<false>(lbuf, ltw, p, R);
    if (job.round_trip) {
        for (int idx = tid; idx < R * N; idx += kThreads) {
            const int r = idx / N, m = idx - r * N;
            const int row = min(row0 + r, job.nrows - 1);
            const cd f = spec[(size_t)(row / job.rows_per_sys) * N + m];
            buf[pad(idx)] = c_mul(buf[pad(idx)], f);
        }
        lds_barrier();
        transform<true>(lbuf, ltw, p, R);
    } else {
        for (int idx = tid; idx < R * N; idx += kThreads)
            buf[pad(idx)] = {buf[pad(idx)].r * job.scale,
                             buf[pad(idx)].i * job.scale};
        lds_barrier();
    }
    for (int idx = tid; idx < R * job.len_out; idx += kThreads) {
        const int r = idx / job.len_out, n = idx - r * job.len_out;
        const int row = row0 + r;
        if (row < job.nrows)
            out[(size_t)row * job.len_out + n] =
                buf[pad(r * N + n + job.out_off)];
    }
}

// kk rows of the convolution plan: k[n] = sum_m pinv[m] exp(+2 pi i m n /
// Nx) (pinv = 1 / (T22s Nx), so k = ifft(1 / T22s)), placed at n + Nx - 1
// and, for n >= 1, at n - 1; zero from 2 Nx - 1 to N.  One block per
// (system, 64 outputs): lane l of every wave takes output n0 + l, the 8
// waves take interleaved eighths of the sum, reduced in LDS in a fixed
// order.  Twiddles exp(2 pi i t / Nx) and the system's pinv row in LDS.
constexpr int kConvWaves = kThreads / kWave;

__global__ __launch_bounds__(kThreads) void conv_kernel_rows_kernel(
    const cd *__restrict__ pinv, int Nx, int N, cd *__restrict__ kk) {
    extern __shared__ cd lds[];
    cd *tw = lds, *sp = lds + Nx, *part = lds + 2 * Nx;
    const int s = blockIdx.y, tid = threadIdx.x;
    const int lane = tid % kWave, wave = tid / kWave;
    for (int t = tid; t < Nx; t += kThreads) {
        double sn, cs;
        sincospi(2.0 * (double)t / (double)Nx, &sn, &cs);
        tw[t] = {cs, sn};
        sp[t] = pinv[(size_t)s * Nx + t];
    }
    __syncthreads();
    const int n = blockIdx.x * kWave + lane;
    double ar = 0.0, ai = 0.0;
    if (n < Nx) {
        // m = wave, wave + 8, ...: index (m n) mod Nx advances by 8 n
        const int stride = (int)(((long long)kConvWaves * n) % Nx);
        int idx = (int)(((long long)wave * n) % Nx);
        for (int m = wave; m < Nx; m += kConvWaves) {
            const cd a = sp[m], w = tw[idx];
            ar = fma(a.r, w.r, fma(-a.i, w.i, ar));
            ai = fma(a.r, w.i, fma(a.i, w.r, ai));
            idx += stride;
            if (idx >= Nx) idx -= Nx;
        }
    }
    part[wave * kWave + lane] = {ar, ai};
    __syncthreads();
    cd *row = kk + (size_t)s * N;
    if (wave == 0 && n < Nx) {
        cd k = part[lane];
#pragma unroll
        for (int w = 1; w < kConvWaves; ++w)
            k = c_add(k, part[w * kWave + lane]);
        row[n + Nx - 1] = k;
        if (n >= 1) row[n - 1] = k;
    }
    for (int j = 2 * Nx - 1 + blockIdx.x * kThreads + tid; j < N;
         j += gridDim.x * kThreads)
        row[j] = {0.0, 0.0};
}

constexpr int kRadices[] = {8, 4, 2, 3, 5, 7, 11, 13};

bool factor(int n, FftPlan *p) {
    p->N = n;
    p->nst = 0;
    for (int r : kRadices) {
        while (n % r == 0) {
            if (p->nst >= kFftMaxPasses) return false;
            p->radix[p->nst++] = r;
            n /= r;
        }
    }
    return n == 1;
}

int rows_per_block(int N, int nrows) {
    // enough workgroups for two per CU where the batch allows, at most
    // kMaxLen elements per workgroup
    const int cap = std::max(1, kMaxLen / N);
    const int want = (nrows + 511) / 512;
    return std::max(1, std::min(cap, want));
}

size_t lds_bytes(int N, int R) {
    const size_t rows = (size_t)N * R;
    return sizeof(cd) * (N + rows + (rows >> kPadShift) + 1);
}

int launch_rows(const FftJob &job, const cd *in, cd *out, const cd *spec,
                hipStream_t stream) {
    const size_t lds = lds_bytes(job.p.N, job.R);
    DYF_HIP(hipFuncSetAttribute((const void *)fft_rows_kernel,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
    const int blocks = (job.nrows + job.R - 1) / job.R;
    fft_rows_kernel<<<blocks, kThreads, lds, stream>>>(job, in, out, spec);
    DYF_LAUNCHED("fft_rows_kernel");
    return DYF_OK;
}

}  // namespace

int fft_plan_for(int Nx, FftPlan *p) {
    if (Nx < 1 || Nx > kMaxLen / 2)
        return fail(DYF_EINVAL, "preconditioner rows of %d points: the "
                    "transform supports 1 .. %d", Nx, kMaxLen / 2);
    if (factor(Nx, p)) {
        p->conv = 0;
        return DYF_OK;
    }
    for (int n = 2 * Nx - 1; n <= kMaxLen; ++n)
        if (factor(n, p)) {
            p->conv = 1;
            return DYF_OK;
        }
    return fail(DYF_EINVAL, "no transform length for %d points", Nx);
}

size_t fft_symbol_elems(const FftPlan &p, int Nx) {
    return (size_t)Nx + (p.conv ? (size_t)p.N : 0);
}

int fft_prepare_spectrum(const FftPlan &p, int Nx, cd *symbol, int nsys,
                         hipStream_t stream) {
    if (!p.conv) return DYF_OK;
    cd *kk = symbol + (size_t)nsys * Nx;
    const size_t lds = sizeof(cd) * (2 * (size_t)Nx + kThreads);
    DYF_HIP(hipFuncSetAttribute((const void *)conv_kernel_rows_kernel,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
    conv_kernel_rows_kernel<<<dim3((Nx + kWave - 1) / kWave, nsys), kThreads,
                              lds, stream>>>(symbol, Nx, p.N, kk);
    DYF_LAUNCHED("conv_kernel_rows_kernel");
    FftJob job;
    job.p = p;
    job.nrows = nsys;
    job.R = rows_per_block(p.N, nsys);
    job.len_in = job.len_out = p.N;
    job.out_off = 0;
    job.rows_per_sys = 1;
    job.round_trip = 0;
    job.scale = 1.0 / (double)p.N;
    return launch_rows(job, kk, kk, nullptr, stream);
}

int fft_precondition(const FftPlan &p, int Nx, const cd *symbol, int nsys,
                     int rows_per_sys, const cd *v, cd *w,
                     hipStream_t stream) {
    FftJob job;
    job.p = p;
    job.nrows = nsys * rows_per_sys;
    job.R = rows_per_block(p.N, job.nrows);
    job.len_in = job.len_out = Nx;
    job.out_off = p.conv ? Nx - 1 : 0;
    job.rows_per_sys = rows_per_sys;
    job.round_trip = 1;
    job.scale = 1.0;
    const cd *spec = p.conv ? symbol + (size_t)nsys * Nx : symbol;
    return launch_rows(job, v, w, spec, stream);
}

}  // namespace dyf
