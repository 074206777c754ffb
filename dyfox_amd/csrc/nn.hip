// Kernels of the learned warm start: a small conv encoder-decoder trained
// and run on gfx950 with BF16 operands on MFMA and FP32 accumulation.
//
// Activations are NHWC; a layer reads a channel slice [off, off+C) of a
// buffer whose pixel stride is its total channel count, so skip
// concatenations cost nothing.  A 3x3 convolution is im2col (BF16) plus
// one GEMM; its two backward products are the same GEMM kernel with
// different operand strides:
//     forward   Y[p][co]   = sum_k Xcol[p][k] W[co][k]
//     data      dXcol[p][k] = sum_co dZ[p][co] W[co][k]
//     weights   dW[co][k]  = sum_p dZ[p][co] Xcol[p][k]   (split over p)
// followed by col2im for the data gradient.  No counterpart in the
// reference.
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace dyf {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));
typedef unsigned short bf16;

__device__ __forceinline__ float bf2f(bf16 v) {
    return __uint_as_float((unsigned)v << 16);
}
__device__ __forceinline__ bf16 f2bf(float f) {   // round to nearest even
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (bf16)(u >> 16);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16)(u >> 16);
}

// ---- GEMM: C(i,j) = sum_k A(i,k) B(j,k), BF16 in, FP32 accumulate ------
// A(i,k) = A[i*a_i + k*a_k], B(j,k) = B[j*b_j + k*b_k].  A 256-thread
// workgroup computes a BM x BN tile; its four waves are laid out WR x
// (4/WR), each owning (BM/WR) x (BN*WR/4) as 16x16 tiles of gfx950's
// v_mfma_f32_16x16x32_bf16 (K = 32 per instruction; lane l supplies
// row l%16, k = 8(l/16) .. 8(l/16)+7 of A and likewise of B, and holds
// column l%16, rows 4(l/16) .. 4(l/16)+3 of the result).  BN follows N
// (16, 32 or 64 columns), so the narrow layers of the network do not
// pay for a 64-wide tile.  K steps of 32 are staged through LDS; each
// operand is loaded from global memory according to its layout:
//   k-contiguous (stride 1 along k): 16-byte loads of 8 k values,
//   row-contiguous (stride 1 along i or j): 16-byte loads of 8 rows at
//     one k, written to LDS transposed,
//   anything else (or misaligned): one element per load.
// grid.z splits K.  Epilogue: 0 = BF16 store of relu?(acc + bias) into
// C(i,j) = C[i*c_i + j], 1 = FP32 store (+bias), 2 = FP32 atomic add.
struct GemmArgs {
    const bf16 *A;
    long long a_i, a_k;
    const bf16 *B;
    long long b_j, b_k;
    int M, N, K, k_chunk;
    int mode, relu;
    const float *bias;
    void *C;
    long long c_i;
};

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int TK = 32, LDK = TK + 8;
enum { LOAD_ANY = 0, LOAD_K = 1, LOAD_ROW = 2 };

// Stage rows [r0, r0 + R) x k [k0, k0 + TK) of X(r, k) = X[r*s_r + k*s_k]
// into S[r - r0][k - k0] (zero outside r < rows, k < ke).
template <int R, int MODE>
__device__ __forceinline__ void stage(bf16 *__restrict__ S,
                                      const bf16 *__restrict__ X,
                                      long long s_r, long long s_k, int r0,
                                      int rows, int k0, int ke, int tid) {
    if constexpr (MODE == LOAD_K) {
        // R x TK / 8 vectors, 4 per row
        for (int v = tid; v < R * (TK / 8); v += 256) {
            const int r = v >> 2, kq = (v & 3) * 8;
            const int gr = r0 + r, gk = k0 + kq;
            uint4 q = {0u, 0u, 0u, 0u};
            if (gr < rows && gk + 7 < ke) {
                q = *(const uint4 *)(X + (long long)gr * s_r + gk);
            } else if (gr < rows) {
                unsigned short e[8];
#pragma unroll
                for (int t = 0; t < 8; ++t)
                    e[t] = gk + t < ke ? X[(long long)gr * s_r + gk + t] : 0;
                q = {e[0] | (unsigned)e[1] << 16, e[2] | (unsigned)e[3] << 16,
                     e[4] | (unsigned)e[5] << 16, e[6] | (unsigned)e[7] << 16};
            }
            *(uint4 *)(S + r * LDK + kq) = q;
        }
    } else if constexpr (MODE == LOAD_ROW) {
        // (R / 8) row groups x TK k values
        for (int v = tid; v < (R / 8) * TK; v += 256) {
            const int k = v % TK, rq = (v / TK) * 8;
            const int gr = r0 + rq, gk = k0 + k;
            unsigned short e[8];
            if (gk < ke && gr + 7 < rows) {
                const uint4 q = *(const uint4 *)(X + gr + (long long)gk * s_k);
                e[0] = q.x & 0xffff, e[1] = q.x >> 16, e[2] = q.y & 0xffff,
                e[3] = q.y >> 16, e[4] = q.z & 0xffff, e[5] = q.z >> 16,
                e[6] = q.w & 0xffff, e[7] = q.w >> 16;
            } else {
#pragma unroll
                for (int t = 0; t < 8; ++t)
                    e[t] = (gk < ke && gr + t < rows)
                               ? X[gr + t + (long long)gk * s_k] : 0;
            }
#pragma unroll
            for (int t = 0; t < 8; ++t) S[(rq + t) * LDK + k] = e[t];
        }
    } else {
        for (int v = tid; v < R * TK; v += 256) {
            const int r = v / TK, k = v % TK;
            const int gr = r0 + r, gk = k0 + k;
            S[r * LDK + k] = (gr < rows && gk < ke)
                                 ? X[(long long)gr * s_r + (long long)gk * s_k]
                                 : (bf16)0;
        }
    }
}

template <int BM, int BN, int WR, int LA, int LB>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs g) {
    constexpr int WC = 4 / WR;
    constexpr int RT = BM / WR / 16, CT = BN / WC / 16;
    static_assert(RT >= 1 && CT >= 1, "tile too small for the wave grid");
    __shared__ __align__(16) bf16 As[BM * LDK];
    __shared__ __align__(16) bf16 Bs[BN * LDK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i0 = blockIdx.x * BM, j0 = blockIdx.y * BN;
    const int kb = blockIdx.z * g.k_chunk;
    const int ke = min(g.K, kb + g.k_chunk);
    const int wr = (wave / WC) * (BM / WR), wc = (wave % WC) * (BN / WC);
    float4_t acc[RT][CT];
#pragma unroll
    for (int a = 0; a < RT; ++a)
#pragma unroll
        for (int b = 0; b < CT; ++b) acc[a][b] = float4_t{0.f, 0.f, 0.f, 0.f};
    const int ko = 8 * (lane >> 4), lr = lane & 15;
    for (int k0 = kb; k0 < ke; k0 += TK) {
        stage<BM, LA>(As, g.A, g.a_i, g.a_k, i0, g.M, k0, ke, tid);
        stage<BN, LB>(Bs, g.B, g.b_j, g.b_k, j0, g.N, k0, ke, tid);
        __syncthreads();
        bf16x8 af[RT], bfr[CT];
#pragma unroll
        for (int a = 0; a < RT; ++a)
            af[a] = *(const bf16x8 *)&As[(wr + 16 * a + lr) * LDK + ko];
#pragma unroll
        for (int b = 0; b < CT; ++b)
            bfr[b] = *(const bf16x8 *)&Bs[(wc + 16 * b + lr) * LDK + ko];
#pragma unroll
        for (int a = 0; a < RT; ++a)
#pragma unroll
            for (int b = 0; b < CT; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    af[a], bfr[b], acc[a][b], 0, 0, 0);
        __syncthreads();
    }
    // lane holds column (lane & 15), rows 4*(lane >> 4) + t of each tile
#pragma unroll
    for (int a = 0; a < RT; ++a)
#pragma unroll
        for (int b = 0; b < CT; ++b)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int i = i0 + wr + 16 * a + 4 * (lane >> 4) + t;
                const int j = j0 + wc + 16 * b + (lane & 15);
                if (i >= g.M || j >= g.N) continue;
                float v = acc[a][b][t];
                if (g.mode == 2) {
                    atomicAdd(&((float *)g.C)[(long long)i * g.c_i + j], v);
                    continue;
                }
                if (g.bias) v += g.bias[j];
                if (g.relu) v = fmaxf(v, 0.f);
                if (g.mode == 0)
                    ((bf16 *)g.C)[(long long)i * g.c_i + j] = f2bf(v);
                else
                    ((float *)g.C)[(long long)i * g.c_i + j] = v;
            }
}

// ---- im2col / col2im for 3x3, stride 1, zero padding -------------------
// X: (B, H, W, Ctot) BF16, channels [off, off+C).  Xcol: (B*H*W, 9*C) with
// column (ky*3 + kx)*C + c.
__global__ void im2col3_kernel(const bf16 *__restrict__ X, int Bn, int H,
                               int W, int Ctot, int off, int C,
                               bf16 *__restrict__ col) {
    const long long n = (long long)Bn * H * W * 9 * C;
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n;
         t += (long long)gridDim.x * 256) {
        const int c = (int)(t % C);
        const long long q = t / C;
        const int tap = (int)(q % 9);
        const long long p = q / 9;
        const int x = (int)(p % W), y = (int)((p / W) % H);
        const long long b = p / ((long long)W * H);
        const int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
        bf16 v = 0;
        if (yy >= 0 && yy < H && xx >= 0 && xx < W)
            v = X[((b * H + yy) * W + xx) * Ctot + off + c];
        col[t] = v;
    }
}

// dX[b,y,x,off+c] += sum over taps of dcol[p'][tap*C + c], p' the pixel
// whose tap lands on (y, x).
__global__ void col2im3_kernel(const float *__restrict__ dcol, int Bn, int H,
                               int W, int C, float *__restrict__ dX, int Ctot,
                               int off) {
    const long long n = (long long)Bn * H * W * C;
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n;
         t += (long long)gridDim.x * 256) {
        const int c = (int)(t % C);
        const long long p = t / C;
        const int x = (int)(p % W), y = (int)((p / W) % H);
        const long long b = p / ((long long)W * H);
        float s = 0.f;
        for (int tap = 0; tap < 9; ++tap) {
            const int ys = y - (tap / 3 - 1), xs = x - (tap % 3 - 1);
            if (ys < 0 || ys >= H || xs < 0 || xs >= W) continue;
            s += dcol[((b * H + ys) * W + xs) * 9 * C + tap * C + c];
        }
        dX[p * Ctot + off + c] += s;
    }
}

// channel slice copy BF16 -> BF16 (1x1 conv input view)
__global__ void slice_kernel(const bf16 *__restrict__ X, long long P, int Ctot,
                             int off, int C, bf16 *__restrict__ out) {
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < P * C;
         t += (long long)gridDim.x * 256)
        out[t] = X[(t / C) * Ctot + off + t % C];
}

// dZ = dY * (Y > 0) (or dY when !relu), FP32 -> BF16; dY is a channel
// slice of a gradient buffer, Y the matching slice of the activation.
__global__ void relu_grad_kernel(const float *__restrict__ dY,
                                 const bf16 *__restrict__ Y, long long P,
                                 int Ctot, int off, int C, int relu,
                                 bf16 *__restrict__ dZ) {
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < P * C;
         t += (long long)gridDim.x * 256) {
        const long long at = (t / C) * Ctot + off + t % C;
        float g = dY[at];
        if (relu && !(bf2f(Y[at]) > 0.f)) g = 0.f;
        dZ[t] = f2bf(g);
    }
}

// db[c] += sum_p dZ[p][c]  (one block per channel)
__global__ void bias_grad_kernel(const bf16 *__restrict__ dZ, long long P,
                                 int C, float *__restrict__ db) {
    __shared__ float red[256];
    const int c = blockIdx.x;
    float s = 0.f;
    for (long long p = threadIdx.x; p < P; p += 256) s += bf2f(dZ[p * C + c]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] += red[0];
}

// 2x2 average pool, BF16 slice -> BF16 buffer (B, H/2, W/2, C)
__global__ void pool_fwd_kernel(const bf16 *__restrict__ X, int Bn, int H,
                                int W, int Ctot, int off, int C,
                                bf16 *__restrict__ Y) {
    const int Ho = H / 2, Wo = W / 2;
    const long long n = (long long)Bn * Ho * Wo * C;
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n;
         t += (long long)gridDim.x * 256) {
        const int c = (int)(t % C);
        const long long p = t / C;
        const int x = (int)(p % Wo), y = (int)((p / Wo) % Ho);
        const long long b = p / ((long long)Wo * Ho);
        float s = 0.f;
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx)
                s += bf2f(X[((b * H + 2 * y + dy) * W + 2 * x + dx) * Ctot + off + c]);
        Y[t] = f2bf(0.25f * s);
    }
}

__global__ void pool_bwd_kernel(const float *__restrict__ dY, int Bn, int H,
                                int W, int C, float *__restrict__ dX, int Ctot,
                                int off) {
    const long long n = (long long)Bn * H * W * C;
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n;
         t += (long long)gridDim.x * 256) {
        const int c = (int)(t % C);
        const long long p = t / C;
        const int x = (int)(p % W), y = (int)((p / W) % H);
        const long long b = p / ((long long)W * H);
        const int Ho = H / 2, Wo = W / 2;
        dX[p * Ctot + off + c] +=
            0.25f * dY[((b * Ho + y / 2) * Wo + x / 2) * C + c];
    }
}

// nearest x2 upsample of (B, H, W, C) into a channel slice of (B, 2H, 2W, Ctot)
__global__ void up_fwd_kernel(const bf16 *__restrict__ X, int Bn, int H, int W,
                              int C, bf16 *__restrict__ Y, int Ctot, int off) {
    const long long n = (long long)Bn * 4 * H * W * C;
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n;
         t += (long long)gridDim.x * 256) {
        const int c = (int)(t % C);
        const long long p = t / C;
        const int x = (int)(p % (2 * W)), y = (int)((p / (2 * W)) % (2 * H));
        const long long b = p / (4LL * W * H);
        Y[p * Ctot + off + c] = X[((b * H + y / 2) * W + x / 2) * C + c];
    }
}

__global__ void up_bwd_kernel(const float *__restrict__ dY, int Bn, int H,
                              int W, int C, int Ctot, int off,
                              float *__restrict__ dX) {
    const long long n = (long long)Bn * H * W * C;
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n;
         t += (long long)gridDim.x * 256) {
        const int c = (int)(t % C);
        const long long p = t / C;
        const int x = (int)(p % W), y = (int)((p / W) % H);
        const long long b = p / ((long long)W * H);
        float s = 0.f;
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx)
                s += dY[((b * 2 * H + 2 * y + dy) * 2 * W + 2 * x + dx) * Ctot + off + c];
        dX[p * C + c] += s;
    }
}

// loss = mean (out - target)^2 over n values; grad = 2 (out - target) / n
__global__ void mse_kernel(const float *__restrict__ out,
                           const float *__restrict__ target, long long n,
                           float *__restrict__ grad, float *__restrict__ loss) {
    __shared__ float red[256];
    float s = 0.f;
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n;
         t += (long long)gridDim.x * 256) {
        const float d = out[t] - target[t];
        grad[t] = 2.f * d / (float)n;
        s += d * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0] / (float)n);
}

// Adam on FP32 master weights, BF16 copy refreshed
__global__ void adam_kernel(float *__restrict__ w, const float *__restrict__ g,
                            float *__restrict__ m, float *__restrict__ v,
                            bf16 *__restrict__ wb, long long n, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n;
         t += (long long)gridDim.x * 256) {
        const float gt = g[t];
        const float mt = b1 * m[t] + (1.f - b1) * gt;
        const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
        m[t] = mt, v[t] = vt;
        const float wt = w[t] - lr * (mt / c1) / (sqrtf(vt / c2) + eps);
        w[t] = wt;
        wb[t] = f2bf(wt);
    }
}

__global__ void f32_to_bf16_kernel(const float *__restrict__ x, long long n,
                                   bf16 *__restrict__ y) {
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n;
         t += (long long)gridDim.x * 256)
        y[t] = f2bf(x[t]);
}

__global__ void fill_f32_kernel(float *__restrict__ x, long long n, float v) {
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n;
         t += (long long)gridDim.x * 256)
        x[t] = v;
}

// Network input from the Bragg e^{iH} stack of nsys instances (one chunk
// of nyc planes each): channels 2k, 2k+1 = Re, Im of e^{iH} at depth
// slice k < nk over columns [x0, x0 + Wn), then s_norm[s] and t_norm[s].
// out: (nsys_total, Hn, Wn, 2 nk + 2) bf16, rows y_off .. y_off + nyc of
// instance `first + s`.
__global__ void bragg_features_kernel(const cx<double> *__restrict__ E, int nsys,
                                      int Nz, int nyc, int Nx, int x0, int Wn,
                                      int nk, const float *__restrict__ s_norm,
                                      const float *__restrict__ t_norm,
                                      bf16 *__restrict__ out, int Hn, int y_off,
                                      int first) {
    const int Cn = 2 * nk + 2;
    const long long n = (long long)nsys * nyc * Wn * (nk + 1);
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n;
         t += (long long)gridDim.x * 256) {
        const int k = (int)(t % (nk + 1));
        const long long q = t / (nk + 1);
        const int x = (int)(q % Wn);
        const int j = (int)((q / Wn) % nyc);
        const int s = (int)(q / ((long long)Wn * nyc));
        bf16 *o = out + (((long long)(first + s) * Hn + y_off + j) * Wn + x) * Cn;
        if (k < nk) {
            const cx<double> e =
                E[(((long long)s * Nz + k) * nyc + j) * Nx + x0 + x];
            o[2 * k] = f2bf((float)e.r);
            o[2 * k + 1] = f2bf((float)e.i);
        } else {
            o[2 * nk] = f2bf(s_norm[s]);
            o[2 * nk + 1] = f2bf(t_norm[s]);
        }
    }
}

static int grid_for(long long n) {
    return (int)std::max<long long>(1, std::min<long long>((n + 255) / 256, 16384));
}

}  // namespace dyf

using namespace dyf;

// How an operand X(r, k) = X[r*s_r + k*s_k] is staged: 16-byte loads
// along k when k is contiguous and rows start 16-byte aligned, along r
// when r is contiguous and k columns start aligned, else element loads.
static int load_mode(const void *X, long long s_r, long long s_k) {
    const bool base = ((uintptr_t)X & 15) == 0;
    if (base && s_k == 1 && s_r % 8 == 0) return LOAD_K;
    if (base && s_r == 1 && s_k % 8 == 0) return LOAD_ROW;
    return LOAD_ANY;
}

using GemmFn = void (*)(GemmArgs);

template <int BM, int BN, int WR>
static GemmFn gemm_for_loads(int la, int lb) {
#define DYF_GEMM_LB(LA)                                               \
    lb == LOAD_K ? gemm_bf16_kernel<BM, BN, WR, LA, LOAD_K>           \
    : lb == LOAD_ROW ? gemm_bf16_kernel<BM, BN, WR, LA, LOAD_ROW>     \
                     : gemm_bf16_kernel<BM, BN, WR, LA, LOAD_ANY>
    if (la == LOAD_K) return DYF_GEMM_LB(LOAD_K);
    if (la == LOAD_ROW) return DYF_GEMM_LB(LOAD_ROW);
    return DYF_GEMM_LB(LOAD_ANY);
#undef DYF_GEMM_LB
}

// tiles: 128 x {16, 32, 64} (waves stacked along i) for the tall products
// of the forward and data-gradient passes, 64 x {16, 32, 64} (2 x 2
// waves, or 4 along i for narrow tiles) for the weight gradients
static GemmFn pick_gemm(int bm, int bn, int la, int lb) {
    if (bm == 128) {
        if (bn == 16) return gemm_for_loads<128, 16, 4>(la, lb);
        if (bn == 32) return gemm_for_loads<128, 32, 4>(la, lb);
        return gemm_for_loads<128, 64, 4>(la, lb);
    }
    if (bn == 16) return gemm_for_loads<64, 16, 4>(la, lb);
    if (bn == 32) return gemm_for_loads<64, 32, 2>(la, lb);
    return gemm_for_loads<64, 64, 2>(la, lb);
}

extern "C" {

int dyf_nn_gemm(const void *A, long long a_i, long long a_k, const void *B,
                long long b_j, long long b_k, int M, int N, int K, int mode,
                int relu, const float *bias, void *C, long long c_i,
                int k_splits, void *stream) {
    if (!A || !B || !C || M < 1 || N < 1 || K < 1 || mode < 0 || mode > 2)
        return fail(DYF_EINVAL, "bad GEMM arguments");
    if (k_splits < 1) k_splits = 1;
    if (mode != 2 && k_splits != 1)
        return fail(DYF_EINVAL, "split-K needs the atomic epilogue");
    GemmArgs g;
    g.A = (const bf16 *)A, g.a_i = a_i, g.a_k = a_k;
    g.B = (const bf16 *)B, g.b_j = b_j, g.b_k = b_k;
    g.M = M, g.N = N, g.K = K;
    g.k_chunk = ((K + k_splits - 1) / k_splits + TK - 1) / TK * TK;
    g.mode = mode, g.relu = relu, g.bias = bias, g.C = C, g.c_i = c_i;
    const int splits = (K + g.k_chunk - 1) / g.k_chunk;
    const int la = load_mode(A, a_i, a_k), lb = load_mode(B, b_j, b_k);
    const int bm = M >= 4096 ? 128 : 64;
    const int bn = N <= 16 ? 16 : N <= 32 ? 32 : 64;
    GemmFn fn = pick_gemm(bm, bn, la, lb);
    dim3 grid((M + bm - 1) / bm, (N + bn - 1) / bn, splits);
    hipLaunchKernelGGL(fn, grid, dim3(256), 0, (hipStream_t)stream, g);
    DYF_LAUNCHED("gemm_bf16_kernel");
    return DYF_OK;
}

int dyf_nn_im2col3(const void *X, int Bn, int H, int W, int Ctot, int off,
                   int C, void *col, void *stream) {
    const long long n = (long long)Bn * H * W * 9 * C;
    im2col3_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>(
        (const bf16 *)X, Bn, H, W, Ctot, off, C, (bf16 *)col);
    DYF_LAUNCHED("im2col3_kernel");
    return DYF_OK;
}

int dyf_nn_col2im3(const float *dcol, int Bn, int H, int W, int C, float *dX,
                   int Ctot, int off, void *stream) {
    const long long n = (long long)Bn * H * W * C;
    col2im3_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>(
        dcol, Bn, H, W, C, dX, Ctot, off);
    DYF_LAUNCHED("col2im3_kernel");
    return DYF_OK;
}

int dyf_nn_slice(const void *X, long long P, int Ctot, int off, int C,
                 void *out, void *stream) {
    slice_kernel<<<grid_for(P * C), 256, 0, (hipStream_t)stream>>>(
        (const bf16 *)X, P, Ctot, off, C, (bf16 *)out);
    DYF_LAUNCHED("slice_kernel");
    return DYF_OK;
}

int dyf_nn_relu_grad(const float *dY, const void *Y, long long P, int Ctot,
                     int off, int C, int relu, void *dZ, void *stream) {
    relu_grad_kernel<<<grid_for(P * C), 256, 0, (hipStream_t)stream>>>(
        dY, (const bf16 *)Y, P, Ctot, off, C, relu, (bf16 *)dZ);
    DYF_LAUNCHED("relu_grad_kernel");
    return DYF_OK;
}

int dyf_nn_bias_grad(const void *dZ, long long P, int C, float *db,
                     void *stream) {
    bias_grad_kernel<<<C, 256, 0, (hipStream_t)stream>>>((const bf16 *)dZ, P,
                                                          C, db);
    DYF_LAUNCHED("bias_grad_kernel");
    return DYF_OK;
}

int dyf_nn_pool(const void *X, int Bn, int H, int W, int Ctot, int off, int C,
                void *Y, void *stream) {
    if ((H & 1) || (W & 1)) return fail(DYF_EINVAL, "pool needs even H, W");
    const long long n = (long long)Bn * (H / 2) * (W / 2) * C;
    pool_fwd_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>(
        (const bf16 *)X, Bn, H, W, Ctot, off, C, (bf16 *)Y);
    DYF_LAUNCHED("pool_fwd_kernel");
    return DYF_OK;
}

int dyf_nn_pool_grad(const float *dY, int Bn, int H, int W, int C, float *dX,
                     int Ctot, int off, void *stream) {
    const long long n = (long long)Bn * H * W * C;
    pool_bwd_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>(
        dY, Bn, H, W, C, dX, Ctot, off);
    DYF_LAUNCHED("pool_bwd_kernel");
    return DYF_OK;
}

int dyf_nn_up(const void *X, int Bn, int H, int W, int C, void *Y, int Ctot,
              int off, void *stream) {
    const long long n = (long long)Bn * 4 * H * W * C;
    up_fwd_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>(
        (const bf16 *)X, Bn, H, W, C, (bf16 *)Y, Ctot, off);
    DYF_LAUNCHED("up_fwd_kernel");
    return DYF_OK;
}

int dyf_nn_up_grad(const float *dY, int Bn, int H, int W, int C, int Ctot,
                   int off, float *dX, void *stream) {
    const long long n = (long long)Bn * H * W * C;
    up_bwd_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>(
        dY, Bn, H, W, C, Ctot, off, dX);
    DYF_LAUNCHED("up_bwd_kernel");
    return DYF_OK;
}

int dyf_nn_mse(const float *out, const float *target, long long n,
               float *grad, float *loss, void *stream) {
    mse_kernel<<<std::min(grid_for(n), 1024), 256, 0, (hipStream_t)stream>>>(
        out, target, n, grad, loss);
    DYF_LAUNCHED("mse_kernel");
    return DYF_OK;
}

int dyf_nn_adam(float *w, const float *g, float *m, float *v, void *wb,
                long long n, float lr, float b1, float b2, float eps, int t,
                void *stream) {
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    adam_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>(
        w, g, m, v, (bf16 *)wb, n, lr, b1, b2, eps, c1, c2);
    DYF_LAUNCHED("adam_kernel");
    return DYF_OK;
}

int dyf_nn_to_bf16(const float *x, long long n, void *y, void *stream) {
    f32_to_bf16_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>(
        x, n, (bf16 *)y);
    DYF_LAUNCHED("f32_to_bf16_kernel");
    return DYF_OK;
}

int dyf_nn_bragg_features(const void *E, int nsys, int Nz, int nyc, int Nx,
                          int x0, int Wn, int nk, const float *s_norm,
                          const float *t_norm, void *out, int Hn, int y_off,
                          int first, void *stream) {
    if (!E || !out || !s_norm || !t_norm || nk < 1 || nk > Nz || x0 < 0 ||
        x0 + Wn > Nx || y_off < 0 || y_off + nyc > Hn)
        return fail(DYF_EINVAL, "bad feature extraction arguments");
    const long long n = (long long)nsys * nyc * Wn * (nk + 1);
    bragg_features_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>(
        (const cx<double> *)E, nsys, Nz, nyc, Nx, x0, Wn, nk, s_norm, t_norm,
        (bf16 *)out, Hn, y_off, first);
    DYF_LAUNCHED("bragg_features_kernel");
    return DYF_OK;
}

int dyf_nn_fill(float *x, long long n, float v, void *stream) {
    fill_f32_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>(x, n, v);
    DYF_LAUNCHED("fill_f32_kernel");
    return DYF_OK;
}

}  // extern "C"
