// Exit wave -> image: pupils, batched 2-D spectra through hipFFT, image
// intensity, and the centre-of-mass rocking accumulation
// (optics.py:55-207, analysis/rocking.py:37-83).
//
// Layout: every field, spectrum, pupil and image is (Ny, Nx) row-major,
// complex interleaved, batches contiguous.  The reference's pupils live
// on the fftshifted frequency grid and its images are
//     I = |fftshift(ifft2(ifftshift(fftshift(fft2(fftshift(f))) P)))|^2 .
// With T the roll by N/2 (integer division) per axis, fftshift = T and
// ifftshift = T^-1, and a pupil multiply is a convolution, which commutes
// with rolls, so
//     I = T^2 |ifft2(fft2(f) . T^-1 P)|^2 :
// one forward and one inverse transform with no shift copies; the pupil
// is read at the ifftshifted index and the image written through T^2
// (the identity for even N, a one-pixel roll for odd N).
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <map>
#include <mutex>
#include <tuple>

#include "common.hpp"

namespace dyf {
namespace {

using cd = cx<double>;
using cf = cx<float>;

// frequency of index i on numpy's fftshift(fftfreq(n, d)) grid: the same
// integer times 1/(n d) product numpy forms
__device__ __forceinline__ double shifted_freq(int i, int n, double val) {
    return (double)(i - n / 2) * val;
}

template <typename T>
__global__ void pupil_kernel(cx<T> *__restrict__ out, int Nx, int Ny,
                             double vx, double vy, double af, double ax,
                             double ay, double sx, double sy, int ell,
                             int levels, int zero_dc) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y;
    if (ix >= Nx) return;
    const double FX = shifted_freq(ix, Nx, vx), FY = shifted_freq(iy, Ny, vy);
    const double dxf = FX - ax, dyf = FY - ay;
    const double rad = sqrt(dxf * dxf + dyf * dyf);
    const double circ = rad <= af ? 1.0 : 0.0;
    double re = circ, im = 0.0;
    if (ell != 0) {
        double ph = (double)ell * atan2(FY - sy, FX - sx);
        if (levels) {
            const double step = 2.0 * kPi / (double)levels;
            double m = fmod(ph, 2.0 * kPi);        // numpy mod: sign of 2 pi
            if (m != 0.0 && m < 0.0) m += 2.0 * kPi;
            ph = floor(m / step) * step;
        }
        double s, c;
        sincos(ph, &s, &c);
        re = circ * c, im = circ * s;
        if (zero_dc && ix == Nx / 2 && iy == Ny / 2) re = im = 0.0;
    }
    out[(size_t)iy * Nx + ix] = cx<T>{(T)re, (T)im};
}

// (re)shift copies for the explicit spectrum functions: out[n] =
// in[(n - s) mod N] per axis (s = N/2: fftshift, s = -(N/2): ifftshift)
template <typename T>
__global__ void roll2_kernel(const cx<T> *__restrict__ in, cx<T> *__restrict__ out,
                             int Nx, int Ny, int sx, int sy, size_t n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const size_t plane = (size_t)Nx * Ny;
    const size_t b = t / plane, p = t % plane;
    const int iy = (int)(p / Nx), ix = (int)(p % Nx);
    int jx = (ix - sx) % Nx, jy = (iy - sy) % Ny;
    jx += jx < 0 ? Nx : 0;
    jy += jy < 0 ? Ny : 0;
    out[t] = in[b * plane + (size_t)jy * Nx + jx];
}

// G[i, p] = F[i] . P'[p] (P' = ifftshift(P)), promoted to TO
template <typename TF, typename TP, typename TO>
__global__ void pupil_multiply_kernel(const cx<TF> *__restrict__ F,
                                      const cx<TP> *__restrict__ P, int nf,
                                      int np, int Nx, int Ny, int shifted,
                                      cx<TO> *__restrict__ G, size_t n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const size_t plane = (size_t)Nx * Ny;
    const size_t q = t / plane, k = t % plane;       // q = i * np + p
    const int i = (int)(q / np), p = (int)(q % np);
    const int ky = (int)(k / Nx), kx = (int)(k % Nx);
    // P is on the fftshifted grid: P'[k] = P[(k + N/2) mod N] (shifted)
    // or already in FFT order (P'[k] = P[k])
    const int jx = shifted ? (kx + Nx / 2) % Nx : kx;
    const int jy = shifted ? (ky + Ny / 2) % Ny : ky;
    const cx<TP> pv = P[(size_t)p * plane + (size_t)jy * Nx + jx];
    const cx<TF> fv = F[(size_t)i * plane + k];
    const TO fr = (TO)fv.r, fi = (TO)fv.i, pr = (TO)pv.r, pi = (TO)pv.i;
    G[t] = cx<TO>{fr * pr - fi * pi, fr * pi + fi * pr};
}

// I[n] = |h[(n - s) mod N]|^2 scale^2 (s = 2 (N/2) per axis: T^2)
template <typename T>
__global__ void intensity_kernel(const cx<T> *__restrict__ h, int Nx, int Ny,
                                 int sx, int sy, T scale, T *__restrict__ I,
                                 size_t n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const size_t plane = (size_t)Nx * Ny;
    const size_t b = t / plane, p = t % plane;
    const int iy = (int)(p / Nx), ix = (int)(p % Nx);
    int jx = (ix - sx) % Nx, jy = (iy - sy) % Ny;
    jx += jx < 0 ? Nx : 0;
    jy += jy < 0 ? Ny : 0;
    const cx<T> v = h[b * plane + (size_t)jy * Nx + jx];
    const T re = v.r * scale, im = v.i * scale;
    I[t] = re * re + im * im;
}

// COM sums over a batch of images, FP64 (one thread per pixel, the batch
// summed in order, so the result is deterministic)
template <typename T>
__global__ void com_accumulate_kernel(const T *__restrict__ I, int nb,
                                      size_t plane, const double *__restrict__ w,
                                      double *__restrict__ swI,
                                      double *__restrict__ sI) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= plane) return;
    double a = swI[p], b = sI[p];
    for (int i = 0; i < nb; ++i) {
        const double v = (double)I[(size_t)i * plane + p];
        a += v * w[i];
        b += v;
    }
    swI[p] = a, sI[p] = b;
}

// rock[i] = sum over the image i, FP64; one workgroup per image, a fixed
// reduction tree (deterministic)
template <typename T>
__global__ __launch_bounds__(1024) void image_sum_kernel(
    const T *__restrict__ I, size_t plane, double *__restrict__ out) {
    __shared__ double red[1024];
    const T *src = I + (size_t)blockIdx.x * plane;
    double acc = 0.0;
    for (size_t p = threadIdx.x; p < plane; p += 1024) acc += (double)src[p];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// ---- hipFFT 2-D plans, cached per (Ny, Nx, batch, type) --------------
std::mutex g_plan_mu;
std::map<std::tuple<int, int, int, int>, hipfftHandle> g_plans;

int plan2d(int Nx, int Ny, int batch, bool f64, hipStream_t stream,
           hipfftHandle *out) {
    std::lock_guard<std::mutex> lock(g_plan_mu);
    const auto key = std::make_tuple(Ny, Nx, batch, (int)f64);
    auto it = g_plans.find(key);
    hipfftHandle h;
    if (it == g_plans.end()) {
        int dims[2] = {Ny, Nx};
        if (hipfftPlanMany(&h, 2, dims, nullptr, 1, Nx * Ny, nullptr, 1,
                           Nx * Ny, f64 ? HIPFFT_Z2Z : HIPFFT_C2C,
                           batch) != HIPFFT_SUCCESS)
            return fail(DYF_EHIP, "hipfftPlanMany (%d x %d, batch %d) failed",
                        Ny, Nx, batch);
        g_plans.emplace(key, h);
    } else {
        h = it->second;
    }
    if (hipfftSetStream(h, stream) != HIPFFT_SUCCESS)
        return fail(DYF_EHIP, "hipfftSetStream failed");
    *out = h;
    return DYF_OK;
}

int fft2(void *data, int Nx, int Ny, int batch, bool f64, int dir,
         hipStream_t stream) {
    if (batch <= 0) return DYF_OK;
    hipfftHandle h;
    int rc = plan2d(Nx, Ny, batch, f64, stream, &h);
    if (rc) return rc;
    hipfftResult r;
    if (f64)
        r = hipfftExecZ2Z(h, (hipfftDoubleComplex *)data,
                          (hipfftDoubleComplex *)data, dir);
    else
        r = hipfftExecC2C(h, (hipfftComplex *)data, (hipfftComplex *)data,
                          dir);
    if (r != HIPFFT_SUCCESS) return fail(DYF_EHIP, "hipfftExec failed (%d)", r);
    return DYF_OK;
}

bool is_f64(int prec) { return prec == DYF_FP64; }

int check_grid(int Nx, int Ny, int n) {
    if (Nx < 1 || Ny < 1 || n < 0) return fail(DYF_EINVAL, "bad image grid");
    return DYF_OK;
}

unsigned blocks_for(size_t n, int threads) {
    return (unsigned)((n + threads - 1) / threads);
}

}  // namespace
}  // namespace dyf

using namespace dyf;

extern "C" int dyf_optics_pupil(const dyf_pupil_args *a, int prec, void *out,
                                void *stream) {
    if (!a || !out) return fail(DYF_EINVAL, "NULL pupil arguments");
    int rc = check_grid(a->Nx, a->Ny, 1);
    if (rc) return rc;
    if (!(a->dx > 0.0) || a->levels < 0)
        return fail(DYF_EINVAL, "bad pupil dx or levels");
    const double vx = 1.0 / ((double)a->Nx * a->dx);
    const double vy = 1.0 / ((double)a->Ny * a->dx);
    dim3 grid(blocks_for(a->Nx, 256), a->Ny);
    if (is_f64(prec))
        pupil_kernel<double><<<grid, 256, 0, (hipStream_t)stream>>>(
            (cd *)out, a->Nx, a->Ny, vx, vy, a->aperture_f, a->ax, a->ay,
            a->sx, a->sy, a->ell, a->levels, a->zero_dc);
    else
        pupil_kernel<float><<<grid, 256, 0, (hipStream_t)stream>>>(
            (cf *)out, a->Nx, a->Ny, vx, vy, a->aperture_f, a->ax, a->ay,
            a->sx, a->sy, a->ell, a->levels, a->zero_dc);
    DYF_LAUNCHED("pupil_kernel");
    return DYF_OK;
}

extern "C" int dyf_optics_spectrum(const void *field, int n, int Nx, int Ny,
                                   int prec, void *work, void *spec,
                                   void *stream_) {
    int rc = check_grid(Nx, Ny, n);
    if (rc) return rc;
    if (n == 0) return DYF_OK;
    if (!field || !work || !spec) return fail(DYF_EINVAL, "NULL spectrum buffer");
    hipStream_t stream = (hipStream_t)stream_;
    const size_t tot = (size_t)n * Nx * Ny;
    const bool f64 = is_f64(prec);
    // spec = T fft2(T f): shift into work, transform, shift out
    if (f64)
        roll2_kernel<double><<<blocks_for(tot, 256), 256, 0, stream>>>(
            (const cd *)field, (cd *)work, Nx, Ny, Nx / 2, Ny / 2, tot);
    else
        roll2_kernel<float><<<blocks_for(tot, 256), 256, 0, stream>>>(
            (const cf *)field, (cf *)work, Nx, Ny, Nx / 2, Ny / 2, tot);
    DYF_LAUNCHED("roll2_kernel");
    rc = fft2(work, Nx, Ny, n, f64, HIPFFT_FORWARD, stream);
    if (rc) return rc;
    if (f64)
        roll2_kernel<double><<<blocks_for(tot, 256), 256, 0, stream>>>(
            (const cd *)work, (cd *)spec, Nx, Ny, Nx / 2, Ny / 2, tot);
    else
        roll2_kernel<float><<<blocks_for(tot, 256), 256, 0, stream>>>(
            (const cf *)work, (cf *)spec, Nx, Ny, Nx / 2, Ny / 2, tot);
    DYF_LAUNCHED("roll2_kernel");
    return DYF_OK;
}

extern "C" int dyf_optics_image_from_spectrum(const void *spec, int n,
                                              int spec_prec, const void *pupil,
                                              int pupil_per_item,
                                              int pupil_prec, int Nx, int Ny,
                                              void *work, void *out,
                                              void *stream_) {
    int rc = check_grid(Nx, Ny, n);
    if (rc) return rc;
    if (n == 0) return DYF_OK;
    if (!spec || !pupil || !work || !out)
        return fail(DYF_EINVAL, "NULL image buffer");
    hipStream_t stream = (hipStream_t)stream_;
    const size_t plane = (size_t)Nx * Ny, tot = (size_t)n * plane;
    const bool o64 = is_f64(spec_prec) || is_f64(pupil_prec);
    // work (2 n planes, output precision): the product spec . P on the
    // shifted grid, then ifftshift(product), transformed in place
#define DYF_MUL(TF, TP, TO)                                                  \
    for (int i = 0; i < n; ++i)                                              \
        pupil_multiply_kernel<TF, TP, TO>                                    \
            <<<blocks_for(plane, 256), 256, 0, stream>>>(                    \
                (const cx<TF> *)spec + (size_t)i * plane,                    \
                (const cx<TP> *)pupil + (pupil_per_item ? (size_t)i * plane  \
                                                        : 0),                \
                1, 1, Nx, Ny, 0, (cx<TO> *)work + (size_t)i * plane, plane)
    if (o64) {
        if (is_f64(spec_prec) && is_f64(pupil_prec)) DYF_MUL(double, double, double);
        else if (is_f64(spec_prec)) DYF_MUL(double, float, double);
        else DYF_MUL(float, double, double);
    } else {
        DYF_MUL(float, float, float);
    }
#undef DYF_MUL
    DYF_LAUNCHED("pupil_multiply_kernel");
    const size_t esz = o64 ? sizeof(cd) : sizeof(cf);
    char *w2 = (char *)work + tot * esz;
    if (o64)
        roll2_kernel<double><<<blocks_for(tot, 256), 256, 0, stream>>>(
            (const cd *)work, (cd *)w2, Nx, Ny, -(Nx / 2), -(Ny / 2), tot);
    else
        roll2_kernel<float><<<blocks_for(tot, 256), 256, 0, stream>>>(
            (const cf *)work, (cf *)w2, Nx, Ny, -(Nx / 2), -(Ny / 2), tot);
    DYF_LAUNCHED("roll2_kernel");
    rc = fft2(w2, Nx, Ny, n, o64, HIPFFT_BACKWARD, stream);
    if (rc) return rc;
    // fftshift of the inverse transform, |.|^2, the 1/(Nx Ny) of ifft2
    if (o64)
        intensity_kernel<double><<<blocks_for(tot, 256), 256, 0, stream>>>(
            (const cd *)w2, Nx, Ny, Nx / 2, Ny / 2, 1.0 / (double)plane,
            (double *)out, tot);
    else
        intensity_kernel<float><<<blocks_for(tot, 256), 256, 0, stream>>>(
            (const cf *)w2, Nx, Ny, Nx / 2, Ny / 2, 1.0f / (float)plane,
            (float *)out, tot);
    DYF_LAUNCHED("intensity_kernel");
    return DYF_OK;
}

extern "C" int dyf_optics_image(const void *field, int n, int field_prec,
                                const void *pupils, int npup, int pupil_prec,
                                int Nx, int Ny, void *work, void *out,
                                void *stream_) {
    int rc = check_grid(Nx, Ny, n);
    if (rc) return rc;
    if (npup < 1) return fail(DYF_EINVAL, "no pupils");
    if (n == 0) return DYF_OK;
    if (!field || !pupils || !work || !out)
        return fail(DYF_EINVAL, "NULL image buffer");
    hipStream_t stream = (hipStream_t)stream_;
    const size_t plane = (size_t)Nx * Ny;
    const bool f64 = is_f64(field_prec);
    const bool o64 = f64 || is_f64(pupil_prec);
    const size_t fsz = f64 ? sizeof(cd) : sizeof(cf);
    // work: F (n planes, field precision), then G (n * npup planes, output
    // precision)
    char *F = (char *)work;
    char *G = F + (size_t)n * plane * fsz;
    DYF_HIP(hipMemcpyAsync(F, field, (size_t)n * plane * fsz,
                           hipMemcpyDeviceToDevice, stream));
    rc = fft2(F, Nx, Ny, n, f64, HIPFFT_FORWARD, stream);
    if (rc) return rc;
    const size_t tot = (size_t)n * npup * plane;
#define DYF_MUL(TF, TP, TO)                                                  \
    pupil_multiply_kernel<TF, TP, TO><<<blocks_for(tot, 256), 256, 0,        \
                                        stream>>>(                           \
        (const cx<TF> *)F, (const cx<TP> *)pupils, n, npup, Nx, Ny, 1,       \
        (cx<TO> *)G, tot)
    if (o64) {
        if (f64 && is_f64(pupil_prec)) DYF_MUL(double, double, double);
        else if (f64) DYF_MUL(double, float, double);
        else DYF_MUL(float, double, double);
    } else {
        DYF_MUL(float, float, float);
    }
#undef DYF_MUL
    DYF_LAUNCHED("pupil_multiply_kernel");
    rc = fft2(G, Nx, Ny, n * npup, o64, HIPFFT_BACKWARD, stream);
    if (rc) return rc;
    const int rx = 2 * (Nx / 2) % Nx, ry = 2 * (Ny / 2) % Ny;
    if (o64)
        intensity_kernel<double><<<blocks_for(tot, 256), 256, 0, stream>>>(
            (const cd *)G, Nx, Ny, rx, ry, 1.0 / (double)plane,
            (double *)out, tot);
    else
        intensity_kernel<float><<<blocks_for(tot, 256), 256, 0, stream>>>(
            (const cf *)G, Nx, Ny, rx, ry, 1.0f / (float)plane, (float *)out,
            tot);
    DYF_LAUNCHED("intensity_kernel");
    return DYF_OK;
}

extern "C" int dyf_optics_image_bytes(int n, int npup, int field_prec,
                                      int pupil_prec, int Nx, int Ny,
                                      size_t *bytes) {
    int rc = check_grid(Nx, Ny, n);
    if (rc) return rc;
    if (!bytes || npup < 1) return fail(DYF_EINVAL, "bad image sizes");
    const size_t plane = (size_t)Nx * Ny;
    const bool f64 = is_f64(field_prec), o64 = f64 || is_f64(pupil_prec);
    *bytes = (size_t)n * plane * (f64 ? sizeof(cd) : sizeof(cf)) +
             (size_t)n * npup * plane * (o64 ? sizeof(cd) : sizeof(cf));
    return DYF_OK;
}

extern "C" int dyf_optics_com_accumulate(const void *images, int n, int prec,
                                         int Nx, int Ny,
                                         const double *weights,
                                         double *sum_wI, double *sum_I,
                                         double *image_sums, void *stream_) {
    int rc = check_grid(Nx, Ny, n);
    if (rc) return rc;
    if (n == 0) return DYF_OK;
    if (!images || !weights || !sum_wI || !sum_I)
        return fail(DYF_EINVAL, "NULL COM buffer");
    hipStream_t stream = (hipStream_t)stream_;
    const size_t plane = (size_t)Nx * Ny;
    if (is_f64(prec)) {
        com_accumulate_kernel<double><<<blocks_for(plane, 256), 256, 0, stream>>>(
            (const double *)images, n, plane, weights, sum_wI, sum_I);
        if (image_sums)
            image_sum_kernel<double><<<n, 1024, 0, stream>>>(
                (const double *)images, plane, image_sums);
    } else {
        com_accumulate_kernel<float><<<blocks_for(plane, 256), 256, 0, stream>>>(
            (const float *)images, n, plane, weights, sum_wI, sum_I);
        if (image_sums)
            image_sum_kernel<float><<<n, 1024, 0, stream>>>(
                (const float *)images, plane, image_sums);
    }
    DYF_LAUNCHED("com_accumulate_kernel");
    return DYF_OK;
}
