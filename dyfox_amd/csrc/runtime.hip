// Runtime entry points of libdyfox: device selection, memory, copies and
// the thread-local error string.  Also the host-side segment reduction
// shared by every Volterra kernel.
#include "common.hpp"

namespace dyf {

static thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
                         ? DYF_ENOMEM
                         : DYF_EHIP;
    return fail(code, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

static inline double dot3(const double *a, const double *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline double norm3(const double *a) { return std::sqrt(dot3(a, a)); }
static inline void cross3(const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// Screw/edge split of b about the line and the (e1, e2) frame, with the
// pure-screw axis choice of dislocations.py:68-90.
static void volterra_frame(const double *xi, const double *b, double *b_screw,
                           double *b_perp, double *e1, double *e2) {
    *b_screw = dot3(b, xi);
    double bp[3] = {b[0] - *b_screw * xi[0], b[1] - *b_screw * xi[1],
                    b[2] - *b_screw * xi[2]};
    *b_perp = norm3(bp);
    if (*b_perp > 1e-20) {
        for (int j = 0; j < 3; ++j) e1[j] = bp[j] / *b_perp;
    } else {
        const double ax[3] = {1.0, 0.0, 0.0}, ay[3] = {0.0, 1.0, 0.0};
        if (std::fabs(xi[2]) > 0.9) {
            e1[0] = 1.0, e1[1] = 0.0, e1[2] = 0.0;
        } else {
            cross3(xi, std::fabs(xi[0]) < 0.9 ? ax : ay, e1);
            const double n = norm3(e1);
            for (int j = 0; j < 3; ++j) e1[j] /= n;
        }
    }
    cross3(xi, e1, e2);
}

int prepare_segments(const double *segs, int nseg, const double *g,
                     double nu, SegParams *out) {
    int kept = 0;
    const double onu = 1.0 - nu;
    for (int n = 0; n < nseg; ++n) {
        const double *ra = segs + 9 * n, *rb = ra + 3, *b = ra + 6;
        double t[3] = {rb[0] - ra[0], rb[1] - ra[1], rb[2] - ra[2]};
        const double L = norm3(t);
        if (L < 1e-20) continue;
        double xi[3] = {t[0] / L, t[1] / L, t[2] / L};
        double b_screw, b_perp, e1[3], e2[3];
        volterra_frame(xi, b, &b_screw, &b_perp, e1, e2);
        SegParams &p = out[kept++];
        p.cx = (ra[0] + rb[0]) / 2.0;
        p.cy = (ra[1] + rb[1]) / 2.0;
        p.cz = (ra[2] + rb[2]) / 2.0;
        p.xx = xi[0], p.xy = xi[1], p.xz = xi[2];
        p.e1x = e1[0], p.e1y = e1[1], p.e1z = e1[2];
        p.e2x = e2[0], p.e2y = e2[1], p.e2z = e2[2];
        const double ge1 = dot3(g, e1), ge2 = dot3(g, e2);
        p.c_atan = ge1 * b_perp + dot3(g, xi) * b_screw;
        const double c1 = b_perp / kTwoPi;
        p.k_xy = kTwoPi * ge1 * c1 / (2.0 * onu);
        p.k_dd = -kTwoPi * ge2 * c1 / (4.0 * onu);
        p.k_log = -kTwoPi * ge2 * c1 * (1.0 - 2.0 * nu) / (4.0 * onu);
        p.edge = b_perp > 1e-20 ? 1.0 : 0.0;
        const double x1 = dot3(e1, xi), x2 = dot3(e2, xi);
        p.f1x = e1[0] - x1 * xi[0], p.f1y = e1[1] - x1 * xi[1],
        p.f1z = e1[2] - x1 * xi[2];
        p.f2x = e2[0] - x2 * xi[0], p.f2y = e2[1] - x2 * xi[1],
        p.f2z = e2[2] - x2 * xi[2];
        const double nw = std::nearbyint(p.c_atan);
        const bool integral =
            std::fabs(p.c_atan - nw) <=
                1e-9 * std::fmax(1.0, std::fabs(p.c_atan)) &&
            std::fabs(nw) <= 8.0;
        p.nwind = integral ? (int32_t)nw : 0;
        p.frac = integral ? 0 : 1;
    }
    return kept;
}

RowBand laue_rows(const dyf_laue_args *a) {
    RowBand band{a->Ny, a->Ny};
    int lo = -1, hi = -1;
    for (int iy = 0; iy < a->Ny; ++iy) {
        const double y = (double)(iy - a->Ny / 2) * a->dx;
        if (std::fabs(y) < a->fov_hy) {
            if (lo < 0) lo = iy;
            hi = iy + 1;
        }
    }
    if (lo >= 0) band = RowBand{lo, hi};
    return band;
}

__global__ __launch_bounds__(256) void copy16_kernel(uint4 *__restrict__ dst,
                                                     const uint4 *__restrict__ src,
                                                     size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (size_t)gridDim.x * 256)
        dst[i] = src[i];
}

}  // namespace dyf

using namespace dyf;

extern "C" {

const char *dyf_last_error(void) { return g_err; }

int dyf_device_count(int *count) {
    if (!count) return fail(DYF_EINVAL, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return DYF_OK;
}

int dyf_init(int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
        return fail(DYF_ENODEV, "no HIP device is visible");
    if (device < 0 || device >= n)
        return fail(DYF_EINVAL, "device %d out of range [0, %d)", device, n);
    DYF_HIP(hipSetDevice(device));
    DYF_HIP(hipFree(nullptr));  // create the context now
    return DYF_OK;
}

int dyf_malloc(void **ptr, size_t bytes) {
    if (!ptr) return fail(DYF_EINVAL, "ptr is NULL");
    *ptr = nullptr;
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        *ptr = nullptr;
        (void)hipGetLastError();
        return fail(DYF_ENOMEM, "hipMalloc(%zu bytes): %s", bytes,
                    hipGetErrorString(e));
    }
    return DYF_OK;
}

int dyf_free(void *ptr) {
    if (ptr) DYF_HIP(hipFree(ptr));
    return DYF_OK;
}

int dyf_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    if (bytes) DYF_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return DYF_OK;
}

int dyf_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    if (bytes) DYF_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return DYF_OK;
}

int dyf_memcpy_d2d(void *dst, const void *src, size_t bytes) {
    if (bytes) DYF_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return DYF_OK;
}

int dyf_memset(void *ptr, int value, size_t bytes) {
    if (bytes) DYF_HIP(hipMemset(ptr, value, bytes));
    return DYF_OK;
}

int dyf_memcpy2d_d2h(void *dst, size_t dpitch, const void *src,
                     size_t spitch, size_t width, size_t height) {
    if (!width || !height) return DYF_OK;
    if (!dst || !src) return fail(DYF_EINVAL, "NULL pointer to dyf_memcpy2d_d2h");
    if (dpitch < width || spitch < width)
        return fail(DYF_EINVAL, "pitch below the row width");
    DYF_HIP(hipMemcpy2D(dst, dpitch, src, spitch, width, height,
                        hipMemcpyDeviceToHost));
    return DYF_OK;
}

int dyf_host_alloc(void **ptr, size_t bytes) {
    if (!ptr) return fail(DYF_EINVAL, "ptr is NULL");
    *ptr = nullptr;
    hipError_t e = hipHostMalloc(ptr, bytes ? bytes : 256, hipHostMallocDefault);
    if (e != hipSuccess) {
        *ptr = nullptr;
        (void)hipGetLastError();
        return fail(DYF_ENOMEM, "hipHostMalloc(%zu bytes): %s", bytes,
                    hipGetErrorString(e));
    }
    return DYF_OK;
}

int dyf_host_free(void *ptr) {
    if (ptr) DYF_HIP(hipHostFree(ptr));
    return DYF_OK;
}

int dyf_mem_info(size_t *free_bytes, size_t *total_bytes) {
    size_t f = 0, t = 0;
    DYF_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return DYF_OK;
}

int dyf_synchronize(void) {
    DYF_HIP(hipDeviceSynchronize());
    return DYF_OK;
}

int dyf_stream_sync(void *stream) {
    DYF_HIP(hipStreamSynchronize((hipStream_t)stream));
    return DYF_OK;
}

int dyf_copy_kernel(void *dst, const void *src, size_t bytes, void *stream) {
    if ((bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15))
        return fail(DYF_EINVAL, "copy needs 16-byte sizes and alignment");
    const size_t n = bytes / 16;
    if (n == 0) return DYF_OK;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 8192);
    copy16_kernel<<<blocks, 256, 0, (hipStream_t)stream>>>(
        (uint4 *)dst, (const uint4 *)src, n);
    DYF_LAUNCHED("copy16_kernel");
    return DYF_OK;
}

int dyf_event_create(void **event) {
    if (!event) return fail(DYF_EINVAL, "event is NULL");
    hipEvent_t e;
    DYF_HIP(hipEventCreate(&e));
    *event = (void *)e;
    return DYF_OK;
}

int dyf_event_record(void *event, void *stream) {
    DYF_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return DYF_OK;
}

int dyf_event_elapsed_ms(void *start, void *stop, float *ms) {
    if (!ms) return fail(DYF_EINVAL, "ms is NULL");
    DYF_HIP(hipEventSynchronize((hipEvent_t)stop));
    DYF_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return DYF_OK;
}

int dyf_event_destroy(void *event) {
    if (event) DYF_HIP(hipEventDestroy((hipEvent_t)event));
    return DYF_OK;
}

}  // extern "C"
