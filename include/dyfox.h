/*
 * dyfox.h -- C ABI of libdyfox.so, the MI355X (gfx950) Takagi-Taupin engine.
 *
 * Every entry point returns an int status (DYF_OK = 0, negative on error);
 * dyf_last_error() gives a thread-local message for the last failure.
 * All pointers named *_dev are device pointers obtained from dyf_malloc
 * (or any HIP allocation); every other pointer is host memory.  `stream`
 * is a hipStream_t (NULL = the legacy default stream).  Launches are
 * asynchronous on `stream`; call dyf_synchronize() or dyf_stream_sync()
 * before reading results on the host.
 *
 * The reference (DYFOX) has no compiled boundary: its kernels are CUDA C
 * strings compiled by CuPy/NVRTC and launched with positional argument
 * tuples.  Each declaration below cites the reference call it replaces.
 */
#ifndef DYFOX_H
#define DYFOX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------- */
enum {
    DYF_OK = 0,
    DYF_EINVAL = -1,   /* bad argument                 -> ValueError  */
    DYF_ENOMEM = -2,   /* device memory / LDS exceeded -> MemoryError */
    DYF_EKEY = -3,     /* unknown mode or precision    -> KeyError    */
    DYF_EHIP = -4,     /* HIP runtime error            -> RuntimeError */
    DYF_ENODEV = -5    /* no usable device             -> RuntimeError */
};

/* ---- precisions and march modes ------------------------------------- */
enum {
    DYF_FP64 = 0,   /* complex128 amplitudes, FP64 phase                   */
    DYF_FP32 = 1,   /* complex64 amplitudes, FP32 state, FP64-reduced phase */
    DYF_MIXED = 2   /* complex64 out; FP32 two-float compensated state      */
};
enum {
    DYF_MODE_GRID = 0,      /* per-voxel H, any segments   (tt3d_fullfield) */
    DYF_MODE_ANALYTIC = 1,  /* one line along x_lab, H(y,z) (tt3d_analytic)  */
    DYF_MODE_PRISTINE = 2   /* H = 0                        (tt3d_pristine)  */
};
enum {
    DYF_PATH_FAST = 0,      /* exact one-pixel shift, register-resident rows */
    DYF_PATH_GENERAL = 1    /* any shift (linear interpolation), LDS rows    */
};

/* ---- runtime (replaces CuPy allocation/copies, backend.py:18-47) ----- */
int dyf_init(int device);
const char *dyf_last_error(void);
int dyf_device_count(int *count);
int dyf_malloc(void **ptr_dev, size_t bytes);
int dyf_free(void *ptr_dev);
int dyf_memcpy_h2d(void *dst_dev, const void *src, size_t bytes);
int dyf_memcpy_d2h(void *dst, const void *src_dev, size_t bytes);
int dyf_memcpy_d2d(void *dst_dev, const void *src_dev, size_t bytes);
int dyf_memset(void *ptr_dev, int value, size_t bytes);
/* Strided device-to-host copy: `height` rows of `width` bytes, rows
 * `spitch` bytes apart on the device and `dpitch` bytes apart on the host
 * (a chunk of planes into its slice of a larger host array).             */
int dyf_memcpy2d_d2h(void *dst, size_t dpitch, const void *src_dev,
                     size_t spitch, size_t width, size_t height);
/* Page-locked host memory (hipHostMalloc): device-to-host copies into it
 * run at link speed instead of through a pageable staging path (CuPy's
 * pinned memory pool plays this part for the reference).                 */
int dyf_host_alloc(void **ptr, size_t bytes);
int dyf_host_free(void *ptr);
int dyf_mem_info(size_t *free_bytes, size_t *total_bytes);
int dyf_synchronize(void);
int dyf_stream_sync(void *stream);
/* Device timers on a stream (hipEvent): kernels of this library are
 * timed on the stream they are launched on.                              */
int dyf_event_create(void **event);
int dyf_event_record(void *event, void *stream);
int dyf_event_elapsed_ms(void *start, void *stop, float *ms);
int dyf_event_destroy(void *event);
/* Device-to-device copy by a kernel (16-byte vector loads/stores), usable
 * under graph capture and as the byte-count calibration for HBM traffic
 * counters.  `bytes` must be a multiple of 16 and both pointers 16-byte
 * aligned.  No counterpart in the reference.                              */
int dyf_copy_kernel(void *dst_dev, const void *src_dev, size_t bytes,
                    void *stream);

/* ---- Volterra displacement phase ------------------------------------ */
/* One straight dislocation segment, lab frame, metres: ra[3], rb[3], b[3]
 * (9 doubles per segment, the reference's (ra, rb, b) tuple).            */
typedef struct {
    int32_t Nx, Nz;   /* x columns, depth planes                  */
    double dx, dz;    /* x pitch, depth step (m)                  */
} dyf_volume;

/* H = 2 pi g.u on rows y_values[0..ny) of a (ny, Nz, Nx) block, x =
 * (ix - Nx/2) dx, z = k dz.  out_prec DYF_FP64 writes double and sums
 * segments in FP64; DYF_FP32 writes float and accumulates each segment's
 * FP64 value in FP32.  Replaces compute_H_grid (dislocations.py:113-168)
 * and H_plane (dislocations.py:171-211).                                  */
int dyf_volterra_h(const double *segs, int nseg, const double *g_vec,
                   double nu, double a_core, const dyf_volume *vol,
                   const double *y_values, int ny, int out_prec,
                   void *H_dev, void *stream);

/* ---- Laue march (laue/kernels.py:48-352, laue/solver.py:34-118) ----- */
typedef struct {
    double s0r, s0i, shr, shi, sbr, sbi;  /* sigma_0, sigma_h, sigma_hbar (1/m) */
    double dz, ds, dx, shift_px;          /* depth step, path step, pitch, shift */
    double sheet_hw;  /* >0: top-hat half width; <=0: minus Gaussian sigma (m) */
    double fov_hy;    /* rows with |(iy - Ny/2) dx| >= fov_hy are zero      */
    int32_t Nz, Nx, Ny, reserved;
} dyf_laue_args;

/* The 12 closed-form scalars of analytic_line_params (dislocations.py:214-242). */
typedef struct {
    double y0, z0, e1y, e1z, e2y, e2z, c_atan, ge1, ge2, b_perp, nu, a2;
} dyf_line_params;

/* Which kernel path a grid takes, its columns per lane, and the marched
 * row band [iy_lo, iy_hi) (FoV test evaluated in FP64).                  */
int dyf_laue_plan(const dyf_laue_args *a, int prec, int *path,
                  int *cols_per_lane, int *iy_lo, int *iy_hi);
/* Bytes of the phase buffer dyf_laue_march reads for (mode, prec): the
 * e^{iH} table on the fast path, 0 on the general path (which reads H).  */
int dyf_laue_phase_bytes(const dyf_laue_args *a, int mode, int prec,
                         size_t *bytes);
/* Bytes of the per-call coefficient workspace for n_e deviations.        */
int dyf_laue_coef_bytes(const dyf_laue_args *a, int n_e, size_t *bytes);
/* Fast path: e^{iH} table from an H grid (Ny, Nz, Nx), H_prec = DYF_FP64
 * (double) or DYF_FP32 (float).  The H_gpu= path of solve_dislocation.    */
int dyf_laue_phase_from_h(const dyf_laue_args *a, int prec, const void *H_dev,
                          int H_prec, void *phase_dev, void *stream);
/* Fast path: e^{iH} table evaluated from segments with no H grid stored
 * (FP64 Volterra, 2 pi reduction in FP64 for the FP32/mixed table).
 * Entries outside the Borrmann fan of a->sheet_hw's incident beam, where
 * the march's fields are exact zeros, are not evaluated: those before a
 * lane's first step in the fan are left unwritten, the lane padding past
 * Nx holds 0.  The table is for marches with the same beam
 * (dyf_laue_march never reads the unwritten entries).                   */
int dyf_laue_phase_from_segments(const dyf_laue_args *a, int prec,
                                 const double *segs, int nseg,
                                 const double *g_vec, double nu,
                                 double a_core, void *phase_dev, void *stream);
/* Fast path, analytic mode: e^{iH(y, z_k)} per marched row and step.     */
int dyf_laue_phase_analytic(const dyf_laue_args *a, int prec,
                            const dyf_line_params *line, void *phase_dev,
                            void *stream);
/* March n_e deviations (s_dev_dev: n_e doubles on the device) and write
 * the exit waves out_dev (n_e, Ny, Nx) interleaved complex: complex128 for
 * DYF_FP64, complex64 otherwise.  `src_dev` is the phase table (fast path)
 * or the H grid with H_prec (general path, GRID mode); `line` is used in
 * ANALYTIC mode on the general path.  Replaces the RawKernel launches of
 * laue/solver.py:73-86 and 107-116.                                       */
int dyf_laue_march(const dyf_laue_args *a, int mode, int prec,
                   const void *src_dev, int H_prec,
                   const dyf_line_params *line, const double *s_dev_dev,
                   int n_e, void *coef_dev, void *out_dev, void *stream);
/* The fraction of the launch's pixel-steps (rows x Nx x (Nz-1) x n_e)
 * that dyf_laue_march executes: below 1 when it marches two angles per
 * wavefront while the beam's fan fits half a wavefront (the rest are
 * exact zeros it does not compute), else 1.  No counterpart in the
 * reference (which computes every pixel-step); for roofline accounting. */
int dyf_laue_march_fraction(const dyf_laue_args *a, int mode, int prec,
                            int n_e, double *fraction);
/* dyf_laue_march without the coefficient workspace: the same kernels and
 * results, with the per-(angle, step) table in a library-held device array
 * of dyf_laue_march_noalloc_max_bytes() (4 MiB: n_e x (Nz-1) up to 87,381)
 * per GPU, kept until the process ends.  Makes no allocation.  Larger
 * calls fail with DYF_EINVAL; calls that overlap on one GPU race on the
 * array.  No counterpart in the reference: it is for callers that cannot
 * allocate (stream capture, serving loops, allocation-free harnesses).   */
int dyf_laue_march_noalloc(const dyf_laue_args *a, int mode, int prec,
                           const void *src_dev, int H_prec,
                           const dyf_line_params *line,
                           const double *s_dev_dev, int n_e, void *out_dev,
                           void *stream);
size_t dyf_laue_march_noalloc_max_bytes(void);

/* ---- Batched Laue solves (persistent launch) -------------------------- */
/* One instance: its grid and crystal, segments (host), angles and output.
 * Batches of these run as one persistent march (no reference
 * counterpart: the reference launches one grid per solve,
 * laue/solver.py:62-90).                                                  */
typedef struct {
    dyf_laue_args grid;
    const double *segs;       /* (nseg, 9) host: ra, rb, b per segment   */
    int32_t nseg, n_e;
    double g_vec[3];
    double nu, a_core;
    const double *s_dev;      /* n_e doubles, device                     */
    void *out;                /* (n_e, Ny, Nx) complex, device           */
} dyf_laue_instance;

typedef struct {
    int32_t waves_per_block;  /* 1, 2 or 4 wavefronts per workgroup      */
    int32_t blocks_per_cu;    /* persistent workgroups per CU            */
} dyf_march_tune;

/* Workspace (e^{iH} + coefficient tables) and meta (segments,
 * descriptors, queue counters) sizes of a batch.                          */
int dyf_laue_batch_bytes(const dyf_laue_instance *inst, int n, int prec,
                         size_t *work_bytes, size_t *meta_bytes);
/* Upload segment parameters and instance descriptors (synchronous).      */
int dyf_laue_batch_prepare(const dyf_laue_instance *inst, int n, int prec,
                           void *work_dev, size_t work_bytes, void *meta_dev,
                           size_t meta_bytes);
/* Enqueue the batch: stages bit 1 = e^{iH} tables, 2 = coefficients and
 * out-of-FoV zero rows, 4 = the persistent march.  Kernels only, so the
 * call can be captured in a graph (dyf_graph_begin/end).                 */
int dyf_laue_batch_run(const dyf_laue_instance *inst, int n, int prec,
                       void *work_dev, void *meta_dev,
                       const dyf_march_tune *tune, int stages, void *stream);
/* hipGraph capture of the work enqueued on `stream` between begin and
 * end; launch replays it.                                                 */
int dyf_graph_begin(void *stream);
int dyf_graph_end(void *stream, void **graph_exec);
int dyf_graph_launch(void *graph_exec, void *stream);
int dyf_graph_destroy(void *graph_exec);
int dyf_stream_create(void **stream);
int dyf_stream_destroy(void *stream);

/* ---- Bragg shooting solve (bragg/solver.py:28-60, 135-379) ---------- */
typedef struct {
    double s0r, s0i, shr, shi, sbr, sbi;  /* sigma_0, sigma_h, sigma_hbar */
    double ds;                            /* path step dz / sin(theta_B)  */
    int32_t Nz, Nx, nyc;                  /* depth planes, columns, y-planes */
    int32_t per_system_phase;  /* 0: one E (Nz, nyc, Nx) shared by all
                                  systems; 1: E is (nsys, Nz, nyc, Nx), one
                                  table per system (different dislocation
                                  fields on the same grid)              */
    int32_t plane_systems;     /* 0: each system is nyc planes of its table.
                                  P > 0 (needs nyc = 1): each system is ONE
                                  plane of a P-plane table, system s = plane
                                  s % P of table s / P (or of the shared
                                  table), so every plane gets its own Krylov
                                  space in dyf_bragg_solve               */
    int32_t reserved;          /* 0 */
} dyf_bragg_args;

typedef struct {
    double tol;        /* SciPy rtol: stop when ||b - A x|| <= tol ||b||   */
    int32_t restart;   /* Krylov dimension of one cycle                    */
    int32_t maxiter;   /* restart cycles                                   */
} dyf_gmres_opts;

typedef struct {
    double residual;   /* final ||b - A x|| / ||b|| (true residual)         */
    int32_t iterations;/* inner iterations (SciPy's pr_norm callbacks)      */
    int32_t code;      /* 0 converged, maxiter otherwise (SciPy's info)     */
} dyf_gmres_info;

/* e^{iH} on the (Nz, ny, Nx) chunk planes y_values[0..ny), x = (ix - Nx/2)
 * dx, z = k dz (the stacked H_plane of bragg/solver.py:256-261).          */
int dyf_bragg_phase(const double *segs, int nseg, const double *g_vec,
                    double nu, double a_core, const dyf_volume *vol,
                    const double *y_values, int ny, void *E_dev,
                    void *stream);
/* Dg(t) of the down-march for nsys systems over the chunk: D0(0) =
 * incident_dev (Nx complex, broadcast; NULL = 0), Dg(0) = Dg0_dev
 * (nsys, nyc, Nx; NULL = 0).  One system per s_dev_dev[s]
 * (_march_down, bragg/solver.py:28-35).                                   */
int dyf_bragg_march(const dyf_bragg_args *a, const double *s_dev_dev,
                     int nsys, const void *E_dev, const void *incident_dev,
                     const void *Dg0_dev, void *out_dev, void *stream);
/* w = M v: ifft(fft(v)/T22s) per plane with the floored perfect-crystal
 * symbol (perfect_symbol + precon, bragg/solver.py:38-60, 163-168,
 * 272-275), one fused transform kernel; pinv_dev (dyf_bragg_symbol_bytes)
 * receives 1/(T22s Nx) (nsys x Nx complex) and, when Nx has a prime factor
 * above 13, the spectrum of the equivalent convolution after it.         */
int dyf_bragg_precondition(const dyf_bragg_args *a, const double *s_dev_dev,
                           int nsys, const void *v_dev, void *w_dev,
                           void *pinv_dev, void *stream);
/* Bytes of dyf_bragg_precondition's symbol workspace for nsys systems
 * (the symbol the reference forms per call, bragg/solver.py:233-238).    */
int dyf_bragg_symbol_bytes(const dyf_bragg_args *a, int nsys, size_t *bytes);
int dyf_bragg_workspace_bytes(const dyf_bragg_args *a, int nsys, int restart,
                              size_t *bytes);
/* Restarted, left-preconditioned GMRES on A v = -d for nsys systems at
 * once (one per s_dev), A v = march(0, v), d = march(incident, 0): the
 * scipy.sparse.linalg.gmres call of bragg/solver.py:330-334 with the
 * Arnoldi basis, orthogonalisation and Givens rotations on the device.
 * x0_dev (nsys, nyc, Nx) warm start or NULL; x_dev receives v.            */
int dyf_bragg_solve(const dyf_bragg_args *a, const double *s_dev_dev,
                    int nsys, const void *E_dev, const void *incident_dev,
                    const dyf_gmres_opts *opts, const void *x0_dev,
                    void *x_dev, void *work_dev, size_t work_bytes,
                    dyf_gmres_info *info, void *stream);
/* Exit wave v e^{-iH(z=0)} (bragg/solver.py:365).                        */
int dyf_bragg_exit_wave(const dyf_bragg_args *a, int nsys, const void *E_dev,
                        const void *v_dev, void *out_dev, void *stream);

/* The fixed-point (trapezoid round-trip) solve, solve_bragg
 * (bragg/solver.py:63-132): forward march of D0 with Dg frozen, backward
 * march of Dg with D0 frozen, until max|Dg(0) - previous| / max|Dg(0)| <
 * tol or n_iter sweeps.  One workgroup per (system, plane) runs every
 * iteration in one launch.  incident_dev: D0 at z = 0 (Nx complex), NULL
 * for the reference's uniform 1.  work_dev: dyf_bragg_fixed_point_bytes
 * (the D0 and Dg fields of every system).  out_dev (nsys, nyc, Nx) gets
 * the exit wave Dg(0) e^{-iH(0)}; info (host, nsys * nyc entries) gets
 * each plane's iterations, last residual and convergence.  Synchronous. */
typedef struct {
    double tol;
    int32_t n_iter;
    int32_t reserved;
} dyf_fixed_opts;

typedef struct {
    double residual;      /* NaN when only one sweep pair ran */
    int32_t iterations;
    int32_t converged;
} dyf_fixed_info;

int dyf_bragg_fixed_point_bytes(const dyf_bragg_args *a, int nsys,
                                size_t *bytes);
int dyf_bragg_fixed_point(const dyf_bragg_args *a, const double *s_dev_dev,
                          int nsys, const void *E_dev,
                          const void *incident_dev, const dyf_fixed_opts *opt,
                          void *work_dev, size_t work_bytes, void *out_dev,
                          dyf_fixed_info *info, void *stream);

/* ---- Exit wave -> image (optics.py:55-207, analysis/rocking.py:37-83) */
/* Fields, spectra, pupils and images are (Ny, Nx) row-major, complex
 * interleaved, batches contiguous; `prec` DYF_FP64 = complex128/float64,
 * otherwise complex64/float32.  Pupils are on the fftshifted frequency
 * grid, as the reference's are.                                         */
typedef struct {
    int32_t Nx, Ny;
    double dx;            /* object pixel (m)                             */
    double aperture_f;    /* NA / lambda: the disc radius (1/m)           */
    double ax, ay;        /* aperture centre offset (1/m)                 */
    double sx, sy;        /* spiral singularity offset (1/m)              */
    int32_t ell;          /* spiral charge                                */
    int32_t levels;       /* phase steps of a quantised plate (0: none)   */
    int32_t zero_dc;      /* 1: zero the bin (Ny/2, Nx/2) (ell != 0 only) */
    int32_t reserved;
} dyf_pupil_args;

/* Optics.pupil (optics.py:67-113): the NA disc times e^{i ell phi}. */
int dyf_optics_pupil(const dyf_pupil_args *a, int prec, void *out_dev,
                     void *stream);
/* object_spectrum (optics.py:172-182): fftshift(fft2(fftshift(f))) of n
 * fields; work_dev holds n planes of the field precision.              */
int dyf_optics_spectrum(const void *field_dev, int n, int Nx, int Ny,
                        int prec, void *work_dev, void *spec_dev,
                        void *stream);
/* image_from_spectrum (optics.py:185-194): |fftshift(ifft2(ifftshift(S P)))|^2
 * for n spectra, one pupil shared (pupil_per_item = 0) or one per item;
 * work_dev holds 2 n planes of the product precision (FP64 if either
 * input is).                                                            */
int dyf_optics_image_from_spectrum(const void *spec_dev, int n,
                                   int spec_prec, const void *pupil_dev,
                                   int pupil_per_item, int pupil_prec,
                                   int Nx, int Ny, void *work_dev,
                                   void *out_dev, void *stream);
/* Optics.image / images, apply_pupils, image_intensity (optics.py:144-
 * 207): n fields through npup pupils, out (n, npup, Ny, Nx) real, one
 * forward transform per field and no shift copies.  work_dev:
 * dyf_optics_image_bytes.                                               */
int dyf_optics_image_bytes(int n, int npup, int field_prec, int pupil_prec,
                           int Nx, int Ny, size_t *bytes);
int dyf_optics_image(const void *field_dev, int n, int field_prec,
                     const void *pupils_dev, int npup, int pupil_prec,
                     int Nx, int Ny, void *work_dev, void *out_dev,
                     void *stream);
/* com_maps' accumulation (analysis/rocking.py:66-75): sum_wI += sum_i
 * w[i] I[i], sum_I += sum_i I[i] (FP64, per pixel, in batch order), and
 * image_sums[i] = sum over I[i] (FP64; NULL to skip).                   */
int dyf_optics_com_accumulate(const void *images_dev, int n, int prec,
                              int Nx, int Ny, const double *weights_dev,
                              double *sum_wI_dev, double *sum_I_dev,
                              double *image_sums_dev, void *stream);

/* ---- Learned warm start: BF16-on-MFMA network kernels --------------- */
/* No counterpart in the reference.  Tensors are
 * NHWC; bf16 is stored as uint16 (bfloat16 bits).  A layer reads a channel
 * slice [off, off+C) of a buffer with Ctot channels.                     */
/* C(i,j) = sum_k A(i,k) B(j,k), A(i,k) = A[i a_i + k a_k] (bf16),
 * B(j,k) = B[j b_j + k b_k] (bf16), FP32 accumulation on MFMA.  mode 0:
 * bf16 store of relu?(acc + bias) at C[i c_i + j]; 1: fp32 store
 * (+ bias); 2: fp32 atomic add (k_splits > 1 allowed).                   */
int dyf_nn_gemm(const void *A, long long a_i, long long a_k, const void *B,
                long long b_j, long long b_k, int M, int N, int K, int mode,
                int relu, const float *bias, void *C, long long c_i,
                int k_splits, void *stream);
int dyf_nn_im2col3(const void *X, int B, int H, int W, int Ctot, int off,
                   int C, void *col, void *stream);
int dyf_nn_col2im3(const float *dcol, int B, int H, int W, int C, float *dX,
                   int Ctot, int off, void *stream);
int dyf_nn_slice(const void *X, long long P, int Ctot, int off, int C,
                 void *out, void *stream);
int dyf_nn_relu_grad(const float *dY, const void *Y, long long P, int Ctot,
                     int off, int C, int relu, void *dZ, void *stream);
int dyf_nn_bias_grad(const void *dZ, long long P, int C, float *db,
                     void *stream);
int dyf_nn_pool(const void *X, int B, int H, int W, int Ctot, int off, int C,
                void *Y, void *stream);
int dyf_nn_pool_grad(const float *dY, int B, int H, int W, int C, float *dX,
                     int Ctot, int off, void *stream);
int dyf_nn_up(const void *X, int B, int H, int W, int C, void *Y, int Ctot,
              int off, void *stream);
int dyf_nn_up_grad(const float *dY, int B, int H, int W, int C, int Ctot,
                   int off, float *dX, void *stream);
int dyf_nn_mse(const float *out, const float *target, long long n,
               float *grad, float *loss, void *stream);
int dyf_nn_adam(float *w, const float *g, float *m, float *v, void *w_bf16,
                long long n, float lr, float b1, float b2, float eps, int t,
                void *stream);
int dyf_nn_to_bf16(const float *x, long long n, void *y, void *stream);
int dyf_nn_fill(float *x, long long n, float v, void *stream);
/* Network input from Bragg e^{iH} stacks (nsys, Nz, nyc, Nx) complex128:
 * Re/Im of the top nk depth slices on columns [x0, x0+Wn), plus the
 * normalised s_dev and thickness planes, written as bf16 NHWC rows
 * y_off.. of samples first..first+nsys of a (*, Hn, Wn, 2nk+2) tensor.  */
int dyf_nn_bragg_features(const void *E, int nsys, int Nz, int nyc, int Nx,
                          int x0, int Wn, int nk, const float *s_norm,
                          const float *t_norm, void *out, int Hn, int y_off,
                          int first, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* DYFOX_H */
