// Batched Laue solves: many instances of different thickness, width,
// crystal, segment list and rocking-scan length in one enqueue - one
// fused Volterra e^{iH} table and one coefficient table per instance,
// then ONE persistent march over every (instance, row, angle) item.
//
// Memory: the caller's workspace holds, per instance, its e^{iH} table and
// coefficient table; the (small) segment parameters and instance
// descriptors live in a separate "meta" buffer.  dyf_laue_batch_prepare
// uploads the meta buffer (host -> device, not capturable);
// dyf_laue_batch_run enqueues kernels only, so it can be captured in a
// hipGraph and replayed.
#include <vector>

#include "internal.hpp"

namespace dyf {

// Work items of an instance's row in the persistent march: angle pairs
// where the pair march runs (C <= pair_max_c(prec)), else angles.
static long long batch_units(int prec, int C, int n_e) {
    return C <= pair_max_c(prec) ? (n_e + 1) / 2 : n_e;
}

static size_t al256(size_t n) { return (n + 255) & ~(size_t)255; }

struct InstLayout {
    PhaseLayout L;
    size_t phase_off, coef_off, seg_off;   // phase/coef in work, segs in meta
    int nseg_kept;
    int bucket, slot;                      // column-count bucket, index in it
};

// Per column-count bucket: its descriptors and queue counters in meta.
struct Bucket {
    int C, n;
    size_t desc_off, ctr_off;
    long long items;
};

static int layout_batch(const dyf_laue_instance *in, int n, int prec,
                        std::vector<InstLayout> *out, size_t *work_bytes,
                        size_t *meta_bytes, std::vector<Bucket> *buckets) {
    if (n < 1 || !in) return fail(DYF_EINVAL, "empty batch");
    if (prec != DYF_FP64 && prec != DYF_FP32 && prec != DYF_MIXED)
        return fail(DYF_EKEY, "unknown precision %d", prec);
    out->resize(n);
    size_t w = 0;
    size_t m = 0;
    for (int i = 0; i < n; ++i) {
        InstLayout &l = (*out)[i];
        int rc = phase_layout(&in[i].grid, DYF_MODE_GRID, prec, &l.L);
        if (rc) return rc;
        if (l.L.path != DYF_PATH_FAST)
            return fail(DYF_EINVAL, "instance %d needs the general path "
                        "(interpolating grid or Nx too large)", i);
        if (in[i].n_e < 0 || in[i].nseg < 0)
            return fail(DYF_EINVAL, "instance %d: bad angle or segment count", i);
        l.phase_off = w;
        w += al256(l.L.bytes);
        l.coef_off = w;
        w += al256(laue_coef_bytes(in[i].n_e, l.L.nsteps));
        l.nseg_kept = 0;
    }
    // buckets in the order of the compiled column counts
    buckets->clear();
    for (int j = 0; j < kNumFastCols; ++j) {
        Bucket b{kFastCols[j], 0, 0, 0, 0};
        for (int i = 0; i < n; ++i)
            if ((*out)[i].L.C == b.C) {
                (*out)[i].bucket = (int)buckets->size();
                (*out)[i].slot = b.n++;
                b.items += (long long)(*out)[i].L.nrows * batch_units(prec, b.C, in[i].n_e);
            }
        if (b.n == 0) continue;
        b.desc_off = m;
        m += al256(sizeof(InstDesc) * b.n);
        b.ctr_off = m;
        m += 256;
        buckets->push_back(b);
    }
    for (int i = 0; i < n; ++i) {
        (*out)[i].seg_off = m;
        m += al256(sizeof(SegParams) * (size_t)std::max(in[i].nseg, 1));
    }
    *work_bytes = w;
    *meta_bytes = m;
    return DYF_OK;
}

}  // namespace dyf

using namespace dyf;

extern "C" int dyf_laue_batch_bytes(const dyf_laue_instance *in, int n,
                                    int prec, size_t *work_bytes,
                                    size_t *meta_bytes) {
    std::vector<InstLayout> lay;
    std::vector<Bucket> bk;
    size_t w = 0, m = 0;
    int rc = layout_batch(in, n, prec, &lay, &w, &m, &bk);
    if (rc) return rc;
    if (work_bytes) *work_bytes = w;
    if (meta_bytes) *meta_bytes = m;
    return DYF_OK;
}

extern "C" int dyf_laue_batch_prepare(const dyf_laue_instance *in, int n,
                                      int prec, void *work, size_t work_bytes,
                                      void *meta, size_t meta_bytes) {
    std::vector<InstLayout> lay;
    std::vector<Bucket> bk;
    size_t w = 0, m = 0;
    int rc = layout_batch(in, n, prec, &lay, &w, &m, &bk);
    if (rc) return rc;
    if (!work || !meta || work_bytes < w || meta_bytes < m)
        return fail(DYF_EINVAL, "workspace too small (%zu/%zu, need %zu/%zu)",
                    work_bytes, meta_bytes, w, m);
    if ((rc = ensure_log_table())) return rc;   // before any capture of run
    for (int i = 0; i < n; ++i)
        if (in[i].n_e > 0 && (!in[i].s_dev || !in[i].out))
            return fail(DYF_EINVAL, "instance %d: NULL angles or output", i);
    std::vector<char> host(m, 0);
    std::vector<long long> base(bk.size(), 0);
    for (int i = 0; i < n; ++i) {
        const dyf_laue_instance &I = in[i];
        InstLayout &l = lay[i];
        if (I.nseg > 0) {
            if (!I.segs) return fail(DYF_EINVAL, "instance %d: NULL segments", i);
            std::vector<SegParams> sp(I.nseg);
            l.nseg_kept = prepare_segments(I.segs, I.nseg, I.g_vec, I.nu, sp.data());
            memcpy(host.data() + l.seg_off, sp.data(), sizeof(SegParams) * l.nseg_kept);
        }
        const Bucket &B = bk[l.bucket];
        InstDesc &d = ((InstDesc *)(host.data() + B.desc_off))[l.slot];
        d.phase = (char *)work + l.phase_off;
        d.coef = (char *)work + l.coef_off;
        d.out = I.out;
        d.n_e = I.n_e, d.nrows = l.L.nrows, d.iy_lo = l.L.iy_lo;
        d.nsteps = l.L.nsteps, d.Nx = I.grid.Nx, d.Ny = I.grid.Ny, d.C = l.L.C;
        d.pad_ = l.nseg_kept;
        d.fan_lo = l.L.fan_lo, d.fan_hi = l.L.fan_hi;
        // pair items: two angles per wavefront while the fan fits half a
        // wavefront, for at least an eighth of the steps (else K = 0: the
        // pair's angles one after the other)
        d.pair_K = 0, d.pair_g0 = 0;
        if (l.L.C <= pair_max_c(prec) && I.n_e >= 2) {
            int K = 0, g0 = 0;
            laue_pair_window(l.L, &K, &g0);
            if (K >= 1 && 8 * K >= l.L.nsteps) d.pair_K = K, d.pair_g0 = g0;
        }
        d.dx = I.grid.dx, d.sheet_hw = I.grid.sheet_hw;
        d.n_items = (long long)l.L.nrows * batch_units(prec, l.L.C, I.n_e);
    }
    // item bases follow slot order inside each bucket
    for (size_t b = 0; b < bk.size(); ++b) {
        InstDesc *dd = (InstDesc *)(host.data() + bk[b].desc_off);
        long long acc = 0;
        for (int j = 0; j < bk[b].n; ++j) {
            dd[j].item_base = acc;
            acc += dd[j].n_items;
        }
    }
    DYF_HIP(hipMemcpy(meta, host.data(), m, hipMemcpyHostToDevice));
    return DYF_OK;
}

extern "C" int dyf_laue_batch_run(const dyf_laue_instance *in, int n, int prec,
                                  void *work, void *meta,
                                  const dyf_march_tune *tune, int stages,
                                  void *stream_) {
    hipStream_t stream = (hipStream_t)stream_;
    std::vector<InstLayout> lay;
    std::vector<Bucket> bk;
    size_t w = 0, m = 0;
    int rc = layout_batch(in, n, prec, &lay, &w, &m, &bk);
    if (rc) return rc;
    if (!work || !meta) return fail(DYF_EINVAL, "NULL workspace");
    for (int i = 0; i < n; ++i)
        if (in[i].n_e > 0 && (!in[i].s_dev || !in[i].out))
            return fail(DYF_EINVAL, "instance %d: NULL angles or output", i);
    int wpb = 1, bpc = 2;
    if (tune) wpb = tune->waves_per_block, bpc = tune->blocks_per_cu;
    int dev = 0, cus = 256;
    DYF_HIP(hipGetDevice(&dev));
    DYF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const char *mb = (const char *)meta;
    for (int i = 0; i < n; ++i) {
        const dyf_laue_instance &I = in[i];
        const InstLayout &l = lay[i];
        if (stages & 1) {
            // nseg_kept lives in the uploaded descriptor; recount here
            std::vector<SegParams> sp(std::max(I.nseg, 1));
            const int kept = I.nseg > 0 ? prepare_segments(I.segs, I.nseg, I.g_vec, I.nu, sp.data()) : 0;
            rc = launch_phase_segments_dev(&I.grid, l.L, prec,
                                           (const SegParams *)(mb + l.seg_off),
                                           kept, I.a_core * I.a_core,
                                           (char *)work + l.phase_off, stream);
            if (rc) return rc;
        }
        if (stages & 2) {
            rc = launch_laue_coef(&I.grid, I.s_dev, I.n_e,
                                  (char *)work + l.coef_off, stream);
            if (rc) return rc;
            rc = launch_zero_rows(&I.grid, l.L, prec, I.n_e, I.out, stream);
            if (rc) return rc;
        }
    }
    if (stages & 4) {
        for (const Bucket &B : bk) {
            rc = launch_laue_persistent((const InstDesc *)(mb + B.desc_off),
                                        B.n, B.items, prec, B.C,
                                        (unsigned *)(mb + B.ctr_off), wpb,
                                        cus * bpc, stream);
            if (rc) return rc;
        }
    }
    return DYF_OK;
}

extern "C" int dyf_graph_begin(void *stream) {
    DYF_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
    return DYF_OK;
}

extern "C" int dyf_graph_end(void *stream, void **exec) {
    if (!exec) return fail(DYF_EINVAL, "exec is NULL");
    hipGraph_t g;
    DYF_HIP(hipStreamEndCapture((hipStream_t)stream, &g));
    hipGraphExec_t ge;
    hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    *exec = (void *)ge;
    return DYF_OK;
}

extern "C" int dyf_graph_launch(void *exec, void *stream) {
    DYF_HIP(hipGraphLaunch((hipGraphExec_t)exec, (hipStream_t)stream));
    return DYF_OK;
}

extern "C" int dyf_graph_destroy(void *exec) {
    if (exec) DYF_HIP(hipGraphExecDestroy((hipGraphExec_t)exec));
    return DYF_OK;
}

extern "C" int dyf_stream_create(void **stream) {
    if (!stream) return fail(DYF_EINVAL, "stream is NULL");
    hipStream_t s;
    DYF_HIP(hipStreamCreate(&s));
    *stream = (void *)s;
    return DYF_OK;
}

extern "C" int dyf_stream_destroy(void *stream) {
    if (stream) DYF_HIP(hipStreamDestroy((hipStream_t)stream));
    return DYF_OK;
}
