"""Many Laue solves in one enqueue: the batched-ensembles path.

    solver = EnsembleSolver([(xtal, grid, segs, s_dev), ...], precision=FP64)
    solver.run()                  # every instance, graph-replayed
    stacks = solver.outputs       # per instance (n_e, Ny, Nx), device

Instances may differ in crystal, thickness (Nz), width (Nx, Ny), segment
list and rocking-scan length.  Per instance the library builds the fused
Volterra e^{iH} table and the coefficient table; then one persistent
march per column-count bucket drains every (instance, row, angle) item
from XCD-local work queues.  Instances are grouped so each group's tables
fit a memory budget (the whole set of FP64 tables of the 256-instance
configuration is ~400 GB); groups share one workspace and run one after
another.  Each group's enqueue is captured once in a hipGraph and
replayed by `run()`.

The persistent launch shape (waves per workgroup, workgroups per CU) is
picked by `autotune`, which times the march stage of the first group for
each candidate and keeps the fastest (cached per precision for the
process).
"""

import ctypes as C
import time

import numpy as np

from .. import _native as N
from ..backend import FP64, precision as _precision
from ..dislocations import pack_segments
from .solver import laue_args

CANDIDATES = [(1, 4), (1, 8), (2, 4), (4, 2), (1, 12), (4, 3), (1, 16), (4, 4)]
_tuned = {}


class _Group:
    def __init__(self, idx):
        self.idx = list(idx)
        self.structs = None
        self.meta = None
        self.graph = None
        self.work_bytes = 0
        self.meta_bytes = 0


class EnsembleSolver:
    """Prepared batch of Laue instances (see module docstring).

    `budget_bytes` bounds the e^{iH} + coefficient workspace of a group
    (default: 45 % of free device memory).  `keep_outputs=False` writes
    every group into one shared output pool (for throughput runs where a
    callback consumes each group); otherwise each instance keeps its own
    exit-wave array.
    """

    def __init__(self, instances, precision=FP64, budget_bytes=None,
                 keep_outputs=True, use_graph=True, tune=None):
        self.prec = _precision(precision)
        self.instances = list(instances)
        self.use_graph = use_graph
        n = len(self.instances)
        if n == 0:
            raise ValueError("no instances")
        free = N.mem_info()[0]
        self.budget = int(budget_bytes or 0.45 * free)
        self._sd = []
        self._segs = []
        self.outputs = [None] * n
        structs = []
        for xtal, grid, segs, s in self.instances:
            s = np.atleast_1d(np.asarray(s, np.float64))
            sd = N.DeviceArray.from_host(s)
            packed = pack_segments(segs)
            self._sd.append(sd)
            self._segs.append(packed)
            st = N.LaueInstance()
            st.grid = laue_args(grid, xtal)
            st.segs = packed.ctypes.data_as(C.POINTER(C.c_double))
            st.nseg = len(packed)
            st.n_e = len(s)
            g = np.asarray(xtal.g_vec, np.float64)
            for j in range(3):
                st.g_vec[j] = float(g[j])
            st.nu = float(xtal.nu)
            st.a_core = float(xtal.a_core)
            st.s_dev = sd.ptr
            structs.append(st)
        self._structs = structs
        # per-instance sizes, then greedy grouping under the budget
        sizes = []
        for st in structs:
            arr = (N.LaueInstance * 1)(st)
            w, m = C.c_size_t(), C.c_size_t()
            N.call("dyf_laue_batch_bytes", arr, 1, self.prec.code,
                   C.byref(w), C.byref(m))
            sizes.append(w.value)
        groups, cur, acc = [], [], 0
        for i, w in enumerate(sizes):
            if cur and acc + w > self.budget:
                groups.append(cur)
                cur, acc = [], 0
            cur.append(i)
            acc += w
        groups.append(cur)
        self.groups = [_Group(g) for g in groups]
        # outputs
        out_dtype = self.prec.complex
        if keep_outputs:
            for i, (xtal, grid, segs, s) in enumerate(self.instances):
                Nx, Ny = grid.geometry()[:2]
                self.outputs[i] = N.DeviceArray.empty(
                    (len(self._sd[i]), Ny, Nx), out_dtype)
        else:
            per_group = []
            for G in self.groups:
                per_group.append(sum(
                    len(self._sd[i]) * int(np.prod(self.instances[i][1].shape))
                    for i in G.idx))
            pool = N.DeviceArray.empty((max(per_group),), out_dtype)
            for G in self.groups:
                off = 0
                for i in G.idx:
                    Ny, Nx = self.instances[i][1].shape
                    cnt = len(self._sd[i]) * Ny * Nx
                    self.outputs[i] = pool[off:off + cnt].reshape(
                        (len(self._sd[i]), Ny, Nx))
                    off += cnt
            self._pool = pool
        for i, st in enumerate(structs):
            st.out = self.outputs[i].ptr
        # one shared workspace sized for the largest group, meta per group
        for G in self.groups:
            G.structs = (N.LaueInstance * len(G.idx))(*[structs[i] for i in G.idx])
            w, m = C.c_size_t(), C.c_size_t()
            N.call("dyf_laue_batch_bytes", G.structs, len(G.idx),
                   self.prec.code, C.byref(w), C.byref(m))
            G.work_bytes, G.meta_bytes = w.value, m.value
            G.meta = N.DeviceArray.empty((max(m.value, 1),), np.uint8)
        self.work = N.DeviceArray.empty(
            (max(G.work_bytes for G in self.groups),), np.uint8)
        for G in self.groups:
            N.call("dyf_laue_batch_prepare", G.structs, len(G.idx),
                   self.prec.code, C.c_void_p(self.work.ptr),
                   C.c_size_t(self.work.nbytes), C.c_void_p(G.meta.ptr),
                   C.c_size_t(G.meta.nbytes))
        self.tune = tune or _tuned.get(self.prec.name) or (1, 2)
        p = C.c_void_p()
        N.call("dyf_stream_create", C.byref(p))
        self.stream = p.value

    @property
    def pixel_steps(self):
        return sum(g.pixel_steps(len(s)) for _, g, _, s in self.instances)

    def _enqueue(self, G, stages=7, tune=None):
        wpb, bpc = tune or self.tune
        t = N.MarchTune(int(wpb), int(bpc))
        N.call("dyf_laue_batch_run", G.structs, len(G.idx), self.prec.code,
               C.c_void_p(self.work.ptr), C.c_void_p(G.meta.ptr), C.byref(t),
               int(stages), C.c_void_p(self.stream))

    def autotune(self, candidates=CANDIDATES, repeats=2):
        """Time the march stage of the first group for each launch shape
        and keep the fastest; returns {shape: ms}."""
        G = self.groups[0]
        self._enqueue(G, stages=3)
        timings = {}
        for cand in candidates:
            self._enqueue(G, stages=4, tune=cand)           # warm
            N.call("dyf_stream_sync", C.c_void_p(self.stream))
            t0 = time.perf_counter()
            for _ in range(repeats):
                self._enqueue(G, stages=4, tune=cand)
            N.call("dyf_stream_sync", C.c_void_p(self.stream))
            timings[cand] = (time.perf_counter() - t0) / repeats * 1e3
        best = min(timings, key=timings.get)
        self.tune = best
        _tuned[self.prec.name] = best
        self._drop_graphs()
        return timings

    def _drop_graphs(self):
        for G in self.groups:
            if G.graph:
                N.call("dyf_graph_destroy", C.c_void_p(G.graph))
                G.graph = None

    def run(self, on_group=None):
        """Solve every instance; `on_group(indices)` is called after each
        group has finished (its outputs are then valid)."""
        for G in self.groups:
            if self.use_graph:
                if G.graph is None:
                    N.call("dyf_graph_begin", C.c_void_p(self.stream))
                    try:
                        self._enqueue(G)
                    finally:
                        ex = C.c_void_p()
                        N.call("dyf_graph_end", C.c_void_p(self.stream),
                               C.byref(ex))
                    G.graph = ex.value
                N.call("dyf_graph_launch", C.c_void_p(G.graph),
                       C.c_void_p(self.stream))
            else:
                self._enqueue(G)
            if on_group is not None or len(self.groups) > 1:
                N.call("dyf_stream_sync", C.c_void_p(self.stream))
                if on_group is not None:
                    on_group(G.idx)
        N.call("dyf_stream_sync", C.c_void_p(self.stream))
        return self.outputs

    def close(self):
        self._drop_graphs()
        if getattr(self, "stream", None):
            N.call("dyf_stream_destroy", C.c_void_p(self.stream))
            self.stream = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_ensembles(instances, precision=FP64, **kw):
    """Exit-wave stacks of every (xtal, grid, segments, s_dev) instance."""
    solver = EnsembleSolver(instances, precision=precision, **kw)
    try:
        return solver.run()
    finally:
        solver.close()
