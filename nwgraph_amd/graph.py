"""Host-side objects over the C ABI: contexts, device-resident graphs, and the
reference-shaped entry points (page_rank, build_adjacency, build_degrees, readers).

Every compute call goes to libnwgraph_hip.so; there is no CPU fallback.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (NWG_INFO_KERNEL_TIMING, NWG_INFO_RELATIVE_THRESHOLD, NWG_MEM_DEVICE, NWG_MEM_HOST, NwgError,
                   PlanParams, PrInfo)

RMAT_ABC = (5700, 1900, 1900)  # A/B/C = 0.57/0.19/0.19 in units of 1/10000 (D = 0.05)


def _ptr(a):
    """ctypes pointer of a C-contiguous numpy array (None passes through)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return ctypes.c_void_p(a.ctypes.data)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


class Context:
    """One HIP device + stream (nwg_ctx)."""

    def __init__(self, device=0):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        rc = self.lib.nwg_create(ctypes.byref(h), int(device))
        if rc != 0:
            raise NwgError(rc, f"nwg_create(device={device}) failed (is a HIP device visible?)")
        self.handle = h
        self.device = device

    def check(self, rc):
        if rc != 0:
            msg = self.lib.nwg_last_error(self.handle)
            raise NwgError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "handle", None):
            self.lib.nwg_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_handle):
        """Run subsequent work on a hipStream_t given as an integer handle (0 = own stream)."""
        self.check(self.lib.nwg_set_stream(self.handle, ctypes.c_void_p(stream_handle or None)))

    # ---- graph construction -------------------------------------------------
    def _wrap(self, h):
        return DeviceGraph(self, h)

    def graph_from_csc(self, indices, col, outdeg=None):
        indices, col = _u32(indices), _u32(col)
        n = indices.size - 1
        od = None if outdeg is None else _u32(outdeg)
        h = ctypes.c_void_p()
        self.check(self.lib.nwg_graph_create_csc(self.handle, n, col.size, _ptr(indices), _ptr(col), _ptr(od),
                                                 NWG_MEM_HOST, ctypes.byref(h)))
        return self._wrap(h)

    def graph_from_edges(self, n, src, dst, clean=True):
        src, dst = _u32(src), _u32(dst)
        assert src.size == dst.size
        h = ctypes.c_void_p()
        self.check(self.lib.nwg_graph_build_from_edges(self.handle, int(n), src.size, _ptr(src), _ptr(dst),
                                                       1 if clean else 0, NWG_MEM_HOST, ctypes.byref(h)))
        return self._wrap(h)

    def rmat(self, scale, edge_factor=16, abc=RMAT_ABC, seed=1):
        h = ctypes.c_void_p()
        self.check(self.lib.nwg_generate_rmat(self.handle, scale, edge_factor, abc[0], abc[1], abc[2], seed,
                                              ctypes.byref(h)))
        return self._wrap(h)

    def rmat_edges(self, scale, first, count, abc=RMAT_ABC, seed=1):
        src = np.empty(count, np.uint32)
        dst = np.empty(count, np.uint32)
        self.check(self.lib.nwg_generate_rmat_edges(self.handle, scale, abc[0], abc[1], abc[2], seed, first, count,
                                                    _ptr(src), _ptr(dst), NWG_MEM_HOST))
        return src, dst

    def grid(self, width, keep=8500, long_range=50, seed=1):
        h = ctypes.c_void_p()
        self.check(self.lib.nwg_generate_grid(self.handle, width, keep, long_range, seed, ctypes.byref(h)))
        return self._wrap(h)

    def multigraph(self, graphs):
        return MultiGraph(self, graphs)

    def grid_edges(self, width, keep=8500, long_range=50, seed=1):
        slots = int(self.lib.nwg_grid_slots(width))
        src = np.empty(slots, np.uint32)
        dst = np.empty(slots, np.uint32)
        self.check(self.lib.nwg_generate_grid_edges(self.handle, width, keep, long_range, seed, _ptr(src), _ptr(dst),
                                                    NWG_MEM_HOST))
        return src, dst


class DeviceGraph:
    """A device-resident CSC (nwg_graph): rows are targets, entries are sources."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.handle = handle
        n = ctypes.c_uint32()
        nnz = ctypes.c_uint64()
        ctx.check(ctx.lib.nwg_graph_info(handle, ctypes.byref(n), ctypes.byref(nnz)))
        self.n = n.value
        self.nnz = nnz.value

    def close(self):
        if getattr(self, "handle", None):
            self.ctx.lib.nwg_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def csc(self):
        ptr = np.empty(self.n + 1, np.uint32)
        idx = np.empty(self.nnz, np.uint32)
        deg = np.empty(self.n, np.uint32)
        self.ctx.check(self.ctx.lib.nwg_graph_get_csc(self.ctx.handle, self.handle, _ptr(ptr), _ptr(idx), _ptr(deg)))
        return ptr, idx, deg

    def csr(self):
        ptr = np.empty(self.n + 1, np.uint32)
        idx = np.empty(self.nnz, np.uint32)
        pos = np.empty(self.nnz, np.uint32)
        self.ctx.check(self.ctx.lib.nwg_graph_get_csr(self.ctx.handle, self.handle, _ptr(ptr), _ptr(idx), _ptr(pos)))
        return ptr, idx, pos

    def device_csc(self):
        a, b, c = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        self.ctx.check(self.ctx.lib.nwg_graph_device_csc(self.handle, ctypes.byref(a), ctypes.byref(b),
                                                         ctypes.byref(c)))
        return a.value, b.value, c.value

    # ---- scheduling plan ----------------------------------------------------
    def get_plan(self):
        p = PlanParams()
        self.ctx.check(self.ctx.lib.nwg_graph_get_plan(self.handle, ctypes.byref(p)))
        return p.as_dict()

    def set_plan(self, **kw):
        p = PlanParams()
        self.ctx.check(self.ctx.lib.nwg_graph_get_plan(self.handle, ctypes.byref(p)))
        for k, v in kw.items():
            setattr(p, k, int(v))
        self.ctx.check(self.ctx.lib.nwg_graph_set_plan(self.ctx.handle, self.handle, ctypes.byref(p)))

    def autotune(self):
        p = PlanParams()
        self.ctx.check(self.ctx.lib.nwg_graph_autotune(self.ctx.handle, self.handle, ctypes.byref(p)))
        return p.as_dict()

    # ---- PageRank -----------------------------------------------------------
    @staticmethod
    def _info(max_iters, kernel_timing=False, relative=False):
        info = PrInfo()
        trace = np.zeros(max(int(max_iters), 1), np.float64)
        info.error_trace = trace.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        info.error_trace_len = trace.size
        info.flags = (NWG_INFO_KERNEL_TIMING if kernel_timing else 0) | (NWG_INFO_RELATIVE_THRESHOLD if relative else 0)
        return info, trace

    @staticmethod
    def _info_dict(info, trace):
        it = int(info.iterations)
        return {
            "iterations": it,
            "final_error": float(info.final_error),
            "ms_total": float(info.ms_total),
            "ms_per_iter": float(info.ms_per_iter),
            "sweep_ms_avg": float(info.sweep_ms_avg),
            "threshold_used": float(info.threshold_used),
            "error_trace": trace[:it].copy(),
        }

    def pagerank(self, damping=0.85, threshold=1e-4, max_iters=20, teleport=None, kernel_timing=False,
                 inplace=False):
        """Plain (or personalised) PageRank; returns (scores float32[n], info dict).  inplace=True
        runs the reference's in-place (v11) schedule, which is not reproducible bit for bit."""
        scores = np.empty(self.n, np.float32)
        tele = None if teleport is None else np.ascontiguousarray(teleport, np.float32)
        info, trace = self._info(max_iters, kernel_timing)
        fn = self.ctx.lib.nwg_graph_pagerank_inplace if inplace else self.ctx.lib.nwg_graph_pagerank
        self.ctx.check(fn(self.ctx.handle, self.handle, _ptr(scores), _ptr(tele), float(damping), float(threshold),
                          int(max_iters), NWG_MEM_HOST, ctypes.byref(info)))
        return scores, self._info_dict(info, trace)

    def pagerank_device(self, scores_ptr, teleport_ptr=None, damping=0.85, threshold=1e-4, max_iters=20,
                        kernel_timing=False, inplace=False):
        """PageRank into a device buffer (integer device pointer); returns the info dict."""
        info, trace = self._info(max_iters, kernel_timing)
        fn = self.ctx.lib.nwg_graph_pagerank_inplace if inplace else self.ctx.lib.nwg_graph_pagerank
        self.ctx.check(fn(self.ctx.handle, self.handle, ctypes.c_void_p(scores_ptr),
                          ctypes.c_void_p(teleport_ptr) if teleport_ptr else None, float(damping), float(threshold),
                          int(max_iters), NWG_MEM_DEVICE, ctypes.byref(info)))
        return self._info_dict(info, trace)

    def pagerank_delta(self, damping=0.85, threshold=1e-4, max_iters=20):
        """The reference's delta PageRank (page_rank_v13, pr --version 13): un-normalised
        scores (about N in total); returns (scores float32[n], info dict)."""
        scores = np.empty(self.n, np.float32)
        info, trace = self._info(max_iters)
        self.ctx.check(self.ctx.lib.nwg_graph_pagerank_delta(self.ctx.handle, self.handle, _ptr(scores),
                                                             float(damping), float(threshold), int(max_iters),
                                                             NWG_MEM_HOST, ctypes.byref(info)))
        return scores, self._info_dict(info, trace)

    # ---- batched personalised PageRank -----------------------------------------
    @staticmethod
    def _seed_csr(seeds):
        seeds = [np.asarray(s, np.uint32).ravel() for s in seeds]
        ptr = np.zeros(len(seeds) + 1, np.uint32)
        ptr[1:] = np.cumsum([s.size for s in seeds])
        flat = np.concatenate(seeds).astype(np.uint32) if seeds else np.zeros(0, np.uint32)
        return len(seeds), ptr, np.ascontiguousarray(flat)

    def ppr(self, seeds, damping=0.85, threshold=1e-4, max_iters=20):
        """k personalised PageRank vectors; seeds[j] = seed ids of teleport vector j (uniform over them).
        Returns (scores float32[n, k], iterations uint64[k], info dict)."""
        k, ptr, flat = self._seed_csr(seeds)
        scores = np.empty((self.n, k), np.float32)
        its = np.zeros(k, np.uint64)
        info, trace = self._info(max_iters)
        self.ctx.check(self.ctx.lib.nwg_ppr_batched(self.ctx.handle, self.handle, k, _ptr(ptr), _ptr(flat),
                                                    _ptr(scores), float(damping), float(threshold), int(max_iters),
                                                    NWG_MEM_HOST, _ptr(its), ctypes.byref(info)))
        return scores, its, self._info_dict(info, trace)

    def ppr_device(self, seeds, scores_ptr, damping=0.85, threshold=1e-4, max_iters=20):
        k, ptr, flat = self._seed_csr(seeds)
        its = np.zeros(k, np.uint64)
        info, trace = self._info(max_iters)
        self.ctx.check(self.ctx.lib.nwg_ppr_batched(self.ctx.handle, self.handle, k, _ptr(ptr), _ptr(flat),
                                                    ctypes.c_void_p(scores_ptr), float(damping), float(threshold),
                                                    int(max_iters), NWG_MEM_DEVICE, _ptr(its), ctypes.byref(info)))
        return its, self._info_dict(info, trace)

    def ppr_autotune(self, seeds):
        k, ptr, flat = self._seed_csr(seeds)
        t = ctypes.c_uint32()
        self.ctx.check(self.ctx.lib.nwg_ppr_autotune(self.ctx.handle, self.handle, k, _ptr(ptr), _ptr(flat),
                                                     ctypes.byref(t)))
        return t.value

    # ---- weighted PageRank and adjoint ----------------------------------------
    def set_weights(self, w_csc, normalize=True):
        w = np.ascontiguousarray(w_csc, np.float32)
        assert w.size == self.nnz
        self.ctx.check(self.ctx.lib.nwg_graph_set_weights(self.ctx.handle, self.handle, _ptr(w), 1 if normalize else 0,
                                                          NWG_MEM_HOST))

    def set_weights_device(self, w_ptr, normalize=True):
        self.ctx.check(self.ctx.lib.nwg_graph_set_weights(self.ctx.handle, self.handle, ctypes.c_void_p(w_ptr),
                                                          1 if normalize else 0, NWG_MEM_DEVICE))

    def random_weights(self, seed=1, normalize=True):
        self.ctx.check(self.ctx.lib.nwg_graph_random_weights(self.ctx.handle, self.handle, int(seed),
                                                             1 if normalize else 0))

    def get_weights(self):
        w = np.empty(self.nnz, np.float32)
        self.ctx.check(self.ctx.lib.nwg_graph_get_weights(self.ctx.handle, self.handle, _ptr(w)))
        return w

    def wpr_forward(self, teleport=None, damping=0.85, threshold=1e-4, max_iters=100, relative=False):
        """relative=True: stop when the L1 change falls below threshold * ||teleport||_1."""
        x = np.empty(self.n, np.float32)
        tele = None if teleport is None else np.ascontiguousarray(teleport, np.float32)
        info, trace = self._info(max_iters, relative=relative)
        self.ctx.check(self.ctx.lib.nwg_wpr_forward(self.ctx.handle, self.handle, _ptr(tele), _ptr(x), float(damping),
                                                    float(threshold), int(max_iters), NWG_MEM_HOST,
                                                    ctypes.byref(info)))
        return x, self._info_dict(info, trace)

    def wpr_backward(self, x, grad_x, damping=0.85, threshold=1e-4, max_iters=100, raw_weights=True, relative=False):
        """relative=True: stop when the L1 change falls below threshold * ||grad_x||_1 / (1 - d)."""
        xa = np.ascontiguousarray(x, np.float32)
        ga = np.ascontiguousarray(grad_x, np.float32)
        gw = np.empty(self.nnz, np.float32)
        gv = np.empty(self.n, np.float32)
        info, trace = self._info(max_iters, relative=relative)
        self.ctx.check(self.ctx.lib.nwg_wpr_backward(self.ctx.handle, self.handle, _ptr(xa), _ptr(ga), _ptr(gw),
                                                     _ptr(gv), float(damping), float(threshold), int(max_iters),
                                                     1 if raw_weights else 0, NWG_MEM_HOST, ctypes.byref(info)))
        return gw, gv, self._info_dict(info, trace)

    def wpr_forward_device(self, teleport_ptr, x_ptr, damping=0.85, threshold=1e-4, max_iters=100, relative=False):
        info, trace = self._info(max_iters, relative=relative)
        self.ctx.check(self.ctx.lib.nwg_wpr_forward(self.ctx.handle, self.handle,
                                                    ctypes.c_void_p(teleport_ptr) if teleport_ptr else None,
                                                    ctypes.c_void_p(x_ptr), float(damping), float(threshold),
                                                    int(max_iters), NWG_MEM_DEVICE, ctypes.byref(info)))
        return self._info_dict(info, trace)

    def wpr_backward_device(self, x_ptr, gx_ptr, gw_ptr, gv_ptr, damping=0.85, threshold=1e-4, max_iters=100,
                            raw_weights=True, relative=False):
        info, trace = self._info(max_iters, relative=relative)
        self.ctx.check(self.ctx.lib.nwg_wpr_backward(self.ctx.handle, self.handle, ctypes.c_void_p(x_ptr),
                                                     ctypes.c_void_p(gx_ptr),
                                                     ctypes.c_void_p(gw_ptr) if gw_ptr else None,
                                                     ctypes.c_void_p(gv_ptr) if gv_ptr else None, float(damping),
                                                     float(threshold), int(max_iters), 1 if raw_weights else 0,
                                                     NWG_MEM_DEVICE, ctypes.byref(info)))
        return self._info_dict(info, trace)


class MultiGraph:
    """Block-diagonal union of device graphs solved by one persistent launch per sweep."""

    def __init__(self, ctx, graphs):
        self.ctx = ctx
        self.graphs = list(graphs)
        arr = (ctypes.c_void_p * len(self.graphs))(*[g.handle.value for g in self.graphs])
        h = ctypes.c_void_p()
        ctx.check(ctx.lib.nwg_multigraph_create(ctx.handle, len(self.graphs), arr, ctypes.byref(h)))
        self.handle = h
        ng, nv, ne = ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_uint64()
        ctx.check(ctx.lib.nwg_multigraph_info(h, ctypes.byref(ng), ctypes.byref(nv), ctypes.byref(ne)))
        self.ngraphs, self.n, self.nnz = ng.value, nv.value, ne.value
        self.offsets = np.concatenate([[0], np.cumsum([g.n for g in self.graphs])]).astype(np.int64)

    def set_plan(self, binned):
        """Schedule of the union's sweeps: the per-item persistent kernel (binned=0) or the
        binned sweep of the union (binned=1)."""
        p = PlanParams()
        p.binned = int(binned)
        self.ctx.check(self.ctx.lib.nwg_multigraph_set_plan(self.ctx.handle, self.handle, ctypes.byref(p)))

    def autotune(self):
        """Times both schedules and keeps the faster; returns 1 if the binned sweep won."""
        b = ctypes.c_uint32(0)
        self.ctx.check(self.ctx.lib.nwg_multigraph_autotune(self.ctx.handle, self.handle, ctypes.byref(b)))
        return int(b.value)

    def pagerank(self, damping=0.85, threshold=1e-4, max_iters=20):
        scores = np.empty(self.n, np.float32)
        its = np.zeros(self.ngraphs, np.uint64)
        errs = np.zeros(self.ngraphs, np.float64)
        info, trace = DeviceGraph._info(max_iters)
        self.ctx.check(self.ctx.lib.nwg_multigraph_pagerank(self.ctx.handle, self.handle, _ptr(scores),
                                                            float(damping), float(threshold), int(max_iters),
                                                            NWG_MEM_HOST, _ptr(its), _ptr(errs), ctypes.byref(info)))
        return scores, its, errs, DeviceGraph._info_dict(info, trace)

    def close(self):
        if getattr(self, "handle", None):
            self.ctx.lib.nwg_multigraph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


# ---------------------------------------------------------------------------
# reference-shaped host objects
# ---------------------------------------------------------------------------
class EdgeList:
    """Directed edge list in struct-of-arrays form (edge_list<directed>, edge_list.hpp:107-312).
    n follows the reference rule: max(given n, max id + 1) (edge_list.hpp:188-199)."""

    def __init__(self, n, src, dst):
        self.src = _u32(src)
        self.dst = _u32(dst)
        m = int(max(self.src.max(initial=0), self.dst.max(initial=0))) + 1 if self.src.size else 0
        self.n = max(int(n), m)

    def __len__(self):
        return int(self.src.size)

    def num_vertices(self):
        return self.n


class Adjacency:
    """Host CSC with the layout of nw::graph::adjacency<1> (compressed.hpp:64-102): indices_ has
    N+1 row offsets, to_be_indexed_ the source id of every in-edge."""

    def __init__(self, indices, to_be_indexed):
        self.indices_ = _u32(indices)
        self.to_be_indexed_ = _u32(to_be_indexed)

    def size(self):
        return int(self.indices_.size - 1)

    def __len__(self):
        return self.size()

    def __getitem__(self, i):
        return self.to_be_indexed_[self.indices_[i]:self.indices_[i + 1]]

    def num_edges(self):
        return int(self.to_be_indexed_.size)

    def serialize(self, outfile_name):
        """index_adjacency::serialize (adjacency.hpp:176-180): the binary form the reference
        writes, through nwg_write_adjacency."""
        lib = _lib.load()
        rc = lib.nwg_write_adjacency(str(outfile_name).encode(), self.size(), self.num_edges(), _ptr(self.indices_),
                                     _ptr(self.to_be_indexed_))
        if rc != 0:
            raise NwgError(rc, f"cannot write {outfile_name}")

    def deserialize(self, infile_name):
        """index_adjacency::deserialize (adjacency.hpp:187-191): replaces this adjacency's
        arrays with the file's, through nwg_read_adjacency (which validates the layout)."""
        lib = _lib.load()
        n = ctypes.c_uint32()
        ne = ctypes.c_uint64()
        pp = ctypes.POINTER(ctypes.c_uint32)()
        ip = ctypes.POINTER(ctypes.c_uint32)()
        rc = lib.nwg_read_adjacency(str(infile_name).encode(), ctypes.byref(n), ctypes.byref(ne), ctypes.byref(pp),
                                    ctypes.byref(ip))
        if rc != 0:
            raise NwgError(rc, f"cannot read {infile_name}")
        try:
            self.indices_ = np.ctypeslib.as_array(pp, shape=(n.value + 1,)).copy()
            m = ne.value
            self.to_be_indexed_ = np.ctypeslib.as_array(ip, shape=(m,)).copy() if m else np.empty(0, np.uint32)
        finally:
            lib.nwg_free(ctypes.cast(pp, ctypes.c_void_p))
            lib.nwg_free(ctypes.cast(ip, ctypes.c_void_p))
        return self


def build_adjacency(el, clean=False, ctx=None):
    """build_adjacency<1> (bench/gapbs/common.hpp:103-108) on the device.  clean=False keeps
    self-loops and duplicates like fill_directed (build.hpp:209-232); rows are sorted
    (sort_to_be_indexed), which fixes the reference's unspecified within-row order."""
    ctx = ctx or default_context()
    g = ctx.graph_from_edges(el.n, el.src, el.dst, clean=clean)
    try:
        ptr, idx, _ = g.csc()
    finally:
        g.close()
    return Adjacency(ptr, idx)


def build_degrees(graph, ctx=None):
    """Out-degree of every vertex of a CSC (bench/gapbs/common.hpp:110-121), computed on the device."""
    ctx = ctx or default_context()
    g = ctx.graph_from_csc(graph.indices_, graph.to_be_indexed_)
    try:
        _, _, deg = g.csc()
    finally:
        g.close()
    return deg


def page_rank(graph, degrees, page_rank, damping_factor=0.85, threshold=1e-4, max_iters=20, num_threads=1,
              ctx=None, info=None):
    """Drop-in for nw::graph::page_rank(graph, degrees, page_rank, damping_factor, threshold,
    max_iters, num_threads) (include/nwgraph/algorithms/page_rank.hpp:89-148).

    ``page_rank`` is a caller-sized float32 array written in place.  ``num_threads`` is accepted
    and ignored (the reference ignores it as well; TBB's global_control sets the width).  Runs
    the v10 Jacobi schedule on the GPU.  If ``info`` is a dict it receives iterations and errors.
    """
    ctx = ctx or default_context()
    if not (isinstance(page_rank, np.ndarray) and page_rank.dtype == np.float32 and page_rank.flags["C_CONTIGUOUS"]):
        raise TypeError("page_rank must be a C-contiguous float32 numpy array")
    n = graph.size()
    if page_rank.size != n:
        raise ValueError("page_rank must be pre-sized to graph.size()")
    deg = None if degrees is None else _u32(degrees)
    if deg is not None and deg.size != n:
        raise ValueError("degrees must have graph.size() entries")
    pi, trace = DeviceGraph._info(max_iters)
    ctx.check(ctx.lib.nwg_pagerank_f32(ctx.handle, n, graph.to_be_indexed_.size, _ptr(graph.indices_),
                                       _ptr(graph.to_be_indexed_), _ptr(deg), _ptr(page_rank), float(damping_factor),
                                       float(threshold), int(max_iters), ctypes.byref(pi)))
    if info is not None:
        info.update(DeviceGraph._info_dict(pi, trace))


def ppr_seeds(n, k, s, seed=1):
    """k teleport seed sets of s distinct ids in [0, n) (nwg_generate_ppr_seeds); shape (k, s)."""
    lib = _lib.load()
    ids = np.empty(k * s, np.uint32)
    rc = lib.nwg_generate_ppr_seeds(int(n), int(k), int(s), int(seed), _ptr(ids))
    if rc != 0:
        raise NwgError(rc, "nwg_generate_ppr_seeds: invalid arguments")
    return ids.reshape(k, s)


# ---------------------------------------------------------------------------
# input formats
# ---------------------------------------------------------------------------
def _read(fn, path):
    lib = _lib.load()
    n = ctypes.c_uint32()
    ne = ctypes.c_uint64()
    s = ctypes.POINTER(ctypes.c_uint32)()
    d = ctypes.POINTER(ctypes.c_uint32)()
    rc = fn(str(path).encode(), ctypes.byref(n), ctypes.byref(ne), ctypes.byref(s), ctypes.byref(d))
    if rc != 0:
        raise NwgError(rc, f"cannot read {path}")
    try:
        m = ne.value
        src = np.ctypeslib.as_array(s, shape=(m,)).copy() if m else np.empty(0, np.uint32)
        dst = np.ctypeslib.as_array(d, shape=(m,)).copy() if m else np.empty(0, np.uint32)
    finally:
        lib.nwg_free(ctypes.cast(s, ctypes.c_void_p))
        lib.nwg_free(ctypes.cast(d, ctypes.c_void_p))
    return EdgeList(n.value, src, dst)


def read_mm(path):
    """read_mm<directed> (io/mmio.hpp:171-217)."""
    return _read(_lib.load().nwg_read_mm, path)


def read_nw(path):
    """edge_list<directed>::deserialize (edge_list.hpp:262-280)."""
    return _read(_lib.load().nwg_read_nw, path)


def write_nw(path, el):
    """edge_list<directed>::serialize (edge_list.hpp:244-255)."""
    lib = _lib.load()
    rc = lib.nwg_write_nw(str(path).encode(), el.n, len(el), _ptr(el.src), _ptr(el.dst))
    if rc != 0:
        raise NwgError(rc, f"cannot write {path}")
