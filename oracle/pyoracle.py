"""ctypes binding of the CPU oracle (oracle/libnwg_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker / baseline, never by the product.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# NWG_ORACLE_LIB names another build of the same source (the sanitizer build of the tests)
LIB_PATH = os.environ.get("NWG_ORACLE_LIB") or os.path.join(_HERE, "libnwg_oracle.so")

u32p = ctypes.c_void_p
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
    L = ctypes.CDLL(LIB_PATH)
    V, U32, U64, F, D, I = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float, ctypes.c_double, ctypes.c_int
    sig = {
        "nwo_philox4x32_10": (None, [V, V, V]),
        "nwo_rmat_edges": (None, [U32, U32, U32, U32, U64, U64, U64, V, V]),
        "nwo_grid_slots": (U64, [U32]),
        "nwo_grid_edges": (None, [U32, U32, U32, U64, V, V]),
        "nwo_ppr_sources": (None, [U32, U32, U32, U64, V]),
        "nwo_edge_weight": (F, [U32, U32, U64]),
        "nwo_normal_vector": (None, [U64, U64, V]),
        "nwo_clean_edges": (U64, [U64, V, V]),
        "nwo_fill_sorted": (None, [I, U32, U64, V, V, V, V]),
        "nwo_out_degrees": (None, [U32, V, V, V]),
        "nwo_pr_v11_blocks": (U64, [U32, V, V, V, V, F, F, U64, U32, V]),
        "nwo_pr_v11_parallel": (U64, [U32, V, V, V, V, F, F, U64, I, V, V]),
        "nwo_pr_v10": (U64, [U32, V, V, V, V, F, F, U64, V, I, V, V]),
        "nwo_pr_v1": (U64, [U32, V, V, V, V, D, D, U64]),
        "nwo_pr_v13": (U64, [U32, V, V, V, V, F, F, U64, V]),
        "nwo_fp64_jacobi": (U64, [U32, V, V, V, V, V, D, D, U64, V]),
        "nwo_fp64_adjoint": (U64, [U32, V, V, V, V, V, D, D, U64, V]),
        "nwo_read_mm": (I, [ctypes.c_char_p, V, V, V, V]),
        "nwo_read_nw": (I, [ctypes.c_char_p, V, V, V, V]),
        "nwo_free": (None, [V]),
    }
    for k, (r, a) in sig.items():
        f = getattr(L, k)
        f.restype = r
        f.argtypes = a
    _lib = L
    return L


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(a.ctypes.data)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    o = np.empty(4, np.uint32)
    load().nwo_philox4x32_10(_p(c), _p(k), _p(o))
    return o


RMAT_ABC = (5700, 1900, 1900)


def rmat_edges(scale, first, count, abc=RMAT_ABC, seed=1):
    s = np.empty(count, np.uint32)
    d = np.empty(count, np.uint32)
    load().nwo_rmat_edges(scale, abc[0], abc[1], abc[2], seed, first, count, _p(s), _p(d))
    return s, d


def grid_edges(width, keep=8500, long_range=50, seed=1):
    L = load()
    m = int(L.nwo_grid_slots(width))
    s = np.empty(m, np.uint32)
    d = np.empty(m, np.uint32)
    L.nwo_grid_edges(width, keep, long_range, seed, _p(s), _p(d))
    return s, d


def ppr_sources(n, k, s, seed=1):
    ids = np.empty(k * s, np.uint32)
    load().nwo_ppr_sources(n, k, s, seed, _p(ids))
    return ids.reshape(k, s)


def edge_weight(u, v, seed=1):
    return load().nwo_edge_weight(int(u), int(v), int(seed))


def normal_vector(n, seed=1):
    out = np.empty(n, np.float64)
    load().nwo_normal_vector(n, seed, _p(out))
    return out


def clean_edges(src, dst):
    s = _u32(src).copy()
    d = _u32(dst).copy()
    m = load().nwo_clean_edges(s.size, _p(s), _p(d))
    return s[:m].copy(), d[:m].copy()


def fill_sorted(idx, n, src, dst):
    s, d = _u32(src), _u32(dst)
    ind = np.empty(n + 1, np.uint32)
    col = np.empty(max(s.size, 1), np.uint32)
    load().nwo_fill_sorted(idx, n, s.size, _p(s), _p(d), _p(ind), _p(col))
    return ind, col[: s.size]


def csc(n, src, dst):
    return fill_sorted(1, n, src, dst)


def out_degrees(n, ind, col):
    deg = np.empty(n, np.uint32)
    load().nwo_out_degrees(n, _p(_u32(ind)), _p(_u32(col)), _p(deg))
    return deg


def rmat_csc(scale, edge_factor=16, abc=RMAT_ABC, seed=1):
    n = 1 << scale
    s, d = rmat_edges(scale, 0, edge_factor * n, abc, seed)
    s, d = clean_edges(s, d)
    ind, col = csc(n, s, d)
    return n, ind, col, out_degrees(n, ind, col)


def grid_csc(width, keep=8500, long_range=50, seed=1):
    n = width * width
    s, d = grid_edges(width, keep, long_range, seed)
    s, d = clean_edges(s, d)
    ind, col = csc(n, s, d)
    return n, ind, col, out_degrees(n, ind, col)


def pr_v11(n, ind, col, deg, damping=0.85, threshold=1e-4, max_iters=20, nblocks=1):
    pr = np.empty(n, np.float32)
    tr = np.zeros(max(max_iters, 1), np.float64)
    it = load().nwo_pr_v11_blocks(n, _p(_u32(ind)), _p(_u32(col)), _p(_u32(deg)), _p(pr), damping, threshold,
                                  max_iters, nblocks, _p(tr))
    return pr, int(it), tr[:it].copy()


def baseline_lib(outdir):
    """The CPU baseline's build of this source (oracle/Makefile target `baseline`: the
    reference's -Ofast -march=native, native to the host running it) in outdir, loaded
    with the v11 loop's signature.  Returns (library, flags) or (None, reason)."""
    import subprocess

    try:
        subprocess.run(["make", "-s", "-C", _HERE, "baseline", f"BASEDIR={outdir}"], check=True,
                       capture_output=True, timeout=120)
        L = ctypes.CDLL(os.path.join(outdir, "libnwg_baseline.so"))
    except (OSError, subprocess.SubprocessError) as e:
        return None, f"native build failed ({type(e).__name__})"
    V, U32, U64, F, I = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float, ctypes.c_int
    L.nwo_pr_v11_parallel.restype = U64
    L.nwo_pr_v11_parallel.argtypes = [U32, V, V, V, V, F, F, U64, I, V, V]
    return L, "-Ofast -march=native -fopenmp"


def pr_v11_parallel(n, ind, col, deg, damping=0.85, threshold=1e-4, max_iters=20, nthreads=1, lib=None):
    pr = np.empty(n, np.float32)
    tr = np.zeros(max(max_iters, 1), np.float64)
    sec = np.zeros(max(max_iters, 1), np.float64)
    it = (lib or load()).nwo_pr_v11_parallel(n, _p(_u32(ind)), _p(_u32(col)), _p(_u32(deg)), _p(pr), damping,
                                             threshold, max_iters, nthreads, _p(tr), _p(sec))
    return pr, int(it), tr[:it].copy(), sec[:it].copy()


def pr_v10(n, ind, col, deg, damping=0.85, threshold=1e-4, max_iters=20, teleport=None, nthreads=1):
    pr = np.empty(n, np.float32)
    tr = np.zeros(max(max_iters, 1), np.float64)
    sec = np.zeros(max(max_iters, 1), np.float64)
    tel = None if teleport is None else np.ascontiguousarray(teleport, np.float32)
    it = load().nwo_pr_v10(n, _p(_u32(ind)), _p(_u32(col)), _p(_u32(deg)), _p(pr), damping, threshold, max_iters,
                           _p(tel), nthreads, _p(tr), _p(sec))
    return pr, int(it), tr[:it].copy(), sec[:it].copy()


def pr_v13(n, ind, col, deg, damping=0.85, threshold=1e-4, max_iters=20):
    pr = np.empty(n, np.float32)
    pushed = np.zeros(max(max_iters, 1), np.uint64)
    it = load().nwo_pr_v13(n, _p(_u32(ind)), _p(_u32(col)), _p(_u32(deg)), _p(pr), damping, threshold, max_iters,
                           _p(pushed))
    return pr, int(it), pushed[:it].copy()


def pr_v1(n, ind, col, deg, damping=0.85, threshold=1e-4, max_iters=100):
    pr = np.empty(n, np.float64)
    it = load().nwo_pr_v1(n, _p(_u32(ind)), _p(_u32(col)), _p(_u32(deg)), _p(pr), damping, threshold, max_iters)
    return pr, int(it)


def fp64_jacobi(n, ind, col, deg, w=None, v=None, damping=0.85, tol=1e-13, max_iters=10000):
    x = np.empty(n, np.float64)
    wa = None if w is None else np.ascontiguousarray(w, np.float64)
    va = None if v is None else np.ascontiguousarray(v, np.float64)
    it = load().nwo_fp64_jacobi(n, _p(_u32(ind)), _p(_u32(col)), _p(_u32(deg)), _p(wa), _p(va), damping, tol,
                                max_iters, _p(x))
    return x, int(it)


def fp64_adjoint(n, ind, col, deg, c, w=None, damping=0.85, tol=1e-13, max_iters=10000):
    lam = np.empty(n, np.float64)
    wa = None if w is None else np.ascontiguousarray(w, np.float64)
    ca = np.ascontiguousarray(c, np.float64)
    it = load().nwo_fp64_adjoint(n, _p(_u32(ind)), _p(_u32(col)), _p(_u32(deg)), _p(wa), _p(ca), damping, tol,
                                 max_iters, _p(lam))
    return lam, int(it)


def _read(fn, path):
    L = load()
    n = ctypes.c_uint32()
    ne = ctypes.c_uint64()
    s = ctypes.c_void_p()
    d = ctypes.c_void_p()
    rc = fn(str(path).encode(), ctypes.byref(n), ctypes.byref(ne), ctypes.byref(s), ctypes.byref(d))
    if rc != 0:
        raise IOError(f"oracle reader failed ({rc}) on {path}")
    m = ne.value
    src = np.ctypeslib.as_array(ctypes.cast(s, ctypes.POINTER(ctypes.c_uint32)), shape=(m,)).copy() if m else np.empty(0, np.uint32)
    dst = np.ctypeslib.as_array(ctypes.cast(d, ctypes.POINTER(ctypes.c_uint32)), shape=(m,)).copy() if m else np.empty(0, np.uint32)
    L.nwo_free(s)
    L.nwo_free(d)
    return n.value, src, dst


def read_mm(path):
    return _read(load().nwo_read_mm, path)


def read_nw(path):
    return _read(load().nwo_read_nw, path)


def dense_matrix(n, ind, col, deg, w=None):
    """Dense FP64 A with A[i, u] = w(u->i) (w None => 1/outdeg(u)), from a CSC."""
    A = np.zeros((n, n), np.float64)
    ind = np.asarray(ind, np.int64)
    col = np.asarray(col, np.int64)
    rows = np.repeat(np.arange(n), np.diff(ind))
    vals = (1.0 / np.asarray(deg, np.float64)[col]) if w is None else np.asarray(w, np.float64)
    np.add.at(A, (rows, col), vals)
    return A
