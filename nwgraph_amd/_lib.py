"""ctypes binding of libnwgraph_hip.so (the C ABI declared in include/nwgraph_hip.h).

The library is the only compute path: if it is missing this module raises
ImportError and nothing falls back to a CPU implementation.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnwgraph_hip.so")

NWG_OK = 0
NWG_EINVAL = -22
NWG_ENOMEM = -12
NWG_EHIP = -5
NWG_ENOTSUP = -95
NWG_EFORMAT = -74

NWG_MEM_HOST = 0
NWG_MEM_DEVICE = 1
NWG_INFO_KERNEL_TIMING = 1
NWG_INFO_RELATIVE_THRESHOLD = 2

c_u32 = ctypes.c_uint32
c_u64 = ctypes.c_uint64
c_int = ctypes.c_int
c_float = ctypes.c_float
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p
P_u32 = ctypes.POINTER(c_u32)
P_u64 = ctypes.POINTER(c_u64)
P_float = ctypes.POINTER(c_float)
P_double = ctypes.POINTER(c_double)


class PrInfo(ctypes.Structure):
    _fields_ = [
        ("iterations", c_u64),
        ("final_error", c_double),
        ("ms_total", c_double),
        ("ms_per_iter", c_double),
        ("error_trace", P_double),
        ("error_trace_len", c_u64),
        ("flags", c_u32),
        ("sweep_ms_avg", c_double),
        ("time_trace", P_double),
        ("time_trace_len", c_u64),
        ("init_ms", c_double * 2),
        ("threshold_used", c_double),
    ]


class PlanParams(ctypes.Structure):
    _fields_ = [
        ("short_max", c_u32),
        ("group_max", c_u32),
        ("hub_min", c_u32),
        ("hub_segment", c_u32),
        ("slice_min", c_u32),
        ("stream_hints", c_u32),
        ("ppr_tile", c_u32),
        ("relabel", c_u32),
        ("binned", c_u32),
        ("bin_bits", c_u32),
    ]

    FIELDS = ("short_max", "group_max", "hub_min", "hub_segment", "slice_min", "stream_hints", "ppr_tile",
              "relabel", "binned", "bin_bits")

    def as_dict(self):
        return {k: int(getattr(self, k)) for k in self.FIELDS}


# name -> (restype, argtypes)
_SIGNATURES = {
    "nwg_version": (c_int, []),
    "nwg_create": (c_int, [ctypes.POINTER(c_void_p), c_int]),
    "nwg_destroy": (None, [c_void_p]),
    "nwg_set_stream": (c_int, [c_void_p, c_void_p]),
    "nwg_last_error": (ctypes.c_char_p, [c_void_p]),
    "nwg_pagerank_f32": (c_int, [c_void_p, c_u32, c_u64, c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_float,
                                 c_u64, ctypes.POINTER(PrInfo)]),
    "nwg_pagerank_f32_noalloc": (c_int, [c_void_p, c_u32, c_u64, c_void_p, c_void_p, c_void_p, c_void_p, c_float,
                                         c_float, c_u64, ctypes.POINTER(PrInfo)]),
    "nwg_graph_create_csc": (c_int, [c_void_p, c_u32, c_u64, c_void_p, c_void_p, c_void_p, c_int,
                                     ctypes.POINTER(c_void_p)]),
    "nwg_graph_build_from_edges": (c_int, [c_void_p, c_u32, c_u64, c_void_p, c_void_p, c_int, c_int,
                                           ctypes.POINTER(c_void_p)]),
    "nwg_graph_destroy": (None, [c_void_p]),
    "nwg_graph_info": (c_int, [c_void_p, P_u32, P_u64]),
    "nwg_graph_get_csc": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "nwg_graph_get_csr": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "nwg_graph_device_csc": (c_int, [c_void_p, ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
                                     ctypes.POINTER(c_void_p)]),
    "nwg_generate_rmat": (c_int, [c_void_p, c_u32, c_u32, c_u32, c_u32, c_u32, c_u64, ctypes.POINTER(c_void_p)]),
    "nwg_generate_rmat_edges": (c_int, [c_void_p, c_u32, c_u32, c_u32, c_u32, c_u64, c_u64, c_u64, c_void_p,
                                        c_void_p, c_int]),
    "nwg_generate_grid": (c_int, [c_void_p, c_u32, c_u32, c_u32, c_u64, ctypes.POINTER(c_void_p)]),
    "nwg_generate_grid_edges": (c_int, [c_void_p, c_u32, c_u32, c_u32, c_u64, c_void_p, c_void_p, c_int]),
    "nwg_grid_slots": (c_u64, [c_u32]),
    "nwg_generate_ppr_seeds": (c_int, [c_u32, c_u32, c_u32, c_u64, c_void_p]),
    "nwg_graph_pagerank": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_float, c_u64, c_int,
                                   ctypes.POINTER(PrInfo)]),
    "nwg_graph_pagerank_inplace": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_float, c_u64, c_int,
                                           ctypes.POINTER(PrInfo)]),
    "nwg_graph_pagerank_delta": (c_int, [c_void_p, c_void_p, c_void_p, c_float, c_float, c_u64, c_int,
                                         ctypes.POINTER(PrInfo)]),
    "nwg_graph_set_plan": (c_int, [c_void_p, c_void_p, ctypes.POINTER(PlanParams)]),
    "nwg_graph_get_plan": (c_int, [c_void_p, ctypes.POINTER(PlanParams)]),
    "nwg_graph_autotune": (c_int, [c_void_p, c_void_p, ctypes.POINTER(PlanParams)]),
    "nwg_ppr_batched": (c_int, [c_void_p, c_void_p, c_u32, c_void_p, c_void_p, c_void_p, c_float, c_float, c_u64,
                                c_int, c_void_p, ctypes.POINTER(PrInfo)]),
    "nwg_ppr_autotune": (c_int, [c_void_p, c_void_p, c_u32, c_void_p, c_void_p, P_u32]),
    "nwg_multigraph_create":(c_int, [c_void_p, c_u32, ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p)]),
    "nwg_multigraph_destroy": (None, [c_void_p]),
    "nwg_multigraph_set_plan": (c_int, [c_void_p, c_void_p, ctypes.POINTER(PlanParams)]),
    "nwg_multigraph_autotune": (c_int, [c_void_p, c_void_p, P_u32]),
    "nwg_multigraph_info": (c_int, [c_void_p, P_u32, P_u64, P_u64]),
    "nwg_multigraph_pagerank": (c_int, [c_void_p, c_void_p, c_void_p, c_float, c_float, c_u64, c_int, c_void_p,
                                        c_void_p, ctypes.POINTER(PrInfo)]),
    "nwg_graph_set_weights": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int]),
    "nwg_graph_random_weights": (c_int, [c_void_p, c_void_p, c_u64, c_int]),
    "nwg_graph_get_weights": (c_int, [c_void_p, c_void_p, c_void_p]),
    "nwg_wpr_forward": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_float, c_u64, c_int,
                                ctypes.POINTER(PrInfo)]),
    "nwg_wpr_backward": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_float, c_u64,
                                 c_int, c_int, ctypes.POINTER(PrInfo)]),
    "nwg_stream_copy": (c_int, [c_void_p, c_void_p, c_void_p, c_u64, P_double]),
    "nwg_stream_probe": (c_int, [c_void_p, c_u32, c_u32, c_u32, c_void_p, c_void_p, c_u64, P_double]),
    "nwg_read_mm":(c_int, [ctypes.c_char_p, P_u32, P_u64, ctypes.POINTER(P_u32), ctypes.POINTER(P_u32)]),
    "nwg_read_nw": (c_int, [ctypes.c_char_p, P_u32, P_u64, ctypes.POINTER(P_u32), ctypes.POINTER(P_u32)]),
    "nwg_write_nw": (c_int, [ctypes.c_char_p, c_u32, c_u64, c_void_p, c_void_p]),
    "nwg_read_adjacency": (c_int, [ctypes.c_char_p, P_u32, P_u64, ctypes.POINTER(P_u32), ctypes.POINTER(P_u32)]),
    "nwg_write_adjacency": (c_int, [ctypes.c_char_p, c_u32, c_u64, c_void_p, c_void_p]),
    "nwg_free": (None, [c_void_p]),
}

_lib = None


def load():
    """Load libnwgraph_hip.so once; raise ImportError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    # One HIP runtime per process: torch bundles its own libamdhip64.  When torch is
    # installed, load it first so that the library and torch share torch's runtime
    # (loading the system runtime first leaves torch unable to initialise its device).
    try:
        import torch  # noqa: F401
    except Exception:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_names():
    return sorted(_SIGNATURES)


class NwgError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"nwgraph_hip error {code}: {message}")
        self.code = code
