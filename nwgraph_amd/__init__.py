"""nwgraph_amd -- MI355X (gfx950) PageRank path of NWGraph, driven through the C ABI
of libnwgraph_hip.so.

Host-side mirror of the reference's interface for this path:
  * ``page_rank(graph, degrees, page_rank, damping_factor, threshold, max_iters,
    num_threads)`` -- nw::graph::page_rank (include/nwgraph/algorithms/page_rank.hpp:89-91)
  * ``Adjacency`` -- the CSC container adjacency<1> (indices_, to_be_indexed_)
  * ``build_degrees`` -- bench/gapbs/common.hpp:110-121
  * ``read_mm`` / ``read_nw`` -- load_graph's two formats (bench/gapbs/common.hpp:82-101)
"""
from ._lib import NwgError, load  # noqa: F401
from .graph import (  # noqa: F401
    Adjacency,
    Context,
    DeviceGraph,
    EdgeList,
    build_adjacency,
    build_degrees,
    default_context,
    page_rank,
    ppr_seeds,
    read_mm,
    read_nw,
    write_nw,
)
