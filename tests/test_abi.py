"""CPU checks of the C ABI: the library loads without a GPU and exports every
entry point include/nwgraph_hip.h declares; host-only entry points work."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

HEADER = os.path.join(ROOT, "include", "nwgraph_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nwg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_surface():
    names = declared_functions()
    for required in ("nwg_pagerank_f32", "nwg_graph_pagerank", "nwg_graph_build_from_edges", "nwg_ppr_batched",
                     "nwg_multigraph_pagerank", "nwg_wpr_forward", "nwg_wpr_backward", "nwg_graph_autotune",
                     "nwg_graph_pagerank_inplace", "nwg_graph_pagerank_delta"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from nwgraph_amd import _lib

    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding covers the same surface
    assert sorted(_lib.exported_names()) == declared_functions()


def test_host_only_entry_points():
    from nwgraph_amd import _lib

    lib = _lib.load()
    assert lib.nwg_version() >= 10000
    assert lib.nwg_grid_slots(4096) == 4 * 4096 * 4095 + 4096 * 4096
    # null-pointer misuse is rejected with a status, not a crash
    assert lib.nwg_create(None, 0) == _lib.NWG_EINVAL
    assert lib.nwg_graph_info(None, None, None) == _lib.NWG_EINVAL
    assert lib.nwg_last_error(None) == b"null context"


def test_ppr_seed_generator_matches_oracle():
    import nwgraph_amd
    import pyoracle as O

    for n, k, s, seed in [(1000, 64, 8, 1), (4194304, 64, 8, 1), (9, 3, 9, 7)]:
        assert np.array_equal(nwgraph_amd.ppr_seeds(n, k, s, seed), O.ppr_sources(n, k, s, seed))
    with pytest.raises(nwgraph_amd.NwgError):
        nwgraph_amd.ppr_seeds(4, 1, 5, 1)  # more seeds than vertices


def test_cpp_dropin_header_compiles(tmp_path):
    """The C++ drop-ins (page_rank and the experimental page_rank_v10 / page_rank_v13
    signatures) compile against a graph with adjacency<1>'s layout (no GPU needed)."""
    import shutil
    import subprocess

    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    src = tmp_path / "use.cpp"
    src.write_text(
        "#include <nwgraph_hip/page_rank.hpp>\n"
        "int main() {\n"
        "  nw::graph::hip::csc_graph g;\n"
        "  g.indices_ = {0, 1, 2};\n"
        "  std::get<0>(g.to_be_indexed_) = {1, 0};\n"
        "  std::vector<uint32_t> deg = {1, 1};\n"
        "  std::vector<float> pr(2);\n"
        "  if (false) nw::graph::page_rank(g, deg, pr, 0.85f, 1e-4f, 20, 1);\n"
        "  if (false) nw::graph::page_rank_v10(g, deg, pr, 0.85f, 1e-4f, 20, 1);\n"
        "  if (false) nw::graph::page_rank_v13(g, deg, pr, 0.85f, 1e-4f, 20, 1);\n"
        "  return (int)g.size() - 2;\n"
        "}\n")
    r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-I" + os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_pr_driver_usage():
    import subprocess

    exe = os.path.join(ROOT, "nwgraph_amd", "bin", "pr")
    assert os.path.exists(exe)
    r = subprocess.run([exe, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "--version ID" in r.stdout and "-f FILE" in r.stdout


def test_readers_match_oracle(tmp_path):
    import nwgraph_amd
    import pyoracle as O

    el = nwgraph_amd.read_mm(os.path.join(GOLDEN, "karate.mtx"))
    n, s, d = O.read_mm(os.path.join(GOLDEN, "karate.mtx"))
    assert el.n == n == 34 and np.array_equal(el.src, s) and np.array_equal(el.dst, d)
    p = tmp_path / "k.nw"
    nwgraph_amd.write_nw(p, el)
    el2 = nwgraph_amd.read_nw(p)
    n2, s2, d2 = O.read_nw(p)
    assert el2.n == n2 == 34 and np.array_equal(el2.src, el.src) and np.array_equal(d2, el.dst)


def test_reader_rejects_bad_files(tmp_path):
    import nwgraph_amd
    from nwgraph_amd import NwgError

    p = tmp_path / "bad.mtx"
    p.write_text("%%MatrixMarket matrix coordinate pattern skew-symmetric\n2 2 1\n1 2\n")
    with pytest.raises(NwgError):
        nwgraph_amd.read_mm(p)
    q = tmp_path / "bad.nw"
    q.write_bytes(b"not a graph")
    with pytest.raises(NwgError):
        nwgraph_amd.read_nw(q)


def _adjacency_bytes(n, ptr, idx, n_field=None, junk=0):
    """The bytes index_adjacency<1>::serialize writes (adjacency.hpp:176-180): graph base
    (graph_base.hpp:72-75: size_t N, bool is_open), indexed struct of arrays
    (compressed.hpp:377-387: 34-byte magic, 8 bytes from the 4-byte index_t N_, size_t N+1,
    size_t 4, offsets), struct of arrays (soa.hpp:333-356: size_t entries, size_t arrays, then
    size_t count, size_t element size and the ids)."""
    ptr = np.asarray(ptr, np.uint32)
    idx = np.asarray(idx, np.uint32)
    nf = n + 1 if n_field is None else n_field
    b = np.uint64(n).tobytes() + b"\0"
    b += b"NW GRAPH indexed_struct_of_arrays\0"
    b += np.uint32(nf).tobytes() + np.uint32(junk).tobytes()
    b += np.uint64(n + 1).tobytes() + np.uint64(4).tobytes() + ptr.tobytes()
    b += np.uint64(idx.size).tobytes() + np.uint64(1).tobytes()
    b += np.uint64(idx.size).tobytes() + np.uint64(4).tobytes() + idx.tobytes()
    return b


def test_adjacency_serialize_layout(tmp_path):
    import nwgraph_amd

    ptr, idx = [0, 1, 3, 3, 4], [2, 0, 3, 1]
    a = nwgraph_amd.Adjacency(ptr, idx)
    p = tmp_path / "a.adj"
    a.serialize(p)
    assert p.read_bytes() == _adjacency_bytes(4, ptr, idx)
    # an empty graph and a graph without edges
    for n, pp, ii in ((0, [0], []), (3, [0, 0, 0, 0], [])):
        q = tmp_path / f"e{n}.adj"
        nwgraph_amd.Adjacency(pp, ii).serialize(q)
        assert q.read_bytes() == _adjacency_bytes(n, pp, ii)
        b = nwgraph_amd.Adjacency([0], []).deserialize(q)
        assert b.size() == n and b.num_edges() == 0


def test_adjacency_roundtrip_and_reference_files(tmp_path):
    import nwgraph_amd
    import pyoracle as O

    n, s, d = O.read_mm(os.path.join(GOLDEN, "karate.mtx"))
    ind, col = O.csc(n, *O.clean_edges(s, d))
    a = nwgraph_amd.Adjacency(ind, col)
    p = tmp_path / "k.adj"
    a.serialize(p)
    b = nwgraph_amd.Adjacency([0], []).deserialize(p)
    assert b.size() == n and np.array_equal(b.indices_, ind) and np.array_equal(b.to_be_indexed_, col)
    # a file as the reference writes it for an adjacency made with the (N, M) constructor
    # (N_ = N) and with memory garbage in the upper half of the 8-byte N_ field
    q = tmp_path / "ref.adj"
    q.write_bytes(_adjacency_bytes(n, ind, col, n_field=n, junk=0xDEADBEEF))
    c = nwgraph_amd.Adjacency([0], []).deserialize(q)
    assert np.array_equal(c.indices_, ind) and np.array_equal(c.to_be_indexed_, col)


def test_adjacency_reader_rejects_bad_files(tmp_path):
    import nwgraph_amd
    from nwgraph_amd import NwgError

    ptr, idx = [0, 1, 3, 3, 4], [2, 0, 3, 1]
    good = _adjacency_bytes(4, ptr, idx)
    bad = {
        "magic": good[:9] + b"NW GRAPH indexed_struct_of_arrayZ\0" + good[43:],
        "truncated": good[:-3],
        "id out of range": _adjacency_bytes(4, ptr, [2, 0, 4, 1]),
        "offsets decrease": _adjacency_bytes(4, [0, 3, 1, 3, 4], idx),
        "offsets end": _adjacency_bytes(4, [0, 1, 3, 3, 3], idx[:4]),
    }
    for name, data in bad.items():
        q = tmp_path / "bad.adj"
        q.write_bytes(data)
        with pytest.raises(NwgError):
            nwgraph_amd.Adjacency([0], []).deserialize(q)
    with pytest.raises(NwgError):  # the writer checks the offsets against the ids
        nwgraph_amd.Adjacency([0, 1, 3], [0]).serialize(tmp_path / "w.adj")


def test_readers_reject_oversized_counts_without_throwing(tmp_path):
    """Counts taken from a file header are checked against the file size before anything is
    allocated for them: an absurd count is a malformed file (NWG_EFORMAT), never an
    exception escaping the C boundary (which would abort the process)."""
    import nwgraph_amd
    from nwgraph_amd import NwgError, _lib

    mm = tmp_path / "huge.mtx"
    mm.write_text("%%MatrixMarket matrix coordinate pattern general\n3 3 4000000000000000000\n1 2\n")
    with pytest.raises(NwgError) as e:
        nwgraph_amd.read_mm(mm)
    assert e.value.code == _lib.NWG_EFORMAT
    mm_sym = tmp_path / "huge_sym.mtx"
    mm_sym.write_text("%%MatrixMarket matrix coordinate pattern symmetric\n3 3 18446744073709551615\n1 2\n")
    with pytest.raises(NwgError) as e:
        nwgraph_amd.read_mm(mm_sym)
    assert e.value.code == _lib.NWG_EFORMAT

    # NW edge list whose header claims 2^61 edges
    hdr = b"NW Graph index_edge_list\0\0\0" + np.int32(1).tobytes() + np.uint64(4).tobytes() + b"\0"
    big = np.uint64(1 << 61)
    nw = tmp_path / "huge.nw"
    nw.write_bytes(hdr + big.tobytes() + np.uint64(2).tobytes() + big.tobytes() + np.uint64(4).tobytes()
                   + np.zeros(8, np.uint32).tobytes())
    with pytest.raises(NwgError) as e:
        nwgraph_amd.read_nw(nw)
    assert e.value.code == _lib.NWG_EFORMAT

    # adjacency whose offsets and ids claim far more entries than the file holds
    good = _adjacency_bytes(4, [0, 1, 3, 3, 4], [2, 0, 3, 1])
    huge_n = (np.uint64(1 << 31).tobytes() + b"\0" + b"NW GRAPH indexed_struct_of_arrays\0"
              + np.uint32(5).tobytes() + np.uint32(0).tobytes() + np.uint64((1 << 31) + 1).tobytes()
              + np.uint64(4).tobytes() + good[67:])
    q = tmp_path / "huge_n.adj"
    q.write_bytes(huge_n)
    with pytest.raises(NwgError) as e:
        nwgraph_amd.Adjacency([0], []).deserialize(q)
    assert e.value.code == _lib.NWG_EFORMAT
    ne_off = len(good) - (4 * 4 + 8 + 8 + 8 + 8)  # start of the struct-of-arrays header
    huge_ne = bytearray(good)
    huge_ne[ne_off:ne_off + 8] = np.uint64(1 << 62).tobytes()
    huge_ne[ne_off + 16:ne_off + 24] = np.uint64(1 << 62).tobytes()
    q.write_bytes(bytes(huge_ne))
    with pytest.raises(NwgError) as e:
        nwgraph_amd.Adjacency([0], []).deserialize(q)
    assert e.value.code == _lib.NWG_EFORMAT
