"""GPU parity of the generators, the on-device builder and the PageRank sweep, all
called through the C ABI, against the CPU oracle on the same seeded inputs."""
import os

import numpy as np
import pytest

import pyoracle as O
from conftest import GOLDEN, golden

pytestmark = pytest.mark.gpu

D = 0.85


def csc_equal(g, n, ind, col, deg):
    ptr, idx, od = g.csc()
    assert g.n == n and g.nnz == col.size
    assert np.array_equal(ptr, ind), "row pointers differ"
    assert np.array_equal(idx, col), "column ids differ"
    assert np.array_equal(od, deg), "out-degrees differ"


# ---------------------------------------------------------------- generators
@pytest.mark.parametrize("scale,first,count", [(12, 0, 65536), (16, 123456, 300000), (25, 10 ** 9, 4096)])
def test_rmat_edges_bit_exact(ctx, scale, first, count):
    s, d = ctx.rmat_edges(scale, first, count, seed=7)
    so, do = O.rmat_edges(scale, first, count, seed=7)
    assert np.array_equal(s, so) and np.array_equal(d, do)


def test_grid_edges_bit_exact(ctx):
    s, d = ctx.grid_edges(96, seed=11)
    so, do = O.grid_edges(96, seed=11)
    assert np.array_equal(s, so) and np.array_equal(d, do)


# ---------------------------------------------------------------- builder
@pytest.mark.parametrize("scale,ef", [(8, 4), (12, 16), (16, 16)])
def test_rmat_build_bit_exact(ctx, scale, ef):
    g = ctx.rmat(scale, ef, seed=3)
    csc_equal(g, *O.rmat_csc(scale, ef, seed=3))


def test_grid_build_bit_exact(ctx):
    g = ctx.grid(128, seed=5)
    csc_equal(g, *O.grid_csc(128, seed=5))


def test_builder_kats_on_device(ctx):
    for case in golden("builder_kat.json")["cases"]:
        e = np.array(case["edges"], np.uint32)
        n = case["n"]
        op = case.get("op")
        if op == "clean":
            g = ctx.graph_from_edges(n, e[:, 0], e[:, 1], clean=True)
            assert g.nnz == case["clean_count"], case["name"]
        elif op == "lexical_sort":
            # CSR of the raw edges = lexical (u, v) order
            g = ctx.graph_from_edges(n, e[:, 1], e[:, 0], clean=False)
            ptr, idx, _ = g.csc()
            rows = np.repeat(np.arange(n), np.diff(ptr.astype(np.int64)))
            pairs = list(zip(rows.tolist(), idx.tolist()))
            assert pairs[: len(case["sorted_prefix"])] == [tuple(p) for p in case["sorted_prefix"]]
        else:
            s, d = (e[:, 0], e[:, 1]) if case["idx"] == 1 else (e[:, 1], e[:, 0])
            g = ctx.graph_from_edges(n, s, d, clean=False)
            ptr, _, _ = g.csc()
            assert np.diff(ptr.astype(np.int64)).tolist() == case["row_sizes"], case["name"]


def test_build_uncleaned_keeps_duplicates_and_loops(ctx):
    rng = np.random.default_rng(4)
    n = 500
    s = rng.integers(0, n, 20000).astype(np.uint32)
    d = rng.integers(0, n, 20000).astype(np.uint32)
    g = ctx.graph_from_edges(n, s, d, clean=False)
    ind, col = O.csc(n, s, d)
    csc_equal(g, n, ind, col, O.out_degrees(n, ind, col))
    gc = ctx.graph_from_edges(n, s, d, clean=True)
    cs, cd = O.clean_edges(s, d)
    ind2, col2 = O.csc(n, cs, cd)
    csc_equal(gc, n, ind2, col2, O.out_degrees(n, ind2, col2))


def test_build_rejects_out_of_range_ids(ctx):
    from nwgraph_amd import NwgError

    with pytest.raises(NwgError):
        ctx.graph_from_edges(4, np.array([0, 9], np.uint32), np.array([1, 2], np.uint32))


def test_csr_bit_exact(ctx):
    g = ctx.rmat(12, 16, seed=9)
    n, ind, col, deg = O.rmat_csc(12, 16, seed=9)
    rows = np.repeat(np.arange(n, dtype=np.uint32), np.diff(ind.astype(np.int64)))
    rind, rcol = O.fill_sorted(0, n, col, rows)
    ptr, idx, pos = g.csr()
    assert np.array_equal(ptr, rind) and np.array_equal(idx, rcol)
    # csr_to_csc maps every CSR entry (u -> v) to the CSC position of the same edge
    assert np.array_equal(col[pos], np.repeat(np.arange(n), np.diff(rind.astype(np.int64))))
    assert np.array_equal(rows[pos], rcol)


# ---------------------------------------------------------------- PageRank
def check_against_oracle(g, n, ind, col, deg, threshold=1e-4, max_iters=20, teleport=None, tol=1e-6):
    x, info = g.pagerank(D, threshold, max_iters, teleport=teleport)
    ref, it, tr, _ = O.pr_v10(n, ind, col, deg, D, threshold, max_iters, teleport=teleport, nthreads=8)
    assert info["iterations"] == it
    assert np.max(np.abs(x.astype(np.float64) - ref)) <= tol
    assert np.allclose(info["error_trace"], tr, rtol=1e-3, atol=1e-12)
    return x, info, ref


@pytest.mark.parametrize("scale", [12, 16, 18])
def test_pagerank_rmat_matches_v10(ctx, scale):
    g = ctx.rmat(scale, 16, seed=1)
    n, ind, col, deg = O.rmat_csc(scale, 16, seed=1)
    x, info, ref = check_against_oracle(g, n, ind, col, deg)
    # relative agreement far inside the 1e-6 absolute gate (FP32 summation order only)
    assert np.max(np.abs(x - ref) / np.maximum(ref, 1e-30)) < 1e-4
    # both the device and the reference's float run against FP64 Jacobi at the same sweep
    # count (same damping and start): within 1e-6 (measured gaps are ~1e-8)
    x64, _ = O.fp64_jacobi(n, ind, col, deg, damping=float(np.float32(D)), tol=0.0, max_iters=info["iterations"])
    assert np.max(np.abs(x - x64)) <= 1e-6 and np.max(np.abs(ref - x64)) <= 1e-6
    # report-only gap to the sequential in-place (v11, one thread) schedule
    v11, it11, _ = O.pr_v11(n, ind, col, deg)
    assert np.max(np.abs(x - v11)) < 5e-5


@pytest.mark.parametrize(
    "plan",
    [
        dict(short_max=0, group_max=0, hub_min=1, hub_segment=256),  # every row a hub
        dict(short_max=0, group_max=64, hub_min=65, hub_segment=256),
        dict(short_max=32, group_max=32, hub_min=33, hub_segment=512),
        dict(short_max=4, group_max=1024, hub_min=100000, hub_segment=4096),  # no hubs
        dict(short_max=100000, group_max=100000, hub_min=100001, hub_segment=4096),  # all short
        dict(short_max=8, group_max=128, hub_min=2048, hub_segment=2048, relabel=1),  # degree-ordered copy
        dict(short_max=2, group_max=64, hub_min=512, hub_segment=512, relabel=1, stream_hints=1),
        # source-sliced rows of every lane class (mean slice length 2..256), hubs above 2047
        dict(short_max=4, group_max=128, hub_min=2048, hub_segment=2048, slice_min=9),
        # sliced rows on the degree-ordered copy, no hubs, every stream hinted
        dict(short_max=8, group_max=64, hub_min=100000, hub_segment=4096, slice_min=65, relabel=1, stream_hints=2),
        # nearly every row sliced: most slices empty or a single edge
        dict(short_max=0, group_max=16, hub_min=1 << 20, hub_segment=256, slice_min=2),
        # binned sweep: one bin and one chunk at this size, bins of 2^13 and 2^14 rows
        dict(binned=1),
        dict(binned=1, relabel=1),
        dict(binned=1, bin_bits=14),
        dict(binned=1, relabel=1, bin_bits=13),
    ],
)
def test_pagerank_plans_agree(ctx, plan):
    g = ctx.rmat(14, 16, seed=2)
    g.set_plan(**plan)
    assert {k: g.get_plan()[k] for k in plan} == plan
    n, ind, col, deg = O.rmat_csc(14, 16, seed=2)
    check_against_oracle(g, n, ind, col, deg)


def test_pagerank_grid_converges_like_oracle(ctx):
    g = ctx.grid(256, seed=5)
    n, ind, col, deg = O.grid_csc(256, seed=5)
    x, info, ref = check_against_oracle(g, n, ind, col, deg, max_iters=100)
    assert info["iterations"] < 100  # converges before the cap


def test_pagerank_karate_golden(ctx):
    import nwgraph_amd

    kg = golden("karate_pagerank.json")
    el = nwgraph_amd.read_mm(os.path.join(GOLDEN, "karate.mtx"))
    adj = nwgraph_amd.build_adjacency(el, ctx=ctx)
    deg = nwgraph_amd.build_degrees(adj, ctx=ctx)
    scores = np.empty(adj.size(), np.float32)
    info = {}
    nwgraph_amd.page_rank(adj, deg, scores, 0.85, 1e-7, 100, 1, ctx=ctx, info=info)
    ans = np.array(kg["pagerank"])
    assert np.all(np.abs(scores - ans) <= kg["rel_tolerance"] * ans)
    ref, it, _, _ = O.pr_v10(34, adj.indices_, adj.to_be_indexed_, deg, 0.85, 1e-7, 100)
    assert info["iterations"] == it and np.max(np.abs(scores - ref)) < 1e-7


@pytest.mark.parametrize("relabel", [0, 1])
def test_pagerank_personalised_teleport(ctx, relabel):
    g = ctx.rmat(13, 16, seed=4)
    if relabel:
        p = g.get_plan()
        p["relabel"] = 1
        g.set_plan(**p)
    n, ind, col, deg = O.rmat_csc(13, 16, seed=4)
    ids = O.ppr_sources(n, 1, 8, seed=4)[0]
    v = np.zeros(n, np.float32)
    v[ids] = 1.0 / 8
    check_against_oracle(g, n, ind, col, deg, teleport=v)


def test_pr_driver_karate(tmp_path):
    """The pr driver (bench/gapbs/pr.cpp counterpart) on karate.mtx: trace, --verify top 10
    and the Times table."""
    import subprocess

    exe = os.path.join(os.path.dirname(GOLDEN), "..", "nwgraph_amd", "bin", "pr")
    out = subprocess.run([exe, "-f", os.path.join(GOLDEN, "karate.mtx"), "-i", "100", "-t", "1e-7", "-n", "2",
                          "--verify", "--version", "11", "--version", "10", "4"], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    text = out.stdout
    assert "iter" in text and "Verifying" in text and "File" in text and "Version" in text
    # top-10 lines "id: score" match the oracle's ordering (adjacency<1> keeps duplicates: none here)
    kg = golden("karate_pagerank.json")
    ind, col = O.csc(34, np.array(kg["src"], np.uint32), np.array(kg["dst"], np.uint32))
    ref, _, _, _ = O.pr_v10(34, ind, col, O.out_degrees(34, ind, col), 0.85, 1e-7, 100)
    top = [int(l.split(":")[0]) for l in text.splitlines() if ": " in l and l.split(":")[0].isdigit()][:10]
    assert top == list(np.argsort(-ref, kind="stable")[:10])
    rows = [l for l in text.splitlines() if "karate.mtx" in l and not l.startswith("Reading")]
    assert len(rows) == 2 and {int(r.split()[1]) for r in rows} == {10, 11}
    assert all(int(r.split()[2]) == 4 for r in rows)
    # every solve prints the reference's init timers around the trace header, then one line
    # per sweep with its own time (page_rank.hpp:99-142)
    lines = text.splitlines()
    heads = [k for k, l in enumerate(lines) if l.split() == ["iter", "error", "time", "outgoing"]]
    assert len(heads) == 4
    for k in heads:
        assert lines[k - 1].startswith("# [ init page rank ]: ") and lines[k - 1].endswith(" ms")
        assert lines[k + 1].startswith("# [ init contrib ]: ") and lines[k + 1].endswith(" ms")
        first = lines[k + 2].split()
        assert first[0] == "0" and float(first[2]) > 0 and first[3] == "0"


def test_pr_driver_log_matches_the_reference_layout(tmp_path):
    """--log / --log-header write the reference's result table (bench/gapbs/Log.hpp:39-179):
    one row per sample with uuid, library, git branch and revision, compiler, build, TBB
    malloc, date, host, benchmark, version, threads, graph stem and the sample's time."""
    import datetime
    import subprocess

    exe = os.path.join(os.path.dirname(GOLDEN), "..", "nwgraph_amd", "bin", "pr")
    log = tmp_path / "logs" / "pr.log"
    out = subprocess.run([exe, "-f", os.path.join(GOLDEN, "karate.mtx"), "-n", "3", "--version", "10",
                          "--log", str(log), "--log-header", "2"], capture_output=True, text=True, timeout=120,
                         cwd=tmp_path)
    assert out.returncode == 0, out.stderr
    lines = log.read_text().splitlines()
    assert len(lines) == 4
    head = lines[0].split()
    assert head == ["UUID", "Library", "Branch", "Revision", "CXX", "CXX_ID", "CXX_VER", "Build", "TBBMALLOC",
                    "Date", "Host", "Benchmark", "Version", "Threads", "Graph", "Time(s)", "Tolerance"]
    # fixed widths: UUID column 26, Library/Branch/Revision 10 each
    assert lines[0].index("Library") == 26 and lines[0].index("Branch") == 36
    uuids = set()
    for row in lines[1:]:
        uuid = row[:26].strip()
        assert len(uuid) == 24 and uuid.isalpha() and uuid.islower()
        uuids.add(uuid)
        assert row[26:36].strip() == "NWG_HIP"
        f = row.split()
        # ... Date, Host, Benchmark, Version, Threads, Graph (stem of karate.mtx), Time
        assert f[-5:-1] == ["pr", "v10", "2", "karate"]
        datetime.date.fromisoformat(f[-7])
        t = f[-1]
        assert "e" in t and float(t) > 0  # scientific, precision 8
        assert len(t.split("e")[0].split(".")[1]) == 8
    assert len(uuids) == 1


def with_plan(g, plan):
    if plan:
        p = g.get_plan()
        p.update(plan)
        g.set_plan(**p)
    return g


@pytest.mark.parametrize("plan", [{}, {"binned": 1}, {"binned": 1, "relabel": 1}, {"binned": 1, "bin_bits": 14},
                                  {"binned": 1, "relabel": 1, "bin_bits": 14}])
def test_edge_cases(ctx, plan):
    # no edges: every score becomes base after the first sweep
    g = with_plan(ctx.graph_from_edges(5, np.zeros(0, np.uint32), np.zeros(0, np.uint32)), plan)
    x, info = g.pagerank(D, 1e-4, 20)
    base = np.float32((1.0 - np.float64(np.float32(D))) / 5)
    assert np.all(x == base) and info["iterations"] == 2
    # one vertex with a self loop removed by cleaning
    g1 = with_plan(ctx.graph_from_edges(1, np.array([0], np.uint32), np.array([0], np.uint32)), plan)
    x1, _ = g1.pagerank(D, 1e-4, 20)
    assert g1.nnz == 0 and x1[0] == np.float32((1.0 - np.float64(np.float32(D))) / 1)
    # max_iters = 0 leaves the initial scores (page_rank.hpp:100-104)
    g2 = with_plan(ctx.rmat(10, 8, seed=1), plan)
    x2, info2 = g2.pagerank(D, 1e-4, 0)
    assert info2["iterations"] == 0 and np.all(x2 == np.float32(1.0 / 1024))
    # a huge threshold stops after one sweep
    n, ind, col, deg = O.rmat_csc(10, 8, seed=1)
    x3, info3 = g2.pagerank(D, 10.0, 20)
    ref, it, _, _ = O.pr_v10(n, ind, col, deg, D, 10.0, 20)
    assert info3["iterations"] == it == 1 and np.max(np.abs(x3 - ref) / ref) < 1e-6
    # dangling vertices lose mass: sum of scores stays below 1 (no redistribution)
    assert x3.sum() < 1.0
    # a star into vertex 0 and a path, n not a multiple of any bin or chunk size, with
    # isolated vertices: rows of every length from 0 to 40000
    n = 70001
    src = np.concatenate([np.arange(1, 40001), np.arange(40001, 69999)]).astype(np.uint32)
    dst = np.concatenate([np.zeros(40000), np.arange(40002, 70000)]).astype(np.uint32)
    g4 = with_plan(ctx.graph_from_edges(n, src, dst), plan)
    ptr, idx, deg = g4.csc()
    x4, info4 = g4.pagerank(D, 0.0, 30)
    # checked against FP64 Jacobi: the float oracle's sequential 40000-term sum into vertex 0
    # drifts by ~1e-5, the device's tree or exact sums do not
    x64, _ = O.fp64_jacobi(n, ptr, idx, deg, damping=float(np.float32(D)), tol=0.0, max_iters=30)
    assert info4["iterations"] == 30 and np.max(np.abs(x4 - x64)) <= 1e-6


@pytest.mark.parametrize("plan", [{"binned": 0}, {"binned": 0, "relabel": 1}, {"binned": 1},
                                  {"binned": 1, "relabel": 1}])
def test_closed_forms_and_invariances(ctx, plan):
    """What the mathematics fixes, on every sweep kind: a regular digraph keeps x = 1/N, a
    Jacobi sweep without dangling vertices keeps the total score 1, a vertex without
    in-edges gets exactly base and no score falls below it, the input edge order does not
    matter (bit-identical), and renaming the vertices renames the scores."""
    base1 = lambda n: np.float32((1.0 - np.float64(np.float32(D))) / n)
    # directed cycle and complete digraph: in- and out-degree equal everywhere
    for n, s, d in [(1000, np.arange(1000), (np.arange(1000) + 1) % 1000),
                    (64, np.repeat(np.arange(64), 64), np.tile(np.arange(64), 64))]:
        g = with_plan(ctx.graph_from_edges(n, s, d), plan)
        x, info = g.pagerank(D, 0.0, 20)
        assert info["iterations"] == 20
        assert np.max(np.abs(x.astype(np.float64) * n - 1.0)) < 1e-6, n
    # no dangling vertices (a Hamiltonian cycle plus random edges): sum of scores stays 1
    rng = np.random.default_rng(21)
    n = 3001
    s = np.concatenate([np.arange(n), rng.integers(0, n, 40000)])
    d = np.concatenate([(np.arange(n) + 1) % n, rng.integers(0, n, 40000)])
    g = with_plan(ctx.graph_from_edges(n, s, d), plan)
    _, _, deg = g.csc()
    assert np.all(deg > 0)
    x, _ = g.pagerank(D, 0.0, 20)
    assert abs(x.sum(dtype=np.float64) - 1.0) < 1e-5
    assert np.all(x >= base1(n))
    # an RMAT graph: isolated and source-only vertices, hubs, dangling vertices
    scale = 12
    n = 1 << scale
    s, d = ctx.rmat_edges(scale, 0, 16 * n, seed=17)
    g = with_plan(ctx.graph_from_edges(n, s, d), plan)
    ptr, _, _ = g.csc()
    x, info = g.pagerank(D, 1e-4, 20)
    no_in = np.diff(ptr.astype(np.int64)) == 0
    assert no_in.any() and np.all(x[no_in] == base1(n)) and np.all(x >= base1(n))
    # the same edges in another order: the same CSC, bit-identical scores
    p = rng.permutation(s.size)
    gp = with_plan(ctx.graph_from_edges(n, s[p], d[p]), plan)
    xp, infop = gp.pagerank(D, 1e-4, 20)
    assert infop["iterations"] == info["iterations"] and np.array_equal(xp, x)
    # vertices renamed by a permutation pi: score of pi(v) = score of v up to FP32 summation order
    pi = rng.permutation(n).astype(np.uint32)
    gr = with_plan(ctx.graph_from_edges(n, pi[s], pi[d]), plan)
    xr, infor = gr.pagerank(D, 0.0, 20)
    x20, _ = g.pagerank(D, 0.0, 20)
    assert infor["iterations"] == 20
    assert np.max(np.abs(xr[pi].astype(np.float64) - x20) / x20) < 1e-5


def test_repeat_calls_bit_identical_and_timing_mode(ctx):
    g = ctx.rmat(15, 16, seed=6)
    a, ia = g.pagerank(D, 1e-4, 20)
    b, ib = g.pagerank(D, 1e-4, 20)
    c, ic = g.pagerank(D, 1e-4, 20, kernel_timing=True)
    assert np.array_equal(a, b) and np.array_equal(a, c)
    assert ia["iterations"] == ib["iterations"] == ic["iterations"]
    assert ic["sweep_ms_avg"] > 0


def test_full_size_properties_rmat22(ctx):
    """Size-independent checks at RMAT-22: one extra sweep equals the oracle's update
    of the previous iterate (sampled rows, FP64) and mass balance over all edges."""
    g = ctx.rmat(22, 16, seed=1)
    ptr, idx, deg = g.csc()
    n = g.n
    # builder invariants: sorted unique rows, no self loops, degree sum
    assert ptr[-1] == g.nnz == idx.size and int(deg.sum(dtype=np.uint64)) == g.nnz
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(ptr.astype(np.int64)))
    keys = rows * n + idx.astype(np.int64)
    assert np.all(np.diff(keys) > 0) and not np.any(rows == idx)
    x19, i19 = g.pagerank(D, 0.0, 19)
    x20, i20 = g.pagerank(D, 0.0, 20)
    assert i19["iterations"] == 19 and i20["iterations"] == 20
    c = x19.astype(np.float64) / np.where(deg > 0, deg, 1)
    base = np.float64(np.float32((1.0 - np.float64(np.float32(D))) / n))
    rng = np.random.default_rng(0)
    sample = np.concatenate([rng.integers(0, n, 2000), np.argsort(np.diff(ptr.astype(np.int64)))[-50:]])
    for i in sample:
        z = c[idx[ptr[i]:ptr[i + 1]]].sum()
        want = base + np.float64(np.float32(D)) * z
        assert abs(x20[i] - want) <= 1e-6 * max(want, 1e-12) + 1e-12
    # mass balance: sum x20 = N base + d * sum_{deg>0} x19
    lhs = x20.astype(np.float64).sum()
    rhs = n * base + np.float64(np.float32(D)) * x19.astype(np.float64)[deg > 0].sum()
    assert abs(lhs - rhs) < 1e-5 * rhs
    # L1 trace equals the host L1 between the two iterates
    assert abs(i20["error_trace"][-1] - np.abs(x20.astype(np.float64) - x19).sum()) < 1e-9


@pytest.mark.parametrize("kind", ["rmat25", "grid4096"])
def test_full_size_properties_binned(ctx, kind):
    """The benchmark's full-size inputs on the binned sweep (the plan the autotuner picks for
    both): one more sweep equals the FP64 update of the previous iterate on sampled rows
    (the 100 highest in-degree rows included), mass balance over all edges, the L1 trace,
    and a repeat that is bit-identical."""
    g = ctx.rmat(25, 16, seed=1) if kind == "rmat25" else ctx.grid(4096, seed=1)
    p = g.get_plan()
    p.update(binned=1, relabel=1 if kind == "rmat25" else 0)
    g.set_plan(**p)
    ptr, idx, deg = g.csc()
    n = g.n
    assert int(deg.sum(dtype=np.uint64)) == g.nnz == idx.size
    x19, i19 = g.pagerank(D, 0.0, 19)
    x20, i20 = g.pagerank(D, 0.0, 20)
    assert i19["iterations"] == 19 and i20["iterations"] == 20
    c = x19.astype(np.float64) / np.where(deg > 0, deg, 1)
    base = np.float64(np.float32((1.0 - np.float64(np.float32(D))) / n))
    dd = np.float64(np.float32(D))
    rng = np.random.default_rng(0)
    sample = np.concatenate([rng.integers(0, n, 4000), np.argsort(np.diff(ptr.astype(np.int64)))[-100:]])
    for i in sample:
        want = base + dd * c[idx[ptr[i]:ptr[i + 1]]].sum()
        assert abs(x20[i] - want) <= 1e-6 * max(want, 1e-12) + 1e-12, (int(i), float(x20[i]), want)
    lhs = x20.astype(np.float64).sum()
    rhs = n * base + dd * x19.astype(np.float64)[deg > 0].sum()
    assert abs(lhs - rhs) < 1e-5 * rhs
    assert abs(i20["error_trace"][-1] - np.abs(x20.astype(np.float64) - x19).sum()) < 1e-6 * i20["error_trace"][-1]
    x20b, _ = g.pagerank(D, 0.0, 20)
    assert np.array_equal(x20, x20b)


def test_full_size_sliced_plan_rmat22(ctx):
    """At RMAT-22 the degree-ordered, source-sliced plan gives the default plan's scores
    (only the FP32 summation order of the sliced rows differs) and leaves the CSC as built."""
    g = ctx.rmat(22, 16, seed=1)
    ptr0, idx0, _ = g.csc()
    x_def, i_def = g.pagerank(D, 1e-4, 20)
    p = g.get_plan()
    p.update(short_max=4, group_max=128, hub_min=65536, hub_segment=8192, slice_min=65, relabel=1, binned=0)
    g.set_plan(**p)
    x_sl, i_sl = g.pagerank(D, 1e-4, 20)
    assert i_sl["iterations"] == i_def["iterations"]
    assert np.max(np.abs(x_sl - x_def) / np.maximum(x_def, 1e-30)) < 1e-5
    ptr1, idx1, _ = g.csc()
    assert np.array_equal(ptr0, ptr1) and np.array_equal(idx0, idx1)
    # and so does the binned sweep (bins split over many items at this size)
    p.update(binned=1)
    g.set_plan(**p)
    x_b, i_b = g.pagerank(D, 1e-4, 20)
    assert i_b["iterations"] == i_def["iterations"]
    assert np.max(np.abs(x_b - x_def) / np.maximum(x_def, 1e-30)) < 1e-5


@pytest.mark.parametrize("kind,plan", [("rmat", {}), ("rmat", {"relabel": 1}), ("rmat", {"binned": 1}),
                                       ("grid", {}), ("rmat16", {})])
def test_delta_schedule_matches_v13(ctx, kind, plan):
    """pr --version 13 (page_rank_v13): the push kernel plus the pull kernels against the
    oracle's restatement.  Pull sums differ from the oracle's sequential float sums only in
    order; a residual that lands within rounding of the threshold can be absorbed by one and
    not the other, so the gate is: same iteration count, and all but 0.1 % of the scores
    within 1e-5 relative, none off by more than 10 x the threshold."""
    thr = 1e-4
    if kind == "grid":
        g = ctx.grid(128, seed=2)
        n, ind, col, deg = O.grid_csc(128, seed=2)
        iters = 100
    else:
        sc = 16 if kind == "rmat16" else 13
        g = ctx.rmat(sc, 16, seed=2)
        n, ind, col, deg = O.rmat_csc(sc, 16, seed=2)
        iters = 60
    if plan:
        p = g.get_plan()
        p.update(plan)
        g.set_plan(**p)
    x, info = g.pagerank_delta(D, thr, iters)
    ref, it, pushed = O.pr_v13(n, ind, col, deg, D, thr, iters)
    assert info["iterations"] == it
    assert np.all(np.asarray(info["error_trace"]) == 0.0)
    diff = np.abs(x.astype(np.float64) - ref)
    assert np.mean(diff > 1e-5 * np.abs(ref) + 1e-7) <= 1e-3
    assert diff.max() <= 10 * thr
    # un-normalised: the scores sum to about N (a little less: dangling vertices and dropped
    # residuals lose mass)
    assert 0.3 * n < x.sum() <= 1.0001 * n
    # deterministic: a second call is bit-identical
    x2, info2 = g.pagerank_delta(D, thr, iters)
    assert np.array_equal(x, x2) and info2["iterations"] == info["iterations"]


def test_delta_schedule_cycle_stops_like_the_closed_form(ctx):
    """A directed cycle: every residual is (1-d) d^t; the device stops after the first
    iteration that pushes nothing (t = 31 at threshold 1e-3), like the oracle."""
    n = 1000
    s = np.arange(n, dtype=np.uint32)
    d = ((s + 1) % n).astype(np.uint32)
    g = ctx.graph_from_edges(n, s, d)
    x, info = g.pagerank_delta(D, 1e-3, 100)
    ind, col = O.csc(n, s, d)
    ref, it, _ = O.pr_v13(n, ind, col, O.out_degrees(n, ind, col), D, 1e-3, 100)
    assert info["iterations"] == it == 32
    assert np.allclose(x, ref, rtol=1e-6)


def test_out_degrees_by_id_range_match_bincount(ctx):
    """Out-degrees of a CSC with at least 2^24 entries are counted per 16384-id range after a
    radix sort over the range bits (no global atomics).  RMAT-22 (65M entries; its low ids
    take more than 2^20 entries, so that range is split over several workgroups): the built
    graph's degrees and those computed for an uploaded CSC without degrees equal bincount."""
    g = ctx.rmat(22, 16, seed=4)
    ptr, col, deg = g.csc()
    n = ptr.size - 1
    assert col.size >= (1 << 24)
    want = np.bincount(col, minlength=n).astype(np.uint32)
    assert np.count_nonzero(col < 16384) > (1 << 20)
    assert np.array_equal(deg, want)
    g2 = ctx.graph_from_csc(ptr, col)
    _, _, deg2 = g2.csc()
    assert np.array_equal(deg2, want)
    g2.close()
    g.close()


def test_device_csc_is_checked_and_degrees_match(ctx):
    """A CSC handed over in device memory is checked on the device (an entry >= n or a
    non-monotone row pointer is rejected before any kernel indexes by it), and the device
    out-degree histogram equals numpy's bincount (hub ids below and above the LDS range)."""
    import ctypes
    import torch
    from nwgraph_amd import _lib

    n, ind, col, deg = O.rmat_csc(15, 16, seed=9)
    dptr = torch.from_numpy(ind.astype(np.int32)).cuda()
    didx = torch.from_numpy(col.astype(np.int32)).cuda()
    h = ctypes.c_void_p()
    rc = ctx.lib.nwg_graph_create_csc(ctx.handle, n, col.size, ctypes.c_void_p(dptr.data_ptr()),
                                      ctypes.c_void_p(didx.data_ptr()), None, _lib.NWG_MEM_DEVICE, ctypes.byref(h))
    assert rc == 0
    g = ctx._wrap(h)
    _, _, od = g.csc()
    assert np.array_equal(od, np.bincount(col, minlength=n).astype(np.uint32))
    g.close()
    bad = didx.clone()
    bad[12345] = n  # one entry out of range
    h2 = ctypes.c_void_p()
    rc = ctx.lib.nwg_graph_create_csc(ctx.handle, n, col.size, ctypes.c_void_p(dptr.data_ptr()),
                                      ctypes.c_void_p(bad.data_ptr()), None, _lib.NWG_MEM_DEVICE, ctypes.byref(h2))
    assert rc == _lib.NWG_EINVAL and not h2.value
    badp = dptr.clone()
    badp[7] = badp[9] + 1  # non-monotone row pointers
    rc = ctx.lib.nwg_graph_create_csc(ctx.handle, n, col.size, ctypes.c_void_p(badp.data_ptr()),
                                      ctypes.c_void_p(didx.data_ptr()), None, _lib.NWG_MEM_DEVICE, ctypes.byref(h2))
    assert rc == _lib.NWG_EINVAL


@pytest.mark.parametrize("kind", ["rmat", "grid"])
def test_autotuned_plan_matches_v10(ctx, kind):
    """The autotuner's plan (pull or binned, whichever timed faster) keeps the v10 gate."""
    if kind == "rmat":
        g = ctx.rmat(17, 16, seed=4)
        n, ind, col, deg = O.rmat_csc(17, 16, seed=4)
        iters = 20
    else:
        g = ctx.grid(512, seed=4)
        n, ind, col, deg = O.grid_csc(512, seed=4)
        iters = 100
    plan = g.autotune()
    assert plan["binned"] in (0, 1) and plan["relabel"] in (0, 1)
    assert g.get_plan() == plan
    check_against_oracle(g, n, ind, col, deg, max_iters=iters)


@pytest.mark.parametrize("relabel", [0, 1])
@pytest.mark.parametrize("bin_bits", [13, 14])
def test_binned_sweep_many_bins_matches_v10(ctx, relabel, bin_bits):
    """Binned sweep at RMAT-18: 32 bins of 2^13 rows or 16 of 2^14, 8 chunks of 2^15 sources,
    the hot bins split over several items (combined through the global accumulators)."""
    g = ctx.rmat(18, 16, seed=3)
    p = g.get_plan()
    p.update(binned=1, relabel=relabel, bin_bits=bin_bits)
    g.set_plan(**p)
    n, ind, col, deg = O.rmat_csc(18, 16, seed=3)
    x, info, ref = check_against_oracle(g, n, ind, col, deg)
    assert np.max(np.abs(x - ref) / np.maximum(ref, 1e-30)) < 1e-5  # FP32 summation order only
    # deterministic and independent of the schedule: a second call is bit-identical
    x2, info2 = g.pagerank(D, 1e-4, 20)
    assert np.array_equal(x, x2) and info2["iterations"] == info["iterations"]


@pytest.mark.parametrize("bin_bits", [13, 14])
def test_binned_sweep_grid_and_personalised(ctx, bin_bits):
    g = ctx.grid(256, seed=5)
    p = g.get_plan()
    p.update(binned=1, bin_bits=bin_bits)
    g.set_plan(**p)
    n, ind, col, deg = O.grid_csc(256, seed=5)
    x, info, ref = check_against_oracle(g, n, ind, col, deg, max_iters=100)
    assert info["iterations"] < 100
    ids = O.ppr_sources(n, 1, 8, seed=9)[0]
    v = np.zeros(n, np.float32)
    v[ids] = 1.0 / 8
    check_against_oracle(g, n, ind, col, deg, teleport=v)


@pytest.mark.parametrize("kind", ["grid", "rmat"])
def test_inplace_schedule_reaches_the_fixed_point(ctx, kind):
    """The reference's in-place (v11) schedule: not reproducible bit for bit, so it is checked
    through what does not depend on the schedule: it reaches the same FP64 fixed point as the
    Jacobi schedule, its reported L1 change is below the threshold at the stop, and on the
    high-diameter grid it takes no more sweeps than Jacobi."""
    if kind == "grid":
        g = ctx.grid(256, seed=5)
        n, ind, col, deg = O.grid_csc(256, seed=5)
    else:
        g = ctx.rmat(14, 16, seed=2)
        n, ind, col, deg = O.rmat_csc(14, 16, seed=2)
    x64, _ = O.fp64_jacobi(n, ind, col, deg, damping=float(np.float32(D)))
    xj, ij = g.pagerank(D, 1e-7, 300)
    xi, ii = g.pagerank(D, 1e-7, 300, inplace=True)
    assert ii["iterations"] < 300 and ii["final_error"] < 1e-7
    assert np.max(np.abs(xj - x64)) <= 1e-6 and np.max(np.abs(xi - x64)) <= 1e-6
    if kind == "grid":
        assert ii["iterations"] <= ij["iterations"]
    # personalised teleport, and the binned sweep in place (one contribution array)
    p = g.get_plan()
    p.update(binned=1)
    g.set_plan(**p)
    v = np.zeros(n, np.float32)
    v[O.ppr_sources(n, 1, 8, seed=3)[0]] = 1.0 / 8
    xv, iv = g.pagerank(D, 1e-7, 300, teleport=v, inplace=True)
    x64v, _ = O.fp64_jacobi(n, ind, col, deg, v=v, damping=float(np.float32(D)))
    assert iv["final_error"] < 1e-7 and np.max(np.abs(xv - x64v)) <= 1e-6


@pytest.mark.parametrize("plan", [{}, {"binned": 1}, {"binned": 1, "relabel": 1}])
def test_uncleaned_graph_with_duplicates_and_loops(ctx, plan):
    """adjacency<1> as pr.cpp builds it keeps duplicate edges and self-loops (build.hpp:209-232):
    every plan sums each copy (a duplicate source appears twice in its row)."""
    rng = np.random.default_rng(12)
    n = 40000
    s = rng.integers(0, n, 600000).astype(np.uint32)
    d = (rng.integers(0, n, 600000) // 7).astype(np.uint32)  # skewed in-degrees
    s[:5000] = d[:5000]  # self-loops
    s[5000:15000] = s[15000:25000]
    d[5000:15000] = d[15000:25000]  # duplicates
    g = with_plan(ctx.graph_from_edges(n, s, d, clean=False), plan)
    ind, col = O.csc(n, s, d)
    deg = O.out_degrees(n, ind, col)
    check_against_oracle(g, n, ind, col, deg)


# ---------------------------------------------------------------- boundary checks
def test_teleport_is_validated(ctx):
    """A personalised teleport must be finite, in [0, 1] and sum to at most 1: the binned
    sweep's exact row sums and the fixed-point L1 change assume scores in [0, 1].  Invalid
    vectors are rejected with NWG_EINVAL on both plans; a valid sub-stochastic one (sum 0.5)
    on a graph of over 2^18 edges (binned sweep by default) matches the oracle."""
    from nwgraph_amd import NwgError, _lib

    g = ctx.rmat(16, 16, seed=4)
    n, ind, col, deg = O.rmat_csc(16, 16, seed=4)
    assert g.nnz >= 1 << 18 and g.get_plan()["binned"] == 1
    bad = {
        "ones": np.ones(n, np.float32),
        "negative": np.full(n, 1.0 / n, np.float32) * np.where(np.arange(n) == 7, -1, 1).astype(np.float32),
        "nan": np.where(np.arange(n) == 3, np.nan, 1.0 / n).astype(np.float32),
        "sum 1.01": np.full(n, 1.01 / n, np.float32),
    }
    for plan in ({"binned": 1}, {"binned": 0}):
        g.set_plan(**plan)
        for name, v in bad.items():
            with pytest.raises(NwgError) as e:
                g.pagerank(D, 1e-4, 20, teleport=v)
            assert e.value.code == _lib.NWG_EINVAL, name
    g.set_plan(binned=1)
    rng = np.random.default_rng(1)
    v = rng.uniform(size=n)
    v = (0.5 * v / v.sum()).astype(np.float32)
    check_against_oracle(g, n, ind, col, deg, teleport=v)


def test_given_outdegrees_are_checked(ctx):
    """nwg_graph_create_csc checks a caller's out-degrees against the CSC: one below a
    vertex's entry count is rejected (its contributions would sum past its score); larger
    ones are accepted and used as given, like the reference's degrees vector."""
    from nwgraph_amd import NwgError, _lib

    n, ind, col, deg = O.rmat_csc(12, 16, seed=5)
    low = deg.copy()
    u = int(np.argmax(deg))
    low[u] -= 1
    with pytest.raises(NwgError) as e:
        ctx.graph_from_csc(ind, col, low)
    assert e.value.code == _lib.NWG_EINVAL
    high = deg + 1
    g = ctx.graph_from_csc(ind, col, high)
    x, info = g.pagerank(D, 1e-4, 20)
    ref, it, _, _ = O.pr_v10(n, ind, col, high, D, 1e-4, 20)
    assert info["iterations"] == it and np.max(np.abs(x - ref)) <= 1e-6


@pytest.mark.parametrize("kind", ["rmat25", "grid4096"])
def test_full_size_matches_oracle_v10(ctx, kind):
    """The headline inputs at full size, row by row: the autotuned plan (what the benchmark
    times) against the oracle's v10 run on the device-built CSC, same sweep count, every row
    within 1e-6, and the device out-degrees against numpy's bincount of the whole CSC."""
    import os

    if kind == "rmat25":
        g, iters = ctx.rmat(25, 16, seed=1), 20
    else:
        g, iters = ctx.grid(4096, seed=1), 100
    g.autotune()
    x, info = g.pagerank(D, 1e-4, iters)
    ptr, idx, deg = g.csc()
    n = g.n
    assert np.array_equal(deg, np.bincount(idx, minlength=n).astype(np.uint32))
    threads = min(16, os.cpu_count() or 1)
    ref, it, tr, _ = O.pr_v10(n, ptr, idx, deg, D, 1e-4, iters, nthreads=threads)
    assert info["iterations"] == it, (info["iterations"], it)
    err = np.abs(x.astype(np.float64) - ref)
    assert float(err.max()) <= 1e-6, (int(err.argmax()), float(err.max()))
    assert np.allclose(info["error_trace"], tr, rtol=1e-3, atol=1e-12)


def test_cpp_dropin_runs_against_the_oracle(tmp_path):
    """The C++ drop-in (include/nwgraph_hip/page_rank.hpp) compiled into a caller that uses it
    the way bench/gapbs/pr.cpp does, on RMAT-16: page_rank, page_rank_v10 and the overload for
    the reference's adjacency<1> (which must win over a generic template) all match the
    oracle's v10 within 1e-6, with the reference's trace on stdout."""
    import re
    import shutil
    import subprocess

    from conftest import ROOT

    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    import nwgraph_amd

    n, ind, col, deg = O.rmat_csc(16, 16, seed=1)
    adj = tmp_path / "rmat16.adj"
    nwgraph_amd.Adjacency(ind, col).serialize(adj)
    exe = tmp_path / "dropin_run"
    libdir = os.path.join(ROOT, "nwgraph_amd")
    subprocess.run(["g++", "-O2", "-std=c++20", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "dropin_run.cpp"), "-o", str(exe), "-L", libdir,
                    "-lnwgraph_hip", f"-Wl,-rpath,{libdir}"], check=True)
    out = subprocess.run([str(exe), str(adj), str(tmp_path / "scores")], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    assert "template calls: 0" in out.stdout
    lines = out.stdout.splitlines()
    assert re.match(r"# \[ init page rank \]: [0-9.e+-]+ ms", lines[0])
    assert lines[1].split() == ["iter", "error", "time", "outgoing"]
    assert re.match(r"# \[ init contrib \]: [0-9.e+-]+ ms", lines[2])
    ref, it, tr, _ = O.pr_v10(n, ind, col, deg, D, 1e-4, 20, nthreads=8)
    rows = [l.split() for l in lines[3:3 + it]]
    assert [int(r[0]) for r in rows] == list(range(it))
    assert np.allclose([float(r[1]) for r in rows], tr, rtol=1e-3)
    assert all(float(r[2]) > 0 for r in rows) and all(r[3] == "0" for r in rows)
    for suffix in ("v11", "v10", "adj"):
        x = np.fromfile(tmp_path / f"scores.{suffix}", np.float32)
        assert x.size == n and np.max(np.abs(x.astype(np.float64) - ref)) <= 1e-6, suffix
