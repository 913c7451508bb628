"""GPU parity of the batched personalised PageRank (SpMM) and of the many-graphs
persistent solve, through the C ABI, against the oracle's per-vector / per-graph runs."""
import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu

D = 0.85


def teleport(n, ids):
    v = np.zeros(n, np.float32)
    v[ids] = np.float32(1.0 / len(ids))
    return v


@pytest.mark.parametrize("k", [1, 5, 16, 37, 64])
def test_ppr_matches_oracle_per_vector(ctx, k):
    g = ctx.rmat(12, 16, seed=8)
    n, ind, col, deg = O.rmat_csc(12, 16, seed=8)
    seeds = O.ppr_sources(n, k, 8, seed=k)
    X, its, info = g.ppr(seeds, D, 1e-4, 30)
    assert X.shape == (n, k)
    for j in range(k):
        ref, it, tr, _ = O.pr_v10(n, ind, col, deg, D, 1e-4, 30, teleport=teleport(n, seeds[j]))
        # The batched solve reads a row's old score back as (x / deg) * deg, within an ulp of x,
        # so its L1 change can differ from the plain one by up to ~2^-23 * sum(x) < 1.2e-7.  No
        # sweep error of these vectors lies that close to the threshold, so every vector must
        # stop at the oracle's sweep.
        assert np.min(np.abs(tr - 1e-4)) > 1.2e-7, (j, tr)
        assert its[j] == it, (j, its[j], it)
        assert np.max(np.abs(X[:, j].astype(np.float64) - ref)) <= 1e-6


@pytest.mark.parametrize("tile", [8, 32])
def test_ppr_degree_ordered_stops_launching_rows_without_in_edges(ctx, tile):
    """On the degree-ordered copy every row without in-edges comes last, and from the third
    sweep on the batched solve launches only the row workgroups before them (their values
    are fixed after one sweep).  Every vector must still stop at the oracle's sweep with the
    oracle's scores, including vectors seeded on rows without in-edges."""
    g = ctx.rmat(12, 16, seed=8)
    plan = g.get_plan()
    plan.update(relabel=1, ppr_tile=tile)
    g.set_plan(**plan)
    n, ind, col, deg = O.rmat_csc(12, 16, seed=8)
    seeds = O.ppr_sources(n, 16, 8, seed=4)
    no_in = np.flatnonzero(np.diff(ind.astype(np.int64)) == 0)
    assert no_in.size > n // 4
    seeds[3] = np.sort(no_in[:8]).astype(seeds[3].dtype)  # a vector seeded only on rows without in-edges
    X, its, _ = g.ppr(seeds, D, 1e-4, 30)
    for j in range(16):
        ref, it, tr, _ = O.pr_v10(n, ind, col, deg, D, 1e-4, 30, teleport=teleport(n, seeds[j]))
        assert np.min(np.abs(tr - 1e-4)) > 1.2e-7, (j, tr)
        assert its[j] == it, (j, its[j], it)
        assert np.max(np.abs(X[:, j].astype(np.float64) - ref)) <= 1e-6


@pytest.mark.parametrize("tile", [8, 16, 32, 64])
def test_ppr_tiles_agree(ctx, tile):
    g = ctx.rmat(11, 16, seed=3)
    plan = g.get_plan()
    plan["ppr_tile"] = tile
    g.set_plan(**plan)
    n, ind, col, deg = O.rmat_csc(11, 16, seed=3)
    seeds = O.ppr_sources(n, 24, 8, seed=1)
    X, its, _ = g.ppr(seeds, D, 1e-5, 40)
    for j in (0, 7, 23):
        ref, it, _, _ = O.pr_v10(n, ind, col, deg, D, 1e-5, 40, teleport=teleport(n, seeds[j]))
        assert its[j] == it and np.max(np.abs(X[:, j] - ref)) <= 1e-6


def test_ppr_with_hub_rows_and_single_solve(ctx):
    # a small hub threshold routes many rows through the segment path
    g = ctx.rmat(13, 16, seed=5)
    g.set_plan(short_max=8, group_max=128, hub_min=256, hub_segment=256)
    n, ind, col, deg = O.rmat_csc(13, 16, seed=5)
    seeds = O.ppr_sources(n, 8, 8, seed=2)
    X, its, _ = g.ppr(seeds, D, 1e-4, 20)
    for j in range(8):
        v = teleport(n, seeds[j])
        single, info = g.pagerank(D, 1e-4, 20, teleport=v)
        assert info["iterations"] == its[j]
        assert np.max(np.abs(X[:, j] - single)) <= 1e-6


def test_ppr_autotune_picks_a_tile(ctx):
    g = ctx.rmat(12, 16, seed=2)
    seeds = O.ppr_sources(g.n, 64, 8, seed=3)
    t = g.ppr_autotune(seeds)
    assert t in (8, 16, 32, 64) and g.get_plan()["ppr_tile"] == t
    X, its, _ = g.ppr(seeds, D, 1e-4, 20)
    n, ind, col, deg = O.rmat_csc(12, 16, seed=2)
    ref, it, _, _ = O.pr_v10(n, ind, col, deg, D, 1e-4, 20, teleport=teleport(n, seeds[11]))
    assert its[11] == it and np.max(np.abs(X[:, 11] - ref)) <= 1e-6


def test_ppr_reuses_state_across_calls_and_plans(ctx):
    """Back-to-back solves on one graph reuse its cached segments and buffers: a larger k,
    then a smaller k, then a relabelled plan (segments rebuilt) and a new tile all agree
    with the oracle."""
    g = ctx.rmat(12, 16, seed=6)
    n, ind, col, deg = O.rmat_csc(12, 16, seed=6)

    def check(seeds, js):
        X, its, _ = g.ppr(seeds, D, 1e-4, 25)
        for j in js:
            ref, it, _, _ = O.pr_v10(n, ind, col, deg, D, 1e-4, 25, teleport=teleport(n, seeds[j]))
            assert its[j] == it and np.max(np.abs(X[:, j] - ref)) <= 1e-6

    s64 = O.ppr_sources(n, 64, 8, seed=11)
    check(s64, (0, 33, 63))
    s5 = O.ppr_sources(n, 5, 3, seed=12)
    check(s5, range(5))
    plan = g.get_plan()
    plan.update(relabel=1, ppr_tile=16, hub_min=512, hub_segment=512)
    g.set_plan(**plan)
    check(s64, (1, 40, 62))
    check(s5, range(5))


def test_ppr_full_size_linearity_rmat22(ctx):
    """k=64 on RMAT-22, 20 sweeps.  A solve is affine in the teleport vector and every vector
    starts from the same 1/N, so the vector over A u B (16 seeds, 1/16 each) must be the mean
    of the vectors over A and over B (8 seeds each): 21 such triples, within 1e-5 in L1.  One
    vector is also checked against the oracle on the device-built CSC."""
    import os

    g = ctx.rmat(22, 16, seed=1)
    n = g.n
    rng = np.random.default_rng(22)
    ids = rng.choice(n, 21 * 16 + 1, replace=False).astype(np.uint32)
    seeds = []
    for t in range(21):
        a, b = np.sort(ids[16 * t:16 * t + 8]), np.sort(ids[16 * t + 8:16 * t + 16])
        seeds += [a, b, np.sort(np.concatenate([a, b]))]
    seeds.append(ids[-1:])
    X, its, _ = g.ppr(seeds, D, 0.0, 20)
    assert X.shape == (n, 64) and np.all(its == 20)
    for t in range(21):
        xa, xb, xab = (X[:, 3 * t + q].astype(np.float64) for q in range(3))
        assert np.abs(xab - 0.5 * (xa + xb)).sum() < 1e-5, t
    ptr, idx, deg = g.csc()
    j = 63
    ref, it, _, _ = O.pr_v10(n, ptr, idx, deg, D, 0.0, 20, teleport=teleport(n, seeds[j]),
                             nthreads=min(16, os.cpu_count() or 1))
    assert it == 20 and np.max(np.abs(X[:, j].astype(np.float64) - ref)) <= 1e-6


def test_ppr_rejects_bad_seeds(ctx):
    from nwgraph_amd import NwgError

    g = ctx.rmat(8, 4, seed=1)
    with pytest.raises(NwgError):
        g.ppr([np.array([1, 1], np.uint32)])  # duplicate seed
    with pytest.raises(NwgError):
        g.ppr([np.array([g.n + 3], np.uint32)])  # out of range


@pytest.mark.parametrize("binned", [0, 1])
def test_multigraph_matches_per_graph_oracle(ctx, binned):
    """24 graphs of scales 4..12; with binned=1 one 2^13-row bin holds many graphs (more
    than the 64 a workgroup counts in LDS), each frozen once converged."""
    rng = np.random.default_rng(11)
    scales = rng.integers(4, 13, 24)
    graphs, refs = [], []
    for q, sc in enumerate(scales):
        graphs.append(ctx.rmat(int(sc), 16, seed=100 + q))
        refs.append(O.rmat_csc(int(sc), 16, seed=100 + q))
    mg = ctx.multigraph(graphs)
    assert mg.ngraphs == len(graphs) and mg.n == sum(g.n for g in graphs)
    mg.set_plan(binned)
    X, its, errs, info = mg.pagerank(D, 1e-4, 50)
    converged = 0
    for q, (n, ind, col, deg) in enumerate(refs):
        ref, it, tr, _ = O.pr_v10(n, ind, col, deg, D, 1e-4, 50)
        xs = X[mg.offsets[q]:mg.offsets[q + 1]]
        assert its[q] == it, (q, its[q], it)
        assert np.max(np.abs(xs.astype(np.float64) - ref)) <= 1e-6
        assert abs(errs[q] - tr[-1]) <= 1e-3 * tr[-1] + 1e-12
        converged += it < 50
    assert converged > 0  # per-graph stopping is exercised


@pytest.mark.parametrize("binned", [0, 1])
def test_multigraph_with_hub_items_and_repeat(ctx, binned):
    # a star whose centre has 5999 in-edges (more than one LDS work item holds) plus a
    # ring, next to an RMAT graph and a grid
    n0 = 6000
    s = np.concatenate([np.arange(1, n0), np.arange(n0)]).astype(np.uint32)
    d = np.concatenate([np.zeros(n0 - 1), (np.arange(n0) + 1) % n0]).astype(np.uint32)
    star = ctx.graph_from_edges(n0, s, d)
    cs, cd = O.clean_edges(s, d)
    ind0, col0 = O.csc(n0, cs, cd)
    graphs = [star, ctx.rmat(9, 8, seed=2), ctx.grid(64, seed=3)]
    refs = [(n0, ind0, col0, O.out_degrees(n0, ind0, col0)), O.rmat_csc(9, 8, seed=2), O.grid_csc(64, seed=3)]
    assert int(np.diff(ind0.astype(np.int64)).max()) > 4096
    mg = ctx.multigraph(graphs)
    mg.set_plan(binned)
    X1, its1, _, _ = mg.pagerank(D, 1e-4, 100)
    X2, its2, _, _ = mg.pagerank(D, 1e-4, 100)
    assert np.array_equal(X1, X2) and np.array_equal(its1, its2)
    for q, (n, ind, col, deg) in enumerate(refs):
        # the stopping sweep follows the FP64 iteration: the oracle's sequential FP32 sum over
        # the 5999-term hub row carries enough rounding noise (propagated round the ring) to
        # keep its L1 change above 1e-4 two sweeps longer than FP64 and the tree-summed GPU
        x64, it64 = O.fp64_jacobi(n, ind, col, deg, tol=1e-4, max_iters=100)
        assert its1[q] == it64
        assert np.max(np.abs(X1[mg.offsets[q]:mg.offsets[q + 1]] - x64)) <= 1e-6
        # at a fixed sweep count the GPU is at least as close to FP64 as the reference's
        # sequential FP32 loop (whose hub-row sum drifts by ~3e-5 on the star)
        k = int(its1[q])
        ref, _, _, _ = O.pr_v10(n, ind, col, deg, D, 0.0, k)
        x64k, _ = O.fp64_jacobi(n, ind, col, deg, tol=0.0, max_iters=k)
        gpu_err = np.max(np.abs(X1[mg.offsets[q]:mg.offsets[q + 1]] - x64k))
        ref_err = np.max(np.abs(ref - x64k))
        assert gpu_err <= 1e-6 and gpu_err <= ref_err + 1e-7, (gpu_err, ref_err)


def test_multigraph_autotune_keeps_results(ctx):
    graphs = [ctx.rmat(sc, 16, seed=40 + sc) for sc in (8, 11, 14, 16)]
    mg = ctx.multigraph(graphs)
    X0, its0, _, _ = mg.pagerank(D, 0.0, 20)  # fixed sweep count: no stopping decision to differ
    choice = mg.autotune()
    assert choice in (0, 1)
    X1, its1, _, _ = mg.pagerank(D, 0.0, 20)
    assert np.all(its0 == 20) and np.all(its1 == 20)
    assert np.max(np.abs(X0 - X1) / np.maximum(X0, 1e-30)) < 1e-5  # summation order only


@pytest.mark.parametrize("binned", [0, 1])
def test_multigraph_256_graphs_match_single_solves(ctx, binned):
    """The benchmark's union: 256 RMAT graphs of scales 10..18 (the same seeds), 20 sweeps.
    Every graph's slice of the union equals its own single-graph solve (another kernel and
    another summation order) within 1e-6, and the largest graph also equals the oracle."""
    import os

    rng = np.random.default_rng(1)
    scales = rng.integers(10, 19, 256)
    graphs = [ctx.rmat(int(s), 16, seed=1000 + q) for q, s in enumerate(scales)]
    mg = ctx.multigraph(graphs)
    mg.set_plan(binned)
    X, its, _, _ = mg.pagerank(D, 0.0, 20)
    assert np.all(its == 20)
    for q, g in enumerate(graphs):
        x, info = g.pagerank(D, 0.0, 20)
        assert info["iterations"] == 20
        assert np.max(np.abs(X[mg.offsets[q]:mg.offsets[q + 1]].astype(np.float64) - x)) <= 1e-6, q
    q = int(np.argmax([g.nnz for g in graphs]))
    ptr, idx, deg = graphs[q].csc()
    ref, it, _, _ = O.pr_v10(graphs[q].n, ptr, idx, deg, D, 0.0, 20, nthreads=min(16, os.cpu_count() or 1))
    assert it == 20 and np.max(np.abs(X[mg.offsets[q]:mg.offsets[q + 1]].astype(np.float64) - ref)) <= 1e-6
    mg.close()
    for g in graphs:
        g.close()
