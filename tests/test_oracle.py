"""Pins the CPU oracle (oracle/nwg_oracle.c) before it is trusted as the checker:
reference golden values, reference known-answer cases, published RNG vectors,
closed forms, invariants and brute force (numpy / scipy)."""
import numpy as np
import pytest
import scipy.sparse as sp

import pyoracle as O
from conftest import GOLDEN, golden


# ---------------------------------------------------------------- Philox
@pytest.mark.parametrize(
    "ctr,key,expect",
    [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ],
)
def test_philox_known_answers(ctr, key, expect):
    # Random123 kat_vectors for philox4x32_10
    assert list(O.philox(ctr, key)) == expect


# ---------------------------------------------------------------- generators
def test_rmat_invariants():
    scale, m = 12, 200000
    s, d = O.rmat_edges(scale, 0, m)
    assert s.max() < (1 << scale) and d.max() < (1 << scale)
    # subrange regeneration is identical (pure function of (index, seed))
    s2, d2 = O.rmat_edges(scale, 1000, 500)
    assert np.array_equal(s2, s[1000:1500]) and np.array_equal(d2, d[1000:1500])
    # a different seed gives a different stream
    s3, _ = O.rmat_edges(scale, 0, 1000, seed=2)
    assert not np.array_equal(s3, s[:1000])
    # top-level quadrant frequencies match A/B/C/D = .57/.19/.19/.05
    top = scale - 1
    sb, db = (s >> top) & 1, (d >> top) & 1
    freq = np.array([np.mean((sb == a) & (db == b)) for a, b in ((0, 0), (0, 1), (1, 0), (1, 1))])
    assert np.allclose(freq, [0.57, 0.19, 0.19, 0.05], atol=0.005)
    # and at the lowest level too (independent draws per level)
    freq0 = np.array([np.mean(((s & 1) == a) & ((d & 1) == b)) for a, b in ((0, 0), (0, 1), (1, 0), (1, 1))])
    assert np.allclose(freq0, [0.57, 0.19, 0.19, 0.05], atol=0.005)


def test_rmat_bit_placement_matches_quadrant_rule():
    # recompute a few edges straight from Philox words: level l uses word l%4 of
    # Philox({e, 0, l/4, 'RMAT'}, seed), most significant bit first
    scale, seed = 9, 77
    tA, tAB, tABC = [(p << 32) // 10000 for p in (5700, 7600, 9500)]
    s, d = O.rmat_edges(scale, 0, 16, seed=seed)
    for e in range(16):
        src = dst = 0
        for l in range(scale):
            if l % 4 == 0:
                w = O.philox([e, 0, l // 4, 0x524D4154], [seed & 0xFFFFFFFF, seed >> 32])
            x = int(w[l % 4])
            sb = 1 if x >= tAB else 0
            db = 0 if x < tA else (1 if x < tAB else (0 if x < tABC else 1))
            src |= sb << (scale - 1 - l)
            dst |= db << (scale - 1 - l)
        assert (src, dst) == (s[e], d[e])


def test_grid_invariants():
    W = 64
    s, d = O.grid_edges(W, keep=8500, long_range=50, seed=3)
    N = W * W
    S = 2 * W * (W - 1)
    assert s.size == 2 * S + N
    lat_s, lat_d = s[: 2 * S], d[: 2 * S]
    kept = lat_s != 0xFFFFFFFF
    # both directions stored for every kept lattice edge
    assert np.array_equal(lat_s[0::2] != 0xFFFFFFFF, lat_s[1::2] != 0xFFFFFFFF)
    assert np.array_equal(lat_s[0::2][kept[0::2]], lat_d[1::2][kept[0::2]])
    # lattice neighbours only
    a, b = lat_s[kept].astype(np.int64), lat_d[kept].astype(np.int64)
    diff = np.abs(a - b)
    assert np.all((diff == 1) | (diff == W))
    assert np.all((diff != 1) | (np.minimum(a, b) % W != W - 1))
    assert abs(kept[0::2].mean() - 0.85) < 0.02
    lr_s, lr_d = s[2 * S:], d[2 * S:]
    has = lr_s != 0xFFFFFFFF
    assert np.all(lr_s[has] == np.arange(N)[has])
    assert np.all(lr_d[has] != lr_s[has]) and np.all(lr_d[has] < N)
    assert 0 < has.sum() < 0.02 * N


def test_ppr_sources_distinct():
    ids = O.ppr_sources(1000, 64, 8, seed=5)
    assert ids.shape == (64, 8)
    for row in ids:
        assert len(set(row.tolist())) == 8 and row.max() < 1000


def test_edge_weight_range_and_formula():
    w = np.array([O.edge_weight(u, v, 9) for u in range(20) for v in range(20)], np.float32)
    assert w.min() >= 0.1 and w.max() < 1.0
    r = O.philox([3, 5, 3, 0x57474854], [9, 0])[0]
    t = np.float32(int(r) >> 8) * np.float32(2.0 ** -24)
    expect = np.float32(np.float64(np.float32(0.9)) * np.float64(t) + np.float64(np.float32(0.1)))
    assert np.float32(O.edge_weight(3, 5, 9)) == expect


# ---------------------------------------------------------------- builder
def test_builder_kats():
    for case in golden("builder_kat.json")["cases"]:
        e = np.array(case["edges"], np.uint32)
        s, d = e[:, 0], e[:, 1]
        n = case["n"]
        op = case.get("op")
        if op == "clean":
            cs, cd = O.clean_edges(s, d)
            assert cs.size == case["clean_count"], case["name"]
            assert np.all(cs != cd)
        elif op == "lexical_sort":
            ind, col = O.fill_sorted(0, n, s, d)
            rows = np.repeat(np.arange(n), np.diff(ind.astype(np.int64)))
            pairs = list(zip(rows.tolist(), col.tolist()))
            assert pairs[: len(case["sorted_prefix"])] == [tuple(p) for p in case["sorted_prefix"]], case["name"]
        else:
            ind, col = O.fill_sorted(case["idx"], n, s, d)
            assert np.diff(ind.astype(np.int64)).tolist() == case["row_sizes"], case["name"]


def test_clean_matches_numpy_bruteforce():
    rng = np.random.default_rng(0)
    s = rng.integers(0, 50, 5000).astype(np.uint32)
    d = rng.integers(0, 50, 5000).astype(np.uint32)
    cs, cd = O.clean_edges(s, d)
    pairs = np.unique(np.stack([s, d], 1), axis=0)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    assert np.array_equal(cs, pairs[:, 0]) and np.array_equal(cd, pairs[:, 1])


@pytest.mark.parametrize("idx", [0, 1])
def test_fill_matches_scipy(idx):
    rng = np.random.default_rng(1)
    n = 300
    s = rng.integers(0, n, 4000).astype(np.uint32)
    d = rng.integers(0, n, 4000).astype(np.uint32)
    ind, col = O.fill_sorted(idx, n, s, d)
    rows, ents = (d, s) if idx == 1 else (s, d)
    m = sp.coo_matrix((np.ones(s.size), (rows, ents)), shape=(n, n)).tocsr()
    m.sum_duplicates()
    # duplicates are kept by fill_directed: compare multiplicities
    assert np.array_equal(ind.astype(np.int64), np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=n))]))
    for r in range(0, n, 37):
        seg = col[ind[r]:ind[r + 1]]
        assert np.all(np.diff(seg.astype(np.int64)) >= 0)
        assert np.array_equal(np.sort(ents[rows == r]), seg)


def test_out_degrees_bincount():
    n, ind, col, deg = O.rmat_csc(10)
    assert np.array_equal(deg, np.bincount(col, minlength=n).astype(np.uint32))


# ---------------------------------------------------------------- PageRank
def karate_csc():
    g = golden("karate_pagerank.json")
    n = g["num_nodes"]
    s = np.array(g["src"], np.uint32)
    d = np.array(g["dst"], np.uint32)
    ind, col = O.csc(n, s, d)
    return g, n, ind, col, O.out_degrees(n, ind, col)


def test_v1_karate_golden():
    g, n, ind, col, deg = karate_csc()
    pr, it = O.pr_v1(n, ind, col, deg, g["damping"], g["threshold"], g["max_iters"])
    ans = np.array(g["pagerank"])
    assert np.all(np.abs(pr - ans) <= g["rel_tolerance"] * np.abs(ans))


def test_v10_float_karate_close_to_golden():
    g, n, ind, col, deg = karate_csc()
    pr, it, tr, _ = O.pr_v10(n, ind, col, deg, 0.85, 1e-7, 100)
    ans = np.array(g["pagerank"])
    assert np.all(np.abs(pr - ans) <= g["rel_tolerance"] * np.abs(ans))
    pr1, _ = O.pr_v1(n, ind, col, deg, 0.85, 1e-7, 100)
    assert np.max(np.abs(pr - pr1)) < 1e-6


def test_v13_karate_golden():
    """v13 (delta PageRank, experimental/algorithms/page_rank.hpp:560-625) keeps un-normalised
    scores: with a tiny threshold, scores / N reach the reference's karate golden values."""
    g, n, ind, col, deg = karate_csc()
    pr, it, pushed = O.pr_v13(n, ind, col, deg, 0.85, 1e-9, 1000)
    ans = np.array(g["pagerank"])
    assert 50 < it < 1000 and pushed[-1] == 0 and np.all(pushed[:-1] > 0)
    assert np.all(np.abs(pr / n - ans) <= g["rel_tolerance"] * np.abs(ans))


@pytest.mark.parametrize("iters", [1, 2, 5, 12])
def test_v13_threshold_zero_is_the_neumann_partial_sum(iters):
    """With threshold 0 every positive residual is absorbed, so after k iterations the scores
    are sum_{t<k} (d A)^t (1-d) 1 (A[i,u] = 1/outdeg(u) per edge u->i): dense matrix powers."""
    n, ind, col, deg = O.rmat_csc(8, 8, seed=5)
    pr, it, _ = O.pr_v13(n, ind, col, deg, 0.85, 0.0, iters)
    assert it == iters
    d = float(np.float32(0.85))
    A = O.dense_matrix(n, ind, col, deg)
    A[:, deg == 0] = 0.0
    r = np.full(n, float(np.float32(1.0 - d)))
    acc = np.zeros(n)
    for _ in range(iters):
        acc += r
        r = d * (A @ r)
    assert np.allclose(pr, acc, rtol=2e-5, atol=1e-7)


def test_v13_cycle_closed_form():
    """On a directed cycle every residual is (1-d) d^t at iteration t; it is absorbed while
    above the threshold, and the loop stops after the first iteration that absorbs nothing."""
    n, thr = 100, 1e-3
    s = np.arange(n, dtype=np.uint32)
    ind, col = O.csc(n, s, ((s + 1) % n).astype(np.uint32))
    deg = O.out_degrees(n, ind, col)
    pr, it, pushed = O.pr_v13(n, ind, col, deg, 0.85, thr, 100)
    d = float(np.float32(0.85))
    rt = [(1.0 - d) * d ** t for t in range(100)]
    stop = next(t for t, r in enumerate(rt) if r <= thr)
    assert it == stop + 1 and np.all(pushed[:stop] == n) and pushed[stop] == 0
    assert np.allclose(pr, sum(rt[:stop]), rtol=1e-5)


def test_v13_drops_unabsorbed_residuals():
    """The pull phase overwrites every residual (page_rank.hpp:608-611), so a residual at or
    below the threshold that was not absorbed is lost.  Vertices: a=0 (out-degree 20, to b
    and 19 sinks), b=1 (a sink), y=2 -> z=3, z (out-degree 10, to b and 9 sinks); threshold
    0.02.  Iteration 0: all absorb r0 = 1 - d; b then holds r0 d/20 + r0 d/10 = 0.0191.
    Iteration 1: b keeps it un-absorbed; z absorbs y's r0 d and pushes r0 d^2/10 = 0.0108,
    which replaces b's 0.0191 (kept, it would have summed to 0.0300 and been absorbed).
    Iteration 2 pushes nothing, so the loop stops after 3 iterations with pr[b] = r0 and
    pr[z] = r0 (1 + d)."""
    s_ = [0] * 20 + [3] * 10 + [2]
    d_ = [1] + list(range(4, 23)) + [1] + list(range(23, 32)) + [3]
    n = 32
    ind, col = O.csc(n, np.array(s_, np.uint32), np.array(d_, np.uint32))
    deg = O.out_degrees(n, ind, col)
    assert deg[0] == 20 and deg[3] == 10 and deg[2] == 1 and deg[1] == 0
    pr, it, pushed = O.pr_v13(n, ind, col, deg, 0.85, 0.02, 20)
    r0 = np.float32(1.0 - float(np.float32(0.85)))
    assert it == 3 and list(pushed) == [3, 1, 0]
    assert pr[1] == r0
    assert np.isclose(pr[3], float(r0) * (1.0 + float(np.float32(0.85))), rtol=1e-6)


def test_v10_regular_digraph_is_uniform():
    # a directed cycle: every vertex has in/out degree 1, PageRank = 1/N exactly each sweep
    n = 1000
    s = np.arange(n, dtype=np.uint32)
    d = ((s + 1) % n).astype(np.uint32)
    ind, col = O.csc(n, s, d)
    deg = O.out_degrees(n, ind, col)
    pr, it, tr, _ = O.pr_v10(n, ind, col, deg, 0.85, 1e-4, 20)
    assert it == 1 and np.allclose(pr, 1.0 / n, rtol=1e-6)


def test_v10_isolated_vertex_gets_base():
    n = 10
    s = np.array([0, 1, 2], np.uint32)
    d = np.array([1, 2, 0], np.uint32)
    ind, col = O.csc(n, s, d)
    deg = O.out_degrees(n, ind, col)
    pr, it, _, _ = O.pr_v10(n, ind, col, deg, 0.85, 1e-12, 5)
    base = np.float32((1.0 - np.float64(np.float32(0.85))) / n)
    assert np.all(pr[3:] == base)


def test_v10_matches_dense_power_iteration():
    n, ind, col, deg = O.rmat_csc(8, 8)
    A = O.dense_matrix(n, ind, col, deg)
    x = np.full(n, 1.0 / n)
    for _ in range(15):
        x = (1 - 0.85) / n + 0.85 * A @ x
    pr, it, _, _ = O.pr_v10(n, ind, col, deg, 0.85, 0.0, 15)
    assert it == 15 and np.max(np.abs(pr - x)) < 1e-6


def test_v10_thread_count_does_not_change_result():
    n, ind, col, deg = O.rmat_csc(12)
    a = O.pr_v10(n, ind, col, deg, nthreads=1)
    b = O.pr_v10(n, ind, col, deg, nthreads=4)
    assert np.array_equal(a[0], b[0]) and a[1] == b[1] and np.array_equal(a[2], b[2])


def test_v11_blocks_limits():
    n, ind, col, deg = O.rmat_csc(10)
    # nblocks = n (every row its own block) is Jacobi = v10, bit for bit
    jac, it_j, tr_j = O.pr_v11(n, ind, col, deg, nblocks=n)
    v10, it_10, tr_10, _ = O.pr_v10(n, ind, col, deg)
    assert it_j == it_10 and np.array_equal(jac, v10)
    # nblocks = 1 is sequential Gauss-Seidel: compare with a plain Python restatement
    n2, ind2, col2, deg2 = O.rmat_csc(7, 8)
    gs, it_gs, _ = O.pr_v11(n2, ind2, col2, deg2, max_iters=6, threshold=0.0)
    x = np.full(n2, np.float32(1.0 / n2), np.float32)
    base = np.float32((1.0 - np.float64(np.float32(0.85))) / n2)
    contrib = (x / deg2.astype(np.float32)).astype(np.float32)
    d32 = np.float32(0.85)
    with np.errstate(divide="ignore"):
        for _ in range(6):
            for i in range(n2):
                z = np.float32(0)
                for p in range(ind2[i], ind2[i + 1]):
                    z = np.float32(z + contrib[col2[p]])
                # exact product, one double rounding of the sum, then float: equals fmaf
                # unless the double sum lands on a float rounding tie (not the case here)
                x[i] = np.float32(np.float64(d32) * np.float64(z) + np.float64(base))
                contrib[i] = np.float32(x[i] / np.float32(deg2[i]))
    assert it_gs == 6 and np.array_equal(gs, x)


@pytest.mark.parametrize("threshold", [1e-4, 0.0])
def test_v11_parallel_one_thread_is_sequential_gauss_seidel(threshold):
    """nwo_pr_v11_parallel (the CPU-baseline loop, page_rank.hpp:124-139) on one thread is the
    sequential in-place sweep: bit for bit nwo_pr_v11_blocks(nblocks=1), which the pure-Python
    Gauss-Seidel restatement above pins.  A Jacobi write (contrib not updated in place), an
    unfused update or a dropped term in the error would each break the equality."""
    n, ind, col, deg = O.rmat_csc(12)
    gs, it_gs, tr_gs = O.pr_v11(n, ind, col, deg, threshold=threshold, max_iters=20, nblocks=1)
    par, it_p, tr_p, sec = O.pr_v11_parallel(n, ind, col, deg, threshold=threshold, max_iters=20, nthreads=1)
    assert it_p == it_gs and np.array_equal(par, gs) and np.array_equal(tr_p, tr_gs)
    assert sec.size == it_p and np.all(sec > 0)
    # and it is not the Jacobi iteration
    jac, _, _, _ = O.pr_v10(n, ind, col, deg, threshold=threshold, max_iters=20)
    assert not np.array_equal(par, jac)


def test_v11_parallel_threads_reach_the_v10_fixed_point():
    """With 8 threads the in-place loop is a schedule-dependent block Gauss-Seidel; it must
    converge to the same fixed point as v10, and its error trace must fall."""
    n, ind, col, deg = O.rmat_csc(12)
    par, it, tr, _ = O.pr_v11_parallel(n, ind, col, deg, threshold=1e-9, max_iters=500, nthreads=8)
    jac, _, _, _ = O.pr_v10(n, ind, col, deg, threshold=1e-9, max_iters=500)
    x64, _ = O.fp64_jacobi(n, ind, col, deg)
    assert it < 500 and tr[-1] < 1e-9
    assert np.max(np.abs(par - jac)) < 1e-7
    assert np.max(np.abs(par - x64)) < 1e-7
    # the error falls sweep after sweep until it reaches float rounding noise (~1e-8 here)
    above = tr[1:][tr[1:] > 1e-6]
    assert above.size >= 10 and np.all(np.diff(above) < 0)


def test_normal_vector_moments():
    """nwo_normal_vector (Box-Muller over Philox words) against the moments of N(0, 1) over
    10^6 draws (bounds ~7 standard errors), a Kolmogorov-Smirnov test, independence of the
    two values of each Box-Muller pair, determinism per seed, and an odd length."""
    import scipy.stats as st

    x = O.normal_vector(1_000_000, seed=3)
    assert np.all(np.isfinite(x))
    assert abs(x.mean()) < 7e-3
    assert abs(x.var() - 1.0) < 1e-2
    assert abs(st.skew(x)) < 1.7e-2
    assert abs(st.kurtosis(x)) < 3.5e-2
    assert st.kstest(x, "norm").pvalue > 1e-3
    assert abs(np.corrcoef(x[0::2], x[1::2])[0, 1]) < 7e-3
    assert np.array_equal(x[:1001], O.normal_vector(1001, seed=3))
    assert not np.array_equal(x[:1000], O.normal_vector(1000, seed=4))


def test_v11_and_v10_converge_to_same_fixed_point():
    n, ind, col, deg = O.rmat_csc(10)
    gs, _, _ = O.pr_v11(n, ind, col, deg, threshold=1e-9, max_iters=500)
    jac, _, _, _ = O.pr_v10(n, ind, col, deg, threshold=1e-9, max_iters=500)
    assert np.max(np.abs(gs - jac)) < 1e-7


def test_fp64_jacobi_and_adjoint_match_dense_solve():
    n, ind, col, deg = O.rmat_csc(9, 8)
    A = O.dense_matrix(n, ind, col, deg)
    M = np.eye(n) - 0.85 * A
    x_dense = np.linalg.solve(M, np.full(n, 0.15 / n))
    x, _ = O.fp64_jacobi(n, ind, col, deg)
    assert np.max(np.abs(x - x_dense)) < 1e-12
    c = O.normal_vector(n, seed=4)
    lam_dense = np.linalg.solve(M.T, c)
    lam, _ = O.fp64_adjoint(n, ind, col, deg, c)
    assert np.max(np.abs(lam - lam_dense)) < 1e-10 * np.max(np.abs(lam_dense))
    # weighted variant with a random row-normalised weight matrix
    rng = np.random.default_rng(2)
    w = rng.uniform(0.1, 1.0, col.size)
    rowsum = np.bincount(col, weights=w, minlength=n)
    wn = w / rowsum[col]
    Aw = O.dense_matrix(n, ind, col, deg, w=wn)
    v = rng.uniform(size=n)
    v /= v.sum()
    xw_dense = np.linalg.solve(np.eye(n) - 0.85 * Aw, 0.15 * v)
    xw, _ = O.fp64_jacobi(n, ind, col, deg, w=wn, v=v)
    assert np.max(np.abs(xw - xw_dense)) < 1e-12


# ---------------------------------------------------------------- readers
def test_read_mm_karate_matches_reference_edge_list():
    g = golden("karate_pagerank.json")
    n, s, d = O.read_mm(GOLDEN + "/karate.mtx")
    assert n == 34 and s.size == 156
    got = sorted(zip(s.tolist(), d.tolist()))
    want = sorted(zip(g["src"], g["dst"]))
    assert got == want


def test_read_nw_layout(tmp_path):
    # writer that follows edge_list.hpp:237-280 byte for byte
    p = tmp_path / "g.nw"
    s = np.array([0, 1, 2, 5], np.uint32)
    d = np.array([1, 2, 0, 3], np.uint32)
    with open(p, "wb") as f:
        f.write(b"NW Graph index_edge_list\0\0\0"[:27])
        f.write(np.int32(1).tobytes())
        f.write(np.uint64(9).tobytes())
        f.write(b"\0")
        f.write(np.uint64(4).tobytes() + np.uint64(2).tobytes())
        for a in (s, d):
            f.write(np.uint64(4).tobytes() + np.uint64(4).tobytes() + a.tobytes())
    n, s2, d2 = O.read_nw(p)
    assert n == 9 and np.array_equal(s, s2) and np.array_equal(d, d2)
