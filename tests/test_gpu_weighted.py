"""GPU weighted PageRank forward, implicit adjoint and edge / teleport gradients against
FP64 dense solves and FP64 central finite differences; autograd wrapper."""
import numpy as np
import pytest
import scipy.linalg as sla

import pyoracle as O

pytestmark = pytest.mark.gpu

D = 0.85


def setup_graph(ctx, scale, seed=1, wseed=5):
    g = ctx.rmat(scale, 16, seed=seed)
    n, ind, col, deg = O.rmat_csc(scale, 16, seed=seed)
    rows = np.repeat(np.arange(n), np.diff(ind.astype(np.int64)))
    w_raw = np.array([O.edge_weight(u, v, wseed) for u, v in zip(col.tolist(), rows.tolist())], np.float32)
    return g, n, ind, col, deg, rows, w_raw


def dense_system(n, ind, col, rows, w_raw):
    s = np.bincount(col, weights=w_raw.astype(np.float64), minlength=n)
    wn = w_raw.astype(np.float64) / s[col]
    A = np.zeros((n, n))
    np.add.at(A, (rows, col), wn)
    return A, wn, s


def test_random_weights_match_oracle_and_normalise(ctx):
    g, n, ind, col, deg, rows, w_raw = setup_graph(ctx, 10)
    g.random_weights(seed=5, normalize=True)
    w = g.get_weights()
    s = np.bincount(col, weights=w_raw.astype(np.float64), minlength=n)
    expect = (w_raw / s[col].astype(np.float32)).astype(np.float32)
    assert np.max(np.abs(w - expect) / expect) < 1e-6
    sums = np.bincount(col, weights=w.astype(np.float64), minlength=n)
    assert np.allclose(sums[deg > 0], 1.0, atol=1e-5)


def test_forward_matches_dense_fp64(ctx):
    g, n, ind, col, deg, rows, w_raw = setup_graph(ctx, 12)
    g.set_weights(w_raw, normalize=True)
    A, wn, _ = dense_system(n, ind, col, rows, w_raw)
    rng = np.random.default_rng(3)
    v = rng.uniform(size=n)
    v = (v / v.sum()).astype(np.float32)
    x, info = g.wpr_forward(v, D, 1e-7, 300)
    x_dense = np.linalg.solve(np.eye(n) - D * A, (1 - D) * v.astype(np.float64))
    assert np.linalg.norm(x - x_dense, 1) / np.linalg.norm(x_dense, 1) < 1e-5
    # same fixed point as the oracle's FP64 Jacobi
    xj, _ = O.fp64_jacobi(n, ind, col, deg, w=wn, v=v.astype(np.float64))
    assert np.max(np.abs(xj - x_dense)) < 1e-12


def test_adjoint_gradients_match_dense_fp64(ctx):
    g, n, ind, col, deg, rows, w_raw = setup_graph(ctx, 12)
    g.set_weights(w_raw, normalize=True)
    A, wn, s = dense_system(n, ind, col, rows, w_raw)
    v = np.full(n, 1.0 / n, np.float32)
    c = O.normal_vector(n, seed=7)
    x, _ = g.wpr_forward(v, D, 1e-7, 300)
    M = np.eye(n) - D * A
    x_d = np.linalg.solve(M, (1 - D) * v.astype(np.float64))
    lam = np.linalg.solve(M.T, c)
    gw_norm = D * lam[rows] * x_d[col]          # dL/dw for the normalised weights
    gv = (1 - D) * lam
    gw_gpu, gv_gpu, info = g.wpr_backward(x, c.astype(np.float32), D, 1e-7, 300, raw_weights=False)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(gw_gpu, gw_norm) < 1e-4
    assert rel(gv_gpu, gv) < 1e-4
    # chained to the raw weights: (g - sum_j g_uj w_uj) / s_u
    t = np.bincount(col, weights=gw_norm * wn, minlength=n)
    gw_raw = (gw_norm - t[col]) / s[col]
    gw_raw_gpu, _, _ = g.wpr_backward(x, c.astype(np.float32), D, 1e-7, 300, raw_weights=True)
    assert rel(gw_raw_gpu, gw_raw) < 1e-4


def test_raw_gradient_matches_central_finite_differences(ctx):
    g, n, ind, col, deg, rows, w_raw = setup_graph(ctx, 12, seed=2, wseed=9)
    g.set_weights(w_raw, normalize=True)
    v = np.full(n, 1.0 / n)
    c = O.normal_vector(n, seed=11)
    x, _ = g.wpr_forward(v.astype(np.float32), D, 1e-7, 300)
    gw_gpu, _, _ = g.wpr_backward(x, c.astype(np.float32), D, 1e-7, 300, raw_weights=True)
    A, wn, s = dense_system(n, ind, col, rows, w_raw)
    lu = sla.lu_factor(np.eye(n) - D * A)
    b = (1 - D) * v
    x0 = sla.lu_solve(lu, b)

    def loss_with(p, delta):
        # raw weight p -> p + delta changes column u = col[p] of A (all out-edges of u)
        u = col[p]
        w2 = w_raw.astype(np.float64).copy()
        w2[p] += delta
        sel = np.nonzero(col == u)[0]
        s2 = w2[sel].sum()
        dcol = np.zeros(n)
        np.add.at(dcol, rows[sel], w2[sel] / s2 - w_raw[sel].astype(np.float64) / s[u])
        # Sherman-Morrison on M' = M - d * dcol e_u^T
        z = sla.lu_solve(lu, D * dcol)
        xn = x0 + z * x0[u] / (1.0 - z[u])
        return float(c @ xn)

    rng = np.random.default_rng(0)
    heavy = np.argsort(-np.abs(gw_gpu))[:10]
    sample = np.concatenate([heavy, rng.integers(0, col.size, 10)])
    h = 1e-6
    fd = np.array([(loss_with(p, h) - loss_with(p, -h)) / (2 * h) for p in sample])
    got = gw_gpu[sample].astype(np.float64)
    assert np.linalg.norm(got - fd) / np.linalg.norm(fd) < 1e-4


def test_gradients_at_tolerance_1e4_match_fp64_with_the_same_sweeps(ctx):
    """The task's setting: forward and adjoint both stopped at L1 tolerance 1e-4.  The FP64
    side runs the same Jacobi iterations for the same numbers of sweeps (forward from 1/n,
    adjoint from grad_x), so the comparison isolates the FP32 arithmetic: edge and teleport
    gradients within 1e-4 relative (L2).  The distance to the converged gradient (the
    truncation of the solves themselves) is checked to be what the tolerance implies."""
    g, n, ind, col, deg, rows, w_raw = setup_graph(ctx, 12, seed=4, wseed=2)
    g.set_weights(w_raw, normalize=True)
    A, wn, s = dense_system(n, ind, col, rows, w_raw)
    v = np.full(n, 1.0 / n, np.float32)
    c = O.normal_vector(n, seed=17)
    x, fi = g.wpr_forward(v, D, 1e-4, 100)
    gw_gpu, gv_gpu, bi = g.wpr_backward(x, c.astype(np.float32), D, 1e-4, 100, raw_weights=False)
    kf, kb = fi["iterations"], bi["iterations"]
    assert 1 < kf < 100 and 1 < kb < 100
    dd = float(np.float32(D))
    x64, itf = O.fp64_jacobi(n, ind, col, deg, w=wn, v=v.astype(np.float64), damping=dd, tol=0.0, max_iters=kf)
    lam64, itb = O.fp64_adjoint(n, ind, col, deg, c, w=wn, damping=dd, tol=0.0, max_iters=kb)
    assert itf == kf and itb == kb
    gw64 = dd * lam64[rows] * x64[col]
    gv64 = (1.0 - dd) * lam64
    rel = lambda a, b: np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b)
    assert rel(gw_gpu, gw64) < 1e-4
    assert rel(gv_gpu, gv64) < 1e-4
    # the stopping decisions of FP64 with the same tolerance (within one sweep: an error
    # within FP32 rounding of the threshold may fall on either side)
    _, itf_tol = O.fp64_jacobi(n, ind, col, deg, w=wn, v=v.astype(np.float64), damping=dd, tol=1e-4, max_iters=100)
    _, itb_tol = O.fp64_adjoint(n, ind, col, deg, c, w=wn, damping=dd, tol=1e-4, max_iters=100)
    assert abs(itf_tol - kf) <= 1 and abs(itb_tol - kb) <= 1
    # and against the converged gradient: the truncation error a 1e-4 L1 tolerance leaves
    M = np.eye(n) - D * A
    lam = np.linalg.solve(M.T, c)
    xd = np.linalg.solve(M, (1 - D) * v.astype(np.float64))
    assert rel(gw_gpu, D * lam[rows] * xd[col]) < 1e-2


def test_adjoint_with_large_gradient_matches_dense_fp64(ctx):
    """A loss whose gradient is large (|grad_x| ~ 1e3 per vertex) with a positive threshold:
    lambda matches the dense FP64 solution and the solve stops within a sweep of an FP64 solve
    with the same threshold (the threshold is set above the FP32 rounding floor of the L1
    change, n * ulp(lambda))."""
    g, n, ind, col, deg, rows, w_raw = setup_graph(ctx, 12, seed=6, wseed=3)
    g.set_weights(w_raw, normalize=True)
    A, wn, _ = dense_system(n, ind, col, rows, w_raw)
    v = np.full(n, 1.0 / n, np.float32)
    c = (1000.0 * O.normal_vector(n, seed=5)).astype(np.float32)
    x, _ = g.wpr_forward(v, D, 1e-7, 300)
    thr = 10.0
    _, gv, bi = g.wpr_backward(x, c, D, thr, 400, raw_weights=False)
    dd = float(np.float32(D))
    lam_it, it64 = O.fp64_adjoint(n, ind, col, deg, c.astype(np.float64), w=wn, damping=dd, tol=thr, max_iters=400)
    assert bi["iterations"] > 40
    assert abs(bi["iterations"] - it64) <= 1, (bi["iterations"], it64)
    assert 0.0 < bi["final_error"] < thr
    lam_gpu = gv.astype(np.float64) / (1.0 - dd)
    # the FP64 iterate after the GPU's own number of sweeps: equal up to FP32 rounding
    lam_k, _ = O.fp64_adjoint(n, ind, col, deg, c.astype(np.float64), w=wn, damping=dd, tol=0.0,
                              max_iters=bi["iterations"])
    assert np.linalg.norm(lam_gpu - lam_k) / np.linalg.norm(lam_k) < 1e-5
    # the exact solution: A is column-stochastic, so d A^T contracts by d in the max norm and
    # |lambda* - lambda_k|_inf <= d / (1 - d) |lambda_k - lambda_(k-1)|_inf < d / (1 - d) * thr
    lam = np.linalg.solve(np.eye(n) - dd * A.T, c.astype(np.float64))
    assert np.max(np.abs(lam_gpu - lam)) <= dd / (1.0 - dd) * thr * 1.01
    assert np.linalg.norm(lam_gpu - lam) / np.linalg.norm(lam) < 1e-4


def test_adjoint_l1_change_does_not_overflow_rmat20(ctx):
    """RMAT-20 (thousands of workgroups per sweep) with |grad_x| ~ 1e3: the first adjoint
    sweeps change lambda by ~1e9 in L1.  The error trace must equal the FP64 iteration's
    (first sweep: |d A^T g|_1) and the solve must stop within a sweep of the FP64 iteration
    with the same threshold, with lambda equal to its result; a bounded-range accumulator
    that saturates or wraps reports a wrong error and stops at the wrong sweep."""
    g = ctx.rmat(20, 16, seed=2)
    n = g.n
    g.random_weights(seed=7, normalize=True)
    ind, col, deg = g.csc()
    wn = g.get_weights().astype(np.float64)
    c = (1000.0 * O.normal_vector(n, seed=9)).astype(np.float32)
    thr = 1e5  # above the FP32 rounding floor of an L1 change over 1M entries of size ~1e3-1e5
    _, gv, bi = g.wpr_backward(np.full(n, 1.0 / n, np.float32), c, D, thr, 200, raw_weights=False)
    dd = float(np.float32(D))
    lam1, _ = O.fp64_adjoint(n, ind, col, deg, c.astype(np.float64), w=wn, damping=dd, tol=0.0, max_iters=1)
    err0 = np.abs(lam1 - c.astype(np.float64)).sum()
    assert err0 > 1e8
    assert abs(bi["error_trace"][0] - err0) <= 1e-5 * err0, (bi["error_trace"][0], err0)
    lam_it, it64 = O.fp64_adjoint(n, ind, col, deg, c.astype(np.float64), w=wn, damping=dd, tol=thr, max_iters=200)
    assert bi["iterations"] > 20 and abs(bi["iterations"] - it64) <= 1, (bi["iterations"], it64)
    lam_gpu = gv.astype(np.float64) / (1.0 - dd)
    assert np.linalg.norm(lam_gpu - lam_it) / np.linalg.norm(lam_it) < 1e-4


def test_relative_threshold_scales_with_the_gradient_norm(ctx):
    """NWG_INFO_RELATIVE_THRESHOLD: the adjoint stops at threshold * ||grad_x||_1 / (1 - d), the
    FP64 iteration with that absolute threshold stops within a sweep of it, scaling grad_x
    scales lambda and keeps the sweep count, and for the uniform teleport the forward is the
    same solve as with the absolute threshold."""
    g, n, ind, col, deg, rows, w_raw = setup_graph(ctx, 12, seed=7, wseed=4)
    g.set_weights(w_raw, normalize=True)
    _, wn, _ = dense_system(n, ind, col, rows, w_raw)
    dd = float(np.float32(D))
    v = np.full(n, 1.0 / n, np.float32)
    x_abs, fa = g.wpr_forward(None, D, 1e-4, 100)
    x_rel, fr = g.wpr_forward(None, D, 1e-4, 100, relative=True)
    assert fr["iterations"] == fa["iterations"] and np.array_equal(x_rel, x_abs)
    assert fr["threshold_used"] == float(np.float32(1e-4))  # the scale of the uniform teleport is 1
    c = O.normal_vector(n, seed=21).astype(np.float32)
    scale = np.abs(c.astype(np.float64)).sum() / (1.0 - dd)
    _, gv1, b1 = g.wpr_backward(x_abs, c, D, 1e-4, 200, raw_weights=False, relative=True)
    assert b1["threshold_used"] == pytest.approx(float(np.float32(1e-4)) * scale, rel=1e-9)
    assert 5 < b1["iterations"] < 200
    _, it64 = O.fp64_adjoint(n, ind, col, deg, c.astype(np.float64), w=wn, damping=dd,
                             tol=b1["threshold_used"], max_iters=200)
    assert abs(b1["iterations"] - it64) <= 1, (b1["iterations"], it64)
    _, gv1000, b1000 = g.wpr_backward(x_abs, 1000.0 * c, D, 1e-4, 200, raw_weights=False, relative=True)
    assert b1000["iterations"] == b1["iterations"]
    assert np.linalg.norm(gv1000 / 1000.0 - gv1) / np.linalg.norm(gv1) < 1e-5


def test_relative_tolerance_gradients_match_fp64_rmat16(ctx):
    """Config-5 form (L = sum c_i x_i, c ~ N(0,1)) at relative L1 tolerance 1e-4 for both solves
    on RMAT-16: edge and teleport gradients within 1e-4 relative of FP64 run for the same
    numbers of sweeps, and both sweep counts within one of the FP64 iterations' own stops."""
    g = ctx.rmat(16, 16, seed=5)
    n = g.n
    g.random_weights(seed=11, normalize=True)
    ind, col, deg = g.csc()
    wn = g.get_weights().astype(np.float64)
    rows = np.repeat(np.arange(n), np.diff(ind.astype(np.int64)))
    c = O.normal_vector(n, seed=31).astype(np.float32)
    x, fi = g.wpr_forward(None, D, 1e-4, 100, relative=True)
    gw, gv, bi = g.wpr_backward(x, c, D, 1e-4, 100, raw_weights=False, relative=True)
    kf, kb = fi["iterations"], bi["iterations"]
    assert 1 < kf < 100 and 1 < kb < 100
    dd = float(np.float32(D))
    x64, _ = O.fp64_jacobi(n, ind, col, deg, w=wn, v=np.full(n, 1.0 / n), damping=dd, tol=0.0, max_iters=kf)
    lam64, _ = O.fp64_adjoint(n, ind, col, deg, c.astype(np.float64), w=wn, damping=dd, tol=0.0, max_iters=kb)
    rel = lambda a, b: np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b)
    assert rel(gw, dd * lam64[rows] * x64[col]) < 1e-4
    assert rel(gv, (1.0 - dd) * lam64) < 1e-4
    _, itf = O.fp64_jacobi(n, ind, col, deg, w=wn, v=np.full(n, 1.0 / n), damping=dd, tol=fi["threshold_used"],
                           max_iters=100)
    _, itb = O.fp64_adjoint(n, ind, col, deg, c.astype(np.float64), w=wn, damping=dd, tol=bi["threshold_used"],
                            max_iters=100)
    assert abs(itf - kf) <= 1 and abs(itb - kb) <= 1, (kf, itf, kb, itb)


@pytest.mark.parametrize("scale,bin_bits", [(12, 0), (15, 13), (15, 14)])
def test_binned_weighted_sweep_matches_pull_and_fp64(ctx, scale, bin_bits):
    """The binned weighted sweep (weights carried per entry in the binned layout, signed fixed
    point scaled by a power of two from the solve's own bound) against the pull sweep on the
    same graph and against FP64 iterations run for the same sweeps: the forward, the adjoint
    of a loss with gradients of both signs at two scales (|grad_x| ~ 1 and ~ 1e3, which moves
    the scale by 2^10), and the edge and teleport gradients.  RMAT-15 has more entries per
    bin than one bin-pass work item takes, so split bins and their combine pass run too; it
    runs with bins of 2^13 and of 2^14 rows."""
    g, n, ind, col, deg, rows, w_raw = setup_graph(ctx, scale, seed=8, wseed=6)
    g.set_weights(w_raw, normalize=True)
    _, wn, _ = dense_system(n, ind, col, rows, w_raw) if scale <= 12 else (None, None, None)
    if wn is None:
        s = np.bincount(col, weights=w_raw.astype(np.float64), minlength=n)
        wn = w_raw.astype(np.float64) / s[col]
    dd = float(np.float32(D))
    rel = lambda a, b: np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b)
    rng = np.random.default_rng(scale)
    v = rng.uniform(size=n)
    v = (v / v.sum()).astype(np.float32)
    out = {}
    for binned in (0, 1):
        g.set_plan(binned=binned, bin_bits=bin_bits)
        x, fi = g.wpr_forward(v, D, 0.0, 60)  # a fixed number of sweeps on both paths
        res = [x, fi]
        for amp in (1.0, 1000.0):
            c = (amp * O.normal_vector(n, seed=40 + scale)).astype(np.float32)
            gw, gv, bi = g.wpr_backward(x, c, D, 1e-4, 200, raw_weights=False, relative=True)
            res += [gw, gv, bi]
        out[binned] = res
    xp, fp = out[0][0], out[0][1]
    xb, fb = out[1][0], out[1][1]
    assert fb["iterations"] == fp["iterations"] == 60
    assert np.max(np.abs(xb.astype(np.float64) - xp)) <= 1e-6 * np.abs(xp).max()
    x64, _ = O.fp64_jacobi(n, ind, col, deg, w=wn, v=v.astype(np.float64), damping=dd, tol=0.0,
                           max_iters=fb["iterations"])
    assert rel(xb, x64) < 1e-5
    for j, amp in enumerate((1.0, 1000.0)):
        gwp, gvp, bip = out[0][2 + 3 * j: 5 + 3 * j]
        gwb, gvb, bib = out[1][2 + 3 * j: 5 + 3 * j]
        assert abs(bib["iterations"] - bip["iterations"]) <= 1, (amp, bib["iterations"], bip["iterations"])
        c = (amp * O.normal_vector(n, seed=40 + scale)).astype(np.float64)
        lam64, _ = O.fp64_adjoint(n, ind, col, deg, c, w=wn, damping=dd, tol=0.0, max_iters=bib["iterations"])
        x64b, _ = O.fp64_jacobi(n, ind, col, deg, w=wn, v=v.astype(np.float64), damping=dd, tol=0.0,
                                max_iters=fb["iterations"])
        assert rel(gvb, (1.0 - dd) * lam64) < 1e-4, amp
        assert rel(gwb, dd * lam64[rows] * x64b[col]) < 1e-4, amp


def test_full_size_adjoint_identities_rmat20(ctx):
    """RMAT-20 with random raw weights, every solve run to its float fixed point (150 sweeps,
    0.85^150 < 1e-10).  Identities that hold for any size; on a dense FP64 system each one
    fails (relative error 3e-2 or more, against the 1e-4 bound) under the mistake named:
    - L = c.x is linear in the teleport, so c.x = (dL/dv).v (an adjoint solved with A instead
      of A^T);
    - sum_p wn[p] dL/dwn[p] = d lambda.(A x) = lambda.x - (dL/dv).v, with lambda = (dL/dv) / (1-d)
      (lambda and x swapped in the edge gradient);
    - L does not change when one source's raw weights are all scaled, so for every source u,
      sum over its out-edges of w~[p] dL/dw~[p] = 0 (the normaliser's mean term dropped)."""
    g = ctx.rmat(20, 16, seed=3)
    n = g.n
    ptr, col, _ = g.csc()
    rng = np.random.default_rng(20)
    w_raw = rng.uniform(0.1, 1.0, col.size).astype(np.float32)
    g.set_weights(w_raw, normalize=True)
    v = np.full(n, 1.0 / n, np.float32)
    c = O.normal_vector(n, seed=13).astype(np.float32)
    x, fi = g.wpr_forward(v, D, 0.0, 150)
    gw, gv, bi = g.wpr_backward(x, c, D, 0.0, 150, raw_weights=False)
    assert fi["iterations"] == 150 and bi["iterations"] == 150
    x64, c64, v64, gv64 = (a.astype(np.float64) for a in (x, c, v, gv))
    loss = c64 @ x64
    assert abs(loss - gv64 @ v64) <= 1e-4 * np.abs(c64 * x64).sum()
    s = np.bincount(col, weights=w_raw.astype(np.float64), minlength=n)
    wn = w_raw.astype(np.float64) / s[col]
    lam = gv64 / (1.0 - np.float64(np.float32(D)))
    lhs = wn @ gw.astype(np.float64)
    assert abs(lhs - (lam @ x64 - gv64 @ v64)) <= 1e-4 * np.abs(wn * gw).sum()
    graw, _, _ = g.wpr_backward(x, c, D, 0.0, 150, raw_weights=True)
    t = w_raw.astype(np.float64) * graw.astype(np.float64)
    per_source = np.bincount(col, weights=t, minlength=n)
    assert np.abs(per_source).sum() <= 1e-4 * np.abs(t).sum()


def test_autograd_function(ctx):
    torch = pytest.importorskip("torch")
    from nwgraph_amd.autograd import weighted_pagerank

    g, n, ind, col, deg, rows, w_raw = setup_graph(ctx, 11)
    w = torch.tensor(w_raw, device="cuda", requires_grad=True)
    v = torch.full((n,), 1.0 / n, device="cuda", requires_grad=True)
    c = torch.tensor(O.normal_vector(n, seed=3), dtype=torch.float32, device="cuda")
    x = weighted_pagerank(w, v, g, D, 1e-7, 300)
    loss = (c * x).sum()
    loss.backward()
    gw_ref, gv_ref, _ = g.wpr_backward(x.detach().cpu().numpy(), c.cpu().numpy(), D, 1e-7, 300, raw_weights=True)
    assert np.allclose(w.grad.cpu().numpy(), gw_ref, rtol=1e-5, atol=1e-9)
    assert np.allclose(v.grad.cpu().numpy(), gv_ref, rtol=1e-5, atol=1e-9)
    # forward equals the C ABI forward
    g.set_weights(w_raw, normalize=True)
    x_ref, _ = g.wpr_forward(np.full(n, 1.0 / n, np.float32), D, 1e-7, 300)
    assert np.allclose(x.detach().cpu().numpy(), x_ref, rtol=1e-6, atol=1e-10)
