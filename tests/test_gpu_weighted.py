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
    assert np.linalg.norm(got - fd) / np.linalg.norm(fd) < 1e-3


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
