"""The allocation-free one-call solve (nwg_pagerank_f32_noalloc) through the C ABI: two
calls in a row (the second on a smaller graph, so stale scratch contents from the first
would show) against nwg_pagerank_f32 on the same CSC and against the oracle's v10, plus
the limits it rejects before launching anything."""
import ctypes

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu

NWG_EINVAL = -22


def _dev_u32(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, np.uint32).view(np.int32)).cuda()


def _solve_both(ctx, torch, n, ind, col, deg, thr, iters, give_deg):
    from nwgraph_amd import _lib

    lib = ctx.lib
    # the existing one-call entry point on host arrays
    x_ref = np.empty(n, np.float32)
    tr_ref = np.zeros(max(iters, 1))
    info_ref = _lib.PrInfo()
    info_ref.error_trace = tr_ref.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    info_ref.error_trace_len = tr_ref.size
    ind32, col32 = np.ascontiguousarray(ind, np.uint32), np.ascontiguousarray(col, np.uint32)
    deg32 = np.ascontiguousarray(deg, np.uint32)
    ctx.check(lib.nwg_pagerank_f32(ctx.handle, n, col32.size, ctypes.c_void_p(ind32.ctypes.data),
                                   ctypes.c_void_p(col32.ctypes.data),
                                   ctypes.c_void_p(deg32.ctypes.data) if give_deg else None,
                                   ctypes.c_void_p(x_ref.ctypes.data), 0.85, thr, iters, ctypes.byref(info_ref)))
    # the allocation-free variant on device arrays
    d_ind, d_col, d_deg = _dev_u32(torch, ind32), _dev_u32(torch, col32), _dev_u32(torch, deg32)
    d_x = torch.full((n,), -1.0, dtype=torch.float32, device="cuda")
    tr = np.zeros(max(iters, 1))
    info = _lib.PrInfo()
    info.error_trace = tr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    info.error_trace_len = tr.size
    torch.cuda.synchronize()
    ctx.check(lib.nwg_pagerank_f32_noalloc(ctx.handle, n, col32.size, ctypes.c_void_p(d_ind.data_ptr()),
                                           ctypes.c_void_p(d_col.data_ptr()),
                                           ctypes.c_void_p(d_deg.data_ptr()) if give_deg else None,
                                           ctypes.c_void_p(d_x.data_ptr()), 0.85, thr, iters, ctypes.byref(info)))
    x = d_x.cpu().numpy()
    return x_ref, info_ref, tr_ref, x, info, tr


def test_noalloc_matches_the_one_call_solve_twice(ctx):
    torch = pytest.importorskip("torch")
    # first call: RMAT-16 (hub rows far above 32 in-edges take the wave path), degrees counted
    # on the device; second call: RMAT-12 to convergence, degrees given, with a threshold
    # halfway (geometric mean) between the oracle's L1 changes of sweeps 40 and 41, so that
    # the stop at sweep 41 is far from FP32 summation noise
    n12, ind12, col12, deg12 = O.rmat_csc(12, 16, seed=3)
    _, _, tr12, _ = O.pr_v10(n12, ind12, col12, deg12, 0.85, 0.0, 60)
    thr12 = float(np.float32(np.sqrt(tr12[39] * tr12[40])))
    cases = [(16, 1e-4, 20, False), (12, thr12, 200, True)]
    for scale, thr, iters, give_deg in cases:
        n, ind, col, deg = O.rmat_csc(scale, 16, seed=3)
        print(f"noalloc: RMAT-{scale} n={n} nnz={col.size} thr={thr} max_iters={iters} given_deg={give_deg}",
              flush=True)
        assert np.diff(ind.astype(np.int64)).max() > 32
        x_ref, info_ref, tr_ref, x, info, tr = _solve_both(ctx, torch, n, ind, col, deg, thr, iters, give_deg)
        it = int(info.iterations)
        assert it == int(info_ref.iterations), (it, int(info_ref.iterations))
        assert np.max(np.abs(x.astype(np.float64) - x_ref)) <= 1e-6
        # the two sweeps add each row's in-edges in different FP32 orders: every row's |x' - x|
        # may differ by a few ulp of its score, so the L1 changes agree to n * 4 ulp(max x)
        floor = n * 4 * float(np.spacing(np.float32(x_ref.max())))
        np.testing.assert_allclose(tr[:it], tr_ref[:it], rtol=1e-3, atol=floor)
        assert abs(info.final_error - info_ref.final_error) <= 1e-3 * info_ref.final_error + floor
        assert info.ms_total > 0.0
        ref, it_o, _, _ = O.pr_v10(n, ind, col, deg, 0.85, thr, iters)
        assert it == it_o
        assert np.max(np.abs(x.astype(np.float64) - ref)) <= 1e-6
        if scale == 12:
            assert it == 41  # the second call stopped on the threshold


def test_noalloc_rejects_what_exceeds_its_scratch(ctx):
    torch = pytest.importorskip("torch")
    from nwgraph_amd import _lib

    lib = ctx.lib
    n, ind, col, deg = O.rmat_csc(8, 16, seed=1)
    d_ind, d_col = _dev_u32(torch, ind), _dev_u32(torch, col)
    d_x = torch.empty(n, dtype=torch.float32, device="cuda")
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    # more vertices than the scratch holds (checked on the host before any launch)
    rc = lib.nwg_pagerank_f32_noalloc(ctx.handle, (1 << 22) + 1, col.size, p(d_ind), p(d_col), None, p(d_x), 0.85,
                                      1e-4, 20, None)
    assert rc == NWG_EINVAL and b"4194304" in lib.nwg_last_error(ctx.handle)
    rc = lib.nwg_pagerank_f32_noalloc(ctx.handle, n, col.size, p(d_ind), p(d_col), None, p(d_x), 0.85, 1e-4, 4097,
                                      None)
    assert rc == NWG_EINVAL and b"4096" in lib.nwg_last_error(ctx.handle)
    info = _lib.PrInfo()
    info.flags = 1  # NWG_INFO_KERNEL_TIMING
    rc = lib.nwg_pagerank_f32_noalloc(ctx.handle, n, col.size, p(d_ind), p(d_col), None, p(d_x), 0.85, 1e-4, 20,
                                      ctypes.byref(info))
    assert rc == NWG_EINVAL
    rc = lib.nwg_pagerank_f32_noalloc(ctx.handle, n, col.size, None, p(d_col), None, p(d_x), 0.85, 1e-4, 20, None)
    assert rc == NWG_EINVAL
    # a CSC whose entry is out of range is rejected by the device check (a flag in scratch)
    bad = col.copy()
    bad[5] = n
    d_bad = _dev_u32(torch, bad)
    rc = lib.nwg_pagerank_f32_noalloc(ctx.handle, n, col.size, p(d_ind), p(d_bad), None, p(d_x), 0.85, 1e-4, 20,
                                      None)
    assert rc == NWG_EINVAL
    # and a valid call right after still works
    x_ref, info_ref, _, x, info, _ = _solve_both(ctx, torch, n, ind, col, deg, 1e-4, 20, False)
    assert int(info.iterations) == int(info_ref.iterations)
    assert np.max(np.abs(x.astype(np.float64) - x_ref)) <= 1e-6
