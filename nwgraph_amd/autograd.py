"""torch.autograd.Function over the weighted PageRank forward / implicit-adjoint backward
pair of the C ABI (nwg_wpr_forward / nwg_wpr_backward).

    x = WeightedPageRank.apply(w_raw, teleport, graph, damping, threshold, max_iters)

w_raw (float32[nnz], CSC order, on the GPU) are raw edge weights, normalised per source
inside the forward; teleport (float32[n]) sums to 1.  backward(dL/dx) solves the
adjoint fixed point lambda = dL/dx + d A^T lambda on the device (no unrolled iterates
are kept) and returns dL/dw_raw (chained through the normaliser) and
dL/dteleport = (1-d) lambda.  With relative=True both solves stop on a relative L1
tolerance (NWG_INFO_RELATIVE_THRESHOLD): the forward at threshold * ||teleport||_1, the
adjoint at threshold * ||dL/dx||_1 / (1-d).  Torch only provides device memory and the
stream: both passes run in libnwgraph_hip on torch's current stream.
"""
import torch


class WeightedPageRank(torch.autograd.Function):
    @staticmethod
    def forward(ctx, w_raw, teleport, graph, damping=0.85, threshold=1e-4, max_iters=100, adjoint_threshold=None,
                relative=False):
        if not (w_raw.is_cuda and teleport.is_cuda):
            raise ValueError("w_raw and teleport must be GPU tensors")
        w_raw = w_raw.detach().contiguous().float()
        teleport = teleport.detach().contiguous().float()
        if w_raw.numel() != graph.nnz or teleport.numel() != graph.n:
            raise ValueError("w_raw must have nnz entries and teleport n entries")
        graph.ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        graph.set_weights_device(w_raw.data_ptr(), normalize=True)
        x = torch.empty(graph.n, dtype=torch.float32, device=w_raw.device)
        info = graph.wpr_forward_device(teleport.data_ptr(), x.data_ptr(), damping, threshold, max_iters,
                                        relative=relative)
        ctx.graph = graph
        ctx.damping = damping
        ctx.adj_threshold = threshold if adjoint_threshold is None else adjoint_threshold
        ctx.max_iters = max_iters
        ctx.relative = relative
        ctx.forward_info = info
        ctx.save_for_backward(x)
        return x

    @staticmethod
    def backward(ctx, grad_x):
        (x,) = ctx.saved_tensors
        g = ctx.graph
        grad_x = grad_x.contiguous().float()
        gw = torch.empty(g.nnz, dtype=torch.float32, device=x.device) if ctx.needs_input_grad[0] else None
        gv = torch.empty(g.n, dtype=torch.float32, device=x.device) if ctx.needs_input_grad[1] else None
        g.ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        g.wpr_backward_device(x.data_ptr(), grad_x.data_ptr(), gw.data_ptr() if gw is not None else None,
                              gv.data_ptr() if gv is not None else None, ctx.damping, ctx.adj_threshold,
                              ctx.max_iters, raw_weights=True, relative=ctx.relative)
        return gw, gv, None, None, None, None, None, None


def weighted_pagerank(w_raw, teleport, graph, damping=0.85, threshold=1e-4, max_iters=100, adjoint_threshold=None,
                      relative=False):
    return WeightedPageRank.apply(w_raw, teleport, graph, damping, threshold, max_iters, adjoint_threshold, relative)
