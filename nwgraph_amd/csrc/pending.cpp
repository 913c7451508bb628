// pending.cpp -- entry points whose device implementation is not built yet; they
// report NWG_ENOTSUP with a message instead of computing anything.
#include "nwg_internal.hpp"

namespace {
int not_yet(nwg_ctx* ctx, const char* what) {
  if (ctx) ctx->last_error = std::string(what) + " is not implemented in this build";
  return NWG_ENOTSUP;
}
}  // namespace

extern "C" {

int nwg_graph_autotune(nwg_ctx* ctx, nwg_graph*, nwg_plan_params*) { return not_yet(ctx, "nwg_graph_autotune"); }

int nwg_ppr_batched(nwg_ctx* ctx, nwg_graph*, uint32_t, const uint32_t*, const uint32_t*, float*, float, float,
                    uint64_t, int, uint64_t*, nwg_pr_info*) {
  return not_yet(ctx, "nwg_ppr_batched");
}

int nwg_multigraph_create(nwg_ctx* ctx, uint32_t, nwg_graph* const*, nwg_multigraph** out) {
  if (out) *out = nullptr;
  return not_yet(ctx, "nwg_multigraph_create");
}
void nwg_multigraph_destroy(nwg_multigraph*) {}
int nwg_multigraph_info(const nwg_multigraph*, uint32_t*, uint64_t*, uint64_t*) { return NWG_ENOTSUP; }
int nwg_multigraph_pagerank(nwg_ctx* ctx, nwg_multigraph*, float*, float, float, uint64_t, int, uint64_t*, double*,
                            nwg_pr_info*) {
  return not_yet(ctx, "nwg_multigraph_pagerank");
}

int nwg_graph_set_weights(nwg_ctx* ctx, nwg_graph*, const float*, int, int) {
  return not_yet(ctx, "nwg_graph_set_weights");
}
int nwg_graph_random_weights(nwg_ctx* ctx, nwg_graph*, uint64_t, int) { return not_yet(ctx, "nwg_graph_random_weights"); }
int nwg_graph_get_weights(nwg_ctx* ctx, const nwg_graph*, float*) { return not_yet(ctx, "nwg_graph_get_weights"); }
int nwg_wpr_forward(nwg_ctx* ctx, nwg_graph*, const float*, float*, float, float, uint64_t, int, nwg_pr_info*) {
  return not_yet(ctx, "nwg_wpr_forward");
}
int nwg_wpr_backward(nwg_ctx* ctx, nwg_graph*, const float*, const float*, float*, float*, float, float, uint64_t, int,
                     int, nwg_pr_info*) {
  return not_yet(ctx, "nwg_wpr_backward");
}

}  // extern "C"
