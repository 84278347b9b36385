// Link-compatible replacement of OpenTimer's GPU entry points over libotg.
//
// OpenTimer's timer calls three functions of its CUDA sources with C++
// linkage and host pointers: rct_bfs_and_compute_cuda(RctEdgeArrayCUDA) from
// FlatRctStorage::_update_timing_cuda (ot/timer/net.cpp:274),
// rct_compute_cuda(RctCUDA) (declared ot/cuda/rct.cuh:44) and
// toposort_compute_cuda(..., std::vector<int>&) from
// Timer::_build_prop_tasks_cuda (ot/timer/timer.cpp:1069-1071).  This file
// defines them with those exact signatures, so the timer links against it in
// place of ot/cuda/*.cu, and forwards each call to the C ABI of libotg
// (otg_rct_bfs_and_compute, otg_rct_compute, otg_toposort_compute).
//
// Method: OTG_RCT_METHOD in the environment picks "plain" (the reference
// kernel's arithmetic), "batched" (FP32) or "compensated" (the default: FP32
// with compensated sums, within 1e-6 of FP64).  Errors print a message and
// abort, as the reference's checkCUDA/assert do (ot/cuda/utils.cuh:34-41):
// the signatures have no status to return.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "otg.h"
#include "ot_cuda_abi.hpp"

namespace {

otg_context* bridge_ctx() {
  // the reference uses the implicit CUDA device 0 (ot/cuda/utils.cuh)
  static std::once_flag once;
  static otg_context* ctx = nullptr;
  std::call_once(once, [] {
    if (otg_create(0, &ctx) != OTG_OK) {
      std::fprintf(stderr, "libotg bridge: otg_create failed: %s\n", otg_last_error(nullptr));
      std::abort();
    }
  });
  return ctx;
}

otg_rct_method bridge_method() {
  const char* m = std::getenv("OTG_RCT_METHOD");
  if (m && !std::strcmp(m, "plain")) return OTG_RCT_PLAIN;
  if (m && !std::strcmp(m, "batched")) return OTG_RCT_BATCHED;
  return OTG_RCT_BATCHED_COMPENSATED;
}

void bridge_check(otg_status s, const char* what) {
  if (s != OTG_OK) {
    std::fprintf(stderr, "libotg bridge: %s: %s: %s\n", what, otg_status_string(s),
                 otg_last_error(bridge_ctx()));
    std::abort();
  }
}

}  // namespace

void rct_bfs_and_compute_cuda(RctEdgeArrayCUDA d) {
  otg_rct_edge_array a{};
  a.num_nets = d.num_nets;
  a.total_num_nodes = d.total_num_nodes;
  a.total_num_edges = d.total_num_edges;
  a.rct_nodes_start = d.rct_nodes_start;
  // {s, t} pairs: the same bytes as int32_t[2E]; oriented in place (s = parent)
  a.rct_edges = reinterpret_cast<int32_t*>(d.rct_edges);
  a.rct_roots = d.rct_roots;
  a.rct_edges_res = d.rct_edges_res;
  a.rct_nodes_cap = d.rct_nodes_cap;
  a.rct_node2bfs_order = d.rct_node2bfs_order;
  a.rct_pid = d.rct_pid;
  a.load = d.load;
  a.delay = d.delay;
  a.ldelay = d.ldelay;
  a.impulse = d.impulse;
  // beta, slew and slew_in stay NULL: the timer forms slews per query
  // (FlatRct::slew, ot/timer/net.cpp:280-283)
  bridge_check(otg_rct_bfs_and_compute(bridge_ctx(), &a, bridge_method()),
               "rct_bfs_and_compute_cuda");
}

void rct_compute_cuda(RctCUDA d) {
  otg_rct_host h{};
  h.num_nets = d.num_nets;
  h.total_num_nodes = d.total_num_nodes;
  h.rct_nodes_start = d.rct_nodes_start;
  h.rct_pid = d.rct_pid;
  h.pres = d.pres;
  h.cap = d.cap;
  h.load = d.load;
  h.delay = d.delay;
  h.ldelay = d.ldelay;
  h.impulse = d.impulse;
  bridge_check(otg_rct_compute(bridge_ctx(), &h, bridge_method()), "rct_compute_cuda");
}

void toposort_compute_cuda(int n, int num_edges, int first_size, int* edgelist_start,
                           int* edgelist, int* out, int* frontiers,
                           std::vector<int>& frontiers_ends) {
  // The caller has already pushed 0 and first_size (timer.cpp:1058-1059);
  // like the reference (ot/cuda/toposort.cu:44-62), one end is appended per
  // further non-empty frontier.
  std::vector<int32_t> ends((size_t)n + 2, 0);
  int32_t num_ends = 0;
  bridge_check(otg_toposort_compute(bridge_ctx(), n, num_edges, first_size, edgelist_start,
                                    edgelist, out, frontiers, ends.data(), &num_ends, nullptr),
               "toposort_compute_cuda");
  for (int32_t k = 2; k < num_ends; ++k) frontiers_ends.push_back(ends[k]);
}
