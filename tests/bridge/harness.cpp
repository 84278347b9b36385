// Test harness for the OpenTimer bridge: calls the bridge's C++ entry points
// the way OpenTimer's timer does, so that a ctypes test can drive them.
//  * rct: a default-constructed RctEdgeArrayCUDA whose fields are assigned
//    from the caller's arrays and which is passed by value
//    (FlatRctStorage::_update_timing_cuda, ot/timer/net.cpp:254-274);
//  * toposort: a std::vector<int> seeded with 0 and first_size before the
//    call (Timer::_build_prop_tasks_cuda, ot/timer/timer.cpp:1058-1071).
// Test infrastructure only; the product is ot_cuda_bridge.cpp.
#include <cstring>
#include <vector>

#include "ot_cuda_abi.hpp"

extern "C" void harness_rct_bfs_and_compute(int num_nets, int total_num_nodes,
                                            int total_num_edges, int* nodes_start, int* edges,
                                            int* roots, float* edges_res, float* nodes_cap,
                                            int* node2bfs, int* pid, float* load, float* delay,
                                            float* ldelay, float* impulse) {
  RctEdgeArrayCUDA rct_cuda;
  rct_cuda.num_nets = num_nets;
  rct_cuda.total_num_nodes = total_num_nodes;
  rct_cuda.total_num_edges = total_num_edges;
  rct_cuda.rct_edges = reinterpret_cast<RctEdgeCUDA*>(edges);
  rct_cuda.rct_roots = roots;
  rct_cuda.rct_nodes_start = nodes_start;
  rct_cuda.rct_node2bfs_order = node2bfs;
  rct_cuda.rct_edges_res = edges_res;
  rct_cuda.rct_nodes_cap = nodes_cap;
  rct_cuda.rct_pid = pid;
  rct_cuda.load = load;
  rct_cuda.delay = delay;
  rct_cuda.ldelay = ldelay;
  rct_cuda.impulse = impulse;
  rct_bfs_and_compute_cuda(rct_cuda);
}

extern "C" void harness_rct_compute(int num_nets, int total_num_nodes, int* nodes_start,
                                    int* pid, float* pres, float* cap, float* load,
                                    float* delay, float* ldelay, float* impulse) {
  RctCUDA d;
  d.num_nets = num_nets;
  d.total_num_nodes = total_num_nodes;
  d.rct_nodes_start = nodes_start;
  d.rct_pid = pid;
  d.pres = pres;
  d.cap = cap;
  d.load = load;
  d.delay = delay;
  d.ldelay = ldelay;
  d.impulse = impulse;
  rct_compute_cuda(d);
}

// Returns the number of entries of frontiers_ends after the call (written to
// ends_out, which has room for n + 2).
extern "C" int harness_toposort(int n, int num_edges, int first_size, int* edgelist_start,
                                int* edgelist, int* out, int* frontiers, int* ends_out) {
  std::vector<int> frontiers_ends;
  frontiers_ends.push_back(0);
  frontiers_ends.push_back(first_size);
  toposort_compute_cuda(n, num_edges, first_size, edgelist_start, edgelist, out, frontiers,
                        frontiers_ends);
  std::memcpy(ends_out, frontiers_ends.data(), sizeof(int) * frontiers_ends.size());
  return (int)frontiers_ends.size();
}
