// The types and functions through which OpenTimer's timer calls its GPU
// code, declared as its callers see them: ot/cuda/rct.cuh:7-47 (RctCUDA,
// RctEdgeCUDA, RctEdgeArrayCUDA, rct_compute_cuda, rct_bfs_and_compute_cuda)
// and ot/cuda/toposort.cuh:1-6 (toposort_compute_cuda).  In OpenTimer's own
// tree the bridge includes those two headers instead; this file exists so
// that the bridge builds, and is tested, without the reference sources.  The
// layout checks below pin the structs to the declaration order of rct.cuh on
// LP64 (int 4 bytes, pointers 8), so that the by-value argument is the same
// object the reference's callers build.
#pragma once

#include <cstddef>
#include <vector>

struct RctCUDA {
  int num_nets;
  int total_num_nodes;
  int* rct_nodes_start;
  int* rct_pid;
  float *pres, *cap, *load, *delay, *ldelay, *impulse;
};

struct RctEdgeCUDA {
  int s;
  int t;
};

struct RctEdgeArrayCUDA {
  int num_nets;
  int total_num_nodes;
  int total_num_edges;
  int* rct_nodes_start = nullptr;
  RctEdgeCUDA* rct_edges = nullptr;
  int* rct_roots = nullptr;
  float* rct_edges_res = nullptr;
  float* rct_nodes_cap = nullptr;
  int* rct_distances = nullptr;    // scratch of the reference's kernels, unused here
  int* rct_sort_counts = nullptr;  // scratch of the reference's kernels, unused here
  float* pres = nullptr;           // scratch of the reference's kernels, unused here
  float* cap = nullptr;            // scratch of the reference's kernels, unused here
  int* rct_node2bfs_order = nullptr;
  int* rct_pid = nullptr;
  float* load = nullptr;
  float* delay = nullptr;
  float* ldelay = nullptr;
  float* impulse = nullptr;
};

static_assert(sizeof(RctCUDA) == 72 && offsetof(RctCUDA, rct_nodes_start) == 8 &&
                  offsetof(RctCUDA, impulse) == 64,
              "RctCUDA layout of ot/cuda/rct.cuh:7-13");
static_assert(sizeof(RctEdgeCUDA) == 8, "RctEdgeCUDA layout of ot/cuda/rct.cuh:15-18");
static_assert(sizeof(RctEdgeArrayCUDA) == 136 && offsetof(RctEdgeArrayCUDA, rct_nodes_start) == 16 &&
                  offsetof(RctEdgeArrayCUDA, rct_nodes_cap) == 48 &&
                  offsetof(RctEdgeArrayCUDA, rct_node2bfs_order) == 88 &&
                  offsetof(RctEdgeArrayCUDA, impulse) == 128,
              "RctEdgeArrayCUDA layout of ot/cuda/rct.cuh:20-42");

// C++ linkage, as the reference declares them (rct.cuh:44-45, toposort.cuh:3-6)
void rct_compute_cuda(RctCUDA data_cpu);
void rct_bfs_and_compute_cuda(RctEdgeArrayCUDA data_cpu);
void toposort_compute_cuda(int n, int num_edges, int first_size, int* edgelist_start,
                           int* edgelist, int* out, int* frontiers,
                           std::vector<int>& frontiers_ends);
