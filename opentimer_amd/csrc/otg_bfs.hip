// Per-net BFS ordering of RC trees plus reordering of R and C into BFS order.
// Replaces compute_net_bfs (reference ot/cuda/rct.cu:120-245), its argsort
// block_couting_sort (ot/cuda/sort.cuh:63-142) and prepare_res_cap
// (rct.cu:314-359).
//
// One wavefront per net.  The net's undirected edges are turned into an
// adjacency list in LDS (small nets) or in global scratch (large nets); the
// BFS runs frontier by frontier with ballot compaction; the ordering is a
// stable counting sort by distance in which node ids are visited from the
// largest down, so that the position of node u is
//   #(nodes with smaller distance) + #(nodes with equal distance, larger id)
// -- the order the reference's bucket-end counting sort produces.  Work is
// O(nodes) per net (the reference rescans every edge once per BFS level).
#include <hip/hip_runtime.h>

#include "otg_internal.hpp"

namespace {

constexpr int kWave = 64;
constexpr int kLdsNodes = 512;  // nets up to this size keep their scratch in LDS

struct BfsParams {
  int32_t num_nets;
  const int32_t* nodes_start;
  const int32_t* edges_in;
  int32_t* edges_out;
  const int32_t* roots;
  const float* edge_res;
  const float4* node_cap;
  int32_t* node2bfs;
  int32_t* pid;
  float* pres;
  float4* cap_bfs;
  int32_t* gscratch;  // 6 ints per node + 2 per net
  int32_t* bad;       // smallest failing net id
};

// Scratch layout per net with n nodes (ints): off[n+1] | adj[2n] | dist[n] |
// frontier[n] | count[n+1]
template <bool GLOBAL>
__device__ void bfs_net(const BfsParams& P, int net, int32_t* base, int lane) {
  const int off = P.nodes_start[net];
  const int n = P.nodes_start[net + 1] - off;
  if (n <= 0) return;
  const int ne = n - 1;
  const int32_t* e = P.edges_in + 2LL * (off - net);
  int32_t* adjoff = base;            // n+1
  int32_t* adj = adjoff + (n + 1);   // 2n
  int32_t* dist = adj + 2 * n;       // n
  int32_t* front = dist + n;         // n
  int32_t* cnt = front + n;          // n+1
  auto sync = [&]() {
    if constexpr (GLOBAL) __threadfence();
    __syncthreads();
  };
  int ok = 1;
  const int root = P.roots[net];
  if (root < 0 || root >= n) ok = 0;
  // degrees
  for (int k = lane; k <= n; k += kWave) adjoff[k] = 0;
  for (int k = lane; k < n; k += kWave) {
    dist[k] = -1;
    cnt[k] = 0;
  }
  sync();
  for (int k = lane; k < ne; k += kWave) {
    const int s = e[2 * k], t = e[2 * k + 1];
    if (s < 0 || s >= n || t < 0 || t >= n || s == t) {
      ok = 0;
    } else {
      atomicAdd(&adjoff[s + 1], 1);
      atomicAdd(&adjoff[t + 1], 1);
    }
  }
  sync();
  ok = __all(ok);
  if (!ok) {
    if (lane == 0) atomicMin(P.bad, net);
    return;
  }
  // exclusive offsets: wave-wide inclusive scan in chunks of 64
  {
    int carry = 0;
    for (int b = 0; b <= n; b += kWave) {
      const int k = b + lane;
      int v = k <= n ? adjoff[k] : 0;
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        const int t = __shfl_up(v, d);
        if (lane >= d) v += t;
      }
      if (k <= n) adjoff[k] = v + carry;
      carry += __shfl(v, kWave - 1);
    }
  }
  sync();
  // fill adjacency (cnt[] as per-node fill cursor)
  for (int k = lane; k < ne; k += kWave) {
    const int s = e[2 * k], t = e[2 * k + 1];
    adj[adjoff[s] + atomicAdd(&cnt[s], 1)] = t;
    adj[adjoff[t] + atomicAdd(&cnt[t], 1)] = s;
  }
  sync();
  // BFS frontier by frontier
  if (lane == 0) {
    dist[root] = 0;
    front[0] = root;
  }
  sync();
  int fb = 0, fe = 1, level = 0, reached = 1;
  while (fb < fe) {
    int tail = fe;
    for (int b = fb; b < fe; b += kWave) {
      const int k = b + lane;
      const int u = k < fe ? front[k] : -1;
      const int jb = u >= 0 ? adjoff[u] : 0, je = u >= 0 ? adjoff[u + 1] : 0;
      for (int j = jb;; ++j) {
        const bool has = j < je;
        if (!__any(has)) break;
        int v = -1;
        bool take = false;
        if (has) {
          v = adj[j];
          // a tree reaches v from exactly one neighbour; the CAS keeps a
          // malformed input (cycle) from queueing a node twice
          if (dist[v] < 0 && atomicCAS(&dist[v], -1, level + 1) == -1) take = true;
        }
        const uint64_t m = __ballot(take);
        if (take) front[tail + __popcll(m & ((1ull << lane) - 1ull))] = v;
        tail += __popcll(m);
      }
    }
    sync();
    fb = fe;
    fe = tail;
    reached = tail;
    ++level;
  }
  if (reached != n) {
    if (lane == 0) atomicMin(P.bad, net);
    return;
  }
  // level sizes -> level start positions (cnt[d])
  for (int k = lane; k <= n; k += kWave) cnt[k] = 0;
  sync();
  for (int k = lane; k < n; k += kWave) atomicAdd(&cnt[dist[k]], 1);
  sync();
  {
    int carry = 0;
    for (int b = 0; b < level; b += kWave) {
      const int k = b + lane;
      int v = k < level ? cnt[k] : 0;
      const int own = v;
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        const int t = __shfl_up(v, d);
        if (lane >= d) v += t;
      }
      if (k < level) cnt[k] = v - own + carry;  // exclusive start of level k
      carry += __shfl(v, kWave - 1);
    }
  }
  sync();
  // stable ranking: ids from the largest down, same-distance ids in
  // descending order take consecutive positions
  for (int top = n - 1; top >= 0; top -= kWave) {
    const int u = top - lane;  // lane 0 holds the largest id of the chunk
    const bool valid = u >= 0;
    const int d = valid ? dist[u] : -1;
    uint64_t remaining = __ballot(valid);
    while (remaining) {
      const int leader = __ffsll((unsigned long long)remaining) - 1;
      const int key = __shfl(d, leader);
      const uint64_t m = __ballot(valid && d == key);
      if (valid && d == key) {
        const int pos = cnt[key] + __popcll(m & ((1ull << lane) - 1ull));
        P.node2bfs[off + u] = pos;
        dist[u] = pos;  // reuse: dist now holds the position
      }
      sync();
      if (lane == leader) cnt[key] += __popcll(m);
      sync();
      remaining &= ~m;
    }
  }
  sync();
  // orient edges, write pid / pres / cap in BFS order
  if (lane == 0) {
    P.pid[off] = 0;
    P.pres[off] = 0.0f;
  }
  int32_t* eo = P.edges_out + 2LL * (off - net);
  const float* res = P.edge_res + (off - net);
  for (int k = lane; k < ne; k += kWave) {
    int s = e[2 * k], t = e[2 * k + 1];
    int ps = dist[s], pt = dist[t];
    if (ps > pt) {
      const int x = s; s = t; t = x;
      const int y = ps; ps = pt; pt = y;
    }
    eo[2 * k] = s;
    eo[2 * k + 1] = t;
    P.pid[off + pt] = pt - ps;
    P.pres[off + pt] = res[k];
  }
  for (int k = lane; k < n; k += kWave) P.cap_bfs[(int64_t)off + dist[k]] = P.node_cap[(int64_t)off + k];
}

__global__ __launch_bounds__(64) void rct_bfs_kernel(BfsParams P) {
  __shared__ int32_t lds[6 * kLdsNodes + 2];
  const int lane = threadIdx.x;
  for (int net = blockIdx.x; net < P.num_nets; net += gridDim.x) {
    const int n = P.nodes_start[net + 1] - P.nodes_start[net];
    if (n <= kLdsNodes) {
      bfs_net<false>(P, net, lds, lane);
    } else {
      bfs_net<true>(P, net, P.gscratch + 6LL * P.nodes_start[net] + 2LL * net, lane);
    }
    __syncthreads();
  }
}

}  // namespace

otg_status otg_rct_bfs_launch(otg_context* ctx, const otg_rct_edges* in, int32_t* node2bfs,
                              int32_t* pid, int32_t* edges_directed, float* pres, float* cap_bfs,
                              int32_t* d_bad, hipStream_t s) {
  BfsParams P{};
  P.num_nets = in->num_nets;
  P.nodes_start = in->nodes_start;
  P.edges_in = in->edges;
  P.edges_out = edges_directed;
  P.roots = in->roots;
  P.edge_res = in->edge_res;
  P.node_cap = reinterpret_cast<const float4*>(in->node_cap);
  P.node2bfs = node2bfs;
  P.pid = pid;
  P.pres = pres;
  P.cap_bfs = reinterpret_cast<float4*>(cap_bfs);
  P.bad = d_bad;
  int32_t* gs = nullptr;
  OTG_HIP(ctx, hipMallocAsync((void**)&gs, sizeof(int32_t) * (6LL * in->total_num_nodes + 2LL * in->num_nets + 8), s));
  P.gscratch = gs;
  const int grid = std::max(1, std::min(in->num_nets, ctx->num_cus * 16));
  hipLaunchKernelGGL(rct_bfs_kernel, dim3(grid), dim3(64), 0, s, P);
  hipError_t le = hipGetLastError();
  OTG_HIP(ctx, hipFreeAsync(gs, s));
  OTG_HIP(ctx, le);
  return OTG_OK;
}
