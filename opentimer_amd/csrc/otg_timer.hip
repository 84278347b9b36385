// A full timing update on device-resident arrays: otg_timer (include/otg_timer.h).
//
// Replaces the GPU flow of Timer::update_timing (reference
// ot/timer/timer.cpp:1239-1334 with cuda(true)) by one call over the
// library's own stages:
//   RC phase      otg_rct_bfs (first update), otg_rct_solve         timer.cpp:925-979
//   net values    RC delay / impulse at each net arc's sink, cell loads
//                 (Net::_delay / _slew / _load, net.cpp:626-699), gathered
//                 on the device
//   levels        fanin CSR without loop breakers, fanout counts, the
//                 first frontier, otg_levelize (first update)     timer.cpp:988-1071
//   propagation   otg_fprop_run over a level plan, then otg_bprop  timer.cpp:1079-1103
// The first update builds everything that depends on the topology only
// (BFS order, RC plan, levels, propagation plan, fanout CSR); later updates
// reorder R and C into the kept BFS order, solve and propagate.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "otg_internal.hpp"
#include "../../include/otg_prop.h"
#include "../../include/otg_timer.h"

struct otg_timer {
  otg_context* ctx = nullptr;
  otg_timer_design d{};
  otg_prop_lib lib{};
  otg_rct_method method = OTG_RCT_BATCHED_COMPENSATED;
  bool built = false;
  // RC forest in BFS order
  int32_t* node2bfs = nullptr;  // [V]
  int32_t* pid = nullptr;       // [V]
  int32_t* edges_dir = nullptr; // [2E] oriented: (parent, child), original ids (build only)
  int32_t* src_pos = nullptr;   // [V] node (global original index) at each BFS position
  int32_t* res_edge = nullptr;  // [V] edge whose child sits at each BFS position; -1 at a root
  float* pres = nullptr;        // [V]
  float* cap_bfs = nullptr;     // [4V]
  float* rc_load = nullptr;     // [4V]
  float* rc_delay = nullptr;    // [4V]
  float* rc_impulse = nullptr;  // [4V]
  otg_rct_plan* rc_plan = nullptr;
  // gathered net values
  float* arc_rc_delay = nullptr;    // [4m]
  float* arc_rc_impulse = nullptr;  // [4m]
  float* cell_load = nullptr;       // [4n]
  // levels
  int32_t* lev_start = nullptr;   // [n+1] fanin CSR without loop breakers
  int32_t* lev_fanin = nullptr;   // [m]
  int32_t* out_degree = nullptr;  // [n]
  int32_t* first = nullptr;       // [n] pins without fanout
  int32_t* frontiers = nullptr;   // [n]
  int32_t* frontier_ends = nullptr;  // [n+2]
  int32_t num_levels = 0;
  otg_fprop_plan* fplan = nullptr;
  // fanout CSR for the backward pass
  int32_t* fanout_start = nullptr;  // [n+1]
  int32_t* fanout_arc = nullptr;    // [m]
  int32_t* fanout_to = nullptr;     // [m] sink of every fanout entry
  int32_t* arc_to = nullptr;        // [m]
  // results
  float* slew = nullptr;       // [4n]
  float* at = nullptr;         // [4n]
  float* rat = nullptr;        // [4n]
  float* slack = nullptr;      // [4n]
  float* arc_delay = nullptr;  // [8m]
};

namespace {

constexpr int kBlock = 256;

int blocks_for(int64_t n) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, 1 << 16));
}

__global__ void timer_fill_kernel(float* a, int64_t n, float v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = v;
}

__global__ void timer_fill_i32_kernel(int32_t* a, int64_t n, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = v;
}

// Per pin p: its fanin arcs that propagate (loop breakers left out, as the
// reference's levelization CSR, timer.cpp:1011-1016, 1031-1036): their
// count (lev_count[p], exclusive-scanned into the CSR offsets), each arc's
// sink, each source pin's count of propagating fanout arcs (out_degree) and
// every arc's key for the fanout sort (its source; n for a loop breaker).
__global__ void timer_arc_count_kernel(int32_t n, int32_t m,
                                       const int32_t* __restrict__ fanin_start,
                                       const int32_t* __restrict__ fanin,
                                       const int32_t* __restrict__ kind, int32_t* __restrict__ lev_count,
                                       int32_t* __restrict__ out_degree, int32_t* __restrict__ arc_to,
                                       int32_t* __restrict__ sort_key, int32_t* __restrict__ bad) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    const int b = fanin_start[p], e = fanin_start[p + 1];
    // a CSR that does not run from 0 to m without decreasing reports the pin
    // (bad[1]) and its range is not walked
    if (b < 0 || e < b || e > m || (p == 0 && b != 0) || (p == n - 1 && e != m)) {
      atomicMin(bad + 1, p);
      lev_count[p] = 0;
      continue;
    }
    int c = 0;
    for (int j = b; j < e; ++j) {
      const int u = fanin[j];
      arc_to[j] = p;
      const bool live = kind[j] != OTG_ARC_SKIP;
      if (u < 0 || u >= n) {
        atomicMin(bad, p);
        sort_key[j] = n;
        continue;
      }
      sort_key[j] = live ? u : n;
      if (live) {
        ++c;
        atomicAdd(&out_degree[u], 1);
      }
    }
    lev_count[p] = c;
  }
}

__global__ void timer_lev_fill_kernel(int32_t n, const int32_t* __restrict__ fanin_start,
                                      const int32_t* __restrict__ fanin,
                                      const int32_t* __restrict__ kind,
                                      const int32_t* __restrict__ lev_start,
                                      int32_t* __restrict__ lev_fanin) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    int q = lev_start[p];
    for (int j = fanin_start[p]; j < fanin_start[p + 1]; ++j)
      if (kind[j] != OTG_ARC_SKIP) lev_fanin[q++] = fanin[j];
  }
}

__global__ void timer_no_fanout_kernel(int32_t n, const int32_t* __restrict__ out_degree,
                                       int32_t* __restrict__ flag) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x)
    flag[p] = out_degree[p] == 0;
}

// Where every BFS position takes its R and C from (prepare_res_cap,
// reference rct.cu:314-359): node u of net k sits at position node2bfs[u],
// and the resistance of each oriented edge (parent, child) belongs to the
// child's position.  One wave per net; run once, at the first update.
__global__ void timer_index_kernel(int32_t N, const int32_t* __restrict__ ns,
                                   const int32_t* __restrict__ node2bfs,
                                   const int32_t* __restrict__ edges_dir,
                                   int32_t* __restrict__ src_pos, int32_t* __restrict__ res_edge) {
  const int warp = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int nw = (gridDim.x * blockDim.x) >> 6;
  for (int k = warp; k < N; k += nw) {
    const int b = ns[k], n = ns[k + 1] - b;
    const int eb = b - k;
    for (int u = lane; u < n; u += 64) src_pos[b + node2bfs[b + u]] = b + u;
    for (int i = lane; i < n - 1; i += 64)
      res_edge[b + node2bfs[b + edges_dir[2 * (eb + i) + 1]]] = eb + i;
  }
}

// R and C of every net into the kept BFS order, one BFS position per lane
// through the position's source indices: a wave takes 256 consecutive
// positions, lane l the positions l, l + 64, l + 128, l + 192, so that every
// index load, capacitance store and resistance store of the wave covers one
// contiguous range (the capacitance and resistance loads gather within the
// few nets those positions belong to).  All loads are issued before the
// first store.
__global__ void timer_reorder_kernel(int64_t V, const int32_t* __restrict__ src_pos,
                                     const int32_t* __restrict__ res_edge,
                                     const float* __restrict__ edge_res,
                                     const float4* __restrict__ cap, float4* __restrict__ cap_bfs,
                                     float* __restrict__ pres) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t waves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t b = wave * 256; b < V; b += waves * 256) {
    if (b + 256 <= V) {
      int sp[4], ep[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        sp[r] = src_pos[b + lane + 64 * r];
        ep[r] = res_edge[b + lane + 64 * r];
      }
      float4 c[4];
      float rr[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        c[r] = cap[sp[r]];
        rr[r] = ep[r] >= 0 ? edge_res[ep[r]] : 0.0f;
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        cap_bfs[b + lane + 64 * r] = c[r];
        pres[b + lane + 64 * r] = rr[r];
      }
    } else {
      for (int64_t i = b + lane; i < V; i += 64) {
        cap_bfs[i] = cap[src_pos[i]];
        const int e = res_edge[i];
        pres[i] = e >= 0 ? edge_res[e] : 0.0f;
      }
    }
  }
}

// Net values the propagation reads: for an RC net arc j, the RC delay and
// impulse at the sink's node (FlatRct::delay / impulse through node2bfs,
// net.cpp:280-287); for a pin that drives an RC net, that net's root load
// (Net::_load, net.cpp:626-642; the root is BFS position 0), else the
// given pin load.  Out-of-range net or node ids give NaN.
__global__ void timer_gather_kernel(int32_t m, int32_t n, int32_t N, const int32_t* __restrict__ ns,
                                    const int32_t* __restrict__ node2bfs,
                                    const int32_t* __restrict__ kind,
                                    const int32_t* __restrict__ arc_net,
                                    const int32_t* __restrict__ arc_node,
                                    const int32_t* __restrict__ pin_net,
                                    const float* __restrict__ pin_load,
                                    const float4* __restrict__ rc_load,
                                    const float4* __restrict__ rc_delay,
                                    const float4* __restrict__ rc_impulse,
                                    float4* __restrict__ arc_d, float4* __restrict__ arc_i,
                                    float4* __restrict__ load) {
  const float q = __builtin_nanf("");
  const float4 nan4 = make_float4(q, q, q, q);
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < m + n; j += gridDim.x * blockDim.x) {
    if (j < m) {
      float4 d = nan4, im = nan4;
      if (kind[j] == OTG_ARC_NET_RC) {
        const int k = arc_net[j], u = arc_node[j];
        if (k >= 0 && k < N) {
          const int b = ns[k];
          if (u >= 0 && u < ns[k + 1] - b) {
            const int i = b + node2bfs[b + u];
            d = rc_delay[i];
            im = rc_impulse[i];
          }
        }
      }
      arc_d[j] = d;
      arc_i[j] = im;
    } else {
      const int p = j - m;
      const int k = pin_net[p];
      float4 l;
      if (k >= 0 && k < N && ns[k + 1] > ns[k]) {
        l = rc_load[ns[k]];
      } else {
        l = make_float4(pin_load[4 * p], pin_load[4 * p + 1], pin_load[4 * p + 2], pin_load[4 * p + 3]);
      }
      load[p] = l;
    }
  }
}

__global__ void timer_report_kernel(int32_t count, const int32_t* __restrict__ pins, int32_t n,
                                    const float* __restrict__ v, float* __restrict__ out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < 4 * count; i += gridDim.x * blockDim.x) {
    const int p = pins[i >> 2];
    out[i] = (p >= 0 && p < n) ? v[4 * p + (i & 3)] : __builtin_nanf("");
  }
}

template <class T>
hipError_t dalloc(T** p, int64_t count) {
  *p = nullptr;
  return hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * std::max<int64_t>(1, count));
}

bool needs(int64_t count, const void* p) { return count <= 0 || p != nullptr; }

otg_status exclusive_scan(otg_context* ctx, const int32_t* in, int32_t* out, int n,
                          hipStream_t s) {
  size_t tb = 0;
  OTG_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, in, out, n, s));
  void* tmp = nullptr;
  OTG_HIP(ctx, hipMallocAsync(&tmp, std::max<size_t>(tb, 1), s));
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(tmp, tb, in, out, n, s);
  (void)hipFreeAsync(tmp, s);
  OTG_HIP(ctx, e);
  return OTG_OK;
}

__global__ void timer_fanout_to_kernel(int32_t m, const int32_t* __restrict__ fanout_arc,
                                       const int32_t* __restrict__ arc_to,
                                       int32_t* __restrict__ fanout_to) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x)
    fanout_to[k] = arc_to[fanout_arc[k]];
}

__global__ void timer_iota_kernel(int32_t* a, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = (int32_t)i;
}

// Stream-ordered working arrays, freed when the scope ends.
struct TimerTmp {
  hipStream_t s;
  std::vector<void*> p;
  ~TimerTmp() {
    for (void* q : p) (void)hipFreeAsync(q, s);
  }
  template <class X>
  hipError_t get(X** out, int64_t count) {
    *out = nullptr;
    const hipError_t e =
        hipMallocAsync(reinterpret_cast<void**>(out), sizeof(X) * std::max<int64_t>(1, count), s);
    if (e == hipSuccess) p.push_back(*out);
    return e;
  }
};

// Everything that depends on the topology only (first update).
otg_status timer_build(otg_timer* T, hipStream_t s) {
  otg_context* ctx = T->ctx;
  const otg_timer_design& d = T->d;
  const int N = d.num_nets, n = d.num_pins, m = d.num_arcs;
  if (N > 0) {
    const otg_rct_edges in{N, d.total_num_nodes, d.nodes_start, d.edges, d.roots, d.edge_res,
                           d.node_cap};
    otg_status st = otg_rct_bfs(ctx, &in, T->node2bfs, T->pid, T->edges_dir, T->pres, T->cap_bfs, s);
    if (st != OTG_OK) return st;
    // source indices of every BFS position for the later updates (in-range
    // defaults first: the identity and "no edge")
    const int64_t V = d.total_num_nodes;
    hipLaunchKernelGGL(timer_iota_kernel, dim3(blocks_for(V)), dim3(kBlock), 0, s, T->src_pos, V);
    hipLaunchKernelGGL(timer_fill_i32_kernel, dim3(blocks_for(V)), dim3(kBlock), 0, s, T->res_edge,
                       V, -1);
    hipLaunchKernelGGL(timer_index_kernel, dim3(blocks_for(64LL * N)), dim3(kBlock), 0, s, N,
                       d.nodes_start, T->node2bfs, T->edges_dir, T->src_pos, T->res_edge);
    OTG_HIP(ctx, hipGetLastError());
    if (T->method != OTG_RCT_PLAIN) {
      const otg_rct_tree topo{N, d.total_num_nodes, d.nodes_start, T->pid, T->pres, T->cap_bfs,
                              nullptr};
      st = otg_rct_plan_create(ctx, &topo, nullptr, T->method, &T->rc_plan);
      if (st != OTG_OK) return st;
    }
  }
  if (n == 0) return OTG_OK;
  TimerTmp W{s, {}};  // working arrays, freed on every path
  int32_t *cnt = nullptr, *bad = nullptr, *key = nullptr, *key2 = nullptr, *seq = nullptr,
          *flag = nullptr, *nsel = nullptr;
  OTG_HIP(ctx, W.get(&cnt, n + 1));
  OTG_HIP(ctx, W.get(&bad, 2));
  OTG_HIP(ctx, W.get(&key, m));
  OTG_HIP(ctx, W.get(&key2, m));
  OTG_HIP(ctx, W.get(&seq, m));
  OTG_HIP(ctx, W.get(&flag, n));
  OTG_HIP(ctx, W.get(&nsel, 1));
  // arcs' sinks, propagating fanin counts, fanout counts
  hipLaunchKernelGGL(timer_fill_i32_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, s, T->out_degree,
                     (int64_t)n, 0);
  hipLaunchKernelGGL(timer_fill_i32_kernel, dim3(1), dim3(kBlock), 0, s, bad, (int64_t)2, INT32_MAX);
  hipLaunchKernelGGL(timer_fill_i32_kernel, dim3(1), dim3(kBlock), 0, s, cnt + n, (int64_t)1, 0);
  hipLaunchKernelGGL(timer_arc_count_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, m,
                     d.fanin_start, d.fanin, d.arc_kind, cnt, T->out_degree, T->arc_to, key, bad);
  OTG_HIP(ctx, hipGetLastError());
  int32_t hbad[2] = {INT32_MAX, INT32_MAX};
  OTG_HIP(ctx, hipMemcpyAsync(hbad, bad, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  OTG_HIP(ctx, hipStreamSynchronize(s));
  if (hbad[1] != INT32_MAX) {
    OTG_SET_ERR(ctx, "pin %d: fanin_start must run from 0 to num_arcs without decreasing",
                hbad[1]);
    return OTG_ERR_INVALID_ARGUMENT;
  }
  if (hbad[0] != INT32_MAX) {
    OTG_SET_ERR(ctx, "pin %d: a fanin arc names a pin out of range", hbad[0]);
    return OTG_ERR_INVALID_ARGUMENT;
  }
  // levelization CSR (exclusive scan of the counts, then the arcs in order)
  otg_status st = exclusive_scan(ctx, cnt, T->lev_start, n + 1, s);
  if (st != OTG_OK) return st;
  hipLaunchKernelGGL(timer_lev_fill_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, s, n,
                     d.fanin_start, d.fanin, d.arc_kind, T->lev_start, T->lev_fanin);
  // fanout CSR: offsets from the fanout counts; the propagating arcs grouped
  // by source, in arc order within a source (a stable sort by source key;
  // loop breakers carry key n and sort last)
  OTG_HIP(ctx, hipMemcpyAsync(cnt, T->out_degree, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
  st = exclusive_scan(ctx, cnt, T->fanout_start, n + 1, s);
  if (st != OTG_OK) return st;
  if (m > 0) {
    hipLaunchKernelGGL(timer_iota_kernel, dim3(blocks_for(m)), dim3(kBlock), 0, s, seq, (int64_t)m);
    const int end_bit = 32 - __builtin_clz((unsigned)n);
    size_t tb = 0;
    OTG_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key, key2, seq, T->fanout_arc, m,
                                                    0, end_bit, s));
    char* tmp = nullptr;
    OTG_HIP(ctx, W.get(&tmp, (int64_t)tb));
    OTG_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(tmp, tb, key, key2, seq, T->fanout_arc, m, 0,
                                                    end_bit, s));
    hipLaunchKernelGGL(timer_fanout_to_kernel, dim3(blocks_for(m)), dim3(kBlock), 0, s, m,
                       T->fanout_arc, T->arc_to, T->fanout_to);
  }
  // first frontier: the pins without fanout, ascending
  hipLaunchKernelGGL(timer_no_fanout_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, s, n,
                     T->out_degree, flag);
  {
    hipcub::CountingInputIterator<int32_t> iota(0);
    size_t tb = 0;
    OTG_HIP(ctx, hipcub::DeviceSelect::Flagged(nullptr, tb, iota, flag, T->first, nsel, n, s));
    char* tmp = nullptr;
    OTG_HIP(ctx, W.get(&tmp, (int64_t)tb));
    OTG_HIP(ctx, hipcub::DeviceSelect::Flagged(tmp, tb, iota, flag, T->first, nsel, n, s));
  }
  int32_t hsz[2] = {0, 0};
  OTG_HIP(ctx, hipMemcpyAsync(&hsz[0], nsel, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  OTG_HIP(ctx, hipMemcpyAsync(&hsz[1], T->lev_start + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  OTG_HIP(ctx, hipStreamSynchronize(s));
  int32_t num_ends = 0;
  st = otg_levelize(ctx, n, hsz[1], T->lev_start, T->lev_fanin, T->out_degree, T->first, hsz[0],
                    T->frontiers, T->frontier_ends, nullptr, &num_ends, s);
  if (st != OTG_OK) return st;
  T->num_levels = std::max(0, num_ends - 1);
  return otg_fprop_plan_create(ctx, n, m, d.fanin_start, d.fanin, T->frontiers, T->frontier_ends,
                               T->num_levels, s, &T->fplan);
}

void timer_free(otg_timer* T) {
  if (T->rc_plan) otg_rct_plan_destroy(T->rc_plan);
  if (T->fplan) otg_fprop_plan_destroy(T->fplan);
  for (void* p : {(void*)T->node2bfs, (void*)T->pid, (void*)T->edges_dir, (void*)T->src_pos,
                  (void*)T->res_edge, (void*)T->pres,
                  (void*)T->cap_bfs, (void*)T->rc_load, (void*)T->rc_delay, (void*)T->rc_impulse,
                  (void*)T->arc_rc_delay, (void*)T->arc_rc_impulse, (void*)T->cell_load,
                  (void*)T->lev_start, (void*)T->lev_fanin, (void*)T->out_degree, (void*)T->first,
                  (void*)T->frontiers, (void*)T->frontier_ends, (void*)T->fanout_start,
                  (void*)T->fanout_arc, (void*)T->fanout_to, (void*)T->arc_to, (void*)T->slew,
                  (void*)T->at,
                  (void*)T->rat, (void*)T->slack, (void*)T->arc_delay})
    if (p) (void)hipFree(p);
  delete T;
}

}  // namespace

extern "C" otg_status otg_timer_create(otg_context* ctx, const otg_timer_design* d,
                                       const otg_prop_lib* lib, otg_rct_method method,
                                       otg_timer** out) {
  if (!ctx || !d || !lib || !out) return OTG_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  if (method != OTG_RCT_PLAIN && method != OTG_RCT_BATCHED && method != OTG_RCT_BATCHED_COMPENSATED)
    return OTG_ERR_INVALID_ARGUMENT;
  const int64_t N = d->num_nets, V = d->total_num_nodes, n = d->num_pins, m = d->num_arcs;
  if (N < 0 || V < N || n < 0 || m < 0) {
    OTG_SET_ERR(ctx, "timer design: sizes must satisfy 0 <= num_nets <= total_num_nodes, "
                     "num_pins >= 0, num_arcs >= 0");
    return OTG_ERR_INVALID_ARGUMENT;
  }
  const bool ok = needs(N, d->nodes_start) && needs(N, d->roots) && needs(N, d->node_cap) &&
                  needs(V - N, d->edges) && needs(V - N, d->edge_res) &&
                  (n == 0 || (d->fanin_start && d->pin_net && d->pin_load && d->pi_slew &&
                              d->pi_at && d->po_rat)) &&
                  needs(m, d->fanin) && needs(m, d->arc_kind) && needs(m, d->arc_timing) &&
                  needs(m, d->arc_net) && needs(m, d->arc_node);
  if (!ok) {
    OTG_SET_ERR(ctx, "timer design: an array of a non-empty part is NULL");
    return OTG_ERR_INVALID_ARGUMENT;
  }
  OTG_HIP(ctx, hipSetDevice(ctx->device));
  otg_timer* T = new otg_timer();
  T->ctx = ctx;
  T->d = *d;
  T->lib = *lib;
  T->method = method;
  auto fail = [&](hipError_t e) {
    timer_free(T);
    OTG_SET_ERR(ctx, "timer allocation failed: %s", hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? OTG_ERR_OUT_OF_MEMORY : OTG_ERR_DEVICE;
  };
  hipError_t e = hipSuccess;
  auto A = [&](auto** p, int64_t count) {
    if (e == hipSuccess) e = dalloc(p, count);
  };
  A(&T->node2bfs, V);
  A(&T->pid, V);
  A(&T->edges_dir, 2 * (V - N));
  A(&T->src_pos, V);
  A(&T->res_edge, V);
  A(&T->pres, V);
  A(&T->cap_bfs, 4 * V);
  A(&T->rc_load, 4 * V);
  A(&T->rc_delay, 4 * V);
  A(&T->rc_impulse, 4 * V);
  A(&T->arc_rc_delay, 4 * m);
  A(&T->arc_rc_impulse, 4 * m);
  A(&T->cell_load, 4 * n);
  A(&T->lev_start, n + 1);
  A(&T->lev_fanin, m);
  A(&T->out_degree, n);
  A(&T->first, n);
  A(&T->frontiers, n);
  A(&T->frontier_ends, n + 2);
  A(&T->fanout_start, n + 1);
  A(&T->fanout_arc, m);
  A(&T->arc_to, m);
  A(&T->fanout_to, m);
  A(&T->slew, 4 * n);
  A(&T->at, 4 * n);
  A(&T->rat, 4 * n);
  A(&T->slack, 4 * n);
  A(&T->arc_delay, 8 * m);
  if (e != hipSuccess) return fail(e);
  // unset until the first update
  const float q = std::nanf("");
  for (float* p : {T->slew, T->at, T->rat, T->slack})
    hipLaunchKernelGGL(timer_fill_kernel, dim3(blocks_for(4 * n)), dim3(kBlock), 0, ctx->stream, p,
                       4 * n, q);
  hipLaunchKernelGGL(timer_fill_kernel, dim3(blocks_for(8 * m)), dim3(kBlock), 0, ctx->stream,
                     T->arc_delay, 8 * m, q);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return fail(e);
  *out = T;
  return OTG_OK;
}

extern "C" otg_status otg_timer_destroy(otg_timer* T) {
  if (!T) return OTG_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(T->ctx->device);
  (void)hipDeviceSynchronize();
  timer_free(T);
  return OTG_OK;
}

extern "C" otg_status otg_timer_update(otg_timer* T, otg_stream stream) {
  if (!T) return OTG_ERR_INVALID_ARGUMENT;
  otg_context* ctx = T->ctx;
  const otg_timer_design& d = T->d;
  const int N = d.num_nets, n = d.num_pins, m = d.num_arcs;
  hipStream_t s = otg_pick_stream(ctx, stream);
  OTG_HIP(ctx, hipSetDevice(ctx->device));
  if (!T->built) {
    const otg_status st = timer_build(T, s);
    if (st != OTG_OK) return st;
    T->built = true;
  } else if (N > 0) {
    // R and C afresh, in the kept BFS order
    const int64_t V = d.total_num_nodes;
    hipLaunchKernelGGL(timer_reorder_kernel, dim3(blocks_for((V + 3) / 4)), dim3(kBlock), 0, s, V,
                       T->src_pos, T->res_edge, d.edge_res,
                       reinterpret_cast<const float4*>(d.node_cap),
                       reinterpret_cast<float4*>(T->cap_bfs), T->pres);
    OTG_HIP(ctx, hipGetLastError());
  }
  // ---- RC phase ----
  if (N > 0) {
    const otg_rct_tree tree{N, d.total_num_nodes, d.nodes_start, T->pid, T->pres, T->cap_bfs,
                            nullptr};
    const otg_rct_result res{T->rc_load, T->rc_delay, nullptr, nullptr, T->rc_impulse, nullptr};
    const otg_status st = otg_rct_solve(ctx, T->rc_plan, &tree, &res, T->method, stream);
    if (st != OTG_OK) return st;
  }
  if (n == 0) return OTG_OK;
  // ---- net values at the arcs' sinks and the cell loads ----
  hipLaunchKernelGGL(timer_gather_kernel, dim3(blocks_for((int64_t)m + n)), dim3(kBlock), 0, s, m, n,
                     N, d.nodes_start, T->node2bfs, d.arc_kind, d.arc_net, d.arc_node, d.pin_net,
                     d.pin_load, reinterpret_cast<const float4*>(T->rc_load),
                     reinterpret_cast<const float4*>(T->rc_delay),
                     reinterpret_cast<const float4*>(T->rc_impulse),
                     reinterpret_cast<float4*>(T->arc_rc_delay),
                     reinterpret_cast<float4*>(T->arc_rc_impulse),
                     reinterpret_cast<float4*>(T->cell_load));
  OTG_HIP(ctx, hipGetLastError());
  // ---- forward and backward propagation over the levels ----
  const otg_prop_graph g{n, m, d.fanin_start, d.fanin, d.arc_kind, d.arc_timing, T->arc_rc_delay,
                         T->arc_rc_impulse, T->cell_load, d.pi_slew, d.pi_at};
  const otg_prop_result out{T->slew, T->at, T->arc_delay};
  otg_status st = otg_fprop_run(ctx, T->fplan, &g, &T->lib, &out, stream);
  if (st != OTG_OK) return st;
  const otg_bprop_graph bg{n,          m,         T->fanout_start, T->fanout_arc, T->arc_to,
                           T->arc_delay, d.po_rat, T->at,           T->fanout_to};
  return otg_bprop(ctx, &bg, T->frontiers, T->frontier_ends, T->num_levels, T->rat, T->slack,
                   stream);
}

extern "C" otg_status otg_timer_results_get(const otg_timer* T, otg_timer_results* r) {
  if (!T || !r) return OTG_ERR_INVALID_ARGUMENT;
  *r = otg_timer_results{T->slew, T->at, T->rat, T->slack, T->arc_delay, T->arc_rc_delay,
                         T->arc_rc_impulse, T->cell_load, T->frontiers, T->frontier_ends,
                         T->num_levels};
  return OTG_OK;
}

extern "C" otg_status otg_timer_report(otg_timer* T, otg_timer_quantity what, int32_t count,
                                       const int32_t* pins, float* out, otg_stream stream) {
  if (!T || count < 0 || (count > 0 && (!pins || !out))) return OTG_ERR_INVALID_ARGUMENT;
  const float* v = what == OTG_REPORT_SLEW ? T->slew
                   : what == OTG_REPORT_AT ? T->at
                   : what == OTG_REPORT_RAT ? T->rat
                   : what == OTG_REPORT_SLACK ? T->slack : nullptr;
  if (!v) return OTG_ERR_INVALID_ARGUMENT;
  if (count == 0) return OTG_OK;
  otg_context* ctx = T->ctx;
  OTG_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = otg_pick_stream(ctx, stream);
  hipLaunchKernelGGL(timer_report_kernel, dim3(blocks_for(4LL * count)), dim3(kBlock), 0, s, count,
                     pins, T->d.num_pins, v, out);
  OTG_HIP(ctx, hipGetLastError());
  return OTG_OK;
}
