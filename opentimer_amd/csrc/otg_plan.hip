// Device-side construction of the work list of the batched P1 solve.
//
// A plan (otg_rct_plan) holds, for one topology (nodes_start, pid):
//  * packed groups: runs of consecutive small nets (at most G nodes each),
//    at most small_target positions and G/2 nets per group, each with a
//    static schedule of one 32-bit word per node (node | depth << 8 | first
//    child << 16 | children << 24, nodes sorted by depth, parent,
//    position), the groups ordered by their number of depth steps, most
//    first;
//  * streamed nets (more than G nodes) in bins by the smallest ring that
//    holds their parent offsets (and, compensated, by whether they are chain
//    nets), largest first, with a scratch region each;
//  * per net: largest parent offset, chain flag, scratch offset, group.
// Everything is computed on the device from the device arrays; the host
// sees only per-segment group counts (to place the groups) and one record
// per streamed net (to bin and order them, a few tens of thousands), so no
// per-node array crosses PCIe.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <numeric>
#include <vector>

#include "otg_internal.hpp"

namespace {

constexpr int kWave = 64;
constexpr int kSeg = 1024;   // nets per packing segment (one thread each)
constexpr int kBatch = 8;    // chunks whose loads are in flight together

struct PlanCtl {
  int32_t bad;          // smallest net whose pid is out of range (INT32_MAX: none)
  int32_t num_large;    // nets larger than a group
  int32_t max_net;
  int32_t packed_nets;  // nets of 1..G nodes
  int32_t start_bad;    // nonzero: nodes_start decreases somewhere
  int32_t pad[3];
};

__device__ __forceinline__ uint64_t lanes_below(int lane) { return (1ull << lane) - 1ull; }

// Sizes: the streamed nets' list, counts, largest net (wave-aggregated atomics).
__global__ __launch_bounds__(256) void plan_size_kernel(int32_t N, const int32_t* __restrict__ ns,
                                                        int32_t G, int32_t* __restrict__ large,
                                                        PlanCtl* c) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const int n = k < N ? ns[k + 1] - ns[k] : 0;
  const bool big = k < N && n > G;
  const bool small = k < N && n > 0 && n <= G;
  const uint64_t mb = __ballot(big), ms = __ballot(small);
  int base = 0;
  if (lane == 0 && mb) base = atomicAdd(&c->num_large, __popcll(mb));
  base = __shfl(base, 0);
  if (big) large[base + __popcll(mb & lanes_below(lane))] = k;
  if (lane == 0 && ms) atomicAdd(&c->packed_nets, __popcll(ms));
  int mx = n;
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) mx = max(mx, __shfl_xor(mx, d));
  if (lane == 0 && mx > 0) atomicMax(&c->max_net, mx);
  if (__any(k < N && n < 0) && lane == 0) atomicOr(&c->start_bad, 1);
}

// Small nets (1..G nodes), one thread each: pid in range, largest offset.
__global__ __launch_bounds__(256) void plan_small_info_kernel(int32_t N, const int32_t* __restrict__ ns,
                                                              const int32_t* __restrict__ pid,
                                                              int32_t G, int32_t* __restrict__ maxpid,
                                                              int32_t* __restrict__ chain, PlanCtl* c) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  const int st = ns[k], ed = ns[k + 1];
  if (ed - st > G) return;  // the streamed nets' kernel
  int m = 0;
  bool ok = true;
  for (int i = st + 1; i < ed; ++i) {
    const int q = pid[i];
    if (q < 1 || q > i - st) ok = false;
    m = max(m, q);
  }
  maxpid[k] = m;
  chain[k] = 0;
  if (!ok) atomicMin(&c->bad, k);
}

// Streamed nets, one wavefront each: pid in range, largest offset, and the
// chain flag (at least half of the net's 64-position windows, and more than
// one window, are paths: every position but the window's first has the
// position before it as its parent).
__global__ __launch_bounds__(64) void plan_large_info_kernel(const int32_t* __restrict__ ns,
                                                             const int32_t* __restrict__ pid,
                                                             const int32_t* __restrict__ large,
                                                             int32_t* __restrict__ maxpid,
                                                             int32_t* __restrict__ chain, PlanCtl* c) {
  const int lane = threadIdx.x;
  const int count = c->num_large;
  for (int idx = blockIdx.x; idx < count; idx += gridDim.x) {
    const int k = large[idx];
    const int st = ns[k], n = ns[k + 1] - st;
    int m = 0, paths = 0, wins = 0;
    bool bad = false;
    for (int w0 = 0; w0 < n; w0 += kBatch * kWave) {
      int q[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int o = w0 + u * kWave + lane;
        q[u] = o < n ? pid[st + o] : 1;
      }
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int o = w0 + u * kWave + lane;
        if (w0 + u * kWave >= n) break;  // uniform
        const bool valid = o < n && o > 0;
        if (valid && (q[u] < 1 || q[u] > o)) bad = true;
        if (valid) m = max(m, q[u]);
        const bool notpath = valid && (o & (kWave - 1)) != 0 && q[u] != 1;
        paths += __ballot(notpath) == 0ull;
        ++wins;
      }
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) m = max(m, __shfl_xor(m, d));
    if (lane == 0) {
      maxpid[k] = m;
      chain[k] = 2 * paths >= wins && wins > 1;
    }
    if (__any(bad)) atomicMin(&c->bad, k);
  }
}

// Greedy packing of consecutive small nets, one thread per segment of kSeg
// nets (a group never crosses a segment boundary).  WRITE = false counts
// the segment's groups and their positions; WRITE = true writes them from
// the segment's offsets.
template <bool WRITE>
__global__ __launch_bounds__(256) void plan_pack_kernel(int32_t N, const int32_t* __restrict__ ns,
                                                        int32_t G, int32_t cap, int32_t target,
                                                        int32_t* __restrict__ seg_groups,
                                                        int64_t* __restrict__ seg_words,
                                                        const int32_t* __restrict__ goff,
                                                        const int64_t* __restrict__ woff,
                                                        RctItem* __restrict__ groups,
                                                        int32_t* __restrict__ group_of_net) {
  const int seg = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = seg * kSeg;
  if (a >= N) return;
  const int b = min(N, a + kSeg);
  int ng = 0;
  int64_t words = 0;
  int gidx = WRITE ? goff[seg] : 0;
  int64_t wpos = WRITE ? woff[seg] : 0;
  RctItem cur{};
  bool open = false;
  auto flush = [&]() {
    if (!open) return;
    if constexpr (WRITE) {
      cur.k = gidx;
      cur.scratch = wpos;
      groups[gidx] = cur;
      for (int j = cur.net_begin; j < cur.net_end; ++j) group_of_net[j] = gidx;
    }
    ++gidx;
    ++ng;
    wpos += cur.pos_end - cur.pos_begin;
    words += cur.pos_end - cur.pos_begin;
    open = false;
  };
  int prev = ns[a];
  for (int k0 = a; k0 < b; k0 += 16) {
    int nxt[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) nxt[u] = k0 + u < b ? ns[k0 + u + 1] : 0;
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int k = k0 + u;
      if (k >= b) break;
      const int st = prev, n = nxt[u] - prev;
      prev = nxt[u];
      if (n <= 0 || n > G) {
        flush();
        continue;
      }
      if (open && (cur.pos_end - cur.pos_begin) + n <= target && cur.net_end - cur.net_begin < cap / 2) {
        cur.pos_end += n;
        cur.net_end = k + 1;
      } else {
        flush();
        cur = RctItem{st, st + n, k, k + 1, ITEM_PACKED, 0, 0};
        open = true;
      }
    }
  }
  flush();
  if constexpr (!WRITE) {
    seg_groups[seg] = ng;
    seg_words[seg] = words;
  }
}

// Static schedule of every group, one wavefront per group (at most 256
// positions, 4 per lane): depths by pointer jumping, the (depth, parent,
// position) order by ranking the packed keys, children's first schedule
// index and count by LDS atomics; the group's number of depth steps (a step
// takes up to 64 entries of one depth) is its cost for ordering.
__global__ __launch_bounds__(64) void plan_sched_kernel(int32_t num_groups,
                                                        const RctItem* __restrict__ groups,
                                                        const int32_t* __restrict__ ns,
                                                        const int32_t* __restrict__ pid,
                                                        uint32_t* __restrict__ sched,
                                                        uint32_t* __restrict__ cost_key) {
  __shared__ int32_t par[256], anc[256], dep[256], ord[256], cf[256], cn[256];
  __shared__ uint32_t key[256];
  const int lane = threadIdx.x;
  for (int g = blockIdx.x; g < num_groups; g += gridDim.x) {
    const RctItem it = groups[g];
    const int pb = it.pos_begin, n = it.pos_end - pb;
    // parents (drivers: -1)
    for (int q = lane; q < 256; q += kWave) {
      par[q] = q < n ? q - pid[pb + q] : -1;
      cn[q] = 0;
      cf[q] = 255;
    }
    __syncthreads();
    for (int j = it.net_begin + lane; j < it.net_end; j += kWave) par[ns[j] - pb] = -1;
    __syncthreads();
    for (int q = lane; q < 256; q += kWave) {
      anc[q] = par[q] < 0 ? q : par[q];
      dep[q] = par[q] < 0 ? 0 : 1;
    }
    __syncthreads();
    // depth = distance to the driver: 8 doubling rounds cover 256 positions
    for (int r = 0; r < 8; ++r) {
      int na[4], nd[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int q = u * kWave + lane;
        const int a = anc[q];
        na[u] = anc[a];
        nd[u] = dep[q] + (a == q ? 0 : dep[a]);
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int q = u * kWave + lane;
        anc[q] = na[u];
        dep[q] = nd[u];
      }
      __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = u * kWave + lane;
      key[q] = q < n ? ((uint32_t)dep[q] << 17) | ((uint32_t)(par[q] + 1) << 8) | (uint32_t)q
                     : 0xffffffffu;
    }
    __syncthreads();
    int rk[4] = {0, 0, 0, 0};
    uint32_t mine[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) mine[u] = key[u * kWave + lane];
    for (int r = 0; r < n; ++r) {
      const uint32_t kr = key[r];
#pragma unroll
      for (int u = 0; u < 4; ++u) rk[u] += kr < mine[u];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = u * kWave + lane;
      if (q < n) {
        ord[rk[u]] = q;
        if (par[q] >= 0) {
          atomicAdd(&cn[par[q]], 1);
          atomicMin(&cf[par[q]], rk[u]);
        }
      }
    }
    __syncthreads();
    int steps = 0;
    int prevd = -1;
    for (int j0 = 0; j0 < n; j0 += kWave) {
      const int j = j0 + lane;
      const int q = j < n ? ord[j] : 0;
      const int d = dep[q];
      const int dprev = j > 0 ? dep[ord[j - 1]] : -1;
      if (j < n) {
        const int c = cn[q];
        sched[it.scratch + j] = (uint32_t)q | ((uint32_t)d << 8) |
                                ((uint32_t)(c ? cf[q] : 0) << 16) | ((uint32_t)c << 24);
      }
      steps += __popcll(__ballot(j < n && (j == 0 || d != dprev || (j & 63) == 0)));
      (void)prevd;
    }
    if (lane == 0) cost_key[g] = 255u - (uint32_t)min(steps, 255);
    __syncthreads();
  }
}

// Groups in cost order: item k is the group of sorted position k.
__global__ __launch_bounds__(256) void plan_permute_kernel(int32_t num_groups,
                                                           const uint32_t* __restrict__ order,
                                                           const RctItem* __restrict__ in,
                                                           RctItem* __restrict__ out,
                                                           int32_t* __restrict__ net_group) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= num_groups) return;
  RctItem it = in[order[i]];
  it.k = i;
  out[i] = it;
  for (int j = it.net_begin; j < it.net_end; ++j) net_group[j] = i;
}

__global__ __launch_bounds__(256) void plan_fill_kernel(int32_t* p, int64_t n, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ __launch_bounds__(256) void plan_fill64_kernel(int64_t* p, int64_t n, int64_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// One record per streamed net: {net, first position, size, 2 * largest
// offset + chain flag}.
__global__ __launch_bounds__(256) void plan_large_gather_kernel(int32_t count,
                                                                const int32_t* __restrict__ large,
                                                                const int32_t* __restrict__ ns,
                                                                const int32_t* __restrict__ maxpid,
                                                                const int32_t* __restrict__ chain,
                                                                int4* __restrict__ rec) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int k = large[i];
  rec[i] = make_int4(k, ns[k], ns[k + 1] - ns[k], 2 * maxpid[k] + (chain[k] ? 1 : 0));
}

__global__ __launch_bounds__(256) void plan_scatter_scratch_kernel(int32_t count,
                                                                   const int64_t* __restrict__ pairs,
                                                                   int64_t* __restrict__ net_scratch) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  net_scratch[pairs[2 * i]] = pairs[2 * i + 1];
}

// Stream-ordered temporaries of one plan build, freed on every path.
struct TmpList {
  std::vector<void*> v;
  hipStream_t s;
  ~TmpList() {
    for (void* q : v) (void)hipFreeAsync(q, s);
  }
  template <class T>
  hipError_t alloc(T** out, int64_t count) {
    void* q = nullptr;
    hipError_t e = hipMallocAsync(&q, std::max<int64_t>(1, count) * (int64_t)sizeof(T), s);
    if (e == hipSuccess) v.push_back(q);
    *out = static_cast<T*>(q);
    return e;
  }
};

int grid_for(int64_t n, int block) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + block - 1) / block, 1 << 20));
}

}  // namespace

// Builds the device arrays of plan p (its tuning, method, bins, G resolved).
// Returns OTG_OK or an error with the context's message set; on error the
// caller destroys the plan.
otg_status otg_rct_plan_build_device(otg_context* ctx, otg_rct_plan* p, const otg_rct_tree* topo,
                                     hipStream_t s) {
  const int N = topo->num_nets;
  const int64_t V = topo->total_num_nodes;
  const int G = p->small_max;     // nets up to G nodes are packed
  const int cap = p->group_cap;   // a group holds at most cap positions and cap/2 nets
  const bool comp = p->method == OTG_RCT_BATCHED_COMPENSATED;
  const int W = p->tuning.ring_slots;
  const int32_t* ns = topo->nodes_start;
  const int32_t* pid = topo->pid;
  if (N == 0) return OTG_OK;
  // temporaries (freed on every path)
  TmpList T{{}, s};
  PlanCtl* ctl = nullptr;
  int32_t* large = nullptr;
  OTG_HIP(ctx, T.alloc(&ctl, 1));
  OTG_HIP(ctx, T.alloc(&large, N));
  OTG_HIP(ctx, hipMalloc(&p->d_net_maxpid, sizeof(int32_t) * N));
  OTG_HIP(ctx, hipMalloc(&p->d_net_chain, sizeof(int32_t) * N));
  OTG_HIP(ctx, hipMalloc(&p->d_net_scratch, sizeof(int64_t) * N));
  {
    const PlanCtl init{INT32_MAX, 0, 0, 0, 0, {0, 0, 0}};
    OTG_HIP(ctx, hipMemcpyAsync(ctl, &init, sizeof(init), hipMemcpyHostToDevice, s));
  }
  hipLaunchKernelGGL(plan_size_kernel, dim3(grid_for(N, 256)), dim3(256), 0, s, N, ns, G, large, ctl);
  hipLaunchKernelGGL(plan_small_info_kernel, dim3(grid_for(N, 256)), dim3(256), 0, s, N, ns, pid, G,
                     p->d_net_maxpid, p->d_net_chain, ctl);
  const int ncu = std::max(1, ctx->num_cus);
  hipLaunchKernelGGL(plan_large_info_kernel, dim3(ncu * 16), dim3(64), 0, s, ns, pid, large,
                     p->d_net_maxpid, p->d_net_chain, ctl);
  const int nseg = (N + kSeg - 1) / kSeg;
  int32_t* seg_groups = nullptr;
  int64_t* seg_words = nullptr;
  OTG_HIP(ctx, T.alloc(&seg_groups, nseg));
  OTG_HIP(ctx, T.alloc(&seg_words, nseg));
  hipLaunchKernelGGL(plan_pack_kernel<false>, dim3(grid_for(nseg, 256)), dim3(256), 0, s, N, ns, G,
                     cap, p->tuning.small_target, seg_groups, seg_words, nullptr, nullptr, nullptr,
                     nullptr);
  OTG_HIP(ctx, hipGetLastError());
  PlanCtl hc{};
  std::vector<int32_t> hsg(nseg);
  std::vector<int64_t> hsw(nseg);
  int32_t ends[2] = {0, 0};
  OTG_HIP(ctx, hipMemcpyAsync(&hc, ctl, sizeof(hc), hipMemcpyDeviceToHost, s));
  OTG_HIP(ctx, hipMemcpyAsync(hsg.data(), seg_groups, sizeof(int32_t) * nseg, hipMemcpyDeviceToHost, s));
  OTG_HIP(ctx, hipMemcpyAsync(hsw.data(), seg_words, sizeof(int64_t) * nseg, hipMemcpyDeviceToHost, s));
  OTG_HIP(ctx, hipMemcpyAsync(&ends[0], ns, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  OTG_HIP(ctx, hipMemcpyAsync(&ends[1], ns + N, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  OTG_HIP(ctx, hipStreamSynchronize(s));
  if (hc.start_bad || ends[0] != 0 || ends[1] != V) {
    OTG_SET_ERR(ctx, "nodes_start must run from 0 to total_num_nodes without decreasing");
    return OTG_ERR_INVALID_ARGUMENT;
  }
  if (hc.bad != INT32_MAX) {
    OTG_SET_ERR(ctx, "net %d: pid must satisfy 1 <= pid[i] <= i - nodes_start[net]", hc.bad);
    return OTG_ERR_BAD_TOPOLOGY;
  }
  // ---- groups: offsets per segment, then the groups and their schedules ----
  std::vector<int32_t> goff(nseg);
  std::vector<int64_t> woff(nseg);
  int64_t ng = 0, nw = 0;
  for (int i = 0; i < nseg; ++i) {
    goff[i] = (int32_t)ng;
    woff[i] = nw;
    ng += hsg[i];
    nw += hsw[i];
  }
  p->num_groups = (int32_t)ng;
  OTG_HIP(ctx, hipMalloc(&p->d_net_group, sizeof(int32_t) * N));
  hipLaunchKernelGGL(plan_fill_kernel, dim3(grid_for(N, 256)), dim3(256), 0, s, p->d_net_group,
                     (int64_t)N, -1);
  if (ng > 0) {
    int32_t* dgoff = nullptr;
    int64_t* dwoff = nullptr;
    RctItem* unsorted = nullptr;
    int32_t* tmp_group = nullptr;
    uint32_t *cost = nullptr, *cost2 = nullptr, *idx = nullptr, *idx2 = nullptr;
    OTG_HIP(ctx, T.alloc(&dgoff, nseg));
    OTG_HIP(ctx, T.alloc(&dwoff, nseg));
    OTG_HIP(ctx, T.alloc(&unsorted, ng));
    OTG_HIP(ctx, T.alloc(&tmp_group, N));
    OTG_HIP(ctx, T.alloc(&cost, ng));
    OTG_HIP(ctx, T.alloc(&cost2, ng));
    OTG_HIP(ctx, T.alloc(&idx, ng));
    OTG_HIP(ctx, T.alloc(&idx2, ng));
    OTG_HIP(ctx, hipMemcpyAsync(dgoff, goff.data(), sizeof(int32_t) * nseg, hipMemcpyHostToDevice, s));
    OTG_HIP(ctx, hipMemcpyAsync(dwoff, woff.data(), sizeof(int64_t) * nseg, hipMemcpyHostToDevice, s));
    OTG_HIP(ctx, hipMalloc(&p->d_sched, sizeof(uint32_t) * std::max<int64_t>(1, nw)));
    OTG_HIP(ctx, hipMalloc(&p->d_groups, sizeof(RctItem) * ng));
    OTG_HIP(ctx, hipMalloc(&p->d_sub_groups, sizeof(RctItem) * ng));
    OTG_HIP(ctx, hipMalloc(&p->d_group_flag, sizeof(int32_t) * ng));
    OTG_HIP(ctx, hipMalloc(&p->d_sub_count, sizeof(int32_t) * 2));
    hipLaunchKernelGGL(plan_pack_kernel<true>, dim3(grid_for(nseg, 256)), dim3(256), 0, s, N, ns,
                       G, cap, p->tuning.small_target, nullptr, nullptr, dgoff, dwoff, unsorted,
                       tmp_group);
    hipLaunchKernelGGL(plan_sched_kernel, dim3((int)std::min<int64_t>(ng, (int64_t)ncu * 32)),
                       dim3(64), 0, s, (int32_t)ng, unsorted, ns, pid, p->d_sched, cost);
    // stable order by cost, most steps first (creation order among equals)
    std::vector<uint32_t> iota_h(ng);
    std::iota(iota_h.begin(), iota_h.end(), 0u);
    OTG_HIP(ctx, hipMemcpyAsync(idx, iota_h.data(), sizeof(uint32_t) * ng, hipMemcpyHostToDevice, s));
    size_t tb = 0;
    OTG_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, tb, cost, cost2, idx, idx2, (int)ng, 0, 8, s));
    void* tsort = nullptr;
    OTG_HIP(ctx, T.alloc(reinterpret_cast<char**>(&tsort), (int64_t)tb));
    OTG_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(tsort, tb, cost, cost2, idx, idx2, (int)ng, 0, 8, s));
    hipLaunchKernelGGL(plan_permute_kernel, dim3(grid_for(ng, 256)), dim3(256), 0, s, (int32_t)ng,
                       idx2, unsorted, p->d_groups, p->d_net_group);
    hipLaunchKernelGGL(plan_fill_kernel, dim3(grid_for(ng, 256)), dim3(256), 0, s, p->d_group_flag,
                       ng, 0);
    OTG_HIP(ctx, hipGetLastError());
    // (the segments' group indices in tmp_group are superseded by the permutation)
    (void)tmp_group;
  }
  // ---- streamed nets: records to the host, binned and ordered there ----
  const int nl = hc.num_large;
  std::vector<int4> rec(nl);
  if (nl > 0) {
    int4* drec = nullptr;
    OTG_HIP(ctx, T.alloc(&drec, nl));
    hipLaunchKernelGGL(plan_large_gather_kernel, dim3(grid_for(nl, 256)), dim3(256), 0, s, nl, large,
                       ns, p->d_net_maxpid, p->d_net_chain, drec);
    OTG_HIP(ctx, hipMemcpyAsync(rec.data(), drec, sizeof(int4) * nl, hipMemcpyDeviceToHost, s));
  }
  hipLaunchKernelGGL(plan_fill64_kernel, dim3(grid_for(N, 256)), dim3(256), 0, s, p->d_net_scratch,
                     (int64_t)N, (int64_t)-1);
  OTG_HIP(ctx, hipStreamSynchronize(s));
  std::sort(rec.begin(), rec.end(), [](const int4& a, const int4& b) { return a.x < b.x; });
  std::vector<RctItem> bin_list[otg_rct_plan::kMaxBins], gitems;
  std::vector<int64_t> pairs;
  pairs.reserve(2 * (size_t)nl);
  const int nb = p->num_bins;
  const int64_t slot_floats = comp ? 16 : 8;
  int64_t scratch = 0, n_stream = 0, n_global = 0;
  for (int i = 0; i < nl; ++i) {
    const int k = rec[i].x, st = rec[i].y, n = rec[i].z, mp = rec[i].w >> 1, ch = rec[i].w & 1;
    RctItem it{st, st + n, k, k + 1, ITEM_STREAM, 1, scratch};
    const int want_chain = comp && ch;
    int bin = 0;
    while (bin < nb && (mp > p->bin_ring[bin] - kWave || p->bin_chain[bin] != want_chain)) ++bin;
    if (bin < nb) {
      scratch += 4LL * n;
      ++n_stream;
      bin_list[bin].push_back(it);
    } else {
      it.kind = ITEM_GLOBAL;
      scratch += slot_floats * n;
      ++n_global;
      gitems.push_back(it);
    }
    pairs.push_back(k);
    pairs.push_back(it.scratch);
  }
  auto larger = [](const RctItem& a, const RctItem& b) {
    return (a.pos_end - a.pos_begin) > (b.pos_end - b.pos_begin);
  };
  int64_t n_items = 0;
  for (int b = 0; b < nb; ++b) {
    std::stable_sort(bin_list[b].begin(), bin_list[b].end(), larger);
    p->bin_items[b] = (int32_t)bin_list[b].size();
    n_items += p->bin_items[b];
    const size_t m = bin_list[b].size();
    if (!m) continue;
    OTG_HIP(ctx, hipMalloc(&p->d_bin_items[b], sizeof(RctItem) * m));
    OTG_HIP(ctx, hipMemcpyAsync(p->d_bin_items[b], bin_list[b].data(), sizeof(RctItem) * m,
                                hipMemcpyHostToDevice, s));
    OTG_HIP(ctx, hipMalloc(&p->d_bin_sub[b], sizeof(RctItem) * m));
  }
  std::stable_sort(gitems.begin(), gitems.end(), larger);
  p->num_gitems = (int32_t)gitems.size();
  if (!gitems.empty()) {
    OTG_HIP(ctx, hipMalloc(&p->d_gitems, sizeof(RctItem) * gitems.size()));
    OTG_HIP(ctx, hipMemcpyAsync(p->d_gitems, gitems.data(), sizeof(RctItem) * gitems.size(),
                                hipMemcpyHostToDevice, s));
  }
  if (nl > 0) {
    int64_t* dpairs = nullptr;
    OTG_HIP(ctx, T.alloc(&dpairs, 2LL * nl));
    OTG_HIP(ctx, hipMemcpyAsync(dpairs, pairs.data(), sizeof(int64_t) * 2 * nl, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(plan_scatter_scratch_kernel, dim3(grid_for(nl, 256)), dim3(256), 0, s, nl,
                       dpairs, p->d_net_scratch);
  }
  p->scratch_floats = scratch;
  if (scratch > 0) OTG_HIP(ctx, hipMalloc(&p->d_scratch, sizeof(float) * scratch));
  p->stats[0] = ng;
  p->stats[1] = hc.packed_nets;
  p->stats[2] = n_stream;
  p->stats[3] = n_global;
  p->stats[4] = hc.max_net;
  p->stats[5] = n_items + p->num_gitems + ng;
  p->stats[6] = scratch;
  p->stats[7] = W;
  OTG_HIP(ctx, hipGetLastError());
  OTG_HIP(ctx, hipStreamSynchronize(s));
  return OTG_OK;
}
