// P2: sink-first levelization of the timing graph.
//
// Semantics of toposort_advance / toposort_compute_cuda (reference
// ot/cuda/toposort.cu:4-72) as driven by Timer::_build_prop_tasks_cuda
// (ot/timer/timer.cpp:1053-1071): frontier 0 is the set of pins without
// (kept) fanout; for every pin u of the current frontier and every fanin v
// of u, the fanout count of v is decremented and v joins the next frontier
// when it reaches zero.  The level of a pin is the index of its frontier,
// i.e. the longest path (in arcs) from the pin to a sink.
//
// The reference launches one kernel per level and, after each, copies the new
// frontier size to the host and a zero back (two synchronous copies per
// level).  Here the levels are launched back-to-back with no host round trip:
// the kernel for level l takes l as an argument, reads where frontier l
// starts from ends[l] (written by the kernel of level l-1), derives where it
// ends from the append counter of level l-1, and appends frontier l+1 with
// one atomic per wavefront and fanin step (ballot compaction).  Kernel
// boundaries give the cross-XCD visibility a grid barrier would need fences
// for.  The host launches levels in batches and reads one flag per batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "otg_internal.hpp"

// Working memory of otg_levelize_noalloc: the consumed fanout counts [n], the
// per-level append counters [n+2] and the end flag [1] of one levelization,
// for at most OTG_LEVELIZE_NOALLOC_MAX_PINS pins.  Each GPU holds its own copy
// for the life of the process; host code reaches it only through the address
// hipGetSymbolAddress returns on that device (the array's name is not a device
// address in host code).
constexpr int64_t kNoAllocInts = 2LL * OTG_LEVELIZE_NOALLOC_MAX_PINS + 8;
static __device__ __align__(256) int32_t scratch[kNoAllocInts];

namespace {

constexpr int kBlock = 256;
constexpr int kBatch = 32;  // level kernels per host check

struct LevelizeParams {
  int32_t n;
  int32_t first_size;
  const int32_t* fanin_start;  // CSR walked from each frontier pin (fanin; fanout for reach)
  const int32_t* fanin;
  int32_t* out;        // Kahn: consumed fanout counts (device copy) | reach: visited flags
  int32_t* frontiers;  // [n], frontier 0 given
  int32_t* ends;       // [levels+2]
  int32_t* level;      // [n] or null
  int32_t* cnt;        // [levels+2] appends made while processing level l
  int32_t* num_ends;   // [1], INT_MAX until the first empty frontier
  const int32_t* first_dev;  // size of frontier 0 on the device (overrides first_size), or null
  const int32_t* cand;       // Kahn: arcs from pins with cand == 0 are skipped; null = all
};

// Kahn: a fanin source joins the next frontier when its last fanout is
// consumed.  Reach: a fanout joins the next frontier the first time it is
// visited (forward closure of frontier 0, the fan-out cone).
enum : int { LV_KAHN = 0, LV_REACH = 1 };

__global__ __launch_bounds__(kBlock) void levelize_init_kernel(LevelizeParams P) {
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gsz = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = gtid; i < (int64_t)P.n + 2; i += gsz) P.cnt[i] = 0;
  if (P.level)
    for (int64_t i = gtid; i < P.n; i += gsz) P.level[i] = -1;
  if (gtid == 0) {
    P.ends[0] = 0;
    P.ends[1] = P.first_size;
    *P.num_ends = P.first_size == 0 ? 2 : INT32_MAX;
  }
}

__global__ __launch_bounds__(kBlock) void levelize_seed_kernel(LevelizeParams P) {
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gsz = (int64_t)gridDim.x * blockDim.x;
  if (P.level)
    for (int64_t i = gtid; i < P.first_size; i += gsz) P.level[P.frontiers[i]] = 0;
}

// Arcs of the frontier are spread evenly over the lanes: a wavefront takes
// 64 frontier pins, scans their fanin counts, and each lane handles arcs
// lane, lane+64, ... of the chunk; up to kArcs arcs per lane are fetched and
// decremented before any result is used, so their latencies overlap.
constexpr int kArcs = 4;

__device__ __forceinline__ int owner_of(int t, int off_mine, int lane_count) {
  // largest lane l with off[l] <= t, off[] exclusive scan held one per lane
  int lo = 0;
#pragma unroll
  for (int step = 32; step > 0; step >>= 1) {
    const int cand = lo + step;
    const int o = __shfl(off_mine, cand < lane_count ? cand : 63);
    if (cand < lane_count && o <= t) lo = cand;
  }
  return lo;
}

// Appends to the next frontier are gathered per workgroup: a wavefront takes
// slots in an LDS staging buffer with an LDS atomic, and after the level's
// work the workgroup reserves its whole run of the frontier with one global
// atomic and copies the buffer out.  (All wavefronts of the grid append to
// one counter; with one global atomic per wavefront and fanin step, those
// same-address atomics serialise and took about half of a level's time.)
// A wavefront whose appends do not fit the buffer appends the rest directly.
constexpr int kStage = 2048;

template <int MODE, int BS>
__global__ __launch_bounds__(BS) void levelize_level_kernel(LevelizeParams P, int lv) {
  __shared__ int32_t s_buf[kStage];
  __shared__ int s_cnt, s_base;
  if (lv > P.n) return;
  const int b = P.ends[lv];
  const int e = lv == 0 ? (P.first_dev ? *P.first_dev : P.first_size) : b + P.cnt[lv - 1];
  if (e == b) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicMin(P.num_ends, lv < 1 ? 2 : lv + 1);
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) P.ends[lv + 1] = e;
  const int last = e - b;
  // (the return above is uniform over the workgroup: b and e are)
  if ((int64_t)blockIdx.x * BS >= last) return;  // no chunk for this workgroup
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  int32_t* next = P.frontiers + e;
  int32_t* ctr = P.cnt + lv;
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  for (int64_t c0 = (int64_t)wave * 64; c0 < last; c0 += (int64_t)nwaves * 64) {
    const int cnt_pins = (int)min<int64_t>(64, last - c0);
    const int u = lane < cnt_pins ? P.frontiers[b + c0 + lane] : -1;
    const int js = u >= 0 ? P.fanin_start[u] : 0;
    const int deg = u >= 0 ? P.fanin_start[u + 1] - js : 0;
    // exclusive scan of the fanin counts over the wave
    int inc = deg;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int t = __shfl_up(inc, d);
      if (lane >= d) inc += t;
    }
    const int off = inc - deg;
    const int total = __shfl(inc, 63);
    for (int t0 = 0; t0 < total; t0 += 64 * kArcs) {
      int v[kArcs];
      int old[kArcs];
#pragma unroll
      for (int k = 0; k < kArcs; ++k) {
        // every lane takes part in the shuffles (a shuffle from an inactive
        // lane reads nothing useful); only in-range arcs are loaded
        const int t = t0 + k * 64 + lane;
        const bool ok = t < total;
        const int tt = ok ? t : 0;
        const int o = owner_of(tt, off, cnt_pins);
        const int oj = __shfl(js, o), oo = __shfl(off, o);
        v[k] = ok ? P.fanin[oj + (tt - oo)] : -1;
      }
      if (MODE == LV_KAHN && P.cand) {
#pragma unroll
        for (int k = 0; k < kArcs; ++k)
          if (v[k] >= 0 && !P.cand[v[k]]) v[k] = -1;
      }
#pragma unroll
      for (int k = 0; k < kArcs; ++k) {
        if (MODE == LV_KAHN) old[k] = v[k] >= 0 ? atomicSub(&P.out[v[k]], 1) : 0;
        else old[k] = v[k] >= 0 ? (atomicExch(&P.out[v[k]], 1) == 0 ? 1 : 0) : 0;
      }
      uint64_t m[kArcs];
      int tot = 0;
#pragma unroll
      for (int k = 0; k < kArcs; ++k) {
        m[k] = __ballot(old[k] == 1);
        tot += __popcll(m[k]);
      }
      if (tot) {
        // slots [off, off + tot) of the staging buffer; those past its end
        // go straight to the frontier
        int so = 0;
        if (lane == 0) so = atomicAdd(&s_cnt, tot);
        so = __shfl(so, 0);
        const int spill = so + tot - max(so, kStage);  // appends past the buffer
        int gb = 0;
        if (spill > 0 && lane == 0) gb = atomicAdd(ctr, spill);
        gb = __shfl(gb, 0) - max(so, kStage);
        int r = so;
#pragma unroll
        for (int k = 0; k < kArcs; ++k) {
          if (old[k] == 1) {
            const int slot = r + __popcll(m[k] & ((1ull << lane) - 1ull));
            if (slot < kStage) {
              s_buf[slot] = v[k];
            } else {
              next[gb + slot] = v[k];
              if (P.level) P.level[v[k]] = lv + 1;
            }
          }
          r += __popcll(m[k]);
        }
      }
    }
  }
  __syncthreads();
  const int ns = min(s_cnt, kStage);
  if (ns == 0) return;  // uniform: every thread reads the same count
  if (threadIdx.x == 0) s_base = atomicAdd(ctr, ns);
  __syncthreads();
  const int base = s_base;
  for (int i = threadIdx.x; i < ns; i += BS) {
    const int v = s_buf[i];
    next[base + i] = v;
    if (P.level) P.level[v] = lv + 1;
  }
}


// ---------------------------------------------------------------------------
// Incremental re-levelization of the fan-out cone of changed nets.
//
// The cone of a set of frontier pins is their forward closure over fanout
// arcs (the FPROP_CAND set of Timer::_build_fprop_cands, reference
// ot/timer/timer.cpp:845-923); only it is re-levelized, with the sink-first
// Kahn sweep restricted to cone pins (the candidate filter of
// Timer::_build_prop_tasks_cuda, timer.cpp:1011-1047).  A cone is forward
// closed, so a cone pin's fanout count within the cone is its full fanout
// count.  Every launch has a fixed grid and the number of level launches is
// fixed by the caller's bound on the graph depth, so one update is a fixed
// sequence of kernels with no host round trip (capturable in a hipGraph).
// ---------------------------------------------------------------------------
struct IncrState {
  int32_t n;
  int32_t num_nets;
  int32_t max_levels;
  const int32_t* net_driver;  // [num_nets] driver pin of each net
  int32_t* flag;              // [n] cone membership, all zero between updates
  int32_t* cone;              // [n] cone pins, frontier by frontier
  int32_t* cone_ends;         // [max_levels+3]
  int32_t* cone_cnt;          // [max_levels+3]
  int32_t* kout;              // [n] fanout counts consumed by the Kahn sweep
  int32_t* kfront;            // [n] cone pins by level
  int32_t* kends;             // [max_levels+3]
  int32_t* kcnt;              // [max_levels+3]
  int32_t* counters;          // [0] seeds [1] sinks [2] cone num_ends [3] kahn num_ends
};

__global__ __launch_bounds__(kBlock) void incr_reset_kernel(IncrState S) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  for (int k = t; k < S.max_levels + 3; k += gridDim.x * blockDim.x) {
    S.cone_cnt[k] = 0;
    S.kcnt[k] = 0;
    S.cone_ends[k] = 0;
    S.kends[k] = 0;
  }
  if (t == 0) {
    S.counters[0] = 0;
    S.counters[1] = 0;
    S.counters[2] = INT32_MAX;
    S.counters[3] = INT32_MAX;
  }
}

// Frontier 0 of the cone: the driver pins of the changed nets, each once.
__global__ __launch_bounds__(kBlock) void incr_seed_kernel(IncrState S, const int32_t* net_ids,
                                                           int32_t count) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  const int net = net_ids[j];
  if (net < 0 || net >= S.num_nets) return;
  const int pin = S.net_driver[net];
  if (pin < 0 || pin >= S.n) return;
  if (atomicExch(&S.flag[pin], 1) == 0) S.cone[atomicAdd(&S.counters[0], 1)] = pin;
}

__device__ __forceinline__ int incr_cone_size(const IncrState& S) {
  const int ne = S.counters[2];
  const int last = ne == INT32_MAX ? S.max_levels + 1 : ne - 1;
  return S.cone_ends[last < S.max_levels + 2 ? last : S.max_levels + 2];
}

// Cone pins: Kahn fanout counts, and frontier 0 of the Kahn sweep (cone pins
// without fanout).
__global__ __launch_bounds__(kBlock) void incr_kahn_init_kernel(IncrState S,
                                                                const int32_t* out_degree,
                                                                int32_t* level) {
  const int size = incr_cone_size(S);
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < size; k += gridDim.x * blockDim.x) {
    const int u = S.cone[k];
    const int d = out_degree[u];
    S.kout[u] = d;
    if (level) level[u] = d == 0 ? 0 : -1;
    if (d == 0) S.kfront[atomicAdd(&S.counters[1], 1)] = u;
  }
}

// Clears the cone flags for the next update.
__global__ __launch_bounds__(kBlock) void incr_cleanup_kernel(IncrState S) {
  const int size = incr_cone_size(S);
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < size; k += gridDim.x * blockDim.x)
    S.flag[S.cone[k]] = 0;
}
}  // namespace

// Launches the whole sweep on `s` and returns the number of ends entries in
// *num_ends_host (synchronises s once per batch of levels).
otg_status otg_levelize_launch_raw(otg_context* ctx, int32_t n, int32_t first_size,
                                   const int32_t* fanin_start, const int32_t* fanin,
                                   int32_t* out_copy, int32_t* frontiers, int32_t* ends,
                                   int32_t* level, int32_t* scratch, int32_t* num_ends_host,
                                   hipStream_t s, bool pipelined) {
  // scratch: [n+2] per-level append counters followed by num_ends
  LevelizeParams P{n, first_size, fanin_start, fanin, out_copy, frontiers, ends, level,
                   scratch, scratch + n + 2, nullptr, nullptr};
  const int init_grid = (int)std::min<int64_t>(4096, ((int64_t)n + 2 + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(levelize_init_kernel, dim3(std::max(1, init_grid)), dim3(kBlock), 0, s, P);
  OTG_HIP(ctx, hipGetLastError());
  if (first_size > 0) {
    const int g = (int)std::min<int64_t>(4096, ((int64_t)first_size + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(levelize_seed_kernel, dim3(std::max(1, g)), dim3(kBlock), 0, s, P);
    OTG_HIP(ctx, hipGetLastError());
  }
  // about one thread per pin on the wide levels, never more than the pins
  const int grid = std::max(1, std::min(4 * std::max(1, ctx->num_cus),
                                        (int)(((int64_t)n + kBlock - 1) / kBlock)));
  // The end flag is checked once per batch of level launches, one batch
  // behind: batch k+1 is queued before the host waits for batch k's flag
  // (copied into pinned memory), so the GPU never idles at a check; the
  // batch queued past the end runs as empty launches.
  // Without `pipelined` (the entry point that allocates nothing: the pinned
  // flags are allocated on first use) each batch is checked synchronously.
  if (pipelined && !ctx->lv_flag) {
    OTG_HIP(ctx, hipHostMalloc((void**)&ctx->lv_flag, 2 * sizeof(int32_t), hipHostMallocDefault));
    for (int k = 0; k < 2; ++k)
      OTG_HIP(ctx, hipEventCreateWithFlags(&ctx->lv_ev[k], hipEventDisableTiming));
  }
  int32_t ne = INT32_MAX;
  if (!pipelined) {
    for (int lv0 = 0; lv0 <= n && ne == INT32_MAX; lv0 += kBatch) {
      for (int lv = lv0; lv < lv0 + kBatch && lv <= n; ++lv)
        hipLaunchKernelGGL((levelize_level_kernel<LV_KAHN, kBlock>), dim3(grid), dim3(kBlock), 0,
                           s, P, lv);
      OTG_HIP(ctx, hipGetLastError());
      OTG_HIP(ctx, hipMemcpyAsync(&ne, P.num_ends, sizeof(int32_t), hipMemcpyDeviceToHost, s));
      OTG_HIP(ctx, hipStreamSynchronize(s));
    }
    if (ne == INT32_MAX) {
      OTG_SET_ERR(ctx, "levelization did not terminate within n levels");
      return OTG_ERR_CYCLE;
    }
    *num_ends_host = ne;
    return OTG_OK;
  }
  int pending = -1;  // slot of the batch whose flag is in flight
  for (int lv0 = 0, slot = 0;; lv0 += kBatch, slot ^= 1) {
    if (lv0 <= n) {
      for (int lv = lv0; lv < lv0 + kBatch && lv <= n; ++lv)
        hipLaunchKernelGGL((levelize_level_kernel<LV_KAHN, kBlock>), dim3(grid), dim3(kBlock), 0,
                           s, P, lv);
      OTG_HIP(ctx, hipGetLastError());
      OTG_HIP(ctx, hipMemcpyAsync(ctx->lv_flag + slot, P.num_ends, sizeof(int32_t),
                                  hipMemcpyDeviceToHost, s));
      OTG_HIP(ctx, hipEventRecord(ctx->lv_ev[slot], s));
    }
    if (pending >= 0) {
      OTG_HIP(ctx, hipEventSynchronize(ctx->lv_ev[pending]));
      ne = ctx->lv_flag[pending];
      if (ne != INT32_MAX) break;
    }
    if (lv0 > n) break;
    pending = slot;
  }
  if (ne == INT32_MAX) {
    // the last batch's flag (n+1 levels launched)
    OTG_HIP(ctx, hipStreamSynchronize(s));
    OTG_HIP(ctx, hipMemcpy(&ne, P.num_ends, sizeof(int32_t), hipMemcpyDeviceToHost));
  }
  if (ne == INT32_MAX) {
    OTG_SET_ERR(ctx, "levelization did not terminate within n levels");
    return OTG_ERR_CYCLE;
  }
  *num_ends_host = ne;
  return OTG_OK;
}

// ---------------------------------------------------------------------------
// Levelization without allocation: otg_levelize on the library's fixed
// per-device array.
// ---------------------------------------------------------------------------
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
int32_t* g_noalloc_ptr[kMaxDevices] = {};

// Device address of `scratch` on ctx's device, looked up once per device with
// that device current.
otg_status noalloc_scratch(otg_context* ctx, int32_t** out) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices) {
    OTG_SET_ERR(ctx, "device %d out of range for the fixed levelization array", ctx->device);
    return OTG_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> g(g_noalloc_mu);
  if (!g_noalloc_ptr[ctx->device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      OTG_SET_ERR(ctx, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      return OTG_ERR_DEVICE;
    }
    g_noalloc_ptr[ctx->device] = static_cast<int32_t*>(p);
  }
  *out = g_noalloc_ptr[ctx->device];
  return OTG_OK;
}
}  // namespace

extern "C" otg_status otg_levelize_noalloc(otg_context* ctx, int32_t n, int32_t num_edges,
                                           const int32_t* fanin_start, const int32_t* fanin,
                                           const int32_t* out_degree,
                                           const int32_t* first_frontier, int32_t first_size,
                                           int32_t* frontiers, int32_t* frontiers_ends,
                                           int32_t* level_of_pin, int32_t* num_ends,
                                           otg_stream stream) {
  if (!ctx || n < 0 || num_edges < 0 || first_size < 0 || first_size > n || !num_ends)
    return OTG_ERR_INVALID_ARGUMENT;
  if (n > OTG_LEVELIZE_NOALLOC_MAX_PINS) {
    OTG_SET_ERR(ctx, "otg_levelize_noalloc: %d pins exceed the fixed array's %d", n,
                OTG_LEVELIZE_NOALLOC_MAX_PINS);
    return OTG_ERR_INVALID_ARGUMENT;
  }
  if (n > 0 && (!fanin_start || !out_degree || !frontiers || !frontiers_ends ||
                (num_edges > 0 && !fanin) || (first_size > 0 && !first_frontier)))
    return OTG_ERR_INVALID_ARGUMENT;
  hipStream_t s = otg_pick_stream(ctx, stream);
  OTG_HIP(ctx, hipSetDevice(ctx->device));
  if (n == 0) {
    // the same answer as otg_levelize (ends = {0, 0}), written by a kernel
    if (frontiers_ends) {
      int32_t* dev = nullptr;
      otg_status st0 = noalloc_scratch(ctx, &dev);
      if (st0 != OTG_OK) return st0;
      LevelizeParams P{};
      P.ends = frontiers_ends;
      P.cnt = dev;
      P.num_ends = dev + 2;
      hipLaunchKernelGGL(levelize_init_kernel, dim3(1), dim3(kBlock), 0, s, P);
      OTG_HIP(ctx, hipGetLastError());
    }
    OTG_HIP(ctx, hipStreamSynchronize(s));
    *num_ends = 2;
    return OTG_OK;
  }
  int32_t* dev = nullptr;
  otg_status st = noalloc_scratch(ctx, &dev);
  if (st != OTG_OK) return st;
  int32_t* out_copy = dev;        // [n]
  int32_t* counters = dev + n;    // [n+3]: append counters, then the end flag
  OTG_HIP(ctx, hipMemcpyAsync(out_copy, out_degree, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
  if (first_size > 0 && first_frontier != frontiers)
    OTG_HIP(ctx, hipMemcpyAsync(frontiers, first_frontier, sizeof(int32_t) * first_size,
                                hipMemcpyDeviceToDevice, s));
  int32_t ne = 0;
  st = otg_levelize_launch_raw(ctx, n, first_size, fanin_start, fanin, out_copy, frontiers,
                               frontiers_ends, level_of_pin, counters, &ne, s, false);
  if (st != OTG_OK) return st;
  *num_ends = ne;
  return OTG_OK;
}

struct otg_incr {
  otg_context* ctx = nullptr;
  IncrState S{};
  const int32_t* fo_start = nullptr;
  const int32_t* fo = nullptr;
  const int32_t* fi_start = nullptr;
  const int32_t* fi = nullptr;
  const int32_t* out_degree = nullptr;
  int32_t* level = nullptr;
  int grid = 1;
};

extern "C" otg_status otg_incr_create(otg_context* ctx, const otg_incr_graph* g, int32_t max_levels,
                                      otg_incr** out) {
  if (!ctx || !g || !out || g->num_pins < 0 || g->num_nets < 0 || max_levels < 0)
    return OTG_ERR_INVALID_ARGUMENT;
  if (g->num_pins > 0 && (!g->fanout_start || !g->fanout || !g->fanin_start || !g->fanin ||
                          !g->out_degree || !g->level))
    return OTG_ERR_INVALID_ARGUMENT;
  if (g->num_nets > 0 && !g->net_driver) return OTG_ERR_INVALID_ARGUMENT;
  OTG_HIP(ctx, hipSetDevice(ctx->device));
  otg_incr* I = new otg_incr();
  I->ctx = ctx;
  I->fo_start = g->fanout_start;
  I->fo = g->fanout;
  I->fi_start = g->fanin_start;
  I->fi = g->fanin;
  I->out_degree = g->out_degree;
  I->level = g->level;
  IncrState& S = I->S;
  S.n = g->num_pins;
  S.num_nets = g->num_nets;
  S.max_levels = max_levels;
  S.net_driver = g->net_driver;
  const int64_t n = std::max<int64_t>(1, g->num_pins), L = (int64_t)max_levels + 3;
  auto fail = [&](hipError_t e) {
    otg_incr_destroy(I);
    otg_set_error(ctx, "otg_incr_create: %s", hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? OTG_ERR_OUT_OF_MEMORY : OTG_ERR_DEVICE;
  };
  hipError_t e;
  if ((e = hipMalloc(&S.flag, sizeof(int32_t) * n)) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&S.cone, sizeof(int32_t) * n)) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&S.kout, sizeof(int32_t) * n)) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&S.kfront, sizeof(int32_t) * n)) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&S.cone_ends, sizeof(int32_t) * L)) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&S.cone_cnt, sizeof(int32_t) * L)) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&S.kends, sizeof(int32_t) * L)) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&S.kcnt, sizeof(int32_t) * L)) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&S.counters, sizeof(int32_t) * 8)) != hipSuccess) return fail(e);
  if ((e = hipMemset(S.flag, 0, sizeof(int32_t) * n)) != hipSuccess) return fail(e);
  if ((e = hipMemset(S.counters, 0, sizeof(int32_t) * 8)) != hipSuccess) return fail(e);
  if ((e = hipDeviceSynchronize()) != hipSuccess) return fail(e);
  I->grid = std::max(1, std::min(4 * std::max(1, ctx->num_cus),
                                 (int)((g->num_pins + kBlock - 1) / kBlock)));
  *out = I;
  return OTG_OK;
}

extern "C" otg_status otg_incr_destroy(otg_incr* I) {
  if (!I) return OTG_OK;
  IncrState& S = I->S;
  for (int32_t* p : {S.flag, S.cone, S.kout, S.kfront, S.cone_ends, S.cone_cnt, S.kends, S.kcnt,
                     S.counters})
    if (p) (void)hipFree(p);
  delete I;
  return OTG_OK;
}

extern "C" otg_status otg_incr_relevelize(otg_incr* I, const int32_t* net_ids, int32_t count,
                                          otg_stream stream) {
  if (!I || count < 0 || (count > 0 && !net_ids)) return OTG_ERR_INVALID_ARGUMENT;
  otg_context* ctx = I->ctx;
  hipStream_t s = otg_pick_stream(ctx, stream);
  const IncrState& S = I->S;
  if (S.n == 0) return OTG_OK;
  const int L = S.max_levels + 1;  // frontiers 0..max_levels, then one empty
  hipLaunchKernelGGL(incr_reset_kernel, dim3(1), dim3(kBlock), 0, s, S);
  if (count > 0)
    hipLaunchKernelGGL(incr_seed_kernel, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, s, S,
                       net_ids, count);
  // forward closure: frontiers of the cone over the fanout CSR
  LevelizeParams R{S.n, 0, I->fo_start, I->fo, S.flag, S.cone, S.cone_ends, nullptr,
                   S.cone_cnt, S.counters + 2, S.counters + 0, nullptr};
  for (int lv = 0; lv <= L; ++lv)
    hipLaunchKernelGGL((levelize_level_kernel<LV_REACH, kBlock>), dim3(I->grid), dim3(kBlock), 0, s,
                       R, lv);
  // sink-first Kahn sweep over the cone
  hipLaunchKernelGGL(incr_kahn_init_kernel, dim3(I->grid), dim3(kBlock), 0, s, S, I->out_degree,
                     I->level);
  LevelizeParams K{S.n, 0, I->fi_start, I->fi, S.kout, S.kfront, S.kends, I->level,
                   S.kcnt, S.counters + 3, S.counters + 1, S.flag};
  for (int lv = 0; lv <= L; ++lv)
    hipLaunchKernelGGL((levelize_level_kernel<LV_KAHN, kBlock>), dim3(I->grid), dim3(kBlock), 0, s,
                       K, lv);
  hipLaunchKernelGGL(incr_cleanup_kernel, dim3(I->grid), dim3(kBlock), 0, s, S);
  OTG_HIP(ctx, hipGetLastError());
  return OTG_OK;
}

extern "C" otg_status otg_incr_query(otg_incr* I, otg_stream stream, int32_t* cone_size,
                                     int32_t* num_levels, const int32_t** cone,
                                     const int32_t** frontiers, const int32_t** frontier_ends) {
  if (!I) return OTG_ERR_INVALID_ARGUMENT;
  otg_context* ctx = I->ctx;
  hipStream_t s = otg_pick_stream(ctx, stream);
  const IncrState& S = I->S;
  int32_t c[4] = {0, 0, 0, 0};
  OTG_HIP(ctx, hipMemcpyAsync(c, S.counters, sizeof(c), hipMemcpyDeviceToHost, s));
  OTG_HIP(ctx, hipStreamSynchronize(s));
  if (c[2] == INT32_MAX || c[3] == INT32_MAX) {
    OTG_SET_ERR(ctx, "cone deeper than max_levels (or a cycle in the cone)");
    return OTG_ERR_CYCLE;
  }
  int32_t csize = 0;
  OTG_HIP(ctx, hipMemcpyAsync(&csize, S.cone_ends + (c[2] - 1), sizeof(int32_t),
                              hipMemcpyDeviceToHost, s));
  OTG_HIP(ctx, hipStreamSynchronize(s));
  if (cone_size) *cone_size = csize;
  if (num_levels) *num_levels = c[3] - 1;  // frontier_ends has num_levels+1 entries
  if (cone) *cone = S.cone;
  if (frontiers) *frontiers = S.kfront;
  if (frontier_ends) *frontier_ends = S.kends;
  return OTG_OK;
}

// ---------------------------------------------------------------------------
// hipGraph capture helpers
// ---------------------------------------------------------------------------
struct otg_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" otg_status otg_graph_begin(otg_context* ctx, otg_stream stream) {
  if (!ctx || !stream) return OTG_ERR_INVALID_ARGUMENT;  // the default stream cannot be captured
  OTG_HIP(ctx, hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
  return OTG_OK;
}

extern "C" otg_status otg_graph_end(otg_context* ctx, otg_stream stream, otg_graph** out) {
  if (!ctx || !stream || !out) return OTG_ERR_INVALID_ARGUMENT;
  otg_graph* g = new otg_graph();
  hipError_t e = hipStreamEndCapture((hipStream_t)stream, &g->graph);
  if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    otg_graph_destroy(g);
    otg_set_error(ctx, "graph capture: %s", hipGetErrorString(e));
    return OTG_ERR_DEVICE;
  }
  *out = g;
  return OTG_OK;
}

extern "C" otg_status otg_graph_launch(otg_context* ctx, otg_graph* g, otg_stream stream) {
  if (!ctx || !g) return OTG_ERR_INVALID_ARGUMENT;
  OTG_HIP(ctx, hipGraphLaunch(g->exec, otg_pick_stream(ctx, stream)));
  return OTG_OK;
}

extern "C" otg_status otg_graph_destroy(otg_graph* g) {
  if (!g) return OTG_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return OTG_OK;
}
