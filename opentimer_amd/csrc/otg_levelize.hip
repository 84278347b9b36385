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

#include "otg_internal.hpp"

namespace {

constexpr int kBlock = 256;
constexpr int kBatch = 32;  // level kernels per host check

struct LevelizeParams {
  int32_t n;
  int32_t first_size;
  const int32_t* fanin_start;
  const int32_t* fanin;
  int32_t* out;        // consumed fanout counts (device copy)
  int32_t* frontiers;  // [n], frontier 0 given
  int32_t* ends;       // [n+2]
  int32_t* level;      // [n] or null
  int32_t* cnt;        // [n+2] appends made while processing level l
  int32_t* num_ends;   // [1], INT_MAX until the first empty frontier
};

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

__global__ __launch_bounds__(kBlock) void levelize_level_kernel(LevelizeParams P, int lv) {
  if (lv > P.n) return;
  const int b = P.ends[lv];
  const int e = lv == 0 ? P.first_size : b + P.cnt[lv - 1];
  if (e == b) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicMin(P.num_ends, lv < 1 ? 2 : lv + 1);
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) P.ends[lv + 1] = e;
  const int last = e - b;
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
#pragma unroll
      for (int k = 0; k < kArcs; ++k) old[k] = v[k] >= 0 ? atomicSub(&P.out[v[k]], 1) : 0;
      uint64_t m[kArcs];
      int tot = 0;
#pragma unroll
      for (int k = 0; k < kArcs; ++k) {
        m[k] = __ballot(old[k] == 1);
        tot += __popcll(m[k]);
      }
      if (tot) {
        int base = 0;
        if (lane == 0) base = atomicAdd(ctr, tot);
        base = __shfl(base, 0);
#pragma unroll
        for (int k = 0; k < kArcs; ++k) {
          if (old[k] == 1) {
            next[base + __popcll(m[k] & ((1ull << lane) - 1ull))] = v[k];
            if (P.level) P.level[v[k]] = lv + 1;
          }
          base += __popcll(m[k]);
        }
      }
    }
  }
}

}  // namespace

// Launches the whole sweep on `s` and returns the number of ends entries in
// *num_ends_host (synchronises s once per batch of levels).
otg_status otg_levelize_launch_raw(otg_context* ctx, int32_t n, int32_t first_size,
                                   const int32_t* fanin_start, const int32_t* fanin,
                                   int32_t* out_copy, int32_t* frontiers, int32_t* ends,
                                   int32_t* level, int32_t* scratch, int32_t* num_ends_host,
                                   hipStream_t s) {
  // scratch: [n+2] per-level append counters followed by num_ends
  LevelizeParams P{n, first_size, fanin_start, fanin, out_copy, frontiers, ends, level,
                   scratch, scratch + n + 2};
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
  int32_t ne = INT32_MAX;
  for (int lv0 = 0; lv0 <= n; lv0 += kBatch) {
    for (int lv = lv0; lv < lv0 + kBatch && lv <= n; ++lv)
      hipLaunchKernelGGL(levelize_level_kernel, dim3(grid), dim3(kBlock), 0, s, P, lv);
    OTG_HIP(ctx, hipGetLastError());
    OTG_HIP(ctx, hipMemcpyAsync(&ne, P.num_ends, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    OTG_HIP(ctx, hipStreamSynchronize(s));
    if (ne != INT32_MAX) break;
  }
  if (ne == INT32_MAX) {
    OTG_SET_ERR(ctx, "levelization did not terminate within n levels");
    return OTG_ERR_CYCLE;
  }
  *num_ends_host = ne;
  return OTG_OK;
}
