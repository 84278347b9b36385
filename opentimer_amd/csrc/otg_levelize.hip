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
// The reference launches one kernel per level and copies the new frontier
// size to the host and a zero back (two synchronous copies per level).  Here
// the whole sweep is one cooperative launch: a grid barrier separates the
// levels, frontier appends are compacted per wavefront with a ballot (one
// atomic per wavefront and fanin step instead of one per pin), and the level
// counters rotate through three slots so that no reset races a reader.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include "otg_internal.hpp"

namespace cg = cooperative_groups;

namespace {

constexpr int kBlock = 256;

struct LevelizeParams {
  int32_t n;
  int32_t first_size;
  const int32_t* fanin_start;
  const int32_t* fanin;
  int32_t* out;        // consumed fanout counts (device copy)
  int32_t* frontiers;  // [n], frontier 0 given
  int32_t* ends;       // [n+2]
  int32_t* level;      // [n] or null
  int32_t* counters;   // [4]: three rotating level counters + num_ends
};

__global__ __launch_bounds__(kBlock) void levelize_coop_kernel(LevelizeParams P) {
  cg::grid_group g = cg::this_grid();
  const int lane = threadIdx.x & 63;
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gsz = (int64_t)gridDim.x * blockDim.x;
  if (P.level)
    for (int64_t i = gtid; i < P.n; i += gsz) P.level[i] = -1;
  if (gtid == 0) {
    P.counters[0] = 0;
    P.counters[1] = 0;
    P.counters[2] = 0;
    P.ends[0] = 0;
    P.ends[1] = P.first_size;
  }
  g.sync();
  __threadfence();
  if (P.level)
    for (int64_t i = gtid; i < P.first_size; i += gsz) P.level[P.frontiers[i]] = 0;
  int wo_last = 0, last = P.first_size, lv = 0;
  if (last == 0) {
    if (gtid == 0) P.counters[3] = 2;
    return;
  }
  for (;;) {
    int32_t* next = P.frontiers + wo_last + last;
    int32_t* ctr = P.counters + (lv % 3);
    // all 64 lanes of a wavefront walk the frontier together (for the ballot)
    const int64_t span = ((int64_t)last + 63) & ~63LL;
    for (int64_t id = gtid; id < span; id += gsz) {
      const int u = id < last ? P.frontiers[wo_last + id] : -1;
      const int jb = u >= 0 ? P.fanin_start[u] : 0;
      const int je = u >= 0 ? P.fanin_start[u + 1] : 0;
      for (int j = jb;; ++j) {
        const bool has = j < je;
        if (!__any(has)) break;
        int v = -1;
        bool app = false;
        if (has) {
          v = P.fanin[j];
          app = atomicSub(&P.out[v], 1) == 1;
        }
        const uint64_t m = __ballot(app);
        if (m) {
          const int leader = __ffsll((unsigned long long)m) - 1;
          int base = 0;
          if (lane == leader) base = atomicAdd(ctr, __popcll(m));
          base = __shfl(base, leader);
          if (app) {
            next[base + __popcll(m & ((1ull << lane) - 1ull))] = v;
            if (P.level) P.level[v] = lv + 1;
          }
        }
      }
    }
    g.sync();
    __threadfence();  // acquire: drop L1 lines of the frontier read before the appends
    const int ns = __hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (gtid == 0) {
      P.counters[(lv + 2) % 3] = 0;  // last read before the previous barrier
      if (ns) P.ends[lv + 2] = wo_last + last + ns;
      else P.counters[3] = lv + 2;
    }
    if (ns == 0) break;
    wo_last += last;
    last = ns;
    ++lv;
  }
}

}  // namespace

otg_status otg_levelize_launch_raw(otg_context* ctx, int32_t n, int32_t first_size,
                                   const int32_t* fanin_start, const int32_t* fanin,
                                   int32_t* out_copy, int32_t* frontiers, int32_t* ends,
                                   int32_t* level, int32_t* counters, hipStream_t s) {
  LevelizeParams P{n, first_size, fanin_start, fanin, out_copy, frontiers, ends, level, counters};
  int per_cu = 0;
  OTG_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                   &per_cu, reinterpret_cast<const void*>(&levelize_coop_kernel), kBlock, 0));
  if (per_cu <= 0) per_cu = 1;
  // a barrier over fewer workgroups is cheaper: two per CU give one thread
  // per pin for frontiers up to ~130k pins; never more than the pins need
  int grid = std::min(per_cu, 2) * std::max(1, ctx->num_cus);
  const int64_t need = ((int64_t)n + kBlock - 1) / kBlock;
  if (grid > need) grid = (int)std::max<int64_t>(1, need);
  void* args[] = {&P};
  OTG_HIP(ctx, hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&levelize_coop_kernel),
                                          dim3(grid), dim3(kBlock), args, 0, s));
  return OTG_OK;
}
