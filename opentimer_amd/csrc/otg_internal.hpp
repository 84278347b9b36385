// Internal declarations shared by the HIP translation units of libotg.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/otg.h"

struct otg_context {
  int device = 0;
  hipStream_t stream = nullptr;  // default stream of the device
  int num_cus = 0;
  int lds_per_cu = 0;
  std::string last_error;
  otg_rct_tuning tuning{};  // defaults for new plans (autotuner may update)
  double last_breakdown[8] = {};  // stages of the last host-pointer P1 call (ms)
  int32_t* lv_flag = nullptr;      // pinned host copies of the levelization end flag [2]
  hipEvent_t lv_ev[2] = {};        // their copy events
  // Auxiliary streams that the launches of one batched solve are spread over
  // (forked from and joined back into the caller's stream), created on first
  // use: up to three for streaming launches, one for packed groups that run
  // beside them.
  static constexpr int kAux = 4;
  hipStream_t aux[kAux] = {};
  hipEvent_t ev_fork = nullptr;
  hipEvent_t ev_join[kAux] = {};
  // Host-pointer entry points: large copies between the caller's (pageable)
  // arrays and the device are split over kStage host threads, each with its
  // own stream and two pinned buffers of kStageBytes (created on first use).
  static constexpr int kStage = 8;
  static constexpr size_t kStageBytes = size_t(8) << 20;
  hipStream_t stage_stream[kStage] = {};
  void* stage_buf[kStage][2] = {};
};

// One unit of work of the batched P1 launch: a contiguous range of whole
// nets [net_begin, net_end) occupying positions [pos_begin, pos_end).
enum RctItemKind : int32_t {
  ITEM_PACKED = 0,  // group of whole small nets, solved on chip from a static schedule
  ITEM_STREAM = 1,  // one larger net whose parent offsets fit the ring (< W-64)
  ITEM_GLOBAL = 2   // one net that needs a per-node window in global memory
};

// Ring sizes (slots per wavefront) of the streaming kernel, ascending.  320
// holds parent offsets up to 256 at 20 KB of LDS in the compensated mode
// (7 workgroups per CU against 4 for 512).
constexpr int kRingLadder[] = {256, 320, 512, 1024, 2048};
constexpr int kRingSteps = 5;

struct RctItem {
  int32_t pos_begin, pos_end;
  int32_t net_begin, net_end;
  int32_t kind;
  int32_t k;        // positions per lane of a window (global-window path: 1)
  int64_t scratch;  // float offset into the plan scratch (stream/global items)
};

struct otg_rct_plan {
  otg_context* ctx = nullptr;
  otg_rct_method method = OTG_RCT_BATCHED;
  otg_rct_tuning tuning{};
  int32_t num_nets = 0;
  int32_t total_nodes = 0;
  // Streamed nets in bins by ring size (ascending, at most the tuned ring):
  // a net goes to the first bin whose ring holds its parent offsets
  // (maxpid <= ring - 64).  A smaller ring needs less LDS, so more
  // workgroups fit on a CU.
  static constexpr int kMaxBins = 10;  // rings of kRingLadder, each for ordinary and
                                       // (compensated) chain nets
  static constexpr int kQueueInts = 4 * kMaxBins + 2;
  int num_bins = 0;
  int bin_ring[kMaxBins] = {};
  int bin_chain[kMaxBins] = {};                // 1: chain nets (chain-window kernel variant)
  int32_t bin_items[kMaxBins] = {};            // nets in the bin
  RctItem* d_bin_items[kMaxBins] = {};         // largest first
  RctItem* d_bin_sub[kMaxBins] = {};           // subset solves: the bin's listed nets
  int bin_grid[kMaxBins] = {};                 // persistent workgroups
  int32_t num_groups = 0;
  int32_t group_cap = 256;         // positions of a packed group (kernel capacity)
  int32_t small_max = 256;         // largest net solved in a packed group (<= group_cap)
  RctItem* d_groups = nullptr;     // packed groups (k: group index, scratch: schedule offset)
  uint32_t* d_sched = nullptr;     // schedule words of the packed groups
  int32_t* d_net_group = nullptr;  // [num_nets] group of a packed net, else -1
  int32_t* d_group_flag = nullptr; // [num_groups] subset solves: group listed in this call
  RctItem* d_sub_groups = nullptr; // [num_groups] groups of a subset solve
  int32_t* d_sub_count = nullptr;  // number of them
  int32_t num_gitems = 0;
  RctItem* d_gitems = nullptr;     // global-window nets
  int32_t* d_queue = nullptr;      // [kQueueInts]: next item and finished workgroups of
                                   // streaming bin b [2b, 2b+1] and of the global-window
                                   // kernel [2B, 2B+1] (B = kMaxBins); [2B + 2 + b]: bin
                                   // b's nets of a subset solve; [3B+2, 3B+3]: packed
                                   // groups that run beside the streaming launches
  float* d_scratch = nullptr;      // window storage of stream/global items
  int64_t scratch_floats = 0;
  int32_t* d_net_maxpid = nullptr;  // per net, for subset solves
  int32_t* d_net_chain = nullptr;   // per net: 1 if at least half its windows are paths
  int64_t* d_net_scratch = nullptr; // per net scratch offset (-1: none)
  int grid = 0;                     // persistent workgroups (global-window)
  int grid_group = 0;               // persistent workgroups (packed groups)
  int64_t stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

#define OTG_SET_ERR(ctx, ...) otg_set_error((ctx), __VA_ARGS__)
void otg_set_error(const otg_context* ctx, const char* fmt, ...);

#define OTG_HIP(ctx, expr)                                                        \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      otg_set_error((ctx), "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                    __FILE__, __LINE__);                                          \
      return e_ == hipErrorOutOfMemory ? OTG_ERR_OUT_OF_MEMORY : OTG_ERR_DEVICE;  \
    }                                                                             \
  } while (0)

// NULL selects the device's default (null) stream.
inline hipStream_t otg_pick_stream(const otg_context* ctx, otg_stream s) {
  (void)ctx;
  return (hipStream_t)s;
}

// Defaults of the batched launch, resolved against the method.
otg_rct_tuning otg_rct_resolve_tuning(const otg_context* ctx, const otg_rct_tuning* t,
                                      otg_rct_method m);

// Launchers implemented in otg_rct.hip
otg_status otg_rct_launch_f64(otg_context* ctx, const otg_rct_tree* tree, double* load,
                              double* delay, double* ldelay, double* beta, double* impulse,
                              double* slew, hipStream_t s);
otg_status otg_rct_launch_plain(otg_context* ctx, const otg_rct_tree* tree,
                                const otg_rct_result* out, hipStream_t s);
otg_status otg_rct_launch_batched(otg_context* ctx, otg_rct_plan* plan, const otg_rct_tree* tree,
                                  const otg_rct_result* out, const int32_t* subset,
                                  int32_t count, hipStream_t s);
int otg_rct_blocks_per_cu(otg_context* ctx, const otg_rct_tuning& t, bool comp);

// otg_plan.hip: the plan's device arrays (groups, schedules, bins, per-net
// data) from the topology, for a plan whose tuning, method, group cap and
// bins are set.
otg_status otg_rct_plan_build_device(otg_context* ctx, otg_rct_plan* p, const otg_rct_tree* topo,
                                     hipStream_t s);

// otg_bfs.hip
otg_status otg_rct_bfs_launch(otg_context* ctx, const otg_rct_edges* in, int32_t* node2bfs,
                              int32_t* pid, int32_t* edges_directed, float* pres,
                              float* cap_bfs, int32_t* d_bad, hipStream_t s);
