/*
 * otg.h -- C ABI of the MI355X-native RC-tree timing and levelization engine.
 *
 * Two data-parallel paths of OpenTimer are served here:
 *   P1  per-net RC-tree Elmore analysis: load, delay, ldelay, beta, impulse
 *       and output slew of every node for the 4 corners (MIN/MAX x RISE/FALL);
 *   P2  sink-first levelization of the timing graph, plus the per-net BFS
 *       ordering of RC trees that P1 consumes.
 *
 * Conventions shared by every entry point:
 *   - corner index c = el*2 + rf: 0 MIN/RISE, 1 MIN/FALL, 2 MAX/RISE,
 *     3 MAX/FALL (reference ot/headerdef.hpp:172-177); per-node quantities
 *     are stored node-major, [node*4 + c] (reference ot/cuda/rct.cu:12-14);
 *   - "BFS order": nodes of a net are stored so that every node comes after
 *     its parent; the parent of position i is i - pid[i] (reference
 *     ot/cuda/rct.cuh:36-37).  Any topological order is accepted by the
 *     solver; otg_rct_bfs produces the reference's exact BFS order;
 *   - units are the caller's: R*C must already be in the time unit of the
 *     driver slew (the TAU/SPEF convention kOhm * fF = ps);
 *   - every call returns an otg_status; on failure otg_last_error() holds a
 *     message.  Entry points that take an otg_stream are asynchronous on that
 *     stream (NULL = the device's default stream) and use device pointers; the
 *     reference-shaped entry points (otg_rct_bfs_and_compute,
 *     otg_toposort_compute) take host pointers and are synchronous, like the
 *     functions they replace.  They move arrays above 32 MB through pinned
 *     buffers with 8 host threads (128 MB of pinned memory held by the
 *     context after first use);
 *   - a context is not thread-safe; calls that overlap on one context race.
 */
#ifndef OTG_H_
#define OTG_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OTG_NUM_CORNERS 4

typedef enum otg_status {
  OTG_OK = 0,
  OTG_ERR_INVALID_ARGUMENT = 1, /* null pointer, negative size, bad enum   */
  OTG_ERR_NOT_A_TREE = 2,       /* a net's edges do not span it as a tree  */
  OTG_ERR_BAD_TOPOLOGY = 3,     /* pid points outside the net or forward   */
  OTG_ERR_DEVICE = 4,           /* HIP runtime error                       */
  OTG_ERR_OUT_OF_MEMORY = 5,
  OTG_ERR_NO_DEVICE = 6,
  OTG_ERR_CYCLE = 7,            /* levelization: a candidate pin on a cycle */
  OTG_ERR_UNSUPPORTED = 8
} otg_status;

typedef struct otg_context otg_context;   /* device, stream, scratch, tuning */
typedef struct otg_rct_plan otg_rct_plan; /* per-topology work list          */
typedef void* otg_stream;                 /* a hipStream_t; NULL = default   */

/* Context.  No reference counterpart: the reference uses the implicit CUDA
 * device 0 and the default stream (ot/cuda/utils.cuh:12-57). */
otg_status otg_create(int device, otg_context** ctx);
otg_status otg_destroy(otg_context* ctx);
otg_status otg_synchronize(otg_context* ctx, otg_stream stream);
const char* otg_status_string(otg_status s);
const char* otg_last_error(const otg_context* ctx);
int otg_version(void);

/* ------------------------------------------------------------------------ */
/* P1: RC-tree Elmore                                                        */
/* ------------------------------------------------------------------------ */

typedef enum otg_rct_method {
  /* One thread per (net, corner), five serial sweeps per net, the exact
   * arithmetic of compute_net_timing (reference ot/cuda/rct.cu:11-59). */
  OTG_RCT_PLAIN = 0,
  /* Size-binned persistent launch, FP32 arithmetic. */
  OTG_RCT_BATCHED = 1,
  /* Size-binned persistent launch, FP32 inputs and outputs with every
   * subtree and path sum (load, L, delay, impulse) accumulated in FP64, and
   * the final per-node combination that forms ldelay, beta and the slew
   * radicand in FP64 (the slew's square root in FP32). */
  OTG_RCT_BATCHED_COMPENSATED = 2
} otg_rct_method;

/* A forest of RC trees in BFS (topological) order, device pointers.
 * Mirrors the device-side fields of RctCUDA (reference ot/cuda/rct.cuh:7-13)
 * plus the driver slew that the reference applies later per query
 * (FlatRct::slew, ot/timer/net.cpp:280-283). */
typedef struct otg_rct_tree {
  int32_t num_nets;
  int32_t total_num_nodes;
  const int32_t* nodes_start; /* [num_nets+1]                                */
  const int32_t* pid;         /* [V] parent of i is i-pid[i]; ignored at net  */
                              /*     starts (the driver)                      */
  const float* pres;          /* [V] resistance of the edge into node i       */
  const float* cap;           /* [4V] node capacitance (ground + sink pin)    */
  const float* slew_in;       /* [4*num_nets] driver slew; NULL = 0           */
} otg_rct_tree;

/* Outputs, device pointers, each [4V] in the tree's order.  A NULL field is
 * not written, except `load`, which the batched methods also use as working
 * storage when a net is larger than one on-chip window (so it must be given
 * whenever the tree holds such nets). */
typedef struct otg_rct_result {
  float* load;    /* downstream capacitance (net.cpp:127-140)                */
  float* delay;   /* Elmore delay (net.cpp:144-157)                          */
  float* ldelay;  /* sum of cap*delay over the subtree (net.cpp:161-175)     */
  float* beta;    /* second moment (net.cpp:179-191)                         */
  float* impulse; /* 2*beta - delay^2 (net.cpp:191)                          */
  float* slew;    /* sign(si)*sqrt(si^2 + impulse) (net.cpp:30-32)           */
} otg_rct_result;

/* Tuning knobs of the batched launch; zero fields take the defaults (or the
 * values persisted by the autotuner, see otg_rct_autotune). */
typedef struct otg_rct_tuning {
  int32_t ring_slots;     /* largest on-chip ring per wavefront: 256, 320,   */
                          /* 512, 1024 or 2048 (default); nets with parent */
                          /* offsets above ring-64 take the slower         */
                          /* global-window path                            */
  int32_t min_ring;       /* smallest ring (same sizes, <= ring_slots): a    */
                          /* net goes to the smallest ring from min_ring up  */
                          /* that holds its parent offsets, one launch per   */
                          /* ring size holding nets; a smaller ring fits     */
                          /* more wavefronts per CU (default: 256 for the    */
                          /* compensated method, 512 for FP32)               */
  int32_t small_target;   /* node budget of one packed group (64..256)      */
  int32_t grid_per_cu;    /* persistent workgroups per CU (0 = occupancy)   */
  int32_t streams;        /* streams the streaming launches of a solve (one */
                          /* per ring size, global-window nets) are spread  */
                          /* over, 1..4 (default 3): forked from the        */
                          /* caller's stream after the packed groups and    */
                          /* joined back, so that one launch's last, largest*/
                          /* nets overlap the others' work; 1 under graph   */
                          /* capture                                        */
  int32_t small_max;      /* largest net solved in a packed group, 2..group */
                          /* capacity (default: the capacity, 128 or 256    */
                          /* from small_target); larger nets are streamed   */
  int32_t ring_mask;      /* ring sizes that may hold nets: bit k allows the */
                          /* k-th of 256, 320, 512, 1024, 2048 (0 = all)     */
  int32_t first_on_caller;/* 1: the largest-ring launch goes to the caller's */
                          /* stream, which is already running and starts    */
                          /* first; 0 (default): the smallest-ring launch   */
  int32_t small_grid_per_cu; /* workgroups per CU of the smallest-ring      */
                          /* launch when several are spread over streams (0: */
                          /* as many as fit), leaving room on every CU for  */
                          /* the larger rings from the start                */
  int32_t group_alongside;/* 0 (default): the packed groups run first, alone,*/
                          /* dealt out to a resident grid; k > 0: they are   */
                          /* taken from a queue by k workgroups per CU on a  */
                          /* stream of their own, launched beside the       */
                          /* streaming launches (1..16; alone under graph   */
                          /* capture, where the launches share one stream)  */
} otg_rct_tuning;

/* Build the work list of a topology (net sizes and parent offsets only):
 * validates pid, bins nets (packed small groups / streamed large nets /
 * global-window nets) and orders the queue largest-first.  The plan stays
 * valid while nodes_start and pid do not change; R, C and slew may change
 * freely between solves.  No reference counterpart (the reference rebuilds
 * everything on every update, ot/cuda/rct.cu:361-397). */
otg_status otg_rct_plan_create(otg_context* ctx, const otg_rct_tree* topology,
                               const otg_rct_tuning* tuning, otg_rct_method method,
                               otg_rct_plan** plan);
otg_status otg_rct_plan_destroy(otg_rct_plan* plan);
/* Counts of the plan's bins: [0] packed groups, [1] nets in them,
 * [2] streamed nets, [3] global-window nets, [4] max net size. */
otg_status otg_rct_plan_stats(const otg_rct_plan* plan, int64_t stats[8]);

/* FP64 reference values of P1 for precision reports: the reference's four
 * sweeps (ot/timer/net.cpp:101-193, impulse = 2 beta - delay^2 as :191)
 * and the output slew (net.cpp:30-32) carried out in double from the FP32
 * inputs, one thread per (net, corner).  No reference counterpart: a
 * measuring instrument that the FP32 methods are compared against on the
 * device when a workload is too large for a host FP64 check (slow: each
 * net is walked sequentially).  Outputs are device double arrays [4V] in
 * the tree's order; slew may be NULL.  Asynchronous on `stream`. */
otg_status otg_rct_reference_f64(otg_context* ctx, const otg_rct_tree* tree, double* load,
                                 double* delay, double* ldelay, double* beta, double* impulse,
                                 double* slew, otg_stream stream);

/* Measuring instrument, no reference counterpart: moves exactly the
 * algorithmic bytes of one batched solve of `tree` (per node: pid, pres and
 * the four capacitances read, the six outputs written; per net: the node
 * offset and the four driver slews read) with coalesced accesses and no tree
 * work, so that its time is the shortest a solve over the same arrays could
 * take if memory traffic alone bounded it.  Overwrites every output array
 * with a fixed function of the node inputs (load = cap, delay = cap * (pres
 * + pid), ...) that has no timing meaning.  Same 16-byte alignment rule as
 * otg_rct_solve.  Asynchronous on `stream`. */
otg_status otg_rct_stream_ceiling(otg_context* ctx, const otg_rct_tree* tree,
                                  const otg_rct_result* out, otg_stream stream);

/* Solve P1 for every net.  Replaces the compute stage of
 * rct_bfs_and_compute_cuda (reference ot/cuda/rct.cu:385-389, kernel
 * compute_net_timing rct.cu:11-59) on device-resident data; adds beta and
 * slew.  plan may be NULL (then one is built and released inside the call,
 * which allocates).  The batched methods move each node's four corners as
 * one 16-byte vector: cap, slew_in and every result array must be 16-byte
 * aligned (OTG_ERR_INVALID_ARGUMENT otherwise; hipMalloc'd arrays are).
 * Asynchronous on `stream`. */
otg_status otg_rct_solve(otg_context* ctx, otg_rct_plan* plan, const otg_rct_tree* tree,
                         const otg_rct_result* out, otg_rct_method method,
                         otg_stream stream);

/* Re-solve the nets listed in net_ids (device array of `count` net indices,
 * any order; duplicates and out-of-range ids are ignored).  Small nets are
 * solved in the packed groups of the plan, so the other nets of a listed
 * net's group are re-solved too, from their current inputs (unchanged
 * inputs give bit-identical outputs); all other nets' outputs are
 * untouched.  Same 16-byte alignment rule as otg_rct_solve.  Used by
 * incremental re-timing.  Capturable in a hipGraph (no
 * allocation, no memset, no host synchronisation); calls on one plan must
 * not overlap.  No reference counterpart (the reference CUDA path
 * re-flattens all nets, ot/timer/timer.cpp:931). */
otg_status otg_rct_solve_subset(otg_context* ctx, otg_rct_plan* plan, const otg_rct_tree* tree,
                                const otg_rct_result* out, otg_rct_method method,
                                const int32_t* net_ids, int32_t count, otg_stream stream);

/* Undirected RC edges of a forest in original node numbering, device
 * pointers (reference RctEdgeArrayCUDA inputs, ot/cuda/rct.cuh:20-28). */
typedef struct otg_rct_edges {
  int32_t num_nets;
  int32_t total_num_nodes;
  const int32_t* nodes_start; /* [N+1]                                        */
  const int32_t* edges;       /* [2E] (s,t) net-local ids; edges of net k     */
                              /*      start at nodes_start[k]-k, E = V-N      */
  const int32_t* roots;       /* [N] net-local driver id                      */
  const float* edge_res;      /* [E]                                          */
  const float* node_cap;      /* [4V], original order                         */
} otg_rct_edges;

/* Per-net BFS ordering plus reordering of R and C into BFS order.
 * Replaces kernels compute_net_bfs (reference ot/cuda/rct.cu:120-245, with
 * block_couting_sort ot/cuda/sort.cuh:63-142) and prepare_res_cap
 * (rct.cu:314-359).  Outputs (device): node2bfs[V] net-local BFS position of
 * each node; pid[V] (0 at the driver); edges_directed[2E] (s = parent),
 * may alias in->edges; pres[V] (0 at the driver); cap_bfs[4V].  Position of
 * node u = #(nodes nearer the driver) + #(nodes at equal distance with a
 * larger id).  Returns OTG_ERR_NOT_A_TREE if a net is not a spanning tree
 * (checked on the device; the call synchronises to report it). */
otg_status otg_rct_bfs(otg_context* ctx, const otg_rct_edges* in, int32_t* node2bfs,
                       int32_t* pid, int32_t* edges_directed, float* pres, float* cap_bfs,
                       otg_stream stream);

/* Timing at pins from a solve's device results: the lookups the reference's
 * timer makes on the flat RC storage, FlatRct::delay / FlatRct::slew
 * (ot/timer/net.cpp:280-287, result[(arr_start + node2bfs[arr_start + id])*4
 * + c]) and Net::_load (net.cpp:626-642, the driver's load), for a batch of
 * pins.  Pin i is node local[i] (original numbering) of net[i]; outputs
 * [4*count] per corner (c = el*2 + rf): delay and slew at the pin, load of
 * its net.  Any output may be NULL.  A pin with a net or node index out of
 * range gets NaN.  All pointers are device pointers; asynchronous on
 * `stream`. */
otg_status otg_rct_query(otg_context* ctx, int32_t num_nets, const int32_t* nodes_start,
                         const int32_t* node2bfs, const otg_rct_result* res, int32_t count,
                         const int32_t* net, const int32_t* local, float* delay, float* slew,
                         float* load, otg_stream stream);

/* Host-pointer drop-in for rct_bfs_and_compute_cuda(RctEdgeArrayCUDA)
 * (reference ot/cuda/rct.cuh:45, rct.cu:361-397).  Same field meaning as
 * RctEdgeArrayCUDA (rct.cuh:20-42); rct_edges is oriented in place;
 * outputs are in BFS order.  beta, slew and slew_in are extensions (NULL
 * skips them).  Synchronous; allocates and frees device memory per call. */
typedef struct otg_rct_edge_array {
  int32_t num_nets;
  int32_t total_num_nodes;
  int32_t total_num_edges;
  int32_t* rct_nodes_start;    /* in  [N+1]  */
  int32_t* rct_edges;          /* in/out [2E] */
  int32_t* rct_roots;          /* in  [N]    */
  float* rct_edges_res;        /* in  [E]    */
  float* rct_nodes_cap;        /* in  [4V]   */
  int32_t* rct_node2bfs_order; /* out [V]    */
  int32_t* rct_pid;            /* out [V]    */
  float* load;                 /* out [4V]   */
  float* delay;                /* out [4V]   */
  float* ldelay;               /* out [4V]   */
  float* impulse;              /* out [4V]   */
  float* beta;                 /* out [4V], optional */
  float* slew;                 /* out [4V], optional */
  const float* slew_in;        /* in  [4N], optional */
} otg_rct_edge_array;

otg_status otg_rct_bfs_and_compute(otg_context* ctx, otg_rct_edge_array* data,
                                   otg_rct_method method);

/* Host-pointer drop-in for rct_compute_cuda(RctCUDA) (reference
 * ot/cuda/rct.cuh:7-13,44, rct.cu:102-118): nets already in BFS order
 * (nodes_start, pid, pres, cap as in RctCUDA), outputs load, delay, ldelay,
 * impulse; beta, slew and slew_in are extensions (NULL skips them).
 * Synchronous; allocates and frees device memory per call. */
typedef struct otg_rct_host {
  int32_t num_nets;
  int32_t total_num_nodes;
  const int32_t* rct_nodes_start; /* in  [N+1] */
  const int32_t* rct_pid;         /* in  [V]   */
  const float* pres;              /* in  [V]   */
  const float* cap;               /* in  [4V]  */
  float* load;                    /* out [4V]  */
  float* delay;                   /* out [4V]  */
  float* ldelay;                  /* out [4V]  */
  float* impulse;                 /* out [4V]  */
  float* beta;                    /* out [4V], optional */
  float* slew;                    /* out [4V], optional */
  const float* slew_in;           /* in  [4N], optional */
} otg_rct_host;

otg_status otg_rct_compute(otg_context* ctx, const otg_rct_host* data, otg_rct_method method);

/* Wall-clock milliseconds of the stages of the last otg_rct_bfs_and_compute
 * or otg_rct_compute call on this context: [0] device allocation, [1] host
 * to device copies, [2] BFS ordering (0 for otg_rct_compute), [3] plan
 * build, [4] solve, [5] device to host copies, [6] release, [7] the whole
 * call.  Each stage ends with a stream synchronisation, as the reference's
 * stages do (ot/cuda/rct.cu:362-396, timed there by its _prof timers). */
otg_status otg_rct_last_breakdown(const otg_context* ctx, double ms[8]);

/* Autotuner: times candidate tunings of the batched launch on `tree`,
 * keeps the fastest in the context (used by later plans) and, when
 * json_path is non-NULL, writes it there.  No reference counterpart (the
 * reference hard-codes chunk=64 and 128 threads, rct.cu:9,371). */
otg_status otg_rct_autotune(otg_context* ctx, const otg_rct_tree* tree,
                            const otg_rct_result* scratch_out, otg_rct_method method,
                            const char* json_path, otg_rct_tuning* best);
otg_status otg_rct_tuning_load(otg_context* ctx, const char* json_path, otg_rct_tuning* t);

/* ------------------------------------------------------------------------ */
/* P2: levelization                                                          */
/* ------------------------------------------------------------------------ */

/* Host-pointer drop-in for toposort_compute_cuda (reference
 * ot/cuda/toposort.cuh:3-6, toposort.cu:25-72), with C types instead of
 * the std::vector& out-parameter.  edgelist_start[n+1]/edgelist[m] is the
 * fanin CSR, out[n] the fanout count (not modified), frontiers[0:first_size)
 * the initial frontier (pins without fanout); on return frontiers[n] holds
 * the pins grouped by level (order inside a level is unspecified, as in the
 * reference) and frontiers_ends[0..*num_ends) the running level ends,
 * starting 0, first_size (timer.cpp:1061-1062).  frontiers_ends needs room
 * for n+2 entries.  level_of_pin (optional, [n]) gets each pin's level or -1.
 * Synchronous. */
otg_status otg_toposort_compute(otg_context* ctx, int32_t n, int32_t num_edges,
                                int32_t first_size, const int32_t* edgelist_start,
                                const int32_t* edgelist, const int32_t* out,
                                int32_t* frontiers, int32_t* frontiers_ends,
                                int32_t* num_ends, int32_t* level_of_pin);

/* Device-resident levelization: same semantics on device arrays.  The level
 * kernels are launched back-to-back on `stream` with no host round trip per
 * level (the host reads one flag per batch of 32 levels).  out_degree is not
 * modified (a device copy is consumed).  frontiers_ends needs n+2 entries;
 * *num_ends is written to host memory.  Synchronous. */
otg_status otg_levelize(otg_context* ctx, int32_t n, int32_t num_edges,
                        const int32_t* fanin_start, const int32_t* fanin,
                        const int32_t* out_degree, const int32_t* first_frontier,
                        int32_t first_size, int32_t* frontiers, int32_t* frontiers_ends,
                        int32_t* level_of_pin, int32_t* num_ends, otg_stream stream);

/* otg_levelize without device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under stream capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * semantics and results as otg_levelize; its working arrays (the consumed
 * fanout counts, the per-level append counters and the end flag, 2n+3 ints)
 * live in a fixed array of the library on each device (about 16 MB, held
 * until the process ends), so n may not exceed
 * OTG_LEVELIZE_NOALLOC_MAX_PINS (OTG_ERR_INVALID_ARGUMENT otherwise).  Every
 * call on one GPU uses that GPU's one array: calls that overlap on a GPU
 * (from any context or stream) race.  Synchronous (it returns *num_ends). */
#define OTG_LEVELIZE_NOALLOC_MAX_PINS (1 << 21)
otg_status otg_levelize_noalloc(otg_context* ctx, int32_t n, int32_t num_edges,
                                const int32_t* fanin_start, const int32_t* fanin,
                                const int32_t* out_degree, const int32_t* first_frontier,
                                int32_t first_size, int32_t* frontiers, int32_t* frontiers_ends,
                                int32_t* level_of_pin, int32_t* num_ends, otg_stream stream);

/* ------------------------------------------------------------------------ */
/* Incremental re-levelization and hipGraph capture                          */
/* ------------------------------------------------------------------------ */

/* The timing graph of an incremental session, device pointers (kept by the
 * caller for the session's lifetime).  level[num_pins] holds each pin's
 * level; an update rewrites the levels of the pins of its cone. */
typedef struct otg_incr_graph {
  int32_t num_pins;
  int32_t num_nets;
  const int32_t* fanout_start; /* [num_pins+1] fanout CSR                      */
  const int32_t* fanout;
  const int32_t* fanin_start;  /* [num_pins+1] fanin CSR                       */
  const int32_t* fanin;
  const int32_t* out_degree;   /* [num_pins] fanout count                      */
  const int32_t* net_driver;   /* [num_nets] driver pin of each RC net         */
  int32_t* level;              /* [num_pins] in/out                            */
} otg_incr_graph;

typedef struct otg_incr otg_incr;
typedef struct otg_graph otg_graph;

/* Allocates the session's work buffers (no allocation happens later).
 * max_levels bounds the depth of any cone (the graph's number of levels).
 * No reference counterpart: the reference rebuilds the candidate set on the
 * host for every update (Timer::_build_prop_cands, ot/timer/timer.cpp:845-923). */
otg_status otg_incr_create(otg_context* ctx, const otg_incr_graph* g, int32_t max_levels,
                           otg_incr** incr);
otg_status otg_incr_destroy(otg_incr* incr);

/* Re-levelizes the fan-out cone of the driver pins of the nets in net_ids
 * (device, `count` entries): the cone is the forward closure over fanout
 * arcs (the FPROP_CAND set, timer.cpp:845-923) and its pins are levelized
 * sink-first among themselves (toposort semantics of toposort.cu:4-72 on the
 * candidate subgraph, timer.cpp:1011-1071).  A fixed sequence of
 * 2*max_levels+8 kernels with no host round trip: capturable in a hipGraph.
 * Asynchronous on `stream`. */
otg_status otg_incr_relevelize(otg_incr* incr, const int32_t* net_ids, int32_t count,
                               otg_stream stream);

/* Reads the result of the last update (synchronises `stream`): cone size,
 * number of levels, and device pointers to the cone pins, the cone pins by
 * level and the level ends (num_levels+1 entries, starting 0). */
otg_status otg_incr_query(otg_incr* incr, otg_stream stream, int32_t* cone_size,
                          int32_t* num_levels, const int32_t** cone, const int32_t** frontiers,
                          const int32_t** frontier_ends);

/* hipGraph capture of any sequence of asynchronous calls on a non-default
 * stream (for example otg_rct_solve_subset + otg_incr_relevelize), then
 * replay.  No reference counterpart. */
otg_status otg_graph_begin(otg_context* ctx, otg_stream stream);
otg_status otg_graph_end(otg_context* ctx, otg_stream stream, otg_graph** graph);
otg_status otg_graph_launch(otg_context* ctx, otg_graph* graph, otg_stream stream);
otg_status otg_graph_destroy(otg_graph* graph);

#ifdef __cplusplus
}
#endif

#endif /* OTG_H_ */
