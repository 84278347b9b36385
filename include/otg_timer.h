/*
 * otg_timer.h -- a full timing update on the device, in one call.
 *
 * Replaces the GPU flow of Timer::update_timing (reference
 * ot/timer/timer.cpp:1239-1334 with Timer::cuda(true)): the RC phase
 * (_build_rc_timing_tasks, timer.cpp:925-979: flatten, BFS-order and solve
 * every net, rct_bfs_and_compute_cuda), the levelization of the candidate
 * pins (_build_prop_tasks_cuda, timer.cpp:988-1071, toposort_compute_cuda)
 * and the level-wise propagation tasks (timer.cpp:1079-1103: _fprop_slew /
 * _fprop_delay / _fprop_at per level from the sources, _bprop_rat per level
 * from the sinks).  Between the RC phase and the propagation, the values the
 * reference's timer reads from each net on demand are gathered on the
 * device: the RC delay and impulse at every net arc's sink (Net::_delay /
 * _slew through FlatRct, ot/timer/net.cpp:646-699, feeding Arc::_fprop_slew
 * / _fprop_delay, ot/timer/arc.cpp:83-135) and the load a cell arc driving
 * a net sees (Net::_load, net.cpp:626-642).
 *
 * Every pin is a candidate (a full update, as the reference's first
 * update_timing); loop-breaker arcs (OTG_ARC_SKIP) are left out of the
 * levelization (timer.cpp:1011-1016) and do not propagate.  Required times
 * come from primary-output values only (see otg_bprop).
 */
#ifndef OTG_TIMER_H_
#define OTG_TIMER_H_

#include <stdint.h>

#include "otg.h"
#include "otg_prop.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A design on the device, kept by the caller for the timer's lifetime (the
 * timer reads R, C and the primary-input / primary-output values again at
 * every update, so the caller may change them between updates).
 * RC nets: the flat input of rct_bfs_and_compute_cuda (ot/cuda/rct.cuh:20-28):
 * node caps include the sink pins' capacitances, not the driver's
 * (net.cpp:433-450).  Timing graph: the fanin CSR of every pin with the arc
 * kinds and timings of otg_prop_graph; a net arc of a net with parasitics
 * (OTG_ARC_NET_RC) names its net and the net-local node (original
 * numbering) of its sink pin, and gets its RC delay and impulse from the
 * solve; a pin that drives an RC net loads its cell arcs with that net's
 * root load, any other pin with pin_load. */
typedef struct otg_timer_design {
  int32_t num_nets;
  int32_t total_num_nodes;
  const int32_t* nodes_start;  /* [N+1]                                       */
  const int32_t* edges;        /* [2(V-N)] undirected, net-local               */
  const int32_t* roots;        /* [N] net-local driver node                    */
  const float* edge_res;       /* [V-N]                                        */
  const float* node_cap;       /* [4V]                                         */
  int32_t num_pins;
  int32_t num_arcs;
  const int32_t* fanin_start;  /* [n+1]                                        */
  const int32_t* fanin;        /* [m] source pin of arc j                      */
  const int32_t* arc_kind;     /* [m] OTG_ARC_*                                */
  const int32_t* arc_timing;   /* [m] cell arcs: timing index                  */
  const int32_t* arc_net;      /* [m] OTG_ARC_NET_RC arcs: the RC net          */
  const int32_t* arc_node;     /* [m] OTG_ARC_NET_RC arcs: the sink's node     */
  const int32_t* pin_net;      /* [n] RC net the pin drives, or -1             */
  const float* pin_load;       /* [4n] load of a pin driving no RC net         */
  const float* pi_slew;        /* [4n] primary-input slew, NaN elsewhere       */
  const float* pi_at;          /* [4n] primary-input arrival, NaN elsewhere    */
  const float* po_rat;         /* [4n] primary-output required time, or NaN    */
} otg_timer_design;

typedef struct otg_timer otg_timer;

/* Validates the sizes, copies nothing, allocates the timer's own arrays
 * (outputs, the RC forest in BFS order, levels, plans).  The library stays
 * the caller's too.  method: the RC solve (OTG_RCT_*). */
otg_status otg_timer_create(otg_context* ctx, const otg_timer_design* design,
                            const otg_prop_lib* lib, otg_rct_method method, otg_timer** timer);
otg_status otg_timer_destroy(otg_timer* timer);

/* Timer::update_timing: RC solve of every net, net-arc RC values and cell
 * loads gathered from it, then forward propagation (slew, arc delays,
 * arrival) and backward propagation (required time, slack) over the levels.
 * The first update also orders the RC trees, levelizes the graph and builds
 * the plans (the topology must not change afterwards; it synchronises
 * `stream` and allocates its working arrays); later updates reuse them,
 * read R, C, pin loads and boundary values afresh, and are asynchronous on
 * `stream`.  A fanin_start that does not run from 0 to num_arcs without
 * decreasing, or a fanin arc naming a pin out of range, gives
 * OTG_ERR_INVALID_ARGUMENT, a net that is not a tree OTG_ERR_NOT_A_TREE
 * (first update). */
otg_status otg_timer_update(otg_timer* timer, otg_stream stream);

/* Device results of the last update (owned by the timer): per pin [4n]
 * slew, at, rat, slack; per arc [8m] arc_delay ([8j + 4el + 2frf + trf]);
 * per arc [4m] the gathered RC delay and impulse; per pin [4n] the cell-arc
 * load; the levels (frontiers [n], frontier_ends [num_levels+1]). */
typedef struct otg_timer_results {
  const float* slew;
  const float* at;
  const float* rat;
  const float* slack;
  const float* arc_delay;
  const float* arc_rc_delay;
  const float* arc_rc_impulse;
  const float* pin_load;
  const int32_t* frontiers;
  const int32_t* frontier_ends;
  int32_t num_levels;
} otg_timer_results;

otg_status otg_timer_results_get(const otg_timer* timer, otg_timer_results* out);

/* Timer::report_slew / report_at / report_rat / report_slack (reference
 * ot/timer/timer.cpp:1635-1710) for a batch of pins: out[4i + c] is the value
 * of pins[i] for corner c = el*2 + rf, NaN where the reference returns no
 * value (or the pin index is out of range).  Device arrays; asynchronous on
 * `stream`. */
typedef enum otg_timer_quantity {
  OTG_REPORT_SLEW = 0,
  OTG_REPORT_AT = 1,
  OTG_REPORT_RAT = 2,
  OTG_REPORT_SLACK = 3
} otg_timer_quantity;

otg_status otg_timer_report(otg_timer* timer, otg_timer_quantity what, int32_t count,
                            const int32_t* pins, float* out, otg_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* OTG_TIMER_H_ */
