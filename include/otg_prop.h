/*
 * otg_prop.h -- level-wise forward propagation of slew, arc delay and
 * arrival time over a levelized timing graph, with NLDM table lookups.
 *
 * Replaces the forward tasks of the reference's GPU flow: for the pins of
 * each level, from the sources down to the sinks (Timer::_build_prop_tasks_cuda,
 * ot/timer/timer.cpp:1079-1086), Timer::_fprop_slew / _fprop_delay /
 * _fprop_at (timer.cpp:733-783) over the pin's fanin arcs, Arc::_fprop_slew /
 * _fprop_delay / _fprop_at (ot/timer/arc.cpp:82-147), with cell-arc values
 * from Timing::slew / Timing::delay (ot/liberty/timing.cpp:197-232, 411-540)
 * and Lut::operator() (ot/liberty/lut.cpp:161-241), net-arc values from the
 * RC results (FlatRct::slew / delay, ot/timer/net.cpp:280-287) and the
 * early/late relaxation of Pin::_relax_slew / _relax_at (ot/timer/pin.cpp:324-360).
 *
 * Corner index c = el*2 + rf as everywhere in this library (el 0 = early/MIN,
 * 1 = late/MAX; rf 0 = RISE, 1 = FALL).  An unset optional value of the
 * reference (no slew, no arrival, no delay) is a NaN here, in inputs and
 * outputs alike.
 */
#ifndef OTG_PROP_H_
#define OTG_PROP_H_

#include <stdint.h>

#include "otg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Arc kinds (one per fanin-CSR entry). */
enum {
  OTG_ARC_NET_RC = 0,    /* net arc of a net with RC results: slew =
                            sign(si) sqrt(si^2 + impulse), delay = RC delay */
  OTG_ARC_NET_EMPTY = 1, /* net arc of a net without parasitics: slew = si,
                            delay = 0 (Net::_slew/_delay on EmptyRct)       */
  OTG_ARC_CELL = 2,      /* cell arc: NLDM lookups of its timing            */
  OTG_ARC_SKIP = 3       /* loop breaker: not propagated                    */
};

/* Timing senses and triggers of a cell arc (Timing::is_transition_defined). */
enum {
  OTG_SENSE_POSITIVE_UNATE = 0,  /* rise->rise, fall->fall */
  OTG_SENSE_NEGATIVE_UNATE = 1,  /* rise->fall, fall->rise */
  OTG_SENSE_NON_UNATE = 2        /* all four               */
};
enum {
  OTG_EDGE_NONE = 0,
  OTG_EDGE_RISING = 1,  /* rising_edge timing type: input rise only  */
  OTG_EDGE_FALLING = 2  /* falling_edge timing type: input fall only */
};

/* A cell library as flat device arrays.
 *
 * Tables (ot::Lut): lut_hdr[4*l .. 4*l+3] = {n1, n2, var1, off}: n1 x n2
 * entries; var1 = 0 when the template's variable_1 is the input transition
 * (index_1 = slew, index_2 = load), 1 when it is the total output net
 * capacitance (index_1 = load, index_2 = slew), -1 for a table without a
 * template (a scalar, n1 = n2 = 1); lut_val[off .. off+n1) = index_1,
 * lut_val[off+n1 .. off+n1+n2) = index_2, then the n1*n2 values row by row
 * (value[i*n2 + j] at index_1[i], index_2[j]).  Indices ascend strictly,
 * n1, n2 >= 1, and every table lies inside the num_lut_vals floats of the
 * pool.  A library of at most 36864 words (4 per table header, 12 per
 * timing, 1 per pool float) is read from LDS, a larger one from device
 * memory; the results are the same.
 *
 * Timings (one per cell arc type): timing_sense[t], timing_edge[t], and
 * timing_lut[8*t + 4*el + k], k = 0 cell_rise, 1 cell_fall, 2 rise_transition,
 * 3 fall_transition: a table index, or -1 where the table is absent.
 * timing_has[2*t + el] = 0 where the arc has no timing view for el
 * (tv[el] empty: nothing propagates for el). */
typedef struct otg_prop_lib {
  int32_t num_luts;
  int32_t num_timings;
  int32_t num_lut_vals;        /* floats in lut_val                  */
  const int32_t* lut_hdr;      /* [4*num_luts]                      */
  const float* lut_val;        /* table pool                        */
  const int32_t* timing_sense; /* [num_timings]                     */
  const int32_t* timing_edge;  /* [num_timings]                     */
  const int32_t* timing_lut;   /* [8*num_timings]                   */
  const int32_t* timing_has;   /* [2*num_timings]                   */
} otg_prop_lib;

/* The graph: the fanin CSR of every pin (fanin_start[n+1], fanin[m]; arc j
 * runs fanin[j] -> its pin, a pin's arcs in Pin::_fanin order, loop
 * breakers kept and marked OTG_ARC_SKIP), the arcs' kinds and data, and
 * per-pin inputs, all device arrays.  This is not the candidate CSR that
 * otg_levelize takes for an incremental update (that one leaves out arcs
 * from pins outside the candidates): a pin of the updated cone relaxes over
 * all its fanin arcs (timer.cpp:733-783), reading the slew and arrival of a
 * source outside the frontiers as the caller's arrays hold them (the
 * previous results).  The frontiers come from the candidate levelization. */
typedef struct otg_prop_graph {
  int32_t num_pins;
  int32_t num_arcs;
  const int32_t* fanin_start;  /* [n+1]                                     */
  const int32_t* fanin;        /* [m]   source pin of arc j                 */
  const int32_t* arc_kind;     /* [m]   OTG_ARC_*                           */
  const int32_t* arc_timing;   /* [m]   cell arcs: timing index             */
  const float* arc_rc_delay;   /* [4m]  RC net arcs: delay at the arc's sink */
  const float* arc_rc_impulse; /* [4m]  RC net arcs: impulse at its sink     */
  const float* pin_load;       /* [4n]  load a cell arc driving the pin sees
                                         (Net::_load of the pin's net; 0 when
                                         the pin has no net)                  */
  const float* pi_slew;        /* [4n]  primary-input slew, NaN elsewhere    */
  const float* pi_at;          /* [4n]  primary-input arrival, NaN elsewhere */
} otg_prop_graph;

/* Outputs, device arrays (NaN = unset).  arc_delay may be NULL. */
typedef struct otg_prop_result {
  float* slew;       /* [4n]                                     */
  float* at;         /* [4n]                                     */
  float* arc_delay;  /* [8m]: arc j, [8j + 4el + 2frf + trf]     */
} otg_prop_result;

/* Forward propagation over the levels that otg_levelize produced for this
 * CSR (frontiers[n], frontier_ends[num_levels+1], sink-first: the pins of
 * level k are frontiers[frontier_ends[k] .. frontier_ends[k+1])), from the
 * last level (sources) to level 0, one launch per level.  Every pin of the
 * candidate graph is assigned; pins outside frontiers are left as they are.
 * Asynchronous on `stream`. */
otg_status otg_fprop(otg_context* ctx, const otg_prop_graph* graph, const otg_prop_lib* lib,
                     const int32_t* frontiers, const int32_t* frontier_ends,
                     int32_t num_levels, const otg_prop_result* out, otg_stream stream);

/* A level-ordered plan of a levelized graph for repeated propagations (no
 * counterpart in the reference, which walks its pins in place: the plan is
 * this library's layout for the same per-level tasks, timer.cpp:1079-1086).
 * The pins of the frontiers are renumbered to their position in the
 * frontiers, so that the pins of a level are consecutive, and their fanin
 * arcs to consecutive slots in that order; the positions of every level are
 * packed into wavefront chunks (at most 16 positions and 16 arc slots, a
 * position with more slots alone).  Built once per levelization (allocates,
 * copies the slot offsets to the host for the packing, synchronises
 * `stream`); the arrays it was built from may be freed afterwards. */
typedef struct otg_fprop_plan otg_fprop_plan;

otg_status otg_fprop_plan_create(otg_context* ctx, int32_t num_pins, int32_t num_arcs,
                                 const int32_t* fanin_start, const int32_t* fanin,
                                 const int32_t* frontiers, const int32_t* frontier_ends,
                                 int32_t num_levels, otg_stream stream, otg_fprop_plan** plan);
otg_status otg_fprop_plan_destroy(otg_fprop_plan* plan);
/* Pins in the plan's levels and their fanin arcs. */
otg_status otg_fprop_plan_sizes(const otg_fprop_plan* plan, int32_t* positions, int32_t* slots);

/* otg_fprop over a plan: the same results, bit for bit.  Each level's
 * wavefront chunks read the graph's per-arc and per-pin inputs through the
 * plan's slot and position indices and write slew, arrival and arc delays
 * in pin and arc order.  graph must have the plan's CSR (its fanin_start /
 * fanin are not read again).  Asynchronous on `stream`; no allocation. */
otg_status otg_fprop_run(otg_context* ctx, otg_fprop_plan* plan, const otg_prop_graph* graph,
                         const otg_prop_lib* lib, const otg_prop_result* out, otg_stream stream);

/* Backward propagation of required times over the same levels, level 0
 * (the sinks) first: the forward tasks' counterpart (timer.cpp:1094-1103 over
 * Timer::_bprop_rat, timer.cpp:812-841, and Arc::_bprop_rat for propagation
 * arcs, arc.cpp:150-175).  A pin's required time starts at its primary-
 * output value (po_rat, NaN elsewhere) and is relaxed over its fanout arcs
 * with rat(sink, el, trf) - delay(el, frf, trf), the delays being the
 * forward pass's arc_delay; MIN keeps the larger value, MAX the smaller
 * (Pin::_relax_rat, pin.cpp:362-380).  Required times from timing checks
 * (setup / hold tests of sequential cells, timer.cpp:823-835) are not
 * computed: po_rat is the only source.  slack (optional, needs at):
 * MIN at - rat, MAX rat - at (Pin::slack, pin.cpp:251-256).  arc_delay and
 * rat must be 8-byte aligned.  Asynchronous on `stream`.
 * Inputs read outside the levels: the rat of a fanout sink that lies
 * outside the frontiers, and the arc_delay of an arc that otg_fprop did not
 * write (an arc into a pin outside its frontiers), are read as the caller
 * left them -- earlier results of an incremental update, or NaN (unset) for
 * a first update; otg_bprop writes only the pins of the frontiers. */
typedef struct otg_bprop_graph {
  int32_t num_pins;
  int32_t num_arcs;
  const int32_t* fanout_start; /* [n+1] arcs grouped by source pin          */
  const int32_t* fanout_arc;   /* [m]   arc ids (fanin-CSR positions)       */
  const int32_t* arc_to;       /* [m]   sink pin of every arc               */
  const float* arc_delay;      /* [8m]  otg_prop_result.arc_delay           */
  const float* po_rat;         /* [4n]  primary-output required time or NaN */
  const float* at;             /* [4n]  arrival (for slack) or NULL         */
  const int32_t* fanout_to;    /* [m]   sink pin of fanout entry k, i.e.
                                *       arc_to[fanout_arc[k]], or NULL: saves
                                *       one dependent load per arc          */
} otg_bprop_graph;

otg_status otg_bprop(otg_context* ctx, const otg_bprop_graph* graph, const int32_t* frontiers,
                     const int32_t* frontier_ends, int32_t num_levels, float* rat, float* slack,
                     otg_stream stream);

/* Host helper: the fanout CSR and arc sinks of a fanin CSR (host arrays;
 * arcs of one source in arc order), as the reference's host builds its
 * per-pin fanout lists. */
otg_status otg_build_fanout_csr(int32_t num_pins, const int32_t* fanin_start,
                                const int32_t* fanin, int32_t* fanout_start,
                                int32_t* fanout_arc, int32_t* arc_to);

#ifdef __cplusplus
}
#endif

#endif /* OTG_PROP_H_ */
