/*
 * otg_gen.h -- seeded synthetic workloads and host-side graph preparation.
 *
 * The generators have no reference counterpart: they produce inputs of the
 * shapes the reference's flattening and levelization code hands to its GPU
 * path (FlatRctStorage, ot/timer/net.hpp:149-166; the fanin CSR of
 * Timer::_build_prop_tasks_cuda, ot/timer/timer.cpp:1000-1063).  Every value
 * is a function of (seed, net or pin index) only, so results do not depend
 * on the thread count.
 */
#ifndef OTG_GEN_H_
#define OTG_GEN_H_

#include <stdint.h>

#include "otg.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
  OTG_SHAPE_RANDOM = 0,   /* node k>0 picks its parent uniformly in [0,k)   */
  OTG_SHAPE_CHAIN = 1,    /* parent k-1 with prob 1-branch_prob, else U[0,k) */
  OTG_SHAPE_BALANCED = 2  /* complete binary tree: parent (k-1)/2          */
};

typedef struct otg_gen_size_class {
  double weight;          /* relative probability of the class              */
  int32_t min_nodes;      /* net size ~ U{min..max}                         */
  int32_t max_nodes;
  int32_t shape;          /* OTG_SHAPE_*                                    */
  int32_t pad;
  double branch_prob;     /* OTG_SHAPE_CHAIN only                           */
} otg_gen_size_class;

typedef struct otg_gen_rc_params {
  uint64_t seed;
  int32_t num_nets;
  int32_t num_classes;          /* 1..4                                     */
  otg_gen_size_class classes[4];
  double r_min, r_max;          /* per-edge R, log-uniform (ohm)            */
  double c_min, c_max;          /* per-node ground C, log-uniform (fF)      */
  double pin_min, pin_max;      /* sink (leaf) pin cap base, log-uniform;   */
                                /* 0,0 = no pin caps                        */
  double corner_scale[4];       /* pin cap multiplier per corner            */
  double slew_min, slew_max;    /* driver slew, uniform (ps)                */
  double r_unit;                /* stored R = R * r_unit (1e-3: ohm->kohm)  */
  int32_t relabel;              /* random node numbering inside each net    */
  int32_t random_orient;        /* random (s,t) order of each edge          */
  int64_t first_net;            /* global index of the first net: the nets  */
                                /* are nets first_net.. of a larger         */
                                /* population (each net is drawn from its   */
                                /* own index), so a large workload can be   */
                                /* generated and solved in batches          */
} otg_gen_rc_params;

/* Net sizes: nodes_start[N+1] and the total node count. */
otg_status otg_gen_rc_sizes(const otg_gen_rc_params* p, int32_t* nodes_start,
                            int64_t* total_nodes);

/* Fill the RctEdgeArray-style inputs (reference ot/cuda/rct.cuh:20-28):
 * edges[2E], roots[N], edge_res[E], node_cap[4V] (ground C on every node,
 * plus the corner-scaled pin cap on leaves; never on the driver, as
 * net.cpp:440), slew_in[4N].  parent_gen (optional, [V]) receives each
 * node's parent in generation order (-1 at the driver), before relabelling.
 * nthreads <= 0 uses the available cores. */
otg_status otg_gen_rc_fill(const otg_gen_rc_params* p, const int32_t* nodes_start,
                           int32_t* edges, int32_t* roots, float* edge_res, float* node_cap,
                           float* slew_in, int32_t* parent_gen, int nthreads);

typedef struct otg_gen_dag_params {
  uint64_t seed;
  int32_t num_pins;
  int32_t num_layers;           /* pins are split evenly over the layers    */
  int64_t num_arcs;             /* exact arc count                          */
  double fanout_alpha;          /* P(k) ~ k^-alpha, 1 <= k <= fanout_cap    */
  int32_t fanout_cap;
  int32_t skip_min, skip_max;   /* skip arcs jump U{skip_min..skip_max}     */
  double skip_prob;             /* fraction of arcs that skip layers        */
  int32_t shuffle_ids;          /* random pin numbering                     */
} otg_gen_dag_params;

/* Layered DAG: arcs from -> to (from in an earlier layer). Every pin outside
 * the last layer has at least one arc into the next layer, so the graph has
 * exactly num_layers levels. */
otg_status otg_gen_dag(const otg_gen_dag_params* p, int32_t* arc_from, int32_t* arc_to,
                       int nthreads);

/* Fanin CSR over candidate pins, as Timer::_build_prop_tasks_cuda builds it
 * (reference ot/timer/timer.cpp:1011-1063): an arc u->v is kept when both
 * ends are candidates (cand == NULL: all pins) and it is not a loop breaker
 * (skip_arc == NULL: none).  Outputs fanin_start[n+1], fanin[*num_kept]
 * (sources of each pin's kept fanin arcs, in arc order), out_degree[n]
 * (kept fanout count) and frontier[*first_size] (candidate pins without kept
 * fanout, ascending id).  fanin must have room for m entries. */
otg_status otg_build_fanin_csr(int32_t n, int64_t m, const int32_t* arc_from,
                               const int32_t* arc_to, const uint8_t* cand,
                               const uint8_t* skip_arc, int32_t* fanin_start, int32_t* fanin,
                               int32_t* out_degree, int32_t* frontier, int32_t* first_size,
                               int64_t* num_kept);

#ifdef __cplusplus
}
#endif

#endif /* OTG_GEN_H_ */
