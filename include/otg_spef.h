/*
 * otg_spef.h -- SPEF (IEEE 1481) parasitics to the flattened RC forest that
 * otg_rct_bfs / otg_rct_bfs_and_compute take.
 *
 * Replaces the reference's SPEF reading and flattening for its CUDA path:
 * spef::Spef::read + expand_name (ot/parser-spef/parser-spef.hpp),
 * Timer::read_spef / _rebase_unit (ot/timer/spef.cpp:6-50,
 * ot/timer/unit.cpp:328-366) and Net::_init_flat_rct / _make_flat_rct
 * (ot/timer/net.cpp:347-461), which fill FlatRctStorage
 * (ot/timer/net.hpp:149-166).  The flattening rules are the reference's:
 *   - a net has (number of *RES entries + 1) nodes (net.cpp:364);
 *   - local node ids are given in order of first appearance in the *RES
 *     list, node1 before node2 of each entry (net.cpp:392-410);
 *   - edge k of a net is its k-th *RES entry, oriented later by the BFS;
 *   - a node's capacitance is the value of the last *CAP entry naming it as
 *     node1, on all four corners (net.cpp:425-431: a coupling entry counts
 *     as ground capacitance on node1 and overwrites earlier entries);
 *   - the root is the net's driver: here the first *CONN entry that drives
 *     the net (a cell pin *I with direction O, or a port *P with direction
 *     I), the pin the reference takes from the netlist (net.cpp:413-419);
 *   - pin capacitances from a cell library are added to every node except
 *     the root (net.cpp:433-450), through otg_spef_add_pin_cap.
 * Values are rebased to the caller's units when they differ from the file's
 * by 1% or more (unit.cpp:328-366).  Parsing runs on host threads, one range
 * of *D_NET sections per thread.
 */
#ifndef OTG_SPEF_H_
#define OTG_SPEF_H_

#include <stddef.h>
#include <stdint.h>

#include "otg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct otg_spef otg_spef;

typedef struct otg_spef_options {
  double cap_unit;   /* capacitance unit of the results, in farads (e.g. 1e-15);
                        0 keeps the file's *C_UNIT                            */
  double res_unit;   /* resistance unit of the results, in ohms (e.g. 1e3);
                        0 keeps the file's *R_UNIT                            */
  int32_t threads;   /* parse threads; 0: the host's hardware threads (<= 64) */
  int32_t reserved;  /* must be 0                                             */
} otg_spef_options;

/* The flattened forest, in the reference's RctEdgeArrayCUDA form
 * (ot/cuda/rct.cuh:20-42) with node ids local to each net.  Pointers stay
 * valid until otg_spef_destroy. */
typedef struct otg_spef_forest {
  int32_t num_nets;
  int32_t total_num_nodes;
  int32_t total_num_edges;
  int32_t num_undriven;        /* nets with no driving *CONN entry (root -1)     */
  int32_t num_dropped_caps;    /* *CAP entries naming no node of their net (the
                                  reference's lookup has no such case)           */
  int32_t reserved;
  const int32_t* nodes_start;  /* [num_nets+1]                                    */
  const int32_t* edges;        /* [2*total_num_edges] (s, t) local ids; the edges
                                  of net k start at nodes_start[k] - k            */
  const int32_t* roots;        /* [num_nets] local id of the driver, or -1       */
  const float* edges_res;      /* [total_num_edges]                               */
  const float* nodes_cap;      /* [4*total_num_nodes], [node*4 + el*2 + rf]       */
  const float* total_cap;      /* [num_nets] the *D_NET header capacitance        */
  double cap_unit;             /* unit of nodes_cap / total_cap, farads           */
  double res_unit;             /* unit of edges_res, ohms                         */
} otg_spef_forest;

/* Reads and flattens a SPEF file.  On failure *out is null and
 * otg_spef_last_error() describes the problem (line number included). */
otg_status otg_spef_read(const char* path, const otg_spef_options* opt, otg_spef** out);

/* Same, from a buffer in memory (not retained). */
otg_status otg_spef_read_buffer(const char* text, size_t len, const otg_spef_options* opt,
                                otg_spef** out);

/* Message of the last failed otg_spef_read* on the calling thread. */
const char* otg_spef_last_error(void);

otg_status otg_spef_forest_get(const otg_spef* s, otg_spef_forest* f);

/* Names (with *NAME_MAP indices expanded) copied into buf (NUL-terminated,
 * truncated to len-1); the return value is the full length, or -1 for an
 * index out of range. */
int32_t otg_spef_net_name(const otg_spef* s, int32_t net, char* buf, size_t len);
int32_t otg_spef_node_name(const otg_spef* s, int32_t net, int32_t local, char* buf, size_t len);

/* Net index by name (expanded form), or -1. */
int32_t otg_spef_find_net(const otg_spef* s, const char* name);

/* Local id of a node of `net` by name (expanded form), or -1. */
int32_t otg_spef_find_node(const otg_spef* s, int32_t net, const char* name);

/* Adds a library pin capacitance (four corners) to node `local` of `net`,
 * unless it is the root (net.cpp:434-449). */
otg_status otg_spef_add_pin_cap(otg_spef* s, int32_t net, int32_t local, const float cap[4]);

/* Writes a forest as SPEF text: net k is "n<k>", its root is the cell pin
 * "d<k>:Z" (direction O), node j > 0 is "n<k>:<j>"; one *RES entry per edge
 * in edge order and one *CAP entry per node (corner 0).  No reference
 * counterpart: it makes SPEF inputs of any size for tests and measurement. */
otg_status otg_spef_write(const char* path, int32_t num_nets, const int32_t* nodes_start,
                          const int32_t* edges, const int32_t* roots, const float* edges_res,
                          const float* nodes_cap);

otg_status otg_spef_destroy(otg_spef* s);

#ifdef __cplusplus
}
#endif

#endif /* OTG_SPEF_H_ */
