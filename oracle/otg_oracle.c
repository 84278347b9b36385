/*
 * otg_oracle.c -- CPU restatement of OpenTimer's RC-tree Elmore path and
 * timing-graph levelization, used ONLY as the checker.
 *
 * TEST INFRASTRUCTURE.  Nothing in the product (the opentimer_amd package and
 * its libotg.so) links, loads or calls this file.  Only tests/, the smoke
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py load the
 * shared library built from it (oracle/_build/liboracle.so).
 *
 * Every function cites the reference file:line whose semantics it restates.
 * The reference's own implementation cannot be compiled here (its CPU timer
 * headers do not build against libstdc++ 11 without source patches and its
 * CUDA files need nvcc), so this restatement is pinned instead by the closed
 * forms, brute-force sums and invariants in tests/test_oracle.py.
 *
 * Layout conventions (reference ot/cuda/rct.cuh:20-42):
 *   nodes_start[N+1]    node offsets per net
 *   edges[2*E]          undirected (s,t) pairs, net-local ids; the edges of
 *                       net k start at nodes_start[k]-k (rct.cu:144)
 *   roots[N]            net-local id of the driver
 *   edge_res[E], node_cap[4*V]   (original node order)
 *   corner index c = el*2 + rf  (ot/headerdef.hpp:172-177)
 * Build with -ffp-contract=off so each float operation rounds once.
 */
#define _POSIX_C_SOURCE 200809L /* pthread barriers */
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NC 4 /* MAX_SPLIT_TRAN */

/* ------------------------------------------------------------------------ */
/* Counting argsort, block_couting_sort (ot/cuda/sort.cuh:63-142) with       */
/* init_counts=true: counts[v-amin] = occurrences (:69-73), inclusive prefix  */
/* sum (:85, block_prefix_sum sort.cuh:14-51), then for i = 0..n-1 in order   */
/* orders[i] = --counts[inputs[i]-amin] and outputs[that] = inputs[i]         */
/* (:102-121).  Equal values therefore take positions in descending index    */
/* order.  counts needs amax-amin+1 entries; orders/outputs may be NULL.      */
/* ------------------------------------------------------------------------ */
void ora_counting_argsort(int n, const int32_t* inputs, int amin, int amax, int32_t* counts,
                          int32_t* orders, int32_t* outputs) {
  const int nb = amax - amin + 1;
  for (int b = 0; b < nb; ++b) counts[b] = 0;
  for (int i = 0; i < n; ++i) counts[inputs[i] - amin]++;
  for (int b = 1; b < nb; ++b) counts[b] += counts[b - 1];
  for (int i = 0; i < n; ++i) {
    const int c = --counts[inputs[i] - amin];
    if (orders) orders[i] = c;
    if (outputs) outputs[c] = inputs[i];
  }
}

/* ------------------------------------------------------------------------ */
/* Per-net BFS ordering.                                                     */
/* Restates compute_net_bfs (ot/cuda/rct.cu:120-245) and the argsort done by  */
/* block_couting_sort (ot/cuda/sort.cuh:63-142):                              */
/*  - distances are BFS hop counts from the root over undirected edges;       */
/*  - every edge is oriented parent->child in place (rct.cu:185-199);         */
/*  - the sort walks node ids in ascending order and hands out positions from */
/*    the END of each distance bucket (--counts[val], sort.cuh:102-121), so   */
/*    position(u) = #(nodes with smaller distance)                            */
/*                + #(nodes with equal distance and larger id);               */
/*  - pid[pos(t)] = pos(t) - pos(s) for every oriented edge (rct.cu:233-243). */
/* The reference leaves pid at the root position unwritten (cudaMalloc'd);    */
/* this restatement writes 0 there.                                           */
/* Returns 0, or -1 if some net is not a connected tree (the reference would  */
/* index counts[INT_MAX] in that case).                                       */
/* ------------------------------------------------------------------------ */
int ora_rct_bfs(int num_nets, const int32_t* nodes_start, int32_t* edges,
                const int32_t* roots, int32_t* node2bfs, int32_t* pid) {
  int maxn = 0;
  for (int k = 0; k < num_nets; ++k) {
    int n = nodes_start[k + 1] - nodes_start[k];
    if (n > maxn) maxn = n;
  }
  if (maxn == 0) return 0;
  int32_t* dist = (int32_t*)malloc(sizeof(int32_t) * maxn);
  int32_t* deg = (int32_t*)malloc(sizeof(int32_t) * (maxn + 1));
  int32_t* adj = (int32_t*)malloc(sizeof(int32_t) * 2 * maxn);
  int32_t* queue = (int32_t*)malloc(sizeof(int32_t) * maxn);
  int32_t* bucket = (int32_t*)malloc(sizeof(int32_t) * (maxn + 1));
  int rc = 0;
  for (int k = 0; k < num_nets && rc == 0; ++k) {
    const int off = nodes_start[k];
    const int n = nodes_start[k + 1] - off;
    const int ne = n - 1;
    int32_t* e = edges + 2 * (int64_t)(off - k);
    if (n <= 0) continue;
    if (roots[k] < 0 || roots[k] >= n) { rc = -1; break; }
    /* adjacency (CSR) */
    memset(deg, 0, sizeof(int32_t) * (n + 1));
    for (int i = 0; i < ne; ++i) {
      if (e[2 * i] < 0 || e[2 * i] >= n || e[2 * i + 1] < 0 || e[2 * i + 1] >= n) { rc = -1; break; }
      deg[e[2 * i] + 1]++;
      deg[e[2 * i + 1] + 1]++;
    }
    if (rc) break;
    for (int i = 0; i < n; ++i) deg[i + 1] += deg[i];
    for (int i = 0; i < ne; ++i) {
      adj[deg[e[2 * i]]++] = e[2 * i + 1];
      adj[deg[e[2 * i + 1]]++] = e[2 * i];
    }
    for (int i = n; i > 0; --i) deg[i] = deg[i - 1];
    deg[0] = 0;
    /* BFS hop distances */
    for (int i = 0; i < n; ++i) dist[i] = -1;
    int head = 0, tail = 0;
    dist[roots[k]] = 0;
    queue[tail++] = roots[k];
    while (head < tail) {
      int u = queue[head++];
      for (int j = deg[u]; j < deg[u + 1]; ++j) {
        int v = adj[j];
        if (dist[v] < 0) { dist[v] = dist[u] + 1; queue[tail++] = v; }
      }
    }
    if (tail != n) { rc = -1; break; }
    /* counting sort by distance, positions handed out from the bucket ends */
    ora_counting_argsort(n, dist, 0, n - 1, bucket, node2bfs + off, NULL);
    /* orient edges and fill pid */
    pid[off + node2bfs[off + roots[k]]] = 0;
    for (int i = 0; i < ne; ++i) {
      int s = e[2 * i], t = e[2 * i + 1];
      if (dist[s] > dist[t]) { e[2 * i] = t; e[2 * i + 1] = s; int x = s; s = t; t = x; }
      if (dist[t] != dist[s] + 1) { rc = -1; break; }
      pid[off + node2bfs[off + t]] = node2bfs[off + t] - node2bfs[off + s];
    }
  }
  free(dist); free(deg); free(adj); free(queue); free(bucket);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Reorder R and C into BFS order: prepare_res_cap (ot/cuda/rct.cu:314-359),  */
/* preceded by the two cudaMemset calls (rct.cu:378-379) that leave the root  */
/* position's resistance at 0.  edges must already be oriented.               */
/* ------------------------------------------------------------------------ */
void ora_rct_prepare(int num_nets, const int32_t* nodes_start, const int32_t* edges,
                     const int32_t* node2bfs, const float* edge_res, const float* node_cap,
                     float* pres, float* cap) {
  for (int k = 0; k < num_nets; ++k) {
    const int off = nodes_start[k];
    const int n = nodes_start[k + 1] - off;
    const int32_t* e = edges + 2 * (int64_t)(off - k);
    const float* r = edge_res + (off - k);
    for (int i = 0; i < n; ++i) pres[off + i] = 0.0f;
    for (int i = 0; i + 1 < n; ++i) pres[off + node2bfs[off + e[2 * i + 1]]] = r[i];
    for (int u = 0; u < n; ++u)
      for (int c = 0; c < NC; ++c)
        cap[(int64_t)(off + node2bfs[off + u]) * NC + c] = node_cap[(int64_t)(off + u) * NC + c];
  }
}

/* ------------------------------------------------------------------------ */
/* Reference GPU kernel arithmetic: compute_net_timing (ot/cuda/rct.cu:11-59) */
/* on BFS-ordered arrays, one net and corner at a time, FP32, same operation  */
/* order.  beta is also returned (the reference overwrites it with impulse,   */
/* rct.cu:49-58).  Any output pointer may be NULL except the work arrays      */
/* load/delay/ldelay/beta, which must be present.                             */
/* ------------------------------------------------------------------------ */
void ora_rct_flat_f32(int num_nets, const int32_t* nodes_start, const int32_t* pid,
                      const float* pres, const float* cap, float* load, float* delay,
                      float* ldelay, float* beta, float* impulse) {
  for (int k = 0; k < num_nets; ++k) {
    const int64_t st = nodes_start[k], ed = nodes_start[k + 1];
    for (int c = 0; c < NC; ++c) {
      for (int64_t i = st; i < ed; ++i) {
        load[i * NC + c] = cap[i * NC + c];
        delay[i * NC + c] = 0.0f; ldelay[i * NC + c] = 0.0f; beta[i * NC + c] = 0.0f;
      }
      for (int64_t i = ed - 1; i > st; --i) load[(i - pid[i]) * NC + c] += load[i * NC + c];
      for (int64_t i = st + 1; i < ed; ++i)
        delay[i * NC + c] += delay[(i - pid[i]) * NC + c] + load[i * NC + c] * pres[i];
      for (int64_t i = ed - 1; i > st; --i) {
        ldelay[i * NC + c] += cap[i * NC + c] * delay[i * NC + c];
        ldelay[(i - pid[i]) * NC + c] += ldelay[i * NC + c];
      }
      if (ed > st) ldelay[st * NC + c] += cap[st * NC + c] * delay[st * NC + c];
      for (int64_t i = st + 1; i < ed; ++i)
        beta[i * NC + c] += beta[(i - pid[i]) * NC + c] + ldelay[i * NC + c] * pres[i];
      if (impulse)
        for (int64_t i = st; i < ed; ++i) {
          float t = delay[i * NC + c];
          impulse[i * NC + c] = 2.0f * beta[i * NC + c] - t * t;
        }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* CPU ot::Timer arithmetic: Rct::update_rc_timing and its four recursive     */
/* sweeps (ot/timer/net.cpp:101-193), restated without recursion on the       */
/* BFS-ordered arrays.  A node's fanout is visited in ascending BFS position  */
/* (the reference visits _fanout in *RES insertion order).  Per node:         */
/*   load   = (sum over children, in order, of child load) + cap  (:127-140)  */
/*   delay  = delay(parent) + res * load                          (:144-157)  */
/*   ldelay = (sum over children of child ldelay) + cap*delay     (:161-175)  */
/*   beta   = beta(parent) + res * ldelay                         (:179-193)  */
/*   impulse = 2.0f*beta - std::pow(delay, 2): std::pow(float,int) returns    */
/*   double, so the subtraction runs in double and rounds once to float.      */
/* ------------------------------------------------------------------------ */
static void children_csr(int64_t st, int64_t ed, const int32_t* pid, int32_t* cstart, int32_t* clist) {
  const int n = (int)(ed - st);
  for (int i = 0; i <= n; ++i) cstart[i] = 0;
  for (int i = 1; i < n; ++i) cstart[i - pid[st + i] + 1]++;
  for (int i = 0; i < n; ++i) cstart[i + 1] += cstart[i];
  for (int i = 1; i < n; ++i) clist[cstart[i - pid[st + i]]++] = i; /* ascending */
  for (int i = n; i > 0; --i) cstart[i] = cstart[i - 1];
  cstart[0] = 0;
}

static int max_net_size(int num_nets, const int32_t* nodes_start) {
  int m = 1;
  for (int k = 0; k < num_nets; ++k) {
    int n = nodes_start[k + 1] - nodes_start[k];
    if (n > m) m = n;
  }
  return m;
}

void ora_rct_dfs_f32(int num_nets, const int32_t* nodes_start, const int32_t* pid,
                     const float* pres, const float* cap, float* load, float* delay,
                     float* ldelay, float* beta, float* impulse) {
  const int maxn = max_net_size(num_nets, nodes_start);
  int32_t* cstart = (int32_t*)malloc(sizeof(int32_t) * (maxn + 1));
  int32_t* clist = (int32_t*)malloc(sizeof(int32_t) * (maxn + 1));
  for (int k = 0; k < num_nets; ++k) {
    const int64_t st = nodes_start[k], ed = nodes_start[k + 1];
    const int n = (int)(ed - st);
    if (n <= 0) continue;
    children_csr(st, ed, pid, cstart, clist);
    for (int c = 0; c < NC; ++c) {
#define A(arr, i) arr[(st + (i)) * NC + c]
      for (int i = n - 1; i >= 0; --i) {
        float acc = 0.0f;
        for (int j = cstart[i]; j < cstart[i + 1]; ++j) acc += A(load, clist[j]);
        A(load, i) = acc + A(cap, i);
      }
      A(delay, 0) = 0.0f;
      for (int i = 1; i < n; ++i) A(delay, i) = A(delay, i - pid[st + i]) + pres[st + i] * A(load, i);
      for (int i = n - 1; i >= 0; --i) {
        float acc = 0.0f;
        for (int j = cstart[i]; j < cstart[i + 1]; ++j) acc += A(ldelay, clist[j]);
        A(ldelay, i) = acc + A(cap, i) * A(delay, i);
      }
      A(beta, 0) = 0.0f;
      for (int i = 1; i < n; ++i) A(beta, i) = A(beta, i - pid[st + i]) + pres[st + i] * A(ldelay, i);
      for (int i = 0; i < n; ++i) {
        double d = (double)A(delay, i);
        A(impulse, i) = (float)((double)(2.0f * A(beta, i)) - d * d);
      }
#undef A
    }
  }
  free(cstart); free(clist);
}

/* FP64 oracle: the same sweeps as ora_rct_dfs_f32 carried out in double     */
/* from the FP32 inputs.  This is the reference value for precision reports.  */
void ora_rct_dfs_f64(int num_nets, const int32_t* nodes_start, const int32_t* pid,
                     const float* pres, const float* cap, double* load, double* delay,
                     double* ldelay, double* beta, double* impulse) {
  const int maxn = max_net_size(num_nets, nodes_start);
  int32_t* cstart = (int32_t*)malloc(sizeof(int32_t) * (maxn + 1));
  int32_t* clist = (int32_t*)malloc(sizeof(int32_t) * (maxn + 1));
  for (int k = 0; k < num_nets; ++k) {
    const int64_t st = nodes_start[k], ed = nodes_start[k + 1];
    const int n = (int)(ed - st);
    if (n <= 0) continue;
    children_csr(st, ed, pid, cstart, clist);
    for (int c = 0; c < NC; ++c) {
#define A(arr, i) arr[(st + (i)) * NC + c]
      for (int i = n - 1; i >= 0; --i) {
        double acc = 0.0;
        for (int j = cstart[i]; j < cstart[i + 1]; ++j) acc += A(load, clist[j]);
        A(load, i) = acc + (double)A(cap, i);
      }
      A(delay, 0) = 0.0;
      for (int i = 1; i < n; ++i) A(delay, i) = A(delay, i - pid[st + i]) + (double)pres[st + i] * A(load, i);
      for (int i = n - 1; i >= 0; --i) {
        double acc = 0.0;
        for (int j = cstart[i]; j < cstart[i + 1]; ++j) acc += A(ldelay, clist[j]);
        A(ldelay, i) = acc + (double)A(cap, i) * A(delay, i);
      }
      A(beta, 0) = 0.0;
      for (int i = 1; i < n; ++i) A(beta, i) = A(beta, i - pid[st + i]) + (double)pres[st + i] * A(ldelay, i);
      for (int i = 0; i < n; ++i) A(impulse, i) = 2.0 * A(beta, i) - A(delay, i) * A(delay, i);
#undef A
    }
  }
  free(cstart); free(clist);
}

/* Output slew: RctNode::slew / FlatRct::slew (ot/timer/net.cpp:30-32,      */
/* :280-283): sign(si) * sqrt(si^2 + impulse), with si the driver-side slew  */
/* of the net for that corner (si[4*net + c]).                               */
void ora_slew_f32(int num_nets, const int32_t* nodes_start, const float* si,
                  const float* impulse, float* slew) {
  for (int k = 0; k < num_nets; ++k)
    for (int64_t i = nodes_start[k]; i < nodes_start[k + 1]; ++i)
      for (int c = 0; c < NC; ++c) {
        float s = si[(int64_t)k * NC + c];
        float v = sqrtf(s * s + impulse[i * NC + c]);
        slew[i * NC + c] = s < 0.0f ? -v : v;
      }
}

void ora_slew_f64(int num_nets, const int32_t* nodes_start, const float* si,
                  const double* impulse, double* slew) {
  for (int k = 0; k < num_nets; ++k)
    for (int64_t i = nodes_start[k]; i < nodes_start[k + 1]; ++i)
      for (int c = 0; c < NC; ++c) {
        double s = si[(int64_t)k * NC + c];
        double v = sqrt(s * s + impulse[i * NC + c]);
        slew[i * NC + c] = s < 0.0 ? -v : v;
      }
}

/* ------------------------------------------------------------------------ */
/* Sink-first levelization: toposort_advance + toposort_compute_cuda          */
/* (ot/cuda/toposort.cu:4-72) driven as in Timer::_build_prop_tasks_cuda      */
/* (ot/timer/timer.cpp:1053-1071).  frontiers[0:first_size) holds the         */
/* initial frontier (pins with no fanout); out[] is the fanout count and is   */
/* consumed.  For u in the current frontier, for v in fanin(u): --out[v] and  */
/* v joins the next frontier when it reaches 0.  ends[] gets 0, first_size,   */
/* then the running end of every non-empty frontier; the return value is the */
/* number of entries written to ends (levels + 1).  level[v] = index of the   */
/* frontier holding v, -1 for pins never reached.                             */
/* ------------------------------------------------------------------------ */
int ora_levelize(int n, const int32_t* fanin_start, const int32_t* fanin, int32_t* out,
                 int first_size, int32_t* frontiers, int32_t* ends, int32_t* level) {
  for (int i = 0; i < n; ++i) level[i] = -1;
  int nends = 0;
  ends[nends++] = 0;
  ends[nends++] = first_size;
  for (int i = 0; i < first_size; ++i) level[frontiers[i]] = 0;
  int wo_last = 0, last = first_size, lv = 0;
  while (last > 0) {
    int new_size = 0;
    int32_t* next = frontiers + wo_last + last;
    for (int id = 0; id < last; ++id) {
      int u = frontiers[wo_last + id];
      for (int j = fanin_start[u]; j < fanin_start[u + 1]; ++j) {
        int v = fanin[j];
        if (out[v]-- == 1) { next[new_size++] = v; level[v] = lv + 1; }
      }
    }
    if (!new_size) break;
    wo_last += last;
    last = new_size;
    ends[nends++] = wo_last + new_size;
    ++lv;
  }
  return nends;
}

/* Parallel sink-first Kahn (a CPU baseline for the levelizer): the same     */
/* frontiers as ora_levelize, each frontier split into nthreads contiguous    */
/* parts with atomic out-degree decrements; a thread appends the pins it     */
/* releases in blocks of up to 1024 (one atomic reservation per block).  The  */
/* order inside a frontier depends on the thread interleaving; the level of   */
/* every pin and the frontier sizes do not.                                   */
typedef struct {
  int n, nthreads;
  const int32_t *fanin_start, *fanin;
  int32_t *out, *frontiers, *ends, *level;
  int beg, size, lv, nends, done;
  int cnt; /* pins appended to the next frontier */
  pthread_barrier_t bar;
} LvPar;

typedef struct {
  LvPar* s;
  int tid;
} LvArg;

static void* lv_worker(void* vp) {
  LvArg* a = (LvArg*)vp;
  LvPar* s = a->s;
  int32_t buf[1024];
  for (;;) {
    const int beg = s->beg, size = s->size, lv = s->lv;
    int32_t* next = s->frontiers + beg + size;
    const int lo = (int)((int64_t)size * a->tid / s->nthreads);
    const int hi = (int)((int64_t)size * (a->tid + 1) / s->nthreads);
    int nb = 0;
    for (int id = lo; id < hi; ++id) {
      const int u = s->frontiers[beg + id];
      for (int j = s->fanin_start[u]; j < s->fanin_start[u + 1]; ++j) {
        const int v = s->fanin[j];
        if (__atomic_fetch_sub(&s->out[v], 1, __ATOMIC_RELAXED) == 1) {
          s->level[v] = lv + 1;
          buf[nb++] = v;
          if (nb == 1024) {
            const int k = __atomic_fetch_add(&s->cnt, nb, __ATOMIC_RELAXED);
            memcpy(next + k, buf, sizeof(int32_t) * nb);
            nb = 0;
          }
        }
      }
    }
    if (nb) {
      const int k = __atomic_fetch_add(&s->cnt, nb, __ATOMIC_RELAXED);
      memcpy(next + k, buf, sizeof(int32_t) * nb);
    }
    pthread_barrier_wait(&s->bar);
    if (a->tid == 0) {
      if (s->cnt == 0) {
        s->done = 1;
      } else {
        s->beg += s->size;
        s->size = s->cnt;
        s->ends[s->nends++] = s->beg + s->size;
        s->lv += 1;
        s->cnt = 0;
      }
    }
    pthread_barrier_wait(&s->bar);
    if (s->done) break;
  }
  return NULL;
}

int ora_levelize_par(int n, const int32_t* fanin_start, const int32_t* fanin, int32_t* out,
                     int first_size, int32_t* frontiers, int32_t* ends, int32_t* level,
                     int nthreads) {
  if (nthreads < 1) nthreads = 1;
  for (int i = 0; i < n; ++i) level[i] = -1;
  for (int i = 0; i < first_size; ++i) level[frontiers[i]] = 0;
  LvPar s;
  s.n = n;
  s.nthreads = nthreads;
  s.fanin_start = fanin_start;
  s.fanin = fanin;
  s.out = out;
  s.frontiers = frontiers;
  s.ends = ends;
  s.level = level;
  s.beg = 0;
  s.size = first_size;
  s.lv = 0;
  s.nends = 0;
  s.done = 0;
  s.cnt = 0;
  ends[s.nends++] = 0;
  ends[s.nends++] = first_size;
  if (first_size == 0) return s.nends;
  pthread_barrier_init(&s.bar, NULL, (unsigned)nthreads);
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
  LvArg* args = (LvArg*)malloc(sizeof(LvArg) * nthreads);
  for (int t = 0; t < nthreads; ++t) {
    args[t].s = &s;
    args[t].tid = t;
    if (t) pthread_create(&th[t], NULL, lv_worker, &args[t]);
  }
  lv_worker(&args[0]);
  for (int t = 1; t < nthreads; ++t) pthread_join(th[t], NULL);
  pthread_barrier_destroy(&s.bar);
  free(th);
  free(args);
  return s.nends;
}

/* Longest-path definition of the same levels, by memoised DFS over fanouts: */
/* level(u) = 0 if u has no fanout, else 1 + max over fanouts.  Used as an    */
/* independent check of ora_levelize on acyclic inputs (fanout CSR given).    */
void ora_levels_longest_path(int n, const int32_t* fanout_start, const int32_t* fanout,
                             int32_t* level) {
  int32_t* stack = (int32_t*)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  int32_t* iter = (int32_t*)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) level[i] = -1;
  for (int r = 0; r < n; ++r) {
    if (level[r] >= 0) continue;
    int sp = 0;
    stack[sp++] = r;
    iter[r] = fanout_start[r];
    level[r] = -2; /* on stack */
    while (sp) {
      int u = stack[sp - 1];
      if (iter[u] < fanout_start[u + 1]) {
        int v = fanout[iter[u]++];
        if (level[v] == -1) { level[v] = -2; iter[v] = fanout_start[v]; stack[sp++] = v; }
      } else {
        int best = 0;
        for (int j = fanout_start[u]; j < fanout_start[u + 1]; ++j)
          if (level[fanout[j]] + 1 > best) best = level[fanout[j]] + 1;
        level[u] = best;
        --sp;
      }
    }
  }
  free(stack); free(iter);
}

/* ------------------------------------------------------------------------ */
/* Forward propagation (slew, arc delay, arrival) over a levelized graph     */
/* with NLDM tables.  NaN stands for the reference's unset std::optional.   */
/* Table layout, arc kinds, senses and triggers as in include/otg_prop.h.   */
/* ------------------------------------------------------------------------ */

/* Lut::operator() (ot/liberty/lut.cpp:161-241): the linear inter/extra-     */
/* polation lambda, lower_bound on each index clamped to [1, size-1], and   */
/* the 1xN, Nx1 and NxN cases.                                              */
static float lut_interpolate(float x, float x1, float x2, float y1, float y2) {
  if (x >= FLT_MAX || x <= -FLT_MAX) return x;
  const float slope = (y2 - y1) / (x2 - x1);
  if (x < x1) return y1 - (x1 - x) * slope;
  else if (x > x2) return y2 + (x - x2) * slope;
  else if (x == x1) return y1;
  else if (x == x2) return y2;
  else return y1 + (x - x1) * slope;
}

static int lut_lower_bound(const float* a, int n, float v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

float ora_lut_eval(const int32_t* hdr, const float* pool, int l, float val1, float val2) {
  const int n1 = hdr[4 * l], n2 = hdr[4 * l + 1], off = hdr[4 * l + 3];
  const float* i1 = pool + off;
  const float* i2 = i1 + n1;
  const float* tab = i2 + n2;
  if (n1 == 1 && n2 == 1) return tab[0];
  int a1 = lut_lower_bound(i1, n1, val1), b1 = lut_lower_bound(i2, n2, val2);
  a1 = a1 < n1 - 1 ? a1 : n1 - 1;
  a1 = a1 > 1 ? a1 : 1;
  b1 = b1 < n2 - 1 ? b1 : n2 - 1;
  b1 = b1 > 1 ? b1 : 1;
  const int a0 = a1 - 1, b0 = b1 - 1;
  if (n1 == 1) return lut_interpolate(val2, i2[b0], i2[b1], tab[b0], tab[b1]);
  if (n2 == 1) return lut_interpolate(val1, i1[a0], i1[a1], tab[a0 * n2], tab[a1 * n2]);
  const float v0 = lut_interpolate(val1, i1[a0], i1[a1], tab[a0 * n2 + b0], tab[a1 * n2 + b0]);
  const float v1 = lut_interpolate(val1, i1[a0], i1[a1], tab[a0 * n2 + b1], tab[a1 * n2 + b1]);
  return lut_interpolate(val2, i2[b0], i2[b1], v0, v1);
}

/* Timing::delay / Timing::slew (ot/liberty/timing.cpp:197-232, 411-540):  */
/* transition rule, table by output transition (k = 0/1 cell_rise/fall,     */
/* 2/3 rise/fall_transition), operands in the template's variable order.   */
static float timing_value(const int32_t* sense, const int32_t* edge, const int32_t* tlut,
                          const int32_t* hdr, const float* pool, int t, int el, int frf, int trf,
                          int kind, float slew, float load) {
  if (edge[t] == 1 && frf != 0) return NAN;
  if (edge[t] == 2 && frf != 1) return NAN;
  if (sense[t] == 0 && frf != trf) return NAN;
  if (sense[t] == 1 && frf == trf) return NAN;
  const int l = tlut[8 * t + 4 * el + kind + trf];
  if (l < 0) return NAN;
  if (hdr[4 * l + 2] == 1) return ora_lut_eval(hdr, pool, l, load, slew);
  return ora_lut_eval(hdr, pool, l, slew, load);
}

/* Pin::_relax_slew / _relax_at (ot/timer/pin.cpp:324-360): MIN keeps the   */
/* smaller value, MAX the larger; an unset value takes any.                  */
static float relax(float cur, float v, int el) {
  if (isnan(v)) return cur;
  if (isnan(cur)) return v;
  return el == 0 ? (v < cur ? v : cur) : (v > cur ? v : cur);
}

/* Timer::_fprop_slew / _fprop_delay / _fprop_at (ot/timer/timer.cpp:      */
/* 733-783) with Arc::_fprop_* (ot/timer/arc.cpp:82-147) for every pin of   */
/* the levels, from the last level (sources) to level 0.                    */
/* One pin's slew, arrival and fanin-arc delays (the body of the level loop
 * of timer.cpp:1079-1086 over _fprop_slew/_delay/_at, timer.cpp:733-783). */
typedef struct fprop_in {
  const int32_t *fs, *fi, *kind, *timing;
  const float *rc_delay, *rc_imp, *load, *pi_slew, *pi_at;
  const int32_t *sense, *edge, *tlut, *thas, *hdr;
  const float* pool;
  float *slew, *at, *arc_delay;
} fprop_in;

static void fprop_pin(const fprop_in* P, int v) {
  for (int el = 0; el < 2; ++el)
    for (int trf = 0; trf < 2; ++trf) {
      const int c = el * 2 + trf;
      float s = P->pi_slew[4 * v + c], a = P->pi_at[4 * v + c];
      for (int j = P->fs[v]; j < P->fs[v + 1]; ++j) {
        const int u = P->fi[j], k = P->kind[j];
        for (int frf = 0; frf < 2; ++frf) {
          const float su = P->slew[4 * u + el * 2 + frf];
          const float au = P->at[4 * u + el * 2 + frf];
          /* every delay of the pin's fanin arcs is reset first (timer.cpp:757-759) */
          float so = NAN, d = NAN;
          if (k == 3) {
            /* loop breaker: nothing propagates */
          } else if (k == 0 || k == 1) {
            /* net arcs carry rf -> rf only */
            if (frf == trf) {
              if (k == 0) {
                const float imp = P->rc_imp[4 * j + c];
                if (!isnan(su)) so = su < 0.0f ? -sqrtf(su * su + imp) : sqrtf(su * su + imp);
                d = P->rc_delay[4 * j + c];
              } else {
                so = su;
                d = 0.0f;
              }
            }
          } else {
            const int t = P->timing[j];
            if (P->thas[2 * t + el] && !isnan(su)) {
              const float lc = P->load[4 * v + c];
              so = timing_value(P->sense, P->edge, P->tlut, P->hdr, P->pool, t, el, frf, trf, 2, su,
                                lc);
              d = timing_value(P->sense, P->edge, P->tlut, P->hdr, P->pool, t, el, frf, trf, 0, su,
                               lc);
            }
          }
          s = relax(s, so, el);
          if (P->arc_delay) P->arc_delay[8 * j + 4 * el + 2 * frf + trf] = d;
          if (!isnan(d) && !isnan(au)) a = relax(a, d + au, el);
        }
      }
      P->slew[4 * v + c] = s;
      P->at[4 * v + c] = a;
    }
}

void ora_fprop(int n, const int32_t* fs, const int32_t* fi, const int32_t* kind,
               const int32_t* timing, const float* rc_delay, const float* rc_imp,
               const float* load, const float* pi_slew, const float* pi_at, const int32_t* sense,
               const int32_t* edge, const int32_t* tlut, const int32_t* thas, const int32_t* hdr,
               const float* pool, const int32_t* frontiers, const int32_t* ends, int num_levels,
               float* slew, float* at, float* arc_delay) {
  (void)n;
  const fprop_in P = {fs, fi, kind, timing, rc_delay, rc_imp, load, pi_slew, pi_at,
                      sense, edge, tlut, thas, hdr, pool, slew, at, arc_delay};
  for (int lv = num_levels - 1; lv >= 0; --lv)
    for (int q = ends[lv]; q < ends[lv + 1]; ++q) fprop_pin(&P, frontiers[q]);
}

/* The same over host threads (the propagation's CPU baseline): the pins of a
 * level are split into contiguous shares, one per thread, with a barrier
 * between levels (the per-level parallel_for of timer.cpp:1079-1086).  Pins
 * of one level write only their own values and their own fanin arcs, so the
 * results equal the serial run's bit for bit. */
typedef struct fprop_par {
  const fprop_in* P;
  const int32_t *frontiers, *ends;
  int num_levels, nthreads;
  pthread_barrier_t bar;
} fprop_par;

typedef struct fprop_arg {
  fprop_par* s;
  int t;
} fprop_arg;

static void* fprop_worker(void* p) {
  fprop_arg* a = (fprop_arg*)p;
  fprop_par* s = a->s;
  for (int lv = s->num_levels - 1; lv >= 0; --lv) {
    const int b = s->ends[lv], e = s->ends[lv + 1];
    const int64_t len = e - b;
    const int lo = b + (int)(len * a->t / s->nthreads), hi = b + (int)(len * (a->t + 1) / s->nthreads);
    for (int q = lo; q < hi; ++q) fprop_pin(s->P, s->frontiers[q]);
    pthread_barrier_wait(&s->bar);
  }
  return NULL;
}

void ora_fprop_par(int n, const int32_t* fs, const int32_t* fi, const int32_t* kind,
                   const int32_t* timing, const float* rc_delay, const float* rc_imp,
                   const float* load, const float* pi_slew, const float* pi_at,
                   const int32_t* sense, const int32_t* edge, const int32_t* tlut,
                   const int32_t* thas, const int32_t* hdr, const float* pool,
                   const int32_t* frontiers, const int32_t* ends, int num_levels, float* slew,
                   float* at, float* arc_delay, int nthreads) {
  (void)n;
  if (nthreads < 1) nthreads = 1;
  const fprop_in P = {fs, fi, kind, timing, rc_delay, rc_imp, load, pi_slew, pi_at,
                      sense, edge, tlut, thas, hdr, pool, slew, at, arc_delay};
  fprop_par s;
  s.P = &P;
  s.frontiers = frontiers;
  s.ends = ends;
  s.num_levels = num_levels;
  s.nthreads = nthreads;
  pthread_barrier_init(&s.bar, NULL, (unsigned)nthreads);
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
  fprop_arg* args = (fprop_arg*)malloc(sizeof(fprop_arg) * nthreads);
  for (int t = 0; t < nthreads; ++t) {
    args[t].s = &s;
    args[t].t = t;
    if (t) pthread_create(&th[t], NULL, fprop_worker, &args[t]);
  }
  fprop_worker(&args[0]);
  for (int t = 1; t < nthreads; ++t) pthread_join(th[t], NULL);
  pthread_barrier_destroy(&s.bar);
  free(th);
  free(args);
}

/* Timer::_bprop_rat (ot/timer/timer.cpp:812-841) with Arc::_bprop_rat    */
/* (ot/timer/arc.cpp:150-175) for propagation arcs, levels first to last  */
/* (timer.cpp:1094-1103): a pin's required time starts at its primary-    */
/* output value and is relaxed over its fanout arcs (in arc order) with   */
/* rat(sink, el, trf) - delay(el, frf, trf); Pin::_relax_rat (pin.cpp:    */
/* 362-380) keeps the larger value for MIN, the smaller for MAX.  The     */
/* arc delays are those of the forward pass (NaN where unset), so net     */
/* arcs (rf -> rf only) and loop breakers (all NaN) need no case of their */
/* own.  Slack (pin.cpp:251-256): MIN at - rat, MAX rat - at.             */
static float relax_rat(float cur, float v, int el) {
  if (isnan(v)) return cur;
  if (isnan(cur)) return v;
  return el == 0 ? (v > cur ? v : cur) : (v < cur ? v : cur);
}

void ora_bprop(const int32_t* fo_start, const int32_t* fo_arc, const int32_t* arc_to,
               const float* arc_delay, const float* po_rat, const float* at,
               const int32_t* frontiers, const int32_t* ends, int num_levels, float* rat,
               float* slack) {
  for (int lv = 0; lv < num_levels; ++lv) {
    for (int q = ends[lv]; q < ends[lv + 1]; ++q) {
      const int u = frontiers[q];
      for (int el = 0; el < 2; ++el)
        for (int frf = 0; frf < 2; ++frf) {
          const int c = el * 2 + frf;
          float r = po_rat[4 * u + c];
          for (int k = fo_start[u]; k < fo_start[u + 1]; ++k) {
            const int j = fo_arc[k], v = arc_to[j];
            for (int trf = 0; trf < 2; ++trf) {
              const float d = arc_delay[8 * j + 4 * el + 2 * frf + trf];
              const float rv = rat[4 * v + el * 2 + trf];
              if (!isnan(d) && !isnan(rv)) r = relax_rat(r, rv - d, el);
            }
          }
          rat[4 * u + c] = r;
          if (slack) {
            const float a = at[4 * u + c];
            slack[4 * u + c] = (isnan(a) || isnan(r)) ? NAN : (el == 0 ? a - r : r - a);
          }
        }
    }
  }
}
