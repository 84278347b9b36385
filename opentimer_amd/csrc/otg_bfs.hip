// Per-net BFS ordering of RC trees plus reordering of R and C into BFS order.
// Replaces compute_net_bfs (reference ot/cuda/rct.cu:120-245), its argsort
// block_couting_sort (ot/cuda/sort.cuh:63-142) and prepare_res_cap
// (rct.cu:314-359).
//
// Nets of up to 64 nodes are packed into wavefronts (one lane per node);
// larger nets take one wavefront each, with their adjacency list, distances
// and BFS queue in LDS as 16-bit ids (nets up to kMaxLds nodes) or in global
// scratch (larger nets).  The BFS runs frontier by frontier with ballot
// compaction; the ordering is a stable counting sort by distance in which
// node ids are visited from the largest down, so that the position of node u
// is
//   #(nodes with smaller distance) + #(nodes with equal distance, larger id)
// -- the order the reference's bucket-end counting sort produces.  Work is
// O(nodes) per net (the reference rescans every edge once per BFS level).
//
// Ordering inside the wavefront: LDS operations of one wavefront execute in
// issue order, so the LDS path only needs the compiler to keep program
// order.  The global-scratch path reads and writes scratch with L2-coherent
// atomic accesses (no stale L1 lines) and waits for its own outstanding
// memory operations between phases.
#include <hip/hip_runtime.h>
#include <rocprim/block/block_radix_sort.hpp>

#include <algorithm>
#include <cstring>

#include "otg_internal.hpp"

namespace {

constexpr int kWave = 64;

struct BfsParams {
  int32_t num_nets;
  const int32_t* nodes_start;
  const int32_t* edges_in;
  int32_t* edges_out;
  const int32_t* roots;
  const float* edge_res;
  const float4* node_cap;
  int32_t* node2bfs;
  int32_t* pid;
  float* pres;
  float4* cap_bfs;
  int32_t* gscratch;  // 6 ints per node + 2 per net
  int32_t* bad;       // smallest failing net id
};

template <bool GLOBAL>
struct Mem {
  __device__ __forceinline__ static int ld(const int32_t* p) {
    if constexpr (GLOBAL) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else return *p;
  }
  __device__ __forceinline__ static void st(int32_t* p, int v) {
    if constexpr (GLOBAL) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *p = v;
  }
  __device__ __forceinline__ static void sync() {
    if constexpr (GLOBAL) __builtin_amdgcn_s_waitcnt(0);  // own loads/stores done
    __builtin_amdgcn_wave_barrier();
    __asm__ __volatile__("" ::: "memory");
  }
};

// Scratch layout per net with n nodes (ints): off[n+1] | adj[2n] | dist[n] |
// frontier[n] | count[n+1]
template <bool GLOBAL>
__device__ __forceinline__ void bfs_net(const BfsParams& P, int net, int32_t* base,
                                        int32_t* lds_free, int lds_cap, int lane) {
  using M = Mem<GLOBAL>;
  const int off = P.nodes_start[net];
  const int n = P.nodes_start[net + 1] - off;
  if (n <= 0) return;
  const int ne = n - 1;
  const int32_t* e = P.edges_in + 2LL * (off - net);
  int32_t* adjoff = base;            // n+1
  int32_t* adj = adjoff + (n + 1);   // 2n
  int32_t* dist = adj + 2 * n;       // n
  int32_t* front = dist + n;         // n
  int32_t* cnt = front + n;          // n+1
  int ok = 1;
  const int root = P.roots[net];
  if (root < 0 || root >= n) ok = 0;
  // degrees
  for (int k = lane; k <= n; k += kWave) {
    M::st(&adjoff[k], 0);
    M::st(&cnt[k], 0);
  }
  for (int k = lane; k < n; k += kWave) M::st(&dist[k], -1);
  M::sync();
  for (int k = lane; k < ne; k += kWave) {
    const int s = e[2 * k], t = e[2 * k + 1];
    if (s < 0 || s >= n || t < 0 || t >= n || s == t) {
      ok = 0;
    } else {
      atomicAdd(&adjoff[s + 1], 1);
      atomicAdd(&adjoff[t + 1], 1);
    }
  }
  M::sync();
  ok = __all(ok);
  if (!ok) {
    atomicMin(P.bad, net);  // every lane: the exit stays uniform
    return;
  }
  // exclusive offsets: wave-wide inclusive scan in chunks of 64
  {
    int carry = 0;
    for (int b = 0; b <= n; b += kWave) {
      const int k = b + lane;
      int v = k <= n ? M::ld(&adjoff[k]) : 0;
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        const int t = __shfl_up(v, d);
        if (lane >= d) v += t;
      }
      if (k <= n) M::st(&adjoff[k], v + carry);
      carry += __shfl(v, kWave - 1);
    }
  }
  M::sync();
  // fill adjacency (cnt[] as per-node fill cursor)
  for (int k = lane; k < ne; k += kWave) {
    const int s = e[2 * k], t = e[2 * k + 1];
    M::st(&adj[M::ld(&adjoff[s]) + atomicAdd(&cnt[s], 1)], t);
    M::st(&adj[M::ld(&adjoff[t]) + atomicAdd(&cnt[t], 1)], s);
  }
  if (lane == 0) {
    M::st(&dist[root], 0);
    M::st(&front[0], root);
  }
  M::sync();
  // BFS frontier by frontier
  int fb = 0, fe = 1, level = 0;
  while (fb < fe) {
    int tail = fe;
    for (int b = fb; b < fe; b += kWave) {
      const int k = b + lane;
      const int u = k < fe ? M::ld(&front[k]) : -1;
      const int jb = u >= 0 ? M::ld(&adjoff[u]) : 0, je = u >= 0 ? M::ld(&adjoff[u + 1]) : 0;
      for (int j = jb;; ++j) {
        const bool has = j < je;
        if (!__any(has)) break;
        int v = -1;
        bool take = false;
        if (has) {
          v = M::ld(&adj[j]);
          // a tree reaches v from exactly one neighbour; the CAS keeps a
          // malformed input (cycle) from queueing a node twice
          if (M::ld(&dist[v]) < 0 && atomicCAS(&dist[v], -1, level + 1) == -1) take = true;
        }
        const uint64_t m = __ballot(take);
        if (take) M::st(&front[tail + __popcll(m & ((1ull << lane) - 1ull))], v);
        tail += __popcll(m);
      }
    }
    M::sync();
    fb = fe;
    fe = tail;
    ++level;
  }
  if (fe != n) {
    atomicMin(P.bad, net);  // every lane: the exit stays uniform
    return;
  }
  // level sizes -> level start positions (cnt[d])
  for (int k = lane; k <= n; k += kWave) M::st(&cnt[k], 0);
  M::sync();
  for (int k = lane; k < n; k += kWave) atomicAdd(&cnt[M::ld(&dist[k])], 1);
  M::sync();
  {
    int carry = 0;
    for (int b = 0; b < level; b += kWave) {
      const int k = b + lane;
      int v = k < level ? M::ld(&cnt[k]) : 0;
      const int own = v;
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        const int t = __shfl_up(v, d);
        if (lane >= d) v += t;
      }
      if (k < level) M::st(&cnt[k], v - own + carry);  // exclusive start of level k
      carry += __shfl(v, kWave - 1);
    }
  }
  M::sync();
  // a large net ranks with its level cursors in LDS when they fit
  bool cur_lds = false;
  if constexpr (GLOBAL) {
    if (level <= lds_cap) {
      for (int k = lane; k < level; k += kWave) lds_free[k] = M::ld(&cnt[k]);
      cur_lds = true;
      M::sync();
    }
  }
  // stable ranking: ids from the largest down; in each chunk the lanes that
  // share a distance take consecutive positions, larger ids first, from a
  // base the group leader claims with one atomic on that level's cursor
  for (int top = n - 1; top >= 0; top -= kWave) {
    const int u = top - lane;  // lane 0 holds the largest id of the chunk
    const bool valid = u >= 0;
    const int d = valid ? M::ld(&dist[u]) : -1;
    uint64_t remaining = __ballot(valid);
    int pos = 0;
    while (remaining) {
      const int leader = __ffsll((unsigned long long)remaining) - 1;
      const int key = __shfl(d, leader);
      const uint64_t m = __ballot(valid && d == key);
      int basepos = 0;
      if (lane == leader)
        basepos = cur_lds ? atomicAdd(&lds_free[key], __popcll(m)) : atomicAdd(&cnt[key], __popcll(m));
      basepos = __shfl(basepos, leader);
      if (valid && d == key) pos = basepos + __popcll(m & ((1ull << lane) - 1ull));
      remaining &= ~m;
    }
    if (valid) {
      P.node2bfs[off + u] = pos;
      M::st(&dist[u], pos);  // dist now holds the position
    }
  }
  M::sync();
  // orient edges, write pid / pres / cap in BFS order
  if (lane == 0) {
    P.pid[off] = 0;
    P.pres[off] = 0.0f;
  }
  int32_t* eo = P.edges_out + 2LL * (off - net);
  const float* res = P.edge_res + (off - net);
  for (int k = lane; k < ne; k += kWave) {
    int s = e[2 * k], t = e[2 * k + 1];
    int ps = M::ld(&dist[s]), pt = M::ld(&dist[t]);
    if (ps > pt) {
      const int x = s; s = t; t = x;
      const int y = ps; ps = pt; pt = y;
    }
    eo[2 * k] = s;
    eo[2 * k + 1] = t;
    P.pid[off + pt] = pt - ps;
    P.pres[off + pt] = res[k];
  }
  for (int k = lane; k < n; k += kWave)
    P.cap_bfs[(int64_t)off + M::ld(&dist[k])] = P.node_cap[(int64_t)off + k];
  M::sync();
}

// ---------------------------------------------------------------------------
// Small nets (at most kSmallNodes nodes): consecutive small nets are packed
// into one wavefront, one lane per node and one lane per edge.  Distances
// come from rounds of edge relaxation (one level per round: a round's reads
// of dist are issued before its writes); the position of node u is counted
// directly from its net's distances, #(smaller distance) + #(equal distance,
// larger id), the rule of the reference's counting sort.
// ---------------------------------------------------------------------------
constexpr int kSmallNodes = 64;

// Orders this wavefront's LDS accesses around it (LDS operations of one
// wavefront execute in issue order; the clobber keeps the compiler from
// carrying LDS values in registers across it).
__device__ __forceinline__ void lds_sync() {
  __builtin_amdgcn_wave_barrier();
  __asm__ __volatile__("" ::: "memory");
}

// Inclusive prefix maximum over the wavefront.
__device__ __forceinline__ int wave_prefix_max(int v, int lane) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int t = __shfl_up(v, d);
    if (lane >= d) v = max(v, t);
  }
  return v;
}

// Smallest failing net of the batch (-1: none); reported like the large-net
// path reports its nets.
__device__ __forceinline__ bool small_batch_failed(const BfsParams& P, int a, int badj, int lane) {
  const int b = __shfl(wave_prefix_max(badj >= 0 ? -badj : INT32_MIN, lane), kWave - 1);
  if (b == INT32_MIN) return false;
  atomicMin(P.bad, a - b);  // every lane: the exit stays uniform
  return true;
}

// Solves nets a .. a+k-1, which start at position `base` and hold at most
// kSmallNodes nodes together.
__device__ __forceinline__ void bfs_small_batch(const BfsParams& P, int a, int k, int base,
                                                int32_t* dist, int32_t* pos, int32_t* mark,
                                                int lane) {
  // this lane as net j of the batch: its first slot and size
  const int j = lane;
  const bool isnet = j < k;
  const int sj = isnet ? P.nodes_start[a + j] - base : 0;
  const int nj = isnet ? P.nodes_start[a + j + 1] - base - sj : 0;
  const int rootj = isnet ? P.roots[a + j] : 0;
  const int m = __shfl(sj + nj, k - 1);  // nodes of the batch
  const int me = m - k;                   // edges of the batch
  const int ebase = base - a;             // first edge of net a
  int badj = -1;                          // a failing net of this lane
  if (isnet && nj > 0 && (rootj < 0 || rootj >= nj)) badj = j;
  // node slot -> net: prefix maximum over marks at each net's first slot
  // (nets without nodes own no slot)
  mark[lane] = -1;
  lds_sync();
  if (isnet && nj > 0) mark[sj] = j;
  lds_sync();
  const int nnet = wave_prefix_max(mark[lane], lane);
  // edge slot -> net: marks at each net's first edge (nets with edges only)
  lds_sync();
  mark[lane] = -1;
  lds_sync();
  if (isnet && nj > 1) mark[sj - j] = j;
  lds_sync();
  const int enet = wave_prefix_max(mark[lane], lane);
  // node: local id, net start slot and size
  const bool isnode = lane < m;
  const int ns0 = __shfl(sj, nnet < 0 ? 0 : nnet);
  const int nn = __shfl(nj, nnet < 0 ? 0 : nnet);
  const int rt = __shfl(rootj, nnet < 0 ? 0 : nnet);
  const int u = lane - ns0;
  dist[lane] = (isnode && u == rt) ? 0 : -1;
  // edge: its endpoints as slots
  const bool isedge = lane < me;
  const int es0 = __shfl(sj, enet < 0 ? 0 : enet);
  const int en = __shfl(nj, enet < 0 ? 0 : enet);
  int es = 0, et = 0, ss = 0, ts = 0;
  if (isedge) {
    es = P.edges_in[2LL * (ebase + lane)];
    et = P.edges_in[2LL * (ebase + lane) + 1];
    if (es < 0 || es >= en || et < 0 || et >= en || es == et) badj = enet;
    ss = es0 + (es < 0 || es >= en ? 0 : es);
    ts = es0 + (et < 0 || et >= en ? 0 : et);
  }
  if (small_batch_failed(P, a, badj, lane)) return;
  lds_sync();
  // relaxation rounds (at most one per node)
  for (int r = 0; r < m; ++r) {
    bool ch = false;
    if (isedge) {
      const int ds = dist[ss], dt = dist[ts];
      if (ds >= 0 && dt < 0) {
        dist[ts] = ds + 1;
        ch = true;
      } else if (dt >= 0 && ds < 0) {
        dist[ss] = dt + 1;
        ch = true;
      }
    }
    lds_sync();
    if (!__any(ch)) break;
  }
  const int du = isnode ? dist[lane] : 0;
  // a node left unreached: not a tree
  if (small_batch_failed(P, a, isnode && du < 0 ? nnet : -1, lane)) return;
  // position inside the net
  int p = 0;
  const int nmax = __shfl(wave_prefix_max(isnode ? nn : 0, lane), kWave - 1);
  for (int v = 0; v < nmax; ++v) {
    const int dv = dist[ns0 + (v < nn ? v : 0)];
    if (isnode && v < nn && (dv < du || (dv == du && v > u))) ++p;
  }
  pos[lane] = p;
  lds_sync();
  if (isnode) {
    const int64_t g = (int64_t)base + lane;
    P.node2bfs[g] = p;
    P.cap_bfs[(int64_t)base + ns0 + p] = P.node_cap[g];
  }
  if (isnet && nj > 0) {
    P.pid[base + sj] = 0;
    P.pres[base + sj] = 0.0f;
  }
  if (isedge) {
    int s0 = es, t0 = et;
    int ps = pos[ss], pt = pos[ts];
    if (ps > pt) {
      const int x = s0; s0 = t0; t0 = x;
      const int y = ps; ps = pt; pt = y;
    }
    P.edges_out[2LL * (ebase + lane)] = s0;
    P.edges_out[2LL * (ebase + lane) + 1] = t0;
    P.pid[(int64_t)base + es0 + pt] = pt - ps;
    P.pres[(int64_t)base + es0 + pt] = P.edge_res[ebase + lane];
  }
  lds_sync();
}

// Each wavefront takes ranges of kSmallRange consecutive nets and packs the
// small ones; larger nets are left to the large-net kernels.
constexpr int kSmallRange = 256;

__global__ __launch_bounds__(64) void rct_bfs_small_kernel(BfsParams P) {
  __shared__ int32_t dist[kWave], pos[kWave], mark[kWave];
  const int lane = threadIdx.x;
  for (int r0 = blockIdx.x * kSmallRange; r0 < P.num_nets; r0 += gridDim.x * kSmallRange) {
    const int r1 = min(P.num_nets, r0 + kSmallRange);
    int a = r0;
    while (a < r1) {
      // nets a, a+1, ... while the batch holds at most 64 nodes
      const int base = P.nodes_start[a];
      const int l = a + lane;
      // (a net without nodes ends the run: it has no edge slots, so the
      // nets after it would start their edges before the batch's first)
      const bool fits = l < r1 && P.nodes_start[l + 1] - base <= kSmallNodes &&
                        P.nodes_start[l + 1] > P.nodes_start[l];
      const uint64_t fm = __ballot(fits);
      // fits is monotone in the lane: the batch is the run of leading lanes
      const int k = fm == ~0ull ? kWave : (int)__builtin_ctzll(~fm);
      if (k == 0) {
        ++a;  // a large net (or one without nodes)
        continue;
      }
      bfs_small_batch(P, a, k, base, dist, pos, mark, lane);
      a += k;
    }
  }
}

// ---------------------------------------------------------------------------
// Large nets (more than 64 nodes, at most kMaxLds): one wavefront per net with
// its whole working set in LDS as 16-bit node ids, 10 bytes per node in size
// classes of CAP nodes (CAP = 1024 ... 8192, 10 KB ... 80 KB), so that smaller
// nets run more wavefronts per CU.  Nets come from a per-class list through
// an atomic counter.  Phases (each bounded by the net's size):
//   1. degrees, counted with 32-bit LDS atomics on packed 16-bit counters;
//   2. exclusive offsets by a wave scan; 3. adjacency filled through packed
//      16-bit cursors;
//   4. frontier BFS over a 16-bit queue with ballot compaction (a level
//      appends at most what the queue holds; at most n levels; a node's
//      neighbours are scanned at most 2(n-1) steps);
//   5. level sizes, their exclusive starts, and the positions: ids from the
//      largest down, lanes of equal distance take consecutive positions from
//      one cursor claim per distinct distance per chunk;
//   6. oriented edges, pid, R and C in BFS order.
// A net that is not a spanning tree (an endpoint out of range, a self loop,
// a node left unreached or reached twice) reports its index.
// ---------------------------------------------------------------------------
constexpr int kMaxLds = 8192;
constexpr uint32_t kUnseen = 0xffffu;
constexpr int kBatchLd = 8;  // edge chunks whose loads are in flight together

template <int CAP>
struct BfsLds16 {
  uint32_t off[CAP / 2 + 2];  // degrees -> exclusive offsets [n+1]; after the BFS: level cursors
  uint32_t adj[CAP];          // [2(n-1)] neighbours
  uint32_t dist[CAP / 2];     // [n] distance, then BFS position
  uint32_t q[CAP / 2];        // [n] fill cursors, then the BFS queue
};

__device__ __forceinline__ int ld16(const uint32_t* a, int i) {
  return (int)reinterpret_cast<const uint16_t*>(a)[i];
}
__device__ __forceinline__ void st16(uint32_t* a, int i, int v) {
  reinterpret_cast<uint16_t*>(a)[i] = (uint16_t)v;
}
// Adds v to 16-bit entry i with a 32-bit atomic on its word; returns the
// entry's old value (entries never exceed 65535, so no carry reaches the
// neighbouring entry).
__device__ __forceinline__ int add16(uint32_t* a, int i, int v) {
  const int sh = (i & 1) * 16;
  const uint32_t old = atomicAdd(&a[i >> 1], (uint32_t)v << sh);
  return (int)((old >> sh) & 0xffffu);
}

__device__ __forceinline__ uint64_t lanes_below(int lane) { return (1ull << lane) - 1ull; }

template <int CAP>
__device__ void bfs_net16(const BfsParams& P, int net, BfsLds16<CAP>& S, int lane) {
  const int off0 = P.nodes_start[net];
  const int n = P.nodes_start[net + 1] - off0;
  if (n <= 0 || n > CAP) {  // not a net of this class (the lists never hold one)
    if (n > CAP) atomicMin(P.bad, net);  // every lane: the exit stays uniform
    return;
  }
  const int ne = n - 1;
  const int32_t* e = P.edges_in + 2LL * (off0 - net);
  const int root = P.roots[net];
  bool bad = root < 0 || root >= n;
  // ---- 1. degrees (entry v+1 counts v's edges) ----
  for (int k = lane; k < n / 2 + 1; k += kWave) S.off[k] = 0u;
  for (int k = lane; k < (n + 1) / 2; k += kWave) {
    S.dist[k] = 0xffffffffu;
    S.q[k] = 0u;
  }
  lds_sync();
  // (edges are read kBatchLd chunks at a time: one memory latency per batch)
  for (int k0 = 0; k0 < ne; k0 += kBatchLd * kWave) {
    int es[kBatchLd], et[kBatchLd];
#pragma unroll
    for (int u = 0; u < kBatchLd; ++u) {
      const int k = k0 + u * kWave + lane;
      es[u] = k < ne ? e[2 * k] : 0;
      et[u] = k < ne ? e[2 * k + 1] : 0;
    }
#pragma unroll
    for (int u = 0; u < kBatchLd; ++u) {
      if (k0 + u * kWave + lane >= ne) continue;
      const int s = es[u], t = et[u];
      if (s < 0 || s >= n || t < 0 || t >= n || s == t) {
        bad = true;
      } else {
        add16(S.off, s + 1, 1);
        add16(S.off, t + 1, 1);
      }
    }
  }
  lds_sync();
  if (__any(bad)) {
    atomicMin(P.bad, net);  // every lane: the exit stays uniform
    return;
  }
  // ---- 2. offsets: inclusive scan of the shifted degrees ----
  {
    int carry = 0;
    for (int b = 0; b <= n; b += kWave) {
      const int k = b + lane;
      int v = k <= n ? ld16(S.off, k) : 0;
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        const int t = __shfl_up(v, d);
        if (lane >= d) v += t;
      }
      if (k <= n) st16(S.off, k, v + carry);
      carry += __shfl(v, kWave - 1);
    }
  }
  lds_sync();
  // ---- 3. adjacency ----
  for (int k0 = 0; k0 < ne; k0 += kBatchLd * kWave) {
    int es[kBatchLd], et[kBatchLd];
#pragma unroll
    for (int u = 0; u < kBatchLd; ++u) {
      const int k = k0 + u * kWave + lane;
      es[u] = k < ne ? e[2 * k] : 0;
      et[u] = k < ne ? e[2 * k + 1] : 0;
    }
#pragma unroll
    for (int u = 0; u < kBatchLd; ++u) {
      if (k0 + u * kWave + lane >= ne) continue;
      const int s = es[u], t = et[u];
      const int cs = add16(S.q, s, 1), ct = add16(S.q, t, 1);
      st16(S.adj, ld16(S.off, s) + cs, t);
      st16(S.adj, ld16(S.off, t) + ct, s);
    }
  }
  lds_sync();
  // ---- 4. BFS ----
  if (lane == 0) {
    st16(S.dist, root, 0);
    st16(S.q, 0, root);
  }
  lds_sync();
  int fb = 0, fe = 1, levels = 0;
  bool over = false;
  while (fb < fe && levels < n) {
    int tail = fe;
    for (int b = fb; b < fe; b += kWave) {
      const int k = b + lane;
      const int u = k < fe ? ld16(S.q, k) : -1;
      const int jb = u >= 0 ? ld16(S.off, u) : 0;
      const int je = u >= 0 ? ld16(S.off, u + 1) : 0;
      for (int it = 0; it < 2 * ne; ++it) {
        const int j = jb + it;
        const bool has = j < je;
        if (!__any(has)) break;
        int v = 0;
        bool take = false;
        if (has) {
          v = ld16(S.adj, j);
          take = (uint32_t)ld16(S.dist, v) == kUnseen;
        }
        const uint64_t m = __ballot(take);
        const int slot = tail + __popcll(m & lanes_below(lane));
        if (take) {
          st16(S.dist, v, levels + 1);
          if (slot < n) st16(S.q, slot, v);
        }
        tail += __popcll(m);
      }
    }
    lds_sync();
    if (tail > n) {
      over = true;
      break;
    }
    fb = fe;
    fe = tail;
    ++levels;
  }
  bad = over || fe != n;
  for (int k = lane; k < n && !bad; k += kWave)
    if ((uint32_t)ld16(S.dist, k) == kUnseen) bad = true;
  if (__any(bad)) {
    atomicMin(P.bad, net);  // every lane: the exit stays uniform
    return;
  }
  // ---- 5. level sizes and starts (in off's space), then positions ----
  for (int k = lane; k < levels / 2 + 1; k += kWave) S.off[k] = 0u;
  lds_sync();
  for (int k = lane; k < n; k += kWave) add16(S.off, ld16(S.dist, k), 1);
  lds_sync();
  {
    int carry = 0;
    for (int b = 0; b < levels; b += kWave) {
      const int k = b + lane;
      const int own = k < levels ? ld16(S.off, k) : 0;
      int v = own;
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        const int t = __shfl_up(v, d);
        if (lane >= d) v += t;
      }
      if (k < levels) st16(S.off, k, v - own + carry);
      carry += __shfl(v, kWave - 1);
    }
  }
  lds_sync();
  for (int top = n - 1; top >= 0; top -= kWave) {
    const int u = top - lane;  // lane 0 holds the chunk's largest id
    const bool valid = u >= 0;
    const int d = valid ? ld16(S.dist, u) : -1;
    uint64_t remaining = __ballot(valid);
    int pos = 0;
    for (int it = 0; it < kWave && remaining; ++it) {
      const int leader = __ffsll((unsigned long long)remaining) - 1;
      const int key = __shfl(d, leader);
      const uint64_t m = __ballot(valid && d == key);
      int base = 0;
      if (lane == leader) base = add16(S.off, key, __popcll(m));
      base = __shfl(base, leader);
      if (valid && d == key) pos = base + __popcll(m & lanes_below(lane));
      remaining &= ~m;
    }
    if (valid) {
      P.node2bfs[off0 + u] = pos;
      st16(S.dist, u, pos);  // this chunk's ids only: later chunks read their own
    }
  }
  lds_sync();
  // ---- 6. oriented edges, pid, R and C in BFS order ----
  if (lane == 0) {
    P.pid[off0] = 0;
    P.pres[off0] = 0.0f;
  }
  int32_t* eo = P.edges_out + 2LL * (off0 - net);
  const float* res = P.edge_res + (off0 - net);
  for (int k0 = 0; k0 < ne; k0 += kBatchLd * kWave) {
    int es[kBatchLd], et[kBatchLd];
    float er[kBatchLd];
#pragma unroll
    for (int u = 0; u < kBatchLd; ++u) {
      const int k = k0 + u * kWave + lane;
      es[u] = k < ne ? e[2 * k] : 0;
      et[u] = k < ne ? e[2 * k + 1] : 0;
      er[u] = k < ne ? res[k] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < kBatchLd; ++u) {
      const int k = k0 + u * kWave + lane;
      if (k >= ne) continue;
      int s = es[u], t = et[u];
      int ps = ld16(S.dist, s), pt = ld16(S.dist, t);
      if (ps > pt) {
        const int x = s; s = t; t = x;
        const int y = ps; ps = pt; pt = y;
      }
      eo[2 * k] = s;
      eo[2 * k + 1] = t;
      P.pid[off0 + pt] = pt - ps;
      P.pres[off0 + pt] = er[u];
    }
  }
  for (int k0 = 0; k0 < n; k0 += kBatchLd * kWave) {
    float4 c[kBatchLd];
#pragma unroll
    for (int u = 0; u < kBatchLd; ++u) {
      const int k = k0 + u * kWave + lane;
      c[u] = k < n ? P.node_cap[(int64_t)off0 + k] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < kBatchLd; ++u) {
      const int k = k0 + u * kWave + lane;
      if (k < n) P.cap_bfs[(int64_t)off0 + ld16(S.dist, k)] = c[u];
    }
  }
  lds_sync();  // the next net reuses the LDS
}

// ---------------------------------------------------------------------------
// Large nets, one workgroup per net: rooting by an Euler tour.
//
// The wavefront above needs one dependent step per BFS level, and the
// chain-like nets of a few thousand nodes have thousands of levels.  This
// path finds every node's distance from the root in O(log n) dependent
// rounds instead: the Euler-tour technique of Tarjan and Vishkin (1985),
// with list ranking by a sparse ruling set (Reid-Miller, 1994).  Phases:
//   1. degrees (packed 16-bit LDS atomics) with the endpoint checks;
//   2. CSR offsets by a block scan;
//   3. adjacency slots and, for every slot (an arc u->v), its twin slot
//      (the arc v->u);
//   4. the tour successor of each arc u->v: the slot after its twin in v's
//      list, cyclically; the arc whose successor would be the root's first
//      slot ends the tour;
//   5. list ranking: every S-th slot (S a power of two, at most T rulers)
//      and the tour's first arc are rulers; one thread per ruler walks to the
//      next ruler, the rulers' list is ranked by pointer jumping
//      (bit_width(R) + 1 rounds), and a second walk writes every arc's tour
//      position.  The successor map is a bijection on the arcs except at the
//      tour's end, so a walk from a ruler ends at the end, at another ruler,
//      or back at its start; every walk is also held to A = 2(n-1) steps;
//   6. an arc is a descent (parent -> child) iff it comes before its twin in
//      the tour; a block scan of +1 (descent) / -1 (ascent) over the tour
//      gives each child its distance at its descent;
//   7. BFS positions: the keys (distance, CAP-1-id) in a block radix sort,
//      so that the position of u is #(smaller distance) + #(equal distance,
//      larger id), the order of the reference's bucket-end counting sort;
//   8. oriented edges, pid, R and C in BFS order, as the wavefront path.
// A net that is not a spanning tree leaves arcs outside the tour from the
// root (the rulers' list from the first arc covers fewer than A arcs), or a
// node without a distance; either rejects it.  Every early exit is decided
// from values that all threads read alike (the net's own header, or LDS
// after a barrier, made scalar), so the workgroup leaves a net together.
// ---------------------------------------------------------------------------
constexpr uint32_t kTourEnd = 0xffffu;

template <int CAP, int T>
struct TourLds {
  uint32_t off[CAP / 2 + 4];  // 16-bit: degrees -> offsets [n+1]; distances, then positions [n]
  uint32_t adj[CAP];          // 16-bit [A]: neighbour at each slot      } from phase 7:
  uint32_t tw[CAP];           // 16-bit [A]: twin slot                   } the sort's
  uint32_t nx[CAP];           // 16-bit [A]: tour successor; +-1 scan    } storage
  uint32_t rk[CAP];           // 16-bit: fill cursors [n]; tour positions [A]
  int32_t rnext[T + 2];       // ruler -> next ruler (-1: the end, -2: walk bound hit)
  int32_t rsuf[T + 2];        // arcs from the ruler to the end
  int32_t part[T / kWave];    // block-scan wave totals
  int32_t ctl[4];             // [0] net index  [1] failure  [2] largest distance
};

template <int T>
__device__ __forceinline__ int block_uniform(const int32_t* p) {
  return __builtin_amdgcn_readfirstlane(*p);
}

// Inclusive scan of L 16-bit entries of a, in place (entries read as signed
// if SIGNED).  Wave w scans its own whole-chunk segment, after a pass that
// sums each segment.  Starts and ends with a barrier; returns the total.
template <int T, bool SIGNED>
__device__ int block_scan16(uint32_t* a, int L, int32_t* part, int tid) {
  constexpr int NW = T / kWave;
  const int lane = tid & (kWave - 1), w = tid / kWave;
  const int seg = ((L + NW - 1) / NW + kWave - 1) & ~(kWave - 1);
  const int b = w * seg, e = min(L, b + seg);
  auto get = [&](int i) {
    const int x = ld16(a, i);
    return SIGNED ? (int)(int16_t)x : x;
  };
  __syncthreads();
  int s = 0;
  for (int i = b + lane; i < e; i += kWave) s += get(i);
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) s += __shfl_xor(s, d);
  if (lane == 0) part[w] = s;
  __syncthreads();
  int carry = 0, total = 0;
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    const int p = part[k];
    carry += k < w ? p : 0;
    total += p;
  }
  for (int c = b; c < e; c += kWave) {
    const int i = c + lane;
    int v = i < e ? get(i) : 0;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const int t = __shfl_up(v, d);
      if (lane >= d) v += t;
    }
    if (i < e) st16(a, i, v + carry);
    carry += __shfl(v, kWave - 1);
  }
  __syncthreads();
  return total;
}

template <int CAP, int T>
__device__ void tour_net(const BfsParams& P, int net, TourLds<CAP, T>& S, int tid) {
  constexpr int IPT = CAP / T;
  constexpr int LOG = __builtin_ctz(CAP);
  const int off0 = __builtin_amdgcn_readfirstlane(P.nodes_start[net]);
  const int n = __builtin_amdgcn_readfirstlane(P.nodes_start[net + 1]) - off0;
  if (n <= 1 || n > CAP) {  // not a net of this class (the lists never hold one)
    if (tid == 0 && n > CAP) atomicMin(P.bad, net);
    return;
  }
  const int ne = n - 1, A = 2 * ne;
  const int32_t* e = P.edges_in + 2LL * (off0 - net);
  const int root = __builtin_amdgcn_readfirstlane(P.roots[net]);
  if (root < 0 || root >= n) {
    if (tid == 0) atomicMin(P.bad, net);
    return;
  }
  // ---- 1. degrees (entry v+1 counts v's edges) ----
  for (int k = tid; k < n / 2 + 2; k += T) S.off[k] = 0u;
  for (int k = tid; k < (n + 1) / 2; k += T) S.rk[k] = 0u;
  if (tid == 0) {
    S.ctl[1] = 0;
    S.ctl[2] = 0;
  }
  __syncthreads();
  bool bad = false;
  for (int k = tid; k < ne; k += T) {
    const int s = e[2 * k], t = e[2 * k + 1];
    if (s < 0 || s >= n || t < 0 || t >= n || s == t) {
      bad = true;
    } else {
      add16(S.off, s + 1, 1);
      add16(S.off, t + 1, 1);
    }
  }
  if (bad) atomicOr(&S.ctl[1], 1);
  // ---- 2. offsets (the scan's barriers publish the failure flag too) ----
  block_scan16<T, false>(S.off, n + 1, S.part, tid);
  if (block_uniform<T>(&S.ctl[1])) {
    if (tid == 0) atomicMin(P.bad, net);
    return;
  }
  // ---- 3. adjacency slots and twins ----
  for (int k = tid; k < ne; k += T) {
    const int s = e[2 * k], t = e[2 * k + 1];
    const int js = ld16(S.off, s) + add16(S.rk, s, 1);
    const int jt = ld16(S.off, t) + add16(S.rk, t, 1);
    st16(S.adj, js, t);
    st16(S.adj, jt, s);
    st16(S.tw, js, jt);
    st16(S.tw, jt, js);
  }
  __syncthreads();
  // ---- 4. tour successors (a root without edges is no tree's root) ----
  const int h = __builtin_amdgcn_readfirstlane(ld16(S.off, root));
  if (__builtin_amdgcn_readfirstlane(ld16(S.off, root + 1)) == h) {
    if (tid == 0) atomicMin(P.bad, net);
    return;
  }
  for (int j = tid; j < A; j += T) {
    const int v = ld16(S.adj, j), t = ld16(S.tw, j);
    const int nxt = t + 1 < ld16(S.off, v + 1) ? t + 1 : ld16(S.off, v);
    st16(S.nx, j, nxt == h ? (int)kTourEnd : nxt);
  }
  __syncthreads();
  // ---- 5. list ranking ----
  int sh = 0;
  while ((T << sh) < A) ++sh;
  const int Sm = (1 << sh) - 1;
  const int Rb = (A + Sm) >> sh;
  const bool hr = (h & Sm) != 0;
  const int R = Rb + (hr ? 1 : 0);  // at most T + 1
  const int H = hr ? Rb : (h >> sh);
  auto is_ruler = [&](int k) { return (k & Sm) == 0 || k == h; };
  auto rid = [&](int k) { return (k & Sm) == 0 ? (k >> sh) : Rb; };
  auto rstart = [&](int r) { return r < Rb ? (r << sh) : h; };
  for (int r = tid; r < R; r += T) {
    int k = ld16(S.nx, rstart(r)), cnt = 1;
    for (int it = 0; it < A && k != (int)kTourEnd && !is_ruler(k); ++it) {
      ++cnt;
      k = ld16(S.nx, k);
    }
    S.rsuf[r] = cnt;
    S.rnext[r] = k == (int)kTourEnd ? -1 : is_ruler(k) ? rid(k) : -2;
  }
  __syncthreads();
  const int rounds = (32 - __clz(R)) + 1;
  for (int rd = 0; rd < rounds; ++rd) {
    int ns[2], nn[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int r = tid + u * T;
      ns[u] = 0;
      nn[u] = -1;
      if (r < R) {
        const int a = S.rnext[r];
        ns[u] = S.rsuf[r];
        nn[u] = a;
        if (a >= 0) {
          ns[u] += S.rsuf[a];
          nn[u] = S.rnext[a];
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int r = tid + u * T;
      if (r < R) {
        S.rsuf[r] = ns[u];
        S.rnext[r] = nn[u];
      }
    }
    __syncthreads();
  }
  // the tour from the root must hold every arc
  if (block_uniform<T>(&S.rsuf[H]) != A || block_uniform<T>(&S.rnext[H]) != -1) {
    if (tid == 0) atomicMin(P.bad, net);
    __syncthreads();  // the next net's set-up writes what this phase read
    return;
  }
  for (int r = tid; r < R; r += T) {
    int j = rstart(r), p = A - S.rsuf[r];
    st16(S.rk, j, p);
    int k = ld16(S.nx, j);
    for (int it = 0; it < A && k != (int)kTourEnd && !is_ruler(k); ++it) {
      st16(S.rk, k, ++p);
      k = ld16(S.nx, k);
    }
  }
  __syncthreads();
  // ---- 6. distances: +1 per descent, -1 per ascent, scanned over the tour ----
  for (int j = tid; j < A; j += T) {
    const int p = ld16(S.rk, j), q = ld16(S.rk, ld16(S.tw, j));
    st16(S.nx, p, p < q ? 1 : 0xffff);
  }
  for (int v = tid; v < n; v += T) st16(S.off, v, (int)kUnseen);
  block_scan16<T, true>(S.nx, A, S.part, tid);
  for (int j = tid; j < A; j += T) {
    const int p = ld16(S.rk, j), q = ld16(S.rk, ld16(S.tw, j));
    if (p < q) st16(S.off, ld16(S.adj, j), ld16(S.nx, p));
  }
  if (tid == 0) st16(S.off, root, 0);
  __syncthreads();
  {
    int mx = 0;
    bool miss = false;
    for (int v = tid; v < n; v += T) {
      const int d = ld16(S.off, v);
      if ((uint32_t)d == kUnseen) miss = true;
      else mx = max(mx, d);
    }
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) mx = max(mx, __shfl_xor(mx, d));
    if ((tid & (kWave - 1)) == 0) atomicMax(&S.ctl[2], mx);
    if (miss) atomicOr(&S.ctl[1], 1);
  }
  __syncthreads();
  if (block_uniform<T>(&S.ctl[1])) {
    if (tid == 0) atomicMin(P.bad, net);
    __syncthreads();
    return;
  }
  const int maxd = block_uniform<T>(&S.ctl[2]);
  // ---- 7. positions: sort (distance, CAP-1-id) ----
  {
    using Sort = rocprim::block_radix_sort<uint32_t, T, IPT>;
    static_assert(sizeof(typename Sort::storage_type) <= sizeof(uint32_t) * 4 * CAP,
                  "the sort's storage must fit in the arc arrays");
    auto& st = *reinterpret_cast<typename Sort::storage_type*>(S.adj);
    uint32_t key[IPT];
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
      const int i = tid * IPT + k;
      key[k] = i < n ? ((uint32_t)ld16(S.off, i) << LOG) | (uint32_t)(CAP - 1 - i) : 0xffffffffu;
    }
    const int end_bit = LOG + (32 - __clz(maxd)) + 1;
    Sort().sort(key, st, 0, end_bit);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
      const int pos = tid * IPT + k;
      if (pos < n) {
        const int id = CAP - 1 - (int)(key[k] & (uint32_t)(CAP - 1));
        st16(S.off, id, pos);
        P.node2bfs[off0 + id] = pos;
      }
    }
  }
  __syncthreads();
  // ---- 8. oriented edges, pid, R and C in BFS order ----
  if (tid == 0) {
    P.pid[off0] = 0;
    P.pres[off0] = 0.0f;
  }
  int32_t* eo = P.edges_out + 2LL * (off0 - net);
  const float* res = P.edge_res + (off0 - net);
  for (int k = tid; k < ne; k += T) {
    int s = e[2 * k], t = e[2 * k + 1];
    int ps = ld16(S.off, s), pt = ld16(S.off, t);
    if (ps > pt) {
      const int x = s; s = t; t = x;
      const int y = ps; ps = pt; pt = y;
    }
    eo[2 * k] = s;
    eo[2 * k + 1] = t;
    P.pid[off0 + pt] = pt - ps;
    P.pres[off0 + pt] = res[k];
  }
  for (int k = tid; k < n; k += T) P.cap_bfs[(int64_t)off0 + ld16(S.off, k)] = P.node_cap[(int64_t)off0 + k];
  __syncthreads();  // the next net reuses the LDS
}

template <int CAP, int T>
__global__ __launch_bounds__(T) void rct_bfs_tour_kernel(BfsParams P, const int32_t* list,
                                                         int32_t count, int32_t* head) {
  __shared__ TourLds<CAP, T> S;
  const int tid = threadIdx.x;
  for (;;) {
    if (tid == 0) S.ctl[0] = atomicAdd(head, 1);
    __syncthreads();
    const int idx = block_uniform<T>(&S.ctl[0]);
    __syncthreads();  // read before the next net's claim overwrites it
    if (idx >= count) break;
    tour_net<CAP, T>(P, list[idx], S, tid);
  }
}

// Size classes of the on-chip kernel, and the class of the nets larger than
// kMaxLds (global scratch).
constexpr int kClasses = 5;
constexpr int kBfsAux = otg_context::kAux;  // auxiliary streams used (small nets on the first)
constexpr int kClassCap[kClasses] = {1024, 2048, 4096, 8192, 0};

__device__ __forceinline__ int bfs_class(int n) {
  return n <= 1024 ? 0 : n <= 2048 ? 1 : n <= 4096 ? 2 : n <= kMaxLds ? 3 : 4;
}

// ctl: [0..4] nets per class, [5..9] list cursors, [10..14] queue heads,
// [16..17] nodes of class-4 nets (int64), [18] largest class-4 net
__global__ __launch_bounds__(256) void rct_bfs_count_kernel(BfsParams P, int32_t* ctl) {
  const int net = blockIdx.x * blockDim.x + threadIdx.x;
  if (net >= P.num_nets) return;
  const int n = P.nodes_start[net + 1] - P.nodes_start[net];
  if (n <= kSmallNodes) return;
  const int c = bfs_class(n);
  atomicAdd(&ctl[c], 1);
  if (c == kClasses - 1) {
    atomicAdd(reinterpret_cast<unsigned long long*>(ctl + 16), (unsigned long long)n);
    atomicMax(&ctl[18], n);
  }
}

struct BfsLists {
  int32_t* list;                // all large nets, class by class
  int32_t first[kClasses];      // start of each class's list
};

__global__ __launch_bounds__(256) void rct_bfs_list_kernel(BfsParams P, int32_t* ctl, BfsLists L) {
  const int net = blockIdx.x * blockDim.x + threadIdx.x;
  if (net >= P.num_nets) return;
  const int n = P.nodes_start[net + 1] - P.nodes_start[net];
  if (n <= kSmallNodes) return;
  const int c = bfs_class(n);
  L.list[L.first[c] + atomicAdd(&ctl[5 + c], 1)] = net;
}

template <int CAP>
__global__ __launch_bounds__(64) void rct_bfs_lds_kernel(BfsParams P, const int32_t* list,
                                                         int32_t count, int32_t* head) {
  __shared__ BfsLds16<CAP> S;
  const int lane = threadIdx.x;
  for (;;) {
    int idx = 0;
    if (lane == 0) idx = atomicAdd(head, 1);
    idx = __builtin_amdgcn_readfirstlane(__shfl(idx, 0));
    if (idx >= count) break;
    bfs_net16<CAP>(P, list[idx], S, lane);
  }
}

// Nets larger than kMaxLds: the global-scratch path, one scratch region of
// 6*maxn+2 ints per persistent wavefront.
__global__ __launch_bounds__(64) void rct_bfs_global_kernel(BfsParams P, const int32_t* list,
                                                            int32_t count, int32_t* head,
                                                            int64_t region) {
  __shared__ int32_t lds[6 * 512 + 2];
  const int lane = threadIdx.x;
  int32_t* mine = P.gscratch + region * blockIdx.x;
  for (;;) {
    int idx = 0;
    if (lane == 0) idx = atomicAdd(head, 1);
    idx = __builtin_amdgcn_readfirstlane(__shfl(idx, 0));
    if (idx >= count) break;
    bfs_net<true>(P, list[idx], mine, lds, 6 * 512 + 2, lane);
  }
}

__global__ void rct_bfs_zero_kernel(int32_t* p, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) p[j] = 0;
}

template <int CAP>
void* bfs_lds_ptr() {
  return reinterpret_cast<void*>(&rct_bfs_lds_kernel<CAP>);
}

template <int CAP, int T>
void* bfs_tour_ptr() {
  return reinterpret_cast<void*>(&rct_bfs_tour_kernel<CAP, T>);
}

}  // namespace

otg_status otg_rct_bfs_launch(otg_context* ctx, const otg_rct_edges* in, int32_t* node2bfs,
                              int32_t* pid, int32_t* edges_directed, float* pres, float* cap_bfs,
                              int32_t* d_bad, hipStream_t s) {
  BfsParams P{};
  P.num_nets = in->num_nets;
  P.nodes_start = in->nodes_start;
  P.edges_in = in->edges;
  P.edges_out = edges_directed;
  P.roots = in->roots;
  P.edge_res = in->edge_res;
  P.node_cap = reinterpret_cast<const float4*>(in->node_cap);
  P.node2bfs = node2bfs;
  P.pid = pid;
  P.pres = pres;
  P.cap_bfs = reinterpret_cast<float4*>(cap_bfs);
  P.bad = d_bad;
  const int N = in->num_nets;
  const int ncu = std::max(1, ctx->num_cus);
  // The size classes are independent (disjoint nets and output ranges): the
  // small nets run on an auxiliary stream forked from `s` while the large
  // nets are counted and listed, and the large classes each take a stream of
  // their own, so that one class's last nets do not leave the CUs idle until
  // the next class starts.  All of it is joined back into `s`.
  for (int k = 0; k < kBfsAux; ++k) {
    if (!ctx->aux[k]) {
      OTG_HIP(ctx, hipStreamCreateWithFlags(&ctx->aux[k], hipStreamNonBlocking));
      OTG_HIP(ctx, hipEventCreateWithFlags(&ctx->ev_join[k], hipEventDisableTiming));
    }
  }
  if (!ctx->ev_fork) OTG_HIP(ctx, hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
  OTG_HIP(ctx, hipEventRecord(ctx->ev_fork, s));
  for (int k = 0; k < kBfsAux; ++k) OTG_HIP(ctx, hipStreamWaitEvent(ctx->aux[k], ctx->ev_fork, 0));
  auto join = [&]() {
    hipError_t r = hipSuccess;
    for (int k = 0; k < kBfsAux; ++k) {
      const hipError_t e1 = hipEventRecord(ctx->ev_join[k], ctx->aux[k]);
      const hipError_t e2 = hipStreamWaitEvent(s, ctx->ev_join[k], 0);
      if (r == hipSuccess) r = e1 != hipSuccess ? e1 : e2;
    }
    return r;
  };
  // small nets: packed into wavefronts
  const int sgrid = std::max(1, std::min((N + kSmallRange - 1) / kSmallRange, ncu * 16));
  hipLaunchKernelGGL(rct_bfs_small_kernel, dim3(sgrid), dim3(64), 0, ctx->aux[0], P);
  {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      (void)join();
      OTG_HIP(ctx, e);
    }
  }
  // large nets: count by class, list, then one launch per class
  int32_t* ctl = nullptr;
  OTG_HIP(ctx, hipMallocAsync((void**)&ctl, sizeof(int32_t) * 32, s));
  hipLaunchKernelGGL(rct_bfs_zero_kernel, dim3(1), dim3(64), 0, s, ctl, 32);
  hipLaunchKernelGGL(rct_bfs_count_kernel, dim3((N + 255) / 256), dim3(256), 0, s, P, ctl);
  int32_t h[32] = {};
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(h, ctl, sizeof(h), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    (void)join();
    (void)hipFreeAsync(ctl, s);
    OTG_HIP(ctx, e);
  }
  int64_t total = 0;
  BfsLists L{};
  for (int c = 0; c < kClasses; ++c) {
    L.first[c] = (int32_t)total;
    total += h[c];
  }
  int64_t nodes4 = 0;
  std::memcpy(&nodes4, h + 16, sizeof(nodes4));
  const int maxn4 = h[18];
  int32_t* list = nullptr;
  int32_t* gs = nullptr;
  hipError_t le = hipSuccess;
  if (total > 0) {
    le = hipMallocAsync((void**)&list, sizeof(int32_t) * total, s);
    L.list = list;
    if (le == hipSuccess)
      hipLaunchKernelGGL(rct_bfs_list_kernel, dim3((N + 255) / 256), dim3(256), 0, s, P, ctl, L);
    // the lists are ready: the classes below the largest go to the other
    // auxiliary streams (largest first, round robin)
    if (le == hipSuccess) le = hipEventRecord(ctx->ev_fork, s);
    for (int k = 1; k < kBfsAux && le == hipSuccess; ++k)
      le = hipStreamWaitEvent(ctx->aux[k], ctx->ev_fork, 0);
    int launched = 0;
    // largest class first
    for (int c = kClasses - 2; c >= 0 && le == hipSuccess; --c) {
      if (h[c] <= 0) continue;
      // nets of up to 1024 nodes: one wavefront each; larger: the Euler
      // tour on a workgroup of CAP/8 threads
      void* fn = nullptr;
      int threads = kWave;
      switch (kClassCap[c]) {
        case 1024: fn = bfs_lds_ptr<1024>(); break;
        case 2048: fn = bfs_tour_ptr<2048, 256>(); threads = 256; break;
        case 4096: fn = bfs_tour_ptr<4096, 512>(); threads = 512; break;
        default: fn = bfs_tour_ptr<kMaxLds, 1024>(); threads = 1024; break;
      }
      int per_cu = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, 0) != hipSuccess ||
          per_cu <= 0)
        per_cu = 1;
      const int grid = std::max(1, std::min(h[c], per_cu * ncu));
      const int32_t* lc = list + L.first[c];
      int32_t cnt = h[c];
      int32_t* head = ctl + 10 + c;
      void* args[] = {&P, &lc, &cnt, &head};
      hipStream_t cs = launched == 0 ? s : ctx->aux[1 + (launched - 1) % (kBfsAux - 1)];
      le = hipLaunchKernel(fn, dim3(grid), dim3(threads), args, 0, cs);
      ++launched;
    }
    const int c4 = kClasses - 1;
    if (h[c4] > 0 && le == hipSuccess) {
      const int grid = std::max(1, std::min(h[c4], ncu * 2));
      const int64_t region = 6LL * maxn4 + 8;
      le = hipMallocAsync((void**)&gs, sizeof(int32_t) * region * grid, s);
      P.gscratch = gs;
      if (le == hipSuccess)
        hipLaunchKernelGGL(rct_bfs_global_kernel, dim3(grid), dim3(64), 0, s, P,
                           list + L.first[c4], h[c4], ctl + 10 + c4, region);
    }
    (void)nodes4;
  }
  if (le == hipSuccess) le = hipGetLastError();
  {
    const hipError_t je = join();
    if (le == hipSuccess) le = je;
  }
  if (gs) (void)hipFreeAsync(gs, s);
  if (list) (void)hipFreeAsync(list, s);
  (void)hipFreeAsync(ctl, s);
  OTG_HIP(ctx, le);
  return OTG_OK;
}
