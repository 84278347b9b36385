// Level-wise forward propagation of slew, arc delay and arrival time with
// NLDM table lookups (the forward tasks of the reference's GPU flow,
// ot/timer/timer.cpp:1079-1086).
//
// One launch per level, from the last level (the sources) to level 0, each
// reading its pin range from the device array of level ends (no host round
// trip between levels).  A thread takes one (pin, corner): corner c = el*2 +
// rf of the pin's slew and arrival, over every fanin arc and both input
// transitions, as Timer::_fprop_slew / _fprop_delay / _fprop_at do for a pin
// (timer.cpp:733-783, arc.cpp:82-147).  The fanin pins belong to later
// levels, whose launches have completed.  Arithmetic is the reference's,
// in FP32 with contraction off (division and square root correctly
// rounded), so results equal the restatement in the oracle bit for bit.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cfloat>
#include <cstdint>

#include "otg_internal.hpp"
#include "otg_prop.h"

namespace {

#pragma clang fp contract(off)

struct FpropParams {
  otg_prop_graph g;
  otg_prop_lib lib;
  const int32_t* frontiers;
  const int32_t* ends;
  otg_prop_result out;
  int32_t pool_words;  // lib.num_lut_vals (the LDS copy)
};

// Library tables, in LDS when they fit (see fprop_level_kernel) or in
// device memory.
struct LibView {
  const int32_t* hdr;
  const float* pool;
  const int32_t* tlut;
  const int32_t* sense;
  const int32_t* edge;
  const int32_t* has;
};

// Words of the LDS copy of a library: hdr 4L, tlut 8T, sense T, edge T,
// has 2T, pool.  Libraries beyond it are read from device memory.
constexpr int kLdsWords = 36 * 1024;  // 144 KiB of the 160 KiB per CU
constexpr int kChunk = 4;             // fanin arcs whose loads are issued together
constexpr int kUnrolledIdx = 8;       // index lengths searched without a loop

// The interpolation lambda of Lut::operator() (ot/liberty/lut.cpp:176-192):
// every case's value is formed with the same operations and the first case
// that holds is selected, without branches (lanes of a wavefront look up
// different tables, and a branch per case diverges).
__device__ __forceinline__ float lut_interpolate(float x, float x1, float x2, float y1, float y2) {
  const float slope = (y2 - y1) / (x2 - x1);
  const float below = y1 - (x1 - x) * slope;
  const float above = y2 + (x - x2) * slope;
  const float inside = y1 + (x - x1) * slope;
  float r = x == x2 ? y2 : inside;
  r = x == x1 ? y1 : r;
  r = x > x2 ? above : r;
  r = x < x1 ? below : r;
  return (x >= FLT_MAX || x <= -FLT_MAX) ? x : r;
}

// std::lower_bound over ascending indices (lut.cpp:196-197) as the count of
// entries below v: the same position for every v, NaN included (0), with
// independent loads instead of a chain of dependent ones.
__device__ __forceinline__ int count_below(const float* a, int n, float v) {
  int c = 0;
  if (n <= kUnrolledIdx) {
#pragma unroll
    for (int k = 0; k < kUnrolledIdx; ++k)
      if (k < n) c += a[k] < v ? 1 : 0;
  } else {
    for (int k = 0; k < n; ++k) c += a[k] < v ? 1 : 0;
  }
  return c;
}

// Lut::operator() (lut.cpp:161-241): clamp to [1, size-1], then the 1xN,
// Nx1 and NxN cases.
__device__ __forceinline__ float lut_eval(const LibView& L, int l, float val1, float val2) {
  const int n1 = L.hdr[4 * l], n2 = L.hdr[4 * l + 1], off = L.hdr[4 * l + 3];
  const float* i1 = L.pool + off;
  const float* i2 = i1 + n1;
  const float* tab = i2 + n2;
  if (n1 == 1 && n2 == 1) return tab[0];
  const int a1 = max(1, min(count_below(i1, n1, val1), n1 - 1));
  const int b1 = max(1, min(count_below(i2, n2, val2), n2 - 1));
  const int a0 = a1 - 1, b0 = b1 - 1;
  if (n1 == 1) return lut_interpolate(val2, i2[b0], i2[b1], tab[b0], tab[b1]);
  if (n2 == 1) return lut_interpolate(val1, i1[a0], i1[a1], tab[a0 * n2], tab[a1 * n2]);
  const float v0 = lut_interpolate(val1, i1[a0], i1[a1], tab[a0 * n2 + b0], tab[a1 * n2 + b0]);
  const float v1 = lut_interpolate(val1, i1[a0], i1[a1], tab[a0 * n2 + b1], tab[a1 * n2 + b1]);
  return lut_interpolate(val2, i2[b0], i2[b1], v0, v1);
}

// Table l (or -1) looked up with its template's operand order: var1 = 1
// puts the load first (Timing::delay / Timing::slew, timing.cpp:444-470).
__device__ __forceinline__ float table_value(const LibView& L, int l, float slew, float load) {
  if (l < 0) return __builtin_nanf("");
  return L.hdr[4 * l + 2] == 1 ? lut_eval(L, l, load, slew) : lut_eval(L, l, slew, load);
}

// table_value of table l at the slews s0 and s1 (the two input transitions)
// and one load, where ok0 / ok1: the header and the indices are read once
// for both.
__device__ __forceinline__ void table_pair(const LibView& L, int l, float s0, float s1,
                                           float load, bool ok0, bool ok1, float& r0, float& r1) {
  r0 = r1 = __builtin_nanf("");
  if (l < 0 || !(ok0 || ok1)) return;
  const int n1 = L.hdr[4 * l], n2 = L.hdr[4 * l + 1], var1 = L.hdr[4 * l + 2],
            off = L.hdr[4 * l + 3];
  const float* i1 = L.pool + off;
  const float* i2 = i1 + n1;
  const float* tab = i2 + n2;
  if (n1 == 1 && n2 == 1) {
    if (ok0) r0 = tab[0];
    if (ok1) r1 = tab[0];
    return;
  }
  if (n1 > kUnrolledIdx || n2 > kUnrolledIdx) {
    if (ok0) r0 = table_value(L, l, s0, load);
    if (ok1) r1 = table_value(L, l, s1, load);
    return;
  }
  float x1[kUnrolledIdx], x2[kUnrolledIdx];
#pragma unroll
  for (int k = 0; k < kUnrolledIdx; ++k) {
    x1[k] = k < n1 ? i1[k] : 0.0f;
    x2[k] = k < n2 ? i2[k] : 0.0f;
  }
  const bool lf = var1 == 1;
  // (both input transitions are looked up, and a transition that is not
  // defined gets NaN afterwards: lanes skipping one would diverge)
#pragma unroll
  for (int f = 0; f < 2; ++f) {
    const float sv = f ? s1 : s0;
    const float val1 = lf ? load : sv, val2 = lf ? sv : load;
    int c1 = 0, c2 = 0;
#pragma unroll
    for (int k = 0; k < kUnrolledIdx; ++k) {
      if (k < n1) c1 += x1[k] < val1 ? 1 : 0;
      if (k < n2) c2 += x2[k] < val2 ? 1 : 0;
    }
    const int a1 = max(1, min(c1, n1 - 1)), b1 = max(1, min(c2, n2 - 1));
    const int a0 = a1 - 1, b0 = b1 - 1;
    float r;
    if (n1 == 1) {
      r = lut_interpolate(val2, i2[b0], i2[b1], tab[b0], tab[b1]);
    } else if (n2 == 1) {
      r = lut_interpolate(val1, i1[a0], i1[a1], tab[a0 * n2], tab[a1 * n2]);
    } else {
      const float v0 =
          lut_interpolate(val1, i1[a0], i1[a1], tab[a0 * n2 + b0], tab[a1 * n2 + b0]);
      const float v1 =
          lut_interpolate(val1, i1[a0], i1[a1], tab[a0 * n2 + b1], tab[a1 * n2 + b1]);
      r = lut_interpolate(val2, i2[b0], i2[b1], v0, v1);
    }
    if (f) r1 = ok1 ? r : __builtin_nanf("");
    else r0 = ok0 ? r : __builtin_nanf("");
  }
}

// Timing::is_transition_defined (ot/liberty/timing.cpp:197-225).
__device__ __forceinline__ bool transition_defined(int sense, int edge, int frf, int trf) {
  if (edge == OTG_EDGE_RISING && frf != 0) return false;
  if (edge == OTG_EDGE_FALLING && frf != 1) return false;
  if (sense == OTG_SENSE_POSITIVE_UNATE && frf != trf) return false;
  if (sense == OTG_SENSE_NEGATIVE_UNATE && frf == trf) return false;
  return true;
}

// Pin::_relax_slew / _relax_at (ot/timer/pin.cpp:324-360).
__device__ __forceinline__ float relax(float cur, float v, int el) {
  if (__builtin_isnan(v)) return cur;
  if (__builtin_isnan(cur)) return v;
  return el == 0 ? (v < cur ? v : cur) : (v > cur ? v : cur);
}

// Slew and delay of one arc for corner (el, trf) from both input
// transitions f (so[f], d[f]; NaN where nothing propagates): Arc::_fprop_slew
// and Arc::_fprop_delay (arc.cpp:82-135) for one arc.
__device__ __forceinline__ void arc_candidates(const LibView& L, int kind, int t, float rd,
                                               float ri, const float su[2], float lc, int el,
                                               int trf, float so[2], float d[2]) {
  // every delay of the pin's fanin arcs is reset first (timer.cpp:757-759)
  so[0] = so[1] = d[0] = d[1] = __builtin_nanf("");
  if (kind == OTG_ARC_NET_RC) {
    // net arcs carry rf -> rf only (arc.cpp:91-98, 120-124)
    const float x = su[trf];
    if (!__builtin_isnan(x)) so[trf] = x < 0.0f ? -sqrtf(x * x + ri) : sqrtf(x * x + ri);
    d[trf] = rd;
  } else if (kind == OTG_ARC_NET_EMPTY) {
    so[trf] = su[trf];
    d[trf] = 0.0f;
  } else if (kind == OTG_ARC_CELL) {
    if (L.has[2 * t + el]) {
      const int sense = L.sense[t], edge = L.edge[t];
      const bool ok0 = !__builtin_isnan(su[0]) && transition_defined(sense, edge, 0, trf);
      const bool ok1 = !__builtin_isnan(su[1]) && transition_defined(sense, edge, 1, trf);
      const int ld = L.tlut[8 * t + 4 * el + trf], ls = L.tlut[8 * t + 4 * el + 2 + trf];
      table_pair(L, ls, su[0], su[1], lc, ok0, ok1, so[0], so[1]);
      table_pair(L, ld, su[0], su[1], lc, ok0, ok1, d[0], d[1]);
    }
  }
}

// Where a pin's fanin arcs and its sources' values are found: the caller's
// arrays in pin / arc order (Direct), or a plan's level-ordered copies
// (Planned).
struct Direct {
  const FpropParams& P;
  __device__ int first(int v) const { return P.g.fanin_start[v]; }
  __device__ int last(int v) const { return P.g.fanin_start[v + 1]; }
  __device__ void arc(int j, int c, int& kind, int& t, int& src, float& rd, float& ri) const {
    kind = P.g.arc_kind[j];
    t = P.g.arc_timing[j];
    src = P.g.fanin[j];
    rd = P.g.arc_rc_delay[4LL * j + c];
    ri = P.g.arc_rc_impulse[4LL * j + c];
  }
  __device__ void source(int u, int el, float su[2], float au[2]) const {
    su[0] = P.out.slew[4 * u + el * 2];
    su[1] = P.out.slew[4 * u + el * 2 + 1];
    au[0] = P.out.at[4 * u + el * 2];
    au[1] = P.out.at[4 * u + el * 2 + 1];
  }
  __device__ int64_t arc_id(int j) const { return j; }
};

// One (pin, corner) of a level: relax its slew and arrival over every fanin
// arc and input transition in arc order (the oracle's order), writing the
// arcs' delays.  Arcs are taken kChunk at a time: their own data, the
// source pins' values and the library entries are loaded for the whole
// chunk before the values are combined.
template <class Src>
__device__ __forceinline__ void fprop_one(const Src& S, const LibView& L, float* arc_delay,
                                          int v, int c, float& s, float& a, float lc) {
  const int el = c >> 1, trf = c & 1;
  const int j0 = S.first(v), j1 = S.last(v);
  for (int jb = j0; jb < j1; jb += kChunk) {
    int kind[kChunk], t[kChunk], u[kChunk];
    float rd[kChunk], ri[kChunk];
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
      const bool live = jb + q < j1;
      S.arc(live ? jb + q : j0, c, kind[q], t[q], u[q], rd[q], ri[q]);
      if (!live) kind[q] = OTG_ARC_SKIP;
    }
    float su[kChunk][2], au[kChunk][2];
#pragma unroll
    for (int q = 0; q < kChunk; ++q) S.source(u[q], el, su[q], au[q]);
    float so[kChunk][2], d[kChunk][2];
#pragma unroll
    for (int q = 0; q < kChunk; ++q)
      arc_candidates(L, kind[q], t[q], rd[q], ri[q], su[q], lc, el, trf, so[q], d[q]);
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
      if (jb + q >= j1) break;
      const int64_t j = S.arc_id(jb + q);
#pragma unroll
      for (int f = 0; f < 2; ++f) {
        s = relax(s, so[q][f], el);
        if (arc_delay) arc_delay[8 * j + 4 * el + 2 * f + trf] = d[q][f];
        if (!__builtin_isnan(d[q][f]) && !__builtin_isnan(au[q][f]))
          a = relax(a, d[q][f] + au[q][f], el);
      }
    }
  }
}

// Copies the library into LDS (the caller checked that it fits) and returns
// the view of it; every thread of the workgroup takes part.
// The library into LDS: the small arrays (table headers, timing entries)
// as one index space of 4-byte words, the table pool as 16-byte vectors when
// its address allows.  Every thread issues all its loads before its first
// store (a store waits for its load: a loop of load-store pairs would pay one
// memory latency per iteration), so a workgroup pays about one latency.
constexpr int kStageSmall = 4;  // small-array words per thread and round
constexpr int kStagePool = 8;   // pool vectors per thread and round

__device__ __forceinline__ LibView stage_library(const otg_prop_lib& lib, int pool_words,
                                                 int32_t* lds) {
  const int nl = lib.num_luts, nt = lib.num_timings;
  const int o_tlut = 4 * nl, o_sense = o_tlut + 8 * nt, o_edge = o_sense + nt,
            o_has = o_edge + nt, o_pool = o_has + 2 * nt;  // o_pool: a multiple of 4
  const int32_t* pool = reinterpret_cast<const int32_t*>(lib.lut_val);
  const bool vec = (reinterpret_cast<uintptr_t>(pool) & 15u) == 0;
  const int nvec = vec ? pool_words / 4 : 0;
  const int B = blockDim.x, tid = threadIdx.x;
  for (int sb = tid, pb = tid; sb < o_pool || pb < nvec; sb += kStageSmall * B, pb += kStagePool * B) {
    int32_t rs[kStageSmall];
    int4 rp[kStagePool];
#pragma unroll
    for (int u = 0; u < kStageSmall; ++u) {
      const int k = sb + u * B;
      rs[u] = 0;
      if (k < o_tlut) rs[u] = lib.lut_hdr[k];
      else if (k < o_sense) rs[u] = lib.timing_lut[k - o_tlut];
      else if (k < o_edge) rs[u] = lib.timing_sense[k - o_sense];
      else if (k < o_has) rs[u] = lib.timing_edge[k - o_edge];
      else if (k < o_pool) rs[u] = lib.timing_has[k - o_has];
    }
#pragma unroll
    for (int u = 0; u < kStagePool; ++u) {
      const int k = pb + u * B;
      rp[u] = k < nvec ? reinterpret_cast<const int4*>(pool)[k] : make_int4(0, 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < kStageSmall; ++u) {
      const int k = sb + u * B;
      if (k < o_pool) lds[k] = rs[u];
    }
#pragma unroll
    for (int u = 0; u < kStagePool; ++u) {
      const int k = pb + u * B;
      if (k < nvec) reinterpret_cast<int4*>(lds + o_pool)[k] = rp[u];
    }
  }
  // pool words past the vectors (all of them when the pool is not 16-byte
  // aligned)
  for (int k = 4 * nvec + tid; k < pool_words; k += B) lds[o_pool + k] = pool[k];
  __syncthreads();
  return LibView{lds, reinterpret_cast<const float*>(lds + o_pool), lds + o_tlut,
                 lds + o_sense, lds + o_edge, lds + o_has};
}

__device__ __forceinline__ LibView device_library(const otg_prop_lib& lib) {
  return LibView{lib.lut_hdr, lib.lut_val, lib.timing_lut, lib.timing_sense, lib.timing_edge,
                 lib.timing_has};
}

// One level in pin order: thread per (pin, corner), grid-stride over the
// level's pins.  kLds: every workgroup first copies the library into LDS (a
// workgroup with no pin of the level returns before that).
template <bool kLds>
__global__ __launch_bounds__(1024) void fprop_level_kernel(FpropParams P, int lv) {
  const int b = P.ends[lv], e = P.ends[lv + 1];
  const int64_t total = 4LL * (e - b);
  LibView L = device_library(P.lib);
  if constexpr (kLds) {
    __shared__ __attribute__((aligned(16))) int32_t lds[kLdsWords];
    if ((int64_t)blockIdx.x * blockDim.x >= total) return;
    L = stage_library(P.lib, P.pool_words, lds);
  }
  const Direct S{P};
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int v = P.frontiers[b + (int)(idx >> 2)], c = (int)(idx & 3);
    float s = P.g.pi_slew[4 * v + c], a = P.g.pi_at[4 * v + c];
    fprop_one(S, L, P.out.arc_delay, v, c, s, a, P.g.pin_load[4 * v + c]);
    P.out.slew[4 * v + c] = s;
    P.out.at[4 * v + c] = a;
  }
}

// ---------------------------------------------------------------------------
// Level plan: the frontiers' pins as positions (the pins of a level are
// consecutive), their fanin arcs as slots in that order, and the positions of
// every level cut into wavefront chunks; the level kernel takes a chunk's
// positions and slots and reads the caller's pin and arc arrays through them.
// ---------------------------------------------------------------------------

struct PlanView {
  const int32_t* frontiers;  // [n_lv] position -> pin
  const int32_t* ends;       // [levels+1]
  const int32_t* aofs;       // [n_lv+1] first slot of a position
  const int32_t* slot_src;   // [m_lv] the arc's source pin
  const int32_t* slot_arc;   // [m_lv] the arc's index in the caller's CSR
  const int32_t* slot_pos;   // [m_lv] position whose fanin arc the slot is
  const int2* wchunk;        // wavefront chunks: positions [x, y)
  const int32_t* wstart;     // [levels+1] first chunk of every level
};

struct PlanParams {
  PlanView pv;
  otg_prop_graph g;
  otg_prop_lib lib;
  otg_prop_result out;
  int32_t pool_words;
};

struct Planned {
  const PlanParams& P;
  __device__ int first(int q) const { return P.pv.aofs[q]; }
  __device__ int last(int q) const { return P.pv.aofs[q + 1]; }
  // slot k: arc j of the caller's CSR, its source pin u
  __device__ void arc(int64_t j, int c, int& kind, int& t, float& rd, float& ri) const {
    kind = P.g.arc_kind[j];
    t = P.g.arc_timing[j];
    rd = P.g.arc_rc_delay[4 * j + c];
    ri = P.g.arc_rc_impulse[4 * j + c];
  }
  __device__ void source(int u, int el, float su[2], float au[2]) const {
    su[0] = P.out.slew[4 * u + el * 2];
    su[1] = P.out.slew[4 * u + el * 2 + 1];
    au[0] = P.out.at[4 * u + el * 2];
    au[1] = P.out.at[4 * u + el * 2 + 1];
  }
};

// One level over the plan by wavefront chunks: a chunk is consecutive
// positions (at most 16) whose fanin arcs fill at most 16 slots (a position
// with more arcs is a chunk of its own, taken 16 slots at a time).  Lane
// 4*i + c computes the candidates of slot i for corner c (both input
// transitions) -- one pass of table work for the chunk's arcs -- and lane
// 4*p + c, the owner of position p, relaxes its slew and arrival over its
// slots in arc order, fetching the candidates by wavefront shuffles: the
// same operations in the same order as fprop_one.  Inputs are read from the
// caller's arrays through the plan's indices (the sink's load once, by its
// owner, and shuffled to the slots), and slew and arrival are written to the
// caller's arrays, where later levels read their sources.
template <bool kLds>
__global__ __launch_bounds__(1024) void fprop_wave_level_kernel(PlanParams P, int lv) {
  const int c0 = P.pv.wstart[lv], nch = P.pv.wstart[lv + 1] - c0;
  LibView L = device_library(P.lib);
  if constexpr (kLds) {
    __shared__ __attribute__((aligned(16))) int32_t lds[kLdsWords];
    if ((int)blockIdx.x >= nch) return;
    L = stage_library(P.lib, P.pool_words, lds);
  }
  const Planned S{P};
  const int lane = threadIdx.x & 63, c = lane & 3, el = c >> 1, trf = c & 1;
  const int twaves = gridDim.x * (blockDim.x >> 6);
  for (int w = blockIdx.x + gridDim.x * (threadIdx.x >> 6); w < nch; w += twaves) {
    const int2 qq = P.pv.wchunk[c0 + w];
    const int q0 = qq.x, q1 = qq.y;
    const int k0 = P.pv.aofs[q0], ns = P.pv.aofs[q1] - k0;
    const int pq = q0 + (lane >> 2);
    const bool owner = pq < q1;
    float s = __builtin_nanf(""), a = __builtin_nanf(""), lown = __builtin_nanf("");
    int ks = 0, ke = 0, v = 0;
    if (owner) {
      v = P.pv.frontiers[pq];
      s = P.g.pi_slew[4 * v + c];
      a = P.g.pi_at[4 * v + c];
      lown = P.g.pin_load[4 * v + c];
      ks = P.pv.aofs[pq] - k0;
      ke = P.pv.aofs[pq + 1] - k0;
    }
    for (int sb = 0; sb < ns; sb += 16) {
      float so[2] = {__builtin_nanf(""), __builtin_nanf("")};
      float ac[2] = {__builtin_nanf(""), __builtin_nanf("")};
      const int sl = sb + (lane >> 2);
      // the sink's load comes from the lane of its owner
      const int kp = sl < ns ? P.pv.slot_pos[k0 + sl] - q0 : 0;
      const float lc = __shfl(lown, 4 * kp + c);
      if (sl < ns) {
        const int k = k0 + sl;
        const int64_t j = P.pv.slot_arc[k];
        int kind, t;
        float rd, ri, su[2], au[2], d[2];
        S.arc(j, c, kind, t, rd, ri);
        S.source(P.pv.slot_src[k], el, su, au);
        arc_candidates(L, kind, t, rd, ri, su, lc, el, trf, so, d);
#pragma unroll
        for (int f = 0; f < 2; ++f) {
          if (P.out.arc_delay) P.out.arc_delay[8 * j + 4 * el + 2 * f + trf] = d[f];
          if (!__builtin_isnan(d[f]) && !__builtin_isnan(au[f])) ac[f] = d[f] + au[f];
        }
      }
      // each owner relaxes over its own slots of this round, in order,
      // fetching their candidates from the lanes that computed them; the
      // round takes as many steps as its owner with the most slots
      const int lo = max(ks, sb), hi = min(ke, sb + 16);
      int steps = owner && hi > lo ? hi - lo : 0;
#pragma unroll
      for (int m = 4; m < 64; m <<= 1) steps = max(steps, __shfl_xor(steps, m));
      for (int i = 0; i < steps; ++i) {
        const int from = 4 * min(lo + i - sb, 15) + c;
        const float s0 = __shfl(so[0], from), s1 = __shfl(so[1], from);
        const float a0 = __shfl(ac[0], from), a1 = __shfl(ac[1], from);
        if (lo + i < hi) {
          s = relax(s, s0, el);
          a = relax(a, a0, el);
          s = relax(s, s1, el);
          a = relax(a, a1, el);
        }
      }
    }
    if (owner) {
      P.out.slew[4 * v + c] = s;
      P.out.at[4 * v + c] = a;
    }
  }
}

// plan build: pin -> position (-1 outside the plan) and the fanin count of
// every position
__global__ void plan_fill_kernel(int32_t* a, int64_t n, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = v;
}

// A frontier pin out of range, or a fanin range outside the arcs, reports the
// position (bad[0], smallest); such a position gets no arcs.
__global__ void plan_pos_kernel(const int32_t* frontiers, int32_t n_lv, int32_t num_pins,
                                int32_t num_arcs, const int32_t* fanin_start, int32_t* deg,
                                int32_t* bad) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n_lv;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int v = frontiers[q];
    int d = 0;
    if (v < 0 || v >= num_pins) {
      atomicMin(bad, (int)q);
    } else {
      const int b = fanin_start[v], e = fanin_start[v + 1];
      if (b < 0 || e < b || e > num_arcs) atomicMin(bad, (int)q);
      else d = e - b;
    }
    deg[q] = d;
  }
}

// A fanin source out of range reports its arc (bad[1], smallest) and is read
// as pin 0 (the plan is then discarded).
__global__ void plan_slot_kernel(const int32_t* frontiers, int32_t n_lv, int32_t num_pins,
                                 const int32_t* fanin_start, const int32_t* fanin,
                                 const int32_t* aofs, int32_t* slot_src,
                                 int32_t* slot_arc, int32_t* slot_pos, int32_t* bad) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n_lv;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int k0 = aofs[q], dn = aofs[q + 1] - k0;
    if (dn == 0) continue;  // (also every position the previous kernel rejected)
    const int v = frontiers[q];
    const int j0 = fanin_start[v];
    for (int r = 0; r < dn; ++r) {
      const int j = j0 + r;
      int u = fanin[j];
      if (u < 0 || u >= num_pins) {
        atomicMin(bad + 1, j);
        u = 0;
      }
      slot_src[k0 + r] = u;
      slot_arc[k0 + r] = j;
      slot_pos[k0 + r] = (int)q;
    }
  }
}

// ---------------------------------------------------------------------------
// Backward propagation of required times (Timer::_bprop_rat,
// ot/timer/timer.cpp:812-841; Arc::_bprop_rat for propagation arcs,
// arc.cpp:150-175), levels first to last: thread per (pin, corner c =
// el*2 + frf), relaxing the primary-output value over the pin's fanout arcs
// in fanout order with rat(sink, el, trf) - delay(el, frf, trf) from the
// forward pass's arc delays (NaN where unset, so net arcs and loop breakers
// need no case of their own); Pin::_relax_rat (pin.cpp:362-380) keeps the
// larger value for MIN and the smaller for MAX.  Slack (pin.cpp:251-256):
// MIN at - rat, MAX rat - at.
// ---------------------------------------------------------------------------

struct BpropParams {
  otg_bprop_graph g;
  const int32_t* frontiers;
  const int32_t* ends;
  float* rat;
  float* slack;
};

__device__ __forceinline__ float relax_rat(float cur, float v, int el) {
  if (__builtin_isnan(v)) return cur;
  if (__builtin_isnan(cur)) return v;
  return el == 0 ? (v > cur ? v : cur) : (v < cur ? v : cur);
}

// A candidate of a pin's relaxation with its place in the sequential order
// (the primary-output value -1, then 2 * fanout position + trf): combining
// candidates keeps the better value and, between equal values (+0 and -0
// included), the earlier one, which is what the sequential relaxation keeps.
__device__ __forceinline__ void rat_pick(float& v, int& i, float v2, int i2, int el) {
  if (__builtin_isnan(v2)) return;
  const bool take = __builtin_isnan(v) || (el == 0 ? v2 > v : v2 < v) || (v2 == v && i2 < i);
  if (take) {
    v = v2;
    i = i2;
  }
}

constexpr int kBpropSmall = 8;  // fanout arcs a pin's own thread walks; more are queued
constexpr int kBpropMid = 64;   // fanout arcs a quarter wavefront takes in one round
constexpr int kBpropStep = 4;   // arcs whose loads a pin's own thread issues together
constexpr int kBpropWide = 16;  // arcs per arc lane whose loads a wavefront issues together
                                // (256 arcs per round: a fanout of 512 takes two rounds of
                                // three dependent loads; 12.8 -> 11.2 ms against 4 per lane)

// Levels first to last, one launch per level: workgroups of 256 threads take
// the level's (pin, corner) items in chunks of 256; a pin with more than
// kBpropSmall fanout arcs is queued in LDS by its corner-0 thread and taken,
// once the chunk is done, by a whole wavefront (16 arc lanes x 4 corners)
// if it has more than kBpropMid arcs, else by a quarter wavefront (4 arc
// lanes x 4 corners, four such pins per wavefront).
// The level's pins are dealt to the chunks transposed (chunk w takes pins
// w, w + B, w + 2B, ... of the level, B chunks of 64 pins), because the
// levelization appends pins of high fanout close together: in consecutive
// runs one workgroup held 30-37 of them and set the level's time.
// TO: the graph lists every fanout entry's sink (fanout_to), read beside the
// arc id instead of through it (a choice per launch, not per load: a load
// under a branch leaves the compiler unable to count what is in flight, and
// it then waits for everything before each dependent load).
// Queued pins of one workgroup, each taken by AL arc lanes x 4 corners (16 /
// AL pins per wavefront; kBpropWide arcs' loads per arc lane and round) with
// an order-preserving reduction over its arc lanes: bit for bit the
// sequential relaxation, +-0 ties included.  Wavefront w takes the groups
// first, first + stride, ...
template <bool TO, int AL>
__device__ __forceinline__ void bprop_queued(const BpropParams& P, const int3* q, int nq,
                                             int first, int stride, int lane) {
  constexpr int PPW = 16 / AL;
  const otg_bprop_graph& G = P.g;
  const int p = lane / (4 * AL), a0 = (lane >> 2) % AL, c = lane & 3;
  const int el = c >> 1, frf = c & 1;
  for (int g = first; g * PPW < nq; g += stride) {
    const int qi = g * PPW + p;
    const bool has = qi < nq;
    const int3 qe = has ? q[qi] : make_int3(0, 0, 0);
    const int u = qe.x, k0 = qe.y, k1 = qe.z;
    float r = (a0 == 0 && has) ? G.po_rat[4 * u + c] : __builtin_nanf("");
    int ri = -1;
    for (int kb = k0 + a0; kb < k1; kb += AL * kBpropWide) {
      int j[kBpropWide];
      float2 d[kBpropWide], rv[kBpropWide];
      int vz[kBpropWide];
#pragma unroll
      for (int z = 0; z < kBpropWide; ++z) {
        const int kk = kb + AL * z < k1 ? kb + AL * z : k0;
        j[z] = G.fanout_arc[kk];
        if constexpr (TO) vz[z] = G.fanout_to[kk];
      }
#pragma unroll
      for (int z = 0; z < kBpropWide; ++z) {
        int v;
        if constexpr (TO) v = vz[z];
        else v = G.arc_to[j[z]];
        d[z] = *reinterpret_cast<const float2*>(G.arc_delay + 8LL * j[z] + 4 * el + 2 * frf);
        rv[z] = *reinterpret_cast<const float2*>(P.rat + 4LL * v + 2 * el);
      }
#pragma unroll
      for (int z = 0; z < kBpropWide; ++z) {
        const int k = kb + AL * z;
        if (k >= k1) break;
        if (!__builtin_isnan(d[z].x) && !__builtin_isnan(rv[z].x))
          rat_pick(r, ri, rv[z].x - d[z].x, 2 * (k - k0), el);
        if (!__builtin_isnan(d[z].y) && !__builtin_isnan(rv[z].y))
          rat_pick(r, ri, rv[z].y - d[z].y, 2 * (k - k0) + 1, el);
      }
    }
#pragma unroll
    for (int m = 4; m < 4 * AL; m <<= 1) {
      const float r2 = __shfl_xor(r, m);
      const int i2 = __shfl_xor(ri, m);
      rat_pick(r, ri, r2, i2, el);
    }
    if (a0 == 0 && has) {
      P.rat[4 * u + c] = r;
      if (P.slack) {
        const float a = G.at[4 * u + c];
        P.slack[4 * u + c] = (__builtin_isnan(a) || __builtin_isnan(r))
                                 ? __builtin_nanf("")
                                 : (el == 0 ? a - r : r - a);
      }
    }
  }
}

template <bool TO>
__global__ __launch_bounds__(256) void bprop_level_kernel(BpropParams P, int lv) {
  __shared__ int3 queue[64];   // (pin, first fanout entry, end): more than kBpropMid arcs
  __shared__ int3 mqueue[64];  // kBpropSmall < arcs <= kBpropMid
  __shared__ int32_t qn, mqn;
  const int b = P.ends[lv], e = P.ends[lv + 1];
  const int64_t nblk = ((int64_t)(e - b) + 63) / 64;
  const int64_t total = 4 * 64 * nblk;
  const otg_bprop_graph& G = P.g;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < total;
       base += (int64_t)gridDim.x * blockDim.x) {
    if (threadIdx.x == 0) qn = mqn = 0;
    __syncthreads();
    const int64_t idx = base + threadIdx.x;
    const int64_t slot = idx >> 2;
    const int64_t pos = (slot & 63) * nblk + (slot >> 6);
    if (pos < e - b) {
      const int u = P.frontiers[b + (int)pos], c = (int)(idx & 3);
      const int el = c >> 1, frf = c & 1;
      const int k0 = G.fanout_start[u], k1 = G.fanout_start[u + 1];
      if (k1 - k0 > kBpropMid) {
        if (c == 0) queue[atomicAdd(&qn, 1)] = make_int3(u, k0, k1);
      } else if (k1 - k0 > kBpropSmall) {
        if (c == 0) mqueue[atomicAdd(&mqn, 1)] = make_int3(u, k0, k1);
      } else {
        float r = G.po_rat[4 * u + c];
        // kBpropStep arcs per step: their ids, then sinks and delays, then
        // the sinks' required times are loaded for all of them before the
        // relaxation in fanout order
        for (int k = k0; k < k1; k += kBpropStep) {
          int j[kBpropStep];
          float2 d[kBpropStep], rv[kBpropStep];
          int vz[kBpropStep];
#pragma unroll
          for (int z = 0; z < kBpropStep; ++z) {
            const int kk = k + z < k1 ? k + z : k0;
            j[z] = G.fanout_arc[kk];
            if constexpr (TO) vz[z] = G.fanout_to[kk];
          }
#pragma unroll
          for (int z = 0; z < kBpropStep; ++z) {
            // (past the pin's last arc these load its first arc again: loads
            // without a branch around them)
            int v;
            if constexpr (TO) v = vz[z];
            else v = G.arc_to[j[z]];
            d[z] = *reinterpret_cast<const float2*>(G.arc_delay + 8LL * j[z] + 4 * el + 2 * frf);
            rv[z] = *reinterpret_cast<const float2*>(P.rat + 4LL * v + 2 * el);
          }
#pragma unroll
          for (int z = 0; z < kBpropStep; ++z) {
            if (k + z >= k1) break;
            if (!__builtin_isnan(d[z].x) && !__builtin_isnan(rv[z].x))
              r = relax_rat(r, rv[z].x - d[z].x, el);
            if (!__builtin_isnan(d[z].y) && !__builtin_isnan(rv[z].y))
              r = relax_rat(r, rv[z].y - d[z].y, el);
          }
        }
        P.rat[4 * u + c] = r;
        if (P.slack) {
          const float a = G.at[4 * u + c];
          P.slack[4 * u + c] = (__builtin_isnan(a) || __builtin_isnan(r))
                                   ? __builtin_nanf("")
                                   : (el == 0 ? a - r : r - a);
        }
      }
    }
    __syncthreads();
    // the pins of highest fanout one per wavefront, then the others four per
    // wavefront, starting after the wavefronts that took one of the former
    const int nq = qn, nm = mqn;
    bprop_queued<TO, 16>(P, queue, nq, wave, nw, lane);
    bprop_queued<TO, 4>(P, mqueue, nm, (wave - nq % nw + nw) % nw, nw, lane);
    __syncthreads();
  }
}

}  // namespace

struct otg_fprop_plan {
  otg_context* ctx = nullptr;
  int32_t num_pins = 0, num_arcs = 0, n_lv = 0, m_lv = 0, num_levels = 0;
  int32_t* frontiers = nullptr;
  int32_t* ends = nullptr;
  int32_t* aofs = nullptr;
  int32_t* slot_src = nullptr;
  int32_t* slot_arc = nullptr;
  int32_t* slot_pos = nullptr;
  int2* wchunk = nullptr;
  int32_t* wstart = nullptr;
  int32_t num_chunks = 0;
};

namespace {

void plan_free(otg_fprop_plan* p) {
  for (void* x : {(void*)p->frontiers, (void*)p->ends, (void*)p->aofs, (void*)p->slot_src,
                  (void*)p->slot_arc, (void*)p->slot_pos, (void*)p->wchunk, (void*)p->wstart})
    if (x) (void)hipFree(x);
  delete p;
}

int grid_for(const otg_context* ctx, int64_t n, int block) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + block - 1) / block, ctx->num_cus * 16LL));
}

bool lib_ok(const otg_prop_lib* lib) {
  return lib->num_luts >= 0 && lib->num_timings >= 0 && lib->num_lut_vals >= 0 &&
         !(lib->num_timings > 0 && (!lib->timing_sense || !lib->timing_edge ||
                                    !lib->timing_lut || !lib->timing_has)) &&
         !(lib->num_luts > 0 && (!lib->lut_hdr || !lib->lut_val || lib->num_lut_vals < 1));
}

bool lib_in_lds(const otg_prop_lib* lib) {
  return 4LL * lib->num_luts + 12LL * lib->num_timings + lib->num_lut_vals <= kLdsWords;
}

}  // namespace

extern "C" otg_status otg_fprop_plan_create(otg_context* ctx, int32_t num_pins, int32_t num_arcs,
                                            const int32_t* fanin_start, const int32_t* fanin,
                                            const int32_t* frontiers,
                                            const int32_t* frontier_ends, int32_t num_levels,
                                            otg_stream stream, otg_fprop_plan** plan) {
  if (!ctx || !plan || num_pins < 0 || num_arcs < 0 || num_levels < 0 ||
      (num_pins > 0 && !fanin_start) || (num_arcs > 0 && !fanin) ||
      (num_levels > 0 && (!frontiers || !frontier_ends)))
    return OTG_ERR_INVALID_ARGUMENT;
  *plan = nullptr;
  OTG_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = otg_pick_stream(ctx, stream);
  auto* p = new otg_fprop_plan();
  p->ctx = ctx;
  p->num_pins = num_pins;
  p->num_arcs = num_arcs;
  p->num_levels = num_levels;
  int32_t* deg = nullptr;
  int32_t* bad = nullptr;
  void* tmp = nullptr;
  auto fail = [&](hipError_t e) {
    (void)hipStreamSynchronize(s);
    if (deg) (void)hipFree(deg);
    if (bad) (void)hipFree(bad);
    if (tmp) (void)hipFree(tmp);
    plan_free(p);
    otg_set_error(ctx, "otg_fprop_plan_create: %s", hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? OTG_ERR_OUT_OF_MEMORY : OTG_ERR_DEVICE;
  };
  hipError_t e;
  int32_t n_lv = 0;
  // the level ends, on the host: from 0, never decreasing, at most num_pins
  std::vector<int32_t> ends(num_levels + 1, 0);
  if (num_levels > 0) {
    if ((e = hipMemcpyAsync(ends.data(), frontier_ends, sizeof(int32_t) * (num_levels + 1),
                            hipMemcpyDeviceToHost, s)) ||
        (e = hipStreamSynchronize(s)))
      return fail(e);
    int32_t bad_lv = ends[0] != 0 ? 0 : -1;
    for (int32_t lv = 0; lv < num_levels && bad_lv < 0; ++lv)
      if (ends[lv + 1] < ends[lv] || ends[lv + 1] > num_pins) bad_lv = lv + 1;
    if (bad_lv >= 0) {
      plan_free(p);
      otg_set_error(ctx, "otg_fprop_plan_create: level end %d is %d (ends must run from 0 "
                         "without decreasing to at most %d pins)", bad_lv, ends[bad_lv], num_pins);
      return OTG_ERR_INVALID_ARGUMENT;
    }
    n_lv = ends[num_levels];
  }
  p->n_lv = n_lv;
  const size_t I = sizeof(int32_t);
  if ((e = hipMalloc(&p->frontiers, I * std::max(n_lv, 1))) ||
      (e = hipMalloc(&p->ends, I * (num_levels + 1))) ||
      (e = hipMalloc(&p->aofs, I * (n_lv + 1))) ||
      (e = hipMalloc(&deg, I * std::max(n_lv, 1))) || (e = hipMalloc(&bad, I * 2)))
    return fail(e);
  hipLaunchKernelGGL(plan_fill_kernel, dim3(1), dim3(64), 0, s, bad, (int64_t)2, INT32_MAX);
  if (n_lv > 0 &&
      (e = hipMemcpyAsync(p->frontiers, frontiers, I * n_lv, hipMemcpyDeviceToDevice, s)))
    return fail(e);
  if (num_levels > 0 &&
      (e = hipMemcpyAsync(p->ends, frontier_ends, I * (num_levels + 1), hipMemcpyDeviceToDevice, s)))
    return fail(e);
  hipLaunchKernelGGL(plan_fill_kernel, dim3(1), dim3(64), 0, s, p->aofs, (int64_t)1, 0);
  if (n_lv > 0) {
    hipLaunchKernelGGL(plan_pos_kernel, dim3(grid_for(ctx, n_lv, 256)), dim3(256), 0, s,
                       p->frontiers, n_lv, num_pins, num_arcs, fanin_start, deg, bad);
    size_t tb = 0;
    if ((e = hipcub::DeviceScan::InclusiveSum(nullptr, tb, deg, p->aofs + 1, n_lv, s)) ||
        (e = hipMalloc(&tmp, std::max<size_t>(tb, 1))) ||
        (e = hipcub::DeviceScan::InclusiveSum(tmp, tb, deg, p->aofs + 1, n_lv, s)))
      return fail(e);
  }
  int32_t m_lv = 0;
  if ((e = hipMemcpyAsync(&m_lv, p->aofs + n_lv, I, hipMemcpyDeviceToHost, s)) ||
      (e = hipStreamSynchronize(s)) || (e = hipGetLastError()))
    return fail(e);
  p->m_lv = m_lv;
  if ((e = hipMalloc(&p->slot_src, I * std::max(m_lv, 1))) ||
      (e = hipMalloc(&p->slot_arc, I * std::max(m_lv, 1))) ||
      (e = hipMalloc(&p->slot_pos, I * std::max(m_lv, 1))))
    return fail(e);
  if (n_lv > 0)
    hipLaunchKernelGGL(plan_slot_kernel, dim3(grid_for(ctx, n_lv, 256)), dim3(256), 0, s,
                       p->frontiers, n_lv, num_pins, fanin_start, fanin, p->aofs,
                       p->slot_src, p->slot_arc, p->slot_pos, bad);
  // wavefront chunks, level by level: consecutive positions, at most 16 of
  // them and at most 16 arc slots (a position with more slots alone)
  {
    std::vector<int32_t> aofs(n_lv + 1);
    int32_t hbad[2] = {INT32_MAX, INT32_MAX};
    if ((e = hipMemcpyAsync(aofs.data(), p->aofs, I * (n_lv + 1), hipMemcpyDeviceToHost, s)) ||
        (e = hipMemcpyAsync(hbad, bad, I * 2, hipMemcpyDeviceToHost, s)) ||
        (e = hipStreamSynchronize(s)))
      return fail(e);
    if (hbad[0] != INT32_MAX || hbad[1] != INT32_MAX) {
      (void)hipFree(deg);
      (void)hipFree(bad);
      if (tmp) (void)hipFree(tmp);
      plan_free(p);
      if (hbad[0] != INT32_MAX)
        otg_set_error(ctx, "otg_fprop_plan_create: frontier position %d names a pin out of range "
                           "or a fanin range outside the arcs", hbad[0]);
      else
        otg_set_error(ctx, "otg_fprop_plan_create: arc %d comes from a pin out of range", hbad[1]);
      return OTG_ERR_INVALID_ARGUMENT;
    }
    std::vector<int2> chunks;
    std::vector<int32_t> wstart(num_levels + 1, 0);
    chunks.reserve(n_lv / 4 + num_levels + 1);
    for (int32_t lv = 0; lv < num_levels; ++lv) {
      wstart[lv] = (int32_t)chunks.size();
      int32_t q = ends[lv];
      while (q < ends[lv + 1]) {
        int32_t q1 = q + 1;
        while (q1 < ends[lv + 1] && q1 - q < 16 && aofs[q1 + 1] - aofs[q] <= 16) ++q1;
        chunks.push_back(make_int2(q, q1));
        q = q1;
      }
    }
    wstart[num_levels] = (int32_t)chunks.size();
    p->num_chunks = (int32_t)chunks.size();
    if ((e = hipMalloc(&p->wchunk, sizeof(int2) * std::max<size_t>(chunks.size(), 1))) ||
        (e = hipMalloc(&p->wstart, I * (num_levels + 1))) ||
        (!chunks.empty() && (e = hipMemcpyAsync(p->wchunk, chunks.data(),
                                                sizeof(int2) * chunks.size(),
                                                hipMemcpyHostToDevice, s))) ||
        (e = hipMemcpyAsync(p->wstart, wstart.data(), I * (num_levels + 1),
                            hipMemcpyHostToDevice, s)))
      return fail(e);
  }
  if ((e = hipStreamSynchronize(s)) || (e = hipGetLastError())) return fail(e);
  (void)hipFree(deg);
  (void)hipFree(bad);
  if (tmp) (void)hipFree(tmp);
  *plan = p;
  return OTG_OK;
}

extern "C" otg_status otg_fprop_plan_destroy(otg_fprop_plan* plan) {
  if (!plan) return OTG_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(plan->ctx->device);
  plan_free(plan);
  return OTG_OK;
}

extern "C" otg_status otg_fprop_plan_sizes(const otg_fprop_plan* plan, int32_t* positions,
                                           int32_t* slots) {
  if (!plan) return OTG_ERR_INVALID_ARGUMENT;
  if (positions) *positions = plan->n_lv;
  if (slots) *slots = plan->m_lv;
  return OTG_OK;
}

extern "C" otg_status otg_fprop_run(otg_context* ctx, otg_fprop_plan* plan,
                                    const otg_prop_graph* graph, const otg_prop_lib* lib,
                                    const otg_prop_result* out, otg_stream stream) {
  if (!ctx || !plan || !graph || !lib || !out || !lib_ok(lib) || plan->ctx != ctx)
    return OTG_ERR_INVALID_ARGUMENT;
  if (graph->num_pins != plan->num_pins || graph->num_arcs != plan->num_arcs) {
    otg_set_error(ctx, "otg_fprop_run: graph of %d pins / %d arcs, plan of %d / %d",
                  graph->num_pins, graph->num_arcs, plan->num_pins, plan->num_arcs);
    return OTG_ERR_INVALID_ARGUMENT;
  }
  if (plan->n_lv == 0) return OTG_OK;
  if (!out->slew || !out->at || !graph->pin_load || !graph->pi_slew || !graph->pi_at ||
      (plan->m_lv > 0 && (!graph->arc_kind || !graph->arc_timing || !graph->arc_rc_delay ||
                          !graph->arc_rc_impulse)))
    return OTG_ERR_INVALID_ARGUMENT;
  OTG_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = otg_pick_stream(ctx, stream);
  PlanView pv{plan->frontiers, plan->ends,     plan->aofs,   plan->slot_src,
              plan->slot_arc,  plan->slot_pos, plan->wchunk, plan->wstart};
  PlanParams P{pv, *graph, *lib, *out, lib->num_lut_vals};
  if (lib_in_lds(lib)) {
    for (int lv = plan->num_levels - 1; lv >= 0; --lv)
      hipLaunchKernelGGL(fprop_wave_level_kernel<true>, dim3(std::max(1, ctx->num_cus)),
                         dim3(1024), 0, s, P, lv);
  } else {
    for (int lv = plan->num_levels - 1; lv >= 0; --lv)
      hipLaunchKernelGGL(fprop_wave_level_kernel<false>, dim3(std::max(1, ctx->num_cus * 2)),
                         dim3(1024), 0, s, P, lv);
  }
  OTG_HIP(ctx, hipGetLastError());
  return OTG_OK;
}

extern "C" otg_status otg_fprop(otg_context* ctx, const otg_prop_graph* graph,
                                const otg_prop_lib* lib, const int32_t* frontiers,
                                const int32_t* frontier_ends, int32_t num_levels,
                                const otg_prop_result* out, otg_stream stream) {
  if (!ctx || !graph || !lib || !out || num_levels < 0 || graph->num_pins < 0 ||
      graph->num_arcs < 0 || lib->num_luts < 0 || lib->num_timings < 0 || lib->num_lut_vals < 0)
    return OTG_ERR_INVALID_ARGUMENT;
  if (num_levels == 0 || graph->num_pins == 0) return OTG_OK;
  if (!frontiers || !frontier_ends || !graph->fanin_start || !out->slew || !out->at ||
      !graph->pin_load || !graph->pi_slew || !graph->pi_at ||
      (graph->num_arcs > 0 && (!graph->fanin || !graph->arc_kind || !graph->arc_timing ||
                               !graph->arc_rc_delay || !graph->arc_rc_impulse)) ||
      (lib->num_timings > 0 && (!lib->timing_sense || !lib->timing_edge || !lib->timing_lut ||
                                !lib->timing_has)) ||
      (lib->num_luts > 0 && (!lib->lut_hdr || !lib->lut_val || lib->num_lut_vals < 1)))
    return OTG_ERR_INVALID_ARGUMENT;
  OTG_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = otg_pick_stream(ctx, stream);
  FpropParams P{*graph, *lib, frontiers, frontier_ends, *out, (int32_t)lib->num_lut_vals};
  // a level's pins are not known on the host: a fixed grid strides over them
  if (lib_in_lds(lib)) {
    const int grid = std::max(1, ctx->num_cus);
    for (int lv = num_levels - 1; lv >= 0; --lv)
      hipLaunchKernelGGL(fprop_level_kernel<true>, dim3(grid), dim3(1024), 0, s, P, lv);
  } else {
    const int grid = std::max(1, ctx->num_cus * 2);
    for (int lv = num_levels - 1; lv >= 0; --lv)
      hipLaunchKernelGGL(fprop_level_kernel<false>, dim3(grid), dim3(1024), 0, s, P, lv);
  }
  OTG_HIP(ctx, hipGetLastError());
  return OTG_OK;
}

extern "C" otg_status otg_bprop(otg_context* ctx, const otg_bprop_graph* graph,
                                const int32_t* frontiers, const int32_t* frontier_ends,
                                int32_t num_levels, float* rat, float* slack, otg_stream stream) {
  if (!ctx || !graph || num_levels < 0 || graph->num_pins < 0 || graph->num_arcs < 0)
    return OTG_ERR_INVALID_ARGUMENT;
  if (num_levels == 0 || graph->num_pins == 0) return OTG_OK;
  if (!frontiers || !frontier_ends || !rat || !graph->fanout_start || !graph->po_rat ||
      (slack && !graph->at) ||
      (graph->num_arcs > 0 && (!graph->fanout_arc || !graph->arc_to || !graph->arc_delay)) ||
      ((uintptr_t)graph->arc_delay & 7) || ((uintptr_t)rat & 7))
    return OTG_ERR_INVALID_ARGUMENT;
  OTG_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = otg_pick_stream(ctx, stream);
  BpropParams P{*graph, frontiers, frontier_ends, rat, slack};
  const int grid = std::max(1, ctx->num_cus * 8);
  for (int lv = 0; lv < num_levels; ++lv) {
    if (graph->fanout_to) hipLaunchKernelGGL(bprop_level_kernel<true>, dim3(grid), dim3(256), 0, s, P, lv);
    else hipLaunchKernelGGL(bprop_level_kernel<false>, dim3(grid), dim3(256), 0, s, P, lv);
  }
  OTG_HIP(ctx, hipGetLastError());
  return OTG_OK;
}
