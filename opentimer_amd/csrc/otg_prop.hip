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

#include <algorithm>
#include <cfloat>

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

// The interpolation lambda of Lut::operator() (ot/liberty/lut.cpp:176-192).
__device__ __forceinline__ float lut_interpolate(float x, float x1, float x2, float y1, float y2) {
  if (x >= FLT_MAX || x <= -FLT_MAX) return x;
  const float slope = (y2 - y1) / (x2 - x1);
  if (x < x1) return y1 - (x1 - x) * slope;
  else if (x > x2) return y2 + (x - x2) * slope;
  else if (x == x1) return y1;
  else if (x == x2) return y2;
  else return y1 + (x - x1) * slope;
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

// One (pin, corner) of a level: relax its slew and arrival over every fanin
// arc and input transition in arc order (the oracle's order), writing the
// arcs' delays.  Arcs are taken kChunk at a time: their own data, the
// source pins' values and the library entries are loaded for the whole
// chunk before the values are combined.
__device__ __forceinline__ void fprop_pin(const FpropParams& P, const LibView& L, int v, int c) {
  const otg_prop_graph& G = P.g;
  const int el = c >> 1, trf = c & 1;
  float s = G.pi_slew[4 * v + c], a = G.pi_at[4 * v + c];
  const float lc = G.pin_load[4 * v + c];
  const int j0 = G.fanin_start[v], j1 = G.fanin_start[v + 1];
  for (int jb = j0; jb < j1; jb += kChunk) {
    int kind[kChunk], t[kChunk], u[kChunk];
    float rd[kChunk], ri[kChunk];
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
      const bool live = jb + q < j1;
      const int j = live ? jb + q : j0;
      kind[q] = live ? G.arc_kind[j] : OTG_ARC_SKIP;
      u[q] = G.fanin[j];
      t[q] = G.arc_timing[j];
      rd[q] = G.arc_rc_delay[4LL * j + c];
      ri[q] = G.arc_rc_impulse[4LL * j + c];
    }
    float su[kChunk][2], au[kChunk][2];
#pragma unroll
    for (int q = 0; q < kChunk; ++q)
#pragma unroll
      for (int f = 0; f < 2; ++f) {
        su[q][f] = P.out.slew[4 * u[q] + el * 2 + f];
        au[q][f] = P.out.at[4 * u[q] + el * 2 + f];
      }
    float so[kChunk][2], d[kChunk][2];
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
      // every delay of the pin's fanin arcs is reset first (timer.cpp:757-759)
#pragma unroll
      for (int f = 0; f < 2; ++f) so[q][f] = d[q][f] = __builtin_nanf("");
      if (kind[q] == OTG_ARC_NET_RC) {
        // net arcs carry rf -> rf only (arc.cpp:91-98, 120-124)
        const float x = su[q][trf];
        if (!__builtin_isnan(x)) so[q][trf] = x < 0.0f ? -sqrtf(x * x + ri[q]) : sqrtf(x * x + ri[q]);
        d[q][trf] = rd[q];
      } else if (kind[q] == OTG_ARC_NET_EMPTY) {
        so[q][trf] = su[q][trf];
        d[q][trf] = 0.0f;
      } else if (kind[q] == OTG_ARC_CELL) {
        const int tt = t[q];
        if (L.has[2 * tt + el]) {
          const int sense = L.sense[tt], edge = L.edge[tt];
          const int ld = L.tlut[8 * tt + 4 * el + trf], ls = L.tlut[8 * tt + 4 * el + 2 + trf];
#pragma unroll
          for (int f = 0; f < 2; ++f) {
            if (!__builtin_isnan(su[q][f]) && transition_defined(sense, edge, f, trf)) {
              so[q][f] = table_value(L, ls, su[q][f], lc);
              d[q][f] = table_value(L, ld, su[q][f], lc);
            }
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
      if (jb + q >= j1) break;
#pragma unroll
      for (int f = 0; f < 2; ++f) {
        s = relax(s, so[q][f], el);
        if (P.out.arc_delay) P.out.arc_delay[8LL * (jb + q) + 4 * el + 2 * f + trf] = d[q][f];
        if (!__builtin_isnan(d[q][f]) && !__builtin_isnan(au[q][f])) a = relax(a, d[q][f] + au[q][f], el);
      }
    }
  }
  P.out.slew[4 * v + c] = s;
  P.out.at[4 * v + c] = a;
}

// One level: thread per (pin, corner), grid-stride over the level's pins.
// kLds: every workgroup first copies the library into LDS (a workgroup with
// no pin of the level returns before that).
template <bool kLds>
__global__ __launch_bounds__(1024) void fprop_level_kernel(FpropParams P, int lv) {
  const int b = P.ends[lv], e = P.ends[lv + 1];
  const int64_t total = 4LL * (e - b);
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  LibView L{P.lib.lut_hdr, P.lib.lut_val, P.lib.timing_lut, P.lib.timing_sense,
            P.lib.timing_edge, P.lib.timing_has};
  if constexpr (kLds) {
    __shared__ int32_t lds[kLdsWords];
    if ((int64_t)blockIdx.x * blockDim.x >= total) return;
    const int nl = P.lib.num_luts, nt = P.lib.num_timings;
    const int o_tlut = 4 * nl, o_sense = o_tlut + 8 * nt, o_edge = o_sense + nt,
              o_has = o_edge + nt, o_pool = o_has + 2 * nt;
    for (int k = threadIdx.x; k < 4 * nl; k += blockDim.x) lds[k] = P.lib.lut_hdr[k];
    for (int k = threadIdx.x; k < 8 * nt; k += blockDim.x) lds[o_tlut + k] = P.lib.timing_lut[k];
    for (int k = threadIdx.x; k < nt; k += blockDim.x) {
      lds[o_sense + k] = P.lib.timing_sense[k];
      lds[o_edge + k] = P.lib.timing_edge[k];
    }
    for (int k = threadIdx.x; k < 2 * nt; k += blockDim.x) lds[o_has + k] = P.lib.timing_has[k];
    const int32_t* pool = reinterpret_cast<const int32_t*>(P.lib.lut_val);
    for (int k = threadIdx.x; k < P.pool_words; k += blockDim.x) lds[o_pool + k] = pool[k];
    __syncthreads();
    L = LibView{lds, reinterpret_cast<const float*>(lds + o_pool), lds + o_tlut, lds + o_sense,
                lds + o_edge, lds + o_has};
  }
  for (; idx < total; idx += (int64_t)gridDim.x * blockDim.x)
    fprop_pin(P, L, P.frontiers[b + (int)(idx >> 2)], (int)(idx & 3));
}

}  // namespace

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
  const int64_t lds_words = 4LL * lib->num_luts + 12LL * lib->num_timings + lib->num_lut_vals;
  // a level's pins are not known on the host: a fixed grid strides over them
  if (lds_words <= kLdsWords) {
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
