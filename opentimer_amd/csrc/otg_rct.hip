// P1: RC-tree Elmore analysis on gfx950.
//
// Two kernels:
//  * rct_plain_kernel -- one thread per (net, corner) walking the net five
//    times in BFS order: the arithmetic of compute_net_timing
//    (reference ot/cuda/rct.cu:11-59), kept as the comparison baseline.
//  * rct_batched_kernel -- persistent, one wavefront per work item pulled
//    from a largest-first queue.  A work item is a packed group of whole
//    small nets or one large net.  Positions are visited in windows of 64
//    (one lane per node): one bottom-up sweep (windows last to first) and
//    one top-down sweep (first to last).  Dependencies inside a window are
//    resolved in rounds (a node goes when its in-window children / parent
//    are done); dependencies across windows go through an on-chip ring of
//    per-node accumulators indexed by position mod W.
//
// Moment formulation used by the batched kernel.  With R_i the resistance
// into node i, load_i the subtree capacitance and P(i) the driver->i path
// (driver excluded):
//    load_i   = C_i + sum_children load_c                      (bottom-up)
//    L_i      = sum_children (L_c + R_c load_c^2)              (bottom-up)
//    delay_i  = sum_{a in P(i)} R_a load_a                     (top-down)
//    impulse_i= sum_{a in P(i)} (R_a load_a)^2 + 2 R_a L_a     (top-down)
//    ldelay_i = delay_i load_i + L_i,  beta_i = (delay_i^2 + impulse_i)/2
// This equals the reference's four sweeps (load, delay, ldelay, beta) and
// impulse = 2 beta - delay^2 (ot/timer/net.cpp:127-191): L_i is the sum of
// C_d (delay_d - delay_i) over the subtree, so ldelay_i = delay_i load_i +
// L_i; and sum_a R_a ldelay_a = sum_a u_a delay_a + sum_a R_a L_a with
// u_a = R_a load_a and sum_{a<=b on the path} u_a u_b = (delay^2 + sum u^2)/2.
// Every term is non-negative, so no subtraction (and no cancellation) is
// left; the sweeps drop from four to two.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

#include "otg_internal.hpp"

namespace {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// FP64 reference values (precision reports): the reference's sweeps
// (ot/timer/net.cpp:101-193; impulse = 2 beta - delay^2 as :191) and the
// output slew (net.cpp:30-32) in double from the FP32 inputs, one thread per
// (net, corner), each net walked in BFS order as the plain kernel walks it.
// A measuring instrument for the FP32 methods, not a solve method.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void rct_f64_kernel(
    int32_t num_nets, const int32_t* __restrict__ nodes_start, const int32_t* __restrict__ pid,
    const float* __restrict__ pres, const float* __restrict__ cap,
    const float* __restrict__ slew_in, double* load, double* delay, double* ldelay, double* beta,
    double* impulse, double* slew) {
  const int net = blockIdx.x * 64 + threadIdx.x;
  const int c = threadIdx.y;
  if (net >= num_nets) return;
  const int64_t st = nodes_start[net], ed = nodes_start[net + 1];
  if (ed <= st) return;
  const int64_t st4 = st * 4 + c, ed4 = ed * 4 + c;
  for (int64_t i = st4; i < ed4; i += 4) load[i] = (double)cap[i];
  for (int64_t i = ed4 - 4, j = ed - 1; i > st4; i -= 4, --j) load[i - (int64_t)pid[j] * 4] += load[i];
  delay[st4] = 0.0;
  for (int64_t i = st4 + 4, j = st + 1; i < ed4; i += 4, ++j)
    delay[i] = delay[i - (int64_t)pid[j] * 4] + (double)pres[j] * load[i];
  for (int64_t i = st4; i < ed4; i += 4) ldelay[i] = (double)cap[i] * delay[i];
  for (int64_t i = ed4 - 4, j = ed - 1; i > st4; i -= 4, --j)
    ldelay[i - (int64_t)pid[j] * 4] += ldelay[i];
  beta[st4] = 0.0;
  for (int64_t i = st4 + 4, j = st + 1; i < ed4; i += 4, ++j)
    beta[i] = beta[i - (int64_t)pid[j] * 4] + (double)pres[j] * ldelay[i];
  const double si = slew_in ? (double)slew_in[(int64_t)net * 4 + c] : 0.0;
  for (int64_t i = st4; i < ed4; i += 4) {
    const double t = delay[i];
    const double imp = 2.0 * beta[i] - t * t;
    impulse[i] = imp;
    if (slew) {
      const double v = sqrt(si * si + imp);
      slew[i] = si < 0.0 ? -v : v;
    }
  }
}

// ---------------------------------------------------------------------------
// Plain kernel: reference arithmetic, FP32, no contraction.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void rct_plain_kernel(
    int32_t num_nets, const int32_t* __restrict__ nodes_start, const int32_t* __restrict__ pid,
    const float* __restrict__ pres, const float* __restrict__ cap,
    const float* __restrict__ slew_in, float* load, float* delay, float* ldelay, float* beta,
    float* impulse, float* slew) {
#pragma clang fp contract(off)
  const int net = blockIdx.x * 64 + threadIdx.x;
  const int c = threadIdx.y;
  if (net >= num_nets) return;
  const int64_t st = nodes_start[net], ed = nodes_start[net + 1];
  if (ed <= st) return;
  const int64_t st4 = st * 4 + c, ed4 = ed * 4 + c;
  for (int64_t i = st4; i < ed4; i += 4) {
    load[i] = cap[i];
    delay[i] = 0.0f;
    ldelay[i] = 0.0f;
    beta[i] = 0.0f;
  }
  for (int64_t i = ed4 - 4, j = ed - 1; i > st4; i -= 4, --j) load[i - (int64_t)pid[j] * 4] += load[i];
  for (int64_t i = st4 + 4, j = st + 1; i < ed4; i += 4, ++j)
    delay[i] += delay[i - (int64_t)pid[j] * 4] + load[i] * pres[j];
  for (int64_t i = ed4 - 4, j = ed - 1; i > st4; i -= 4, --j) {
    ldelay[i] += cap[i] * delay[i];
    ldelay[i - (int64_t)pid[j] * 4] += ldelay[i];
  }
  ldelay[st4] += cap[st4] * delay[st4];
  for (int64_t i = st4 + 4, j = st + 1; i < ed4; i += 4, ++j)
    beta[i] += beta[i - (int64_t)pid[j] * 4] + ldelay[i] * pres[j];
  const float si = slew_in ? slew_in[(int64_t)net * 4 + c] : 0.0f;
  for (int64_t i = st4; i < ed4; i += 4) {
    const float t = delay[i];
    const float imp = 2.0f * beta[i] - t * t;
    if (impulse) impulse[i] = imp;
    if (slew) {
      const float v = sqrtf(si * si + imp);
      slew[i] = si < 0.0f ? -v : v;
    }
  }
}

// ---------------------------------------------------------------------------
// Compensated mode: every sum (subtree sums of load and L, path sums of delay
// and impulse) is kept in FP64; the inputs and the spilled per-node load and
// L stay FP32.  gfx950 executes v_add_f64 / v_fma_f64 at the rate of
// unpacked FP32 operations, so a compensated addition is one instruction
// (a double-float two-sum took six to ten).  All terms are non-negative, so
// the sums carry no cancellation, and FP64 keeps their relative error near
// 1e-16 whatever their length or order.
//
// A double travels through the rings and the (hi, lo) float pairs of the
// window code as its two 32-bit halves (hi: the upper word, lo: the lower),
// so the FP32 method's one float per value and the compensated method's
// pair share the code paths; converting between the halves and the double
// is a register reinterpretation.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double as_f64(float hi, float lo) {
  return __hiloint2double(__float_as_int(hi), __float_as_int(lo));
}
__device__ __forceinline__ void split_f64(double v, float& hi, float& lo) {
  hi = __int_as_float(__double2hiint(v));
  lo = __int_as_float(__double2loint(v));
}

__device__ __forceinline__ void unpack(float4 v, float (&a)[4]) {
  a[0] = v.x;
  a[1] = v.y;
  a[2] = v.z;
  a[3] = v.w;
}
__device__ __forceinline__ float4 pack(const float (&a)[4]) { return make_float4(a[0], a[1], a[2], a[3]); }

// Compensated mode in LDS: four FP64 values (one per corner) in the entries
// k of two float4 arrays, corners 0-1 in A and 2-3 in B, so that a 16-byte
// LDS access moves two doubles to or from the register pairs the FP64
// instructions use (a value's words split over two arrays cost a register
// move per word).  The _hl forms take and give the (hi, lo) words.
__device__ __forceinline__ void lds_get_d4(const float4* A, const float4* B, int k,
                                           double (&v)[4]) {
  const double2 p = reinterpret_cast<const double2*>(A)[k];
  const double2 q = reinterpret_cast<const double2*>(B)[k];
  v[0] = p.x;
  v[1] = p.y;
  v[2] = q.x;
  v[3] = q.y;
}
__device__ __forceinline__ void lds_put_d4(float4* A, float4* B, int k, const double (&v)[4]) {
  reinterpret_cast<double2*>(A)[k] = make_double2(v[0], v[1]);
  reinterpret_cast<double2*>(B)[k] = make_double2(v[2], v[3]);
}
__device__ __forceinline__ void lds_get_hl(const float4* A, const float4* B, int k, float (&h)[4],
                                           float (&l)[4]) {
  double v[4];
  lds_get_d4(A, B, k, v);
#pragma unroll
  for (int c = 0; c < 4; ++c) split_f64(v[c], h[c], l[c]);
}
__device__ __forceinline__ void lds_put_hl(float4* A, float4* B, int k, const float (&h)[4],
                                           const float (&l)[4]) {
  double v[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) v[c] = as_f64(h[c], l[c]);
  lds_put_d4(A, B, k, v);
}

// Global-memory accesses through address-space-1 pointers: the compiler then
// always emits global_load/global_store.  (A generic pointer it cannot prove
// global becomes a FLAT access, which also counts against the LDS counter,
// so every LDS wait would wait for HBM latency too.)
#define OTG_GLOBAL __attribute__((address_space(1)))
typedef float otg_f4v __attribute__((ext_vector_type(4)));
typedef int otg_i4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ int gld(const int32_t* p) { return *(const OTG_GLOBAL int32_t*)p; }
__device__ __forceinline__ int64_t gld(const int64_t* p) { return *(const OTG_GLOBAL int64_t*)p; }
__device__ __forceinline__ float gld(const float* p) { return *(const OTG_GLOBAL float*)p; }
__device__ __forceinline__ float4 gld(const float4* p) {
  const otg_f4v v = *(const OTG_GLOBAL otg_f4v*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gst(float4* p, const float4& v) {
  otg_f4v t = {v.x, v.y, v.z, v.w};
  *(OTG_GLOBAL otg_f4v*)p = t;
}
__device__ __forceinline__ RctItem gld(const RctItem* p) {
  static_assert(sizeof(RctItem) == 32, "RctItem layout");
  const otg_i4v a = *(const OTG_GLOBAL otg_i4v*)p;
  const otg_i4v b = *((const OTG_GLOBAL otg_i4v*)p + 1);
  RctItem it;
  it.pos_begin = a.x;
  it.pos_end = a.y;
  it.net_begin = a.z;
  it.net_end = a.w;
  it.kind = b.x;
  it.k = b.y;
  it.scratch = (int64_t)(((uint64_t)(uint32_t)b.w << 32) | (uint32_t)b.z);
  return it;
}

// ---------------------------------------------------------------------------
// Batched kernel.
// ---------------------------------------------------------------------------
struct RctParams {
  const int32_t* nodes_start;
  const int32_t* pid;
  const float* pres;
  const float4* cap;
  const float4* slew_in;
  float4* load;
  float4* delay;
  float4* ldelay;
  float4* beta;
  float4* impulse;
  float4* slew;
  const RctItem* items;
  int32_t num_items;
  int32_t num_nets_total;
  int32_t subset_count;
  const int32_t* subset;
  const int32_t* net_maxpid;
  const int64_t* net_scratch;
  float* scratch;
  int32_t* queue;
  int32_t long_maxpid;             // largest parent offset the streaming kernel's ring takes
  const uint32_t* sched;           // packed groups: schedule words
  const int32_t* num_items_dev;    // item count on the device (subset solves), or null
  int32_t* group_flag;             // subset solves: groups listed in this call
  int32_t group_max;               // largest net of a packed group (G)
};


template <int W, bool COMP>
struct RctLds {
  static constexpr int KP = (W < 256 ? W : 256) / 64;  // positions per lane, packed items
  int32_t spid[W];             // parent offsets of a packed item, kept for the top-down sweep
  float sres[W];               // resistances of a packed item
  uint32_t rootbits[W / 32];   // driver positions of a packed group
  int32_t cnt[kWave * KP];     // per-position child counters / done flags of a window
};

// Window storage of the global-window path: a per-node region in global
// memory (no wrap).  FP32 method: a and b hold one float4 per node.
// Compensated method: a and b hold four doubles per node (each twice the
// size), handed to the window code as (hi, lo) halves; children add into
// their parent with FP64 atomics.
template <bool GLOBAL, bool COMP>
struct Ring {
  float4* a;
  float4* b;
  int32_t base;
  int32_t mask;  // -1 (no wrap) for global memory
  __device__ __forceinline__ int slot(int i) const { return (i - base) & mask; }

  __device__ __forceinline__ static float ld32(const float* q) {
    if constexpr (GLOBAL) return __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else return *q;
  }
  __device__ __forceinline__ static double ld64(const double* q) {
    if constexpr (GLOBAL) {
      const uint64_t v = __hip_atomic_load(reinterpret_cast<const uint64_t*>(q), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
      return __builtin_bit_cast(double, v);
    } else {
      return *q;
    }
  }
  __device__ __forceinline__ static void get1(const float4* p, int k, float (&x)[4], float (&xl)[4]) {
    if constexpr (COMP) {
      const double* q = reinterpret_cast<const double*>(p) + 4LL * k;
#pragma unroll
      for (int c = 0; c < 4; ++c) split_f64(ld64(q + c), x[c], xl[c]);
    } else {
      const float* q = reinterpret_cast<const float*>(p + k);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        x[c] = ld32(q + c);
        xl[c] = 0.f;
      }
    }
  }
  __device__ __forceinline__ static void put1(float4* p, int k, const float (&x)[4],
                                              const float (&xl)[4]) {
    if constexpr (COMP) {
      double* q = reinterpret_cast<double*>(p) + 4LL * k;
#pragma unroll
      for (int c = 0; c < 4; ++c) q[c] = as_f64(x[c], xl[c]);
    } else {
      p[k] = pack(x);
    }
  }
  __device__ __forceinline__ static void add1(float4* p, int k, const float (&x)[4],
                                              const float (&xl)[4]) {
    if constexpr (COMP) {
      double* q = reinterpret_cast<double*>(p) + 4LL * k;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if constexpr (GLOBAL) unsafeAtomicAdd(q + c, as_f64(x[c], xl[c]));
        else atomicAdd(q + c, as_f64(x[c], xl[c]));
      }
    } else {
      float* q = reinterpret_cast<float*>(p + k);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if constexpr (GLOBAL) unsafeAtomicAdd(q + c, x[c]);
        else atomicAdd(q + c, x[c]);
      }
    }
  }
  __device__ __forceinline__ void get(int k, float (&x)[4], float (&y)[4], float (&xl)[4],
                                      float (&yl)[4]) const {
    get1(a, k, x, xl);
    get1(b, k, y, yl);
  }
  __device__ __forceinline__ void put(int k, const float (&x)[4], const float (&y)[4],
                                      const float (&xl)[4], const float (&yl)[4]) const {
    put1(a, k, x, xl);
    put1(b, k, y, yl);
  }
  __device__ __forceinline__ void zero(int k) const {
    const float z[4] = {0.f, 0.f, 0.f, 0.f};
    put(k, z, z, z, z);
  }
  // acc[k] += (x, xl), acc'[k] += (y, yl)
  __device__ __forceinline__ void add(int k, const float (&x)[4], const float (&y)[4],
                                      const float (&xl)[4], const float (&yl)[4]) const {
    add1(a, k, x, xl);
    add1(b, k, y, yl);
  }
  __device__ __forceinline__ void fence() const {
    if constexpr (GLOBAL) __threadfence();
  }
};

// The workgroup is one wavefront, and LDS operations of a wavefront execute
// in issue order, so ordering LDS traffic between rounds only needs the
// compiler to keep program order.  (__syncthreads() would also wait for every
// outstanding global load and store, serialising memory latency into each
// round.)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_wave_barrier();
  __asm__ __volatile__("" ::: "memory");
}

// Inputs of one window of 64*K positions; lane l holds positions
// wk + k*64 + l, k = 0..K-1 (each k is a coalesced 64-position block).
template <int K>
struct BuIn {
  int32_t pid[K];
  float r[K];
  float4 c[K];
};
template <int K, bool KEEP>
struct TdIn {
  int32_t pid[K];
  float r[K];
  float4 ld[KEEP ? 1 : K];   // streamed items: load and L read back
  float4 L[KEEP ? 1 : K];
  float4 si[KEEP ? K : 1];   // driver slew: per position for a packed group
};

// Counter slot of window-relative position q: each lane's K counters are
// contiguous, so a lane reads them with wide LDS loads.
template <int K>
__device__ __forceinline__ int cidx(int q) {
  return (q & (kWave - 1)) * K + (q >> 6);
}

// Buffer resources (V#) over the item's slice of each array.  Loads past the
// slice read 0 and stores past it are dropped by the hardware range check,
// so window code needs no per-lane branch around its memory operations and
// every window issues a fixed number of them.  A null array gets a
// zero-size slice (all its stores dropped).
typedef __amdgpu_buffer_rsrc_t Rsrc;
typedef unsigned int otg_u4v __attribute__((ext_vector_type(4)));
constexpr int kRsrcWord3 = 0x00020000;  // raw, untyped buffer (gfx9 family)

__device__ __forceinline__ Rsrc make_rsrc(const void* base, int64_t elem0, int elem_bytes,
                                          int count) {
  if (!base) return __builtin_amdgcn_make_buffer_rsrc(nullptr, 0, 0, kRsrcWord3);
  return __builtin_amdgcn_make_buffer_rsrc(
      const_cast<char*>(static_cast<const char*>(base) + elem0 * elem_bytes), 0,
      count * elem_bytes, kRsrcWord3);
}
__device__ __forceinline__ float4 bload4(Rsrc r, int byte_off) {
  const otg_u4v v = __builtin_amdgcn_raw_buffer_load_b128(r, byte_off, 0, 0);
  const otg_f4v f = __builtin_bit_cast(otg_f4v, v);
  return make_float4(f.x, f.y, f.z, f.w);
}
__device__ __forceinline__ int bload_i(Rsrc r, int byte_off) {
  return (int)__builtin_amdgcn_raw_buffer_load_b32(r, byte_off, 0, 0);
}
__device__ __forceinline__ float bload_f(Rsrc r, int byte_off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, byte_off, 0, 0));
}
__device__ __forceinline__ void bstore4(Rsrc r, int byte_off, const float (&a)[4]) {
  const otg_f4v f = {a[0], a[1], a[2], a[3]};
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(otg_u4v, f), r, byte_off, 0, 0);
}

struct ItemBufs {
  Rsrc pid, res, cap, load, gl, delay, ldelay, beta, imp, slew;
};

template <int W, bool COMP>
struct ItemCtx {
  RctLds<W, COMP>& S;
  int pb, pe;
  bool multi;
  int lane;
  ItemBufs B;
  __device__ __forceinline__ bool is_root(int i) const {
    if (multi) return (S.rootbits[(i - pb) >> 5] >> ((i - pb) & 31)) & 1u;
    return i == pb;
  }
};

// ---- bottom-up window: load_i = C_i + sum load_c, L_i = sum (L_c + R_c load_c^2)
// Each round reads every position's counter and accumulator slot
// unconditionally (one LDS round trip for the whole window), computes every
// position, and writes only the positions that are ready (their in-window
// children are all done).  Reads of a not-yet-ready slot are discarded.
template <int K, bool KEEP, int W, bool COMP, bool GLOBAL>
__device__ __forceinline__ void bu_window(const ItemCtx<W, COMP>& X, const Ring<GLOBAL, COMP>& R,
                                          const int wk, const BuIn<K>& in) {
  constexpr bool STASH = KEEP && !GLOBAL;  // a packed group keeps pid/R in LDS
  const int lane = X.lane;
  int par[K], slot[K];
  uint32_t pending = 0, inwin = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = wk + k * kWave + lane;
    const bool valid = i < X.pe;
    const bool root = valid && X.is_root(i);
    par[k] = (valid && !root) ? i - in.pid[k] : -1;
    slot[k] = R.slot(i);
    if (valid) pending |= 1u << k;
    if (par[k] >= wk) inwin |= 1u << k;
    X.S.cnt[lane * K + k] = 0;
    if (STASH && valid) {
      X.S.spid[i - X.pb] = in.pid[k];
      X.S.sres[i - X.pb] = in.r[k];
    }
  }
  wave_sync();
#pragma unroll
  for (int k = 0; k < K; ++k)
    if ((inwin >> k) & 1u) atomicAdd(&X.S.cnt[cidx<K>(par[k] - wk)], 1);
  wave_sync();
  for (int round = 0; round <= kWave * K; ++round) {
    int cn[K];
    float xa[K][4], xb[K][4], xal[K][4], xbl[K][4];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      cn[k] = X.S.cnt[lane * K + k];
      R.get(slot[k], xa[k], xb[k], xal[k], xbl[k]);
    }
    uint32_t ready = 0;
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (((pending >> k) & 1u) && cn[k] == 0) ready |= 1u << k;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      float cap4[4];
      unpack(in.c[k], cap4);
      float ld[4], ldl[4], L[4], Ll[4], w[4], wl[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if constexpr (COMP) {
          // FP64 sums and products
          const double t = as_f64(xa[k][c], xal[k][c]) + (double)cap4[c];
          split_f64(t, ld[c], ldl[c]);
          L[c] = xb[k][c];
          Ll[c] = xbl[k][c];
          split_f64(as_f64(L[c], Ll[c]) + (double)in.r[k] * t * t, w[c], wl[c]);
        } else {
          ld[c] = cap4[c] + xa[k][c];
          ldl[c] = 0.f;
          L[c] = xb[k][c];
          Ll[c] = 0.f;
          w[c] = L[c] + in.r[k] * ld[c] * ld[c];
          wl[c] = 0.f;
        }
      }
      if (!((ready >> k) & 1u)) continue;
      R.put(slot[k], ld, L, ldl, Ll);  // final: no child adds to it any more
      if (par[k] >= 0) {
        R.add(R.slot(par[k]), ld, w, ldl, wl);
        if ((inwin >> k) & 1u) atomicSub(&X.S.cnt[cidx<K>(par[k] - wk)], 1);
      }
    }
    pending &= ~ready;
    R.fence();
    wave_sync();
    if (__all(pending == 0)) break;
  }
  if constexpr (!KEEP) {
    // streamed item: flush the window's load and L to global memory (a fixed
    // number of stores per window, issued after the rounds, so the next
    // window's prefetched loads can be waited for without waiting for these
    // stores) and free the slots for positions W further on
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = wk + k * kWave + lane;
      float ld[4], L[4], ldl[4], Ll[4];
      R.get(slot[k], ld, L, ldl, Ll);
      float o1[4], o2[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        o1[c] = COMP ? (float)as_f64(ld[c], ldl[c]) : ld[c];
        o2[c] = COMP ? (float)as_f64(L[c], Ll[c]) : L[c];
      }
      bstore4(X.B.load, (i - X.pb) * 16, o1);
      bstore4(X.B.gl, (i - X.pb) * 16, o2);
    }
    wave_sync();
#pragma unroll
    for (int k = 0; k < K; ++k) R.zero(slot[k]);
  }
  wave_sync();
}

// ---- top-down window: delay_i = delay_p + R_i load_i,
//      impulse_i = impulse_p + (R_i load_i)^2 + 2 R_i L_i.  Each round reads
//      every position's parent flag and parent slot together and lets the
//      ready positions (parent done) write their own slot; the outputs of the
//      whole window are computed and stored after the rounds, a fixed number
//      of stores per window.
template <int K, bool KEEP, int W, bool COMP, bool GLOBAL>
__device__ __forceinline__ void td_window(const ItemCtx<W, COMP>& X, const Ring<GLOBAL, COMP>& R,
                                          const int wk, const TdIn<K, KEEP>& in) {
  const int lane = X.lane;
  int par[K], slot[K], pslot[K], flag_idx[K];
  float rk[K];
  float ld[K][4], ldl[K][4], L[K][4], Ll[K][4];
  uint32_t pending = 0, inwin = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = wk + k * kWave + lane;
    const bool valid = i < X.pe;
    const bool root = valid && X.is_root(i);
    par[k] = (valid && !root) ? i - in.pid[k] : -1;
    slot[k] = R.slot(i);
    pslot[k] = par[k] >= 0 ? R.slot(par[k]) : slot[k];
    rk[k] = root ? 0.f : in.r[k];
    if (valid) pending |= 1u << k;
    if (par[k] >= wk) inwin |= 1u << k;
    flag_idx[k] = par[k] >= wk ? cidx<K>(par[k] - wk) : lane * K + k;
    X.S.cnt[lane * K + k] = 0;
    // own load and L: from the slot (whole item on chip) or the prefetch
    if constexpr (KEEP) {
      R.get(slot[k], ld[k], L[k], ldl[k], Ll[k]);
    } else {
      unpack(in.ld[KEEP ? 0 : k], ld[k]);
      unpack(in.L[KEEP ? 0 : k], L[k]);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        ldl[k][c] = Ll[k][c] = 0.f;
        if constexpr (COMP) {  // FP32 values read back, as (hi, lo) halves of doubles
          split_f64((double)ld[k][c], ld[k][c], ldl[k][c]);
          split_f64((double)L[k][c], L[k][c], Ll[k][c]);
        }
      }
    }
  }
  wave_sync();
  for (int round = 0; round <= kWave * K; ++round) {
    int fl[K];
    float pd[K][4], pdl[K][4], pm[K][4], pml[K][4];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      fl[k] = X.S.cnt[flag_idx[k]];
      R.get(pslot[k], pd[k], pm[k], pdl[k], pml[k]);
    }
    uint32_t ready = 0;
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (((pending >> k) & 1u) && (!((inwin >> k) & 1u) || fl[k] != 0)) ready |= 1u << k;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (!((ready >> k) & 1u)) continue;
      const float r = rk[k];
      const bool has_par = par[k] >= 0;
      float d[4], dl[4], m[4], ml[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const float qd = has_par ? pd[k][c] : 0.f, qdl = has_par ? pdl[k][c] : 0.f;
        const float qm = has_par ? pm[k][c] : 0.f, qml = has_par ? pml[k][c] : 0.f;
        if constexpr (COMP) {
          // FP64 products and path sums
          const double u = (double)r * as_f64(ld[k][c], ldl[k][c]);
          const double z = u * u + 2.0 * (double)r * as_f64(L[k][c], Ll[k][c]);
          split_f64(as_f64(qd, qdl) + u, d[c], dl[c]);
          split_f64(as_f64(qm, qml) + z, m[c], ml[c]);
        } else {
          const float u = r * ld[k][c];
          d[c] = qd + u;
          m[c] = qm + (u * u + 2.0f * r * L[k][c]);
          dl[c] = ml[c] = 0.f;
        }
      }
      R.put(slot[k], d, m, dl, ml);
      X.S.cnt[lane * K + k] = 1;
    }
    pending &= ~ready;
    R.fence();
    wave_sync();
    if (__all(pending == 0)) break;
  }
  // outputs of the window
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = wk + k * kWave + lane;
    float d[4], dl[4], m[4], ml[4];
    R.get(slot[k], d, m, dl, ml);
    float sv[4];
    unpack(in.si[KEEP ? k : 0], sv);
    float o_ld[4], o_d[4], o_ldel[4], o_b[4], o_m[4], o_s[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if constexpr (COMP) {
        // final per-node combination in FP64; the slew's radicand is formed
        // in FP64 and its square root taken in FP32, as in node_outputs
        const double dd = as_f64(d[c], dl[c]);
        const double md = as_f64(m[c], ml[c]);
        const double ldd = as_f64(ld[k][c], ldl[k][c]);
        const double Ld = as_f64(L[k][c], Ll[k][c]);
        o_ld[c] = (float)ldd;
        o_d[c] = (float)dd;
        o_m[c] = (float)md;
        o_ldel[c] = (float)(dd * ldd + Ld);
        o_b[c] = (float)(0.5 * (dd * dd + md));
        const float q = sqrtf((float)((double)sv[c] * (double)sv[c] + md));
        o_s[c] = sv[c] < 0.f ? -q : q;
      } else {
        o_ld[c] = ld[k][c];
        o_d[c] = d[c];
        o_m[c] = m[c];
        o_ldel[c] = d[c] * ld[k][c] + L[k][c];
        o_b[c] = 0.5f * (d[c] * d[c] + m[c]);
        const float q = sqrtf(sv[c] * sv[c] + m[c]);
        o_s[c] = sv[c] < 0.f ? -q : q;
      }
    }
    const int off = (i - X.pb) * 16;
    if constexpr (KEEP) bstore4(X.B.load, off, o_ld);
    bstore4(X.B.delay, off, o_d);
    bstore4(X.B.ldelay, off, o_ldel);
    bstore4(X.B.beta, off, o_b);
    bstore4(X.B.imp, off, o_m);
    bstore4(X.B.slew, off, o_s);
  }
  wave_sync();
}

template <int K, int W, bool COMP>
__device__ __forceinline__ void load_bu(const ItemCtx<W, COMP>& X, int wk, BuIn<K>& in) {
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int q = wk + k * kWave + X.lane - X.pb;  // item-relative position
    in.pid[k] = bload_i(X.B.pid, q * 4);
    in.r[k] = bload_f(X.B.res, q * 4);
    in.c[k] = bload4(X.B.cap, q * 16);
  }
}

// One work item: windows of 64*K positions, bottom-up over the windows from
// last to first, then top-down from first to last.  The next window's
// global inputs are requested before the current window is processed.
// KEEP: the whole item fits the on-chip window (packed groups, global-window
// nets), so load and L stay on chip between the sweeps; otherwise (streamed
// nets) they go to global memory and are read back.
template <int K, bool KEEP, int W, bool COMP, bool GLOBAL>
__device__ __forceinline__ void process_item(const RctParams& P, RctLds<W, COMP>& S,
                                             const Ring<GLOBAL, COMP>& R, const RctItem& it,
                                             const int lane) {
  constexpr int WIN = kWave * K;
  constexpr bool STASH = KEEP && !GLOBAL;
  const int pb = it.pos_begin, pe = it.pos_end, n = pe - pb;
  const bool multi = it.net_end - it.net_begin > 1;
  const float* gl = P.scratch + (KEEP ? 0 : it.scratch);
  ItemCtx<W, COMP> X{S, pb, pe, multi, lane,
                     ItemBufs{make_rsrc(P.pid, pb, 4, n), make_rsrc(P.pres, pb, 4, n),
                              make_rsrc(P.cap, pb, 16, n), make_rsrc(P.load, pb, 16, n),
                              make_rsrc(KEEP ? nullptr : gl, 0, 16, n),
                              make_rsrc(P.delay, pb, 16, n), make_rsrc(P.ldelay, pb, 16, n),
                              make_rsrc(P.beta, pb, 16, n), make_rsrc(P.impulse, pb, 16, n),
                              make_rsrc(P.slew, pb, 16, n)}};
  const Rsrc rsi = make_rsrc(P.slew_in, 0, 16, P.num_nets_total);
  const int nwin = (n + WIN - 1) / WIN;
  // first window's inputs go out before the set-up work
  BuIn<K> cur, nxt;
  load_bu<K>(X, pb + (nwin - 1) * WIN, cur);
  // ---- set-up: clear the window, mark drivers ------------------------------
  {
    const int nz = KEEP ? n : W;
    for (int k = lane; k < nz; k += kWave) R.zero(k);
    if (multi) {
      for (int k = lane; k < W / 32; k += kWave) S.rootbits[k] = 0u;
      wave_sync();
      for (int j = it.net_begin + lane; j < it.net_end; j += kWave) {
        const int r = gld(P.nodes_start + j) - pb;
        atomicOr(&S.rootbits[r >> 5], 1u << (r & 31));
      }
    }
    R.fence();
    wave_sync();
  }
  // ---- bottom-up sweep ----------------------------------------------------
  for (int w = nwin - 1; w >= 0; --w) {
    load_bu<K>(X, pb + (w > 0 ? w - 1 : 0) * WIN, nxt);
    bu_window<K, KEEP, W, COMP, GLOBAL>(X, R, pb + w * WIN, cur);
    cur = nxt;
  }
  R.fence();
  wave_sync();
  // ---- top-down sweep -----------------------------------------------------
  // driver slew of each position's net: counted from the driver bitmap in
  // window order for a packed group; the single net otherwise
  int net_carry = 0;
  const float4 si0 = bload4(rsi, it.net_begin * 16);
  auto load_td = [&](int w, TdIn<K, KEEP>& in) {
    if constexpr (!KEEP) in.si[0] = si0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = pb + w * WIN + k * kWave + lane;
      const int q = i - pb;
      if constexpr (KEEP) {
        const bool root = i < pe && X.is_root(i);
        const uint64_t rm = __ballot(root);
        const int net = multi ? it.net_begin + net_carry + __popcll(rm & ((2ull << lane) - 1ull)) - 1
                              : it.net_begin;
        net_carry += __popcll(rm);
        in.si[k] = bload4(rsi, (i < pe ? net : P.num_nets_total) * 16);
      }
      if constexpr (STASH) {
        in.pid[k] = i < pe ? S.spid[q] : 0;
        in.r[k] = i < pe ? S.sres[q] : 0.f;
      } else {
        in.pid[k] = bload_i(X.B.pid, q * 4);
        in.r[k] = bload_f(X.B.res, q * 4);
      }
      if constexpr (!KEEP) {
        in.ld[KEEP ? 0 : k] = bload4(X.B.load, q * 16);
        in.L[KEEP ? 0 : k] = bload4(X.B.gl, q * 16);
      }
    }
  };
  TdIn<K, KEEP> tc, tn;
  load_td(0, tc);
  for (int w = 0; w < nwin; ++w) {
    load_td(w + 1 < nwin ? w + 1 : w, tn);
    td_window<K, KEEP, W, COMP, GLOBAL>(X, R, pb + w * WIN, tc);
    tc = tn;
  }
  R.fence();
  wave_sync();
}

// The last workgroup out resets its queue for the next launch.
__device__ __forceinline__ void queue_release(int32_t* queue, int lane) {
  __syncthreads();
  if (lane == 0) {
    __threadfence();
    const int old = atomicAdd(&queue[1], 1);
    if (old == (int)gridDim.x - 1) {
      atomicExch(&queue[0], 0);
      atomicExch(&queue[1], 0);
    }
  }
}

// Nets whose parent offsets exceed what the streaming kernel's ring can hold:
// windows of 64 with the per-node window storage in global memory.  In subset
// mode every listed net is classified and only those nets are taken.
template <bool COMP>
__global__ __launch_bounds__(64) void rct_global_kernel(RctParams P) {
  constexpr int W = 256;  // LDS of this path: counters only, the ring is in global memory
  __shared__ RctLds<W, COMP> S;
  const int lane = threadIdx.x;
  const int total = P.subset ? P.subset_count : P.num_items;
  for (;;) {
    int idx = 0;
    if (lane == 0) idx = atomicAdd(&P.queue[0], 1);
    idx = __shfl(idx, 0);
    if (idx >= total) break;
    RctItem it;
    if (P.subset) {
      const int net = gld(P.subset + idx);
      it.pos_begin = gld(P.nodes_start + net);
      it.pos_end = gld(P.nodes_start + net + 1);
      it.net_begin = net;
      it.net_end = net + 1;
      const int nn = it.pos_end - it.pos_begin;
      const int mp = gld(P.net_maxpid + net);
      it.kind = (nn <= P.group_max || mp <= P.long_maxpid) ? ITEM_PACKED : ITEM_GLOBAL;
      it.k = 1;
      it.scratch = gld(P.net_scratch + net);
    } else {
      it = gld(P.items + idx);
    }
    if (it.pos_end <= it.pos_begin || it.kind != ITEM_GLOBAL) continue;
    const int64_t nn = it.pos_end - it.pos_begin;
    float4* base = reinterpret_cast<float4*>(P.scratch + it.scratch);
    Ring<true, COMP> R{base, base + (COMP ? 2 : 1) * nn, it.pos_begin, -1};
    process_item<1, true, W, COMP, true>(P, S, R, it, lane);
  }
  queue_release(P.queue, lane);
}

// ===========================================================================
// Streaming kernel: nets larger than a packed group.
//
// Persistent, one wavefront per workgroup, items from a largest-first queue.
// Positions are visited in windows of 64, one lane per position.  Inside a
// window the dependency order is resolved by in-window depth: a lane whose
// parent lies before the window has depth 0, otherwise its parent's depth
// plus one (computed with one ballot per level, no memory traffic).  The
// bottom-up sweep visits depths from the deepest up, the top-down sweep from
// 0 down; every step touches only the lanes of that depth.  Parents in
// earlier windows are reached through a ring of W per-position slots in LDS
// (slot = position mod W), which holds every parent as long as parent
// offsets stay below W - 64.
//
// Global inputs arrive in chunks of four windows (256 positions) that are
// requested one chunk ahead.  A wavefront that both loads and stores can
// only wait for all of its memory operations at once, so the prefetch
// distance is one chunk of work rather than one window.  The bottom-up sweep
// runs over chunks last to first, spilling the (rounded) load into the load
// output and L into scratch; the top-down sweep runs first to last and reads
// them back.
// ===========================================================================
constexpr int kChunkWin = 4;
constexpr int kChunkPos = kChunkWin * kWave;  // 256
constexpr int kClaim = 256;                   // claim table entries (parent slot mod 256)

// Slot W of each ring array is a zero slot, never written after the ring is
// cleared: the driver's top-down parent.
template <int W, bool COMP>
struct StreamLds {
  // bottom-up: load and L accumulators | top-down: delay and impulse; FP32,
  // or in the compensated mode FP64 with corners 0-1 in a/b, 2-3 in al/bl
  float4 a[W + 1];
  float4 b[W + 1];
  float4 al[COMP ? W + 1 : 1];
  float4 bl[COMP ? W + 1 : 1];
  int32_t claim[kClaim];  // one writer per parent slot per round
};

// Inputs of one chunk: lane l holds positions q0 + 64 s + l, s = 0..3.
struct Chunk {
  int32_t pid[kChunkWin];
  float r[kChunkWin];
  float4 x[kChunkWin];  // cap (bottom-up) | load read back (top-down)
  float4 y[kChunkWin];  // L read back (top-down)
  float4 si;            // the net's driver slew (top-down)
};

struct Item {
  int pb, pe, nb, ne, kind;
  int k;            // packed group: its index in the plan's group list
  int64_t scratch;  // streamed net: float offset of its L scratch;
                    // packed group: offset of its schedule words
};

enum : int { PH_BU = 1, PH_TD = 2 };

__device__ __forceinline__ int num_chunks(const Item& it) {
  return (it.pe - it.pb + kChunkPos - 1) / kChunkPos;
}

// Loads that must not hit a stale L1 line: the top-down sweep reads back
// what the same lanes stored during the bottom-up sweep (glc: read from L2).
__device__ __forceinline__ float4 bload4_l2(Rsrc r, int byte_off) {
  const otg_u4v v = __builtin_amdgcn_raw_buffer_load_b128(r, byte_off, 0, 1);
  const otg_f4v f = __builtin_bit_cast(otg_f4v, v);
  return make_float4(f.x, f.y, f.z, f.w);
}

// Every field of the chunk is assigned on every path (a struct filled in
// part on one path and in part on another is kept in scratch memory).
template <int PH>
__device__ __forceinline__ void load_chunk(const RctParams& P, const Item& it, int c, Chunk& ch,
                                           int lane) {
  const int n = it.pe - it.pb;
  const int q0 = c * kChunkPos;
  const Rsrc rp = make_rsrc(P.pid, it.pb, 4, n);
  const Rsrc rr = make_rsrc(P.pres, it.pb, 4, n);
  const Rsrc rx = PH == PH_TD ? make_rsrc(P.load, it.pb, 16, n) : make_rsrc(P.cap, it.pb, 16, n);
  const Rsrc ry = make_rsrc(PH == PH_TD ? P.scratch + it.scratch : nullptr, 0, 16, n);
#pragma unroll
  for (int s = 0; s < kChunkWin; ++s) {
    const int q = q0 + s * kWave + lane;
    ch.pid[s] = bload_i(rp, q * 4);
    ch.r[s] = bload_f(rr, q * 4);
    if constexpr (PH == PH_TD) {
      ch.x[s] = bload4_l2(rx, q * 16);
      ch.y[s] = bload4_l2(ry, q * 16);
    } else {
      ch.x[s] = bload4(rx, q * 16);
      // y is not used in this phase and left as it is: a value written here
      // could land in a register that a load of the other phase still has
      // in flight, and would wait for that load
    }
  }
  // (read in both phases: chunk 0 of the bottom-up sweep becomes the
  // top-down sweep's first chunk in registers)
  ch.si = bload4(make_rsrc(P.slew_in, it.nb, 16, 1), 0);
}

__device__ __forceinline__ void load_chunk(const RctParams& P, const Item& it, int ph, int c,
                                           Chunk& ch, int lane) {
  if (ph == PH_TD) load_chunk<PH_TD>(P, it, c, ch, lane);
  else load_chunk<PH_BU>(P, it, c, ch, lane);
}

// In-window depth of this lane's position (-1: no position).  Returns the
// largest depth in the window.  One ballot per level; a malformed parent
// (cycle) leaves its lanes at -1 instead of looping.
__device__ __forceinline__ int window_depth(bool valid, int par, int ws, int& dw) {
  const bool inw = valid && par >= ws;
  const int pl = (par - ws) & (kWave - 1);
  dw = (valid && !inw) ? 0 : -1;
  uint64_t front = __ballot(valid && !inw);
  int t = 0;
  for (;;) {
    const bool mine = inw && dw < 0 && ((front >> pl) & 1ull);
    const uint64_t nf = __ballot(mine);
    if (!nf) break;
    ++t;
    if (mine) dw = t;
    front = nf;
  }
  return t;
}

template <int W, bool COMP>
struct SRing {
  StreamLds<W, COMP>& S;
  int base;
  // position i >= base of the item; rings need not be powers of two
  __device__ __forceinline__ int slot(int i) const {
    if constexpr ((W & (W - 1)) == 0) return (i - base) & (W - 1);
    else return (int)((unsigned)(i - base) % (unsigned)W);
  }
  // compensated mode: a slot holds the FP64 values, corners 0-1 in a/b and
  // 2-3 in al/bl (lds_get_d4)
  __device__ __forceinline__ void get(int k, float (&x)[4], float (&y)[4], float (&xl)[4],
                                      float (&yl)[4]) const {
    if constexpr (COMP) {
      lds_get_hl(S.a, S.al, k, x, xl);
      lds_get_hl(S.b, S.bl, k, y, yl);
    } else {
      unpack(S.a[k], x);
      unpack(S.b[k], y);
#pragma unroll
      for (int c = 0; c < 4; ++c) xl[c] = yl[c] = 0.f;
    }
  }
  __device__ __forceinline__ void put(int k, const float (&x)[4], const float (&y)[4],
                                      const float (&xl)[4], const float (&yl)[4]) const {
    if constexpr (COMP) {
      lds_put_hl(S.a, S.al, k, x, xl);
      lds_put_hl(S.b, S.bl, k, y, yl);
    } else {
      S.a[k] = pack(x);
      S.b[k] = pack(y);
    }
  }
  __device__ __forceinline__ void zero(int k) const {
    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    S.a[k] = z;
    S.b[k] = z;
    if constexpr (COMP) {
      S.al[k] = z;
      S.bl[k] = z;
    }
  }
};

// Chain windows (compensated mode).  When every position of a window but
// the first has the previous position as its parent, the window is a path:
// its in-window subtree sums are suffix sums over the lanes and its path
// sums are prefix sums.  Both are taken as wave-wide scans of FP64
// values (6 shuffle stages) instead of one depth step per position (64 for
// a chain).  The compensated sums keep their accuracy in any summation
// order; the FP32 method keeps the sequential order of the CPU timer.
__device__ __forceinline__ bool chain_window(bool valid, int par, int i, int lane) {
  const uint64_t vm = __ballot(valid);
  const uint64_t cm = __ballot(valid && (lane == 0 || par == i - 1));
  return vm != 0 && cm == vm && (vm & 1ull) && __popcll(vm) > 8;
}

// inclusive suffix sums over lanes (lane l: sum of lanes l..63)
__device__ __forceinline__ void suffix_scan(double (&v)[4], int lane) {
  // (stages not unrolled: the scan is inlined into every window of a chunk,
  // and unrolled stages would triple the kernel's code)
#pragma unroll 1
  for (int off = 1; off < kWave; off <<= 1) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double o = __shfl_down(v[c], off);
      if (lane + off < kWave) v[c] += o;
    }
  }
}
// inclusive prefix sums over lanes (lane l: sum of lanes 0..l)
__device__ __forceinline__ void prefix_scan(double (&v)[4], int lane) {
#pragma unroll 1
  for (int off = 1; off < kWave; off <<= 1) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double o = __shfl_up(v[c], off);
      if (lane >= off) v[c] += o;
    }
  }
}

// Bottom-up window: load_i = C_i + sum load_c, L_i = sum (L_c + R_c load_c^2).
// Returns the lane's final load and L rounded to FP32 (ldo, Lo).  The slot of
// every position of the window is cleared once at the window's end (for the
// position W further back): a slot is read only at its own step, and its
// next user gets its first child only in a later window.
//
// A child adds into its parent's slot with a plain read-modify-write (LDS
// float atomics cost several cycles per lane).  Siblings of one step would
// race, so each step runs in rounds: every pending lane writes its lane id
// to claim[parent slot mod 256], and the lanes that read their own id back
// own their parent for the round.  Each round retires at least one lane, and
// two winners never share a parent.
template <int W, bool COMP, bool CHAIN>
__device__ __forceinline__ void bu_window(const SRing<W, COMP>& R, int ws, int pe, int lane,
                                          int pidv, float r, float4 capv, bool root,
                                          float (&ldo)[4], float (&Lo)[4]) {
  const int i = ws + lane;
  const bool valid = i < pe;
  const int par = (valid && !root) ? i - pidv : -1;
  const int sl = R.slot(i);
  const int ps = R.slot(par);
  float cap4[4];
  unpack(capv, cap4);
  if (CHAIN && COMP && chain_window(valid, par, i, lane)) {
    // load_l = sum over lanes j >= l of (C_j + external children of j);
    // w_l = sum over j >= l of (external L of j + R_j load_j^2); lane 0
    // pushes (load, w) to its parent in an earlier window
    float xa[4], xb[4], xal[4], xbl[4];
    R.get(sl, xa, xb, xal, xbl);
    double S[4], V[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) S[c] = valid ? as_f64(xa[c], xal[c]) + (double)cap4[c] : 0.0;
    suffix_scan(S, lane);
#pragma unroll
    for (int c = 0; c < 4; ++c) V[c] = valid ? as_f64(xb[c], xbl[c]) + (double)r * S[c] * S[c] : 0.0;
    suffix_scan(V, lane);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double nx = __shfl_down(V[c], 1);  // w of the in-window child
      if (lane == kWave - 1) nx = 0.0;
      ldo[c] = (float)S[c];
      Lo[c] = (float)(nx + as_f64(xb[c], xbl[c]));
    }
    if (lane == 0 && par >= 0) {
      float pa[4], pb[4], pal[4], pbl[4];
      R.get(ps, pa, pb, pal, pbl);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        split_f64(as_f64(pa[c], pal[c]) + S[c], pa[c], pal[c]);
        split_f64(as_f64(pb[c], pbl[c]) + V[c], pb[c], pbl[c]);
      }
      R.put(ps, pa, pb, pal, pbl);
    }
    wave_sync();
    R.zero(sl);
    wave_sync();
    return;
  }
  int dw;
  const int maxd = window_depth(valid, par, ws, dw);
  const int cs = ps & (kClaim - 1);
#pragma unroll
  for (int c = 0; c < 4; ++c) ldo[c] = Lo[c] = 0.f;
  for (int t = maxd; t >= 0; --t) {
    float ld[4], ldl[4], w[4], wl[4];
    const bool act = dw == t;
    if (act) {
      float xa[4], xb[4], xal[4], xbl[4];
      R.get(sl, xa, xb, xal, xbl);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if constexpr (COMP) {
          // sums and products in FP64
          const double tt = as_f64(xa[c], xal[c]) + (double)cap4[c];
          const double Ld = as_f64(xb[c], xbl[c]);
          split_f64(tt, ld[c], ldl[c]);
          split_f64(Ld + (double)r * tt * tt, w[c], wl[c]);
          ldo[c] = (float)tt;
          Lo[c] = (float)Ld;
        } else {
          ld[c] = cap4[c] + xa[c];
          w[c] = xb[c] + r * ld[c] * ld[c];
          ldl[c] = wl[c] = 0.f;
          ldo[c] = ld[c];
          Lo[c] = xb[c];
        }
      }
    }
    uint64_t pend = __ballot(act && par >= 0);
    while (pend) {
      const bool mine = (pend >> lane) & 1ull;
      if (mine) R.S.claim[cs] = lane;
      wave_sync();
      const bool win = mine && R.S.claim[cs] == lane;
      if (win) {
        float pa[4], pb[4], pal[4], pbl[4];
        R.get(ps, pa, pb, pal, pbl);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          if constexpr (COMP) {
            split_f64(as_f64(pa[c], pal[c]) + as_f64(ld[c], ldl[c]), pa[c], pal[c]);
            split_f64(as_f64(pb[c], pbl[c]) + as_f64(w[c], wl[c]), pb[c], pbl[c]);
          } else {
            pa[c] += ld[c];
            pb[c] += w[c];
          }
        }
        R.put(ps, pa, pb, pal, pbl);
      }
      pend &= ~__ballot(win);
      wave_sync();
    }
  }
  R.zero(sl);
  wave_sync();
}

// Final per-node outputs from delay/impulse (d, m; in the compensated mode
// with dl, ml the (hi, lo) halves of FP64 values), the rounded load and L,
// and the driver slew.
template <bool COMP>
__device__ __forceinline__ void node_outputs(const float (&d)[4], const float (&dl)[4],
                                             const float (&m)[4], const float (&ml)[4],
                                             const float (&ld)[4], const float (&L)[4],
                                             const float4 siv, float (&o_d)[4],
                                             float (&o_ldel)[4], float (&o_b)[4],
                                             float (&o_m)[4], float (&o_s)[4]) {
  float sv[4];
  unpack(siv, sv);
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    if constexpr (COMP) {
      // final per-node combination in FP64
      const double dd = as_f64(d[c], dl[c]);
      const double md = as_f64(m[c], ml[c]);
      const double ldd = (double)ld[c];
      o_d[c] = (float)dd;
      o_m[c] = (float)md;
      o_ldel[c] = (float)(dd * ldd + (double)L[c]);
      o_b[c] = (float)(0.5 * (dd * dd + md));
      // the radicand is formed in FP64 from the compensated impulse; its
      // square root is taken in FP32 (a square root halves the relative
      // error of the rounded radicand)
      const float q = sqrtf((float)((double)sv[c] * (double)sv[c] + md));
      o_s[c] = sv[c] < 0.f ? -q : q;
    } else {
      o_d[c] = d[c];
      o_m[c] = m[c];
      o_ldel[c] = d[c] * ld[c] + L[c];
      o_b[c] = 0.5f * (d[c] * d[c] + m[c]);
      const float q = sqrtf(sv[c] * sv[c] + m[c]);
      o_s[c] = sv[c] < 0.f ? -q : q;
    }
  }
}

// Top-down increments of one node: u = R load (added to the delay) and
// z = (R load)^2 + 2 R L (added to the impulse); FP64 in the compensated
// mode (the product of two floats is exact in FP64).  They depend on the
// node alone, so a window computes them once for all its lanes before its
// depth steps.
template <bool COMP>
struct TdIncr {
  using T = typename std::conditional<COMP, double, float>::type;
  T u[4], z[4];
};

template <bool COMP>
__device__ __forceinline__ TdIncr<COMP> td_incr(float rr, const float (&ld)[4], const float (&L)[4]) {
  TdIncr<COMP> t;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    if constexpr (COMP) {
      t.u[c] = (double)rr * (double)ld[c];
      t.z[c] = t.u[c] * t.u[c] + 2.0 * (double)rr * (double)L[c];
    } else {
      t.u[c] = rr * ld[c];
      t.z[c] = t.u[c] * t.u[c] + 2.0f * rr * L[c];
    }
  }
  return t;
}

// delay = delay_parent + u, impulse = impulse_parent + z
template <bool COMP>
__device__ __forceinline__ void td_accum(const TdIncr<COMP>& t, const float (&pd)[4],
                                         const float (&pdl)[4], const float (&pm)[4],
                                         const float (&pml)[4], float (&d)[4], float (&dl)[4],
                                         float (&m)[4], float (&ml)[4]) {
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    if constexpr (COMP) {
      split_f64(as_f64(pd[c], pdl[c]) + t.u[c], d[c], dl[c]);
      split_f64(as_f64(pm[c], pml[c]) + t.z[c], m[c], ml[c]);
    } else {
      d[c] = pd[c] + t.u[c];
      m[c] = pm[c] + t.z[c];
      dl[c] = ml[c] = 0.f;
    }
  }
}

// Top-down window: delay_i = delay_p + R_i load_i,
// impulse_i = impulse_p + (R_i load_i)^2 + 2 R_i L_i; then the lane's
// outputs.  ld/L are the lane's rounded bottom-up results.
template <int W, bool COMP, bool CHAIN>
__device__ __forceinline__ void td_window(const SRing<W, COMP>& R, int ws, int pe, int lane,
                                          int pidv, float r, bool root, const float (&ld)[4],
                                          const float (&L)[4], float (&o_d)[4],
                                          float (&o_ldel)[4], float (&o_b)[4], float (&o_m)[4],
                                          float (&o_s)[4], const float4 siv) {
  const int i = ws + lane;
  const bool valid = i < pe;
  const int par = (valid && !root) ? i - pidv : -1;
  const float rr = root ? 0.f : r;
  const int sl = R.slot(i);
  const int ps = R.slot(par);
  const TdIncr<COMP> inc = td_incr<COMP>(rr, ld, L);
  if (CHAIN && COMP && chain_window(valid, par, i, lane)) {
    // delay_l = delay of lane 0's parent + sum over lanes j <= l of u_j
    // (impulse likewise with z)
    double D[4], M[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      D[c] = valid ? (double)inc.u[c] : 0.0;
      M[c] = valid ? (double)inc.z[c] : 0.0;
    }
    if (lane == 0 && par >= 0) {
      float pd[4], pm[4], pdl[4], pml[4];
      R.get(ps, pd, pm, pdl, pml);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        D[c] += as_f64(pd[c], pdl[c]);
        M[c] += as_f64(pm[c], pml[c]);
      }
    }
    prefix_scan(D, lane);
    prefix_scan(M, lane);
    float d[4], dl[4], m[4], ml[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      split_f64(D[c], d[c], dl[c]);
      split_f64(M[c], m[c], ml[c]);
    }
    if (valid) R.put(sl, d, m, dl, ml);
    wave_sync();
    node_outputs<COMP>(d, dl, m, ml, ld, L, siv, o_d, o_ldel, o_b, o_m, o_s);
    return;
  }
  int dw;
  const int maxd = window_depth(valid, par, ws, dw);
  float d[4], dl[4], m[4], ml[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) d[c] = dl[c] = m[c] = ml[c] = 0.f;
  // depth 0: parents in earlier windows (the driver reads the ring's zero slot)
  if (dw == 0) {
    float pd[4], pm[4], pdl[4], pml[4];
    R.get(par >= 0 ? ps : W, pd, pm, pdl, pml);
    td_accum<COMP>(inc, pd, pdl, pm, pml, d, dl, m, ml);
    R.put(sl, d, m, dl, ml);
  }
  wave_sync();
  for (int t = 1; t <= maxd; ++t) {
    if (dw == t) {
      float pd[4], pm[4], pdl[4], pml[4];
      R.get(ps, pd, pm, pdl, pml);
      td_accum<COMP>(inc, pd, pdl, pm, pml, d, dl, m, ml);
      R.put(sl, d, m, dl, ml);
    }
    wave_sync();
  }
  node_outputs<COMP>(d, dl, m, ml, ld, L, siv, o_d, o_ldel, o_b, o_m, o_s);
}

struct OutBufs {
  Rsrc load, delay, ldelay, beta, imp, slew, gl;
};

__device__ __forceinline__ OutBufs out_bufs(const RctParams& P, const Item& it) {
  const int n = it.pe - it.pb;
  return OutBufs{make_rsrc(P.load, it.pb, 16, n),    make_rsrc(P.delay, it.pb, 16, n),
                 make_rsrc(P.ldelay, it.pb, 16, n),  make_rsrc(P.beta, it.pb, 16, n),
                 make_rsrc(P.impulse, it.pb, 16, n), make_rsrc(P.slew, it.pb, 16, n),
                 it.kind == ITEM_STREAM ? make_rsrc(P.scratch + it.scratch, 0, 16, n)
                                        : make_rsrc(nullptr, 0, 16, 0)};
}

// One chunk of a streamed net's bottom-up sweep (windows last to first).  If
// `take`, `ch` becomes the next chunk `nx` (requested a chunk earlier) after
// the chunk's wait.
template <int W, bool COMP, bool CHAIN>
__device__ __forceinline__ void run_bu_chunk(StreamLds<W, COMP>& S, const RctParams& P,
                                             const Item& it, int c, Chunk& ch, bool take,
                                             const Chunk& nx, int lane) {
  const SRing<W, COMP> R{S, it.pb};
  const OutBufs B = out_bufs(P, it);
  const int n = it.pe - it.pb;
#pragma unroll
  for (int s = kChunkWin - 1; s >= 0; --s) {
    const int q0 = c * kChunkPos + s * kWave;
    // window 0 exists in every chunk; it is processed last
    if (s == 0 || q0 < n) {
      float ldo[4], Lo[4];
      bu_window<W, COMP, CHAIN>(R, it.pb + q0, it.pe, lane, ch.pid[s], ch.r[s], ch.x[s], q0 + lane == 0,
                         ldo, Lo);
      // chunk 0 stays in registers as the top-down sweep's first chunk
      ch.x[s] = make_float4(ldo[0], ldo[1], ldo[2], ldo[3]);
      ch.y[s] = make_float4(Lo[0], Lo[1], Lo[2], Lo[3]);
      if (s == 0) {
        // the chunk's one wait, on every path: before its last window's
        // stores (the next chunk was requested a whole chunk earlier,
        // earlier windows' stores long issued); the next chunk is taken
        // over right after it
        __builtin_amdgcn_s_waitcnt(0);
        if (take) ch = nx;
      }
      bstore4(B.load, (q0 + lane) * 16, ldo);
      bstore4(B.gl, (q0 + lane) * 16, Lo);
    }
  }
}

// One chunk of a streamed net's top-down sweep (windows first to last).  The
// last window's outputs are stored after the chunk's one wait, which sits on
// every path; the next chunk is taken over right after that wait.
template <int W, bool COMP, bool CHAIN>
__device__ __forceinline__ void run_td_chunk(StreamLds<W, COMP>& S, const RctParams& P,
                                             const Item& it, int c, Chunk& ch, bool take,
                                             const Chunk& nx, int lane) {
  const SRing<W, COMP> R{S, it.pb};
  const OutBufs B = out_bufs(P, it);
  const int n = it.pe - it.pb;
  // the last window's outputs (every chunk has one; left undefined until
  // then, so that no instruction writes these registers early)
  const float u = __builtin_nondeterministic_value(0.f);
  float p_d[4] = {u, u, u, u}, p_ldel[4] = {u, u, u, u}, p_b[4] = {u, u, u, u},
        p_m[4] = {u, u, u, u}, p_s[4] = {u, u, u, u};
  int p_off = __builtin_nondeterministic_value(0);
#pragma unroll
  for (int s = 0; s < kChunkWin; ++s) {
    const int q0 = c * kChunkPos + s * kWave;
    if (s == 0 || q0 < n) {
      float ld[4], L[4];
      unpack(ch.x[s], ld);
      unpack(ch.y[s], L);
      float o_d[4], o_ldel[4], o_b[4], o_m[4], o_s[4];
      td_window<W, COMP, CHAIN>(R, it.pb + q0, it.pe, lane, ch.pid[s], ch.r[s], q0 + lane == 0, ld, L,
                         o_d, o_ldel, o_b, o_m, o_s, ch.si);
      const int off = (q0 + lane) * 16;
      if (s < kChunkWin - 1 && q0 + kWave < n) {
        bstore4(B.delay, off, o_d);
        bstore4(B.ldelay, off, o_ldel);
        bstore4(B.beta, off, o_b);
        bstore4(B.imp, off, o_m);
        bstore4(B.slew, off, o_s);
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          p_d[j] = o_d[j];
          p_ldel[j] = o_ldel[j];
          p_b[j] = o_b[j];
          p_m[j] = o_m[j];
          p_s[j] = o_s[j];
        }
        p_off = off;
      }
    }
  }
  __builtin_amdgcn_s_waitcnt(0);
  if (take) ch = nx;
  bstore4(B.delay, p_off, p_d);
  bstore4(B.ldelay, p_off, p_ldel);
  bstore4(B.beta, p_off, p_b);
  bstore4(B.imp, p_off, p_m);
  bstore4(B.slew, p_off, p_s);
}

// Work items come from the largest-first list in one of two ways:
//  * dealt out (packed groups, whose costs are alike): in round r workgroup
//    w takes item r*G + w for even r and r*G + G-1-w for odd r (G
//    workgroups), which evens out the totals of a sorted list without a
//    shared counter;
//  * from a queue (streamed nets, whose costs vary with their shapes): a
//    counter in global memory, incremented once per item.  The increment
//    for an item is issued one item before its index is read, so the wait
//    for it is the chunk wait that the item in between has anyway.
// Descriptors are read through the scalar cache (the list is read-only
// during a launch) two items ahead, so the next item is always known when
// its first loads are requested.  Scalar loads count against lgkmcnt, which
// every LDS wait drains in passing; a vector load of the descriptor would be
// waited for with vmcnt, i.e. together with every store still in flight.
typedef const __attribute__((address_space(4))) RctItem ConstItem;

template <bool QUEUE>
struct ItemPipe {
  const RctItem* items;
  int total;
  int32_t* queue;  // QUEUE: next item index, finished workgroups
  int lane;
  int round = 0;    // dealt: round of n2
  int q = 0;        // queue: index after n2 (lane 0; requested, read one item later)
  Item n1{}, n2{};  // the next two items of this workgroup (past the end: empty)
  bool v1 = false, v2 = false;

  __device__ __forceinline__ int dealt(int r) const {
    const int g = gridDim.x, w = blockIdx.x;
    return r * g + ((r & 1) ? g - 1 - w : w);
  }
  __device__ __forceinline__ int pop() const {
    int v = 0;
    if (lane == 0) v = atomicAdd(queue, 1);
    return v;
  }
  __device__ __forceinline__ Item fetch(int i) const {
    Item it{};
    if (i < total) {
      ConstItem* d = (ConstItem*)items + i;
      it.pb = d->pos_begin;
      it.pe = d->pos_end;
      it.nb = d->net_begin;
      it.ne = d->net_end;
      it.kind = d->kind;
      it.k = d->k;
      it.scratch = d->scratch;
    }
    return it;
  }
  // Starts the pipeline; returns the first item (valid if *have).
  __device__ __forceinline__ Item start(bool* have) {
    int i0, i1, i2;
    if constexpr (QUEUE) {
      int v = 0;
      if (lane == 0) v = atomicAdd(queue, 3);
      const int b = __builtin_amdgcn_readfirstlane(v);
      i0 = b;
      i1 = b + 1;
      i2 = b + 2;
    } else {
      i0 = dealt(0);
      i1 = dealt(1);
      i2 = dealt(2);
      round = 2;
    }
    *have = i0 < total;
    const Item first = fetch(i0);
    v1 = i1 < total;
    n1 = fetch(i1);
    v2 = i2 < total;
    n2 = fetch(i2);
    return first;
  }
  __device__ __forceinline__ bool has_next() const { return v1; }
  // Moves to the next item; false at the end of this workgroup's share.
  // (Dealt indices grow with the round and queue indices with every
  // request, so the first index past the end ends the share.)
  __device__ __forceinline__ bool next(Item& it) {
    if (!v1) return false;
    it = n1;
    n1 = n2;
    v1 = v2;
    int i;
    if constexpr (QUEUE) {
      i = __builtin_amdgcn_readfirstlane(pop());
    } else {
      i = dealt(++round);
    }
    v2 = i < total;
    n2 = fetch(i);
    return true;
  }
};

__device__ __forceinline__ bool stream_first_chunk(const Item& it, int& ph, int& c) {
  if (it.pe <= it.pb || it.kind != ITEM_STREAM) return false;
  ph = PH_BU;
  c = num_chunks(it) - 1;
  return true;
}

template <int W, bool COMP, bool CHAIN>
__global__ __launch_bounds__(64) void rct_stream_kernel(RctParams P) {
  __shared__ StreamLds<W, COMP> S;
  const int lane = threadIdx.x;
  const int total = P.num_items_dev ? gld(P.num_items_dev) : P.num_items;
  ItemPipe<true> pipe{P.items, total, P.queue, lane};
  bool have = false;
  Item it = pipe.start(&have);
  int ph = 0, c = 0;
  while (have && !stream_first_chunk(it, ph, c)) have = pipe.next(it);
  // Chunk data: `cur` is being processed, `nxt` is in flight.  A wavefront
  // that has both loads and stores outstanding can only wait for all of
  // them, so each chunk waits once, just before the stores of its last
  // window, on every path through the chunk: by then the next chunk's loads
  // and the earlier windows' stores were issued long ago, and the last
  // window's stores stay in flight during the next chunk.  `nxt` is copied
  // to `cur` right after that wait, where nothing is outstanding, so the
  // copy needs no wait of its own.  The bottom-up sweep's chunk 0 (the last
  // one it processes) becomes the top-down sweep's chunk 0 in registers.
  Chunk cur, nxt;
  int nxt_pb = -1, nxt_ph = -1, nxt_c = -1;  // what nxt holds
  if (have) {
    load_chunk(P, it, ph, c, cur, lane);
    __builtin_amdgcn_s_waitcnt(0);
  }
  while (have) {
    const int nc = num_chunks(it);
    // ---- request the chunk after this one --------------------------------
    Item pit = it;
    int pph = -1, pc = -1;
    bool pv = false;
    if (ph == PH_BU && c > 0) {
      pph = PH_BU;
      pc = c - 1;
      pv = true;
    } else if (ph == PH_TD && c + 1 < nc) {
      pph = PH_TD;
      pc = c + 1;
      pv = true;
    } else if (ph == PH_TD && pipe.has_next()) {
      // the current item's last chunk: the next item's first chunk
      // (top-down chunk 1 of a longer net is requested after the bottom-up
      // chunk 0, whose wait completes the stores it reads; the bottom-up
      // chunk 0 requests nothing, since its registers become the top-down
      // sweep's chunk 0 and a chunk taken over there would replace them --
      // for a net of one chunk too)
      pit = pipe.n1;
      pv = stream_first_chunk(pit, pph, pc);
    }
    if (pv && !(pit.pb == nxt_pb && pph == nxt_ph && pc == nxt_c)) {
      load_chunk(P, pit, pph, pc, nxt, lane);
      nxt_pb = pit.pb;
      nxt_ph = pph;
      nxt_c = pc;
    }
    // the next chunk is in nxt (always, except after the bottom-up chunk 0,
    // and at an item's end when the next item was not known in time)
    const bool take = pv && pit.pb == nxt_pb && pph == nxt_ph && pc == nxt_c;
    // ---- process ----------------------------------------------------------
    if (ph == PH_BU) {
      if (c == nc - 1) {
        const SRing<W, COMP> R{S, it.pb};
        for (int k = lane; k <= W; k += kWave) R.zero(k);
        wave_sync();
      }
      run_bu_chunk<W, COMP, CHAIN>(S, P, it, c, cur, take, nxt, lane);
    } else {
      run_td_chunk<W, COMP, CHAIN>(S, P, it, c, cur, take, nxt, lane);
    }
    if (take) nxt_pb = -1;
    // ---- advance ------------------------------------------------------------
    if (ph == PH_BU && c == 0) {
      if (nc > 1) {
        load_chunk<PH_TD>(P, it, 1, nxt, lane);
        nxt_pb = it.pb;
        nxt_ph = PH_TD;
        nxt_c = 1;
      }
      ph = PH_TD;  // chunk 0 is in cur already
      c = 0;
      continue;
    }
    if (ph == PH_BU) {
      --c;
    } else if (c + 1 < nc) {
      ++c;
    } else {
      bool found = false;
      while (pipe.next(it)) {
        if (stream_first_chunk(it, ph, c)) {
          found = true;
          break;
        }
      }
      if (!found) break;
    }
    if (!take) {
      // not requested ahead (an item that was not known in time)
      __builtin_amdgcn_s_waitcnt(0);
      load_chunk(P, it, ph, c, cur, lane);
      __builtin_amdgcn_s_waitcnt(0);
    }
  }
  queue_release(P.queue, lane);
}

// ===========================================================================
// Packed-group kernel: groups of whole small nets (at most G positions).
//
// A group is processed entirely on chip in a static schedule built with the
// plan: its nodes sorted by depth, then by parent, then by position.  One
// 32-bit word per schedule index j holds
//     bits  0..7  the node (group-relative position) at j
//     bits  8..15 its depth
//     bits 16..23 schedule index of its first child
//     bits 24..31 its number of children
// (children of a node are consecutive in the schedule: same depth, same
// parent).  A step takes up to 64 consecutive schedule indices of one depth,
// so every lane of a step has work:
//  * bottom-up, deepest first: a node pulls the (load, w) of its children
//    from their schedule slots, w_c = L_c + R_c load_c^2, and writes its own
//    (no read-modify-write, no conflicts);
//  * top-down, depth 0 first: a node reads its parent's delay and impulse
//    (stored by position) and adds its own increments;
//  * a last pass with lanes on consecutive positions forms the six outputs
//    and stores them coalesced.
// Steps per group: about (group size / 64) + depth per sweep, against one
// step per depth level per 64-position window otherwise.
// ===========================================================================
template <int G, bool COMP>
struct GroupLds {
  static constexpr int K = G / kWave;        // positions per lane
  static constexpr int NS = G / (2 * kWave);  // driver slews per lane (<= G/2 nets)
  // In the compensated mode an FP64 quantity takes two arrays, corners 0-1
  // in the first and 2-3 in the second (lds_get_d4):
  // load (ld, ldl), w (w, wl), delay (cap, wl), impulse (w, ldl).
  float4 cap[G];   // bottom-up: cap by position   | top-down: delay by position
  float4 w[G];     // bottom-up: w by schedule      | top-down: impulse by position
  float4 ld[G];    // load by schedule index, rounded to FP32 after the sweep
  float4 L[G];     // L rounded to FP32 by schedule index
  float4 ldl[COMP ? G : 1];  // bottom-up: load by schedule index |
                             // top-down: impulse by position
  float4 wl[COMP ? G : 1];   // bottom-up: w | top-down: delay by position
  float r[G];
  int32_t pid[G];
  uint32_t pw[G];
  int32_t net[G];           // group-relative net of each position (top-down)
  int32_t pos2j[G];         // schedule index of each position
  float4 si[G / 2];
};

template <int G>
struct GroupChunk {
  static constexpr int K = G / kWave;
  static constexpr int NS = G / (2 * kWave);
  int32_t pid[K];
  float r[K];
  float4 cap[K];
  uint32_t pw[K];
  float4 si[NS];
};

template <int G>
__device__ __forceinline__ void load_group(const RctParams& P, const Item& it, GroupChunk<G>& ch,
                                           int lane) {
  constexpr int K = GroupChunk<G>::K;
  const int n = it.pe - it.pb;
  const int nn = it.ne - it.nb;
  const Rsrc rp = make_rsrc(P.pid, it.pb, 4, n);
  const Rsrc rr = make_rsrc(P.pres, it.pb, 4, n);
  const Rsrc rc = make_rsrc(P.cap, it.pb, 16, n);
  const Rsrc rw = make_rsrc(P.sched, it.scratch, 4, n);
  const Rsrc rs = make_rsrc(P.slew_in, it.nb, 16, nn);
#pragma unroll
  for (int s = 0; s < K; ++s) {
    const int q = s * kWave + lane;
    ch.pid[s] = bload_i(rp, q * 4);
    ch.r[s] = bload_f(rr, q * 4);
    ch.cap[s] = bload4(rc, q * 16);
    ch.pw[s] = (uint32_t)bload_i(rw, q * 4);
  }
#pragma unroll
  for (int s = 0; s < GroupChunk<G>::NS; ++s) ch.si[s] = bload4(rs, (s * kWave + lane) * 16);
}

template <int G, bool COMP>
__device__ __forceinline__ void group_stage(GroupLds<G, COMP>& S, const Item& it,
                                            const GroupChunk<G>& ch, int lane) {
  constexpr int K = GroupChunk<G>::K;
  const int n = it.pe - it.pb;
  // ---- stage the group ------------------------------------------------------
#pragma unroll
  for (int s = 0; s < K; ++s) {
    const int q = s * kWave + lane;
    S.cap[q] = ch.cap[s];
    S.r[q] = ch.r[s];
    S.pid[q] = ch.pid[s];
    S.pw[q] = ch.pw[s];
    if (q < n) S.pos2j[ch.pw[s] & 0xffu] = q;
  }
#pragma unroll
  for (int s = 0; s < GroupChunk<G>::NS; ++s) S.si[s * kWave + lane] = ch.si[s];
  wave_sync();
}

// Both sweeps of a staged group (LDS only).
template <int G, bool COMP>
__device__ __forceinline__ void group_sweeps(GroupLds<G, COMP>& S, const Item& it, int lane) {
  constexpr int kPull = 1;  // children fetched per batch (a step loops to its largest child count)
  const int n = it.pe - it.pb;
  // ---- bottom-up, deepest schedule entries first ------------------------------
  uint32_t pwj = n - kWave + lane >= 0 ? S.pw[n - kWave + lane] : 0u;
  for (int e = n; e > 0;) {
    const int j = e - kWave + lane;
    const int dj = (int)((pwj >> 8) & 0xffu);
    const int dlast = __builtin_amdgcn_readlane(dj, kWave - 1);  // depth of schedule entry e-1
    const uint64_t m = __ballot(j >= 0 && dj == dlast);
    const int b = e - __popcll(m);
    // schedule words of the next step, requested before this step's work
    const uint32_t pwn = b - kWave + lane >= 0 ? S.pw[b - kWave + lane] : 0u;
    if (j >= b) {
      const int q = (int)(pwj & 0xffu);
      const int cf = (int)((pwj >> 16) & 0xffu);
      const int cn = (int)(pwj >> 24);
      float cap4[4];
      unpack(S.cap[q], cap4);
      const float r = S.r[q];
      using T = typename std::conditional<COMP, double, float>::type;
      T ld[4], sw[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        ld[c] = (T)cap4[c];
        sw[c] = (T)0;
      }
      for (int k0 = 0; k0 < cn; k0 += kPull) {
#pragma unroll
        for (int u = 0; u < kPull; ++u) {
          const int jc = cf + k0 + u;
          if (k0 + u < cn) {
            if constexpr (COMP) {
              double a1[4], a2[4];
              lds_get_d4(S.ld, S.ldl, jc, a1);
              lds_get_d4(S.w, S.wl, jc, a2);
#pragma unroll
              for (int c = 0; c < 4; ++c) {
                ld[c] += a1[c];
                sw[c] += a2[c];
              }
            } else {
              float a1[4], a2[4];
              unpack(S.ld[jc], a1);
              unpack(S.w[jc], a2);
#pragma unroll
              for (int c = 0; c < 4; ++c) {
                ld[c] += a1[c];
                sw[c] += a2[c];
              }
            }
          }
        }
      }
      float o_L[4];
      if constexpr (COMP) {
        double o_w[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          o_w[c] = sw[c] + (double)r * ld[c] * ld[c];
          o_L[c] = (float)sw[c];
        }
        lds_put_d4(S.ld, S.ldl, j, ld);
        lds_put_d4(S.w, S.wl, j, o_w);
      } else {
        float o_ld[4], o_w[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          o_ld[c] = ld[c];
          o_w[c] = sw[c] + r * ld[c] * ld[c];
          o_L[c] = sw[c];
        }
        S.ld[j] = pack(o_ld);
        S.w[j] = pack(o_w);
      }
      S.L[j] = pack(o_L);
    }
    wave_sync();
    e = b;
    pwj = pwn;
  }
  // compensated mode: each node's load rounded to FP32 for the top-down sweep
  // and the outputs (which reuse ldl by position)
  if constexpr (COMP) {
    for (int j = lane; j < n; j += kWave) {
      double v[4];
      float o[4];
      lds_get_d4(S.ld, S.ldl, j, v);
#pragma unroll
      for (int c = 0; c < 4; ++c) o[c] = (float)v[c];
      S.ld[j] = pack(o);
    }
    wave_sync();
  }
  // ---- top-down, depth 0 first: delay, impulse and net by position -----------
  pwj = lane < n ? S.pw[lane] : 0xffffffffu;
  for (int b = 0; b < n;) {
    const int j = b + lane;
    const int dj = (int)((pwj >> 8) & 0xffu);
    const int dfirst = __builtin_amdgcn_readfirstlane(dj);  // depth of schedule entry b
    const uint64_t m = __ballot(j < n && dj == dfirst);
    const int e = b + __popcll(m);
    const uint32_t pwn = e + lane < n ? S.pw[e + lane] : 0xffffffffu;
    if (j < e) {
      const int q = (int)(pwj & 0xffu);
      const bool root = dfirst == 0;
      const float rr = root ? 0.f : S.r[q];
      const int p = root ? 0 : q - S.pid[q];
      // drivers are the depth-0 schedule entries, in position (= net) order
      const int net = root ? j : S.net[p];
      float ldv[4], Lv[4];
      unpack(S.ld[j], ldv);
      unpack(S.L[j], Lv);
      float pd[4], pm[4], pdl[4], pml[4];
      if (!root) {
        if constexpr (COMP) {
          lds_get_hl(S.cap, S.wl, p, pd, pdl);
          lds_get_hl(S.w, S.ldl, p, pm, pml);
        } else {
          unpack(S.cap[p], pd);
          unpack(S.w[p], pm);
#pragma unroll
          for (int c = 0; c < 4; ++c) pdl[c] = pml[c] = 0.f;
        }
      } else {
#pragma unroll
        for (int c = 0; c < 4; ++c) pd[c] = pm[c] = pdl[c] = pml[c] = 0.f;
      }
      float d[4], dl[4], mm[4], ml[4];
      td_accum<COMP>(td_incr<COMP>(rr, ldv, Lv), pd, pdl, pm, pml, d, dl, mm, ml);
      if constexpr (COMP) {
        lds_put_hl(S.cap, S.wl, q, d, dl);
        lds_put_hl(S.w, S.ldl, q, mm, ml);
      } else {
        S.cap[q] = pack(d);
        S.w[q] = pack(mm);
      }
      S.net[q] = net;
    }
    wave_sync();
    b = e;
    pwj = pwn;
  }
}

// Outputs of a swept group, lanes on consecutive positions (coalesced stores).
template <int G, bool COMP>
__device__ __forceinline__ void group_outputs(GroupLds<G, COMP>& S, const RctParams& P,
                                              const Item& it, int lane) {
  constexpr int K = GroupChunk<G>::K;
  const int n = it.pe - it.pb;
  const Rsrc rout[6] = {make_rsrc(P.load, it.pb, 16, n),   make_rsrc(P.delay, it.pb, 16, n),
                        make_rsrc(P.ldelay, it.pb, 16, n), make_rsrc(P.beta, it.pb, 16, n),
                        make_rsrc(P.impulse, it.pb, 16, n), make_rsrc(P.slew, it.pb, 16, n)};
#pragma unroll
  for (int s = 0; s < K; ++s) {
    const int q = s * kWave + lane;
    if (s * kWave < n) {
      const int j = q < n ? S.pos2j[q] : 0;
      const int net = q < n ? S.net[q] : 0;
      float ldv[4], Lv[4], d[4], dl[4], mm[4], ml[4];
      unpack(S.ld[j], ldv);
      unpack(S.L[j], Lv);
      if constexpr (COMP) {
        lds_get_hl(S.cap, S.wl, q, d, dl);
        lds_get_hl(S.w, S.ldl, q, mm, ml);
      } else {
        unpack(S.cap[q], d);
        unpack(S.w[q], mm);
#pragma unroll
        for (int c = 0; c < 4; ++c) dl[c] = ml[c] = 0.f;
      }
      float o_d[4], o_ldel[4], o_b[4], o_m[4], o_s[4];
      node_outputs<COMP>(d, dl, mm, ml, ldv, Lv, S.si[net < G / 2 ? net : G / 2 - 1], o_d,
                         o_ldel, o_b, o_m, o_s);
      const int off = q * 16;
      bstore4(rout[0], off, ldv);
      bstore4(rout[1], off, o_d);
      bstore4(rout[2], off, o_ldel);
      bstore4(rout[3], off, o_b);
      bstore4(rout[4], off, o_m);
      bstore4(rout[5], off, o_s);
    }
  }
  wave_sync();
}

// QUEUE: groups are taken from a queue (the kernel runs beside streaming
// launches, so its workgroups start at different times); otherwise they are
// dealt out to a grid that is resident from the start.
template <int G, bool COMP, bool QUEUE>
__global__ __launch_bounds__(64) void rct_group_kernel(RctParams P) {
  __shared__ GroupLds<G, COMP> S;
  const int lane = threadIdx.x;
  const int total = P.num_items_dev ? gld(P.num_items_dev) : P.num_items;
  ItemPipe<QUEUE> pipe{P.items, total, P.queue, lane};
  bool have = false;
  Item it = pipe.start(&have);
  // One register buffer: a group's inputs are staged into LDS first, after
  // which the buffer takes the next group's inputs while this group is
  // swept.  The single wait per group sits between the sweeps and the
  // output stores, so it covers loads and stores issued a whole group
  // earlier, and this group's stores stay in flight during the next group.
  GroupChunk<G> R;
  if (have) {
    load_group<G>(P, it, R, lane);
    __builtin_amdgcn_s_waitcnt(0);
  }
  while (have) {
    group_stage<G, COMP>(S, it, R, lane);
    // (every listed group is non-empty, and an empty one would be solved as
    // a no-op, so the next item is taken without a skip loop: a loop here
    // would make the compiler wait for the descriptor loads it issues)
    Item nit = it;
    const bool nhave = pipe.next(nit);
    if (nhave) load_group<G>(P, nit, R, lane);
    group_sweeps<G, COMP>(S, it, lane);
    __builtin_amdgcn_s_waitcnt(0);
    group_outputs<G, COMP>(S, P, it, lane);
    // a group listed by a subset solve is released for the next call
    if (P.group_flag && lane == 0) P.group_flag[it.k] = 0;
    it = nit;
    have = nhave;
  }
  if constexpr (QUEUE) queue_release(P.queue, lane);
}

// Group capacity (positions) of a plan: 128 or 256, from small_target.
inline int group_cap(int small_target) { return small_target <= 128 ? 128 : 256; }

// Subset solves, streamed nets: the listed nets that a streaming kernel
// solves, appended to the list of the first bin of their kind (chain net or
// not) whose ring holds their parent offsets (lim[b] = ring - 64; counts[b]
// counts bin b's list).
struct SubBins {
  RctItem* items[otg_rct_plan::kMaxBins];
  int32_t lim[otg_rct_plan::kMaxBins];
  int32_t chain[otg_rct_plan::kMaxBins];
  int32_t num_bins;
};

__global__ void rct_subset_items_kernel(int32_t count, const int32_t* __restrict__ subset,
                                        int32_t num_nets, const int32_t* __restrict__ nodes_start,
                                        const int32_t* __restrict__ net_maxpid,
                                        const int64_t* __restrict__ net_scratch,
                                        const int32_t* __restrict__ net_chain,
                                        int32_t group_max, SubBins bins,
                                        int32_t* __restrict__ counts) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  const int net = subset[j];
  if (net < 0 || net >= num_nets) return;
  const int pb = nodes_start[net], pe = nodes_start[net + 1];
  if (pe - pb <= group_max) return;
  const int mp = net_maxpid[net];
  const int ch = net_chain[net];
  for (int b = 0; b < bins.num_bins; ++b) {
    if (mp <= bins.lim[b] && bins.chain[b] == ch) {
      bins.items[b][atomicAdd(counts + b, 1)] =
          RctItem{pb, pe, net, net + 1, ITEM_STREAM, 1, net_scratch[net]};
      return;
    }
  }
}

// Subset solves, packed nets: the groups holding the listed nets, each once.
// A group is claimed through its flag (0 -> 1); the group kernel clears the
// flag after solving it, so every call starts from all-clear flags.
__global__ void rct_subset_groups_kernel(int32_t count, const int32_t* __restrict__ subset,
                                         int32_t num_nets, const int32_t* __restrict__ net_group,
                                         int32_t* __restrict__ group_flag,
                                         const RctItem* __restrict__ groups,
                                         RctItem* __restrict__ out, int32_t* __restrict__ out_count) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  const int net = subset[j];
  if (net < 0 || net >= num_nets) return;
  const int g = net_group[net];
  if (g < 0) return;
  if (atomicCAS(&group_flag[g], 0, 1) != 0) return;
  out[atomicAdd(out_count, 1)] = groups[g];
}

// Clears a small device counter (a kernel rather than a memset node, so that
// captured graphs replay it correctly).
__global__ void rct_zero_kernel(int32_t* p, int32_t n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) p[j] = 0;
}

template <int W, bool COMP, bool CHAIN>
void* stream_kernel_ptr() {
  return reinterpret_cast<void*>(&rct_stream_kernel<W, COMP, CHAIN>);
}

template <bool COMP, bool CHAIN>
void* pick_stream_t(int W) {
  switch (W) {
    case 256: return stream_kernel_ptr<256, COMP, CHAIN>();
    case 320: return stream_kernel_ptr<320, COMP, CHAIN>();
    case 1024: return stream_kernel_ptr<1024, COMP, CHAIN>();
    case 2048: return stream_kernel_ptr<2048, COMP, CHAIN>();
    default: return stream_kernel_ptr<512, COMP, CHAIN>();
  }
}

// The chain variant (wave-wide scans for windows that are paths) exists for
// the compensated method only; it is a separate kernel so that the other
// nets run the smaller code.
void* pick_stream(int W, bool comp, bool chain = false) {
  if (!comp) return pick_stream_t<false, false>(W);
  return chain ? pick_stream_t<true, true>(W) : pick_stream_t<true, false>(W);
}

template <bool QUEUE>
void* pick_group_t(int G, bool comp) {
  if (comp)
    return G <= 128 ? reinterpret_cast<void*>(&rct_group_kernel<128, true, QUEUE>)
                    : reinterpret_cast<void*>(&rct_group_kernel<256, true, QUEUE>);
  return G <= 128 ? reinterpret_cast<void*>(&rct_group_kernel<128, false, QUEUE>)
                  : reinterpret_cast<void*>(&rct_group_kernel<256, false, QUEUE>);
}

void* pick_group(int G, bool comp, bool queue = false) {
  return queue ? pick_group_t<true>(G, comp) : pick_group_t<false>(G, comp);
}

void* pick_global(bool comp) {
  return comp ? reinterpret_cast<void*>(&rct_global_kernel<true>)
              : reinterpret_cast<void*>(&rct_global_kernel<false>);
}

int blocks_per_cu(void* fn) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, 64, 0) != hipSuccess || nb <= 0) nb = 1;
  return nb;
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
otg_rct_tuning otg_rct_resolve_tuning(const otg_context* ctx, const otg_rct_tuning* t,
                                      otg_rct_method m) {
  otg_rct_tuning r{};
  if (t) r = *t;
  const otg_rct_tuning& d = ctx->tuning;
  auto ring_ok = [](int w) {
    for (int k = 0; k < kRingSteps; ++k)
      if (kRingLadder[k] == w) return true;
    return false;
  };
  // the largest ring by default: a bin holding no nets costs nothing, and
  // nets with parent offsets up to 1984 (high-fanout nets) stay on chip
  if (!r.ring_slots) r.ring_slots = d.ring_slots ? d.ring_slots : 2048;
  if (!ring_ok(r.ring_slots)) r.ring_slots = 2048;
  if (!r.small_target) r.small_target = d.small_target ? d.small_target : 256;
  // a packed group holds at most 256 positions (8-bit schedule fields)
  r.small_target = std::max(64, std::min(r.small_target, 256));
  if (!r.grid_per_cu) r.grid_per_cu = d.grid_per_cu;
  if (!r.streams) r.streams = d.streams ? d.streams : 3;
  r.streams = std::max(1, std::min(r.streams, otg_context::kAux));
  // the compensated kernel's ring entries are twice as large, and its
  // occupancy with the 512-slot ring (4 wavefronts per CU) leaves LDS
  // latency exposed: its default splits the streamed nets by ring size
  if (!r.min_ring) r.min_ring = d.min_ring ? d.min_ring
                                           : (m == OTG_RCT_BATCHED_COMPENSATED ? 256 : 512);
  if (!ring_ok(r.min_ring)) r.min_ring = 256;
  r.min_ring = std::min(r.min_ring, r.ring_slots);
  const int cap = group_cap(r.small_target);
  if (!r.small_max) r.small_max = d.small_max ? d.small_max : cap;
  r.small_max = std::max(2, std::min(r.small_max, cap));
  if (!r.ring_mask) r.ring_mask = d.ring_mask ? d.ring_mask : (1 << kRingSteps) - 1;
  r.ring_mask &= (1 << kRingSteps) - 1;
  if (!r.ring_mask) r.ring_mask = (1 << kRingSteps) - 1;
  if (!r.first_on_caller) r.first_on_caller = d.first_on_caller;
  r.first_on_caller = r.first_on_caller ? 1 : 0;
  if (!r.small_grid_per_cu) r.small_grid_per_cu = d.small_grid_per_cu;
  r.small_grid_per_cu = std::max(0, r.small_grid_per_cu);
  if (!r.group_alongside) r.group_alongside = d.group_alongside;
  r.group_alongside = std::max(0, std::min(r.group_alongside, 16));
  return r;
}

int otg_rct_blocks_per_cu(otg_context* ctx, const otg_rct_tuning& t, bool comp) {
  (void)ctx;
  return blocks_per_cu(pick_stream(t.ring_slots, comp));
}

otg_status otg_rct_launch_f64(otg_context* ctx, const otg_rct_tree* tree, double* load,
                              double* delay, double* ldelay, double* beta, double* impulse,
                              double* slew, hipStream_t s) {
  if (!load || !delay || !ldelay || !beta || !impulse) {
    OTG_SET_ERR(ctx, "FP64 reference needs load, delay, ldelay, beta and impulse");
    return OTG_ERR_INVALID_ARGUMENT;
  }
  if (tree->num_nets == 0) return OTG_OK;
  dim3 grid((tree->num_nets + 63) / 64), block(64, 4);
  hipLaunchKernelGGL(rct_f64_kernel, grid, block, 0, s, tree->num_nets, tree->nodes_start,
                     tree->pid, tree->pres, tree->cap, tree->slew_in, load, delay, ldelay, beta,
                     impulse, slew);
  OTG_HIP(ctx, hipGetLastError());
  return OTG_OK;
}

otg_status otg_rct_launch_plain(otg_context* ctx, const otg_rct_tree* tree,
                                const otg_rct_result* out, hipStream_t s) {
  if (!out->load || !out->delay || !out->ldelay || !out->beta) {
    OTG_SET_ERR(ctx, "plain method needs load, delay, ldelay and beta (work arrays)");
    return OTG_ERR_INVALID_ARGUMENT;
  }
  if (tree->num_nets == 0) return OTG_OK;
  dim3 grid((tree->num_nets + 63) / 64), block(64, 4);
  hipLaunchKernelGGL(rct_plain_kernel, grid, block, 0, s, tree->num_nets, tree->nodes_start,
                     tree->pid, tree->pres, tree->cap, tree->slew_in, out->load, out->delay,
                     out->ldelay, out->beta, out->impulse, out->slew);
  OTG_HIP(ctx, hipGetLastError());
  return OTG_OK;
}

otg_status otg_rct_launch_batched(otg_context* ctx, otg_rct_plan* plan, const otg_rct_tree* tree,
                                  const otg_rct_result* out, const int32_t* subset,
                                  int32_t count, hipStream_t s) {
  const bool comp = plan->method == OTG_RCT_BATCHED_COMPENSATED;
  if (plan->stats[2] > 0 && !out->load) {
    OTG_SET_ERR(ctx, "load output is required: the tree holds nets larger than a group");
    return OTG_ERR_INVALID_ARGUMENT;
  }
  if (subset && count > plan->num_nets) {
    OTG_SET_ERR(ctx, "subset of %d nets exceeds the plan's %d nets", count, plan->num_nets);
    return OTG_ERR_INVALID_ARGUMENT;
  }
  if (subset && count <= 0) return OTG_OK;
  RctParams P{};
  P.nodes_start = tree->nodes_start;
  P.pid = tree->pid;
  P.pres = tree->pres;
  P.cap = reinterpret_cast<const float4*>(tree->cap);
  P.slew_in = reinterpret_cast<const float4*>(tree->slew_in);
  P.load = reinterpret_cast<float4*>(out->load);
  P.delay = reinterpret_cast<float4*>(out->delay);
  P.ldelay = reinterpret_cast<float4*>(out->ldelay);
  P.beta = reinterpret_cast<float4*>(out->beta);
  P.impulse = reinterpret_cast<float4*>(out->impulse);
  P.slew = reinterpret_cast<float4*>(out->slew);
  P.num_nets_total = tree->num_nets;
  P.net_maxpid = plan->d_net_maxpid;
  P.net_scratch = plan->d_net_scratch;
  P.scratch = plan->d_scratch;
  P.sched = plan->d_sched;
  P.long_maxpid = plan->tuning.ring_slots - kWave;
  P.group_max = plan->small_max;
  const int G = plan->group_cap;
  // ---- packed groups ------------------------------------------------------------
  // Either first and alone on the caller's stream, dealt out to a grid that
  // is resident from the start, or (group_alongside) from a queue on a stream
  // of their own beside the streaming launches, so that the small nets fill
  // the CUs that the largest nets' launches leave idle.
  hipStreamCaptureStatus cap = hipStreamCaptureStatusActive;
  const bool can_fork =
      hipStreamIsCapturing(s, &cap) == hipSuccess && cap == hipStreamCaptureStatusNone;
  const bool along = plan->num_groups > 0 && plan->tuning.group_alongside > 0 && can_fork;
  RctParams QG = P;
  int grid_g = 0;
  if (plan->num_groups > 0) {
    if (subset) {
      hipLaunchKernelGGL(rct_zero_kernel, dim3(1), dim3(64), 0, s, plan->d_sub_count, 1);
      hipLaunchKernelGGL(rct_subset_groups_kernel, dim3((count + 255) / 256), dim3(256), 0, s,
                         count, subset, plan->num_nets, plan->d_net_group, plan->d_group_flag,
                         plan->d_groups, plan->d_sub_groups, plan->d_sub_count);
      OTG_HIP(ctx, hipGetLastError());
      QG.items = plan->d_sub_groups;
      QG.num_items = plan->num_groups;  // bound; the count is read on the device
      QG.num_items_dev = plan->d_sub_count;
      QG.group_flag = plan->d_group_flag;
    } else {
      QG.items = plan->d_groups;
      QG.num_items = plan->num_groups;
    }
    const int most = subset ? std::min(count, plan->num_groups) : plan->num_groups;
    if (along) {
      QG.queue = plan->d_queue + 3 * otg_rct_plan::kMaxBins + 2;
      grid_g = std::min(plan->tuning.group_alongside * std::max(1, ctx->num_cus), most);
    } else {
      grid_g = std::min(plan->grid_group, most);
      void* args[] = {&QG};
      OTG_HIP(ctx, hipLaunchKernel(pick_group(G, comp), dim3(std::max(1, grid_g)), dim3(64),
                                   args, 0, s));
    }
  }
  // ---- streamed nets, one launch per ring size (largest ring first), and the
  // nets whose parent offsets exceed every ring --------------------------------
  int32_t* n_sub = plan->d_queue + 2 * otg_rct_plan::kMaxBins + 2;
  if (plan->num_bins > 0 && subset && plan->stats[2] > 0) {
    SubBins sb{};
    sb.num_bins = plan->num_bins;
    for (int b = 0; b < plan->num_bins; ++b) {
      sb.items[b] = plan->d_bin_sub[b];
      sb.lim[b] = plan->bin_ring[b] - kWave;
      sb.chain[b] = plan->bin_chain[b];
    }
    hipLaunchKernelGGL(rct_zero_kernel, dim3(1), dim3(64), 0, s, n_sub, plan->num_bins);
    hipLaunchKernelGGL(rct_subset_items_kernel, dim3((count + 255) / 256), dim3(256), 0, s,
                       count, subset, plan->num_nets, tree->nodes_start, plan->d_net_maxpid,
                       plan->d_net_scratch, plan->d_net_chain, plan->small_max, sb, n_sub);
    OTG_HIP(ctx, hipGetLastError());
  }
  struct Launch {
    void* fn;
    int grid;
    RctParams Q;
  };
  Launch L[otg_rct_plan::kMaxBins + 1];
  int nl = 0;
  for (int b = plan->num_bins - 1; b >= 0; --b) {
    if (plan->bin_items[b] <= 0) continue;
    RctParams Q = P;
    Q.queue = plan->d_queue + 2 * b;
    int total = plan->bin_items[b];
    Q.items = plan->d_bin_items[b];
    if (subset) {
      Q.items = plan->d_bin_sub[b];
      Q.num_items_dev = n_sub + b;
      total = std::min(count, plan->bin_items[b]);  // bound; the count is read on the device
    }
    Q.num_items = total;
    L[nl++] = Launch{pick_stream(plan->bin_ring[b], comp, plan->bin_chain[b]),
                     std::min(plan->bin_grid[b], total), Q};
  }
  if (plan->stats[3] > 0) {
    RctParams Q = P;
    Q.queue = plan->d_queue + 2 * otg_rct_plan::kMaxBins;
    if (subset) {
      Q.subset = subset;
      Q.subset_count = count;
      Q.items = nullptr;
      Q.num_items = 0;
    } else {
      Q.items = plan->d_gitems;
      Q.num_items = plan->num_gitems;
    }
    const int gt = subset ? count : plan->num_gitems;
    L[nl++] = Launch{pick_global(comp), std::min(plan->grid, std::max(1, gt)), Q};
  }
  // These launches take their nets from queues, so they may share the GPU.
  // Launch i goes to stream n-1-(i mod n) of n: the caller's stream and up
  // to three auxiliary streams forked from it (one stream under graph
  // capture).  The launches then start together: one holds the CUs first,
  // and as its workgroups run out of nets the others take their place, so
  // the last, largest nets of one launch overlap the work of the others
  // instead of leaving CUs idle.  Packed groups that run beside them take
  // the third auxiliary stream.
  hipStream_t st[otg_context::kAux];
  int ns = 1;
  st[0] = s;
  // (at least one: with no streamed nets, packed groups beside them still
  // take their own stream)
  const int want =
      can_fork ? std::max(1, std::min(std::min(plan->tuning.streams, nl), otg_context::kAux)) : 1;
  const int naux = (want - 1) + (along ? 1 : 0);
  hipStream_t gs = s;  // stream of packed groups that run beside the streaming launches
  if (naux > 0) {
    if (!ctx->ev_fork)
      OTG_HIP(ctx, hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
    for (int k = 0; k < otg_context::kAux; ++k) {
      if ((k < want - 1 || (along && k == otg_context::kAux - 1)) && !ctx->aux[k]) {
        OTG_HIP(ctx, hipStreamCreateWithFlags(&ctx->aux[k], hipStreamNonBlocking));
        OTG_HIP(ctx, hipEventCreateWithFlags(&ctx->ev_join[k], hipEventDisableTiming));
      }
    }
    OTG_HIP(ctx, hipEventRecord(ctx->ev_fork, s));
    for (int k = 0; k + 1 < want; ++k) {
      OTG_HIP(ctx, hipStreamWaitEvent(ctx->aux[k], ctx->ev_fork, 0));
      st[ns++] = ctx->aux[k];
    }
    if (along) {
      gs = ctx->aux[otg_context::kAux - 1];
      OTG_HIP(ctx, hipStreamWaitEvent(gs, ctx->ev_fork, 0));
    }
  }
  // (the smallest-ring launch, L[nl-1] unless global-window nets follow,
  // may be held to fewer workgroups per CU so that the larger rings find
  // room from the start)
  if (ns > 1 && plan->tuning.small_grid_per_cu > 0 && plan->num_bins > 0) {
    for (int i = nl - 1; i >= 0; --i) {
      if (L[i].fn == pick_global(comp)) continue;
      L[i].grid = std::min(L[i].grid, plan->tuning.small_grid_per_cu * std::max(1, ctx->num_cus));
      break;
    }
  }
  hipError_t err = hipSuccess;
  if (along) {
    void* args[] = {&QG};
    err = hipLaunchKernel(pick_group(G, comp, true), dim3(std::max(1, grid_g)), dim3(64), args,
                          0, gs);
  }
  for (int i = 0; i < nl && err == hipSuccess; ++i) {
    void* args[] = {&L[i].Q};
    const int si = plan->tuning.first_on_caller ? i % ns : ns - 1 - i % ns;
    err = hipLaunchKernel(L[i].fn, dim3(L[i].grid), dim3(64), args, 0, st[si]);
  }
  // joined on every path, so that the caller's stream orders after all of it
  for (int k = 0; k < otg_context::kAux; ++k) {
    const bool used = (k < ns - 1) || (along && k == otg_context::kAux - 1);
    if (!used) continue;
    const hipError_t e1 = hipEventRecord(ctx->ev_join[k], ctx->aux[k]);
    const hipError_t e2 = hipStreamWaitEvent(s, ctx->ev_join[k], 0);
    if (err == hipSuccess) err = e1 != hipSuccess ? e1 : e2;
  }
  OTG_HIP(ctx, err);
  return OTG_OK;
}

extern "C" otg_status otg_rct_plan_create(otg_context* ctx, const otg_rct_tree* topo,
                                          const otg_rct_tuning* tuning, otg_rct_method method,
                                          otg_rct_plan** out) {
  if (!ctx || !topo || !out || topo->num_nets < 0 || topo->total_num_nodes < 0 ||
      (topo->num_nets > 0 && (!topo->nodes_start || !topo->pid)))
    return OTG_ERR_INVALID_ARGUMENT;
  if (method != OTG_RCT_PLAIN && method != OTG_RCT_BATCHED && method != OTG_RCT_BATCHED_COMPENSATED)
    return OTG_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  OTG_HIP(ctx, hipSetDevice(ctx->device));
  otg_rct_plan* p = new otg_rct_plan();
  p->ctx = ctx;
  p->method = method;
  p->tuning = otg_rct_resolve_tuning(ctx, tuning, method);
  p->num_nets = topo->num_nets;
  p->total_nodes = topo->total_num_nodes;
  const int W = p->tuning.ring_slots;
  const int G = group_cap(p->tuning.small_target);
  p->group_cap = G;
  p->small_max = std::min(p->tuning.small_max, G);
  const bool comp = method == OTG_RCT_BATCHED_COMPENSATED;
  auto fail = [&](otg_status st) {
    otg_rct_plan_destroy(p);
    return st;
  };
#define PLAN_HIP(expr)                                                                        \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) {                                                                   \
      otg_set_error(ctx, "%s failed: %s", #expr, hipGetErrorString(e_));                      \
      return fail(e_ == hipErrorOutOfMemory ? OTG_ERR_OUT_OF_MEMORY : OTG_ERR_DEVICE);        \
    }                                                                                         \
  } while (0)
  // ---- bins: per ring size a bin of ordinary nets and (compensated) one
  // of chain nets, which take the kernel variant with the chain-window scans
  int nb = 0;
  for (int k = 0; k < kRingSteps; ++k) {
    const int ring = kRingLadder[k];
    if (ring < p->tuning.min_ring || ring > W || !((p->tuning.ring_mask >> k) & 1)) continue;
    for (int ch = 0; ch <= (comp ? 1 : 0); ++ch) {
      p->bin_ring[nb] = ring;
      p->bin_chain[nb] = ch;
      ++nb;
    }
  }
  p->num_bins = nb;
  // ---- work list, schedules and bins, built on the device ----
  {
    const otg_status st = otg_rct_plan_build_device(ctx, p, topo, ctx->stream);
    if (st != OTG_OK) return fail(st);
  }
  PLAN_HIP(hipMalloc(&p->d_queue, sizeof(int32_t) * otg_rct_plan::kQueueInts));
  PLAN_HIP(hipMemsetAsync(p->d_queue, 0, sizeof(int32_t) * otg_rct_plan::kQueueInts, ctx->stream));
  PLAN_HIP(hipStreamSynchronize(ctx->stream));
  const int ncu = std::max(1, ctx->num_cus);
  const int g1 = p->tuning.grid_per_cu;
  p->grid = std::max(1, (g1 > 0 ? g1 : blocks_per_cu(pick_stream(W, comp))) * ncu);
  for (int b = 0; b < nb; ++b)
    p->bin_grid[b] =
        std::max(1, (g1 > 0 ? g1 : blocks_per_cu(pick_stream(p->bin_ring[b], comp,
                                                                 p->bin_chain[b]))) * ncu);
  p->grid_group = std::max(1, (g1 > 0 ? g1 : blocks_per_cu(pick_group(G, comp))) * ncu);
#undef PLAN_HIP
  *out = p;
  return OTG_OK;
}

extern "C" otg_status otg_rct_plan_destroy(otg_rct_plan* p) {
  if (!p) return OTG_OK;
  // (hipFree of a null pointer is a no-op; failures here cannot be acted on)
  for (int b = 0; b < otg_rct_plan::kMaxBins; ++b) {
    (void)hipFree(p->d_bin_items[b]);
    (void)hipFree(p->d_bin_sub[b]);
  }
  for (void* q : {(void*)p->d_gitems, (void*)p->d_groups, (void*)p->d_sub_groups,
                  (void*)p->d_sched, (void*)p->d_group_flag, (void*)p->d_sub_count,
                  (void*)p->d_net_group, (void*)p->d_queue, (void*)p->d_scratch,
                  (void*)p->d_net_maxpid, (void*)p->d_net_chain, (void*)p->d_net_scratch})
    (void)hipFree(q);
  delete p;
  return OTG_OK;
}

extern "C" otg_status otg_rct_plan_stats(const otg_rct_plan* p, int64_t stats[8]) {
  if (!p || !stats) return OTG_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < 8; ++i) stats[i] = p->stats[i];
  return OTG_OK;
}
