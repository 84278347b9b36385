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
#include <vector>

#include "otg_internal.hpp"

namespace {

constexpr int kWave = 64;

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
// Double-float helpers (compensated mode).
// ---------------------------------------------------------------------------
struct ff {
  float hi, lo;
};
__device__ __forceinline__ ff two_sum(float a, float b) {
  const float s = a + b;
  const float bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}
__device__ __forceinline__ float two_sum_err(float a, float b, float s) {
  const float bb = s - a;
  return (a - (s - bb)) + (b - bb);
}
__device__ __forceinline__ ff quick_two_sum(float a, float b) {
  const float s = a + b;
  return {s, b - (s - a)};
}
__device__ __forceinline__ ff ff_add(ff a, ff b) {
  ff s = two_sum(a.hi, b.hi);
  return quick_two_sum(s.hi, s.lo + (a.lo + b.lo));
}
__device__ __forceinline__ ff ff_add_f(ff a, float b) {
  ff s = two_sum(a.hi, b);
  return quick_two_sum(s.hi, s.lo + a.lo);
}
__device__ __forceinline__ ff ff_mul_f(ff a, float b) {
  const float p = a.hi * b;
  const float e = fmaf(a.hi, b, -p);
  return quick_two_sum(p, fmaf(a.lo, b, e));
}
__device__ __forceinline__ ff ff_mul(ff a, ff b) {
  const float p = a.hi * b.hi;
  const float e = fmaf(a.hi, b.hi, -p);
  return quick_two_sum(p, e + (a.hi * b.lo + a.lo * b.hi));
}

__device__ __forceinline__ void unpack(float4 v, float (&a)[4]) {
  a[0] = v.x;
  a[1] = v.y;
  a[2] = v.z;
  a[3] = v.w;
}
__device__ __forceinline__ float4 pack(const float (&a)[4]) { return make_float4(a[0], a[1], a[2], a[3]); }

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
  int32_t subset_count;
  const int32_t* subset;
  const int32_t* net_maxpid;
  const int64_t* net_scratch;
  float* scratch;
  int32_t* queue;
};

template <int W, bool COMP>
struct RctLds {
  float4 ra[W];                // bottom-up: load acc (hi) | top-down: delay (hi)
  float4 rb[W];                // bottom-up: L acc (hi)    | top-down: impulse (hi)
  float4 ral[COMP ? W : 1];    // low parts (compensated mode)
  float4 rbl[COMP ? W : 1];
  int32_t spid[W];             // parent offsets of a packed item, kept for the top-down sweep
  float sres[W];               // resistances of a packed item
  uint32_t rootbits[W / 32];   // driver positions of a packed group
  int32_t cnt[kWave];          // per-lane child counters / done flags of a window
};

// Window storage: either the LDS ring (index mod W) or a per-node region in
// global memory (no wrap).  lo arrays are used only in compensated mode.
template <bool GLOBAL, bool COMP>
struct Ring {
  float4* a;
  float4* b;
  float4* al;
  float4* bl;
  int32_t base;
  int32_t mask;  // W-1 for the LDS ring, -1 (no wrap) for global
  __device__ __forceinline__ int slot(int i) const { return (i - base) & mask; }

  __device__ __forceinline__ float4 lda(const float4* p, int k) const {
    if constexpr (GLOBAL) {
      const float* q = reinterpret_cast<const float*>(p + k);
      return make_float4(__hip_atomic_load(q + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                         __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                         __hip_atomic_load(q + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                         __hip_atomic_load(q + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    } else {
      return p[k];
    }
  }
  __device__ __forceinline__ void get(int k, float (&x)[4], float (&y)[4], float (&xl)[4],
                                      float (&yl)[4]) const {
    unpack(lda(a, k), x);
    unpack(lda(b, k), y);
    if constexpr (COMP) {
      unpack(lda(al, k), xl);
      unpack(lda(bl, k), yl);
    }
  }
  __device__ __forceinline__ void put(int k, const float (&x)[4], const float (&y)[4],
                                      const float (&xl)[4], const float (&yl)[4]) const {
    a[k] = pack(x);
    b[k] = pack(y);
    if constexpr (COMP) {
      al[k] = pack(xl);
      bl[k] = pack(yl);
    }
  }
  __device__ __forceinline__ void zero(int k) const {
    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    a[k] = z;
    b[k] = z;
    if constexpr (COMP) {
      al[k] = z;
      bl[k] = z;
    }
  }
  __device__ __forceinline__ static float atomic_add_rtn(float* p, float v) {
    if constexpr (GLOBAL) return unsafeAtomicAdd(p, v);
    else return atomicAdd(p, v);
  }
  __device__ __forceinline__ static void atomic_add(float* p, float v) {
    if constexpr (GLOBAL) unsafeAtomicAdd(p, v);
    else atomicAdd(p, v);
  }
  // acc[k] += (x, xl), acc'[k] += (y, yl); compensated: the rounding error of
  // the high-part add is recovered from the returned old value (two-sum) and
  // added to the low part.
  __device__ __forceinline__ void add(int k, const float (&x)[4], const float (&y)[4],
                                      const float (&xl)[4], const float (&yl)[4]) const {
    float* pa = reinterpret_cast<float*>(a + k);
    float* pb = reinterpret_cast<float*>(b + k);
    if constexpr (COMP) {
      float* pal = reinterpret_cast<float*>(al + k);
      float* pbl = reinterpret_cast<float*>(bl + k);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const float oa = atomic_add_rtn(pa + c, x[c]);
        atomic_add(pal + c, two_sum_err(oa, x[c], oa + x[c]) + xl[c]);
        const float ob = atomic_add_rtn(pb + c, y[c]);
        atomic_add(pbl + c, two_sum_err(ob, y[c], ob + y[c]) + yl[c]);
      }
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        atomic_add(pa + c, x[c]);
        atomic_add(pb + c, y[c]);
      }
    }
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

struct WinIn {
  int32_t pid;
  float r;
  float4 c;
};

template <int W, bool COMP, bool GLOBAL>
__device__ __forceinline__ void process_item(const RctParams& P, RctLds<W, COMP>& S,
                                             const Ring<GLOBAL, COMP>& R, const RctItem& it,
                                             const int lane) {
  const int pb = it.pos_begin, pe = it.pos_end, n = pe - pb;
  const bool keep = it.kind != ITEM_STREAM;             // window holds the whole item
  const bool stash = it.kind == ITEM_PACKED;            // pid/R kept in LDS
  const bool multi = it.net_end - it.net_begin > 1;     // packed group of several nets
  float4* gL = reinterpret_cast<float4*>(P.scratch + (it.kind == ITEM_STREAM ? it.scratch : 0));

  // ---- set-up: clear the window, mark drivers -------------------------------
  {
    const int nz = keep ? n : W;
    for (int k = lane; k < nz; k += kWave) R.zero(k);
    if (multi) {
      for (int k = lane; k < W / 32; k += kWave) S.rootbits[k] = 0u;
      wave_sync();
      for (int j = it.net_begin + lane; j < it.net_end; j += kWave) {
        const int r = P.nodes_start[j] - pb;
        atomicOr(&S.rootbits[r >> 5], 1u << (r & 31));
      }
    }
    R.fence();
    wave_sync();
  }
  auto is_root = [&](int i) -> bool {
    if (multi) return (S.rootbits[(i - pb) >> 5] >> ((i - pb) & 31)) & 1u;
    return i == pb;
  };

  const int nwin = (n + kWave - 1) / kWave;

  // ---- bottom-up sweep --------------------------------------------------------
  auto load_bu = [&](int w, WinIn& in) {
    const int i = pb + w * kWave + lane;
    if (i < pe) {
      in.pid = P.pid[i];
      in.r = P.pres[i];
      in.c = P.cap[i];
    } else {
      in.pid = 0;
      in.r = 0.f;
      in.c = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  };
  WinIn cur, nxt;
  load_bu(nwin - 1, cur);
  for (int w = nwin - 1; w >= 0; --w) {
    if (w > 0) load_bu(w - 1, nxt);
    const int wk = pb + w * kWave;
    const int i = wk + lane;
    const bool valid = i < pe;
    const bool root = valid && is_root(i);
    const int p = (valid && !root) ? i - cur.pid : -1;
    const bool pin = p >= wk;
    if (stash && valid) {
      S.spid[i - pb] = cur.pid;
      S.sres[i - pb] = cur.r;
    }
    S.cnt[lane] = 0;
    wave_sync();
    if (pin) atomicAdd(&S.cnt[p - wk], 1);
    wave_sync();
    bool done = !valid;
    float cap4[4];
    unpack(cur.c, cap4);
    for (int round = 0; round <= kWave; ++round) {
      const bool ready = !done && S.cnt[lane] == 0;
      if (ready) {
        const int s = R.slot(i);
        float xa[4], xb[4], xal[4], xbl[4];
        R.get(s, xa, xb, xal, xbl);
        float ld[4], ldl[4], L[4], Ll[4], w[4], wl[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          if constexpr (COMP) {
            const ff t = ff_add(two_sum(cap4[c], xa[c]), ff{xal[c], 0.f});
            ld[c] = t.hi;
            ldl[c] = t.lo;
            L[c] = xb[c];
            Ll[c] = xbl[c];
            const ff q = ff_add(ff{L[c], Ll[c]}, ff_mul_f(ff_mul(t, t), cur.r));
            w[c] = q.hi;
            wl[c] = q.lo;
          } else {
            ld[c] = cap4[c] + xa[c];
            ldl[c] = 0.f;
            L[c] = xb[c];
            Ll[c] = 0.f;
            w[c] = L[c] + cur.r * ld[c] * ld[c];
            wl[c] = 0.f;
          }
        }
        if (keep) {
          R.put(s, ld, L, ldl, Ll);
        } else {
          float o1[4], o2[4];
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            o1[c] = COMP ? ld[c] + ldl[c] : ld[c];
            o2[c] = COMP ? L[c] + Ll[c] : L[c];
          }
          P.load[i] = pack(o1);
          gL[i - pb] = pack(o2);
        }
        if (p >= 0) {
          R.add(R.slot(p), ld, w, ldl, wl);
          if (pin) atomicSub(&S.cnt[p - wk], 1);
        }
        done = true;
      }
      R.fence();
      wave_sync();
      if (__all(done)) break;
    }
    if (!keep && valid) R.zero(R.slot(i));
    wave_sync();
    cur = nxt;
  }
  R.fence();
  wave_sync();

  // ---- top-down sweep ---------------------------------------------------------
  struct TdIn {
    int32_t pid;
    float r;
    float4 ld, L;
  };
  auto load_td = [&](int w, TdIn& in) {
    const int i = pb + w * kWave + lane;
    in.ld = make_float4(0.f, 0.f, 0.f, 0.f);
    in.L = in.ld;
    if (i >= pe) {
      in.pid = 0;
      in.r = 0.f;
      return;
    }
    if (stash) {
      in.pid = S.spid[i - pb];
      in.r = S.sres[i - pb];
    } else {
      in.pid = P.pid[i];
      in.r = P.pres[i];
    }
    if (!keep) {
      in.ld = P.load[i];
      in.L = gL[i - pb];
    }
  };
  TdIn tc, tn;
  load_td(0, tc);
  int carry = 0;
  for (int w = 0; w < nwin; ++w) {
    if (w + 1 < nwin) load_td(w + 1, tn);
    const int wk = pb + w * kWave;
    const int i = wk + lane;
    const bool valid = i < pe;
    const bool root = valid && is_root(i);
    const int p = (valid && !root) ? i - tc.pid : -1;
    const bool pin = p >= wk;
    int net = it.net_begin;
    if (multi) {
      const uint64_t rm = __ballot(root);
      net += carry + __popcll(rm & ((2ull << lane) - 1ull)) - 1;
      carry += __popcll(rm);
    }
    // own load and L
    float ld[4], ldl[4], L[4], Ll[4];
    if (keep) {
      if (valid) {
        R.get(R.slot(i), ld, L, ldl, Ll);
      } else {
#pragma unroll
        for (int c = 0; c < 4; ++c) ld[c] = ldl[c] = L[c] = Ll[c] = 0.f;
      }
    } else {
      unpack(tc.ld, ld);
      unpack(tc.L, L);
#pragma unroll
      for (int c = 0; c < 4; ++c) ldl[c] = Ll[c] = 0.f;
    }
    const float r = root ? 0.f : tc.r;
    // per-node path terms u = R load, z = u^2 + 2 R L
    float u[4], ul[4], z[4], zl[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if constexpr (COMP) {
        const ff uu = ff_mul_f(ff{ld[c], ldl[c]}, r);
        const ff zz = ff_add(ff_mul(uu, uu), ff_mul_f(ff{L[c], Ll[c]}, 2.0f * r));
        u[c] = uu.hi;
        ul[c] = uu.lo;
        z[c] = zz.hi;
        zl[c] = zz.lo;
      } else {
        u[c] = r * ld[c];
        z[c] = u[c] * u[c] + 2.0f * r * L[c];
        ul[c] = zl[c] = 0.f;
      }
    }
    S.cnt[lane] = 0;
    wave_sync();
    bool done = !valid;
    float d[4], dl[4], m[4], ml[4];
    for (int round = 0; round <= kWave; ++round) {
      const bool ready = !done && (!pin || S.cnt[p - wk] != 0);
      if (ready) {
        float pd[4] = {0.f, 0.f, 0.f, 0.f}, pm[4] = {0.f, 0.f, 0.f, 0.f};
        float pdl[4] = {0.f, 0.f, 0.f, 0.f}, pml[4] = {0.f, 0.f, 0.f, 0.f};
        if (p >= 0) R.get(R.slot(p), pd, pm, pdl, pml);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          if constexpr (COMP) {
            const ff dd = ff_add(ff{pd[c], pdl[c]}, ff{u[c], ul[c]});
            const ff mm = ff_add(ff{pm[c], pml[c]}, ff{z[c], zl[c]});
            d[c] = dd.hi;
            dl[c] = dd.lo;
            m[c] = mm.hi;
            ml[c] = mm.lo;
          } else {
            d[c] = pd[c] + u[c];
            m[c] = pm[c] + z[c];
            dl[c] = ml[c] = 0.f;
          }
        }
        R.put(R.slot(i), d, m, dl, ml);
        S.cnt[lane] = 1;
        done = true;
      }
      R.fence();
      wave_sync();
      if (__all(done)) break;
    }
    if (valid) {
      float4 si4 = P.slew_in ? P.slew_in[net] : make_float4(0.f, 0.f, 0.f, 0.f);
      float si[4];
      unpack(si4, si);
      float o_ld[4], o_d[4], o_ldel[4], o_b[4], o_m[4], o_s[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if constexpr (COMP) {
          // final per-node combination in FP64
          const double dd = (double)d[c] + (double)dl[c];
          const double md = (double)m[c] + (double)ml[c];
          const double ldd = (double)ld[c] + (double)ldl[c];
          const double Ld = (double)L[c] + (double)Ll[c];
          o_ld[c] = (float)ldd;
          o_d[c] = (float)dd;
          o_m[c] = (float)md;
          o_ldel[c] = (float)(dd * ldd + Ld);
          o_b[c] = (float)(0.5 * (dd * dd + md));
          const double sv = sqrt((double)si[c] * (double)si[c] + md);
          o_s[c] = (float)(si[c] < 0.f ? -sv : sv);
        } else {
          o_ld[c] = ld[c];
          o_d[c] = d[c];
          o_m[c] = m[c];
          o_ldel[c] = d[c] * ld[c] + L[c];
          o_b[c] = 0.5f * (d[c] * d[c] + m[c]);
          const float sv = sqrtf(si[c] * si[c] + m[c]);
          o_s[c] = si[c] < 0.f ? -sv : sv;
        }
      }
      if (keep && P.load) P.load[i] = pack(o_ld);
      if (P.delay) P.delay[i] = pack(o_d);
      if (P.ldelay) P.ldelay[i] = pack(o_ldel);
      if (P.beta) P.beta[i] = pack(o_b);
      if (P.impulse) P.impulse[i] = pack(o_m);
      if (P.slew) P.slew[i] = pack(o_s);
    }
    tc = tn;
  }
  R.fence();
  wave_sync();
}

template <int W, bool COMP>
__global__ __launch_bounds__(64) void rct_batched_kernel(RctParams P) {
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
      const int net = P.subset[idx];
      it.pos_begin = P.nodes_start[net];
      it.pos_end = P.nodes_start[net + 1];
      it.net_begin = net;
      it.net_end = net + 1;
      const int nn = it.pos_end - it.pos_begin;
      it.kind = nn <= W ? ITEM_PACKED : (P.net_maxpid[net] <= W - kWave ? ITEM_STREAM : ITEM_GLOBAL);
      it.pad = 0;
      it.scratch = P.net_scratch[net];
    } else {
      it = P.items[idx];
    }
    if (it.pos_end <= it.pos_begin) continue;
    if (it.kind == ITEM_GLOBAL) {
      const int64_t nn = it.pos_end - it.pos_begin;
      float4* base = reinterpret_cast<float4*>(P.scratch + it.scratch);
      Ring<true, COMP> R{base, base + nn, COMP ? base + 2 * nn : base, COMP ? base + 3 * nn : base,
                         it.pos_begin, -1};
      process_item<W, COMP, true>(P, S, R, it, lane);
    } else {
      Ring<false, COMP> R{S.ra, S.rb, S.ral, S.rbl, it.pos_begin, W - 1};
      process_item<W, COMP, false>(P, S, R, it, lane);
    }
  }
  // The last workgroup out resets the queue for the next launch.
  __syncthreads();
  if (lane == 0) {
    __threadfence();
    const int old = atomicAdd(&P.queue[1], 1);
    if (old == (int)gridDim.x - 1) {
      atomicExch(&P.queue[0], 0);
      atomicExch(&P.queue[1], 0);
    }
  }
}

// Per-net validation and largest parent offset (plan building).
__global__ void rct_netinfo_kernel(int32_t num_nets, const int32_t* __restrict__ nodes_start,
                                   const int32_t* __restrict__ pid, int32_t* __restrict__ maxpid,
                                   int32_t* __restrict__ bad) {
  const int net = blockIdx.x * blockDim.x + threadIdx.x;
  if (net >= num_nets) return;
  const int32_t st = nodes_start[net], ed = nodes_start[net + 1];
  int32_t m = 0;
  bool ok = ed >= st;
  for (int32_t i = st + 1; i < ed; ++i) {
    const int32_t q = pid[i];
    if (q < 1 || q > i - st) ok = false;
    m = q > m ? q : m;
  }
  maxpid[net] = m;
  if (!ok) atomicMin(bad, net);
}

template <int W, bool COMP>
void* batched_kernel_ptr() {
  return reinterpret_cast<void*>(&rct_batched_kernel<W, COMP>);
}

void* pick_batched(int W, bool comp) {
  if (comp) {
    if (W == 256) return batched_kernel_ptr<256, true>();
    if (W == 1024) return batched_kernel_ptr<1024, true>();
    return batched_kernel_ptr<512, true>();
  }
  if (W == 256) return batched_kernel_ptr<256, false>();
  if (W == 1024) return batched_kernel_ptr<1024, false>();
  return batched_kernel_ptr<512, false>();
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
  if (!r.ring_slots) r.ring_slots = d.ring_slots ? d.ring_slots : 512;
  if (r.ring_slots != 256 && r.ring_slots != 512 && r.ring_slots != 1024) r.ring_slots = 512;
  if (!r.small_target) r.small_target = d.small_target ? d.small_target : r.ring_slots;
  r.small_target = std::max(64, std::min(r.small_target, r.ring_slots));
  if (!r.grid_per_cu) r.grid_per_cu = d.grid_per_cu;
  r.waves_per_wg = 1;
  (void)m;
  return r;
}

int otg_rct_blocks_per_cu(otg_context* ctx, const otg_rct_tuning& t, bool comp) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pick_batched(t.ring_slots, comp), 64, 0) !=
          hipSuccess ||
      nb <= 0)
    nb = 1;
  (void)ctx;
  return nb;
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
    OTG_SET_ERR(ctx, "load output is required: the tree holds nets larger than the window");
    return OTG_ERR_INVALID_ARGUMENT;
  }
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
  P.items = plan->d_items;
  P.num_items = plan->num_items;
  P.subset = subset;
  P.subset_count = count;
  P.net_maxpid = plan->d_net_maxpid;
  P.net_scratch = plan->d_net_scratch;
  P.scratch = plan->d_scratch;
  P.queue = plan->d_queue;
  const int total = subset ? count : plan->num_items;
  if (total <= 0) return OTG_OK;
  int grid = plan->grid;
  if (grid > total) grid = total;
  void* fn = pick_batched(plan->tuning.ring_slots, comp);
  void* args[] = {&P};
  OTG_HIP(ctx, hipLaunchKernel(fn, dim3(grid), dim3(64), args, 0, s));
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
  const int N = topo->num_nets;
  const int W = p->tuning.ring_slots;
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
  std::vector<int32_t> hstart(N + 1, 0), hmax(N, 0);
  if (N > 0) {
    PLAN_HIP(hipMalloc(&p->d_net_maxpid, sizeof(int32_t) * N));
    PLAN_HIP(hipMalloc(&p->d_net_scratch, sizeof(int64_t) * N));
    int32_t* d_bad = nullptr;
    PLAN_HIP(hipMalloc(&d_bad, sizeof(int32_t)));
    const int32_t big = INT32_MAX;
    PLAN_HIP(hipMemcpyAsync(d_bad, &big, sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(rct_netinfo_kernel, dim3((N + 255) / 256), dim3(256), 0, ctx->stream, N,
                       topo->nodes_start, topo->pid, p->d_net_maxpid, d_bad);
    PLAN_HIP(hipGetLastError());
    int32_t bad = 0;
    PLAN_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    PLAN_HIP(hipMemcpyAsync(hstart.data(), topo->nodes_start, sizeof(int32_t) * (N + 1),
                            hipMemcpyDeviceToHost, ctx->stream));
    PLAN_HIP(hipMemcpyAsync(hmax.data(), p->d_net_maxpid, sizeof(int32_t) * N,
                            hipMemcpyDeviceToHost, ctx->stream));
    PLAN_HIP(hipStreamSynchronize(ctx->stream));
    hipFree(d_bad);
    if (bad != INT32_MAX) {
      otg_set_error(ctx, "net %d: pid must satisfy 1 <= pid[i] <= i - nodes_start[net]", bad);
      return fail(OTG_ERR_BAD_TOPOLOGY);
    }
    if (hstart[0] != 0 || hstart[N] != topo->total_num_nodes) {
      otg_set_error(ctx, "nodes_start must run from 0 to total_num_nodes");
      return fail(OTG_ERR_INVALID_ARGUMENT);
    }
  }
  // ---- binning --------------------------------------------------------------
  std::vector<RctItem> big_items, packed;
  std::vector<int64_t> net_scratch(N, -1);
  int64_t scratch = 0;
  const int64_t slot_floats = comp ? 16 : 8;
  int64_t n_packed_nets = 0, n_stream = 0, n_global = 0, maxn = 0;
  RctItem cur{};
  bool open = false;
  auto flush = [&]() {
    if (open) packed.push_back(cur);
    open = false;
  };
  for (int k = 0; k < N; ++k) {
    const int32_t n = hstart[k + 1] - hstart[k];
    maxn = std::max<int64_t>(maxn, n);
    if (n <= 0) {
      flush();
      continue;
    }
    if (n <= W) {
      ++n_packed_nets;
      if (open && (cur.pos_end - cur.pos_begin) + n <= p->tuning.small_target) {
        cur.pos_end += n;
        cur.net_end = k + 1;
      } else {
        flush();
        cur = RctItem{hstart[k], hstart[k + 1], k, k + 1, ITEM_PACKED, 0, 0};
        open = true;
      }
      continue;
    }
    flush();
    RctItem it{hstart[k], hstart[k + 1], k, k + 1, ITEM_STREAM, 0, 0};
    if (hmax[k] <= W - kWave) {
      it.kind = ITEM_STREAM;
      it.scratch = scratch;
      scratch += 4LL * n;
      ++n_stream;
    } else {
      it.kind = ITEM_GLOBAL;
      it.scratch = scratch;
      scratch += slot_floats * n;
      ++n_global;
    }
    net_scratch[k] = it.scratch;
    big_items.push_back(it);
  }
  flush();
  std::stable_sort(big_items.begin(), big_items.end(), [](const RctItem& a, const RctItem& b) {
    return (a.pos_end - a.pos_begin) > (b.pos_end - b.pos_begin);
  });
  std::vector<RctItem> items;
  items.reserve(big_items.size() + packed.size());
  items.insert(items.end(), big_items.begin(), big_items.end());
  items.insert(items.end(), packed.begin(), packed.end());
  p->num_items = (int32_t)items.size();
  p->stats[0] = (int64_t)packed.size();
  p->stats[1] = n_packed_nets;
  p->stats[2] = n_stream;
  p->stats[3] = n_global;
  p->stats[4] = maxn;
  p->stats[5] = p->num_items;
  p->stats[6] = scratch;
  p->stats[7] = W;
  if (!items.empty()) {
    PLAN_HIP(hipMalloc(&p->d_items, sizeof(RctItem) * items.size()));
    PLAN_HIP(hipMemcpyAsync(p->d_items, items.data(), sizeof(RctItem) * items.size(),
                            hipMemcpyHostToDevice, ctx->stream));
  }
  if (N > 0)
    PLAN_HIP(hipMemcpyAsync(p->d_net_scratch, net_scratch.data(), sizeof(int64_t) * N,
                            hipMemcpyHostToDevice, ctx->stream));
  p->scratch_floats = scratch;
  if (scratch > 0) PLAN_HIP(hipMalloc(&p->d_scratch, sizeof(float) * scratch));
  PLAN_HIP(hipMalloc(&p->d_queue, sizeof(int32_t) * 4));
  PLAN_HIP(hipMemsetAsync(p->d_queue, 0, sizeof(int32_t) * 4, ctx->stream));
  PLAN_HIP(hipStreamSynchronize(ctx->stream));
  const int per_cu =
      p->tuning.grid_per_cu > 0 ? p->tuning.grid_per_cu : otg_rct_blocks_per_cu(ctx, p->tuning, comp);
  p->grid = std::max(1, per_cu * std::max(1, ctx->num_cus));
#undef PLAN_HIP
  *out = p;
  return OTG_OK;
}

extern "C" otg_status otg_rct_plan_destroy(otg_rct_plan* p) {
  if (!p) return OTG_OK;
  if (p->d_items) hipFree(p->d_items);
  if (p->d_queue) hipFree(p->d_queue);
  if (p->d_scratch) hipFree(p->d_scratch);
  if (p->d_net_maxpid) hipFree(p->d_net_maxpid);
  if (p->d_net_scratch) hipFree(p->d_net_scratch);
  delete p;
  return OTG_OK;
}

extern "C" otg_status otg_rct_plan_stats(const otg_rct_plan* p, int64_t stats[8]) {
  if (!p || !stats) return OTG_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < 8; ++i) stats[i] = p->stats[i];
  return OTG_OK;
}
