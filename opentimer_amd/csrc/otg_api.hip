// C ABI of libotg: context management, P1 / P2 entry points and the
// host-pointer drop-ins for the reference's GPU functions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "otg_internal.hpp"

otg_status otg_levelize_launch_raw(otg_context* ctx, int32_t n, int32_t first_size,
                                   const int32_t* fanin_start, const int32_t* fanin,
                                   int32_t* out_copy, int32_t* frontiers, int32_t* ends,
                                   int32_t* level, int32_t* scratch, int32_t* num_ends_host,
                                   hipStream_t s, bool pipelined);

namespace {
thread_local std::string g_last_error;  // errors raised before a context exists
}

void otg_set_error(const otg_context* ctx, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (ctx) const_cast<otg_context*>(ctx)->last_error = buf;
  else g_last_error = buf;
}

extern "C" int otg_version(void) { return 1; }

extern "C" const char* otg_status_string(otg_status s) {
  switch (s) {
    case OTG_OK: return "ok";
    case OTG_ERR_INVALID_ARGUMENT: return "invalid argument";
    case OTG_ERR_NOT_A_TREE: return "net is not a spanning tree";
    case OTG_ERR_BAD_TOPOLOGY: return "bad parent offsets";
    case OTG_ERR_DEVICE: return "HIP runtime error";
    case OTG_ERR_OUT_OF_MEMORY: return "out of device memory";
    case OTG_ERR_NO_DEVICE: return "no HIP device";
    case OTG_ERR_CYCLE: return "cycle in the timing graph";
    case OTG_ERR_UNSUPPORTED: return "unsupported";
  }
  return "unknown status";
}

extern "C" const char* otg_last_error(const otg_context* ctx) {
  return ctx ? ctx->last_error.c_str() : g_last_error.c_str();
}

extern "C" otg_status otg_create(int device, otg_context** out) {
  if (!out) return OTG_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
    otg_set_error(nullptr, "no HIP device visible");
    return OTG_ERR_NO_DEVICE;
  }
  if (device < 0 || device >= count) return OTG_ERR_INVALID_ARGUMENT;
  otg_context* ctx = new otg_context();
  ctx->device = device;
  auto fail = [&](const char* what, hipError_t e) {
    otg_set_error(nullptr, "%s: %s", what, hipGetErrorString(e));
    delete ctx;
    return OTG_ERR_DEVICE;
  };
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return fail("hipSetDevice", e);
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) return fail("hipGetDeviceProperties", e);
  ctx->num_cus = prop.multiProcessorCount;
  ctx->lds_per_cu = (int)prop.maxSharedMemoryPerMultiProcessor;
  // Work the library issues for itself (plan building, host-pointer entry
  // points) goes to the device's default stream, so it is ordered with the
  // caller's default-stream copies and kernels.
  ctx->stream = nullptr;
  *out = ctx;
  return OTG_OK;
}

extern "C" otg_status otg_destroy(otg_context* ctx) {
  if (!ctx) return OTG_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (int k = 0; k < otg_context::kAux; ++k) {
    if (ctx->aux[k]) {
      (void)hipStreamSynchronize(ctx->aux[k]);
      (void)hipStreamDestroy(ctx->aux[k]);
    }
    if (ctx->ev_join[k]) (void)hipEventDestroy(ctx->ev_join[k]);
  }
  if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
  for (hipEvent_t e : ctx->lv_ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->lv_flag) (void)hipHostFree(ctx->lv_flag);
  for (int t = 0; t < otg_context::kStage; ++t) {
    if (ctx->stage_stream[t]) {
      (void)hipStreamSynchronize(ctx->stage_stream[t]);
      (void)hipStreamDestroy(ctx->stage_stream[t]);
    }
    for (void* b : ctx->stage_buf[t])
      if (b) (void)hipHostFree(b);
  }
  delete ctx;
  return OTG_OK;
}

extern "C" otg_status otg_synchronize(otg_context* ctx, otg_stream stream) {
  if (!ctx) return OTG_ERR_INVALID_ARGUMENT;
  OTG_HIP(ctx, hipStreamSynchronize(otg_pick_stream(ctx, stream)));
  return OTG_OK;
}

// ---------------------------------------------------------------------------
// P1
// ---------------------------------------------------------------------------
static otg_status check_tree(otg_context* ctx, const otg_rct_tree* t) {
  if (!t || t->num_nets < 0 || t->total_num_nodes < 0) return OTG_ERR_INVALID_ARGUMENT;
  if (t->num_nets > 0 && (!t->nodes_start || !t->pid || !t->pres || !t->cap)) {
    OTG_SET_ERR(ctx, "tree arrays must be non-null");
    return OTG_ERR_INVALID_ARGUMENT;
  }
  return OTG_OK;
}

// The batched kernels move the per-corner arrays as float4 (16-byte vector
// loads and stores).
static bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

static otg_status check_vec_alignment(otg_context* ctx, const otg_rct_tree* t,
                                      const otg_rct_result* o) {
  const void* arrays[] = {t->cap, t->slew_in, o->load, o->delay, o->ldelay, o->beta,
                          o->impulse, o->slew};
  for (const void* p : arrays) {
    if (p && !aligned16(p)) {
      OTG_SET_ERR(ctx, "cap, slew_in and the result arrays must be 16-byte aligned "
                       "(per-node corner quadruples are moved as float4)");
      return OTG_ERR_INVALID_ARGUMENT;
    }
  }
  return OTG_OK;
}

extern "C" otg_status otg_rct_reference_f64(otg_context* ctx, const otg_rct_tree* tree,
                                            double* load, double* delay, double* ldelay,
                                            double* beta, double* impulse, double* slew,
                                            otg_stream stream) {
  if (!ctx) return OTG_ERR_INVALID_ARGUMENT;
  otg_status st = check_tree(ctx, tree);
  if (st != OTG_OK) return st;
  OTG_HIP(ctx, hipSetDevice(ctx->device));
  return otg_rct_launch_f64(ctx, tree, load, delay, ldelay, beta, impulse, slew,
                            otg_pick_stream(ctx, stream));
}

extern "C" otg_status otg_rct_solve(otg_context* ctx, otg_rct_plan* plan, const otg_rct_tree* tree,
                                    const otg_rct_result* out, otg_rct_method method,
                                    otg_stream stream) {
  if (!ctx || !out) return OTG_ERR_INVALID_ARGUMENT;
  otg_status st = check_tree(ctx, tree);
  if (st != OTG_OK) return st;
  hipStream_t s = otg_pick_stream(ctx, stream);
  OTG_HIP(ctx, hipSetDevice(ctx->device));
  if (method == OTG_RCT_PLAIN) return otg_rct_launch_plain(ctx, tree, out, s);
  if (method != OTG_RCT_BATCHED && method != OTG_RCT_BATCHED_COMPENSATED)
    return OTG_ERR_INVALID_ARGUMENT;
  st = check_vec_alignment(ctx, tree, out);
  if (st != OTG_OK) return st;
  if (plan) {
    if (plan->num_nets != tree->num_nets || plan->total_nodes != tree->total_num_nodes) {
      OTG_SET_ERR(ctx, "plan was built for a different topology");
      return OTG_ERR_INVALID_ARGUMENT;
    }
    if (plan->method != method) {
      OTG_SET_ERR(ctx, "plan was built for a different method");
      return OTG_ERR_INVALID_ARGUMENT;
    }
    return otg_rct_launch_batched(ctx, plan, tree, out, nullptr, 0, s);
  }
  otg_rct_plan* tmp = nullptr;
  st = otg_rct_plan_create(ctx, tree, nullptr, method, &tmp);
  if (st != OTG_OK) return st;
  st = otg_rct_launch_batched(ctx, tmp, tree, out, nullptr, 0, s);
  hipError_t e = hipStreamSynchronize(s);
  otg_rct_plan_destroy(tmp);
  if (st == OTG_OK && e != hipSuccess) {
    OTG_SET_ERR(ctx, "solve failed: %s", hipGetErrorString(e));
    st = OTG_ERR_DEVICE;
  }
  return st;
}

extern "C" otg_status otg_rct_solve_subset(otg_context* ctx, otg_rct_plan* plan,
                                           const otg_rct_tree* tree, const otg_rct_result* out,
                                           otg_rct_method method, const int32_t* net_ids,
                                           int32_t count, otg_stream stream) {
  if (!ctx || !plan || !out || count < 0 || (count > 0 && !net_ids))
    return OTG_ERR_INVALID_ARGUMENT;
  otg_status st = check_tree(ctx, tree);
  if (st != OTG_OK) return st;
  if (method == OTG_RCT_PLAIN || plan->method != method) {
    OTG_SET_ERR(ctx, "subset solves need a batched plan of the same method");
    return OTG_ERR_INVALID_ARGUMENT;
  }
  if (plan->num_nets != tree->num_nets) return OTG_ERR_INVALID_ARGUMENT;
  st = check_vec_alignment(ctx, tree, out);
  if (st != OTG_OK) return st;
  return otg_rct_launch_batched(ctx, plan, tree, out, net_ids, count, otg_pick_stream(ctx, stream));
}

extern "C" otg_status otg_rct_bfs(otg_context* ctx, const otg_rct_edges* in, int32_t* node2bfs,
                                  int32_t* pid, int32_t* edges_directed, float* pres,
                                  float* cap_bfs, otg_stream stream) {
  if (!ctx || !in || in->num_nets < 0 || in->total_num_nodes < 0) return OTG_ERR_INVALID_ARGUMENT;
  if (in->num_nets == 0) return OTG_OK;
  if (!in->nodes_start || !in->edges || !in->roots || !in->edge_res || !in->node_cap || !node2bfs ||
      !pid || !edges_directed || !pres || !cap_bfs)
    return OTG_ERR_INVALID_ARGUMENT;
  hipStream_t s = otg_pick_stream(ctx, stream);
  OTG_HIP(ctx, hipSetDevice(ctx->device));
  int32_t* d_bad = nullptr;
  OTG_HIP(ctx, hipMallocAsync((void**)&d_bad, sizeof(int32_t), s));
  const int32_t big = INT32_MAX;
  OTG_HIP(ctx, hipMemcpyAsync(d_bad, &big, sizeof(int32_t), hipMemcpyHostToDevice, s));
  otg_status st = otg_rct_bfs_launch(ctx, in, node2bfs, pid, edges_directed, pres, cap_bfs, d_bad, s);
  int32_t bad = INT32_MAX;
  if (st == OTG_OK) {
    OTG_HIP(ctx, hipMemcpyAsync(&bad, d_bad, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  }
  OTG_HIP(ctx, hipFreeAsync(d_bad, s));
  OTG_HIP(ctx, hipStreamSynchronize(s));
  if (st != OTG_OK) return st;
  if (bad != INT32_MAX) {
    OTG_SET_ERR(ctx, "net %d: edges do not form a tree spanning its nodes from its root", bad);
    return OTG_ERR_NOT_A_TREE;
  }
  return OTG_OK;
}

namespace {
// One thread per (pin, corner): four consecutive lanes read the four corners
// of one node, 16 contiguous bytes.
__global__ void rct_query_kernel(int32_t num_nets, const int32_t* __restrict__ nodes_start,
                                 const int32_t* __restrict__ node2bfs, const float* __restrict__ d,
                                 const float* __restrict__ sl, const float* __restrict__ ld,
                                 int32_t count, const int32_t* __restrict__ net,
                                 const int32_t* __restrict__ local, float* delay, float* slew,
                                 float* load) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 4 * (int64_t)count) return;
  const int64_t q = i >> 2;
  const int c = (int)(i & 3);
  const int32_t k = net[q];
  float vd = __builtin_nanf(""), vs = vd, vl = vd;
  if (k >= 0 && k < num_nets) {
    const int32_t b = nodes_start[k], n = nodes_start[k + 1] - b, u = local[q];
    if (u >= 0 && u < n) {
      const int64_t pos = (int64_t)b + node2bfs[(int64_t)b + u];
      if (d) vd = d[4 * pos + c];
      if (sl) vs = sl[4 * pos + c];
      if (ld) vl = ld[4 * (int64_t)b + c];  // the driver is BFS position 0
    }
  }
  if (delay) delay[i] = vd;
  if (slew) slew[i] = vs;
  if (load) load[i] = vl;
}
}  // namespace

extern "C" otg_status otg_rct_query(otg_context* ctx, int32_t num_nets, const int32_t* nodes_start,
                                    const int32_t* node2bfs, const otg_rct_result* res,
                                    int32_t count, const int32_t* net, const int32_t* local,
                                    float* delay, float* slew, float* load, otg_stream stream) {
  if (!ctx || !res || count < 0 || num_nets < 0) return OTG_ERR_INVALID_ARGUMENT;
  if (count == 0) return OTG_OK;
  if (!nodes_start || !node2bfs || !net || !local) return OTG_ERR_INVALID_ARGUMENT;
  if ((delay && !res->delay) || (slew && !res->slew) || (load && !res->load)) {
    OTG_SET_ERR(ctx, "query asks for an output the result does not hold");
    return OTG_ERR_INVALID_ARGUMENT;
  }
  hipStream_t s = otg_pick_stream(ctx, stream);
  OTG_HIP(ctx, hipSetDevice(ctx->device));
  const int64_t threads = 4 * (int64_t)count;
  hipLaunchKernelGGL(rct_query_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s,
                     num_nets, nodes_start, node2bfs, res->delay, res->slew, res->load, count,
                     net, local, delay, slew, load);
  OTG_HIP(ctx, hipGetLastError());
  return OTG_OK;
}

namespace {
// RAII list of device buffers for the host-pointer entry points
struct DevBufs {
  std::vector<void*> p;
  ~DevBufs() {
    for (void* q : p) (void)hipFree(q);
  }
  template <class T>
  hipError_t alloc(T** out, size_t count) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(1, count) * sizeof(T));
    if (e == hipSuccess) p.push_back(q);
    *out = reinterpret_cast<T*>(q);
    return e;
  }
};
}  // namespace

namespace {
// Stage clock of the host-pointer entry points: each mark synchronises the
// stream and records the milliseconds since the previous mark.
struct StageClock {
  otg_context* ctx;
  hipStream_t s;
  std::chrono::steady_clock::time_point t0, last;
  explicit StageClock(otg_context* c, hipStream_t st) : ctx(c), s(st) {
    for (double& v : ctx->last_breakdown) v = 0.0;
    t0 = last = std::chrono::steady_clock::now();
  }
  hipError_t mark(int stage) {
    const hipError_t e = hipStreamSynchronize(s);
    const auto now = std::chrono::steady_clock::now();
    ctx->last_breakdown[stage] += std::chrono::duration<double, std::milli>(now - last).count();
    ctx->last_breakdown[7] = std::chrono::duration<double, std::milli>(now - t0).count();
    last = now;
    return e;
  }
};

// Solve on device arrays with a plan built for this call (none for the
// plain method), timing plan and solve separately.
otg_status solve_timed(otg_context* ctx, const otg_rct_tree* tree, const otg_rct_result* out,
                       otg_rct_method method, hipStream_t s, StageClock& clk) {
  otg_rct_plan* plan = nullptr;
  if (method != OTG_RCT_PLAIN) {
    otg_status st = otg_rct_plan_create(ctx, tree, nullptr, method, &plan);
    if (st != OTG_OK) return st;
  }
  OTG_HIP(ctx, clk.mark(3));
  otg_status st = otg_rct_solve(ctx, plan, tree, out, method, s);
  const hipError_t e = clk.mark(4);
  otg_rct_plan_destroy(plan);
  if (st == OTG_OK && e != hipSuccess) {
    OTG_SET_ERR(ctx, "solve failed: %s", hipGetErrorString(e));
    st = OTG_ERR_DEVICE;
  }
  return st;
}
}  // namespace

extern "C" otg_status otg_rct_last_breakdown(const otg_context* ctx, double ms[8]) {
  if (!ctx || !ms) return OTG_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < 8; ++i) ms[i] = ctx->last_breakdown[i];
  return OTG_OK;
}

// A copy between a caller's host array and device memory, staged through
// pinned memory by otg_context::kStage host threads: each takes a contiguous
// share and moves it in kStageBytes chunks, the DMA of one chunk overlapping
// its own copy of the chunk before (a copy to or from pageable memory is
// otherwise staged by the runtime one chunk at a time, ~17 GB/s device to
// host measured on MI355X).  Synchronous: the device side of `dst` / `src`
// must be ready (the caller synchronises stream `s` first).  Copies below
// four chunks, and every copy if the pinned buffers cannot be allocated, are
// queued on `s` unchanged.
static hipError_t staged_copy_raw(otg_context* ctx, void* dst, const void* src, size_t bytes,
                                  hipMemcpyKind kind, hipStream_t s) {
  constexpr int T = otg_context::kStage;
  constexpr size_t C = otg_context::kStageBytes;
  bool staged = bytes >= 4 * C;
  for (int t = 0; t < T && staged; ++t) {
    if (!ctx->stage_stream[t] &&
        hipStreamCreateWithFlags(&ctx->stage_stream[t], hipStreamNonBlocking) != hipSuccess) {
      ctx->stage_stream[t] = nullptr;
      staged = false;
    }
    for (void*& b : ctx->stage_buf[t])
      if (staged && !b && hipHostMalloc(&b, C, hipHostMallocDefault) != hipSuccess) {
        b = nullptr;
        staged = false;
      }
  }
  if (!staged) {
    // small copies, or no pinned memory to be had: the runtime's own staging
    (void)hipGetLastError();
    return hipMemcpyAsync(dst, src, bytes, kind, s);
  }
  const bool d2h = kind == hipMemcpyDeviceToHost;
  const size_t share = ((bytes + T - 1) / T + 4095) & ~size_t(4095);
  hipError_t err[T];
  std::vector<std::thread> pool;
  pool.reserve(T);
  for (int t = 0; t < T; ++t) {
    err[t] = hipSuccess;
    const size_t a = std::min(bytes, t * share), b = std::min(bytes, a + share);
    pool.emplace_back([=, &err]() {
      hipError_t e = hipSetDevice(ctx->device);
      hipStream_t st = ctx->stage_stream[t];
      char* const* buf = reinterpret_cast<char* const*>(ctx->stage_buf[t]);
      char* out = static_cast<char*>(dst);
      const char* in = static_cast<const char*>(src);
      size_t prev_off = 0, prev_len = 0;
      int cur = 0;
      for (size_t off = a; off < b && e == hipSuccess; off += C, cur ^= 1) {
        const size_t len = std::min(C, b - off);
        if (d2h) {
          // chunk k lands in one buffer while chunk k-1 leaves the other
          e = hipMemcpyAsync(buf[cur], in + off, len, hipMemcpyDeviceToHost, st);
          if (prev_len) std::memcpy(out + prev_off, buf[cur ^ 1], prev_len);
          if (e == hipSuccess) e = hipStreamSynchronize(st);
          prev_off = off;
          prev_len = len;
        } else {
          // chunk k is filled while chunk k-1 is on its way (waited for
          // before its buffer is reused)
          std::memcpy(buf[cur], in + off, len);
          e = hipStreamSynchronize(st);
          if (e == hipSuccess) e = hipMemcpyAsync(out + off, buf[cur], len, hipMemcpyHostToDevice, st);
        }
      }
      if (e == hipSuccess) {
        if (d2h && prev_len) std::memcpy(out + prev_off, buf[cur ^ 1], prev_len);
        if (!d2h) e = hipStreamSynchronize(st);
      }
      err[t] = e;
    });
  }
  for (std::thread& th : pool) th.join();
  for (int t = 0; t < T; ++t)
    if (err[t] != hipSuccess) return err[t];
  return hipSuccess;
}

static otg_status staged_copy(otg_context* ctx, void* dst, const void* src, size_t bytes,
                              hipMemcpyKind kind, hipStream_t s) {
  OTG_HIP(ctx, staged_copy_raw(ctx, dst, src, bytes, kind, s));
  return OTG_OK;
}

// Device-to-host copies run by a host thread beside the caller's further
// work (staged_copy_raw, which leaves the context's error message alone);
// joined by wait() or, on every early return, by the destructor.
struct BackgroundCopies {
  std::thread th;
  hipError_t err = hipSuccess;
  template <class F>
  void start(F f) {
    th = std::thread([this, f]() { err = f(); });
  }
  hipError_t wait() {
    if (th.joinable()) th.join();
    return err;
  }
  ~BackgroundCopies() {
    if (th.joinable()) th.join();
  }
};

extern "C" otg_status otg_rct_bfs_and_compute(otg_context* ctx, otg_rct_edge_array* d,
                                              otg_rct_method method) {
  if (!ctx || !d || d->num_nets < 0 || d->total_num_nodes < 0) return OTG_ERR_INVALID_ARGUMENT;
  if (d->num_nets == 0) return OTG_OK;
  if (d->total_num_edges != d->total_num_nodes - d->num_nets) {
    OTG_SET_ERR(ctx, "total_num_edges must equal total_num_nodes - num_nets (a forest)");
    return OTG_ERR_INVALID_ARGUMENT;
  }
  if (!d->rct_nodes_start || !d->rct_edges || !d->rct_roots || !d->rct_edges_res ||
      !d->rct_nodes_cap || !d->rct_node2bfs_order || !d->rct_pid || !d->load || !d->delay ||
      !d->ldelay || !d->impulse)
    return OTG_ERR_INVALID_ARGUMENT;
  OTG_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  StageClock clk(ctx, s);
  const int64_t N = d->num_nets, V = d->total_num_nodes, E = d->total_num_edges;
  otg_status st = OTG_OK;
  {
    DevBufs B;
    int32_t *ns, *ed, *rt, *n2b, *pid;
    float *res, *ncap, *pres, *cap, *si = nullptr;
    float *load, *delay, *ldelay, *beta, *imp, *slew = nullptr;
    OTG_HIP(ctx, B.alloc(&ns, N + 1));
    OTG_HIP(ctx, B.alloc(&ed, 2 * E));
    OTG_HIP(ctx, B.alloc(&rt, N));
    OTG_HIP(ctx, B.alloc(&res, E));
    OTG_HIP(ctx, B.alloc(&ncap, 4 * V));
    OTG_HIP(ctx, B.alloc(&n2b, V));
    OTG_HIP(ctx, B.alloc(&pid, V));
    OTG_HIP(ctx, B.alloc(&pres, V));
    OTG_HIP(ctx, B.alloc(&cap, 4 * V));
    OTG_HIP(ctx, B.alloc(&load, 4 * V));
    OTG_HIP(ctx, B.alloc(&delay, 4 * V));
    OTG_HIP(ctx, B.alloc(&ldelay, 4 * V));
    OTG_HIP(ctx, B.alloc(&beta, 4 * V));
    OTG_HIP(ctx, B.alloc(&imp, 4 * V));
    if (d->slew) OTG_HIP(ctx, B.alloc(&slew, 4 * V));
    if (d->slew_in) OTG_HIP(ctx, B.alloc(&si, 4 * N));
    OTG_HIP(ctx, clk.mark(0));
    auto h2d = [&](void* dst, const void* src, size_t bytes) {
      return staged_copy(ctx, dst, src, bytes, hipMemcpyHostToDevice, s);
    };
    if ((st = h2d(ns, d->rct_nodes_start, sizeof(int32_t) * (N + 1))) != OTG_OK ||
        (st = h2d(ed, d->rct_edges, sizeof(int32_t) * 2 * E)) != OTG_OK ||
        (st = h2d(rt, d->rct_roots, sizeof(int32_t) * N)) != OTG_OK ||
        (st = h2d(res, d->rct_edges_res, sizeof(float) * E)) != OTG_OK ||
        (st = h2d(ncap, d->rct_nodes_cap, sizeof(float) * 4 * V)) != OTG_OK ||
        (si && (st = h2d(si, d->slew_in, sizeof(float) * 4 * N)) != OTG_OK))
      return st;
    OTG_HIP(ctx, clk.mark(1));
    otg_rct_edges in{d->num_nets, d->total_num_nodes, ns, ed, rt, res, ncap};
    st = otg_rct_bfs(ctx, &in, n2b, pid, ed, pres, cap, s);
    if (st != OTG_OK) return st;
    OTG_HIP(ctx, clk.mark(2));
    // the ordering's outputs go back to the host while the plan is built and
    // the forest solved (neither writes them; the thread is joined before
    // the device buffers are released, on every path)
    BackgroundCopies bg;
    bg.start([=]() {
      hipError_t e = hipSetDevice(ctx->device);  // staging streams are made on the current device
      if (e == hipSuccess)
        e = staged_copy_raw(ctx, d->rct_edges, ed, sizeof(int32_t) * 2 * E, hipMemcpyDeviceToHost,
                            s);
      if (e == hipSuccess)
        e = staged_copy_raw(ctx, d->rct_node2bfs_order, n2b, sizeof(int32_t) * V,
                            hipMemcpyDeviceToHost, s);
      if (e == hipSuccess)
        e = staged_copy_raw(ctx, d->rct_pid, pid, sizeof(int32_t) * V, hipMemcpyDeviceToHost, s);
      return e;
    });
    otg_rct_tree tree{d->num_nets, d->total_num_nodes, ns, pid, pres, cap, si};
    otg_rct_result out{load, delay, ldelay, beta, imp, slew};
    st = solve_timed(ctx, &tree, &out, method, s, clk);
    if (st != OTG_OK) return st;
    OTG_HIP(ctx, hipStreamSynchronize(s));
    OTG_HIP(ctx, bg.wait());
    auto d2h = [&](void* dst, const void* src, size_t bytes) {
      return staged_copy(ctx, dst, src, bytes, hipMemcpyDeviceToHost, s);
    };
    if ((st = d2h(d->load, load, sizeof(float) * 4 * V)) != OTG_OK ||
        (st = d2h(d->delay, delay, sizeof(float) * 4 * V)) != OTG_OK ||
        (st = d2h(d->ldelay, ldelay, sizeof(float) * 4 * V)) != OTG_OK ||
        (st = d2h(d->impulse, imp, sizeof(float) * 4 * V)) != OTG_OK ||
        (d->beta && (st = d2h(d->beta, beta, sizeof(float) * 4 * V)) != OTG_OK) ||
        (d->slew && (st = d2h(d->slew, slew, sizeof(float) * 4 * V)) != OTG_OK))
      return st;
    OTG_HIP(ctx, clk.mark(5));
  }  // device buffers released here
  OTG_HIP(ctx, clk.mark(6));
  return st;
}

extern "C" otg_status otg_rct_compute(otg_context* ctx, const otg_rct_host* d,
                                      otg_rct_method method) {
  if (!ctx || !d || d->num_nets < 0 || d->total_num_nodes < 0) return OTG_ERR_INVALID_ARGUMENT;
  if (d->num_nets == 0) return OTG_OK;
  if (!d->rct_nodes_start || !d->rct_pid || !d->pres || !d->cap || !d->load || !d->delay ||
      !d->ldelay || !d->impulse)
    return OTG_ERR_INVALID_ARGUMENT;
  OTG_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  StageClock clk(ctx, s);
  const int64_t N = d->num_nets, V = d->total_num_nodes;
  otg_status st = OTG_OK;
  {
    DevBufs B;
    int32_t *ns, *pid;
    float *pres, *cap, *si = nullptr;
    float *load, *delay, *ldelay, *beta, *imp, *slew = nullptr;
    OTG_HIP(ctx, B.alloc(&ns, N + 1));
    OTG_HIP(ctx, B.alloc(&pid, V));
    OTG_HIP(ctx, B.alloc(&pres, V));
    OTG_HIP(ctx, B.alloc(&cap, 4 * V));
    OTG_HIP(ctx, B.alloc(&load, 4 * V));
    OTG_HIP(ctx, B.alloc(&delay, 4 * V));
    OTG_HIP(ctx, B.alloc(&ldelay, 4 * V));
    OTG_HIP(ctx, B.alloc(&beta, 4 * V));
    OTG_HIP(ctx, B.alloc(&imp, 4 * V));
    if (d->slew) OTG_HIP(ctx, B.alloc(&slew, 4 * V));
    if (d->slew_in) OTG_HIP(ctx, B.alloc(&si, 4 * N));
    OTG_HIP(ctx, clk.mark(0));
    auto h2d = [&](void* dst, const void* src, size_t bytes) {
      return staged_copy(ctx, dst, src, bytes, hipMemcpyHostToDevice, s);
    };
    if ((st = h2d(ns, d->rct_nodes_start, sizeof(int32_t) * (N + 1))) != OTG_OK ||
        (st = h2d(pid, d->rct_pid, sizeof(int32_t) * V)) != OTG_OK ||
        (st = h2d(pres, d->pres, sizeof(float) * V)) != OTG_OK ||
        (st = h2d(cap, d->cap, sizeof(float) * 4 * V)) != OTG_OK ||
        (si && (st = h2d(si, d->slew_in, sizeof(float) * 4 * N)) != OTG_OK))
      return st;
    OTG_HIP(ctx, clk.mark(1));
    otg_rct_tree tree{d->num_nets, d->total_num_nodes, ns, pid, pres, cap, si};
    otg_rct_result out{load, delay, ldelay, beta, imp, slew};
    st = solve_timed(ctx, &tree, &out, method, s, clk);
    if (st != OTG_OK) return st;
    OTG_HIP(ctx, hipStreamSynchronize(s));
    auto d2h = [&](void* dst, const void* src, size_t bytes) {
      return staged_copy(ctx, dst, src, bytes, hipMemcpyDeviceToHost, s);
    };
    if ((st = d2h(d->load, load, sizeof(float) * 4 * V)) != OTG_OK ||
        (st = d2h(d->delay, delay, sizeof(float) * 4 * V)) != OTG_OK ||
        (st = d2h(d->ldelay, ldelay, sizeof(float) * 4 * V)) != OTG_OK ||
        (st = d2h(d->impulse, imp, sizeof(float) * 4 * V)) != OTG_OK ||
        (d->beta && (st = d2h(d->beta, beta, sizeof(float) * 4 * V)) != OTG_OK) ||
        (d->slew && (st = d2h(d->slew, slew, sizeof(float) * 4 * V)) != OTG_OK))
      return st;
    OTG_HIP(ctx, clk.mark(5));
  }
  OTG_HIP(ctx, clk.mark(6));
  return st;
}

// ---------------------------------------------------------------------------
// P2
// ---------------------------------------------------------------------------
extern "C" otg_status otg_levelize(otg_context* ctx, int32_t n, int32_t num_edges,
                                   const int32_t* fanin_start, const int32_t* fanin,
                                   const int32_t* out_degree, const int32_t* first_frontier,
                                   int32_t first_size, int32_t* frontiers, int32_t* frontiers_ends,
                                   int32_t* level_of_pin, int32_t* num_ends, otg_stream stream) {
  if (!ctx || n < 0 || num_edges < 0 || first_size < 0 || first_size > n || !num_ends)
    return OTG_ERR_INVALID_ARGUMENT;
  if (n > 0 && (!fanin_start || !out_degree || !frontiers || !frontiers_ends ||
                (num_edges > 0 && !fanin) || (first_size > 0 && !first_frontier)))
    return OTG_ERR_INVALID_ARGUMENT;
  hipStream_t s = otg_pick_stream(ctx, stream);
  OTG_HIP(ctx, hipSetDevice(ctx->device));
  if (n == 0) {
    *num_ends = 2;
    const int32_t z[2] = {0, 0};
    if (frontiers_ends) OTG_HIP(ctx, hipMemcpyAsync(frontiers_ends, z, sizeof(z), hipMemcpyHostToDevice, s));
    OTG_HIP(ctx, hipStreamSynchronize(s));
    return OTG_OK;
  }
  int32_t *out_copy = nullptr, *scratch = nullptr;
  OTG_HIP(ctx, hipMallocAsync((void**)&out_copy, sizeof(int32_t) * n, s));
  OTG_HIP(ctx, hipMallocAsync((void**)&scratch, sizeof(int32_t) * ((int64_t)n + 3), s));
  OTG_HIP(ctx, hipMemcpyAsync(out_copy, out_degree, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
  if (first_size > 0 && first_frontier != frontiers)
    OTG_HIP(ctx, hipMemcpyAsync(frontiers, first_frontier, sizeof(int32_t) * first_size,
                                hipMemcpyDeviceToDevice, s));
  int32_t ne = 0;
  otg_status st = otg_levelize_launch_raw(ctx, n, first_size, fanin_start, fanin, out_copy, frontiers,
                                          frontiers_ends, level_of_pin, scratch, &ne, s, true);
  OTG_HIP(ctx, hipFreeAsync(out_copy, s));
  OTG_HIP(ctx, hipFreeAsync(scratch, s));
  OTG_HIP(ctx, hipStreamSynchronize(s));
  if (st != OTG_OK) return st;
  *num_ends = ne;
  return OTG_OK;
}

extern "C" otg_status otg_toposort_compute(otg_context* ctx, int32_t n, int32_t num_edges,
                                           int32_t first_size, const int32_t* edgelist_start,
                                           const int32_t* edgelist, const int32_t* out,
                                           int32_t* frontiers, int32_t* frontiers_ends,
                                           int32_t* num_ends, int32_t* level_of_pin) {
  if (!ctx || n < 0 || num_edges < 0 || first_size < 0 || first_size > n || !num_ends ||
      !frontiers_ends)
    return OTG_ERR_INVALID_ARGUMENT;
  if (n > 0 && (!edgelist_start || !out || !frontiers || (num_edges > 0 && !edgelist)))
    return OTG_ERR_INVALID_ARGUMENT;
  if (n == 0) {
    frontiers_ends[0] = 0;
    frontiers_ends[1] = 0;
    *num_ends = 2;
    return OTG_OK;
  }
  OTG_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  DevBufs B;
  int32_t *st_d, *el_d, *out_d, *fr_d, *ends_d, *lv_d = nullptr;
  OTG_HIP(ctx, B.alloc(&st_d, (size_t)n + 1));
  OTG_HIP(ctx, B.alloc(&el_d, (size_t)num_edges));
  OTG_HIP(ctx, B.alloc(&out_d, (size_t)n));
  OTG_HIP(ctx, B.alloc(&fr_d, (size_t)n));
  OTG_HIP(ctx, B.alloc(&ends_d, (size_t)n + 2));
  if (level_of_pin) OTG_HIP(ctx, B.alloc(&lv_d, (size_t)n));
  otg_status st = OTG_OK;
  auto copy = [&](void* dst, const void* src, size_t count, hipMemcpyKind kind) {
    return staged_copy(ctx, dst, src, sizeof(int32_t) * count, kind, s);
  };
  constexpr hipMemcpyKind H2D = hipMemcpyHostToDevice, D2H = hipMemcpyDeviceToHost;
  if ((st = copy(st_d, edgelist_start, (size_t)n + 1, H2D)) != OTG_OK ||
      (num_edges && (st = copy(el_d, edgelist, (size_t)num_edges, H2D)) != OTG_OK) ||
      (st = copy(out_d, out, (size_t)n, H2D)) != OTG_OK ||
      (st = copy(fr_d, frontiers, (size_t)n, H2D)) != OTG_OK)
    return st;
  int32_t ne = 0;
  st = otg_levelize(ctx, n, num_edges, st_d, el_d, out_d, fr_d, first_size, fr_d, ends_d, lv_d, &ne, s);
  if (st != OTG_OK) return st;
  OTG_HIP(ctx, hipStreamSynchronize(s));
  if ((st = copy(frontiers, fr_d, (size_t)n, D2H)) != OTG_OK ||
      (st = copy(frontiers_ends, ends_d, (size_t)ne, D2H)) != OTG_OK ||
      (lv_d && (st = copy(level_of_pin, lv_d, (size_t)n, D2H)) != OTG_OK))
    return st;
  OTG_HIP(ctx, hipStreamSynchronize(s));
  *num_ends = ne;
  return OTG_OK;
}

// ---------------------------------------------------------------------------
// Autotuner
// ---------------------------------------------------------------------------
static bool write_tuning(const char* path, const otg_rct_tuning& t, otg_rct_method m, double ms,
                         int evaluated) {
  FILE* f = fopen(path, "w");
  if (!f) return false;
  fprintf(f,
          "{\"ring_slots\": %d, \"min_ring\": %d, \"small_target\": %d, "
          "\"grid_per_cu\": %d, \"streams\": %d, \"small_max\": %d, \"ring_mask\": %d, "
          "\"first_on_caller\": %d, \"small_grid_per_cu\": %d, \"group_alongside\": %d, "
          "\"method\": %d, \"ms\": %.6f, \"candidates\": %d}\n",
          t.ring_slots, t.min_ring, t.small_target, t.grid_per_cu, t.streams, t.small_max,
          t.ring_mask, t.first_on_caller, t.small_grid_per_cu, t.group_alongside, (int)m, ms,
          evaluated);
  fclose(f);
  return true;
}

static int json_int(const std::string& s, const char* key, int dflt) {
  std::string k = std::string("\"") + key + "\"";
  size_t p = s.find(k);
  if (p == std::string::npos) return dflt;
  p = s.find(':', p);
  if (p == std::string::npos) return dflt;
  return atoi(s.c_str() + p + 1);
}

extern "C" otg_status otg_rct_tuning_load(otg_context* ctx, const char* path, otg_rct_tuning* t) {
  if (!ctx || !path) return OTG_ERR_INVALID_ARGUMENT;
  FILE* f = fopen(path, "r");
  if (!f) {
    OTG_SET_ERR(ctx, "cannot open %s", path);
    return OTG_ERR_INVALID_ARGUMENT;
  }
  std::string s;
  char buf[512];
  size_t k;
  while ((k = fread(buf, 1, sizeof(buf), f)) > 0) s.append(buf, k);
  fclose(f);
  otg_rct_tuning r{};
  r.ring_slots = json_int(s, "ring_slots", 0);
  r.min_ring = json_int(s, "min_ring", 0);
  r.small_target = json_int(s, "small_target", 0);
  r.grid_per_cu = json_int(s, "grid_per_cu", 0);
  r.streams = json_int(s, "streams", 0);
  r.small_max = json_int(s, "small_max", 0);
  r.ring_mask = json_int(s, "ring_mask", 0);
  r.first_on_caller = json_int(s, "first_on_caller", 0);
  r.small_grid_per_cu = json_int(s, "small_grid_per_cu", 0);
  r.group_alongside = json_int(s, "group_alongside", 0);
  ctx->tuning = r;
  if (t) *t = r;
  return OTG_OK;
}

extern "C" otg_status otg_rct_autotune(otg_context* ctx, const otg_rct_tree* tree,
                                       const otg_rct_result* scratch_out, otg_rct_method method,
                                       const char* json_path, otg_rct_tuning* best) {
  if (!ctx || !tree || !scratch_out) return OTG_ERR_INVALID_ARGUMENT;
  if (method != OTG_RCT_BATCHED && method != OTG_RCT_BATCHED_COMPENSATED)
    return OTG_ERR_INVALID_ARGUMENT;
  otg_status st = check_tree(ctx, tree);
  if (st != OTG_OK) return st;
  OTG_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  // zero fields of a candidate mean the built-in defaults while searching
  // (not a tuning loaded into the context earlier)
  const otg_rct_tuning saved = ctx->tuning;
  ctx->tuning = otg_rct_tuning{};
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    ctx->tuning = saved;
    OTG_SET_ERR(ctx, "autotune: hipEventCreate failed");
    return OTG_ERR_DEVICE;
  }
  int evaluated = 0;
  // ms of one solve with tuning t (plan built and released here); < 0: the
  // candidate did not run
  auto measure = [&](const otg_rct_tuning& t, otg_rct_tuning* used) -> double {
    otg_rct_plan* plan = nullptr;
    if (otg_rct_plan_create(ctx, tree, &t, method, &plan) != OTG_OK) return -1.0;
    double ms = -1.0;
    if (otg_rct_launch_batched(ctx, plan, tree, scratch_out, nullptr, 0, s) == OTG_OK) {  // warm-up
      const int reps = 3;
      bool ok = hipEventRecord(e0, s) == hipSuccess;
      for (int r = 0; r < reps && ok; ++r)
        ok = otg_rct_launch_batched(ctx, plan, tree, scratch_out, nullptr, 0, s) == OTG_OK;
      ok = ok && hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
      float f = 0.f;
      if (ok && hipEventElapsedTime(&f, e0, e1) == hipSuccess) ms = f / reps;
    }
    if (used) *used = plan->tuning;
    otg_rct_plan_destroy(plan);
    ++evaluated;
    return ms;
  };
  // Coordinate descent from the defaults, two passes: every knob in turn
  // takes each of its candidate values with the others at their best so far.
  // The knobs: the smallest ring (LDS per wavefront, hence occupancy, against
  // one more launch), the ring ladder (320 on or off), the largest net of a
  // packed group (the group/stream threshold), the group size, which launch
  // takes the caller's stream, the smallest-ring launch's workgroups per CU,
  // the workgroups per CU of every launch (0: as many as fit), and whether
  // the packed groups run first or beside the streaming launches (with how
  // many workgroups per CU).
  otg_rct_tuning cur{};
  cur.ring_slots = 2048;
  otg_rct_tuning resolved{};
  double best_ms = measure(cur, &resolved);
  if (best_ms < 0) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->tuning = saved;
    OTG_SET_ERR(ctx, "autotune: the default tuning did not run");
    return OTG_ERR_DEVICE;
  }
  cur = resolved;
  const int all = (1 << kRingSteps) - 1;
  struct Knob {
    int32_t otg_rct_tuning::*field;
    std::vector<int32_t> values;
  };
  const std::vector<Knob> knobs = {
      {&otg_rct_tuning::min_ring, {256, 320, 512}},
      {&otg_rct_tuning::ring_mask, {all, all & ~2}},
      {&otg_rct_tuning::small_max, {64, 128, 256}},
      {&otg_rct_tuning::small_target, {128, 256}},
      {&otg_rct_tuning::first_on_caller, {0, 1}},
      {&otg_rct_tuning::small_grid_per_cu, {0, 4, 6}},
      {&otg_rct_tuning::grid_per_cu, {0, 4}},
      {&otg_rct_tuning::group_alongside, {0, 1, 2, 4}},
  };
  for (int pass = 0; pass < 2; ++pass) {
    for (const Knob& k : knobs) {
      for (int32_t v : k.values) {
        if (cur.*k.field == v) continue;
        otg_rct_tuning t = cur;
        t.*k.field = v;
        otg_rct_tuning used{};
        const double ms = measure(t, &used);
        if (ms >= 0 && ms < best_ms) {
          best_ms = ms;
          cur = used;
        }
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  ctx->tuning = cur;
  if (best) *best = cur;
  if (json_path && !write_tuning(json_path, cur, method, best_ms, evaluated)) {
    OTG_SET_ERR(ctx, "cannot write %s", json_path);
    return OTG_ERR_INVALID_ARGUMENT;
  }
  return OTG_OK;
}
