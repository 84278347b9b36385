// Measuring instrument: the streaming ceiling of one P1 solve's traffic.
//
// A batched solve must read, per node, the parent offset, the edge
// resistance and the four corner capacitances (24 B), per net the node
// offset and the four driver slews (20 B), and write the six outputs of every
// node (96 B).  This kernel moves exactly those bytes over the same arrays,
// coalesced and with no tree work between them, so its time is what a solve
// over those arrays could reach if it were limited by memory traffic alone.
// Its outputs are a fixed function of the inputs (they depend on every load,
// so none is elided) and carry no timing meaning.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "otg_internal.hpp"

namespace {

constexpr int kCeilBlock = 256;

__global__ __launch_bounds__(kCeilBlock) void rct_ceiling_kernel(
    int32_t num_nodes, int32_t num_nets, const int32_t* __restrict__ nodes_start,
    const int32_t* __restrict__ pid, const float* __restrict__ pres,
    const float4* __restrict__ cap, const float4* __restrict__ slew_in, float4* __restrict__ load,
    float4* __restrict__ delay, float4* __restrict__ ldelay, float4* __restrict__ beta,
    float4* __restrict__ impulse, float4* __restrict__ slew) {
  const int64_t stride = (int64_t)gridDim.x * kCeilBlock;
  const int64_t t0 = (int64_t)blockIdx.x * kCeilBlock + threadIdx.x;
  for (int64_t i = t0; i < num_nodes; i += stride) {
    const float f = pres[i] + (float)pid[i];
    const float4 c = cap[i];
    load[i] = c;
    delay[i] = make_float4(c.x * f, c.y * f, c.z * f, c.w * f);
    ldelay[i] = make_float4(c.x + f, c.y + f, c.z + f, c.w + f);
    beta[i] = make_float4(c.x - f, c.y - f, c.z - f, c.w - f);
    impulse[i] = make_float4(f, c.x, c.y, c.z);
    slew[i] = make_float4(c.w, f, c.x, c.y);
  }
  // per net: the node offset and the driver slews, read; the store below
  // depends on both but never happens for a valid forest (node offsets are
  // not negative), so the outputs stay a function of the node data alone
  for (int64_t k = t0; k < num_nets; k += stride) {
    const int32_t p = nodes_start[k];
    const float4 si = slew_in ? slew_in[k] : make_float4(0.f, 0.f, 0.f, 0.f);
    if (p < 0 && num_nodes > 0) load[0] = si;
  }
}

}  // namespace

extern "C" otg_status otg_rct_stream_ceiling(otg_context* ctx, const otg_rct_tree* tree,
                                             const otg_rct_result* out, otg_stream stream) {
  if (!ctx || !tree || !out || tree->num_nets < 0 || tree->total_num_nodes < 0)
    return OTG_ERR_INVALID_ARGUMENT;
  if (tree->total_num_nodes > 0 &&
      (!tree->pid || !tree->pres || !tree->cap || !out->load || !out->delay || !out->ldelay ||
       !out->beta || !out->impulse || !out->slew))
    return OTG_ERR_INVALID_ARGUMENT;
  if (tree->num_nets > 0 && !tree->nodes_start) return OTG_ERR_INVALID_ARGUMENT;
  for (const void* p : {(const void*)tree->cap, (const void*)tree->slew_in, (const void*)out->load,
                        (const void*)out->delay, (const void*)out->ldelay, (const void*)out->beta,
                        (const void*)out->impulse, (const void*)out->slew}) {
    if (reinterpret_cast<uintptr_t>(p) & 15u) {
      OTG_SET_ERR(ctx, "otg_rct_stream_ceiling: corner arrays must be 16-byte aligned");
      return OTG_ERR_INVALID_ARGUMENT;
    }
  }
  if (tree->total_num_nodes == 0 && tree->num_nets == 0) return OTG_OK;
  OTG_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = otg_pick_stream(ctx, stream);
  const int64_t work = std::max<int64_t>(tree->total_num_nodes, tree->num_nets);
  // 32 workgroups per CU: the best of 4..32 for a stream of this shape on
  // MI355X
  const int grid = (int)std::max<int64_t>(
      1, std::min<int64_t>(32LL * std::max(1, ctx->num_cus), (work + kCeilBlock - 1) / kCeilBlock));
  hipLaunchKernelGGL(rct_ceiling_kernel, dim3(grid), dim3(kCeilBlock), 0, s, tree->total_num_nodes,
                     tree->num_nets, tree->nodes_start, tree->pid, tree->pres,
                     reinterpret_cast<const float4*>(tree->cap),
                     reinterpret_cast<const float4*>(tree->slew_in),
                     reinterpret_cast<float4*>(out->load), reinterpret_cast<float4*>(out->delay),
                     reinterpret_cast<float4*>(out->ldelay), reinterpret_cast<float4*>(out->beta),
                     reinterpret_cast<float4*>(out->impulse), reinterpret_cast<float4*>(out->slew));
  OTG_HIP(ctx, hipGetLastError());
  return OTG_OK;
}
