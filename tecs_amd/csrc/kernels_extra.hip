// kernels_extra.hip - the generic wall forces, the link forces and cell division
// (Integrator::evaluate_force_impl for GenericForceTag / LinkForceTag, base.hpp:116-230, with the
// functors of examples/growth_w_wall.cpp, growth_w_squeeze.cpp and intercalation.cpp).
//
// Within an evaluation the reference applies the generic forces before the pairwise force and
// the link forces after it (fuse_forces, kernel_fuser.hpp:148-156). The pair sweep writes its
// result F = (0 + acc) + 0 into the delta buffer; k_extra then forms the walls' sum W the way
// the reference's delta does (from zero, wall after wall) and stores W + F, which is bitwise
// (W + acc) + 0 (addition commutes, and neither W nor F is -0). The same thread then adds
// the cell's link contributions, in link order (k_extra).
#include "tecs_internal.hpp"

namespace tecs_amd {
namespace {

// PlaneN::signed_distance (types/plane.hpp:36-42): d + a x + b y + c z, in that order; the
// walls' sum W formed from zero, wall after wall.
__device__ __forceinline__ void wall_sum(const InstanceDesc& in, float4 p, float* W) {
  for (int w = 0; w < in.n_walls; ++w) {
    const float* c = in.wall[w];
    float sd = c[3];
    sd = sd + c[0] * p.x;
    sd = sd + c[1] * p.y;
    sd = sd + c[2] * p.z;
    if (sd < 0.0f) {
      const float f = -sd * c[4];
      W[0] = W[0] + c[0] * f;
      W[1] = W[1] + c[1] * f;
      W[2] = W[2] + c[2] * f;
    }
  }
}

// One thread per cell: the walls (W + F), then its incident links in link order, rows[slot_begin
// + agent] .. rows[slot_begin + agent + 1] of ent (the other endpoint, +1 where the cell is the
// link's a, -1 where it is its b). u = p_b - p_a, d = |u|; the a end adds 0 + (k u) / d, the b
// end 0 + (-k u) / d, as the reference's zeroed pack per endpoint (base.hpp:210-226). Each end
// forms u and d itself, with the same operations, so both see the same values. The reference
// adds the packs with atomics in no fixed order; here each cell adds them in link order.
template <bool LINKS>
__global__ __launch_bounds__(kBlock) void k_extra(const InstanceDesc* inst, const int* block_inst,
                                                  const int* id, const int* inv, const int* rows,
                                                  const int2* ent, const float4* xs, float4* d) {
  const int lb = (int)blockIdx.x;
  const InstanceDesc& in = inst[block_inst[lb]];
  const int s = lb * kBlock + (int)threadIdx.x;
  if (s - in.slot_begin >= in.n) return;
  int e0 = 0, e1 = 0;
  if (LINKS) {
    const int me = id[s];
    e0 = rows[in.slot_begin + me];
    e1 = rows[in.slot_begin + me + 1];
  }
  if (in.n_walls == 0 && e0 == e1) return;
  const float4 p = xs[s];
  float4 F = d[s];
  if (in.n_walls > 0) {
    float W[3] = {0.0f, 0.0f, 0.0f};
    wall_sum(in, p, W);
    F = make_float4(W[0] + F.x, W[1] + F.y, W[2] + F.z, F.w);
  }
  const float k = in.link_k;
  for (int e = e0; e < e1; ++e) {
    const int2 L = ent[e];
    if (L.x >= in.n) continue;  // an endpoint beyond the active count
    const float4 o = xs[inv[in.slot_begin + L.x]];
    const bool is_a = L.y > 0;
    // u = p_b - p_a
    const float u0 = is_a ? o.x - p.x : p.x - o.x;
    const float u1 = is_a ? o.y - p.y : p.y - o.y;
    const float u2 = is_a ? o.z - p.z : p.z - o.z;
    const float dist = sqrtf(len2(u0, u1, u2));
    const float kk = is_a ? k : -k;
    F.x = F.x + (0.0f + (kk * u0) / dist);
    F.y = F.y + (0.0f + (kk * u1) / dist);
    F.z = F.z + (0.0f + (kk * u2) / dist);
  }
  d[s] = F;
}

// inv[slot_begin + agent] = slot, over every slot of the batch (the ids of a padded range are a
// permutation of it)
__global__ __launch_bounds__(kBlock) void k_inverse_ids(const InstanceDesc* inst,
                                                        const int* block_inst, const int* id,
                                                        int* inv) {
  const int lb = (int)blockIdx.x;
  const InstanceDesc& in = inst[block_inst[lb]];
  const int s = lb * kBlock + (int)threadIdx.x;
  const int a = id[s];
  if ((unsigned)a < (unsigned)in.padded) inv[in.slot_begin + a] = s;
}

// Daughter q of a division: parent agent index parents[q], direction dirs[3q..3q+2]; the
// daughter is agent first + q, inactive, so its slot is slot_begin + first + q in both buffer
// generations (inactive cells sit in index order there).
__global__ __launch_bounds__(kBlock) void k_divide(const InstanceDesc* inst, int ii, const int* inv,
                                                   const int* parents, const float* dirs, int count,
                                                   int first, float offset, float4* x_cur,
                                                   float4* x_other) {
  const int q = (int)blockIdx.x * kBlock + (int)threadIdx.x;
  if (q >= count) return;
  const InstanceDesc& in = inst[ii];
  const float4 p = x_cur[inv[in.slot_begin + parents[q]]];
  const int s = in.slot_begin + first + q;
  const float3 v = make_float3(p.x + offset * dirs[3 * q], p.y + offset * dirs[3 * q + 1],
                               p.z + offset * dirs[3 * q + 2]);
  x_cur[s] = make_float4(v.x, v.y, v.z, x_cur[s].w);
  x_other[s] = make_float4(v.x, v.y, v.z, x_other[s].w);
}

}  // namespace

void launch_inverse_ids(const InstanceDesc* inst, const int* block_inst, const int* id, int* inv,
                        int n_blocks, hipStream_t stream) {
  hipLaunchKernelGGL(k_inverse_ids, dim3(n_blocks), dim3(kBlock), 0, stream, inst, block_inst, id,
                     inv);
}

void launch_extra(const InstanceDesc* inst, const int* block_inst, const int* id, const int* inv,
                  const int* rows, const int2* ent, const float4* xs, float4* d, int n_blocks,
                  hipStream_t stream) {
  if (rows)
    hipLaunchKernelGGL(k_extra<true>, dim3(n_blocks), dim3(kBlock), 0, stream, inst, block_inst,
                       id, inv, rows, ent, xs, d);
  else
    hipLaunchKernelGGL(k_extra<false>, dim3(n_blocks), dim3(kBlock), 0, stream, inst, block_inst,
                       id, inv, rows, ent, xs, d);
}

void launch_divide(const InstanceDesc* inst, int ii, const int* inv, const int* parents,
                   const float* dirs, int count, int first, float offset, float4* x_cur,
                   float4* x_other, hipStream_t stream) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_divide, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, inst,
                     ii, inv, parents, dirs, count, first, offset, x_cur, x_other);
}

}  // namespace tecs_amd
