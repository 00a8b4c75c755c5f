// kernels_extra.hip - the generic wall forces, the link forces and cell division
// (Integrator::evaluate_force_impl for GenericForceTag / LinkForceTag, base.hpp:116-230, with the
// functors of examples/growth_w_wall.cpp, growth_w_squeeze.cpp and intercalation.cpp).
//
// Within an evaluation the reference applies the generic forces before the pairwise force and
// the link forces after it (fuse_forces, kernel_fuser.hpp:148-156). The pair sweep writes its
// result F = (0 + acc) + 0 into the delta buffer; k_walls then forms the walls' sum W the way
// the reference's delta does (from zero, wall after wall) and stores W + F, which is bitwise
// (W + acc) + 0 (addition commutes, and neither W nor F is -0). k_links adds each link's
// contribution to both endpoints with float atomics, as the reference does.
#include "tecs_internal.hpp"

namespace tecs_amd {
namespace {

__global__ __launch_bounds__(kBlock) void k_walls(const InstanceDesc* inst, const int* block_inst,
                                                  const float4* xs, float4* d) {
  const int lb = (int)blockIdx.x;
  const InstanceDesc& in = inst[block_inst[lb]];
  const int s = lb * kBlock + (int)threadIdx.x;
  if (in.n_walls == 0 || s - in.slot_begin >= in.n) return;
  const float4 p = xs[s];
  float W[3] = {0.0f, 0.0f, 0.0f};
  for (int w = 0; w < in.n_walls; ++w) {
    const float* c = in.wall[w];
    // PlaneN::signed_distance (types/plane.hpp:36-42): d + a x + b y + c z, in that order
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
  const float4 F = d[s];
  d[s] = make_float4(W[0] + F.x, W[1] + F.y, W[2] + F.z, F.w);
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

// One thread per link: (instance, a, b, strength bits). u = p_b - p_a, d = |u|;
// delta_a += 0 + (k u) / d, delta_b += 0 + (-k u) / d (a zeroed pack per endpoint).
__global__ __launch_bounds__(kBlock) void k_links(const InstanceDesc* inst, const int4* links,
                                                  int n_links, const int* inv,
                                                  const float4* xs, float4* d) {
  const int q = (int)blockIdx.x * kBlock + (int)threadIdx.x;
  if (q >= n_links) return;
  const int4 L = links[q];
  const InstanceDesc& in = inst[L.w];
  if (L.x == L.y || L.x >= in.n || L.y >= in.n) return;
  const int sa = inv[in.slot_begin + L.x], sb = inv[in.slot_begin + L.y];
  const float k = __int_as_float(L.z);
  const float4 pa = xs[sa], pb = xs[sb];
  const float u[3] = {pb.x - pa.x, pb.y - pa.y, pb.z - pa.z};
  const float dist = sqrtf(len2(u[0], u[1], u[2]));
  float* da = &d[sa].x;
  float* db = &d[sb].x;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    unsafeAtomicAdd(da + c, 0.0f + (k * u[c]) / dist);
    unsafeAtomicAdd(db + c, 0.0f + (-k * u[c]) / dist);
  }
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

void launch_walls(const InstanceDesc* inst, const int* block_inst, const float4* xs, float4* d,
                  int n_blocks, hipStream_t stream) {
  hipLaunchKernelGGL(k_walls, dim3(n_blocks), dim3(kBlock), 0, stream, inst, block_inst, xs, d);
}

void launch_inverse_ids(const InstanceDesc* inst, const int* block_inst, const int* id, int* inv,
                        int n_blocks, hipStream_t stream) {
  hipLaunchKernelGGL(k_inverse_ids, dim3(n_blocks), dim3(kBlock), 0, stream, inst, block_inst, id,
                     inv);
}

void launch_links(const InstanceDesc* inst, const int4* links, int n_links, const int* inv,
                  const float4* xs, float4* d, hipStream_t stream) {
  if (n_links <= 0) return;
  hipLaunchKernelGGL(k_links, dim3((n_links + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, inst,
                     links, n_links, inv, xs, d);
}

void launch_divide(const InstanceDesc* inst, int ii, const int* inv, const int* parents,
                   const float* dirs, int count, int first, float offset, float4* x_cur,
                   float4* x_other, hipStream_t stream) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_divide, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, inst,
                     ii, inv, parents, dirs, count, first, offset, x_cur, x_other);
}

}  // namespace tecs_amd
