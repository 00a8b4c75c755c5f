// kernels_query.hip - batched grid queries (acceleration::Grid, grid.hpp:413-564): one thread
// per query point, on the grid of the last rebuild (slot order and bin offsets) with the current
// positions, visiting bins and members in the grid's order (bins x-fastest, members ascending in
// agent index), so the first minimum and the k-th candidate are the reference's.
#include "tecs_internal.hpp"

#include <cfloat>
#include <mutex>

namespace tecs_amd {

// Staging of the allocation-free nearest query (tecs_sim_nearest_noalloc): the query points
// (3 floats each) from offset 0, the results (one int32 each) from kNoAllocOutOffset. Each GPU
// holds one copy of the array for the life of the process; every call on that GPU uses it, so
// calls that overlap on one GPU race. Host code never passes the array's name to a kernel or a
// copy: it is a host-side stand-in there; the device address comes from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

int noalloc_scratch(void** out) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static void* addr[kMaxDevices] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= kMaxDevices) return (int)hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(mu);
  if (!addr[dev]) {
    // once per device, with that device current: this GPU's copy of the array
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    addr[dev] = p;
  }
  *out = addr[dev];
  return 0;
}

namespace {

// The contiguous slot range of every x-run of bins within `radius_bins` of the point's bin, in
// visiting order (z outer, y, then x within the run), handed to f(start, end).
template <typename F>
__device__ __forceinline__ void for_rows(const QueryArgs& a, const InstanceDesc& in, float4 p,
                                         int radius_bins, F&& f) {
  const BinRanges<true> offb = bin_ranges<true>(in, a.off, a.rel);
  int bc[3];
  bin_coords_div(in, p.x, p.y, p.z, bc);  // calc_bin_coords_from_point, grid.hpp:222-230
  const int x_lo = max(bc[0] - radius_bins, 0);
  const int x_hi = min(bc[0] + radius_bins, in.ext[0] - 1);
  for (int dz = -radius_bins; dz <= radius_bins; ++dz) {
    const int z = bc[2] + dz;
    if (z < 0 || z >= in.ext[2]) continue;
    for (int dy = -radius_bins; dy <= radius_bins; ++dy) {
      const int y = bc[1] + dy;
      if (y < 0 || y >= in.ext[1]) continue;
      f(offb.at(x_lo, y, z), offb.at(x_hi + 1, y, z));
    }
  }
}

// distance_to_squared of the query point to a cell (vector.hpp:335-343)
__device__ __forceinline__ float dist2_to(float4 p, float4 c) {
  return len2(p.x - c.x, p.y - c.y, p.z - c.z);
}

__global__ __launch_bounds__(kBlock) void k_nearest(QueryArgs a) {
  const int q = (int)(blockIdx.x * kBlock + threadIdx.x);
  if (q >= a.n) return;
  const InstanceDesc& in = a.inst[0];
  const float4 p = make_float4(a.points[3 * q], a.points[3 * q + 1], a.points[3 * q + 2], 0.0f);
  int best = -1;
  float best_d2 = FLT_MAX;  // finite_max_v<float>
  for_rows(a, in, p, a.radius_bins, [&](int start, int end) {
    for (int m = start; m < end; ++m) {
      const float d2 = dist2_to(p, a.xs[m]);
      if (d2 < best_d2) {
        best_d2 = d2;
        best = a.id[m];
      }
    }
  });
  a.out[q] = best;
}

// The two random picks: count the candidates, then return candidate floor(u * total).
// BY_RADIUS: cells with d2 <= radius^2 (grid.hpp:507-564); else every cell of the bins
// (grid.hpp:453-495).
template <bool BY_RADIUS>
__global__ __launch_bounds__(kBlock) void k_random_pick(QueryArgs a) {
  const int q = (int)(blockIdx.x * kBlock + threadIdx.x);
  if (q >= a.n) return;
  const InstanceDesc& in = a.inst[0];
  const float4 p = make_float4(a.points[3 * q], a.points[3 * q + 1], a.points[3 * q + 2], 0.0f);
  const float r2 = a.radius * a.radius;
  int total = 0;
  for_rows(a, in, p, a.radius_bins, [&](int start, int end) {
    if (!BY_RADIUS) {
      total += end - start;
      return;
    }
    for (int m = start; m < end; ++m)
      if (dist2_to(p, a.xs[m]) <= r2) ++total;
  });
  int pick = -1;
  if (total > 0) {
    const int target = min((int)(a.uniforms[q] * (float)total), total - 1);
    int seen = 0;
    for_rows(a, in, p, a.radius_bins, [&](int start, int end) {
      if (pick >= 0) return;
      for (int m = start; m < end; ++m) {
        if (BY_RADIUS && !(dist2_to(p, a.xs[m]) <= r2)) continue;
        if (seen == target) {
          pick = a.id[m];
          return;
        }
        ++seen;
      }
    });
  }
  a.out[q] = pick;
}

}  // namespace

void launch_query(int kind, const QueryArgs& a, hipStream_t stream) {
  if (a.n <= 0) return;
  const dim3 g((a.n + kBlock - 1) / kBlock), b(kBlock);
  if (kind == kQueryNearest)
    hipLaunchKernelGGL(k_nearest, g, b, 0, stream, a);
  else if (kind == kQueryRandomInBins)
    hipLaunchKernelGGL(k_random_pick<false>, g, b, 0, stream, a);
  else
    hipLaunchKernelGGL(k_random_pick<true>, g, b, 0, stream, a);
}

}  // namespace tecs_amd
