// kernels_delaunay.hip - NaiveDelaunay, 3-D (naive_delaunay.hpp:157-293).
//
// The reference runs one work item per (i, j, k) - i a cell, j another cell, k a cell other
// than i and j - and, when d(i, j) < r, looks for the first l > k (l not i, j) whose sphere
// through i, j, k, l holds no other cell; an item that finds one calls the functor for (i, j)
// on a zeroed accumulator pack and adds the pack to out(i) with atomics. An edge (i, j) is
// therefore evaluated once per qualifying k, with the same value each time.
//
// Two passes per evaluation:
//   k_dl_count: the items. Workgroup (instance, i, split) stages the instance's positions in
//     LDS (at most kDelaunayMaxCells), lists the partners j within the cutoff in index order,
//     and takes every dl_splits-th slice of kBlock items (j, k) of that list, one item per
//     lane, consecutive lanes on consecutive k (similar l ranges). A lane runs the l and m loops
//     serially over LDS, in the reference's FP32 operation order, and adds 1 to the (i, j)
//     counter when the item finds an l. Counters are integers, so the result does not depend
//     on which lane took which item.
//   k_dl_cells: one thread per cell. For each partner j in index order with count c > 0 it
//     evaluates the functor once on a zeroed pack and adds that pack c times (the items'
//     adds, in (j, k) order; the reference's atomics have no fixed order), clears the counter,
//     and writes the delta (Delta), the pair count (Count) or j c times (Emit).
#include "tecs_internal.hpp"

namespace tecs_amd {
namespace {

constexpr float kFltEps = 1.1920928955078125e-07f;  // FLT_EPSILON

// Whether item (i, j, k) finds an l (naive_delaunay.hpp:187-257): P holds the n positions.
__device__ __forceinline__ bool dl_item_found(const float4* P, int n, int i, int j, int k) {
  const float4 pi = P[i], pj = P[j], pk = P[k];
  const float e1[3] = {pj.x - pi.x, pj.y - pi.y, pj.z - pi.z};
  const float e2[3] = {pk.x - pi.x, pk.y - pi.y, pk.z - pi.z};
  const float li = len2(pi.x, pi.y, pi.z);
  const float b1 = len2(pj.x, pj.y, pj.z) - li;
  const float b2 = len2(pk.x, pk.y, pk.z) - li;
  for (int l = k + 1; l < n; ++l) {
    if (l == i || l == j) continue;
    const float4 pl = P[l];
    const float e3[3] = {pl.x - pi.x, pl.y - pi.y, pl.z - pi.z};
    // the three minors of the edge matrix that the denominator and the b1 terms share
    const float m0 = e2[1] * e3[2] - e2[2] * e3[1];
    const float m1 = e2[0] * e3[2] - e2[2] * e3[0];
    const float m2 = e2[0] * e3[1] - e2[1] * e3[0];
    const float den = 2.0f * ((e1[0] * m0 - e1[1] * m1) + e1[2] * m2);
    if (fabsf(den) <= kFltEps) continue;
    const float b3 = len2(pl.x, pl.y, pl.z) - li;
    const float c0 = pi.x + ((b1 * m0 - e1[1] * (b2 * e3[2] - e2[2] * b3)) +
                             e1[2] * (b2 * e3[1] - e2[1] * b3)) / den;
    const float c1 = pi.y + ((e1[0] * (b2 * e3[2] - e2[2] * b3) - b1 * m1) +
                             e1[2] * (e2[0] * b3 - b2 * e3[0])) / den;
    const float c2 = pi.z + ((e1[0] * (e2[1] * b3 - b2 * e3[1]) - e1[1] * (e2[0] * b3 - b2 * e3[0])) +
                             b1 * m2) / den;
    const float rad2 = len2(pi.x - c0, pi.y - c1, pi.z - c2);
    const float rad = sqrtf(rad2);
    bool valid = true;
    for (int m = 0; m < n; ++m) {
      if (m == i || m == j || m == k || m == l) continue;
      const float4 pm = P[m];
      const float dx = pm.x - c0, dy = pm.y - c1, dz = pm.z - c2;
      if (fabsf(dx) > rad || fabsf(dy) > rad || fabsf(dz) > rad) continue;
      if (len2(dx, dy, dz) <= rad2) {
        valid = false;
        break;
      }
    }
    if (valid) return true;
  }
  return false;
}

__global__ __launch_bounds__(kBlock) void k_dl_count(SweepArgs a) {
  __shared__ float4 P[kDelaunayMaxCells];
  __shared__ int J[kDelaunayMaxCells];
  __shared__ int wave_n[kBlock / 64];
  const int S = a.dl_splits, R = a.dl_row;
  const int b = (int)blockIdx.x;
  const int w = b % S;
  const int i = (b / S) % R;
  const int ii = b / (S * R);
  const InstanceDesc& in = a.inst[ii];
  const int n = in.n;
  if (i >= n) return;  // the whole workgroup
  const float4* xs = a.xs + in.slot_begin;
  for (int t = (int)threadIdx.x; t < n; t += kBlock) P[t] = xs[t];
  __syncthreads();
  // partners within the cutoff (naive_delaunay.hpp:177-182), compacted in index order: the
  // workgroups of one cell split the items by position in this list, so it must be the same
  // list in each of them
  const float4 pi = P[i];
  const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
  int nj = 0;
  for (int base = 0; base + 1 < n; base += kBlock) {
    const int jj = base + (int)threadIdx.x;
    const int j = jj + (jj >= i);
    bool in_range = false;
    if (jj + 1 < n) {
      const float4 pj = P[j];
      in_range = len2(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z) < in.cutoff2;
    }
    const unsigned long long mask = __ballot(in_range);
    if (lane == 0) wave_n[wave] = __popcll(mask);
    __syncthreads();
    int at = nj;
    for (int q = 0; q < wave; ++q) at += wave_n[q];
    at += __popcll(mask & ((1ull << lane) - 1ull));
    if (in_range) J[at] = j;
    for (int q = 0; q < kBlock / 64; ++q) nj += wave_n[q];
    __syncthreads();
  }
  const int items = nj * n;
  uint32_t* row = a.dl_cnt + ((size_t)ii * R + i) * R;
  for (int t = w * kBlock + (int)threadIdx.x; t < items; t += S * kBlock) {
    const int j = J[t / n];
    // k skips i and j: shifted past the smaller, then past the larger (naive_delaunay.hpp:
    // 165-175); a k that lands on n has no l
    int k = t % n;
    k += k >= min(i, j);
    k += k >= max(i, j);
    if (k >= n) continue;
    if (dl_item_found(P, n, i, j, k)) atomicAdd(row + j, 1u);
  }
}

template <bool HAS_TYPE, bool HAS_POL, int MODE>
__global__ __launch_bounds__(kBlock) void k_dl_cells(SweepArgs a) {
  __shared__ unsigned lds_u[4];
  const int lb = logical_block();
  const int ii = a.block_inst[lb];
  const InstanceDesc& in = a.inst[ii];
  const int s = lb * kBlock + (int)threadIdx.x;
  const int i = s - in.slot_begin;
  const bool active = i < in.n;
  unsigned pairs = 0;
  float acc_pos[3] = {0.0f, 0.0f, 0.0f}, acc_pol[2] = {0.0f, 0.0f};
  if (active) {
    uint32_t* row = a.dl_cnt + ((size_t)ii * a.dl_row + i) * a.dl_row;
    const float4 me = a.xs[s];
    Pol pme{};
    if (HAS_POL) pme = make_pol(a.ps[s], a.trig[s]);
    int64_t at = MODE == kModeEmit ? a.row_ptr[a.id[s]] : 0;
    for (int j = 0; j < in.n; ++j) {
      const unsigned c = row[j];
      if (c == 0) continue;
      row[j] = 0u;
      pairs += c;
      const int m = in.slot_begin + j;
      if (MODE == kModeEmit) {
        const int32_t jid = a.id[m];
        for (unsigned r = 0; r < c; ++r) a.cols[at++] = jid;
      }
      if (MODE != kModeDelta) continue;
      const float4 pj = a.xs[m];
      const float disp[3] = {me.x - pj.x, me.y - pj.y, me.z - pj.z};
      const float d = sqrtf(len2(disp[0], disp[1], disp[2]));
      Pol po{};
      if (HAS_POL) po = make_pol(a.ps[m], a.trig[m]);
      float vp[3] = {0.0f, 0.0f, 0.0f}, vl[2] = {0.0f, 0.0f};
      apply_force<HAS_TYPE, HAS_POL>(in.force, disp, d, __float_as_int(me.w), __float_as_int(pj.w),
                                     pme, po, vp, vl);
      for (unsigned r = 0; r < c; ++r) {
#pragma unroll
        for (int q = 0; q < 3; ++q) acc_pos[q] += vp[q];
        acc_pol[0] += vl[0];
        acc_pol[1] += vl[1];
      }
    }
  }
  if (MODE == kModeCount) {
    if (active) a.row_count[a.id[s]] = (int)pairs;
    return;
  }
  if (MODE == kModeEmit) return;
  if (active) {
    // out(i) holds the items' packs; the drag pass adds total_velocity / fmax(total_drag, eps),
    // which is +0 because no force term writes drag (naive_delaunay.hpp:282-289)
    a.o_pos[s] = make_float4((0.0f + acc_pos[0]) + 0.0f, (0.0f + acc_pos[1]) + 0.0f,
                             (0.0f + acc_pos[2]) + 0.0f, 0.0f);
    if (HAS_POL) a.o_pol[s] = make_float2(0.0f + acc_pol[0], 0.0f + acc_pol[1]);
  }
  const unsigned tp = block_sum_u(active ? pairs : 0u, lds_u);
  if (threadIdx.x == 0) a.block_pairs[lb] += tp;
}

template <int MODE>
void launch_cells(bool has_type, bool has_pol, const SweepArgs& a, int n_blocks, hipStream_t stream) {
  const dim3 g(n_blocks), b(kBlock);
  if (MODE != kModeDelta)
    hipLaunchKernelGGL((k_dl_cells<false, false, MODE>), g, b, 0, stream, a);
  else if (has_type && has_pol)
    hipLaunchKernelGGL((k_dl_cells<true, true, MODE>), g, b, 0, stream, a);
  else if (has_type)
    hipLaunchKernelGGL((k_dl_cells<true, false, MODE>), g, b, 0, stream, a);
  else if (has_pol)
    hipLaunchKernelGGL((k_dl_cells<false, true, MODE>), g, b, 0, stream, a);
  else
    hipLaunchKernelGGL((k_dl_cells<false, false, MODE>), g, b, 0, stream, a);
}

}  // namespace

int launch_delaunay(bool has_type, bool has_pol, int mode, const SweepArgs& a, int n_blocks,
                    hipStream_t stream) {
  if (mode != kModeDelta && mode != kModeCount && mode != kModeEmit) return -1;
  const int64_t groups = (int64_t)a.n_inst * a.dl_row * a.dl_splits;
  if (groups > 0)
    hipLaunchKernelGGL(k_dl_count, dim3((unsigned)groups), dim3(kBlock), 0, stream, a);
  if (mode == kModeDelta) launch_cells<kModeDelta>(has_type, has_pol, a, n_blocks, stream);
  else if (mode == kModeCount) launch_cells<kModeCount>(has_type, has_pol, a, n_blocks, stream);
  else launch_cells<kModeEmit>(has_type, has_pol, a, n_blocks, stream);
  return 0;
}

}  // namespace tecs_amd
