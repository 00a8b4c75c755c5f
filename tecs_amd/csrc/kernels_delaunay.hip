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
//     lane, partner fastest (the lanes of a wavefront share a few consecutive k, so their l
//     ranges nearly coincide). A lane runs the l and m loops
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

// The cells of an instance as one workgroup stages them for cell i: sorted by squared
// distance to p_i (ties by index), each with its index in .w, and each cell's place in that
// order.
struct NearOrder {
  const float4* pos;  // pos[r]: (x, y, z, index as int bits) of the r-th nearest cell, r < n,
                      // readable up to r = n + 2
  const int* rank;    // rank[m]: place of cell m
};

// Whether item (i, j, k) finds an l (naive_delaunay.hpp:187-257). The item's result is
// "some l > k (l not i, j) passes the denominator test and has no cell m (not i, j, k, l) that
// passes the box prefilter and lies in or on its sphere", which does not depend on the order
// in which l and m are visited; each per-(l, m) test is evaluated with the reference's FP32
// operations. Both loops visit the cells nearest to p_i first: the sphere passes through p_i,
// so a cell near p_i is the likeliest to lie inside and end the m loop early, and spheres
// through near cells are the likeliest to be empty. The m loop is one predicate per cell and
// one exit, so a wavefront's lanes stay converged until they leave.
__device__ __forceinline__ bool dl_item_found(const NearOrder& o, int n, int i, int j, int k) {
  const float4 pi = o.pos[o.rank[i]], pj = o.pos[o.rank[j]], pk = o.pos[o.rank[k]];
  const float e1[3] = {pj.x - pi.x, pj.y - pi.y, pj.z - pi.z};
  const float e2[3] = {pk.x - pi.x, pk.y - pi.y, pk.z - pi.z};
  const float li = len2(pi.x, pi.y, pi.z);
  const float b1 = len2(pj.x, pj.y, pj.z) - li;
  const float b2 = len2(pk.x, pk.y, pk.z) - li;
  for (int rl = 0; rl < n; ++rl) {
    const float4 pl = o.pos[rl];
    const int l = __float_as_int(pl.w);
    if (l <= k || l == i || l == j) continue;
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
    // two cells per iteration (the entry after the last one is padding, masked out), the
    // next pair loaded while this one is tested
    auto hit = [&](float4 pm) {
      const int m = __float_as_int(pm.w);
      const float dx = pm.x - c0, dy = pm.y - c1, dz = pm.z - c2;
      const bool in_box = !(fabsf(dx) > rad || fabsf(dy) > rad || fabsf(dz) > rad);
      const bool own = (m == i) | (m == j) | (m == k) | (m == l);
      return !own & in_box & (len2(dx, dy, dz) <= rad2);
    };
    bool inside = false;
    float4 qa = o.pos[0], qb = o.pos[1];
    for (int rm = 0; rm < n && !inside; rm += 2) {
      const float4 ca = qa, cb = qb;
      qa = o.pos[rm + 2];
      qb = o.pos[rm + 3];
      inside = hit(ca) | ((rm + 1 < n) & hit(cb));
    }
    if (!inside) return true;
  }
  return false;
}

__global__ __launch_bounds__(kBlock) void k_dl_count(SweepArgs a) {
  // positions and indices, nearest to p_i first (+ 3 entries of padding read ahead by the m
  // loop and never used)
  __shared__ float4 P[kDelaunayMaxCells + 3];
  __shared__ int Rk[kDelaunayMaxCells];    // place of each cell in that order
  __shared__ float D[kDelaunayMaxCells];   // squared distance of each cell to p_i
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
  const float4 pi = xs[i];
  // d^2 of (i, m) as the reference's cutoff test forms it (naive_delaunay.hpp:177-182)
  for (int m = (int)threadIdx.x; m < n; m += kBlock) {
    const float4 pm = xs[m];
    D[m] = len2(pi.x - pm.x, pi.y - pm.y, pi.z - pm.z);
  }
  __syncthreads();
  for (int m = (int)threadIdx.x; m < n; m += kBlock) {
    const float dm = D[m];
    int r = 0;
    for (int q = 0; q < n; ++q) {
      const float dq = D[q];
      r += (dq < dm) || (dq == dm && q < m);
    }
    Rk[m] = r;
    const float4 pm = xs[m];
    P[r] = make_float4(pm.x, pm.y, pm.z, __int_as_float(m));
  }
  // partners within the cutoff, compacted in index order: the workgroups of one cell split
  // the items by position in this list, so it must be the same list in each of them
  const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
  int nj = 0;
  for (int base = 0; base + 1 < n; base += kBlock) {
    const int jj = base + (int)threadIdx.x;
    const int j = jj + (jj >= i);
    const bool in_range = jj + 1 < n && D[j] < in.cutoff2;
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
  __syncthreads();
  const NearOrder o{P, Rk};
  const int items = nj * n;
  uint32_t* row = a.dl_cnt + ((size_t)ii * R + i) * R;
  for (int t = w * kBlock + (int)threadIdx.x; t < items; t += S * kBlock) {
    // partner fastest: the lanes of a wavefront share a few consecutive k, so their l ranges
    // (l > k) nearly coincide
    const int j = J[t % nj];
    // k skips i and j: shifted past the smaller, then past the larger (naive_delaunay.hpp:
    // 165-175); a k that lands on n has no l
    int k = t / nj;
    k += k >= min(i, j);
    k += k >= max(i, j);
    if (k >= n) continue;
    if (dl_item_found(o, n, i, j, k)) atomicAdd(row + j, 1u);
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
  constexpr bool kForce = MODE == kModeDelta || MODE == kModeDeltaDrag;
  unsigned pairs = 0;
  float acc_pos[3] = {0.0f, 0.0f, 0.0f}, acc_pol[2] = {0.0f, 0.0f};
  float tdrag = 0.0f, tvel[3] = {0.0f, 0.0f, 0.0f};
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
      if (!kForce) continue;
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
      if (MODE == kModeDeltaDrag) {
        // naive_delaunay.hpp:256-276: each item adds its drag and drag * old_vel(j)
        const float g = pair_drag<HAS_TYPE>(in.force, __float_as_int(me.w), __float_as_int(pj.w));
        const float4 ov = a.ov_in[m];
        for (unsigned r = 0; r < c; ++r) {
          tdrag = tdrag + g;
          tvel[0] = tvel[0] + g * ov.x;
          tvel[1] = tvel[1] + g * ov.y;
          tvel[2] = tvel[2] + g * ov.z;
        }
      }
    }
  }
  if (MODE == kModeCount) {
    if (active) a.row_count[a.id[s]] = (int)pairs;
    return;
  }
  if (MODE == kModeEmit) return;
  if (active) {
    // out(i) holds the items' packs; the drag pass adds total_velocity / fmax(total_drag, eps)
    // (naive_delaunay.hpp:282-289), +0 when no pair has drag
    const float den = fmaxf(tdrag, 1.1920928955078125e-07f);
    const float e0 = MODE == kModeDeltaDrag ? tvel[0] / den : 0.0f;
    const float e1 = MODE == kModeDeltaDrag ? tvel[1] / den : 0.0f;
    const float e2 = MODE == kModeDeltaDrag ? tvel[2] / den : 0.0f;
    a.o_pos[s] = make_float4((0.0f + acc_pos[0]) + e0, (0.0f + acc_pos[1]) + e1,
                             (0.0f + acc_pos[2]) + e2, 0.0f);
    if (HAS_POL) a.o_pol[s] = make_float2(0.0f + acc_pol[0], 0.0f + acc_pol[1]);
  }
  const unsigned tp = block_sum_u(active ? pairs : 0u, lds_u);
  if (threadIdx.x == 0) a.block_pairs[lb] += tp;
}

template <int MODE>
void launch_cells(bool has_type, bool has_pol, const SweepArgs& a, int n_blocks, hipStream_t stream) {
  const dim3 g(n_blocks), b(kBlock);
  if (MODE != kModeDelta && MODE != kModeDeltaDrag)
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
  if (mode != kModeDelta && mode != kModeDeltaDrag && mode != kModeCount && mode != kModeEmit) return -1;
  const int64_t groups = (int64_t)a.n_inst * a.dl_row * a.dl_splits;
  if (groups > 0)
    hipLaunchKernelGGL(k_dl_count, dim3((unsigned)groups), dim3(kBlock), 0, stream, a);
  if (mode == kModeDelta) launch_cells<kModeDelta>(has_type, has_pol, a, n_blocks, stream);
  else if (mode == kModeDeltaDrag) launch_cells<kModeDeltaDrag>(has_type, has_pol, a, n_blocks, stream);
  else if (mode == kModeCount) launch_cells<kModeCount>(has_type, has_pol, a, n_blocks, stream);
  else launch_cells<kModeEmit>(has_type, has_pol, a, n_blocks, stream);
  return 0;
}

}  // namespace tecs_amd
