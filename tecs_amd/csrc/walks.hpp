// walks.hpp - device walks over the binned grid shared by the thread and team sweeps: the
// Gabriel blocker test on the stale grid and the two roles of the symmetric finders.
#pragma once

#include "tecs_internal.hpp"

namespace tecs_amd {
namespace walks {

// Gabriel emptiness test of the ball with diameter c*|p_i - p_j| around the midpoint over the
// members of the (stale) grid bins overlapping its bounding box, binned_gabriel.hpp:133-209.
template <class BR>
__device__ __forceinline__ bool gabriel_blocked(const InstanceDesc& in, const float4* __restrict__ xs,
                                                const BR& offb, int s, int m,
                                                float4 me, const float* disp, float d2) {
  const float mid0 = me.x - disp[0] * 0.5f;
  const float mid1 = me.y - disp[1] * 0.5f;
  const float mid2 = me.z - disp[2] * 0.5f;
  const float rho2 = d2 * in.gabriel_factor;
  const float rho = sqrtf(rho2);
  const float mid[3] = {mid0, mid1, mid2};
  int mn[3], mx[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    mn[d] = clampi((int)floorf((mid[d] - rho - in.minb[d]) * in.inv_bin), in.ext[d] - 1);
    mx[d] = clampi((int)floorf((mid[d] + rho - in.minb[d]) * in.inv_bin), in.ext[d] - 1);
  }
  for (int b2 = mn[2]; b2 <= mx[2]; ++b2) {
    for (int b1 = mn[1]; b1 <= mx[1]; ++b1) {
      const int start = offb.at(mn[0], b1, b2);
      const int end = offb.at(mx[0] + 1, b1, b2);
      for (int k = start; k < end; ++k) {
        if (k == s || k == m) continue;
        const float4 pk = xs[k];
        if (len2(pk.x - mid0, pk.y - mid1, pk.z - mid2) < rho2) return true;
      }
    }
  }
  return false;
}

// ------------------------------------------------------------------ symmetric finders
// The reference evaluates each unordered pair once, from the cell with the larger agent index
// i, and applies the result v to both cells with atomics: delta(i) += v, delta(j) -= v
// (binned_all_pairs_symmetric.hpp:176-200). Scattered float atomic adds on MI355X sustain
// about 2.5e10 per second (measured with that scheme here: 2.3 ms per 4M-cell sweep, 7x the
// ordered sweep), far less than evaluating a pair a second time costs. So the engine gathers:
// every cell walks its own neighbourhood and, for each accepted candidate m,
//   own:    id(m) < id(c): c adds v(c, m);
//   gather: id(m) > id(c): c adds -v(m, c), with v computed exactly as at m (m as self, m's
//           Gabriel ball) - the value m's walk applies to c in the reference.
// This is exact for pairs of stayers, cells whose current bin is their bin at the last rebuild
// (on a fresh grid, every cell): m's walk reaches c's slot iff their bins are within the
// search radius, which for stayers is the same as c's walk reaching m's slot. A mover's pairs
// need one of two rules (a sweep applies one rule to every cell):
//   two walks (thread sweep): a mover k adds v(k, j) for its own pairs and scatters -v into
//     j's second buffer q with a float atomic; a cell c never gathers from a mover, and a
//     mover c gathers from the stayers k > c in a second walk around its rebuild bin (k's walk
//     reaches c's slot iff k's bin is within the search radius of c's rebuild bin);
//   scatter (team sweep): every pair with a mover is delivered by its larger-index cell, which
//     scatters -v to the other cell's q; the smaller-index cell skips it. One walk per cell.
// The integrator update (or k_fold_q) adds q and clears it. The contributions of a cell are
// summed in its walk order, the scattered ones in atomic order; the reference's atomics have no
// fixed order either.
enum SymRole { kRoleNone = 0, kRoleOwn = 1, kRoleGather = 2, kRoleOwnScatter = 3 };
// which roles a walk collects
enum SymWalk { kWalkBoth = 3, kWalkOwn = 1, kWalkGather = 2 };

// Slot k sits outside its rebuild bin (skey null: the grid was built from these positions).
// mover8: the predictor's flags of these positions, when it wrote them.
__device__ __forceinline__ bool is_mover(const InstanceDesc& in, const int* skey,
                                         const unsigned char* mover8, int k, float4 p) {
  if (skey == nullptr) return false;
  return mover8 ? mover8[k] != 0 : bin_key(in, p) != skey[k];
}

// The role candidate m plays in a walk of cell s (skey: the rebuild keys on a stale grid, else
// null; me_mover: s itself is a mover; SCATTER: the scatter rule, else the two-walk rule). The
// distance test is shared: d2 of (m, s) is bitwise d2 of (s, m). m_mover: whether m is a mover
// when the caller loaded its flag already (0 or 1; -1: looked up here).
template <bool GABRIEL, bool SCATTER, class BR>
__device__ __forceinline__ int sym_role(const InstanceDesc& in, const int* skey,
                                        const unsigned char* mover8,
                                        const float4* __restrict__ xs,
                                        const BR& offb, int s, int me_id,
                                        bool me_mover, float4 me, int m, float4 pm, int m_id,
                                        int walk = kWalkBoth, int m_mover = -1) {
  const float disp[3] = {me.x - pm.x, me.y - pm.y, me.z - pm.z};
  const float d2 = len2(disp[0], disp[1], disp[2]);
  if (d2 >= in.cutoff2 || m == s) return kRoleNone;
  if (m_id < me_id) {
    if (!(walk & kWalkOwn)) return kRoleNone;
    if (GABRIEL && gabriel_blocked(in, xs, offb, s, m, me, disp, d2)) return kRoleNone;
    if (me_mover || (SCATTER && (m_mover >= 0 ? m_mover != 0 : is_mover(in, skey, mover8, m, pm))))
      return kRoleOwnScatter;
    return kRoleOwn;
  }
  if (!(walk & kWalkGather) || (SCATTER && me_mover) ||
      (m_mover >= 0 ? m_mover != 0 : is_mover(in, skey, mover8, m, pm)))
    return kRoleNone;
  if (GABRIEL) {
    const float back[3] = {pm.x - me.x, pm.y - me.y, pm.z - me.z};
    if (gabriel_blocked(in, xs, offb, m, s, pm, back, d2)) return kRoleNone;
  }
  return kRoleGather;
}

// The functor on a zeroed accumulator pack for the pair of the role, m's position pm already
// loaded: own v(s, m) added (and -v scattered to m's q), gather -v(m, s) added. A gathered pair
// is m's own evaluation with the two cells swapped, so both roles run one evaluation on
// operands selected by the role (a wavefront whose lanes hold both roles does not run the
// force twice) and the gathered result is negated, exactly, before it is added.
template <bool HAS_TYPE, bool HAS_POL, bool MAY_SCATTER = true>
__device__ __forceinline__ void sym_eval_at(const SweepArgs& a, const InstanceDesc& in, float4 me,
                                            const Pol& pme, int m, float4 pm, int role,
                                            float* acc_pos, float* acc_pol) {
  Pol po{};
  if (HAS_POL) po = make_pol(a.ps[m], a.trig[m]);
  const bool g = role == kRoleGather;
  const float4 pa = g ? pm : me, pb = g ? me : pm;
  Pol qa{}, qb{};
  if (HAS_POL) {
    qa = g ? po : pme;
    qb = g ? pme : po;
  }
  float vp[3] = {0.0f, 0.0f, 0.0f}, vl[2] = {0.0f, 0.0f};
  const float disp[3] = {pa.x - pb.x, pa.y - pb.y, pa.z - pb.z};
  const float d = sqrtf(len2(disp[0], disp[1], disp[2]));
  apply_force<HAS_TYPE, HAS_POL>(in.force, disp, d, __float_as_int(pa.w), __float_as_int(pb.w),
                                 qa, qb, vp, vl);
#pragma unroll
  for (int c = 0; c < 3; ++c) acc_pos[c] += g ? -vp[c] : vp[c];
  acc_pol[0] += g ? -vl[0] : vl[0];
  acc_pol[1] += g ? -vl[1] : vl[1];
  if (MAY_SCATTER && role == kRoleOwnScatter) {
    float* q = &a.q_pos[m].x;
    unsafeAtomicAdd(q + 0, -vp[0]);
    unsafeAtomicAdd(q + 1, -vp[1]);
    unsafeAtomicAdd(q + 2, -vp[2]);
    if (HAS_POL) {
      float* r = &a.q_pol[m].x;
      unsafeAtomicAdd(r + 0, -vl[0]);
      unsafeAtomicAdd(r + 1, -vl[1]);
    }
  }
}

template <bool HAS_TYPE, bool HAS_POL, bool MAY_SCATTER = true>
__device__ __forceinline__ void sym_eval(const SweepArgs& a, const InstanceDesc& in, float4 me,
                                         const Pol& pme, int m, int role, float* acc_pos,
                                         float* acc_pol) {
  sym_eval_at<HAS_TYPE, HAS_POL, MAY_SCATTER>(a, in, me, pme, m, a.xs[m], role, acc_pos, acc_pol);
}

// A serial walk of cell s over the bins within the search radius of bin bc (z, y, x order),
// evaluating the roles `walk` collects.
template <bool GABRIEL, bool HAS_TYPE, bool HAS_POL, bool SCATTER, bool ROWS>
__device__ void sym_walk_serial(const SweepArgs& a, const InstanceDesc& in, int s, int me_id,
                                bool me_mover, float4 me, const Pol& pme, const int* bc, int walk,
                                float* acc_pos, float* acc_pol, unsigned& pairs) {
  const float4* __restrict__ xs = a.xs;
  const BinRanges<ROWS> offb = bin_ranges<ROWS>(in, a.off, a.rel);
  const int R = in.search_radius;
  const int x_lo = max(bc[0] - R, 0);
  const int x_hi = min(bc[0] + R, in.ext[0] - 1);
  for (int dz = -R; dz <= R; ++dz) {
    const int z = bc[2] + dz;
    if (z < 0 || z >= in.ext[2]) continue;
    for (int dy = -R; dy <= R; ++dy) {
      const int y = bc[1] + dy;
      if (y < 0 || y >= in.ext[1]) continue;
      const int start = offb.at(x_lo, y, z);
      const int end = offb.at(x_hi + 1, y, z);
      for (int m = start; m < end; ++m) {
        const float4 pm = xs[m];
        const int role = sym_role<GABRIEL, SCATTER>(in, a.skey, a.mover8, xs, offb, s, me_id, me_mover, me,
                                                    m, pm, a.id[m], walk);
        if (role == kRoleNone) continue;
        sym_eval_at<HAS_TYPE, HAS_POL>(a, in, me, pme, m, pm, role, acc_pos, acc_pol);
        pairs += role != kRoleGather;
      }
    }
  }
}

// One cell of a symmetric sweep walked serially under the sweep's rule: one walk around its
// bin bc, or, for a mover under the two-walk rule, the own walk around bc and the gather walk
// around its rebuild bin.
template <bool GABRIEL, bool HAS_TYPE, bool HAS_POL, bool SCATTER, bool ROWS>
__device__ void sym_cell_serial(const SweepArgs& a, const InstanceDesc& in, int s, int me_id,
                                bool me_mover, float4 me, const Pol& pme, const int* bc,
                                float* acc_pos, float* acc_pol, unsigned& pairs) {
  if (SCATTER || !me_mover) {
    sym_walk_serial<GABRIEL, HAS_TYPE, HAS_POL, SCATTER, ROWS>(a, in, s, me_id, me_mover, me, pme, bc,
                                                         kWalkBoth, acc_pos, acc_pol, pairs);
    return;
  }
  sym_walk_serial<GABRIEL, HAS_TYPE, HAS_POL, SCATTER, ROWS>(a, in, s, me_id, true, me, pme, bc,
                                                       kWalkOwn, acc_pos, acc_pol, pairs);
  const int key = a.skey[s];
  const int sc[3] = {key % in.ext[0], (key / in.ext[0]) % in.ext[1], key / (in.ext[0] * in.ext[1])};
  sym_walk_serial<GABRIEL, HAS_TYPE, HAS_POL, SCATTER, ROWS>(a, in, s, me_id, true, me, pme, sc,
                                                       kWalkGather, acc_pos, acc_pol, pairs);
}

// The movers of a wavefront on a stale grid, under the two-walk rule (the own walk around the
// cell's bin, the gather walk around its rebuild bin), one mover after another over all 64
// lanes instead of each serially in its own lane while the rest of the wavefront waits: a
// walk's candidates (the 9 rows around the bin) are spread over the lanes, each lane adds the
// contributions of its candidates (scattering as the serial walk does), and a butterfly over
// the wavefront sums them into the mover's lane. The symmetric finders are held to the FP64
// bar, not to an order of summation, and the butterfly's order is fixed, so the result is
// deterministic. Requires search_radius == 1; every lane of the wavefront calls it (the
// lanes without a cell as well). Returns the pairs this lane counted.
template <bool GABRIEL, bool HAS_TYPE, bool HAS_POL, bool ROWS>
__device__ unsigned sym_movers_wave(const SweepArgs& a, const InstanceDesc& in, int s, bool mover,
                                    float* acc_pos, float* acc_pol) {
  const int lane = (int)(threadIdx.x & 63);
  const float4* __restrict__ xs = a.xs;
  const BinRanges<ROWS> offb = bin_ranges<ROWS>(in, a.off, a.rel);
  unsigned pairs = 0;
  unsigned long long mv = __ballot(mover);
  while (mv != 0ull) {
    const int src = __ffsll((long long)mv) - 1;
    mv &= mv - 1ull;
    const int ms = s - lane + src;  // the mover's slot
    const float4 mme = xs[ms];
    Pol mpol{};
    if (HAS_POL) mpol = make_pol(a.ps[ms], a.trig[ms]);
    const int mid = a.id[ms];
    int bc[3];
    bin_coords_div(in, mme.x, mme.y, mme.z, bc);
    const int key = a.skey[ms];
    const int sc[3] = {key % in.ext[0], (key / in.ext[0]) % in.ext[1], key / (in.ext[0] * in.ext[1])};
    float p[3] = {0.0f, 0.0f, 0.0f}, l[2] = {0.0f, 0.0f};
#pragma unroll
    for (int w = 0; w < 2; ++w) {
      const int cx = w == 0 ? bc[0] : sc[0], cy = w == 0 ? bc[1] : sc[1], cz = w == 0 ? bc[2] : sc[2];
      const int walk = w == 0 ? kWalkOwn : kWalkGather;
      const int x_lo = max(cx - 1, 0), x_hi = min(cx + 1, in.ext[0] - 1);
      int rs[9], pre[10];
      pre[0] = 0;
#pragma unroll
      for (int r = 0; r < 9; ++r) {
        const int z = cz + r / 3 - 1, y = cy + r % 3 - 1;
        const bool ok = z >= 0 && z < in.ext[2] && y >= 0 && y < in.ext[1];
        rs[r] = ok ? offb.at(x_lo, y, z) : 0;
        pre[r + 1] = pre[r] + (ok ? offb.at(x_hi + 1, y, z) - rs[r] : 0);
      }
      for (int k = lane; k < pre[9]; k += 64) {
        // candidate k of the walk: row r holds candidates pre[r] .. pre[r + 1] - 1
        int m = 0;
#pragma unroll
        for (int r = 0; r < 9; ++r)
          if (k >= pre[r] && k < pre[r + 1]) m = rs[r] + (k - pre[r]);
        const float4 pm = xs[m];
        const int role =
            sym_role<GABRIEL, false>(in, a.skey, a.mover8, xs, offb, ms, mid, true, mme, m, pm, a.id[m], walk);
        if (role == kRoleNone) continue;
        sym_eval_at<HAS_TYPE, HAS_POL>(a, in, mme, mpol, m, pm, role, p, l);
        pairs += role != kRoleGather;
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
      for (int c = 0; c < 3; ++c) p[c] = p[c] + __shfl_xor(p[c], o, 64);
      if (HAS_POL) {
        l[0] = l[0] + __shfl_xor(l[0], o, 64);
        l[1] = l[1] + __shfl_xor(l[1], o, 64);
      }
    }
    if (lane == src) {
#pragma unroll
      for (int c = 0; c < 3; ++c) acc_pos[c] += p[c];
      acc_pol[0] += l[0];
      acc_pol[1] += l[1];
    }
  }
  return pairs;
}

}  // namespace walks
}  // namespace tecs_amd
