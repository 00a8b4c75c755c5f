// kernels_sweep.hip - pair-finder sweeps with the fused force and integrator updates.
//
// One thread per slot (cell i). The binned sweep walks the neighbour bins of i in the order of
// BinnedAllPairs / BinnedGabriel (binned_all_pairs.hpp:116-171, binned_gabriel.hpp:108-227):
// z outer, y middle, x inner. Because slots are sorted by (bin, agent index), the x-run of
// neighbour bins in one (y, z) row is one contiguous slot range, visited front to back, which
// is exactly the reference's sequence of (bin, member) visits. Neighbouring threads of a
// wavefront hold neighbouring cells, so their neighbour ranges overlap and are served from
// L1/L2. The per-cell sum is accumulated in that visiting order, so it is bitwise the serial
// reference sum.
//
// Stale-grid semantics (binned_all_pairs.hpp:77-84, heun.hpp:206): the Heun corrector sweep
// reads predicted positions while the slot order and bin offsets stay those of the last
// rebuild; the centre bin of i is recomputed from its current position, exactly like the
// reference.
//
// Symmetric finders (binned_all_pairs_symmetric.hpp:148-219 and its three siblings): the pair
// export walks keep only partners with a smaller agent index (SYM); the force evaluation runs
// in the k_sym_* kernels below.
#include <type_traits>

#include "tecs_internal.hpp"
#include "walks.hpp"

namespace tecs_amd {
namespace {

using namespace walks;

// Per-pair visit context; `Acc` carries the force accumulators of cell i.
struct CellState {
  float acc_pos[3];
  float acc_pol[2];
  unsigned pairs;
  int64_t emit_at;
  float drag;    // DeltaDrag: sum of the pairs' drag (binned_all_pairs.hpp:113, 168)
  float vel[3];  //            sum of drag * old_velocity_j (:114, 169)
};

// `tab` is the instance's adhesion table (the InstanceDesc's, or a copy in LDS).
template <bool HAS_TYPE, bool HAS_POL, int MODE, bool SYM>
__device__ __forceinline__ void visit(const SweepArgs& a, const InstanceDesc& in, CellState& cs,
                                      float4 me, const Pol& pme, int m, float4 pj,
                                      const float* disp, float d2, const float* tab) {
  cs.pairs++;
  if (MODE == kModeCount) return;
  if (MODE == kModeEmit) {
    a.cols[cs.emit_at++] = a.id[m];
    return;
  }
  const float d = sqrtf(d2);
  Pol po{};
  if (HAS_POL) po = make_pol(a.ps[m], a.trig[m]);
  static_assert(!SYM || MODE == kModeCount || MODE == kModeEmit, "symmetric walks export pairs only");
  apply_force_k<HAS_POL>(in.force, disp, d,
                         pair_kadh<HAS_TYPE>(in.force, tab, __float_as_int(me.w), __float_as_int(pj.w)),
                         pme, po, cs.acc_pos, cs.acc_pol);
  if (MODE == kModeDeltaDrag) {
    // binned_all_pairs.hpp:168-169: total_drag += drag; total_velocity += drag * old_vel(j)
    const float g = pair_drag<HAS_TYPE>(in.force, __float_as_int(me.w), __float_as_int(pj.w));
    const float4 ov = a.ov_in[m];
    cs.drag = cs.drag + g;
    cs.vel[0] = cs.vel[0] + g * ov.x;
    cs.vel[1] = cs.vel[1] + g * ov.y;
    cs.vel[2] = cs.vel[2] + g * ov.z;
  }
}

template <bool GABRIEL, bool HAS_TYPE, bool HAS_POL, int MODE, bool SYM, bool ROWS>
__device__ __forceinline__ void walk_binned(const SweepArgs& a, const InstanceDesc& in, int s,
                                            int me_id, float4 me, const Pol& pme, CellState& cs) {
  const float4* __restrict__ xs = a.xs;
  const BinRanges<ROWS> offb = bin_ranges<ROWS>(in, a.off, a.rel);
  int bc[3];
  if (GABRIEL || SYM)
    bin_coords_div(in, me.x, me.y, me.z, bc);  // binned_gabriel.hpp:101, *_symmetric.hpp
  else
    bin_coords_mul(in, me.x, me.y, me.z, bc);  // binned_all_pairs.hpp:103-109
  const int R = in.search_radius;
  const int x_lo = max(bc[0] - R, 0);
  const int x_hi = min(bc[0] + R, in.ext[0] - 1);
  const float r2 = in.cutoff2;
  for (int dz = -R; dz <= R; ++dz) {
    const int z = bc[2] + dz;
    if (z < 0 || z >= in.ext[2]) continue;
    for (int dy = -R; dy <= R; ++dy) {
      const int y = bc[1] + dy;
      if (y < 0 || y >= in.ext[1]) continue;
      const int start = offb.at(x_lo, y, z);
      const int end = offb.at(x_hi + 1, y, z);
      for (int m = start; m < end; ++m) {
        if (m == s) continue;
        const float4 pj = xs[m];
        const float disp[3] = {me.x - pj.x, me.y - pj.y, me.z - pj.z};
        const float d2 = len2(disp[0], disp[1], disp[2]);
        if (d2 >= r2) continue;
        if (SYM && a.id[m] >= me_id) continue;  // j < i only
        if (GABRIEL) {
          if (gabriel_blocked(in, xs, offb, s, m, me, disp, d2)) continue;
        }
        visit<HAS_TYPE, HAS_POL, MODE, SYM>(a, in, cs, me, pme, m, pj, disp, d2, in.force.table);
      }
    }
  }
}

// Accepted neighbours a thread of the deferred-force sweep keeps in LDS.
constexpr int kDeferMax = 16;

// One candidate m of cell s: the reference's tests in its order (distance, for the symmetric
// finders the partner index, then the Gabriel ball; binned_all_pairs.hpp:140-151,
// binned_gabriel.hpp:123-209). A NaN distance is accepted, as `d2 >= r2` is false for it.
template <bool GABRIEL, bool SYM, class BR>
__device__ __forceinline__ bool accepted(const SweepArgs& a, const InstanceDesc& in,
                                         const float4* __restrict__ xs,
                                         const BR& offb, int s, int me_id,
                                         float4 me, int m, float4 pj) {
  const float disp[3] = {me.x - pj.x, me.y - pj.y, me.z - pj.z};
  const float d2 = len2(disp[0], disp[1], disp[2]);
  if (d2 >= in.cutoff2 || m == s) return false;
  if (SYM && a.id[m] >= me_id) return false;
  if (GABRIEL && gabriel_blocked(in, xs, offb, s, m, me, disp, d2)) return false;
  return true;
}

// The binned sweep with the force evaluations taken out of the candidate loop (search radius
// 1, force modes). In the one-thread-per-cell walk, the lanes of a wavefront accept neighbours
// at different iterations, so the force code runs in nearly every iteration with few lanes
// active. Here
//   1. each lane walks its candidates in the reference's order, two per iteration (both
//      positions are loaded before either is tested), and appends the accepted slots to its
//      LDS list; it counts every accepted candidate, the list keeps the first kDeferMax;
//   2. the wavefront evaluates the forces list entry by list entry, every lane its own e-th
//      pair, so the force code runs max(list length) times with most lanes active;
//   3. a lane that accepted more than kDeferMax walks again and evaluates the ones after the
//      first kDeferMax inline.
// Each lane still adds its contributions in its own visiting order, so the sums are bitwise
// those of walk_binned. Called by every lane of the wavefront (inactive lanes have no cell).
template <bool GABRIEL, bool HAS_TYPE, bool HAS_POL, int MODE, bool SYM, bool ROWS>
__device__ __forceinline__ void walk_binned_deferred(const SweepArgs& a, const InstanceDesc& in,
                                                     int s, int me_id, bool active, float4 me,
                                                     const Pol& pme, CellState& cs, int* list) {
  const float4* __restrict__ xs = a.xs;
  const BinRanges<ROWS> offb = bin_ranges<ROWS>(in, a.off, a.rel);
  int na = 0;  // accepted candidates, also those beyond the list
  int rs[9], re[9];
  if (active) {
    int bc[3];
    if (GABRIEL || SYM)
      bin_coords_div(in, me.x, me.y, me.z, bc);  // binned_gabriel.hpp:101, *_symmetric.hpp
    else
      bin_coords_mul(in, me.x, me.y, me.z, bc);  // binned_all_pairs.hpp:103-109
    const int x_lo = max(bc[0] - 1, 0);
    const int x_hi = min(bc[0] + 1, in.ext[0] - 1);
#pragma unroll
    for (int r = 0; r < 9; ++r) {
      const int z = bc[2] + r / 3 - 1, y = bc[1] + r % 3 - 1;
      const bool ok = z >= 0 && z < in.ext[2] && y >= 0 && y < in.ext[1];
      rs[r] = ok ? offb.at(x_lo, y, z) : 0;
      re[r] = ok ? offb.at(x_hi + 1, y, z) : 0;
    }
    for (int r = 0; r < 9; ++r) {
      const int end = re[r];
      int m = rs[r];
      for (; m + 1 < end; m += 2) {
        const float4 p0 = xs[m];
        const float4 p1 = xs[m + 1];
        const bool a0 = accepted<GABRIEL, SYM>(a, in, xs, offb, s, me_id, me, m, p0);
        const bool a1 = accepted<GABRIEL, SYM>(a, in, xs, offb, s, me_id, me, m + 1, p1);
        if (a0) {
          if (na < kDeferMax) list[na * kBlock] = m;
          ++na;
        }
        if (a1) {
          if (na < kDeferMax) list[na * kBlock] = m + 1;
          ++na;
        }
      }
      if (m < end) {
        const float4 p0 = xs[m];
        if (accepted<GABRIEL, SYM>(a, in, xs, offb, s, me_id, me, m, p0)) {
          if (na < kDeferMax) list[na * kBlock] = m;
          ++na;
        }
      }
    }
  }
  // 2. forces, one list entry per lane per round
  const int nl = min(na, kDeferMax);
  for (int e = 0; __ballot(e < nl) != 0ull; ++e) {
    if (e < nl) {
      const int m = list[e * kBlock];
      const float4 pj = xs[m];
      const float disp[3] = {me.x - pj.x, me.y - pj.y, me.z - pj.z};
      const float d2 = len2(disp[0], disp[1], disp[2]);
      visit<HAS_TYPE, HAS_POL, MODE, SYM>(a, in, cs, me, pme, m, pj, disp, d2, in.force.table);
    }
  }
  // 3. the accepted candidates after the first kDeferMax, inline and in order
  if (na > kDeferMax) {
    int k = 0;
    for (int r = 0; r < 9; ++r) {
      const int end = re[r];
      for (int m = rs[r]; m < end; ++m) {
        const float4 pj = xs[m];
        if (!accepted<GABRIEL, SYM>(a, in, xs, offb, s, me_id, me, m, pj)) continue;
        if (k++ < kDeferMax) continue;
        const float disp[3] = {me.x - pj.x, me.y - pj.y, me.z - pj.z};
        const float d2 = len2(disp[0], disp[1], disp[2]);
        visit<HAS_TYPE, HAS_POL, MODE, SYM>(a, in, cs, me, pme, m, pj, disp, d2, in.force.table);
      }
    }
  }
}

// The all-pairs walk with one candidate stream per lane (search radius 1, force modes).
// walk_binned_deferred runs its candidate loop row by row, so a wavefront spends on each of the
// 9 rows as many iterations as its longest lane there; here every lane walks the slot ranges
// of its rows back to back, and the wavefront runs as many iterations as its longest lane's
// whole candidate set. Each lane's ranges (its 9 rows, empty ones dropped) are kept in LDS:
//   1. each lane walks its ranges in the reference's order, two candidates per iteration, and
//      appends the accepted ones to its LDS list (the first kDeferMax; it counts all of them)
//      as (range << 8) | offset in the range. The cell's own slot is walked like the others
//      and is accepted (distance 0, or NaN, neither >= r2); it takes a list entry;
//   2. the wavefront evaluates the forces list entry by list entry, skipping the own slot;
//   3. a lane that accepted more than kDeferMax walks its ranges again and evaluates the ones
//      after the first kDeferMax inline.
// Each lane adds its contributions in its own visiting order without its own slot, so the
// sums are bitwise those of walk_binned. A range must hold fewer than 256 slots (flat_ranges
// reports it); the LDS (about 19.3 KiB per workgroup) keeps 8 workgroups on a CU.
constexpr int kFlatRanges = 9;

// LDS of the flattened walk, lane-interleaved ([k * kBlock + lane]).
struct FlatLds {
  int start[kFlatRanges * kBlock];
  unsigned short list[kDeferMax * kBlock];
  unsigned char len[kFlatRanges * kBlock];
  float tab[kMaxTypes * kMaxTypes];  // adhesion scale by (type_i * n + type_j), or [0] untyped
};
// the row-by-row walk's list fits in it (its fallback)
static_assert(sizeof(FlatLds) >= sizeof(int) * kDeferMax * kBlock, "FlatLds too small");

// Step 0: the lane's ranges into LDS; returns their count. `fits` is cleared when a range holds
// 256 slots or more; `self` is set when a range holds the lane's own slot s and the walk will
// accept it (its distance to itself, 0 or NaN, is not >= r2).
template <bool ROWS>
__device__ __forceinline__ int flat_ranges(const SweepArgs& a, const InstanceDesc& in, int s,
                                           float4 me, FlatLds& L, bool& fits, bool& self) {
  const BinRanges<ROWS> offb = bin_ranges<ROWS>(in, a.off, a.rel);
  const int lane = (int)threadIdx.x;
  int bc[3];
  bin_coords_mul(in, me.x, me.y, me.z, bc);  // binned_all_pairs.hpp:103-109
  const int x_lo = max(bc[0] - 1, 0);
  const int x_hi = min(bc[0] + 1, in.ext[0] - 1);
  int rs[9], re[9];
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    const int z = bc[2] + r / 3 - 1, y = bc[1] + r % 3 - 1;
    const bool ok = z >= 0 && z < in.ext[2] && y >= 0 && y < in.ext[1];
    rs[r] = ok ? offb.at(x_lo, y, z) : 0;
    re[r] = ok ? offb.at(x_hi + 1, y, z) : 0;
  }
  int nr = 0;
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    const int b = rs[r], e = re[r];
    if (e - b >= 256) fits = false;
    if (b <= s && s < e) self = !(len2(me.x - me.x, me.y - me.y, me.z - me.z) >= in.cutoff2);
    if (b < e) {
      L.start[nr * kBlock + lane] = b;
      L.len[(nr++) * kBlock + lane] = (unsigned char)(e - b);
    }
  }
  return nr;
}

template <bool HAS_TYPE, bool HAS_POL, int MODE>
__device__ __forceinline__ void walk_binned_flat(const SweepArgs& a, const InstanceDesc& in, int s,
                                                 int nr, bool self, float4 me, const Pol& pme,
                                                 CellState& cs, FlatLds& L) {
  const float4* __restrict__ xs = a.xs;
  const int lane = (int)threadIdx.x;
  const float r2 = in.cutoff2;
  int na = 0;  // accepted candidates (the own slot among them), also those beyond the list
  if (nr > 0) {
    // one loop over the lane's candidates: q walks range k, o is the offset in it, n its
    // length; the second load of an iteration may read one slot past the range (positions
    // carry a slot past the last one), its result used only when o + 1 < n
    const int* sp = L.start + lane;
    const unsigned char* lp = L.len + lane;
    const int* const sp_end = sp + nr * kBlock;
    const float4* q = xs + *sp;
    int n = (int)*lp, o = 0, tag = 0;
    // each iteration's pair is loaded in the iteration before (the first before the loop), so
    // its loads are in flight while that iteration tests its own pair
    float4 q0 = q[0], q1 = q[1];
    while (o < n) {
      const float4 c0 = q0, c1 = q1;
      const int entry = tag | o;      // the list entry of c0 (c1's: entry + 1)
      const bool has1 = o + 1 < n;
      o += 2;
      q += 2;
      if (o >= n) {
        sp += kBlock;
        lp += kBlock;
        tag += 1 << 8;
        if (sp < sp_end) {
          q = xs + *sp;
          n = (int)*lp;
          o = 0;
        }
      }
      // past the lane's last candidate it loads its own slot, unused
      const float4* qn = o < n ? q : xs + s;
      q0 = qn[0];
      q1 = qn[1];
      // binned_all_pairs.hpp:140-151: a NaN distance is accepted, as `d2 >= r2` is false
      const float d0 = len2(me.x - c0.x, me.y - c0.y, me.z - c0.z);
      const float d1 = len2(me.x - c1.x, me.y - c1.y, me.z - c1.z);
      if (!(d0 >= r2)) {
        if (na < kDeferMax) L.list[na * kBlock + lane] = (unsigned short)entry;
        ++na;
      }
      if (has1 && !(d1 >= r2)) {
        if (na < kDeferMax) L.list[na * kBlock + lane] = (unsigned short)(entry + 1);
        ++na;
      }
    }
  }
  // 2. forces, one list entry per lane per round; the rounds run to the wavefront's longest
  // list, a lane past its own list evaluates its own slot and adds nothing
  const int nl = min(na, kDeferMax);
  int rounds = nl;
#pragma unroll
  for (int w = 32; w > 0; w >>= 1) rounds = max(rounds, __shfl_xor(rounds, w, 64));
  if (!HAS_POL) {
    // PiecewiseLinear alone (predefined.hpp:107-144), its parameters in registers, the
    // adhesion table in LDS, a loop with no branch in its body
    cs.pairs += (unsigned)(na - (self ? 1 : 0));  // the own slot's acceptance is not a pair
    const ForceParams& f = in.force;
    const float h = f.h, krep = f.krep;
    const int tn = (HAS_TYPE && f.n_types > 0) ? f.n_types : 0;
    const int row = tn * __float_as_int(me.w);
    const int tmask = tn > 0 ? kMaxTypes * kMaxTypes - 1 : 0;
    if ((f.flags & kForceLinear) != 0) {
      // round e's partner slot (a lane past its list: its own slot)
      auto partner = [&](int e) {
        const int en = (int)L.list[min(e, kDeferMax - 1) * kBlock + lane];
        return e < nl ? L.start[(en >> 8) * kBlock + lane] + (en & 255) : s;
      };
      // the rounds are software-pipelined: the next round's partner is looked up and its
      // position loaded while this round's force is computed (the loads, not the arithmetic,
      // bound the rounds)
      int m_next = partner(0);
      float4 pj_next = xs[m_next];
      for (int e = 0; e < rounds; ++e) {
        const int m = m_next;
        const float4 pj = pj_next;
        m_next = partner(e + 1);
        pj_next = xs[m_next];
        const float kadh = L.tab[(row + (__float_as_int(pj.w) & tmask)) & (kMaxTypes * kMaxTypes - 1)];
        const float disp[3] = {me.x - pj.x, me.y - pj.y, me.z - pj.z};
        const float d = sqrtf(len2(disp[0], disp[1], disp[2]));
        const float F = krep * fmaxf(h - d, 0.0f) - kadh * fmaxf(d - h, 0.0f);
        // the own slot, or d == 0 (a zero force): adding +0 changes no sum; such a lane divides
        // by 1 (its quotient is not used, its displacement is 0), so it never sends the
        // wavefront into the IEEE path
        const bool z = m == s || d == 0.0f;
        float v[3];
        pl_div3_rn(disp, F, z ? 1.0f : d, v);  // (disp * F) / d
        cs.acc_pos[0] = cs.acc_pos[0] + (z ? 0.0f : v[0]);
        cs.acc_pos[1] = cs.acc_pos[1] + (z ? 0.0f : v[1]);
        cs.acc_pos[2] = cs.acc_pos[2] + (z ? 0.0f : v[2]);
      }
    }
  } else {
    for (int e = 0; e < rounds; ++e) {
      if (e < nl) {
        const int en = (int)L.list[e * kBlock + lane];
        const int m = L.start[(en >> 8) * kBlock + lane] + (en & 255);
        if (m != s) {
          const float4 pj = xs[m];
          const float disp[3] = {me.x - pj.x, me.y - pj.y, me.z - pj.z};
          const float d2 = len2(disp[0], disp[1], disp[2]);
          visit<HAS_TYPE, HAS_POL, MODE, false>(a, in, cs, me, pme, m, pj, disp, d2, in.force.table);
        }
      }
    }
  }
  // 3. the accepted candidates after the first kDeferMax, inline and in order, resuming the
  // walk right after the last listed one (its range and offset are in its list entry)
  if (na > kDeferMax) {
    const int last = (int)L.list[(kDeferMax - 1) * kBlock + lane];
    int o0 = (last & 255) + 1;
    for (int r = last >> 8; r < nr; ++r, o0 = 0) {
      const int b = L.start[r * kBlock + lane];
      const int end = b + (int)L.len[r * kBlock + lane];
      for (int m = b + o0; m < end; ++m) {
        const float4 pj = xs[m];
        const float disp[3] = {me.x - pj.x, me.y - pj.y, me.z - pj.z};
        const float d2 = len2(disp[0], disp[1], disp[2]);
        if (d2 >= r2 || m == s) continue;
        if (!HAS_POL) cs.pairs--;  // counted above; visit counts it again
        visit<HAS_TYPE, HAS_POL, MODE, false>(a, in, cs, me, pme, m, pj, disp, d2, L.tab);
      }
    }
  }
}

template <bool HAS_POL>
__device__ __forceinline__ CorrIn load_corr(const SweepArgs& a, int s) {
  CorrIn c{};
  c.d1 = load_stream(a.d1_pos + s);
  c.x0 = load_stream(a.o2_pos + s);
  if (HAS_POL) {
    c.e1 = a.d1_pol[s];
    c.p0 = a.o2_pol[s];
  }
  return c;
}

// Epilogue shared by the binned and naive sweeps: turn the accumulators into the evaluation
// result and apply the mode's integrator update.
// The key at the last rebuild of the cell in slot s, loaded with the cell's other inputs before
// the walk when the step's final sweep counts the next rebuild's keys incrementally (else -1).
template <int MODE>
__device__ __forceinline__ int load_old_key(const SweepArgs& a, int s, bool active) {
  if ((MODE == kModeCorr || MODE == kModeEuler) && a.kc.keys && a.kc.skey && active)
    return a.kc.skey[s];
  return -1;
}

template <bool HAS_POL, int MODE, bool SYM>
__device__ __forceinline__ void finish(const SweepArgs& a, const InstanceDesc& in, int ii, int lb,
                                       int s, bool active, float4 me, float2 mp, CellState& cs,
                                       const CorrIn& ci, int old_key = -1) {
  __shared__ float lds_f[4];
  __shared__ unsigned lds_u[4];
  if (MODE == kModeCount) {
    if (active) a.row_count[a.id[s]] = (int)cs.pairs;
    return;
  }
  if (MODE == kModeEmit) return;
  // out(i) += total (the delta buffer is zero before an evaluation), then
  // out.pos(i) += total_velocity / fmax(total_drag, eps) (binned_all_pairs.hpp:174-180), which
  // is +0 when no pair has drag (the modes other than DeltaDrag).
  float F[3], G[2];
  if (MODE == kModeDeltaDrag) {
    const float den = fmaxf(cs.drag, 1.1920928955078125e-07f);
#pragma unroll
    for (int c = 0; c < 3; ++c) F[c] = (0.0f + cs.acc_pos[c]) + cs.vel[c] / den;
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c) F[c] = (0.0f + cs.acc_pos[c]) + 0.0f;
  }
  G[0] = 0.0f + cs.acc_pol[0];
  G[1] = 0.0f + cs.acc_pol[1];
  const float dt = in.dt;
  float4 xnew = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // Corr / Euler: the new position
  float2 pnew = make_float2(0.0f, 0.0f);               // Corr / Euler: the new polarity
  if (active) {
    if (MODE == kModeDelta || MODE == kModeDeltaDrag) {
      a.o_pos[s] = make_float4(F[0], F[1], F[2], 0.0f);
      if (HAS_POL) a.o_pol[s] = make_float2(G[0], G[1]);
    } else if (MODE == kModePred) {
      // apply_euler_predictor, heun.hpp:146-156: predicted = current + delta * dt (the
      // delta is read once, by the corrector; the predicted positions by every walk)
      store_stream(a.o_pos + s, make_float4(F[0], F[1], F[2], 0.0f));
      a.o2_pos[s] = make_float4(me.x + F[0] * dt, me.y + F[1] * dt, me.z + F[2] * dt, me.w);
      if (HAS_POL) {
        a.o_pol[s] = make_float2(G[0], G[1]);
        const float2 np = make_float2(mp.x + G[0] * dt, mp.y + G[1] * dt);
        a.o2_pol[s] = np;
        a.o_trig[s] = polar_trig(np);
      }
    } else if (MODE == kModeCorr) {
      // apply_heun_corrector, heun.hpp:170-186
      const float4 d1 = ci.d1;
      const float s0 = d1.x + F[0], s1 = d1.y + F[1], s2 = d1.z + F[2];
      if (a.old_vel) store_stream(a.old_vel + s, make_float4(s0 * 0.5f, s1 * 0.5f, s2 * 0.5f, 0.0f));
      const float4 x0 = ci.x0;
      xnew = make_float4(x0.x + s0 * 0.5f * dt, x0.y + s1 * 0.5f * dt, x0.z + s2 * 0.5f * dt, x0.w);
      store_stream(a.o2_pos + s, xnew);  // read once more, by the next rebuild
      if (HAS_POL) {
        const float2 e1 = ci.e1;
        const float2 p0 = ci.p0;
        const float2 np = make_float2(p0.x + (e1.x + G[0]) * 0.5f * dt,
                                      p0.y + (e1.y + G[1]) * 0.5f * dt);
        a.o2_pol[s] = np;
        a.o_trig[s] = polar_trig(np);
        pnew = np;
      }
    } else if (MODE == kModeEuler) {
      // apply_euler, euler.hpp:52-68
      if (a.old_vel) a.old_vel[s] = make_float4(F[0], F[1], F[2], 0.0f);
      xnew = make_float4(me.x + F[0] * dt, me.y + F[1] * dt, me.z + F[2] * dt, me.w);
      a.o_pos[s] = xnew;
      if (HAS_POL) {
        const float2 np = make_float2(mp.x + G[0] * dt, mp.y + G[1] * dt);
        a.o_pol[s] = np;
        a.o_trig[s] = polar_trig(np);
        pnew = np;
      }
    }
    if ((MODE == kModeCorr || MODE == kModeEuler) && a.nonfinite && not_finite(xnew, pnew, HAS_POL))
      a.nonfinite[ii] = 1;
  }
  // keys of the new positions for the rebuild that starts the next step
  if ((MODE == kModeCorr || MODE == kModeEuler) && a.kc.keys)
    count_key(in, a.kc, s, active, s - in.slot_begin == in.n, xnew, old_key);
  if (MODE == kModeDelta && a.com_partial) {
    const float f0 = active ? F[0] : 0.0f, f1 = active ? F[1] : 0.0f, f2 = active ? F[2] : 0.0f;
    const float t0 = block_sum_f(f0, lds_f);
    const float t1 = block_sum_f(f1, lds_f);
    const float t2 = block_sum_f(f2, lds_f);
    if (threadIdx.x == 0) a.com_partial[lb] = make_float4(t0, t1, t2, 0.0f);
  }
  const unsigned tp = block_sum_u(active ? cs.pairs : 0u, lds_u);
  if (threadIdx.x == 0) a.block_pairs[lb] += tp;
}

template <bool GABRIEL, bool HAS_TYPE, bool HAS_POL, int MODE, bool SYM, bool ROWS>
__device__ __forceinline__ void sweep_binned(const SweepArgs& a) {
  const int lb = a.chunks ? a.chunks[logical_block()] : logical_block();
  const int ii = a.block_inst[lb];
  const InstanceDesc& in = a.inst[ii];
  const int s = lb * kBlock + (int)threadIdx.x;
  const bool active = (s - in.slot_begin) < in.n;
  CellState cs{{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f}, 0u, 0, 0.0f, {0.0f, 0.0f, 0.0f}};
  float4 me = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float2 mp = make_float2(0.0f, 0.0f);
  Pol pme{};
  CorrIn ci{};
  int me_id = 0;
  if (active) {
    me = a.xs[s];
    if (SYM) me_id = a.id[s];
    if (HAS_POL) {
      mp = a.ps[s];
      pme = make_pol(mp, a.trig[s]);
    }
    if (MODE == kModeEmit) cs.emit_at = a.row_ptr[a.id[s]];
    if (MODE == kModeCorr) ci = load_corr<HAS_POL>(a, s);
  }
  const int old_key = load_old_key<MODE>(a, s, active);
  // the drag mode walks inline: its drag sums run in the same visiting order either way
  constexpr bool kDefer = MODE != kModeCount && MODE != kModeEmit && MODE != kModeDeltaDrag;
  constexpr bool kFlat = kDefer && !GABRIEL && !SYM;
  if (kFlat && in.search_radius == 1) {
    // the instance decides the path for the whole workgroup (all its slots are one instance)
    __shared__ FlatLds flat;
    if (threadIdx.x < kMaxTypes * kMaxTypes) {
      const ForceParams& f = in.force;
      flat.tab[threadIdx.x] = (HAS_TYPE && f.n_types > 0) ? f.table[threadIdx.x] : f.kadh;
    }
    bool fits = true, self = false;
    const int nr = active ? flat_ranges<ROWS>(a, in, s, me, flat, fits, self) : 0;
    // the table, and whether every range fits an 8-bit length (a workgroup barrier)
    if (__syncthreads_and(fits)) {
      walk_binned_flat<HAS_TYPE, HAS_POL, MODE>(a, in, s, nr, self, me, pme, cs, flat);
    } else {
      // a range of 256 slots or more: the row-by-row walk, its list in the ranges' LDS (the
      // workgroup has passed the barrier, so no lane reads the ranges any more)
      walk_binned_deferred<GABRIEL, HAS_TYPE, HAS_POL, MODE, SYM, ROWS>(a, in, s, me_id, active, me, pme,
                                                                  cs, reinterpret_cast<int*>(&flat) + threadIdx.x);
    }
  } else if (kDefer && in.search_radius == 1) {
    __shared__ int list[kDeferMax * kBlock];
    walk_binned_deferred<GABRIEL, HAS_TYPE, HAS_POL, MODE, SYM, ROWS>(a, in, s, me_id, active, me, pme,
                                                                cs, list + threadIdx.x);
  } else if (active) {
    walk_binned<GABRIEL, HAS_TYPE, HAS_POL, MODE, SYM, ROWS>(a, in, s, me_id, me, pme, cs);
  }
  finish<HAS_POL, MODE, SYM>(a, in, ii, lb, s, active, me, mp, cs, ci, old_key);
}

template <bool GABRIEL, bool HAS_TYPE, bool HAS_POL, int MODE, bool SYM, bool ROWS>
__global__ __launch_bounds__(kBlock) void k_sweep_binned(SweepArgs a) {
  sweep_binned<GABRIEL, HAS_TYPE, HAS_POL, MODE, SYM, ROWS>(a);
}

// The all-pairs sweep of models without polarity, held to 80 scalar registers: with 84 or more
// the stale-grid sweep kept fewer wavefronts resident on gfx950 (measured with SQ_WAVE_CYCLES /
// SQ_BUSY_CYCLES, although the compiler reports 8 wavefronts per SIMD either way) and ran
// slower; at 80 it needs no spills. The other variants need more and keep the default budget.
template <bool HAS_TYPE, int MODE, bool ROWS>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_num_sgpr(80))) void k_sweep_binned_ap(SweepArgs a) {
  sweep_binned<false, HAS_TYPE, false, MODE, false, ROWS>(a);
}

// Naive finders (naive_all_pairs.hpp:41-103, naive_gabriel.hpp:47-119): j runs over all agents
// in ascending order; tiles of j positions are staged in LDS and shared by the workgroup.
// Slots of a naive-finder instance are never re-sorted, so slot order is agent order (and the
// symmetric finders' j < i is m < s).
template <bool GABRIEL, bool HAS_TYPE, bool HAS_POL, int MODE, bool SYM>
__global__ __launch_bounds__(kBlock) void k_sweep_naive(SweepArgs a) {
  __shared__ float4 tile[kBlock];
  const int lb = logical_block();
  const int ii = a.block_inst[lb];
  const InstanceDesc& in = a.inst[ii];
  const int s = lb * kBlock + (int)threadIdx.x;
  const bool active = (s - in.slot_begin) < in.n;
  CellState cs{{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f}, 0u, 0, 0.0f, {0.0f, 0.0f, 0.0f}};
  float4 me = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float2 mp = make_float2(0.0f, 0.0f);
  Pol pme{};
  CorrIn ci{};
  if (active) {
    me = a.xs[s];
    if (HAS_POL) {
      mp = a.ps[s];
      pme = make_pol(mp, a.trig[s]);
    }
    if (MODE == kModeEmit) cs.emit_at = a.row_ptr[a.id[s]];
    if (MODE == kModeCorr) ci = load_corr<HAS_POL>(a, s);
  }
  const int begin = in.slot_begin, n = in.n;
  const float r2 = in.cutoff2;
  for (int base = 0; base < n; base += kBlock) {
    __syncthreads();
    if (base + (int)threadIdx.x < n) tile[threadIdx.x] = a.xs[begin + base + threadIdx.x];
    __syncthreads();
    const int cnt = min(kBlock, n - base);
    if (!active) continue;
    for (int t = 0; t < cnt; ++t) {
      const int m = begin + base + t;
      if (SYM ? m >= s : m == s) continue;
      const float4 pj = tile[t];
      const float disp[3] = {me.x - pj.x, me.y - pj.y, me.z - pj.z};
      const float d2 = len2(disp[0], disp[1], disp[2]);
      if (d2 >= r2) continue;
      if (GABRIEL) {
        // naive_gabriel.hpp:68-80: rho^2 = (d2 * 0.25) * c^2, strict test over every k
        const float mid0 = me.x - disp[0] * 0.5f;
        const float mid1 = me.y - disp[1] * 0.5f;
        const float mid2 = me.z - disp[2] * 0.5f;
        const float rho2 = d2 * 0.25f * in.gabriel_c2;
        bool blocked = false;
        for (int k = begin; k < begin + n; ++k) {
          if (k == s || k == m) continue;
          const float4 pk = a.xs[k];
          if (len2(pk.x - mid0, pk.y - mid1, pk.z - mid2) < rho2) {
            blocked = true;
            break;
          }
        }
        if (blocked) continue;
      }
      visit<HAS_TYPE, HAS_POL, MODE, SYM>(a, in, cs, me, pme, m, pj, disp, d2, in.force.table);
    }
  }
  finish<HAS_POL, MODE, SYM>(a, in, ii, lb, s, active, me, mp, cs, ci);
}

// ------------------------------------------------------------------ symmetric finders
// Binned symmetric finders, Delta mode, the two-walk rule: o_pos / o_pol receive each cell's
// own and gathered contributions (plain stores), q_pos / q_pol the movers' scattered ones.
// STALE: a.skey is set (a stale grid, movers possible). Search radius 1: the stayers defer
// their forces as in walk_binned_deferred, list entries carrying the role in bits 30-31; the
// rare movers are walked by k_sym_movers, a wavefront per mover, after the launch.
template <bool GABRIEL, bool HAS_TYPE, bool HAS_POL, bool STALE, bool ROWS>
__global__ __launch_bounds__(kBlock) void k_sym_binned(SweepArgs a) {
  __shared__ int list_all[kDeferMax * kBlock];
  __shared__ unsigned lds_u[4];
  const int lb = a.chunks ? a.chunks[logical_block()] : logical_block();
  const InstanceDesc& in = a.inst[a.block_inst[lb]];
  const int s = lb * kBlock + (int)threadIdx.x;
  const bool active = (s - in.slot_begin) < in.n;
  const float4* __restrict__ xs = a.xs;
  const BinRanges<ROWS> offb = bin_ranges<ROWS>(in, a.off, a.rel);
  const int* skey = STALE ? a.skey : nullptr;
  int* list = list_all + threadIdx.x;
  float4 me = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  Pol pme{};
  int me_id = 0;
  int bc[3] = {0, 0, 0};
  bool mover = false;
  if (active) {
    me = xs[s];
    if (HAS_POL) pme = make_pol(a.ps[s], a.trig[s]);
    me_id = a.id[s];
    bin_coords_div(in, me.x, me.y, me.z, bc);  // calc_bin_coords_from_point, *_symmetric.hpp
    if (STALE) mover = flatten(in, bc[0], bc[1], bc[2]) != skey[s];
  }
  float acc_pos[3] = {0.0f, 0.0f, 0.0f}, acc_pol[2] = {0.0f, 0.0f};
  unsigned pairs = 0;
  const bool stayer = active && !mover;
  if (in.search_radius == 1) {
    int na = 0;
    int rs[9], re[9];
    if (stayer) {
      const int x_lo = max(bc[0] - 1, 0);
      const int x_hi = min(bc[0] + 1, in.ext[0] - 1);
#pragma unroll
      for (int r = 0; r < 9; ++r) {
        const int z = bc[2] + r / 3 - 1, y = bc[1] + r % 3 - 1;
        const bool ok = z >= 0 && z < in.ext[2] && y >= 0 && y < in.ext[1];
        rs[r] = ok ? offb.at(x_lo, y, z) : 0;
        re[r] = ok ? offb.at(x_hi + 1, y, z) : 0;
      }
      for (int r = 0; r < 9; ++r) {
        const int end = re[r];
        int m = rs[r];
        for (; m + 1 < end; m += 2) {
          // positions, agent indices and (stale grid) mover flags of both candidates are loaded
          // before any test
          const float4 p0 = xs[m];
          const float4 p1 = xs[m + 1];
          const int i0 = a.id[m], i1 = a.id[m + 1];
          const int f0 = STALE && a.mover8 ? (int)a.mover8[m] : -1;
          const int f1 = STALE && a.mover8 ? (int)a.mover8[m + 1] : -1;
          const int r0 = sym_role<GABRIEL, false>(in, skey, a.mover8, xs, offb, s, me_id, false, me, m, p0, i0,
                                                  kWalkBoth, f0);
          const int r1 = sym_role<GABRIEL, false>(in, skey, a.mover8, xs, offb, s, me_id, false, me, m + 1, p1, i1,
                                                  kWalkBoth, f1);
          if (r0 != kRoleNone) {
            if (na < kDeferMax) list[na * kBlock] = (int)((unsigned)m | ((unsigned)r0 << 30));
            ++na;
          }
          if (r1 != kRoleNone) {
            if (na < kDeferMax) list[na * kBlock] = (int)((unsigned)(m + 1) | ((unsigned)r1 << 30));
            ++na;
          }
        }
        if (m < end) {
          const int r0 =
              sym_role<GABRIEL, false>(in, skey, a.mover8, xs, offb, s, me_id, false, me, m, xs[m], a.id[m]);
          if (r0 != kRoleNone) {
            if (na < kDeferMax) list[na * kBlock] = (int)((unsigned)m | ((unsigned)r0 << 30));
            ++na;
          }
        }
      }
    }
    // stayers have no role that scatters; each round loads the next round's entry and partner
    // position before evaluating its own
    const int nl = min(na, kDeferMax);
    unsigned ent_next = 0u;
    float4 pm_next = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (nl > 0) {
      ent_next = (unsigned)list[0];
      pm_next = xs[ent_next & 0x3fffffffu];
    }
    for (int e = 0; __ballot(e < nl) != 0ull; ++e) {
      const unsigned ent = ent_next;
      const float4 pm = pm_next;
      if (e + 1 < nl) {
        ent_next = (unsigned)list[(e + 1) * kBlock];
        pm_next = xs[ent_next & 0x3fffffffu];
      }
      if (e < nl) {
        const int role = (int)(ent >> 30);
        sym_eval_at<HAS_TYPE, HAS_POL, false>(a, in, me, pme, (int)(ent & 0x3fffffffu), pm, role,
                                              acc_pos, acc_pol);
        pairs += role != kRoleGather;
      }
    }
    if (na > kDeferMax) {
      int k = 0;
      for (int r = 0; r < 9; ++r) {
        const int end = re[r];
        for (int m = rs[r]; m < end; ++m) {
          const int role =
              sym_role<GABRIEL, false>(in, skey, a.mover8, xs, offb, s, me_id, false, me, m, xs[m], a.id[m]);
          if (role == kRoleNone || k++ < kDeferMax) continue;
          sym_eval<HAS_TYPE, HAS_POL, false>(a, in, me, pme, m, role, acc_pos, acc_pol);
          pairs += role != kRoleGather;
        }
      }
    }
  } else if (stayer) {
    sym_cell_serial<GABRIEL, HAS_TYPE, HAS_POL, false, ROWS>(a, in, s, me_id, false, me, pme, bc,
                                                       acc_pos, acc_pol, pairs);
  }
  // the movers: with search radius 1, k_sym_movers walks them after this launch (a mover's
  // delta is written as zero here and replaced there); otherwise serially in their lanes
  if (STALE && mover && in.search_radius != 1)
    sym_cell_serial<GABRIEL, HAS_TYPE, HAS_POL, false, ROWS>(a, in, s, me_id, true, me, pme, bc,
                                                       acc_pos, acc_pol, pairs);
  if (active) {
    // read once, by the update (streaming hint)
    store_stream(a.o_pos + s, make_float4(acc_pos[0], acc_pos[1], acc_pos[2], 0.0f));
    if (HAS_POL) a.o_pol[s] = make_float2(acc_pol[0], acc_pol[1]);
    if (!STALE && a.o2_pos) {
      // the Heun predictor fused (heun.hpp:146-156), asked for on a fresh grid without the COM
      // fix: no cell moved out of its bin, so no contribution arrives after this walk and the
      // delta is final. The same arithmetic as the separate update, and the flags of the cells
      // the predicted positions take out of their rebuild bin (flatten(bc), the key the grid
      // was just built with).
      const float dt = in.dt;
      const float4 xp = make_float4(me.x + acc_pos[0] * dt, me.y + acc_pos[1] * dt,
                                    me.z + acc_pos[2] * dt, me.w);
      a.o2_pos[s] = xp;
      if (a.mover8_out) a.mover8_out[s] = bin_key(in, xp) != flatten(in, bc[0], bc[1], bc[2]) ? 1 : 0;
      if (HAS_POL) {
        const float2 np = make_float2(pme.t + acc_pol[0] * dt, pme.p + acc_pol[1] * dt);
        a.o2_pol[s] = np;
        a.o_trig[s] = polar_trig(np);
      }
    }
  }
  const unsigned tp = block_sum_u(active ? pairs : 0u, lds_u);
  if (threadIdx.x == 0) a.block_pairs[lb] += tp;
}

// The movers of a stale-grid symmetric sweep (search radius 1), after k_sym_binned: each
// wavefront takes the movers among its 64 slots, one after another over all its lanes
// (sym_movers_wave), and writes their deltas; a wavefront without a mover ends at once. Its
// own register budget keeps the stayers' sweep at eight wavefronts per SIMD.
template <bool GABRIEL, bool HAS_TYPE, bool HAS_POL, bool ROWS>
__global__ __launch_bounds__(kBlock) void k_sym_movers(SweepArgs a) {
  const int lb = a.chunks ? a.chunks[logical_block()] : logical_block();
  const InstanceDesc& in = a.inst[a.block_inst[lb]];
  if (in.search_radius != 1) return;  // k_sym_binned walked them
  const int s = lb * kBlock + (int)threadIdx.x;
  const bool active = (s - in.slot_begin) < in.n;
  bool mover = false;
  if (active) {
    const float4 me = a.xs[s];
    int bc[3];
    bin_coords_div(in, me.x, me.y, me.z, bc);  // as k_sym_binned decides it
    mover = flatten(in, bc[0], bc[1], bc[2]) != a.skey[s];
  }
  if (__ballot(mover) == 0ull) return;
  float acc_pos[3] = {0.0f, 0.0f, 0.0f}, acc_pol[2] = {0.0f, 0.0f};
  const unsigned pairs = sym_movers_wave<GABRIEL, HAS_TYPE, HAS_POL, ROWS>(a, in, s, mover, acc_pos, acc_pol);
  if (mover) {
    a.o_pos[s] = make_float4(acc_pos[0], acc_pos[1], acc_pos[2], 0.0f);
    if (HAS_POL) a.o_pol[s] = make_float2(acc_pol[0], acc_pol[1]);
  }
  unsigned tp = pairs;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) tp += __shfl_xor(tp, o, 64);
  if ((threadIdx.x & 63) == 0 && tp != 0u) atomicAdd(&a.block_pairs[lb], (unsigned long long)tp);
}

// Naive symmetric finders, Delta mode (naive_all_pairs_symmetric.hpp:100-143,
// naive_gabriel_symmetric.hpp:107-168): slot order is agent order, so cell s plays the own
// role for m < s and the gather role for m > s; no grid, no movers.
template <bool GABRIEL, bool HAS_TYPE, bool HAS_POL>
__global__ __launch_bounds__(kBlock) void k_sym_naive(SweepArgs a) {
  __shared__ float4 tile[kBlock];
  __shared__ unsigned lds_u[4];
  const int lb = logical_block();
  const InstanceDesc& in = a.inst[a.block_inst[lb]];
  const int s = lb * kBlock + (int)threadIdx.x;
  const bool active = (s - in.slot_begin) < in.n;
  float4 me = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  Pol pme{};
  if (active) {
    me = a.xs[s];
    if (HAS_POL) pme = make_pol(a.ps[s], a.trig[s]);
  }
  float acc_pos[3] = {0.0f, 0.0f, 0.0f}, acc_pol[2] = {0.0f, 0.0f};
  unsigned pairs = 0;
  const int begin = in.slot_begin, n = in.n;
  for (int base = 0; base < n; base += kBlock) {
    __syncthreads();
    if (base + (int)threadIdx.x < n) tile[threadIdx.x] = a.xs[begin + base + threadIdx.x];
    __syncthreads();
    const int cnt = min(kBlock, n - base);
    if (!active) continue;
    for (int t = 0; t < cnt; ++t) {
      const int m = begin + base + t;
      if (m == s) continue;
      const float4 pm = tile[t];
      const float disp[3] = {me.x - pm.x, me.y - pm.y, me.z - pm.z};
      const float d2 = len2(disp[0], disp[1], disp[2]);
      if (d2 >= in.cutoff2) continue;
      const bool own = m < s;
      if (GABRIEL) {
        // the ball of the pair as its larger-index cell sees it: rho^2 = (d2 * 0.25) * c^2,
        // strict test over every other cell (naive_gabriel_symmetric.hpp:123-139)
        const float4 pi = own ? me : pm;
        const float sgn = own ? 1.0f : -1.0f;
        const float mid0 = pi.x - (sgn * disp[0]) * 0.5f;
        const float mid1 = pi.y - (sgn * disp[1]) * 0.5f;
        const float mid2 = pi.z - (sgn * disp[2]) * 0.5f;
        const float rho2 = d2 * 0.25f * in.gabriel_c2;
        bool blocked = false;
        for (int k = begin; k < begin + n; ++k) {
          if (k == s || k == m) continue;
          const float4 pk = a.xs[k];
          if (len2(pk.x - mid0, pk.y - mid1, pk.z - mid2) < rho2) {
            blocked = true;
            break;
          }
        }
        if (blocked) continue;
      }
      sym_eval_at<HAS_TYPE, HAS_POL, false>(a, in, me, pme, m, pm, own ? kRoleOwn : kRoleGather,
                                            acc_pos, acc_pol);
      pairs += own;
    }
  }
  if (active) {
    // read once, by the update (streaming hint)
    store_stream(a.o_pos + s, make_float4(acc_pos[0], acc_pos[1], acc_pos[2], 0.0f));
    if (HAS_POL) a.o_pol[s] = make_float2(acc_pol[0], acc_pol[1]);
  }
  const unsigned tp = block_sum_u(active ? pairs : 0u, lds_u);
  if (threadIdx.x == 0) a.block_pairs[lb] += tp;
}

template <bool GABRIEL, bool BINNED, bool HAS_TYPE, bool HAS_POL, int MODE, bool SYM>
void launch_one(const SweepArgs& a, int n_blocks, hipStream_t stream) {
  // a.rel: the batch has sparse instances (their bin lookups need the ROWS kernels)
  if (BINNED && !GABRIEL && !HAS_POL && !SYM) {
    if (a.rel)
      hipLaunchKernelGGL((k_sweep_binned_ap<HAS_TYPE, MODE, true>), dim3(n_blocks), dim3(kBlock), 0, stream, a);
    else
      hipLaunchKernelGGL((k_sweep_binned_ap<HAS_TYPE, MODE, false>), dim3(n_blocks), dim3(kBlock), 0, stream, a);
  } else if (BINNED && a.rel) {
    hipLaunchKernelGGL((k_sweep_binned<GABRIEL, HAS_TYPE, HAS_POL, MODE, SYM, true>), dim3(n_blocks),
                       dim3(kBlock), 0, stream, a);
  } else if (BINNED) {
    hipLaunchKernelGGL((k_sweep_binned<GABRIEL, HAS_TYPE, HAS_POL, MODE, SYM, false>), dim3(n_blocks),
                       dim3(kBlock), 0, stream, a);
  }
  else
    hipLaunchKernelGGL((k_sweep_naive<GABRIEL, HAS_TYPE, HAS_POL, MODE, SYM>), dim3(n_blocks),
                       dim3(kBlock), 0, stream, a);
}

template <bool GABRIEL, bool BINNED, int MODE, bool SYM>
void launch_force(bool has_type, bool has_pol, const SweepArgs& a, int n_blocks,
                  hipStream_t stream) {
  if (MODE == kModeCount || MODE == kModeEmit) {
    launch_one<GABRIEL, BINNED, false, false, MODE, SYM>(a, n_blocks, stream);
    return;
  }
  if (has_type) {
    if (has_pol)
      launch_one<GABRIEL, BINNED, true, true, MODE, SYM>(a, n_blocks, stream);
    else
      launch_one<GABRIEL, BINNED, true, false, MODE, SYM>(a, n_blocks, stream);
  } else {
    if (has_pol)
      launch_one<GABRIEL, BINNED, false, true, MODE, SYM>(a, n_blocks, stream);
    else
      launch_one<GABRIEL, BINNED, false, false, MODE, SYM>(a, n_blocks, stream);
  }
}

// Returns nonzero for an unknown mode.
template <bool GABRIEL, bool BINNED>
int launch_mode(int mode, bool has_type, bool has_pol, const SweepArgs& a, int n_blocks,
                hipStream_t stream) {
  switch (mode) {
    case kModeDelta: launch_force<GABRIEL, BINNED, kModeDelta, false>(has_type, has_pol, a, n_blocks, stream); break;
    case kModePred: launch_force<GABRIEL, BINNED, kModePred, false>(has_type, has_pol, a, n_blocks, stream); break;
    case kModeCorr: launch_force<GABRIEL, BINNED, kModeCorr, false>(has_type, has_pol, a, n_blocks, stream); break;
    case kModeEuler: launch_force<GABRIEL, BINNED, kModeEuler, false>(has_type, has_pol, a, n_blocks, stream); break;
    case kModeDeltaDrag: launch_force<GABRIEL, BINNED, kModeDeltaDrag, false>(has_type, has_pol, a, n_blocks, stream); break;
    case kModeCount: launch_force<GABRIEL, BINNED, kModeCount, false>(has_type, has_pol, a, n_blocks, stream); break;
    case kModeEmit: launch_force<GABRIEL, BINNED, kModeEmit, false>(has_type, has_pol, a, n_blocks, stream); break;
    default: return -1;
  }
  return 0;
}

template <bool GABRIEL, bool BINNED, bool HAS_TYPE, bool HAS_POL>
void launch_sym(const SweepArgs& a, int n_blocks, hipStream_t stream) {
  if (BINNED && a.skey && a.rel) {
    hipLaunchKernelGGL((k_sym_binned<GABRIEL, HAS_TYPE, HAS_POL, true, true>), dim3(n_blocks),
                       dim3(kBlock), 0, stream, a);
    hipLaunchKernelGGL((k_sym_movers<GABRIEL, HAS_TYPE, HAS_POL, true>), dim3(n_blocks),
                       dim3(kBlock), 0, stream, a);
  } else if (BINNED && a.skey) {
    hipLaunchKernelGGL((k_sym_binned<GABRIEL, HAS_TYPE, HAS_POL, true, false>), dim3(n_blocks),
                       dim3(kBlock), 0, stream, a);
    hipLaunchKernelGGL((k_sym_movers<GABRIEL, HAS_TYPE, HAS_POL, false>), dim3(n_blocks),
                       dim3(kBlock), 0, stream, a);
  }
  else if (BINNED && a.rel)
    hipLaunchKernelGGL((k_sym_binned<GABRIEL, HAS_TYPE, HAS_POL, false, true>), dim3(n_blocks),
                       dim3(kBlock), 0, stream, a);
  else if (BINNED)
    hipLaunchKernelGGL((k_sym_binned<GABRIEL, HAS_TYPE, HAS_POL, false, false>), dim3(n_blocks),
                       dim3(kBlock), 0, stream, a);
  else
    hipLaunchKernelGGL((k_sym_naive<GABRIEL, HAS_TYPE, HAS_POL>), dim3(n_blocks), dim3(kBlock), 0,
                       stream, a);
}

// Symmetric finders: Delta (the k_sym_* kernels), Count and Emit (the ordered walks restricted
// to partners j < i).
template <bool GABRIEL, bool BINNED>
int launch_mode_sym(int mode, bool has_type, bool has_pol, const SweepArgs& a, int n_blocks,
                    hipStream_t stream) {
  switch (mode) {
    case kModeDelta:
      if (has_type) {
        if (has_pol) launch_sym<GABRIEL, BINNED, true, true>(a, n_blocks, stream);
        else launch_sym<GABRIEL, BINNED, true, false>(a, n_blocks, stream);
      } else {
        if (has_pol) launch_sym<GABRIEL, BINNED, false, true>(a, n_blocks, stream);
        else launch_sym<GABRIEL, BINNED, false, false>(a, n_blocks, stream);
      }
      return 0;
    case kModeCount: launch_force<GABRIEL, BINNED, kModeCount, true>(has_type, has_pol, a, n_blocks, stream); return 0;
    case kModeEmit: launch_force<GABRIEL, BINNED, kModeEmit, true>(has_type, has_pol, a, n_blocks, stream); return 0;
    default: return -1;
  }
}

// d += q and q = 0 for every slot (the movers' contributions of a symmetric evaluation that no
// integrator update consumes).
__global__ __launch_bounds__(kBlock) void k_fold_q(float4* d, float2* dp, float4* q, float2* qp,
                                                   int total) {
  const int s = blockIdx.x * kBlock + (int)threadIdx.x;
  if (s >= total) return;
  const float4 v = q[s];
  const float4 x = d[s];
  d[s] = make_float4(x.x + v.x, x.y + v.y, x.z + v.z, x.w);
  q[s] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (dp) {
    const float2 w = qp[s];
    const float2 y = dp[s];
    dp[s] = make_float2(y.x + w.x, y.y + w.y);
    qp[s] = make_float2(0.0f, 0.0f);
  }
}

// ------------------------------------------------------------------ COM fix and updates
// Per-workgroup sums of the finished position deltas (the symmetric sweeps cannot sum them in
// their epilogue: other cells still add to a cell's delta until the sweep ends).
__global__ __launch_bounds__(kBlock) void k_com_partials(const InstanceDesc* inst,
                                                         const int* block_inst, const float4* d,
                                                         const float4* q, float4* partial) {
  __shared__ float lds[4];
  const int lb = logical_block();
  const InstanceDesc& in = inst[block_inst[lb]];
  const int s = lb * kBlock + (int)threadIdx.x;
  const bool active = (s - in.slot_begin) < in.n;
  float4 v = active ? d[s] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (active && q) {
    const float4 w = q[s];
    v = make_float4(v.x + w.x, v.y + w.y, v.z + w.z, 0.0f);
  }
  const float t0 = block_sum_f(v.x, lds);
  const float t1 = block_sum_f(v.y, lds);
  const float t2 = block_sum_f(v.z, lds);
  if (threadIdx.x == 0) partial[lb] = make_float4(t0, t1, t2, 0.0f);
}

// GlobalComFixer::fix, com_fixers.hpp:35-43: sum of the position deltas over the capacity,
// divided by the capacity; deterministic fixed-order reduction of the per-block partials.
__global__ __launch_bounds__(kBlock) void k_com_finalize(const InstanceDesc* inst,
                                                         const float4* partial, int per_block,
                                                         float4* com) {
  __shared__ float lds[4];
  const InstanceDesc& in = inst[blockIdx.x];
  float v[3] = {0.0f, 0.0f, 0.0f};
  const int first = in.block_begin * per_block, count = in.n_blocks * per_block;
  for (int b = threadIdx.x; b < count; b += kBlock) {
    const float4 p = partial[first + b];
    v[0] = v[0] + p.x;
    v[1] = v[1] + p.y;
    v[2] = v[2] + p.z;
  }
  const float t0 = block_sum_f(v[0], lds);
  const float t1 = block_sum_f(v[1], lds);
  const float t2 = block_sum_f(v[2], lds);
  if (threadIdx.x == 0) {
    const float n = (float)in.cap;
    com[blockIdx.x] = in.cap == 0 ? make_float4(0.0f, 0.0f, 0.0f, 0.0f)
                                  : make_float4(t0 / n, t1 / n, t2 / n, 0.0f);
  }
}

// The update of slot s (a cell); xnew: its new position (corrector, Euler).
template <int KIND, bool HAS_POL>
__device__ __forceinline__ void update_cell(const UpdateArgs& a, const InstanceDesc& in, int ii,
                                            int s, float4& xnew) {
  const float4 com = a.com[ii];
  const float dt = in.dt;
  // a symmetric evaluation on a stale grid: the movers' contributions in q join the evaluated
  // deltas (d1, or d3 for the corrector) and q is cleared for the next one
  float4 qv = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float2 ql = make_float2(0.0f, 0.0f);
  if (a.q_pos) {
    qv = a.q_pos[s];
    if (qv.x != 0.0f || qv.y != 0.0f || qv.z != 0.0f) a.q_pos[s] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (HAS_POL) {
      ql = a.q_pol[s];
      if (ql.x != 0.0f || ql.y != 0.0f) a.q_pol[s] = make_float2(0.0f, 0.0f);
    }
  }
  // every cell's state is read and written once here (streaming hints), except the predicted
  // positions and mover flags, which the next sweep's walks read many times
  const float4 d1_in = load_stream(a.d1_pos + s);
  float4 d1 = d1_in;
  if (KIND != kUpdateCorr && a.q_pos) d1 = make_float4(d1.x + qv.x, d1.y + qv.y, d1.z + qv.z, 0.0f);
  d1 = make_float4(d1.x - com.x, d1.y - com.y, d1.z - com.z, 0.0f);
  if (KIND == kUpdatePred) {
    // heun.hpp:150-154
    store_stream(a.d1_pos + s, d1);
    if (HAS_POL && a.q_pos) {
      const float2 e = a.d1_pol[s];
      a.d1_pol[s] = make_float2(e.x + ql.x, e.y + ql.y);
    }
    const float4 x = load_stream(a.x + s);
    const float4 xp = make_float4(x.x + d1.x * dt, x.y + d1.y * dt, x.z + d1.z * dt, x.w);
    a.xp[s] = xp;
    if (a.mover8) a.mover8[s] = bin_key(in, xp) != a.mover_key[s] ? 1 : 0;
    if (HAS_POL) {
      const float2 p = a.pol[s], e = a.d1_pol[s];
      const float2 np = make_float2(p.x + e.x * dt, p.y + e.y * dt);
      a.polp[s] = np;
      a.trigp[s] = polar_trig(np);
    }
  } else if (KIND == kUpdateCorr) {
    // heun.hpp:174-183 (d1 was COM-fixed by the predictor; com here is the corrector's)
    const float4 d1f = d1_in;
    float4 d3 = load_stream(a.d3_pos + s);
    if (a.q_pos) d3 = make_float4(d3.x + qv.x, d3.y + qv.y, d3.z + qv.z, 0.0f);
    d3 = make_float4(d3.x - com.x, d3.y - com.y, d3.z - com.z, 0.0f);
    const float s0 = d1f.x + d3.x, s1 = d1f.y + d3.y, s2 = d1f.z + d3.z;
    if (a.old_vel) store_stream(a.old_vel + s, make_float4(s0 * 0.5f, s1 * 0.5f, s2 * 0.5f, 0.0f));
    const float4 x = load_stream(a.x + s);
    xnew = make_float4(x.x + s0 * 0.5f * dt, x.y + s1 * 0.5f * dt, x.z + s2 * 0.5f * dt, x.w);
    store_stream(a.x + s, xnew);
    if (HAS_POL) {
      float2 e3 = a.d3_pol[s];
      e3 = make_float2(e3.x + ql.x, e3.y + ql.y);  // ql = 0 without q
      const float2 p = a.pol[s], e1 = a.d1_pol[s];
      const float2 np = make_float2(p.x + (e1.x + e3.x) * 0.5f * dt, p.y + (e1.y + e3.y) * 0.5f * dt);
      a.pol[s] = np;
      a.trig[s] = polar_trig(np);
    }
  } else {
    // euler.hpp:56-66
    if (a.old_vel) store_stream(a.old_vel + s, d1);
    const float4 x = load_stream(a.x + s);
    xnew = make_float4(x.x + d1.x * dt, x.y + d1.y * dt, x.z + d1.z * dt, x.w);
    store_stream(a.x + s, xnew);
    if (HAS_POL) {
      float2 e = a.d1_pol[s];
      e = make_float2(e.x + ql.x, e.y + ql.y);
      const float2 p = a.pol[s];
      const float2 np = make_float2(p.x + e.x * dt, p.y + e.y * dt);
      a.pol[s] = np;
      a.trig[s] = polar_trig(np);
    }
  }
}


template <int KIND, bool HAS_POL>
__global__ __launch_bounds__(kBlock) void k_update(UpdateArgs a) {
  const int lb = logical_block();
  const int ii = a.block_inst[lb];
  const InstanceDesc& in = a.inst[ii];
  const int s = lb * kBlock + (int)threadIdx.x;
  const bool cell = (s - in.slot_begin) < in.n;
  // the corrector and Euler updates count the next rebuild's keys when asked to, which every
  // lane of the wavefront takes part in
  const bool keys = KIND != kUpdatePred && a.kc.keys != nullptr;
  if (!cell && !keys) return;
  float4 xnew = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (cell) {
    update_cell<KIND, HAS_POL>(a, in, ii, s, xnew);
    if (KIND != kUpdatePred && a.nonfinite) {
      const float2 pn = HAS_POL ? a.pol[s] : make_float2(0.0f, 0.0f);
      if (not_finite(xnew, pn, HAS_POL)) a.nonfinite[ii] = 1;
    }
  }
  if (keys) count_key(in, a.kc, s, cell, s - in.slot_begin == in.n, xnew);
}

}  // namespace

int launch_sweep(int finder, bool has_type, bool has_pol, int mode, const SweepArgs& a,
                 int n_blocks, hipStream_t stream) {
  switch (finder) {
    case kBinnedAllPairs: return launch_mode<false, true>(mode, has_type, has_pol, a, n_blocks, stream);
    case kBinnedGabriel: return launch_mode<true, true>(mode, has_type, has_pol, a, n_blocks, stream);
    case kNaiveAllPairs: return launch_mode<false, false>(mode, has_type, has_pol, a, n_blocks, stream);
    case kNaiveGabriel: return launch_mode<true, false>(mode, has_type, has_pol, a, n_blocks, stream);
    case kBinnedAllPairsSym: return launch_mode_sym<false, true>(mode, has_type, has_pol, a, n_blocks, stream);
    case kBinnedGabrielSym: return launch_mode_sym<true, true>(mode, has_type, has_pol, a, n_blocks, stream);
    case kNaiveAllPairsSym: return launch_mode_sym<false, false>(mode, has_type, has_pol, a, n_blocks, stream);
    case kNaiveGabrielSym: return launch_mode_sym<true, false>(mode, has_type, has_pol, a, n_blocks, stream);
    case kNaiveDelaunay: return launch_delaunay(has_type, has_pol, mode, a, n_blocks, stream);
    default: return -1;
  }
}

void launch_fold_q(float4* d, float2* dp, float4* q, float2* qp, int total_slots,
                   hipStream_t stream) {
  hipLaunchKernelGGL(k_fold_q, dim3((total_slots + kBlock - 1) / kBlock), dim3(kBlock), 0, stream,
                     d, dp, q, qp, total_slots);
}

void launch_com_partials(const InstanceDesc* inst, const int* block_inst, const float4* d,
                         const float4* q, float4* partial, int n_blocks, hipStream_t stream) {
  hipLaunchKernelGGL(k_com_partials, dim3(n_blocks), dim3(kBlock), 0, stream, inst, block_inst, d,
                     q, partial);
}

void launch_com_finalize(const InstanceDesc* inst, int n_inst, const float4* partial,
                         int partials_per_block, float4* com, hipStream_t stream) {
  hipLaunchKernelGGL(k_com_finalize, dim3(n_inst), dim3(kBlock), 0, stream, inst, partial,
                     partials_per_block, com);
}

void launch_update(int kind, bool has_pol, const UpdateArgs& a, int n_blocks, hipStream_t stream) {
  const dim3 g(n_blocks), b(kBlock);
  if (kind == kUpdatePred) {
    if (has_pol) hipLaunchKernelGGL((k_update<kUpdatePred, true>), g, b, 0, stream, a);
    else hipLaunchKernelGGL((k_update<kUpdatePred, false>), g, b, 0, stream, a);
  } else if (kind == kUpdateCorr) {
    if (has_pol) hipLaunchKernelGGL((k_update<kUpdateCorr, true>), g, b, 0, stream, a);
    else hipLaunchKernelGGL((k_update<kUpdateCorr, false>), g, b, 0, stream, a);
  } else {
    if (has_pol) hipLaunchKernelGGL((k_update<kUpdateEuler, true>), g, b, 0, stream, a);
    else hipLaunchKernelGGL((k_update<kUpdateEuler, false>), g, b, 0, stream, a);
  }
}

}  // namespace tecs_amd
