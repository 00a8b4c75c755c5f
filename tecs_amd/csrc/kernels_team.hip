// kernels_team.hip - binned sweeps with a team of T lanes per cell (T = 4, 8 or 16).
//
// A wavefront holds 64 / T cells. The lanes of a cell's team split its work:
//   1. rows: the 9 neighbour rows (x-runs of bins, z outer, y inner - the order of
//      binned_all_pairs.hpp:117-138) are looked up by the team's lanes into an LDS row table;
//   2. candidates: the concatenated row ranges are the reference's visiting sequence; lane t
//      takes candidates t, t+T, ...; each computes d2 and the cutoff test;
//   3. compaction: accepted candidates go to LDS in visiting order (ballot + popcount);
//   4. Gabriel (binned_gabriel.hpp:133-209): each accepted pair is tested against the team's
//      accepted list. Every reference blocker k of (i, j) lies within (c/2 + 1/2) d(i, j) of i
//      (the ball has radius c d / 2 around the midpoint; c <= 0.95 gives < 0.98 d) and inside
//      the 27 bins around i, so it is an accepted candidate of i; the reference only counts k
//      whose bin (of the last rebuild) lies in the ball's bounding box, which is checked with
//      the candidate's row and x-bin. The result is the reference's blocked flag exactly;
//   5. forces of the surviving pairs are computed by their lanes and stored in LDS;
//   6. the cell's force components are summed in visiting order, one component per lane
//      (x, y, z and the polarity's two), so each sum is bitwise the serial one.
// A team whose accepted list overflows the LDS capacity (or, defensively, whose Gabriel box
// leaves the 27 bins) hands the cell to its leader, which walks it serially like the
// one-thread-per-cell kernel. Requires search_radius == 1 and, for Gabriel,
// gabriel_coefficient <= 0.95 (checked on the host).
//
// T trades lane utilisation against per-team overhead: the force phase runs one round per
// T accepted pairs, the row and candidate phases one per T rows / candidates. The host's
// autotuner times T = 4, 8, 16 (and the thread-per-cell sweep) on the live state.
//
// Symmetric finders (SYM, Delta mode; the roles and the scatter rule used here are described
// in walks.hpp): an accepted
// candidate is an own pair (smaller agent index; scattered when a mover is involved), a gather
// pair (larger index, both cells stayers) or only a potential blocker (a pair with a mover
// that the other cell delivers). The ball of a gather pair (m, i) is the one m's walk tests;
// it lies within (c/2 + 1/2) d(m, i) of i as well, so its blockers are accepted candidates of
// i too.
#include <type_traits>

#include "tecs_internal.hpp"
#include "walks.hpp"

namespace tecs_amd {
namespace {

using namespace walks;

// Accepted neighbours held in LDS per cell: a Gabriel cell on a sheet or in a tissue has ~5-8
// neighbours within the cutoff, an all-pairs cell in a dense tissue ~6-10; a cell with more
// takes the serial path.
template <bool GABRIEL>
struct TeamCap {
  static constexpr int kAMax = GABRIEL ? 16 : 24;
};

// SPLIT: an ordered sweep whose force can have two terms per field (polarity models): each
// pair's terms are kept apart, so the leader adds them to the cell's running sums one by one in
// the composition's order, as the reference's fused forces add to one pack
// (kernel_fuser.hpp:40-55). The symmetric finders evaluate a pair on a zeroed pack and add the
// pack (binned_all_pairs_symmetric.hpp:176-200), so they keep one value per pair.
template <int T, bool GABRIEL, bool HAS_POL, bool SPLIT = false>
struct TeamLds {
  static constexpr int TB = kBlock / T;  // teams (cells) per workgroup
  static constexpr int AM = TeamCap<GABRIEL>::kAMax;
  static constexpr int TG = GABRIEL ? TB : 1;
  static constexpr int TP = HAS_POL ? TB : 1;
  // all-pairs computes forces while the rows are still read, so its second terms need arrays
  // of their own; Gabriel computes them after the rows are dead and keeps them there
  static constexpr int TS = SPLIT && !GABRIEL ? TB : 1;
  // per accepted neighbour, in visiting order:
  //   ent: Gabriel - its position (w = type bits) until the Gabriel tests are done, then
  //        the pair's position contribution; all-pairs - the contribution. w = 1 when the
  //        pair counts.
  // Rows are padded by one entry: the teams of a wavefront read the same entry index of
  // their own rows at once, and an unpadded row stride (a multiple of 128 bytes for ent and
  // cpol) would put all of them in the same LDS banks.
  float4 ent[TB][AM + 1];
  int aslot[TB][AM + 1];    // its slot
  int abit[TG][AM + 1];     // one-hot bit rx + 3 ry + 9 rz of its (stale) bin relative to
                            // i's centre bin - 1 (r* in 0..2)
  float2 cpol[TP][AM + 1];  // polarity contribution (SPLIT: its first term)
  float4 ent2[TS][AM + 1];  // SPLIT all-pairs: the second position term (xyz)
  float2 cpol2[TS][AM + 1]; // SPLIT all-pairs: the second polarity term
  // The 9 neighbour rows of each team's cell. A Gabriel SPLIT force phase, which runs after the
  // last read of the rows, keeps each entry's second position term in their place (and the
  // second polarity term in the entry's aslot / abit, which are dead by then as well).
  struct RowsG {
    int start[9];
    int len[9];
    int base[9];  // start - the row's first candidate index (slot = base + c)
    int b1[9];    // first slot of the row's second bin
    int b2[9];    // first slot of the row's third bin
    int xlo[9];
  };
  struct RowsA {
    int start[9];
    int len[9];
    int base[9];
    int b1[1], b2[1], xlo[1];  // unused without Gabriel
  };
  union TeamRows {
    typename std::conditional<GABRIEL, RowsG, RowsA>::type r;
    float p2[GABRIEL ? 3 * AM : 1];
  };
  static_assert(!GABRIEL || sizeof(RowsG) >= sizeof(float) * 3 * AM, "second terms fit the rows");
  TeamRows rows[TB];
  int unsafe[TB];
};

// ---------------------------------------------------------------- serial fallback (leader)
template <bool GABRIEL, bool HAS_TYPE, bool HAS_POL, int MODE, bool ROWS>
__device__ void serial_cell(const SweepArgs& a, const InstanceDesc& in, int s, float4 me,
                            const Pol& pme, float* acc_pos, float* acc_pol, unsigned& pairs) {
  int64_t emit_at = 0;
  if (MODE == kModeEmit) emit_at = a.row_ptr[a.id[s]];
  const float4* __restrict__ xs = a.xs;
  const BinRanges<ROWS> offb = bin_ranges<ROWS>(in, a.off, a.rel);
  int bc[3];
  if (GABRIEL)
    bin_coords_div(in, me.x, me.y, me.z, bc);
  else
    bin_coords_mul(in, me.x, me.y, me.z, bc);
  const int x_lo = max(bc[0] - 1, 0), x_hi = min(bc[0] + 1, in.ext[0] - 1);
  for (int dz = -1; dz <= 1; ++dz) {
    const int z = bc[2] + dz;
    if (z < 0 || z >= in.ext[2]) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      const int y = bc[1] + dy;
      if (y < 0 || y >= in.ext[1]) continue;
      int start, end;
      offb.row(x_lo, x_hi, y, z, start, end);
      for (int m = start; m < end; ++m) {
        if (m == s) continue;
        const float4 pj = xs[m];
        const float disp[3] = {me.x - pj.x, me.y - pj.y, me.z - pj.z};
        const float d2 = len2(disp[0], disp[1], disp[2]);
        if (d2 >= in.cutoff2) continue;
        if (GABRIEL) {
          const float mid0 = me.x - disp[0] * 0.5f, mid1 = me.y - disp[1] * 0.5f,
                      mid2 = me.z - disp[2] * 0.5f;
          const float rho2 = d2 * in.gabriel_factor;
          const float rho = sqrtf(rho2);
          const float mid[3] = {mid0, mid1, mid2};
          int mn[3], mx[3];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            mn[d] = clampi((int)floorf((mid[d] - rho - in.minb[d]) * in.inv_bin), in.ext[d] - 1);
            mx[d] = clampi((int)floorf((mid[d] + rho - in.minb[d]) * in.inv_bin), in.ext[d] - 1);
          }
          bool blocked = false;
          for (int b2 = mn[2]; b2 <= mx[2] && !blocked; ++b2)
            for (int b1 = mn[1]; b1 <= mx[1] && !blocked; ++b1) {
              int s2, e2;
              offb.row(mn[0], mx[0], b1, b2, s2, e2);
              for (int k = s2; k < e2; ++k) {
                if (k == s || k == m) continue;
                const float4 pk = xs[k];
                if (len2(pk.x - mid0, pk.y - mid1, pk.z - mid2) < rho2) {
                  blocked = true;
                  break;
                }
              }
            }
          if (blocked) continue;
        }
        pairs++;
        if (MODE == kModeCount) continue;
        if (MODE == kModeEmit) {
          a.cols[emit_at++] = a.id[m];
          continue;
        }
        const float d = sqrtf(d2);
        Pol po{};
        if (HAS_POL) po = make_pol(a.ps[m], a.trig[m]);
        apply_force<HAS_TYPE, HAS_POL>(in.force, disp, d, __float_as_int(me.w),
                                       __float_as_int(pj.w), pme, po, acc_pos, acc_pol);
      }
    }
  }
}

__device__ __forceinline__ float tblock_sum_f(float v, float* lds) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) lds[w] = v;
  __syncthreads();
  float r = 0.0f;
  if (threadIdx.x == 0) r = ((lds[0] + lds[1]) + lds[2]) + lds[3];
  return r;
}
__device__ __forceinline__ unsigned wave_sum_u(unsigned v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// A team's LDS rows are touched only by the team's lanes and a team never spans two
// wavefronts (T divides 64), so the phases synchronise the wavefront, not the workgroup: the
// fences wait for the wavefront's outstanding LDS accesses and keep the compiler from moving
// LDS accesses across this point; the four wavefronts of a workgroup run their cells
// independently.
__device__ __forceinline__ void team_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// abit bits 28-29: the entry's role in a symmetric sweep (SymRole; 0 = blocker only)
constexpr int kRoleShift = 28;

template <int T, bool GABRIEL, bool HAS_TYPE, bool HAS_POL, int MODE, bool SYM, bool ROWS>
__global__ __launch_bounds__(kBlock, 6) void k_sweep_team(SweepArgs a) {
  static_assert(!SYM || MODE == kModeDelta, "symmetric team sweeps evaluate forces only");
  constexpr bool SPLIT = HAS_POL && !SYM;
  // the cell's force components: x, y, z and, with polarity, the polarity's two
  constexpr int NC = HAS_POL ? 5 : 3;
  // the epilogue by component (each lane updates the component it summed) for the integrator
  // modes; Delta, the symmetric sweeps' mode, keeps the leader's (measured faster there)
  constexpr bool LANE_EPI = T >= NC && (MODE == kModePred || MODE == kModeCorr || MODE == kModeEuler);
  constexpr int TB = TeamLds<T, GABRIEL, HAS_POL, SPLIT>::TB;
  constexpr int AM = TeamLds<T, GABRIEL, HAS_POL, SPLIT>::AM;
  __shared__ TeamLds<T, GABRIEL, HAS_POL, SPLIT> L;
  __shared__ float lds_f[4];

  // team-block -> slots: each 256-slot chunk (one instance) holds T team-blocks;
  // XCD-aware remap (see logical_block): contiguous runs of team-blocks per XCD
  const int p = (int)blockIdx.x;
  const int nb = (int)gridDim.x;
  const int q = nb >> 3, r = nb & 7, xcd = p & 7, idx = p >> 3;
  const int tl = (xcd < r) ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
  // a launch over part of the batch: team-block tl is team-block (tl mod T) of listed chunk tl / T
  const int tb = a.chunks ? a.chunks[tl / T] * T + tl % T : tl;
  const int team = (int)threadIdx.x / T;
  const int lane = (int)threadIdx.x % T;
  const int s = tb * TB + team;
  // the 256-slot chunk, one instance; TB divides kBlock, so every cell of the workgroup is in
  // the chunk of its first one (computed from that, the instance and its descriptor stay in
  // scalar registers)
  static_assert(kBlock % TB == 0, "a workgroup's cells lie in one chunk");
  const int chunk = (tb * TB) / kBlock;
  const int ii = a.block_inst[chunk];
  const InstanceDesc& in = a.inst[ii];
  const bool active = (s - in.slot_begin) < in.n;

  float4 me = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float2 mp = make_float2(0.0f, 0.0f);
  Pol pme{};
  int bc[3] = {0, 0, 0};
  int me_id = 0;
  bool mover = false;  // symmetric: this cell left its rebuild bin
  // corrector inputs: the leader's, or with the epilogue by component the lane's (d1 / e1 of
  // its force component, x0 of lanes 0-3 (lane 3: w), p0 of the polarity lanes)
  float4 cin_d1 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), cin_x0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float2 cin_e1 = make_float2(0.0f, 0.0f), cin_p0 = make_float2(0.0f, 0.0f);
  float cin_q = 0.0f, cin_x = 0.0f, cin_p = 0.0f;
  if (active) {
    me = a.xs[s];
    if (HAS_POL) {
      mp = a.ps[s];
      pme = make_pol(mp, a.trig[s]);
    }
    if (GABRIEL || SYM)
      bin_coords_div(in, me.x, me.y, me.z, bc);  // binned_gabriel.hpp:101, *_symmetric.hpp
    else
      bin_coords_mul(in, me.x, me.y, me.z, bc);  // binned_all_pairs.hpp:103-109
    if (SYM) {
      me_id = a.id[s];
      mover = a.skey != nullptr && flatten(in, bc[0], bc[1], bc[2]) != a.skey[s];
    }
    // the corrector's inputs, loaded now so that their latency hides behind the sweep rather
    // than stalling the epilogue (by component with LANE_EPI, else by the leader)
    if (MODE == kModeCorr) {
      if constexpr (LANE_EPI) {
        if (lane < 3)
          cin_q = reinterpret_cast<const float*>(a.d1_pos + s)[lane];
        else if (HAS_POL && lane < 5)
          cin_q = reinterpret_cast<const float*>(a.d1_pol + s)[lane - 3];
        if (lane < 4) cin_x = reinterpret_cast<const float*>(a.o2_pos + s)[lane];
        if (HAS_POL && lane >= 3 && lane < 5)
          cin_p = reinterpret_cast<const float*>(a.o2_pol + s)[lane - 3];
      } else if (lane == 0) {
        cin_d1 = a.d1_pos[s];
        cin_x0 = a.o2_pos[s];
        if (HAS_POL) {
          cin_e1 = a.d1_pol[s];
          cin_p0 = a.o2_pol[s];
        }
      }
    }
    // row rr = (dz + 1) * 3 + (dy + 1), looked up by lane rr mod T
    const int x_lo = max(bc[0] - 1, 0), x_hi = min(bc[0] + 1, in.ext[0] - 1);
    const BinRanges<ROWS> offb = bin_ranges<ROWS>(in, a.off, a.rel);
    for (int rr = lane; rr < 9; rr += T) {
      const int z = bc[2] + rr / 3 - 1, y = bc[1] + rr % 3 - 1;
      int start = 0, len = 0, b1 = 0x7fffffff, b2 = 0x7fffffff;
      if (z >= 0 && z < in.ext[2] && y >= 0 && y < in.ext[1]) {
        int end = 0;
        offb.row(x_lo, x_hi, y, z, start, end);
        len = end - start;
        if (GABRIEL) {
          if (x_lo + 1 <= x_hi) b1 = offb.at(x_lo + 1, y, z);
          if (x_lo + 2 <= x_hi) b2 = offb.at(x_lo + 2, y, z);
        }
      }
      L.rows[team].r.start[rr] = start;
      L.rows[team].r.len[rr] = len;
      if (GABRIEL) {
        L.rows[team].r.b1[rr] = b1;
        L.rows[team].r.b2[rr] = b2;
        L.rows[team].r.xlo[rr] = x_lo;
      }
    }
  }
  if (lane == 0) L.unsafe[team] = 0;
  team_sync();

  int row_pre[9], row_start[9];
  int total = 0;
#pragma unroll
  for (int rr = 0; rr < 9; ++rr) {
    row_pre[rr] = total;
    row_start[rr] = L.rows[team].r.start[rr];
    total += L.rows[team].r.len[rr];
  }
  if (!active) total = 0;
  if (lane == 0) {
#pragma unroll
    for (int rr = 0; rr < 9; ++rr) L.rows[team].r.base[rr] = row_start[rr] - row_pre[rr];
  }
  team_sync();
  const float cut2 = in.cutoff2;

  // candidates -> accepted list in LDS, in visiting order
  int base = 0;
  // pairs this lane's entries count for the reference (ordered, or own), kept when no serial
  // walk replaces the team's list
  unsigned lane_pairs = 0;
  const int team_shift = ((int)(threadIdx.x & 63) / T) * T;
  const unsigned team_bits = (T == 32) ? 0xffffffffu : ((1u << T) - 1u);
  const int n_iter = (total + T - 1) / T;
  // candidate c of iteration it: its row and slot (-1 past the lane's last candidate); the
  // row holding c is the last row whose prefix is <= c, i.e. the number of rows 1..8 whose
  // prefix is <= c (prefixes never decrease; an empty row shares its prefix with the next one)
  auto candidate = [&](int it, int& rsel) {
    const int c = it * T + lane;
    rsel = 0;
    if (c >= total) return -1;
#pragma unroll
    for (int rr = 1; rr < 9; ++rr) rsel += c >= row_pre[rr] ? 1 : 0;
    return L.rows[team].r.base[rsel] + c;
  };
  // each iteration loads the next one's candidate position before it works on its own, so
  // that the load's latency hides behind this iteration's tests and LDS stores
  int rsel_next = 0;
  int m_next = candidate(0, rsel_next);
  float4 pj_next = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (m_next >= 0 && m_next != s) pj_next = a.xs[m_next];
  for (int it = 0; it < n_iter; ++it) {
    bool acc = false;
    const int m = m_next, rr_sel = rsel_next;
    float4 pj = pj_next;
    m_next = candidate(it + 1, rsel_next);
    if (m_next >= 0 && m_next != s) pj_next = a.xs[m_next];
    float disp[3] = {0.0f, 0.0f, 0.0f};
    float d2 = 0.0f;
    if (m >= 0) {
      if (m != s) {
        disp[0] = me.x - pj.x;
        disp[1] = me.y - pj.y;
        disp[2] = me.z - pj.z;
        d2 = len2(disp[0], disp[1], disp[2]);
        acc = !(d2 >= cut2);
      }
    }
    int role = kRoleOwn;
    if (SYM && acc) {
      // the scatter rule: a pair with a mover goes through its larger-index cell
      const bool other_mover = is_mover(in, a.skey, a.mover8, m, pj);
      if (a.id[m] < me_id)
        role = mover || other_mover ? kRoleOwnScatter : kRoleOwn;
      else
        role = mover || other_mover ? kRoleNone : kRoleGather;
      // all-pairs has no blockers: a candidate without a role is dropped
      if (!GABRIEL && role == kRoleNone) acc = false;
    }
    const unsigned long long bal = __ballot(acc);
    const unsigned tm = (unsigned)(bal >> team_shift) & team_bits;
    const int pos = base + __popc(tm & ((1u << lane) - 1u));
    if (acc && pos < AM) {
      L.aslot[team][pos] = m;
      if (!GABRIEL) lane_pairs += role != kRoleGather;
      if (GABRIEL) {
        // stale bin of m relative to (bc - 1): the row gives y and z, the row's bin
        // boundaries give x
        const int bx = L.rows[team].r.xlo[rr_sel] + (m >= L.rows[team].r.b1[rr_sel] ? 1 : 0) +
                       (m >= L.rows[team].r.b2[rr_sel] ? 1 : 0);
        L.ent[team][pos] = pj;
        // row rr_sel = 3 rz + ry, so 3 ry + 9 rz = 3 rr_sel
        L.abit[team][pos] = (1 << ((bx - bc[0] + 1) + 3 * rr_sel)) | (SYM ? role << kRoleShift : 0);
      } else if (SYM) {
        // all-pairs, symmetric: own v(i, m) or gather -v(m, i), computed here
        float cp[3] = {0.0f, 0.0f, 0.0f}, cl[2] = {0.0f, 0.0f};
        sym_eval_at<HAS_TYPE, HAS_POL>(a, in, me, pme, m, pj, role, cp, cl);
        L.ent[team][pos] = make_float4(cp[0], cp[1], cp[2], role == kRoleGather ? 2.0f : 1.0f);
        if (HAS_POL) L.cpol[team][pos] = make_float2(cl[0], cl[1]);
      } else {
        // all-pairs: the force needs nothing else, compute it here
        float cp[3] = {0.0f, 0.0f, 0.0f}, cl[2] = {0.0f, 0.0f};
        float cp2[3] = {0.0f, 0.0f, 0.0f}, cl2[2] = {0.0f, 0.0f};
        if (MODE != kModeCount && MODE != kModeEmit) {
          const float d = sqrtf(d2);
          Pol po{};
          if (HAS_POL) po = make_pol(a.ps[m], a.trig[m]);
          if (SPLIT)
            apply_force_terms<HAS_TYPE, HAS_POL>(in.force, disp, d, __float_as_int(me.w),
                                                 __float_as_int(pj.w), pme, po, cp, cp2, cl, cl2);
          else
            apply_force<HAS_TYPE, HAS_POL>(in.force, disp, d, __float_as_int(me.w),
                                           __float_as_int(pj.w), pme, po, cp, cl);
        }
        L.ent[team][pos] = make_float4(cp[0], cp[1], cp[2], 1.0f);
        if (HAS_POL) L.cpol[team][pos] = make_float2(cl[0], cl[1]);
        if (SPLIT) {
          L.ent2[team][pos] = make_float4(cp2[0], cp2[1], cp2[2], 0.0f);
          L.cpol2[team][pos] = make_float2(cl2[0], cl2[1]);
        }
      }
    }
    base += __popc(tm);
  }
  const bool overflow = base > AM;
  const int na = min(base, AM);
  team_sync();

  // Gabriel tests of every accepted pair (bit r of `blocked_bits`: entry lane + r T)
  unsigned blocked_bits = 0u;
  if (GABRIEL && active && !overflow) {
    int r = 0;
    for (int e = lane; e < na; e += T, ++r) {
      const float4 pj = L.ent[team][e];
      const int erole = SYM ? (L.abit[team][e] >> kRoleShift) : kRoleOwn;
      if (erole == kRoleNone) {  // a blocker only: no pair to test
        blocked_bits |= 1u << r;
        continue;
      }
      // the ball as the pair's larger-index cell tests it: i's for an own pair, m's for a
      // gather pair (mid = p_m - (p_m - p_i) * 0.5; d2 is the same either way)
      const float4 ps = erole != kRoleGather ? me : pj;
      const float4 po = erole != kRoleGather ? pj : me;
      const float disp[3] = {ps.x - po.x, ps.y - po.y, ps.z - po.z};
      const float d2 = len2(disp[0], disp[1], disp[2]);
      const float mid0 = ps.x - disp[0] * 0.5f;
      const float mid1 = ps.y - disp[1] * 0.5f;
      const float mid2 = ps.z - disp[2] * 0.5f;
      const float rho2 = d2 * in.gabriel_factor;
      const float rho = sqrtf(rho2);
      // bounding-box bins (binned_gabriel.hpp:142-153), relative to bc - 1
      const float mid[3] = {mid0, mid1, mid2};
      int dmask[3];  // per dimension: bits lo..hi of the box
      bool out_of_reach = false;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int mn =
            clampi((int)floorf((mid[d] - rho - in.minb[d]) * in.inv_bin), in.ext[d] - 1) - bc[d] + 1;
        const int mx =
            clampi((int)floorf((mid[d] + rho - in.minb[d]) * in.inv_bin), in.ext[d] - 1) - bc[d] + 1;
        out_of_reach = out_of_reach || mn < 0 || mx > 2;
        const int lo = max(mn, 0), hi = min(mx, 2);
        dmask[d] = lo <= hi ? ((2 << hi) - (1 << lo)) : 0;
      }
      if (out_of_reach) L.unsafe[team] = 1;
      // the box as a set of the 27 relative bins (bit rx + 3 ry + 9 rz): the x mask repeated in
      // the three y groups, ANDed with 3-bit runs at the y bits, repeated in the three z groups,
      // ANDed with 9-bit runs at the z bits (all products fit 24-bit multiplies)
      const unsigned dx = (unsigned)dmask[0], dy = (unsigned)dmask[1], dz = (unsigned)dmask[2];
      const unsigned x9 = (dx & 7u) * 0x49u;                                          // bits x + 3 y
      const unsigned y9 = ((dy & 1u) | ((dy & 2u) << 2) | ((dy & 4u) << 4)) * 7u;  // 3 ones per y
      const unsigned z27 = ((dz & 1u) | ((dz & 2u) << 8) | ((dz & 4u) << 16)) * 0x1ffu;
      const int box = (int)(((x9 & y9) * 0x40201u) & z27);
      bool blocked = false;
      for (int k = 0; k < na; ++k) {
        const int bits = SYM ? (L.abit[team][k] & ((1 << 27) - 1)) : L.abit[team][k];
        const float4 pk = L.ent[team][k];
        const bool in_box = k != e && (bits & box) != 0;
        const bool in_ball = len2(pk.x - mid0, pk.y - mid1, pk.z - mid2) < rho2;
        blocked = blocked | (in_box & in_ball);
      }
      if (blocked) blocked_bits |= 1u << r;
    }
  }
  team_sync();

  // forces of the surviving pairs; each lane overwrites only its own entries
  if (GABRIEL && active && !overflow) {
    int r = 0;
    for (int e = lane; e < na; e += T, ++r) {
      const float4 pj = L.ent[team][e];
      const bool blocked = (blocked_bits >> r) & 1u;
      float cp[3] = {0.0f, 0.0f, 0.0f}, cl[2] = {0.0f, 0.0f};
      float cp2[3] = {0.0f, 0.0f, 0.0f}, cl2[2] = {0.0f, 0.0f};
      const int erole = SYM ? (L.abit[team][e] >> kRoleShift) : kRoleOwn;
      if (SYM && !blocked) {
        // the entry holds m's position (written by the candidate phase)
        sym_eval_at<HAS_TYPE, HAS_POL>(a, in, me, pme, L.aslot[team][e], pj, erole, cp, cl);
      } else if (!blocked && MODE != kModeCount && MODE != kModeEmit) {
        const int m = L.aslot[team][e];
        const float disp[3] = {me.x - pj.x, me.y - pj.y, me.z - pj.z};
        const float d = sqrtf(len2(disp[0], disp[1], disp[2]));
        Pol po{};
        if (HAS_POL) po = make_pol(a.ps[m], a.trig[m]);
        if (SPLIT)
          apply_force_terms<HAS_TYPE, HAS_POL>(in.force, disp, d, __float_as_int(me.w),
                                               __float_as_int(pj.w), pme, po, cp, cp2, cl, cl2);
        else
          apply_force<HAS_TYPE, HAS_POL>(in.force, disp, d, __float_as_int(me.w),
                                         __float_as_int(pj.w), pme, po, cp, cl);
      }
      // w: 0 no pair, 1 a pair the reference counts (ordered, or own), 2 a gathered pair;
      // a blocked entry keeps zero terms, which the ordered sums add as the identity
      lane_pairs += !blocked && erole != kRoleGather;
      L.ent[team][e] = make_float4(cp[0], cp[1], cp[2],
                                   blocked ? 0.0f : (erole == kRoleGather ? 2.0f : 1.0f));
      if (HAS_POL) L.cpol[team][e] = make_float2(cl[0], cl[1]);
      if (SPLIT) {
        // the rows were last read by the candidate phase and aslot / abit of this entry by
        // this lane above and by the Gabriel tests before the last team_sync
        L.rows[team].p2[3 * e + 0] = cp2[0];
        L.rows[team].p2[3 * e + 1] = cp2[1];
        L.rows[team].p2[3 * e + 2] = cp2[2];
        L.aslot[team][e] = __float_as_int(cl2[0]);
        L.abit[team][e] = __float_as_int(cl2[1]);
      }
    }
  }
  team_sync();

  // Ordered sums. The cell's force is NC independent sums (x, y, z and, with polarity, the
  // two polarity components), each over the list in visiting order with a pair's first term
  // before its second (kernel_fuser.hpp:40-55). Lane q of the team runs sum q (and q + T when
  // T < NC), so each add is one of the reference's adds in its order; the lane then updates
  // that component of the cell (LANE_EPI), or the leader collects the sums. Entries without a
  // pair hold zero terms: a running sum that starts at +0 is
  // never -0, so adding +0 leaves it unchanged, as skipping the entry would. After an
  // overflow or an unsafe Gabriel box the leader walks the cell serially instead.
  constexpr int NQ = (NC + T - 1) / T;  // sums per lane
  float acc_pos[3] = {0.0f, 0.0f, 0.0f}, acc_pol[2] = {0.0f, 0.0f};
  const bool leader = lane == 0 && active;
  const bool serial = active && (overflow || L.unsafe[team]);
  unsigned pairs = active && !serial ? lane_pairs : 0u;
  if (leader && serial) {
    if (SYM)
      sym_cell_serial<GABRIEL, HAS_TYPE, HAS_POL, true, ROWS>(a, in, s, me_id, mover, me, pme, bc,
                                                        acc_pos, acc_pol, pairs);
    else
      serial_cell<GABRIEL, HAS_TYPE, HAS_POL, MODE, ROWS>(a, in, s, me, pme, acc_pos, acc_pol, pairs);
  }
  if (MODE == kModeEmit) {
    if (leader && !serial) {
      int64_t emit_at = a.row_ptr[a.id[s]];
      for (int e = 0; e < na; ++e)
        if (L.ent[team][e].w != 0.0f) a.cols[emit_at++] = a.id[L.aslot[team][e]];
    }
    return;
  }
  if (MODE == kModeCount) {
#pragma unroll
    for (int w = 1; w < T; w <<= 1) pairs += __shfl_xor(pairs, w, 64);
    if (leader) a.row_count[a.id[s]] = (int)pairs;
    return;
  }
  float lane_sum[NQ];
#pragma unroll
  for (int j = 0; j < NQ; ++j) {
    const int q = lane + j * T;
    lane_sum[j] = 0.0f;
    if (active && !serial && q < NC) {
      // the first term of sum q at entry e is t1[e * s1], the second t2[e * s2]
      const float* t1 = !HAS_POL || q < 3 ? &L.ent[team][0].x + q : &L.cpol[team][0].x + (q - 3);
      const int s1 = q < 3 ? 4 : 2;
      const float* t2 = nullptr;
      int s2 = 0;
      if (SPLIT && GABRIEL) {
        t2 = q < 3 ? &L.rows[team].p2[q]
                   : reinterpret_cast<const float*>(q == 3 ? &L.aslot[team][0] : &L.abit[team][0]);
        s2 = q < 3 ? 3 : 1;
      } else if (SPLIT) {
        t2 = q < 3 ? &L.ent2[team][0].x + q : &L.cpol2[team][0].x + (q - 3);
        s2 = q < 3 ? 4 : 2;
      }
      float acc = 0.0f;
      for (int e = 0; e < na; ++e) {
        acc = acc + t1[e * s1];
        if (SPLIT) acc = acc + t2[e * s2];
      }
      lane_sum[j] = acc;
    }
  }
  const int base_lane = (int)(threadIdx.x & 63) - lane;
  float F[3] = {0.0f, 0.0f, 0.0f};  // the leader's force sums where the epilogue needs them
  if constexpr (LANE_EPI) {
    // Epilogue by component (Pred / Corr / Euler): lane q < NC holds sum q and updates that
    // component of the cell (lanes 0-2 the position, 3-4 the polarity; lane 3 also writes the
    // position's w), with the same arithmetic the integrators apply per component.
    float fq = lane_sum[0];
    if (__ballot(serial) != 0ull) {
      const float la[5] = {acc_pos[0], acc_pos[1], acc_pos[2], acc_pol[0], acc_pol[1]};
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const float v = __shfl(la[c], base_lane, 64);
        if (serial && lane == c) fq = v;
      }
    }
    // out(i) += total, then the +0 drag correction (binned_all_pairs.hpp:174-180)
    fq = lane < 3 ? (0.0f + fq) + 0.0f : 0.0f + fq;
    const bool pos_lane = active && lane < 3;
    const bool w_lane = active && lane == 3;
    const bool pol_lane = HAS_POL && active && lane >= 3 && lane < 5;
    const int pc = lane - 3;  // a polarity lane's component
    const float me_q = lane == 0 ? me.x : lane == 1 ? me.y : lane == 2 ? me.z : me.w;
    const float mp_c = lane == 3 ? mp.x : mp.y;
    const float dt = in.dt;
    float* const pos_out = reinterpret_cast<float*>(MODE == kModeCorr ? a.o2_pos + s : a.o_pos + s);
    float xq = 0.0f;   // the new position's component (lanes 0-2)
    float npc = 0.0f;  // the new polarity's component (lanes 3-4)
    if (MODE == kModePred) {
      // heun.hpp:146-156
      if (pos_lane || w_lane) {
        pos_out[lane] = pos_lane ? fq : 0.0f;
        reinterpret_cast<float*>(a.o2_pos + s)[lane] = pos_lane ? me_q + fq * dt : me.w;
      }
      if (pol_lane) {
        reinterpret_cast<float*>(a.o_pol + s)[pc] = fq;
        npc = mp_c + fq * dt;
        reinterpret_cast<float*>(a.o2_pol + s)[pc] = npc;
      }
    } else if (MODE == kModeCorr) {
      // heun.hpp:170-186
      const float sq = cin_q + fq;
      if (a.old_vel && (pos_lane || w_lane))
        reinterpret_cast<float*>(a.old_vel + s)[lane] = pos_lane ? sq * 0.5f : 0.0f;
      xq = cin_x + sq * 0.5f * dt;
      if (pos_lane || w_lane) pos_out[lane] = pos_lane ? xq : cin_x;
      if (pol_lane) {
        npc = cin_p + (cin_q + fq) * 0.5f * dt;
        reinterpret_cast<float*>(a.o2_pol + s)[pc] = npc;
      }
    } else if (MODE == kModeEuler) {
      // euler.hpp:52-68
      if (a.old_vel && (pos_lane || w_lane))
        reinterpret_cast<float*>(a.old_vel + s)[lane] = pos_lane ? fq : 0.0f;
      xq = me_q + fq * dt;
      if (pos_lane || w_lane) pos_out[lane] = pos_lane ? xq : me.w;
      if (pol_lane) {
        npc = mp_c + fq * dt;
        reinterpret_cast<float*>(a.o_pol + s)[pc] = npc;
      }
    }
    if (MODE != kModePred && a.nonfinite &&
        ((pos_lane && !isfinite(xq)) || (pol_lane && !isfinite(npc))))
      a.nonfinite[ii] = 1;
    if (HAS_POL && pol_lane) {
      // polar_trig of the new polarity: each polarity lane its angle's sine and cosine
      float sv, cv;
      sincosf(npc, &sv, &cv);
      reinterpret_cast<float2*>(a.o_trig + s)[pc] = make_float2(sv, cv);
    }
    // keys of the new positions for the rebuild that starts the next step: lane d < 3 bins
    // dimension d (Grid::rebuild's key, grid.hpp:394-395)
    if (MODE != kModePred && a.kc.keys) {
      const int d = min(lane, 2);
      const int cd = bin_coord_div(in, xq, d == 0 ? in.minb[0] : d == 1 ? in.minb[1] : in.minb[2],
                                   d == 0 ? in.ext[0] : d == 1 ? in.ext[1] : in.ext[2]);
      const int c1 = __shfl(cd, base_lane + 1, 64), c2 = __shfl(cd, base_lane + 2, 64);
      count_key_of(in, a.kc, s, leader, lane == 0 && s - in.slot_begin == in.n,
                   flatten(in, cd, c1, c2));
    }
  } else {
    // the leader collects the sums and updates the cell
    float got[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) got[q] = __shfl(lane_sum[q / T], base_lane + q % T, 64);
    if (leader && !serial) {
      acc_pos[0] = got[0];
      acc_pos[1] = got[1];
      acc_pos[2] = got[2];
      if (HAS_POL) {
        acc_pol[0] = got[NC - 2];
        acc_pol[1] = got[NC - 1];
      }
    }
    // out(i) += total, then the +0 drag correction (binned_all_pairs.hpp:174-180)
    float G[2];
#pragma unroll
    for (int c = 0; c < 3; ++c) F[c] = (0.0f + acc_pos[c]) + 0.0f;
    G[0] = 0.0f + acc_pol[0];
    G[1] = 0.0f + acc_pol[1];
    const float dt = in.dt;
    float4 xnew = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // Corr / Euler: the new position
    float2 npol = make_float2(0.0f, 0.0f);               // Pred / Corr / Euler: the new polarity
    if (leader) {
      if (MODE == kModeDelta) {
        a.o_pos[s] = make_float4(F[0], F[1], F[2], 0.0f);
        if (HAS_POL) a.o_pol[s] = make_float2(G[0], G[1]);
      } else if (MODE == kModePred) {
        // heun.hpp:146-156
        a.o_pos[s] = make_float4(F[0], F[1], F[2], 0.0f);
        a.o2_pos[s] = make_float4(me.x + F[0] * dt, me.y + F[1] * dt, me.z + F[2] * dt, me.w);
        if (HAS_POL) {
          a.o_pol[s] = make_float2(G[0], G[1]);
          const float2 np = make_float2(mp.x + G[0] * dt, mp.y + G[1] * dt);
          a.o2_pol[s] = np;
          npol = np;
        }
      } else if (MODE == kModeCorr) {
        // heun.hpp:170-186
        const float4 d1 = cin_d1;
        const float s0 = d1.x + F[0], s1 = d1.y + F[1], s2 = d1.z + F[2];
        if (a.old_vel) a.old_vel[s] = make_float4(s0 * 0.5f, s1 * 0.5f, s2 * 0.5f, 0.0f);
        const float4 x0 = cin_x0;
        xnew = make_float4(x0.x + s0 * 0.5f * dt, x0.y + s1 * 0.5f * dt, x0.z + s2 * 0.5f * dt, x0.w);
        a.o2_pos[s] = xnew;
        if (HAS_POL) {
          const float2 e1 = cin_e1;
          const float2 p0 = cin_p0;
          const float2 np = make_float2(p0.x + (e1.x + G[0]) * 0.5f * dt,
                                        p0.y + (e1.y + G[1]) * 0.5f * dt);
          a.o2_pol[s] = np;
          npol = np;
        }
      } else if (MODE == kModeEuler) {
        // euler.hpp:52-68
        if (a.old_vel) a.old_vel[s] = make_float4(F[0], F[1], F[2], 0.0f);
        xnew = make_float4(me.x + F[0] * dt, me.y + F[1] * dt, me.z + F[2] * dt, me.w);
        a.o_pos[s] = xnew;
        if (HAS_POL) {
          const float2 np = make_float2(mp.x + G[0] * dt, mp.y + G[1] * dt);
          a.o_pol[s] = np;
          npol = np;
        }
      }
    }
    if ((MODE == kModeCorr || MODE == kModeEuler) && leader && a.nonfinite &&
        not_finite(xnew, npol, HAS_POL))
      a.nonfinite[ii] = 1;
    if (HAS_POL && (MODE == kModePred || MODE == kModeCorr || MODE == kModeEuler)) {
      // polar_trig of the new polarity, split over the team: lane 0 takes theta, lane 1 phi
      const float th = __shfl(npol.x, base_lane, 64);
      const float ph = __shfl(npol.y, base_lane, 64);
      float sv, cv;
      sincosf(lane == 1 ? ph : th, &sv, &cv);
      const float sp = __shfl(sv, base_lane + 1, 64);
      const float cp = __shfl(cv, base_lane + 1, 64);
      if (leader) a.o_trig[s] = make_float4(sv, cv, sp, cp);
    }
    // keys of the new positions for the rebuild that starts the next step
    if ((MODE == kModeCorr || MODE == kModeEuler) && a.kc.keys)
      count_key(in, a.kc, s, leader, lane == 0 && s - in.slot_begin == in.n, xnew);
  }
  if (MODE == kModeDelta && a.com_partial) {
    const float t0 = tblock_sum_f(leader ? F[0] : 0.0f, lds_f);
    const float t1 = tblock_sum_f(leader ? F[1] : 0.0f, lds_f);
    const float t2 = tblock_sum_f(leader ? F[2] : 0.0f, lds_f);
    if (threadIdx.x == 0) a.com_partial[tb] = make_float4(t0, t1, t2, 0.0f);
  }
  const unsigned tp = wave_sum_u(pairs);
  if ((threadIdx.x & 63) == 0 && tp != 0u) atomicAdd(&a.block_pairs[tb], (unsigned long long)tp);
}

template <int T, bool GABRIEL, bool HAS_TYPE, bool HAS_POL, int MODE, bool SYM = false>
void launch_t(const SweepArgs& a, int n_blocks, hipStream_t stream) {
  // a.rel: the batch has sparse instances (their bin lookups need the ROWS kernels)
  if (a.rel)
    hipLaunchKernelGGL((k_sweep_team<T, GABRIEL, HAS_TYPE, HAS_POL, MODE, SYM, true>), dim3(n_blocks * T),
                       dim3(kBlock), 0, stream, a);
  else
    hipLaunchKernelGGL((k_sweep_team<T, GABRIEL, HAS_TYPE, HAS_POL, MODE, SYM, false>), dim3(n_blocks * T),
                       dim3(kBlock), 0, stream, a);
}

template <int T, bool GABRIEL, int MODE, bool SYM = false>
void launch_force_t(bool has_type, bool has_pol, const SweepArgs& a, int nb, hipStream_t st) {
  if (MODE == kModeCount || MODE == kModeEmit) {
    launch_t<T, GABRIEL, false, false, MODE, SYM>(a, nb, st);
  } else if (has_type) {
    if (has_pol) launch_t<T, GABRIEL, true, true, MODE, SYM>(a, nb, st);
    else launch_t<T, GABRIEL, true, false, MODE, SYM>(a, nb, st);
  } else {
    if (has_pol) launch_t<T, GABRIEL, false, true, MODE, SYM>(a, nb, st);
    else launch_t<T, GABRIEL, false, false, MODE, SYM>(a, nb, st);
  }
}

// Returns nonzero for a mode the team sweep does not have (the drag mode runs on the thread
// sweep).
template <int T, bool GABRIEL>
int launch_mode_t(int mode, bool has_type, bool has_pol, const SweepArgs& a, int nb,
                  hipStream_t st) {
  switch (mode) {
    case kModeDelta: launch_force_t<T, GABRIEL, kModeDelta>(has_type, has_pol, a, nb, st); return 0;
    case kModePred: launch_force_t<T, GABRIEL, kModePred>(has_type, has_pol, a, nb, st); return 0;
    case kModeCorr: launch_force_t<T, GABRIEL, kModeCorr>(has_type, has_pol, a, nb, st); return 0;
    case kModeEuler: launch_force_t<T, GABRIEL, kModeEuler>(has_type, has_pol, a, nb, st); return 0;
    case kModeCount: launch_force_t<T, GABRIEL, kModeCount>(has_type, has_pol, a, nb, st); return 0;
    case kModeEmit: launch_force_t<T, GABRIEL, kModeEmit>(has_type, has_pol, a, nb, st); return 0;
    default: return -1;
  }
}

template <int T>
int launch_team_width(bool gabriel, bool sym, bool has_type, bool has_pol, int mode,
                      const SweepArgs& a, int n_blocks, hipStream_t stream) {
  if (sym) {  // Delta only (checked by the caller)
    if (gabriel)
      launch_force_t<T, true, kModeDelta, true>(has_type, has_pol, a, n_blocks, stream);
    else
      launch_force_t<T, false, kModeDelta, true>(has_type, has_pol, a, n_blocks, stream);
    return 0;
  }
  if (gabriel) return launch_mode_t<T, true>(mode, has_type, has_pol, a, n_blocks, stream);
  return launch_mode_t<T, false>(mode, has_type, has_pol, a, n_blocks, stream);
}

}  // namespace

int launch_sweep_team(int team_lanes, bool gabriel, bool sym, bool has_type, bool has_pol,
                      int mode, const SweepArgs& a, int n_blocks, hipStream_t stream) {
  if (sym && mode != kModeDelta) return -1;
  switch (team_lanes) {
    case 4: return launch_team_width<4>(gabriel, sym, has_type, has_pol, mode, a, n_blocks, stream);
    case 8: return launch_team_width<8>(gabriel, sym, has_type, has_pol, mode, a, n_blocks, stream);
    case 16: return launch_team_width<16>(gabriel, sym, has_type, has_pol, mode, a, n_blocks, stream);
    default: return -1;
  }
}

}  // namespace tecs_amd
