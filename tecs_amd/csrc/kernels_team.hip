// kernels_team.hip - binned sweeps with a team of kTeam lanes per cell.
//
// A wavefront holds 64 / kTeam cells. The lanes of a cell's team split its work:
//   1. rows: the (2R+1)^2 = 9 neighbour rows (x-runs of bins, z outer, y inner - the order of
//      binned_all_pairs.hpp:117-138) are loaded one per lane and broadcast with shuffles;
//   2. candidates: the concatenated row ranges are the reference's visiting sequence; lane t
//      takes candidates t, t+kTeam, ...; each computes d2 and the cutoff test;
//   3. compaction: accepted candidates go to LDS in visiting order (ballot + popcount);
//   4. Gabriel (binned_gabriel.hpp:133-209): each accepted pair is tested against the team's
//      accepted list. Every reference blocker k of (i, j) lies within 0.9 d(i, j) of i (the ball
//      has radius c d / 2 <= 0.475 d around the midpoint, for c <= 0.95) and inside the 27 bins
//      around i, so it is an accepted candidate of i; the reference only counts k whose bin (of
//      the last rebuild) lies in the ball's bounding box, which is checked with the candidate's
//      row and x-bin. The result is the reference's blocked flag exactly;
//   5. forces of the surviving pairs are computed by their lanes and stored in LDS;
//   6. the team leader sums them in visiting order, so the sum is bitwise the serial one.
// A team whose accepted list overflows the LDS capacity (or, defensively, whose Gabriel box
// leaves the 27 bins) hands the cell to its leader, which walks it serially like the
// one-thread-per-cell kernel. Requires search_radius == 1 and, for
// Gabriel, gabriel_coefficient <= 0.95 (checked on the host).
#include "tecs_internal.hpp"

namespace tecs_amd {
namespace {

constexpr int kTeam = kTeamLanes;
constexpr int kTeamsPerBlock = kBlock / kTeam;
constexpr int kAMax = 32;  // accepted neighbours per cell held in LDS

struct TeamLds {
  float4 apos[kTeamsPerBlock][kAMax];  // accepted neighbour position (w = type bits)
  int aslot[kTeamsPerBlock][kAMax];    // its slot
  int arel[kTeamsPerBlock][kAMax];     // its (stale) bin relative to i's centre bin - 1, packed
  float4 cpos[kTeamsPerBlock][kAMax];  // position contribution, w = 1 when the pair counts
  float2 cpol[kTeamsPerBlock][kAMax];  // polarity contribution
  int unsafe[kTeamsPerBlock];  // a bounding box left the 27 bins: leader walks serially
};

__device__ __forceinline__ int pack_rel(int rx, int ry, int rz) { return rx | (ry << 2) | (rz << 4); }

// ---------------------------------------------------------------- serial fallback (leader)
template <bool GABRIEL, bool HAS_TYPE, bool HAS_POL, int MODE>
__device__ void serial_cell(const SweepArgs& a, const InstanceDesc& in, int s, float4 me,
                            float2 mp, float* acc_pos, float* acc_pol, unsigned& pairs) {
  int64_t emit_at = 0;
  if (MODE == kModeEmit) emit_at = a.row_ptr[a.id[s]];
  const float4* __restrict__ xs = a.xs;
  const int* __restrict__ offb = a.off + in.bin_begin;
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
      const int start = offb[flatten(in, x_lo, y, z)];
      const int end = offb[flatten(in, x_hi, y, z) + 1];
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
              const int s2 = offb[flatten(in, mn[0], b1, b2)];
              const int e2 = offb[flatten(in, mx[0], b1, b2) + 1];
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
        float2 q = make_float2(0.0f, 0.0f);
        if (HAS_POL) q = a.ps[m];
        apply_force<HAS_TYPE, HAS_POL>(in.force, disp, d, __float_as_int(me.w),
                                       __float_as_int(pj.w), mp.x, mp.y, q.x, q.y, acc_pos,
                                       acc_pol);
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
__device__ __forceinline__ unsigned tblock_sum_u(unsigned v, unsigned* lds) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) lds[w] = v;
  __syncthreads();
  unsigned r = 0;
  if (threadIdx.x == 0) r = lds[0] + lds[1] + lds[2] + lds[3];
  return r;
}

template <bool GABRIEL, bool HAS_TYPE, bool HAS_POL, int MODE>
__global__ __launch_bounds__(kBlock) void k_sweep_team(SweepArgs a) {
  __shared__ TeamLds L;
  __shared__ float lds_f[4];
  __shared__ unsigned lds_u[4];

  // team-block -> slots: each 256-slot chunk (one instance) holds kTeam team-blocks
  const int p = (int)blockIdx.x;
  const int nb = (int)gridDim.x;
  // XCD-aware remap (see logical_block): contiguous runs of team-blocks per XCD
  const int q = nb >> 3, r = nb & 7, xcd = p & 7, idx = p >> 3;
  const int tb = (xcd < r) ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
  const int team = (int)threadIdx.x / kTeam;
  const int lane = (int)threadIdx.x % kTeam;
  const int s = tb * kTeamsPerBlock + team;
  const int chunk = s / kBlock;  // the 256-slot chunk, one instance
  const int ii = a.block_inst[chunk];
  const InstanceDesc& in = a.inst[ii];
  const bool active = (s - in.slot_begin) < in.n;

  float4 me = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float2 mp = make_float2(0.0f, 0.0f);
  int bc[3] = {0, 0, 0};
  // rows: lane r < 9 owns row r = (dz + 1) * 3 + (dy + 1)
  int r_start = 0, r_len = 0, r_b1 = 0x7fffffff, r_b2 = 0x7fffffff, r_xlo = 0;
  if (active) {
    me = a.xs[s];
    if (HAS_POL) mp = a.ps[s];
    if (GABRIEL)
      bin_coords_div(in, me.x, me.y, me.z, bc);  // binned_gabriel.hpp:101
    else
      bin_coords_mul(in, me.x, me.y, me.z, bc);  // binned_all_pairs.hpp:103-109
    if (lane < 9) {
      const int dz = lane / 3 - 1, dy = lane % 3 - 1;
      const int z = bc[2] + dz, y = bc[1] + dy;
      if (z >= 0 && z < in.ext[2] && y >= 0 && y < in.ext[1]) {
        const int x_lo = max(bc[0] - 1, 0), x_hi = min(bc[0] + 1, in.ext[0] - 1);
        const int* offb = a.off + in.bin_begin;
        const int b_lo = flatten(in, x_lo, y, z);
        const int b_hi = flatten(in, x_hi, y, z);
        r_start = offb[b_lo];
        r_len = offb[b_hi + 1] - r_start;
        r_xlo = x_lo;
        if (GABRIEL) {
          if (x_lo + 1 <= x_hi) r_b1 = offb[b_lo + 1];
          if (x_lo + 2 <= x_hi) r_b2 = offb[b_lo + 2];
        }
      }
    }
  }
  // broadcast the row table inside the team
  int row_start[9], row_pre[9], row_b1[9], row_b2[9], row_xlo[9];
  int total = 0;
#pragma unroll
  for (int rr = 0; rr < 9; ++rr) {
    row_start[rr] = __shfl(r_start, rr, kTeam);
    const int len = __shfl(r_len, rr, kTeam);
    row_pre[rr] = total;
    total += len;
    if (GABRIEL) {
      row_b1[rr] = __shfl(r_b1, rr, kTeam);
      row_b2[rr] = __shfl(r_b2, rr, kTeam);
      row_xlo[rr] = __shfl(r_xlo, rr, kTeam);
    }
  }
  if (!active) total = 0;

  // candidates -> accepted list in LDS (visiting order)
  int base = 0;
  bool overflow = false;
  const unsigned long long team_mask_shift = (unsigned long long)((threadIdx.x & 63) / kTeam) * kTeam;
  const int n_iter = (total + kTeam - 1) / kTeam;
  float acc_pos[3] = {0.0f, 0.0f, 0.0f}, acc_pol[2] = {0.0f, 0.0f};
  for (int it = 0; it < n_iter; ++it) {
    const int c = it * kTeam + lane;
    bool acc = false;
    int m = -1, rr_sel = 0;
    float4 pj = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float disp[3] = {0.0f, 0.0f, 0.0f};
    float d2 = 0.0f;
    if (c < total) {
#pragma unroll
      for (int rr = 1; rr < 9; ++rr)
        if (c >= row_pre[rr]) rr_sel = rr;
      // rows of zero length share a prefix with the next row; pick the last row starting at c
      m = row_start[rr_sel] + (c - row_pre[rr_sel]);
      if (m != s) {
        pj = a.xs[m];
        disp[0] = me.x - pj.x;
        disp[1] = me.y - pj.y;
        disp[2] = me.z - pj.z;
        d2 = len2(disp[0], disp[1], disp[2]);
        acc = !(d2 >= in.cutoff2);
      }
    }
    const unsigned long long bal = __ballot(acc);
    const unsigned tm = (unsigned)((bal >> team_mask_shift) & ((1ull << kTeam) - 1ull));
    const int pos = base + __popc(tm & ((1u << lane) - 1u));
    if (acc) {
      if (pos < kAMax) {
        if (GABRIEL) {
          // stale bin of m relative to (bc - 1): row gives y, z; the row's bin boundaries x
          const int bx = row_xlo[rr_sel] + (m >= row_b1[rr_sel] ? 1 : 0) + (m >= row_b2[rr_sel] ? 1 : 0);
          const int rel = pack_rel(bx - bc[0] + 1, rr_sel % 3, rr_sel / 3);
          L.apos[team][pos] = pj;
          L.aslot[team][pos] = m;
          L.arel[team][pos] = rel;
        } else {
          // all-pairs: the force needs nothing else, compute it here
          const float d = sqrtf(d2);
          float cp[3] = {0.0f, 0.0f, 0.0f}, cl[2] = {0.0f, 0.0f};
          float2 qj = make_float2(0.0f, 0.0f);
          if (HAS_POL) qj = a.ps[m];
          if (MODE != kModeCount && MODE != kModeEmit)
            apply_force<HAS_TYPE, HAS_POL>(in.force, disp, d, __float_as_int(me.w),
                                           __float_as_int(pj.w), mp.x, mp.y, qj.x, qj.y, cp, cl);
          L.cpos[team][pos] = make_float4(cp[0], cp[1], cp[2], 1.0f);
          L.cpol[team][pos] = make_float2(cl[0], cl[1]);
          L.aslot[team][pos] = m;
        }
      }
    }
    base += __popc(tm);
  }
  overflow = base > kAMax;
  const int na = min(base, kAMax);
  if (lane == 0) L.unsafe[team] = 0;
  __syncthreads();

  if (GABRIEL && active && !overflow) {
    for (int e = lane; e < na; e += kTeam) {
      const float4 pj = L.apos[team][e];
      const int m = L.aslot[team][e];
      const float disp[3] = {me.x - pj.x, me.y - pj.y, me.z - pj.z};
      const float d2 = len2(disp[0], disp[1], disp[2]);
      const float mid0 = me.x - disp[0] * 0.5f;
      const float mid1 = me.y - disp[1] * 0.5f;
      const float mid2 = me.z - disp[2] * 0.5f;
      const float rho2 = d2 * in.gabriel_factor;
      const float rho = sqrtf(rho2);
      // bounding-box bins (binned_gabriel.hpp:142-153), relative to bc - 1
      const float mid[3] = {mid0, mid1, mid2};
      int lo = 0, hi = 0;  // packed 2-bit fields per dimension
      bool out_of_reach = false;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int mn = clampi((int)floorf((mid[d] - rho - in.minb[d]) * in.inv_bin), in.ext[d] - 1) - bc[d] + 1;
        const int mx = clampi((int)floorf((mid[d] + rho - in.minb[d]) * in.inv_bin), in.ext[d] - 1) - bc[d] + 1;
        out_of_reach = out_of_reach || mn < 0 || mx > 2;
        lo |= max(mn, 0) << (2 * d);
        hi |= min(mx, 2) << (2 * d);
      }
      if (out_of_reach) L.unsafe[team] = 1;
      bool blocked = false;
      for (int k = 0; k < na; ++k) {
        if (k == e) continue;
        const int rel = L.arel[team][k];
        bool inside = true;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const int v = (rel >> (2 * d)) & 3;
          inside = inside && v >= ((lo >> (2 * d)) & 3) && v <= ((hi >> (2 * d)) & 3);
        }
        if (!inside) continue;
        const float4 pk = L.apos[team][k];
        if (len2(pk.x - mid0, pk.y - mid1, pk.z - mid2) < rho2) {
          blocked = true;
          break;
        }
      }
      float cp[3] = {0.0f, 0.0f, 0.0f}, cl[2] = {0.0f, 0.0f};
      if (!blocked && MODE != kModeCount && MODE != kModeEmit) {
        const float d = sqrtf(d2);
        float2 qj = make_float2(0.0f, 0.0f);
        if (HAS_POL) qj = a.ps[m];
        apply_force<HAS_TYPE, HAS_POL>(in.force, disp, d, __float_as_int(me.w),
                                       __float_as_int(pj.w), mp.x, mp.y, qj.x, qj.y, cp, cl);
      }
      L.cpos[team][e] = make_float4(cp[0], cp[1], cp[2], blocked ? 0.0f : 1.0f);
      L.cpol[team][e] = make_float2(cl[0], cl[1]);
    }
  }
  __syncthreads();

  // leader: ordered sum (or the serial walk after an overflow)
  unsigned pairs = 0;
  const bool leader = lane == 0 && active;
  if (leader) {
    if (overflow || L.unsafe[team]) {
      serial_cell<GABRIEL, HAS_TYPE, HAS_POL, MODE>(a, in, s, me, mp, acc_pos, acc_pol, pairs);
    } else {
      int64_t emit_at = 0;
      if (MODE == kModeEmit) emit_at = a.row_ptr[a.id[s]];
      for (int e = 0; e < na; ++e) {
        const float4 c = L.cpos[team][e];
        if (c.w == 0.0f) continue;
        pairs++;
        if (MODE == kModeEmit) {
          a.cols[emit_at++] = a.id[L.aslot[team][e]];
          continue;
        }
        acc_pos[0] = acc_pos[0] + c.x;
        acc_pos[1] = acc_pos[1] + c.y;
        acc_pos[2] = acc_pos[2] + c.z;
        if (HAS_POL) {
          const float2 g = L.cpol[team][e];
          acc_pol[0] = acc_pol[0] + g.x;
          acc_pol[1] = acc_pol[1] + g.y;
        }
      }
    }
  }
  if (MODE == kModeCount) {
    if (leader) a.row_count[a.id[s]] = (int)pairs;
    return;
  }
  if (MODE == kModeEmit) return;

  float F[3], G[2];
#pragma unroll
  for (int c = 0; c < 3; ++c) F[c] = (0.0f + acc_pos[c]) + 0.0f;
  G[0] = 0.0f + acc_pol[0];
  G[1] = 0.0f + acc_pol[1];
  const float dt = in.dt;
  if (leader) {
    if (MODE == kModeDelta) {
      a.o_pos[s] = make_float4(F[0], F[1], F[2], 0.0f);
      if (HAS_POL) a.o_pol[s] = make_float2(G[0], G[1]);
    } else if (MODE == kModePred) {
      a.o_pos[s] = make_float4(F[0], F[1], F[2], 0.0f);
      a.o2_pos[s] = make_float4(me.x + F[0] * dt, me.y + F[1] * dt, me.z + F[2] * dt, me.w);
      if (HAS_POL) {
        a.o_pol[s] = make_float2(G[0], G[1]);
        a.o2_pol[s] = make_float2(mp.x + G[0] * dt, mp.y + G[1] * dt);
      }
    } else if (MODE == kModeCorr) {
      const float4 d1 = a.d1_pos[s];
      const float s0 = d1.x + F[0], s1 = d1.y + F[1], s2 = d1.z + F[2];
      if (a.old_vel) a.old_vel[s] = make_float4(s0 * 0.5f, s1 * 0.5f, s2 * 0.5f, 0.0f);
      const float4 x0 = a.o2_pos[s];
      a.o2_pos[s] = make_float4(x0.x + s0 * 0.5f * dt, x0.y + s1 * 0.5f * dt,
                                x0.z + s2 * 0.5f * dt, x0.w);
      if (HAS_POL) {
        const float2 e1 = a.d1_pol[s];
        const float2 p0 = a.o2_pol[s];
        a.o2_pol[s] = make_float2(p0.x + (e1.x + G[0]) * 0.5f * dt,
                                  p0.y + (e1.y + G[1]) * 0.5f * dt);
      }
    } else if (MODE == kModeEuler) {
      if (a.old_vel) a.old_vel[s] = make_float4(F[0], F[1], F[2], 0.0f);
      a.o_pos[s] = make_float4(me.x + F[0] * dt, me.y + F[1] * dt, me.z + F[2] * dt, me.w);
      if (HAS_POL) a.o_pol[s] = make_float2(mp.x + G[0] * dt, mp.y + G[1] * dt);
    }
  }
  if (MODE == kModeDelta && a.com_partial) {
    const float t0 = tblock_sum_f(leader ? F[0] : 0.0f, lds_f);
    const float t1 = tblock_sum_f(leader ? F[1] : 0.0f, lds_f);
    const float t2 = tblock_sum_f(leader ? F[2] : 0.0f, lds_f);
    if (threadIdx.x == 0) a.com_partial[tb] = make_float4(t0, t1, t2, 0.0f);
  }
  const unsigned tp = tblock_sum_u(pairs, lds_u);
  if (threadIdx.x == 0) a.block_pairs[tb] += tp;
}

template <bool GABRIEL, bool HAS_TYPE, bool HAS_POL, int MODE>
void launch_t(const SweepArgs& a, int n_team_blocks, hipStream_t stream) {
  hipLaunchKernelGGL((k_sweep_team<GABRIEL, HAS_TYPE, HAS_POL, MODE>), dim3(n_team_blocks),
                     dim3(kBlock), 0, stream, a);
}

template <bool GABRIEL, int MODE>
void launch_force_t(bool has_type, bool has_pol, const SweepArgs& a, int nb, hipStream_t st) {
  if (MODE == kModeCount || MODE == kModeEmit) {
    launch_t<GABRIEL, false, false, MODE>(a, nb, st);
  } else if (has_type) {
    if (has_pol) launch_t<GABRIEL, true, true, MODE>(a, nb, st);
    else launch_t<GABRIEL, true, false, MODE>(a, nb, st);
  } else {
    if (has_pol) launch_t<GABRIEL, false, true, MODE>(a, nb, st);
    else launch_t<GABRIEL, false, false, MODE>(a, nb, st);
  }
}

template <bool GABRIEL>
void launch_mode_t(int mode, bool has_type, bool has_pol, const SweepArgs& a, int nb,
                   hipStream_t st) {
  switch (mode) {
    case kModeDelta: launch_force_t<GABRIEL, kModeDelta>(has_type, has_pol, a, nb, st); break;
    case kModePred: launch_force_t<GABRIEL, kModePred>(has_type, has_pol, a, nb, st); break;
    case kModeCorr: launch_force_t<GABRIEL, kModeCorr>(has_type, has_pol, a, nb, st); break;
    case kModeEuler: launch_force_t<GABRIEL, kModeEuler>(has_type, has_pol, a, nb, st); break;
    case kModeCount: launch_force_t<GABRIEL, kModeCount>(has_type, has_pol, a, nb, st); break;
    default: launch_force_t<GABRIEL, kModeEmit>(has_type, has_pol, a, nb, st); break;
  }
}

}  // namespace

int team_blocks_for(int n_blocks) { return n_blocks * kTeam; }

void launch_sweep_team(bool gabriel, bool has_type, bool has_pol, int mode, const SweepArgs& a,
                       int n_blocks, hipStream_t stream) {
  const int nb = team_blocks_for(n_blocks);
  if (gabriel)
    launch_mode_t<true>(mode, has_type, has_pol, a, nb, stream);
  else
    launch_mode_t<false>(mode, has_type, has_pol, a, nb, stream);
}

}  // namespace tecs_amd
