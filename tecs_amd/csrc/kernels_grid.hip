// kernels_grid.hip - grid rebuild as a deterministic device counting sort, plus state I/O.
//
// Grid::rebuild (grid.hpp:389-401) computes one bin key per cell and Kokkos::BinSort orders the
// cells by key; on the Serial backend the members of a bin end up in ascending agent index.
// The device version reproduces that order exactly and moves the cell state into it. Cells are
// counted per segment: per bin for dense grids, per row of bins (fixed y, z) for sparse grids
// (e.g. a cell sheet on a sphere with ~40 bins per cell, where a scan over every bin would cost
// more than the cells), see seg_of:
//   1. keys:    key = flatten(clamp(int((p - min) / bin))), cnt[seg(key)] += 1 (atomics);
//               done by the epilogue of the sweep that wrote the positions when the host knows
//               the next step rebuilds (count_key), else by k_bin_keys
//   2. scan:    coff = exclusive_scan(cnt) over all instances' segments (+ pad segments)
//   3. scatter: slot = coff[seg] + (atomic decrement of cnt[seg]) - 1 -> tmp; this leaves cnt
//               at zero for the next rebuild, so no clearing pass is needed
//   4. gather:  each cell's final slot is coff[seg] + #(segment members before it in (key,
//               agent index) order); its state (position, polarity and its sines/cosines, id,
//               old velocity) moves
//   5. rows:    (sparse grids) each row's bin offsets relative to its first slot,
//               rel[x] = #(members with key < row key + x), for every row after a full rebuild
//               and for the rows whose members changed bins after an incremental one; it reads
//               the members as step 3 left them, so it runs in the same launch as step 4.
// Step 3's order within a segment is arbitrary, step 4 makes it canonical, so the result does
// not depend on atomic arrival order.
#include "tecs_internal.hpp"

namespace tecs_amd {
namespace {

constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;
static_assert(kScanTile == kScanTileSegments, "count_key keeps the sums of these tiles");

__global__ __launch_bounds__(kBlock) void k_bin_keys(RebuildArgs a) {
  const int lb = logical_block();
  const InstanceDesc& in = a.inst[a.block_inst[lb]];
  const int s = lb * kBlock + (int)threadIdx.x;
  const int local = s - in.slot_begin;
  const bool cell = local < in.n;
  count_key(in, a.kc, s, cell, local == in.n, cell ? a.x_old[s] : make_float4(0.0f, 0.0f, 0.0f, 0.0f));
}

// Full rebuild, step 3: each cell takes a temporary slot in its segment's new range.
__global__ __launch_bounds__(kBlock) void k_scatter(RebuildArgs a) {
  const int lb = logical_block();
  const InstanceDesc& in = a.inst[a.block_inst[lb]];
  const int s = lb * kBlock + (int)threadIdx.x;
  const int local = s - in.slot_begin;
  const bool valid = local < in.n;
  int g = 0, key = 0;
  if (valid) {
    key = a.kc.keys[s];
    g = in.cbin_begin + seg_of(in, key);
  } else if (local == in.n) {
    a.kc.cnt[in.cbin_begin + in.cbin_count - 1] = 0;
  }
  const int old = aggregated_atomic(a.kc.cnt, g, valid, -1);
  if (valid) {
    const int pos = a.coff[g] + old - 1;
    a.tmp[pos] = s;
    a.tkey[pos] = key;
    a.tid[pos] = a.id_old[s];
  }
}

__device__ __forceinline__ bool before(int ka, int ia, int kb, int ib) {
  return ka < kb || (ka == kb && ia < ib);
}

// The state of one cell, loaded before its new slot is known (the loads overlap the rank
// count). The sines and cosines of the polarity are not moved but computed again from it
// (polar_trig, the function every writer of a polarity uses, so they are the same values).
struct CellMove {
  float4 x, ov;
  float2 pol;
};
__device__ __forceinline__ CellMove load_cell(const RebuildArgs& a, int o) {
  CellMove c{};
  // read once here (streaming); the new slots are written with ordinary stores, as the sweeps
  // read them next
  c.x = load_stream(a.x_old + o);
  if (a.pol_new) c.pol = load_stream(a.pol_old + o);
  if (a.ov_new) c.ov = a.ov_old[o];
  return c;
}
__device__ __forceinline__ void store_cell(const RebuildArgs& a, const CellMove& c, int dst, int id, int key) {
  a.id_new[dst] = id;
  a.skey_new[dst] = key;
  a.x_new[dst] = c.x;
  if (a.pol_new) {
    a.pol_new[dst] = c.pol;
    a.trig_new[dst] = polar_trig(c.pol);
  }
  if (a.ov_new) a.ov_new[dst] = c.ov;
}

// #(members u in [s0, s1) with key in [k0, k0 + nb) before (key, id)); four members per
// iteration, their loads issued together.
__device__ __forceinline__ int rank_in_range(const int* __restrict__ keys, const int* __restrict__ ids,
                                             int s0, int s1, int k0, int nb, int key, int id) {
  int rank = 0;
  int u = s0;
  for (; u + 3 < s1; u += 4) {
    int k[4], i[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      k[q] = keys[u + q];
      i[q] = ids[u + q];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) rank += ((unsigned)(k[q] - k0) < (unsigned)nb && before(k[q], i[q], key, id)) ? 1 : 0;
  }
  for (; u < s1; ++u) {
    const int ku = keys[u];
    rank += ((unsigned)(ku - k0) < (unsigned)nb && before(ku, ids[u], key, id)) ? 1 : 0;
  }
  return rank;
}

// Full rebuild, step 4: final slot = segment start + #(segment members before it in (key,
// agent index) order), the members being the cells the scatter put in the segment's range.
__device__ __forceinline__ void gather_block(const RebuildArgs& a, int lb) {
  const InstanceDesc& in = a.inst[a.block_inst[lb]];
  const int t = lb * kBlock + (int)threadIdx.x;
  if ((t - in.slot_begin) >= in.n) return;
  const int key = a.tkey[t];
  const int my_id = a.tid[t];
  const CellMove c = load_cell(a, a.tmp[t]);
  const int g = in.cbin_begin + seg_of(in, key);
  const int start = a.coff[g], end = a.coff[g + 1];
  // every cell of the range is a member: the key filter passes all of them
  const int rank = rank_in_range(a.tkey, a.tid, start, end, 0, 0x7fffffff, key, my_id);
  store_cell(a, c, start + rank, my_id, key);
}

// Incremental rebuild: the members of segment g are the cells of its range at the last
// rebuild whose segment did not change, plus the cells on its arrival list. Calls f(key, id)
// for each member (old slot order, so keys and ids are read from the current arrays).
template <class F>
__device__ __forceinline__ void for_members(const RebuildArgs& a, const InstanceDesc& in, int g,
                                            int s0, int s1, int head, F f) {
  const int nb = seg_bins(in);
  const int k0 = (g - in.cbin_begin) * nb;  // the segment's keys are [k0, k0 + nb)
  for (int u = s0; u < s1; ++u) {
    const int ku = a.kc.keys[u];
    if ((unsigned)(ku - k0) < (unsigned)nb) f(ku, a.id_old[u]);
  }
  for (int u = head; u >= 0; u = a.kc.nxt[u]) f(a.kc.keys[u], a.id_old[u]);
}

// Incremental rebuild, step 4 (no scatter): cells are read in the current order.
__device__ __forceinline__ void gather_block_inc(const RebuildArgs& a, int lb) {
  const InstanceDesc& in = a.inst[a.block_inst[lb]];
  const int t = lb * kBlock + (int)threadIdx.x;
  if ((t - in.slot_begin) >= in.n) return;
  const int key = a.kc.keys[t];
  const int my_id = a.id_old[t];
  const CellMove c = load_cell(a, t);
  const int g = in.cbin_begin + seg_of(in, key);
  const int s0 = a.coff_old[g], s1 = a.coff_old[g + 1], head = a.kc.head[g], dst0 = a.coff[g];
  int rank;
  if (in.rel_begin < 0 && head < 0 && a.coff[g + 1] - dst0 == s1 - s0) {
    // a bin (dense grid: one segment per bin) with no arrivals and as many members as at the
    // last rebuild lost none either, and its members' keys are its own: they keep their order
    rank = t - s0;
  } else {
    // for_members, with the old range counted four members at a time
    const int nb = seg_bins(in);
    rank = rank_in_range(a.kc.keys, a.id_old, s0, s1, (g - in.cbin_begin) * nb, nb, key, my_id);
    for (int u = head; u >= 0; u = a.kc.nxt[u]) rank += before(a.kc.keys[u], a.id_old[u], key, my_id) ? 1 : 0;
  }
  store_cell(a, c, dst0 + rank, my_id, key);
}

// Step 5, one sparse row per wavefront pass: rel[x] = #(members with key < row key + x) for
// x = 0..ext[0] (the last entry is the row's count). Member order does not matter, so the rows
// read the members as the scatter left them (full) or as for_members lists them (incremental:
// the old range, lanes in parallel, then the arrival list) and do not wait for the gather. The
// members are counted into an LDS histogram of the row's bins (hist[x + 1] for a member in bin
// x), whose inclusive scan is rel. A fill workgroup covers kFillRows consecutive rows of one
// instance, a quarter per wavefront; after an incremental rebuild only the rows marked dirty by
// count_key are written (and their marks cleared).
constexpr int kRowHist = 1024;  // longest row (ext[0] + 1 entries) the histogram holds

template <bool INC>
__device__ __forceinline__ void fill_rows(const RebuildArgs& a, int blk, int* hist_all) {
  const InstanceDesc& in = a.inst[a.fill_inst[blk]];
  const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
  int* hist = hist_all + w * kRowHist;
  const int rows = in.cbin_count - 1;
  const int e0 = in.ext[0];
  constexpr int kWaveRows = kFillRows / (kBlock / 64);
  const int r0 = a.fill_base[blk] + w * kWaveRows;
  bool need = lane < kWaveRows && r0 + lane < rows;
  if (INC && need) need = a.dirty[in.cbin_begin + r0 + lane] != 0;
  unsigned long long todo = __ballot(need);
  while (todo != 0ull) {
    const int row = r0 + __ffsll((long long)todo) - 1;
    todo &= todo - 1ull;
    const int g = in.cbin_begin + row;
    const int k0 = row * e0;
    int* out = a.rel + in.rel_begin + row * (e0 + 1);
    if (e0 + 1 <= kRowHist) {
      for (int x = lane; x <= e0; x += 64) hist[x] = 0;
      __builtin_amdgcn_wave_barrier();
      if (INC) {
        const int s1 = a.coff_old[g + 1];
        for (int u = a.coff_old[g] + lane; u < s1; u += 64) {
          const int ku = a.kc.keys[u] - k0;
          if ((unsigned)ku < (unsigned)e0) atomicAdd(hist + ku + 1, 1);
        }
        if (lane == 0)
          for (int u = a.kc.head[g]; u >= 0; u = a.kc.nxt[u]) atomicAdd(hist + (a.kc.keys[u] - k0) + 1, 1);
      } else {
        const int s1 = a.coff[g + 1];
        for (int u = a.coff[g] + lane; u < s1; u += 64) atomicAdd(hist + (a.tkey[u] - k0) + 1, 1);
      }
      __builtin_amdgcn_wave_barrier();
      int carry = 0;
      for (int x0 = 0; x0 <= e0; x0 += 64) {
        const int x = x0 + lane;
        int v = x <= e0 ? hist[x] : 0;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const int y = __shfl_up(v, o, 64);
          if (lane >= o) v += y;
        }
        if (x <= e0) out[x] = carry + v;
        carry += __shfl(v, 63, 64);
      }
      __builtin_amdgcn_wave_barrier();
    } else {
      // a row too long for the histogram: every lane counts the members below its bins
      for (int x = lane; x <= e0; x += 64) {
        int less = 0;
        if (INC)
          for_members(a, in, g, a.coff_old[g], a.coff_old[g + 1], a.kc.head[g],
                      [&](int k, int) { less += k < k0 + x ? 1 : 0; });
        else
          for (int u = a.coff[g]; u < a.coff[g + 1]; ++u) less += a.tkey[u] < k0 + x ? 1 : 0;
        out[x] = less;
      }
    }
    if (lane == 0) a.dirty[g] = 0;
  }
}

// After a full rebuild the scatter has left cnt at zero; an incremental rebuild starts from the
// counts of the last rebuild: cnt[g] = coff[g + 1] - coff[g].
__global__ __launch_bounds__(kBlock) void k_counts_from_offsets(const int* coff, int n, int* cnt) {
  const int g = (int)blockIdx.x * kBlock + (int)threadIdx.x;
  if (g < n) cnt[g] = coff[g + 1] - coff[g];
}

// Incremental rebuild, before the scan: every cell marked as a mover (count_key) joins its new
// segment's arrival list. One thread per 64-slot word of marks (the 64 slots are in one
// workgroup, so one instance); the marks are cleared for the next rebuild. The list order is
// free: the gather and the fill only count a segment's members.
__device__ __forceinline__ void link_movers_block(const RebuildArgs& a, int n_words, int blk) {
  const int w = blk * kBlock + (int)threadIdx.x;
  if (w >= n_words) return;
  unsigned long long m = a.kc.movers[w];
  if (m == 0ull) return;
  a.kc.movers[w] = 0ull;
  const InstanceDesc& in = a.inst[a.block_inst[(w * 64) / kBlock]];
  while (m != 0ull) {
    const int s = w * 64 + (__ffsll((long long)m) - 1);
    m &= m - 1ull;
    const int g = in.cbin_begin + seg_of(in, a.kc.keys[s]);
    a.kc.nxt[s] = atomicExch(a.kc.head + g, s);
  }
}

__global__ __launch_bounds__(kBlock) void k_link_movers(RebuildArgs a, int n_words) {
  link_movers_block(a, n_words, (int)blockIdx.x);
}

// Steps 4 and 5 in one launch: workgroups [0, n_gather) gather (XCD-aware order), the rest fill.
template <bool INC>
__global__ __launch_bounds__(kBlock) void k_gather_fill(RebuildArgs a, int n_gather) {
  __shared__ int hist[(kBlock / 64) * kRowHist];
  const int p = (int)blockIdx.x;
  if (p < n_gather) {
    if (INC)
      gather_block_inc(a, remap_block(p, n_gather));
    else
      gather_block(a, remap_block(p, n_gather));
  } else {
    fill_rows<INC>(a, p - n_gather, hist);
  }
}

// ------------------------------------------------------------------ exclusive scan (int)
__device__ __forceinline__ int block_exclusive_scan(int v, int* lds_w, int* total) {
  // inclusive scan inside the wavefront
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  __syncthreads();
  if (lane == 63) lds_w[w] = x;
  __syncthreads();
  int wave_off = 0;
  for (int k = 0; k < w; ++k) wave_off += lds_w[k];
  *total = lds_w[0] + lds_w[1] + lds_w[2] + lds_w[3];
  return wave_off + x - v;
}

// Loads 8 consecutive ints per thread; full aligned tiles use two 16-byte loads.
__device__ __forceinline__ void load8(const int* in, int n, int base, int* v) {
  if (base + kScanItems <= n && ((reinterpret_cast<uintptr_t>(in + base) & 15) == 0)) {
    const int4 a = *reinterpret_cast<const int4*>(in + base);
    const int4 b = *reinterpret_cast<const int4*>(in + base + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else {
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) v[k] = (base + k < n) ? in[base + k] : 0;
  }
}

// Sets the 8 entries of an item range to -1 (the arrival heads of the following rebuild).
__device__ __forceinline__ void clear_items(int* head_clear, int n, int base) {
  if (!head_clear) return;
  if (base + kScanItems <= n && ((reinterpret_cast<uintptr_t>(head_clear + base) & 15) == 0)) {
    const int4 m = make_int4(-1, -1, -1, -1);
    *reinterpret_cast<int4*>(head_clear + base) = m;
    *reinterpret_cast<int4*>(head_clear + base + 4) = m;
  } else {
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
      if (base + k < n) head_clear[base + k] = -1;
  }
}

__global__ __launch_bounds__(kBlock) void k_scan_reduce(const int* in, int n, int* bsum) {
  __shared__ int lds[4];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int items[kScanItems];
  load8(in, n, base, items);
  int v = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) v += items[k];
  int total;
  block_exclusive_scan(v, lds, &total);
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// Single workgroup: exclusive scan of the tile sums in place; writes the grand total. Each
// thread owns a contiguous run of the tile sums, so one block scan covers them all.
__global__ __launch_bounds__(kBlock) void k_scan_top(int* bsum, int nt, int* total_out) {
  __shared__ int lds[4];
  const int per = (nt + kBlock - 1) / kBlock;
  const int first = (int)threadIdx.x * per;
  const int last = min(first + per, nt);
  int sum = 0;
  for (int i = first; i < last; ++i) sum += bsum[i];
  int total;
  int run = block_exclusive_scan(sum, lds, &total);
  for (int i = first; i < last; ++i) {
    const int v = bsum[i];
    bsum[i] = run;
    run += v;
  }
  if (threadIdx.x == 0) *total_out = total;
}

// Scans of at most this many tiles skip k_scan_top: every k_scan_apply_fused workgroup sums
// the tile sums before its own tile (at most 4096 reads from L2 per workgroup).
constexpr int kScanFusedTiles = kScanFusedTileLimit;

__device__ __forceinline__ void scan_apply_fused_tile(const int* in, int n, const int* bsum, int nt,
                                                      int* out, int* head_clear, int tile) {
  __shared__ int lds[4];
  __shared__ int lds_p[4];
  int part = 0;
  for (int i = (int)threadIdx.x; i < tile; i += kBlock) part += bsum[i];
  int prefix;
  block_exclusive_scan(part, lds_p, &prefix);
  const int base = tile * kScanTile + threadIdx.x * kScanItems;
  int v[kScanItems];
  load8(in, n, base, v);
  clear_items(head_clear, n, base);
  int sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) sum += v[k];
  int total;
  int run = prefix + block_exclusive_scan(sum, lds, &total);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (tile == nt - 1 && threadIdx.x == 0) out[n] = prefix + total;
}

__global__ __launch_bounds__(kBlock) void k_scan_apply_fused(const int* in, int n, const int* bsum,
                                                             int nt, int* out, int* head_clear) {
  scan_apply_fused_tile(in, n, bsum, nt, out, head_clear, (int)blockIdx.x);
}

// An incremental rebuild's scan (its tiles' sums are current) and the linking of its movers in
// one launch: the two touch different arrays (the scan writes the new segment offsets and
// clears the next rebuild's arrival heads, the movers join this rebuild's lists).
__global__ __launch_bounds__(kBlock) void k_scan_link(RebuildArgs a, int nt, int n_words) {
  const int p = (int)blockIdx.x;
  if (p < nt)
    scan_apply_fused_tile(a.kc.cnt, a.total_cbins, a.scan_bsum, nt, a.coff, a.head_clear, p);
  else
    link_movers_block(a, n_words, p - nt);
}

__global__ __launch_bounds__(kBlock) void k_scan_apply(const int* in, int n, const int* bsum, int* out,
                                                       int* head_clear) {
  __shared__ int lds[4];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int v[kScanItems];
  load8(in, n, base, v);
  clear_items(head_clear, n, base);
  int sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) sum += v[k];
  int total;
  int run = bsum[blockIdx.x] + block_exclusive_scan(sum, lds, &total);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
}

// ------------------------------------------------------------------------ state I/O
__global__ __launch_bounds__(kBlock) void k_init_ids(const InstanceDesc* inst,
                                                     const int* block_inst, int* id) {
  const int lb = blockIdx.x;
  const InstanceDesc& in = inst[block_inst[lb]];
  const int s = lb * kBlock + (int)threadIdx.x;
  // every slot starts with the cell of its own index; inactive slots keep that order
  id[s] = s - in.slot_begin;
}

// Host arrays hold cells first .. first + count - 1 in agent order; slot s holds agent id[s]
// (an inactive slot: the cell of its own index, kept in both buffer generations x / x2).
__global__ __launch_bounds__(kBlock) void k_upload(UploadArgs a) {
  const InstanceDesc& in = a.inst[a.ii];
  const int s = (a.block_begin + (int)blockIdx.x) * kBlock + (int)threadIdx.x;
  const int local = s - in.slot_begin;
  const bool active = local < in.n;
  const int i = (active ? a.id[s] : local) - a.first;
  if (i < 0 || i >= a.count) return;
  float w = a.x[s].w;
  if (a.type) w = __int_as_float((int)a.type[i]);
  const float4 v = make_float4(a.pos[3 * i], a.pos[3 * i + 1], a.pos[3 * i + 2], w);
  a.x[s] = v;
  if (!active) a.x2[s] = v;
  if (a.pol) {
    const float2 q = make_float2(a.pol[2 * i], a.pol[2 * i + 1]);
    a.p[s] = q;
    const float4 t = polar_trig(q);
    a.trig[s] = t;
    if (!active) {
      a.p2[s] = q;
      a.trig2[s] = t;
    }
  }
}

// A population change of instance `target` (set_agent_count): every slot of the batch moves
// into the other buffer generation. The target's cells go to the slot of their agent index
// (the order its inactive slots always have, so its ids are a permutation of the padded range);
// the other instances' slots are copied in place.
__global__ __launch_bounds__(kBlock) void k_reorder_identity(ReorderArgs a) {
  const int lb = blockIdx.x;
  const int ii = a.block_inst[lb];
  const InstanceDesc& in = a.inst[ii];
  const int s = lb * kBlock + (int)threadIdx.x;
  const int id = a.id_old[s];
  if (ii == a.target && (unsigned)id >= (unsigned)in.padded) return;  // never: ids are a permutation
  const int dst = ii == a.target ? in.slot_begin + id : s;
  a.id_new[dst] = id;
  a.x_new[dst] = a.x_old[s];
  if (a.pol_old) {
    a.pol_new[dst] = a.pol_old[s];
    a.trig_new[dst] = a.trig_old[s];
  }
  if (a.ov_old) a.ov_new[dst] = a.ov_old[s];
}

// Cells [from, to) of instance ii (inactive, in index order) lose their state, in both buffer
// generations: Kokkos::resize to a smaller extent drops them, a larger one brings zeros.
__global__ __launch_bounds__(kBlock) void k_clear_cells(ClearArgs a) {
  const InstanceDesc& in = a.inst[a.ii];
  const int local = a.from + (int)(blockIdx.x * kBlock + threadIdx.x);
  if (local >= a.to) return;
  const int s = in.slot_begin + local;
  for (int g = 0; g < 2; ++g) {
    a.x[g][s] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (a.p[g]) {
      a.p[g][s] = make_float2(0.0f, 0.0f);
      a.trig[g][s] = polar_trig(make_float2(0.0f, 0.0f));
    }
    if (a.ov[g]) a.ov[g][s] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
}

__global__ __launch_bounds__(kBlock) void k_download(const InstanceDesc* inst, int ii,
                                                     int block_begin, const int* id,
                                                     const float4* x, const float2* p,
                                                     const float4* ov, float* pos, float* pol,
                                                     float* old_vel) {
  const InstanceDesc& in = inst[ii];
  const int s = (block_begin + (int)blockIdx.x) * kBlock + (int)threadIdx.x;
  if ((s - in.slot_begin) >= in.n) return;
  const int i = id[s];
  if (pos) {
    const float4 v = x[s];
    pos[3 * i] = v.x;
    pos[3 * i + 1] = v.y;
    pos[3 * i + 2] = v.z;
  }
  if (pol) {
    const float2 v = p[s];
    pol[2 * i] = v.x;
    pol[2 * i + 1] = v.y;
  }
  if (old_vel) {
    const float4 v = ov[s];
    old_vel[3 * i] = v.x;
    old_vel[3 * i + 1] = v.y;
    old_vel[3 * i + 2] = v.z;
  }
}

__global__ __launch_bounds__(kBlock) void k_polar_trig(const float2* __restrict__ pol,
                                                       float4* __restrict__ trig, int n) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= n) return;
  trig[s] = polar_trig(pol[s]);
}

__global__ __launch_bounds__(kBlock) void k_copy_probe(const float4* __restrict__ src,
                                                       float4* __restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    dst[i] = src[i];
}

}  // namespace

void launch_polar_trig(const float2* pol, float4* trig, int total_slots, hipStream_t stream) {
  hipLaunchKernelGGL(k_polar_trig, dim3((total_slots + kBlock - 1) / kBlock), dim3(kBlock), 0,
                     stream, pol, trig, total_slots);
}

void launch_copy_probe(const float4* src, float4* dst, int64_t n, hipStream_t stream) {
  hipLaunchKernelGGL(k_copy_probe, dim3(4096), dim3(kBlock), 0, stream, src, dst, n);
}

int scan_tiles_for(int n) { return (n + kScanTile - 1) / kScanTile; }

// out[0..n] = exclusive scan of in[0..n), out[n] = total; head_clear (may be null) is set to -1
// over the same range. bsum needs scan_tiles_for(n) ints and holds the tiles' sums afterwards
// when the scan fits the fused path; tile_sums_ready: it holds them already (count_key kept
// them current), so the reduce pass is skipped.
void launch_exclusive_scan(const int* in, int n, int* out, int* bsum, int* head_clear,
                           hipStream_t stream, bool tile_sums_ready) {
  const int nt = scan_tiles_for(n);
  if (nt <= kScanFusedTiles) {
    if (!tile_sums_ready)
      hipLaunchKernelGGL(k_scan_reduce, dim3(nt), dim3(kBlock), 0, stream, in, n, bsum);
    hipLaunchKernelGGL(k_scan_apply_fused, dim3(nt), dim3(kBlock), 0, stream, in, n, bsum, nt, out,
                       head_clear);
    return;
  }
  hipLaunchKernelGGL(k_scan_reduce, dim3(nt), dim3(kBlock), 0, stream, in, n, bsum);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kBlock), 0, stream, bsum, nt, out + n);
  hipLaunchKernelGGL(k_scan_apply, dim3(nt), dim3(kBlock), 0, stream, in, n, bsum, out, head_clear);
}

void launch_rebuild(const RebuildArgs& a, int n_blocks, hipStream_t stream) {
  if (!a.keys_ready) hipLaunchKernelGGL(k_bin_keys, dim3(n_blocks), dim3(kBlock), 0, stream, a);
  const int nt = scan_tiles_for(a.total_cbins);
  if (a.incremental && a.kc.tsum != nullptr && nt <= kScanFusedTiles) {
    // the counts moved with their tiles' sums (count_key): scan and movers in one launch
    const int n_words = n_blocks * (kBlock / 64);
    hipLaunchKernelGGL(k_scan_link, dim3(nt + (n_words + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, a,
                       nt, n_words);
  } else {
    if (a.incremental) {
      const int n_words = n_blocks * (kBlock / 64);
      hipLaunchKernelGGL(k_link_movers, dim3((n_words + kBlock - 1) / kBlock), dim3(kBlock), 0, stream,
                         a, n_words);
    }
    launch_exclusive_scan(a.kc.cnt, a.total_cbins, a.coff, a.scan_bsum, a.head_clear, stream,
                          a.incremental && a.kc.tsum != nullptr);
  }
  const dim3 gf(n_blocks + (a.fill ? a.fill_blocks : 0));
  if (a.incremental) {
    hipLaunchKernelGGL(k_gather_fill<true>, gf, dim3(kBlock), 0, stream, a, n_blocks);
  } else {
    hipLaunchKernelGGL(k_scatter, dim3(n_blocks), dim3(kBlock), 0, stream, a);
    hipLaunchKernelGGL(k_gather_fill<false>, gf, dim3(kBlock), 0, stream, a, n_blocks);
    hipLaunchKernelGGL(k_counts_from_offsets, dim3((a.total_cbins + kBlock - 1) / kBlock),
                       dim3(kBlock), 0, stream, a.coff, a.total_cbins, a.kc.cnt);
  }
}

void launch_init_ids(const InstanceDesc* inst, const int* block_inst, int* id, int n_blocks,
                     hipStream_t stream) {
  hipLaunchKernelGGL(k_init_ids, dim3(n_blocks), dim3(kBlock), 0, stream, inst, block_inst, id);
}

void launch_upload(const UploadArgs& a, int n_blocks, hipStream_t stream) {
  hipLaunchKernelGGL(k_upload, dim3(n_blocks), dim3(kBlock), 0, stream, a);
}

void launch_reorder_identity(const ReorderArgs& a, int n_blocks, hipStream_t stream) {
  hipLaunchKernelGGL(k_reorder_identity, dim3(n_blocks), dim3(kBlock), 0, stream, a);
}

void launch_clear_cells(const ClearArgs& a, hipStream_t stream) {
  if (a.to <= a.from) return;
  hipLaunchKernelGGL(k_clear_cells, dim3((a.to - a.from + kBlock - 1) / kBlock), dim3(kBlock), 0,
                     stream, a);
}

void launch_download(const InstanceDesc* inst, int ii, int block_begin, int n_blocks,
                     const int* id, const float4* x, const float2* p, const float4* ov,
                     float* pos, float* pol, float* old_vel, hipStream_t stream) {
  hipLaunchKernelGGL(k_download, dim3(n_blocks), dim3(kBlock), 0, stream, inst, ii, block_begin,
                     id, x, p, ov, pos, pol, old_vel);
}

}  // namespace tecs_amd
