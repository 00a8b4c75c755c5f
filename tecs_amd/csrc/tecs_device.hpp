// tecs_device.hpp - device data layout and the per-pair arithmetic of the engine.
//
// Layout (HBM, per batch of simulations; a single simulation is a batch of one):
//   * Cells live in "slots". Each instance owns a slot range [slot_begin, slot_begin + padded)
//     with padded a multiple of kBlock, so every workgroup belongs to exactly one instance.
//   * Between grid rebuilds the slots of an instance are ordered by (bin, agent index): the
//     members of a bin are contiguous and ascending in agent index, which is the order
//     Kokkos::BinSort gives on the Serial backend (grid.hpp:399-400). Walking a bin's slot range
//     therefore visits j in the reference's order.
//   * x4[s]   = float4(x, y, z, type bits)  - position and the read-only cell type
//     pol[s]  = float2(theta, phi)           - polarity (when present)
//     id[s]   = agent index of the cell in slot s (within its instance)
//   * Bin offsets of all instances are concatenated; instance bins are followed by one pad bin
//     that holds the instance's pad slots, so one global exclusive scan yields global slot
//     offsets for every instance.
//
// All expressions below keep the reference's operand order; the library is compiled with
// -ffp-contract=off so no multiply-add is fused and every operation rounds once, as in
// IEEE binary32 evaluation of the reference source. Division and square root are the
// correctly rounded sequences (hipcc's default for HIP single precision).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tecs_amd {

constexpr int kBlock = 256;  // slots per workgroup (4 wavefronts of 64)
constexpr int kMaxTypes = 4;
constexpr int kMaxWalls = 8;  // TECS_MAX_WALLS

// Term order flags of the fused force: only the relative order of terms that write the same
// field changes a sum (PiecewiseLinear and bending both write the position delta, bending and
// alignment both write the polarity delta).
constexpr int kForceLinear = 1;      // the PiecewiseLinear term is applied
constexpr int kForceBendFirst = 2;   // bending's position term is added before PiecewiseLinear's
constexpr int kForceAlignFirst = 4;  // alignment is added before bending's polarity term
constexpr int kForceDrag = 8;        // some pair has nonzero drag (the drag correction runs)

struct ForceParams {
  float h, krep, kadh;
  int n_types;
  float table[kMaxTypes * kMaxTypes];
  float bend, align;
  int flags;
  float drag;
  int drag_by_type;
  float drag_table[kMaxTypes * kMaxTypes];
};

struct InstanceDesc {
  int slot_begin;  // first global slot
  int n;           // active cells
  int cap;         // capacity (GlobalComFixer divisor)
  int padded;      // slot range length, multiple of kBlock
  int block_begin; // first logical workgroup of the instance
  int n_blocks;    // padded / kBlock
  int bin_count;   // real bins; the pad bin has local index bin_count
  int cbin_begin;  // first entry of the instance in the concatenated segment arrays
  int cbin_count;  // segments including the pad segment (the last one)
  int rel_begin;   // sparse grid: first entry of its row-relative bin offsets (-1: dense grid,
                   // one counting segment per bin)
  unsigned row_mul;  // sparse grid: key / ext[0] == umulhi(key, row_mul) >> row_shift for every
  int row_shift;     // key < 2^31 (row_mul 0: ext[0] == 1)
  int ext[3];
  float minb[3];
  float bin_size, inv_bin;
  float cutoff2;
  float gabriel_factor;  // c*c*0.25 (binned_gabriel.hpp:42)
  float gabriel_c2;      // c*c     (naive_gabriel.hpp:392)
  int search_radius;
  int bin_pow2;          // bin_size is a power of two: (p - min) / bin == (p - min) * inv_bin
  float dt;              // float(dt) of the current step
  ForceParams force;
  int n_walls;           // generic wall forces: normalised (a, b, c, d) and strength
  float wall[kMaxWalls][5];
  float link_k;          // link force strength
};

// Logical workgroup index with an XCD-aware remap: the hardware hands consecutive workgroup
// ids to the 8 XCDs round-robin; remapping gives each XCD a contiguous run of logical blocks,
// so spatially neighbouring slots (which share neighbour bins) meet in the same L2.
// The map is a bijection on [0, nb) for any nb.
__device__ __forceinline__ int remap_block(int p, int nb) {
  const int q = nb >> 3, r = nb & 7;
  const int xcd = p & 7, idx = p >> 3;
  return (xcd < r) ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}
__device__ __forceinline__ int logical_block() { return remap_block((int)blockIdx.x, (int)gridDim.x); }

// ---------------------------------------------------------------- bin arithmetic (grid.hpp)
__device__ __forceinline__ int clampi(int v, int hi) { return max(0, min(v, hi)); }

// Grid::calc_bin_coords_from_point, grid.hpp:222-230: (p - min) / bin, truncation, clamp.
// Dividing by a power of two and multiplying by its exact reciprocal round the same exact
// value, so the division is replaced by the product when bin_pow2 is set.
__device__ __forceinline__ void bin_coords_div(const InstanceDesc& in, float x, float y, float z,
                                               int* c) {
  if (in.bin_pow2) {
    c[0] = clampi((int)((x - in.minb[0]) * in.inv_bin), in.ext[0] - 1);
    c[1] = clampi((int)((y - in.minb[1]) * in.inv_bin), in.ext[1] - 1);
    c[2] = clampi((int)((z - in.minb[2]) * in.inv_bin), in.ext[2] - 1);
    return;
  }
  c[0] = clampi((int)((x - in.minb[0]) / in.bin_size), in.ext[0] - 1);
  c[1] = clampi((int)((y - in.minb[1]) / in.bin_size), in.ext[1] - 1);
  c[2] = clampi((int)((z - in.minb[2]) / in.bin_size), in.ext[2] - 1);
}
// One coordinate of bin_coords_div: dimension d's position v, minimum mn and extent ext.
__device__ __forceinline__ int bin_coord_div(const InstanceDesc& in, float v, float mn, int ext) {
  return clampi((int)(in.bin_pow2 ? (v - mn) * in.inv_bin : (v - mn) / in.bin_size), ext - 1);
}
// binned_all_pairs.hpp:103-109: (p - min) * inv_bin, truncation, clamp.
__device__ __forceinline__ void bin_coords_mul(const InstanceDesc& in, float x, float y, float z,
                                               int* c) {
  c[0] = clampi((int)((x - in.minb[0]) * in.inv_bin), in.ext[0] - 1);
  c[1] = clampi((int)((y - in.minb[1]) * in.inv_bin), in.ext[1] - 1);
  c[2] = clampi((int)((z - in.minb[2]) * in.inv_bin), in.ext[2] - 1);
}
// Grid::flatten_bin_index, grid.hpp:243-245.
__device__ __forceinline__ int flatten(const InstanceDesc& in, int c0, int c1, int c2) {
  return c0 + c1 * in.ext[0] + c2 * (in.ext[0] * in.ext[1]);
}

// Wavefront-aggregated counter update: lanes holding the same counter index in a run of
// consecutive lanes (cells arrive in the previous bin order, so runs are common) issue one
// atomic for the run. Returns the counter's value before the update, minus the lane's rank in
// the run, so that with delta = -1 every lane gets a distinct position (old - 1 - rank).
struct LaneRun {
  bool head;    // the lane issues the run's atomic
  int my_head;  // lane of the run's head
  int run;      // lanes in the run
};
__device__ __forceinline__ LaneRun lane_run(int idx, bool valid) {
  const int lane = (int)(threadIdx.x & 63);
  const int key = valid ? idx : -1;
  const int prev = __shfl_up(key, 1, 64);
  const bool head = valid && (lane == 0 || prev != key);
  const unsigned long long heads = __ballot(head);
  const unsigned long long valids = __ballot(valid);
  // my run's head: the highest head at or below this lane
  const unsigned long long below = heads & ((lane == 63) ? ~0ull : ((2ull << lane) - 1ull));
  const int my_head = valid ? 63 - __clzll((long long)below) : 0;
  // run length: lanes from the head up to (not including) the next head or invalid lane
  const unsigned long long after = (heads | ~valids) & ~((2ull << my_head) - 1ull);
  const int next = after ? __ffsll((long long)after) - 1 : 64;
  return LaneRun{head, my_head, next - my_head};
}
__device__ __forceinline__ int aggregated_atomic(int* counter_base, int idx, bool valid, int delta) {
  const int lane = (int)(threadIdx.x & 63);
  const LaneRun lr = lane_run(idx, valid);
  const bool head = lr.head;
  const int my_head = lr.my_head, run = lr.run;
  int old = 0;
  if (head) old = atomicAdd(counter_base + idx, delta * run);
  old = __shfl(old, my_head, 64);
  return old - (lane - my_head);
}

// Counting segments and bin offsets. A dense grid counts its cells per bin, so the scanned
// segment offsets are the bin offsets. A sparse grid (many bins per cell, e.g. a cell sheet on
// a sphere) counts per row of bins (fixed y and z) and keeps, per row, the offset of every bin
// relative to the row's first slot (ext[0] + 1 entries; the last one is the row's count). A
// rebuild then rewrites the relative offsets of the rows whose members changed bins only,
// instead of one offset per bin of the whole grid.
__device__ __forceinline__ int seg_of(const InstanceDesc& in, int key) {
  if (in.rel_begin < 0 || in.row_mul == 0u) return key;
  return (int)(__umulhi((unsigned)key, in.row_mul) >> in.row_shift);
}
// keys of the bins of local segment g: [g * seg_bins, (g + 1) * seg_bins)
__device__ __forceinline__ int seg_bins(const InstanceDesc& in) {
  return in.rel_begin >= 0 ? in.ext[0] : 1;
}

// The slot ranges of an instance's bins in the current order (see seg_of): the first slot of
// bin (x, y, z) is base(row) + tab[row * stride + x], row = y + z * ext[1], where a dense grid
// reads its segment (bin) offsets directly (base 0, stride ext[0]) and a sparse one its
// row-relative offsets (base = the row's segment offset, stride ext[0] + 1). ROWS: the kernel
// may meet sparse instances (a batch without any is swept by kernels built with ROWS false,
// whose lookups are the dense ones only).
template <bool ROWS>
struct BinRanges {
  const int* seg;  // segment offsets of the instance (global slots)
  const int* tab;  // dense: seg; sparse: the instance's row-relative offsets
  int stride, e1;
  bool sparse;
  __device__ __forceinline__ const int* row_of(int y, int z, int& base) const {
    const int r = y + z * e1;
    base = 0;
    if (ROWS) {
      // the row's segment offset is a valid entry for a dense instance too (row < bins), so
      // it is loaded unconditionally and the lookups of several rows can be issued together
      const int b = seg[r];
      base = sparse ? b : 0;
    }
    return tab + r * stride;
  }
  // first slot of bin (x, y, z); x == ext[0] is the end of the row
  __device__ __forceinline__ int at(int x, int y, int z) const {
    int base;
    const int* p = row_of(y, z, base);
    return base + p[x];
  }
  // slots [start, end) of bins x_lo..x_hi of row (y, z)
  __device__ __forceinline__ void row(int x_lo, int x_hi, int y, int z, int& start, int& end) const {
    int base;
    const int* p = row_of(y, z, base);
    start = base + p[x_lo];
    end = base + p[x_hi + 1];
  }
};
// seg_off: the segment offsets of all instances, rel: the row-relative offsets of all sparse
// instances (null when there are none; then ROWS may be false)
template <bool ROWS>
__device__ __forceinline__ BinRanges<ROWS> bin_ranges(const InstanceDesc& in, const int* seg_off, const int* rel) {
  const bool sparse = ROWS && in.rel_begin >= 0;
  const int* seg = seg_off + in.cbin_begin;
  return BinRanges<ROWS>{seg, sparse ? rel + in.rel_begin : seg, sparse ? in.ext[0] + 1 : in.ext[0],
                         in.ext[1], sparse};
}

// Grid::rebuild's key of a position (grid.hpp:394-395) and its counting-sort segment.
__device__ __forceinline__ int bin_key(const InstanceDesc& in, float4 p) {
  int c[3];
  bin_coords_div(in, p.x, p.y, p.z, c);
  return flatten(in, c[0], c[1], c[2]);
}

// Step 1 of the rebuild (count_key): the key of every cell and its count per segment. Full
// rebuild: cnt starts at zero and every cell counts. Incremental rebuild: cnt holds the counts
// of the last rebuild and only a cell whose segment changed moves its count (old segment - 1,
// new segment + 1) and joins the new segment's arrival list head[] / nxt[].
// Segments per tile of the rebuild's exclusive scan (256 threads x 8 items).
constexpr int kScanTileSegments = 2048;

struct KeyCount {
  int* keys;        // key per slot (current order)
  int* cnt;         // per-segment counts
  const int* skey;  // incremental: key per slot at the last rebuild (null = full)
  int* head;        // incremental: arrival list heads per segment (-1 = empty) ...
  int* nxt;         //              ... and links per slot, built by k_link_movers
  int* tsum;        // incremental: the scan's per-tile sums of cnt, kept current with it
  unsigned long long* movers;  // incremental: bit (s & 63) of word s >> 6 marks a cell that
                               // changed segment; k_link_movers puts it on its arrival list
  int* dirty;       // incremental, sparse grids: 1 for a row whose members changed bins (its
                    // relative offsets are rewritten by the rebuild, which clears the flag)
};

// Step 1 of the rebuild (kernels_grid.hip) for one slot: store the key of the cell's position
// and count it. Full rebuild: every cell counts in its segment (wavefront-aggregated atomics)
// and the slot just past the instance's cells sets the pad segment's count. Incremental
// rebuild: only a cell whose segment differs from its segment at the last rebuild moves its
// count (and, between scan tiles, the tiles' sums, so the scan need not sum the tiles again)
// and marks itself as a mover; the rebuild links the movers into their new segments' arrival
// lists (k_link_movers), so no wavefront here waits for an atomic's result. Called by every
// lane of the wavefront; `cell` marks lanes that hold a cell, `pad` the lane of the slot just
// past the cells.
// `old_key` is k.skey[s] when the caller has loaded it already (a sweep loads it with the
// cell's other inputs), else -1.
// count_key_of: the same with the cell's key computed by the caller (ignored where !cell).
__device__ __forceinline__ void count_key_of(const InstanceDesc& in, const KeyCount& k, int s,
                                             bool cell, bool pad, int key, int old_key = -1) {
  if (k.skey) {
    if (!cell) return;
    k.keys[s] = key;
    const int ko = old_key >= 0 ? old_key : k.skey[s];
    const int gn = in.cbin_begin + seg_of(in, key);
    const int go = in.cbin_begin + seg_of(in, ko);
    if (in.rel_begin >= 0 && key != ko) {
      k.dirty[gn] = 1;
      k.dirty[go] = 1;
    }
    if (gn != go) {
      __hip_atomic_fetch_add(k.cnt + go, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(k.cnt + gn, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const int to = go / kScanTileSegments, tn = gn / kScanTileSegments;
      if (k.tsum && to != tn) {
        __hip_atomic_fetch_add(k.tsum + to, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(k.tsum + tn, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      __hip_atomic_fetch_or(k.movers + (s >> 6), 1ull << (s & 63), __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_AGENT);
    }
    return;
  }
  int g = 0;
  if (cell) {
    k.keys[s] = key;
    g = in.cbin_begin + seg_of(in, key);
  } else if (pad) {
    // the pad segment holds the instance's pad slots, so the global scan lands on
    // slot_begin of the next instance
    k.cnt[in.cbin_begin + in.cbin_count - 1] = in.padded - in.n;
  }
  const LaneRun lr = lane_run(g, cell);
  if (lr.head) __hip_atomic_fetch_add(k.cnt + g, lr.run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void count_key(const InstanceDesc& in, const KeyCount& k, int s,
                                          bool cell, bool pad, float4 p, int old_key = -1) {
  count_key_of(in, k, s, cell, pad, cell ? bin_key(in, p) : 0, old_key);
}

// Streaming accesses: data a kernel touches once (the corrector's inputs, the step's outputs
// that the next kernel reads once) is loaded and stored with the non-temporal hint, so that it
// does not push out of the caches the positions the pair walks read many times over.
typedef float StreamF4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 load_stream(const float4* p) {
  const StreamF4 v = __builtin_nontemporal_load(reinterpret_cast<const StreamF4*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}
typedef float StreamF2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 load_stream(const float2* p) {
  const StreamF2 v = __builtin_nontemporal_load(reinterpret_cast<const StreamF2*>(p));
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ void store_stream(float4* p, float4 v) {
  const StreamF4 w = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(w, reinterpret_cast<StreamF4*>(p));
}

// ------------------------------------------------------------------- vector.hpp:312-343
__device__ __forceinline__ float len2(float a, float b, float c) {
  float r = 0.0f;
  r = r + a * a;
  r = r + b * b;
  r = r + c * c;
  return r;
}
// VectorN::dot, vector.hpp:312-318.
__device__ __forceinline__ float dot3(const float* a, const float* b) {
  float r = 0.0f;
  r = r + a[0] * b[0];
  r = r + a[1] * b[1];
  r = r + a[2] * b[2];
  return r;
}

// ------------------------------------------- correctly rounded division by a positive value
// RN(x / d) and RN(1 / d) from the hardware reciprocal estimate (v_rcp_f32) by Markstein's
// method (P. Markstein, "Computation of elementary functions on the IBM RISC System/6000
// processor", IBM J. Res. Develop. 34(1), 1990): one Newton-Raphson step y = y0 + y0 (1 - d y0)
// on the estimate gives y = RN(1 / d); with q0 = RN(x y) the remainder r = d q0 - x is exact in
// one FMA, and q = RN(q0 - r y) = RN(x / d). Checked against IEEE division on gfx950
// (tests/test_gpu_fast_div.py): the reciprocal for every d with |d| in [2^-16, 2^16], the
// quotient for every pair of significands (x, d in [1, 2)). Scaling x and d by powers of two,
// or flipping the sign of x, maps every step of the sequence exactly while all its values stay
// normal, which holds on the quotient's domain d in [2^-16, 2^16], |x| in [2^-96, 2^96] or
// x = +-0 (then q = x: the remainder is +0 and q0 - (+0) y keeps q0's sign, which needs y > 0,
// so the quotient's domain has d > 0). Outside its domain each function divides as IEEE.
constexpr float kDivDMin = 0x1p-16f, kDivDMax = 0x1p16f;
constexpr float kDivXMin = 0x1p-96f, kDivXMax = 0x1p96f;

__device__ __forceinline__ float rcp_refined(float d) {
  const float y0 = __builtin_amdgcn_rcpf(d);
  return __builtin_fmaf(__builtin_fmaf(-d, y0, 1.0f), y0, y0);
}
// the reciprocal's domain (either sign) and the quotient's (d > 0)
// (bitwise operators on ints: no short-circuit branches)
__device__ __forceinline__ bool recip_in_domain(float d) {
  const float a = fabsf(d);
  return (int)(a >= kDivDMin) & (int)(a <= kDivDMax);
}
__device__ __forceinline__ bool div_d_in_domain(float d) {
  return (int)(d >= kDivDMin) & (int)(d <= kDivDMax);
}
__device__ __forceinline__ bool div_x_in_domain(float x) {
  const float a = fabsf(x);
  return ((int)(a >= kDivXMin) | (int)(a == 0.0f)) & (int)(a <= kDivXMax);
}
// RN(x / d) for x, d in the domain, y = rcp_refined(d).
__device__ __forceinline__ float div_by_rcp(float x, float d, float y) {
  const float q0 = x * y;
  const float r = __builtin_fmaf(d, q0, -x);
  return __builtin_fmaf(-r, y, q0);
}
// RN(1 / d).
__device__ __forceinline__ float recip_rn(float d) {
  const float y = rcp_refined(d);
  return recip_in_domain(d) ? y : 1.0f / d;
}
// out[c] = RN(x[c] / d), c < 3.
__device__ __forceinline__ void div3_rn(const float* x, float d, float* out) {
  const float y = rcp_refined(d);
#pragma unroll
  for (int c = 0; c < 3; ++c) out[c] = div_by_rcp(x[c], d, y);
  const bool in = (int)div_d_in_domain(d) & (int)div_x_in_domain(x[0]) &
                  (int)div_x_in_domain(x[1]) & (int)div_x_in_domain(x[2]);
  if (__builtin_expect(!in, 0)) {
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c] = x[c] / d;
  }
}

// out[c] = RN(v[c] * F / d), c < 3, the quotients of PiecewiseLinear, for d > 0 and
// |v[c]| <= d (v a displacement and d its length). The domain test is cheaper than div3_rn's:
// with d <= 2^16 and |F| <= 2^64 every |v[c] * F| is below 2^81, so only the smallest
// numerator needs a test, through its binary exponent (frexp: 0 for a zero numerator, which
// is in the domain, at most -96 below 2^-96, subnormals included).
__device__ __forceinline__ void pl_div3_rn(const float* v, float F, float d, float* out) {
  const float x[3] = {v[0] * F, v[1] * F, v[2] * F};
  const float y = rcp_refined(d);
#pragma unroll
  for (int c = 0; c < 3; ++c) out[c] = div_by_rcp(x[c], d, y);
  const int emin = min(min(__builtin_amdgcn_frexp_expf(x[0]), __builtin_amdgcn_frexp_expf(x[1])),
                       __builtin_amdgcn_frexp_expf(x[2]));
  const bool in = (int)(d >= kDivDMin) & (int)(d <= kDivDMax) & (int)(fabsf(F) <= 0x1p64f) &
                  (int)(emin >= -95);
  if (__builtin_expect(!in, 0)) {
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c] = x[c] / d;
  }
}

// ---------------------------------------------------------- force laws (predefined.hpp)
// PiecewiseLinear, predefined.hpp:107-144.
__device__ __forceinline__ void piecewise_linear(const float* disp, float d, float h, float krep,
                                                 float kadh, float* out) {
  if (d == 0.0f) {
    out[0] = out[1] = out[2] = 0.0f;
    return;
  }
  const float F = krep * fmaxf(h - d, 0.0f) - kadh * fmaxf(d - h, 0.0f);
  pl_div3_rn(disp, F, d, out);  // (disp * F) / d
}

// ---------------------------------------------------------------- polarity.hpp:111-198
// Polarity of a cell with the sines and cosines of its angles: trig = (sin theta, cos theta,
// sin phi, cos phi), computed once per substep per slot (k_polar_trig).
// The Heun corrector's per-cell inputs besides the evaluation: the first evaluation d1 and the
// stage-0 state it updates in place. The sweeps load them before walking the neighbours, so
// the loads overlap the walk instead of stalling the epilogue.
struct CorrIn {
  float4 d1;
  float4 x0;
  float2 e1;
  float2 p0;
};

struct Pol {
  float t, p;    // theta, phi
  float4 trig;   // sin t, cos t, sin p, cos p
};
__device__ __forceinline__ Pol make_pol(float2 q, float4 trig) { return Pol{q.x, q.y, trig}; }
// (sin theta, cos theta, sin phi, cos phi) of a polarity, from sincosf: the sines and cosines
// the reference's force evaluates (to_vector3 polarity.hpp:111-118, dot :131-136, the
// sin theta of :153). Every kernel that writes a polarity also writes this of it, so a
// polarity buffer and its trig buffer always agree.
__device__ __forceinline__ float4 polar_trig(float2 q) {
  float st, ct, sp, cp;
  sincosf(q.x, &st, &ct);
  sincosf(q.y, &sp, &cp);
  return make_float4(st, ct, sp, cp);
}

// The arithmetic of the fused pairwise force's terms (see tecs_force_desc in tecs_amd.h).
//
// The PiecewiseLinear part is evaluated with the reference's operations in its order, so it is
// bitwise the reference's FP32 result. The polarity terms are evaluated from the same
// mathematical expressions with the angle functions replaced by exact identities over the
// precomputed sines and cosines (checked against an FP64 evaluation of the reference's
// formulas, 1e-5 relative):
//   bending_force (polarity.hpp:181-198): prod = v.disp / d, the coefficients -prod/d and
//     pow(prod, 2)/pow(d, 2) = (prod/d)^2 from one reciprocal of d;
//   from_vector3(disp, d) (polarity.hpp:62-66) gives q = (acos(dz/d), atan2(dy, dx)), so
//     sin q0 cos(p - q1) = (cos p dx + sin p dy) / d, sin q0 sin(p - q1) = (sin p dx - cos p dy) / d
//     and cos q0 = dz / d; unidirectional_polarization_force(q) (polarity.hpp:145-157) needs
//     only these three (sin q0 = sqrt(dx^2 + dy^2) / d >= 0 is never formed, which avoids the
//     cancellation of acos near the poles);
//   bidirectional_polarization_force (polarity.hpp:165-168, dot :131-136) uses
//     cos(p_i - p_j) = cos p_i cos p_j + sin p_i sin p_j and
//     sin(p_i - p_j) = sin p_i cos p_j - cos p_i sin p_j.
// The guard |sin theta_i| > FLT_EPSILON (polarity.hpp:154) is kept on the same value.
// The values of the terms of one pair, each on its own (zero when the term is not applied):
// PiecewiseLinear's position term v, bending's position term bpos = bend * b.vector and
// polarity term bpol = bend * b.polarity, alignment's polarity term al.
struct PairTerms {
  float v[3], bpos[3], bpol[2], al[2];
};

// The adhesion scale of pair (i, j): the type table's entry, or the scalar without a table.
template <bool HAS_TYPE>
__device__ __forceinline__ float pair_kadh(const ForceParams& f, const float* table, int type_i,
                                           int type_j) {
  if (HAS_TYPE && f.n_types > 0) return table[type_i * f.n_types + type_j];
  return f.kadh;
}

// `kadh` is the pair's adhesion scale (pair_kadh).
template <bool HAS_POL>
__device__ __forceinline__ void pair_terms_k(const ForceParams& f, const float* disp, float d,
                                             float kadh, const Pol& pi, const Pol& pj,
                                             PairTerms& t) {
  t.v[0] = t.v[1] = t.v[2] = 0.0f;
  t.bpos[0] = t.bpos[1] = t.bpos[2] = 0.0f;
  t.bpol[0] = t.bpol[1] = t.al[0] = t.al[1] = 0.0f;
  if (f.flags & kForceLinear) piecewise_linear(disp, d, f.h, f.krep, kadh, t.v);
  if (!HAS_POL) return;
  const float sti = pi.trig.x, cti = pi.trig.y, spi = pi.trig.z, cpi = pi.trig.w;
  const float stj = pj.trig.x, ctj = pj.trig.y, spj = pj.trig.z, cpj = pj.trig.w;
  const bool pole = !(fabsf(sti) > 1.1920928955078125e-07f);
  const float inv_sti = pole ? 0.0f : recip_rn(sti);
  if (f.bend != 0.0f) {
    const float inv_d = recip_rn(d);
    const float tv[3] = {sti * cpi, sti * spi, cti};  // to_vector3, polarity.hpp:111-118
    const float ov[3] = {stj * cpj, stj * spj, ctj};
    const float prod_i = dot3(tv, disp) * inv_d;
    const float prod_j = dot3(ov, disp) * inv_d;
    const float ri = prod_i * inv_d, rj = prod_j * inv_d;
    const float a = -ri, b = ri * ri, c = -rj, e = rj * rj;
#pragma unroll
    for (int k = 0; k < 3; ++k) t.bpos[k] = f.bend * (((a * tv[k] + b * disp[k]) + c * ov[k]) + e * disp[k]);
    const float sc = (cpi * disp[0] + spi * disp[1]) * inv_d;  // sin q0 cos(p_i - q1)
    const float ss = (spi * disp[0] - cpi * disp[1]) * inv_d;  // sin q0 sin(p_i - q1)
    const float cq0 = disp[2] * inv_d;
    const float u0 = cti * sc - sti * cq0;
    const float u1 = pole ? 0.0f : -ss * inv_sti;
    t.bpol[0] = f.bend * (-prod_i * u0);
    t.bpol[1] = f.bend * (-prod_i * u1);
  }
  if (f.align != 0.0f) {
    const float cdp = cpi * cpj + spi * spj;  // cos(p_i - p_j)
    const float sdp = spi * cpj - cpi * spj;  // sin(p_i - p_j)
    const float u0 = cti * stj * cdp - sti * ctj;
    const float u1 = pole ? 0.0f : -stj * sdp * inv_sti;
    const float sdot = sti * stj * cdp + cti * ctj;
    t.al[0] = f.align * (sdot * u0);
    t.al[1] = f.align * (sdot * u1);
  }
}

template <bool HAS_TYPE, bool HAS_POL>
__device__ __forceinline__ void pair_terms(const ForceParams& f, const float* disp, float d,
                                           int type_i, int type_j, const Pol& pi, const Pol& pj,
                                           PairTerms& t) {
  pair_terms_k<HAS_POL>(f, disp, d, pair_kadh<HAS_TYPE>(f, f.table, type_i, type_j), pi, pj, t);
}

// The terms of one pair in the order they add to a field: p1 then p2 for the position, q1 then
// q2 for the polarity. A term that is not applied is +0, and adding it changes nothing: a sum
// that starts at +0 never becomes -0 under round-to-nearest, and adding +0 to any other value
// leaves it unchanged. So acc = (acc + p1) + p2, pair after pair, is the reference's running
// sum with only the applied terms, in the composition's order. The order flags are set on the
// host only when both terms of a field are applied.
template <bool HAS_POL>
__device__ __forceinline__ void apply_force_terms_k(const ForceParams& f, const float* disp, float d,
                                                    float kadh, const Pol& pi, const Pol& pj,
                                                    float* p1, float* p2, float* q1, float* q2) {
  PairTerms t;
  pair_terms_k<HAS_POL>(f, disp, d, kadh, pi, pj, t);
  const bool bend_first = (f.flags & kForceBendFirst) != 0;
  const bool align_first = (f.flags & kForceAlignFirst) != 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    p1[k] = bend_first ? t.bpos[k] : t.v[k];
    p2[k] = bend_first ? t.v[k] : t.bpos[k];
  }
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    q1[k] = align_first ? t.al[k] : t.bpol[k];
    q2[k] = align_first ? t.bpol[k] : t.al[k];
  }
}

// apply_force_terms_k with the adhesion scale from the type table.
template <bool HAS_TYPE, bool HAS_POL>
__device__ __forceinline__ void apply_force_terms(const ForceParams& f, const float* disp, float d,
                                                  int type_i, int type_j, const Pol& pi,
                                                  const Pol& pj, float* p1, float* p2, float* q1,
                                                  float* q2) {
  apply_force_terms_k<HAS_POL>(f, disp, d, pair_kadh<HAS_TYPE>(f, f.table, type_i, type_j), pi, pj,
                               p1, p2, q1, q2);
}

// One invocation of the fused force, added term by term to the cell's running accumulators in
// the composition's order, as the reference's fused forces add to the same pack
// (kernel_fuser.hpp:40-55). Without polarity the only term is PiecewiseLinear (+0 when it is
// not applied). `kadh` is the pair's adhesion scale (pair_kadh).
template <bool HAS_POL>
__device__ __forceinline__ void apply_force_k(const ForceParams& f, const float* disp, float d,
                                              float kadh, const Pol& pi, const Pol& pj,
                                              float* acc_pos, float* acc_pol) {
  if (!HAS_POL) {
    PairTerms t;
    pair_terms_k<HAS_POL>(f, disp, d, kadh, pi, pj, t);
#pragma unroll
    for (int k = 0; k < 3; ++k) acc_pos[k] = acc_pos[k] + t.v[k];
    return;
  }
  float p1[3], p2[3], q1[2], q2[2];
  apply_force_terms_k<HAS_POL>(f, disp, d, kadh, pi, pj, p1, p2, q1, q2);
#pragma unroll
  for (int k = 0; k < 3; ++k) acc_pos[k] = (acc_pos[k] + p1[k]) + p2[k];
#pragma unroll
  for (int k = 0; k < 2; ++k) acc_pol[k] = (acc_pol[k] + q1[k]) + q2[k];
}

template <bool HAS_TYPE, bool HAS_POL>
__device__ __forceinline__ void apply_force(const ForceParams& f, const float* disp, float d,
                                            int type_i, int type_j, const Pol& pi, const Pol& pj,
                                            float* acc_pos, float* acc_pol) {
  apply_force_k<HAS_POL>(f, disp, d, pair_kadh<HAS_TYPE>(f, f.table, type_i, type_j), pi, pj,
                         acc_pos, acc_pol);
}

// The drag output of the fused force for pair (i, j) (binned_all_pairs.hpp:155-169: the
// functor adds to a zeroed pairwise drag).
template <bool HAS_TYPE>
__device__ __forceinline__ float pair_drag(const ForceParams& f, int type_i, int type_j) {
  if (HAS_TYPE && f.drag_by_type) return 0.0f + f.drag_table[type_i * f.n_types + type_j];
  return 0.0f + f.drag;
}

// A step's new state of a cell is not finite (a NaN or an infinity: overflow, or a NaN upload).
__device__ __forceinline__ bool not_finite(float4 x, float2 p, bool has_pol) {
  bool ok = isfinite(x.x) && isfinite(x.y) && isfinite(x.z);
  if (has_pol) ok = ok && isfinite(p.x) && isfinite(p.y);
  return !ok;
}

// Deterministic block sum of one float per thread (fixed butterfly + fixed wave order).
__device__ __forceinline__ float block_sum_f(float v, float* lds) {
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
__device__ __forceinline__ unsigned block_sum_u(unsigned v, unsigned* lds) {
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

}  // namespace tecs_amd
