// tecs_internal.hpp - host-side view of the device state and the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tecs_device.hpp"

namespace tecs_amd {

enum Finder {
  kNaiveAllPairs = 0,
  kBinnedAllPairs = 1,
  kNaiveGabriel = 2,
  kBinnedGabriel = 3,
  kNaiveAllPairsSym = 4,  // the symmetric finders: partners j < i, out(i) += v, out(j) -= v
  kBinnedAllPairsSym = 5,
  kNaiveGabrielSym = 6,
  kBinnedGabrielSym = 7,
  kNaiveDelaunay = 8  // work items (i, j, k), functor once per item with an empty sphere
};
// NaiveDelaunay keeps every position of an instance in LDS: capacity at most this
constexpr int kDelaunayMaxCells = 512;
inline bool is_symmetric(int finder) { return finder >= kNaiveAllPairsSym && finder <= kBinnedGabrielSym; }
// finders whose force evaluation only writes deltas (Delta mode); the COM fix and the
// integrator update run as kernels of their own
inline bool delta_only(int finder) { return is_symmetric(finder) || finder == kNaiveDelaunay; }
inline bool is_binned(int finder) {
  return finder == kBinnedAllPairs || finder == kBinnedGabriel || finder == kBinnedAllPairsSym ||
         finder == kBinnedGabrielSym;
}
inline bool is_gabriel(int finder) {
  return finder == kNaiveGabriel || finder == kBinnedGabriel || finder == kNaiveGabrielSym ||
         finder == kBinnedGabrielSym;
}

// What a sweep does with the force sum F of each cell (dt and the COM fix come from the
// instance descriptor / com buffer).
enum Mode {
  kModeDelta = 0,  // write F (and per-block COM partials when requested)
  kModePred = 1,   // Heun substep 1 without COM fix: d1 = F, x~ = x + d1*dt
  kModeCorr = 2,   // Heun substep 2 without COM fix: x += ((d1 + F)*0.5)*dt, old_vel
  kModeEuler = 3,  // Euler without COM fix: old_vel = F, x' = x + F*dt (ping-pong buffer)
  kModeCount = 4,  // pair export: number of pairs per agent
  kModeEmit = 5,   // pair export: the pairs, in visiting order
  kModeDeltaDrag = 6  // Delta with the drag correction (reads SweepArgs::ov_in)
};

struct SweepArgs {
  const InstanceDesc* inst;
  const int* block_inst;
  // binned sweeps: the 256-slot chunks this launch covers, one per workgroup (team sweeps: per T
  // workgroups), when it covers part of the batch (a size class); null: all chunks in order
  const int* chunks;
  const float4* xs;  // positions of the evaluated stage
  const float2* ps;  // polarities of the evaluated stage (may be null)
  const float4* trig;  // polar_trig of ps (kept current by every writer of a polarity buffer)
  const int* off;    // segment offsets of all instances (global slots; bin_ranges)
  const int* rel;    // row-relative bin offsets of the sparse instances (bin_ranges)
  const int* id;     // slot -> agent index
  float4* o_pos;     // Delta: F; Pred: d1; Euler: new positions
  float2* o_pol;
  float4* o2_pos;    // Pred: predicted positions; Corr: stage-0 positions (updated in place)
  float2* o2_pol;
  const float4* d1_pos;  // Corr: first evaluation
  const float2* d1_pol;
  float4* o_trig;        // Pred / Corr / Euler: polar_trig of the polarity written
  const int* skey;       // symmetric Delta: bin key of every slot at the last rebuild (null: the
                         // grid was built from these positions)
  const unsigned char* mover8;  // symmetric Delta on a stale grid: 1 for a slot outside its bin
  unsigned char* mover8_out = nullptr;  // the fused symmetric predictor's flags of these positions
                                // of the last rebuild (null: computed from skey per candidate)
  float4* q_pos;         // symmetric Delta on a stale grid: the movers' scattered contributions
  float2* q_pol;
  KeyCount kc;           // Corr / Euler: count_key of the new positions for the next rebuild
                         // (kc.keys null = not this step)
  float4* old_vel;       // Corr / Euler, may be null
  const float4* ov_in;   // DeltaDrag: the old velocities of the evaluated slots
  int* nonfinite;        // Corr / Euler: set to 1 for an instance whose new state is not
                         // finite (per instance; may be null)
  float4* com_partial;   // Delta: per-block sums of F (may be null)
  unsigned long long* block_pairs;  // per-block pair-interaction counters (accumulated)
  int* row_count;        // Count
  const int64_t* row_ptr;  // Emit
  int32_t* cols;           // Emit
  // NaiveDelaunay: the work items found per (cell, partner), row (ii * dl_row + i) of dl_row
  // counters per instance ii and cell i; zero between evaluations (the cell pass clears them)
  uint32_t* dl_cnt;
  int dl_row;
  int dl_splits;  // workgroups per cell of the counting pass
  int n_inst;
};

// Returns nonzero for a mode the finder does not have (symmetric finders and NaiveDelaunay:
// Delta, Count, Emit). A symmetric Delta writes o_pos / o_pol and, on a stale grid (skey set),
// adds the contributions of cells outside their rebuild bin to q_pos / q_pol, which must hold
// zeros.
int launch_sweep(int finder, bool has_type, bool has_pol, int mode, const SweepArgs& a,
                 int n_blocks, hipStream_t stream);
// NaiveDelaunay (kernels_delaunay.hip): the counting pass, then the per-cell pass of the mode
int launch_delaunay(bool has_type, bool has_pol, int mode, const SweepArgs& a, int n_blocks,
                    hipStream_t stream);
// d += q, q = 0 over all slots (dp / qp may be null)
void launch_fold_q(float4* d, float2* dp, float4* q, float2* qp, int total_slots,
                   hipStream_t stream);
// com partial of every 256-slot workgroup from the finished deltas d (+ q when non-null)
void launch_com_partials(const InstanceDesc* inst, const int* block_inst, const float4* d,
                         const float4* q, float4* partial, int n_blocks, hipStream_t stream);

struct UpdateArgs {
  const InstanceDesc* inst;
  const int* block_inst;
  const float4* com;  // per instance
  float4* x;          // stage-0 positions
  float2* pol;
  float4* d1_pos;     // first (or only) evaluation, COM-fixed in place
  float2* d1_pol;
  float4* d3_pos;        // second evaluation (Heun corrector)
  float2* d3_pol;
  float4* xp;            // predicted positions (Heun predictor)
  float2* polp;
  float4* trig;          // polar_trig of pol (corrector, Euler) ...
  float4* trigp;         // ... and of polp (predictor)
  float4* old_vel;
  float4* q_pos;         // symmetric evaluation on a stale grid: movers' contributions to add to
  float2* q_pol;         // the evaluated deltas (then cleared); null otherwise
  KeyCount kc;           // corrector / Euler: the next rebuild's keys of the new positions
                         // (kc.keys null: not counted here)
  int* nonfinite;        // corrector / Euler: 1 for an instance whose new state is not finite
  // predictor before a symmetric evaluation on a stale grid: mover8[s] = 1 for a cell whose
  // predicted position lies outside its bin of the last rebuild (keys mover_key), so the sweep
  // reads one byte per candidate instead of computing its bin
  const int* mover_key;
  unsigned char* mover8;
};
enum UpdateKind { kUpdatePred = 0, kUpdateCorr = 1, kUpdateEuler = 2 };
void launch_update(int kind, bool has_pol, const UpdateArgs& a, int n_blocks, hipStream_t stream);
// partials_per_block: COM partials written per 256-slot block (1, or the team count of the team
// sweep)
void launch_com_finalize(const InstanceDesc* inst, int n_inst, const float4* partial,
                         int partials_per_block, float4* com, hipStream_t stream);

// Team sweep (binned finders, search radius 1): team_lanes (4, 8 or 16) lanes per cell, launched
// as n_blocks * team_lanes workgroups; block_pairs and com_partial need
// n_blocks * kMaxTeamLanes entries.
constexpr int kMaxTeamLanes = 16;
// sym: a symmetric finder (Delta mode only; returns nonzero for another mode)
int launch_sweep_team(int team_lanes, bool gabriel, bool sym, bool has_type, bool has_pol,
                      int mode, const SweepArgs& a, int n_blocks, hipStream_t stream);

struct RebuildArgs {
  const InstanceDesc* inst;
  const int* block_inst;
  const float4* x_old;
  const float2* pol_old;
  const float4* trig_old;
  const int* id_old;
  const float4* ov_old;
  float4* x_new;
  float2* pol_new;
  float4* trig_new;
  int* id_new;
  float4* ov_new;
  KeyCount kc;     // keys, counts and (incremental) moves of the cells in the current order
  int* tmp;
  int* tkey;       // bin key of the cell scattered to each temporary slot
  int* tid;        // agent index of the cell scattered to each temporary slot
  bool keys_ready; // a sweep epilogue already ran count_key for these positions
  bool incremental;     // derive the new order from the last one (the moved counts and the
                        // arrival lists of kc) instead of counting and scattering every cell
  const int* coff_old;  // segment offsets of the last rebuild (global slots)
  int* coff;            // segment offsets being built (the bin offsets of dense grids)
  int* skey_new;        // key of every slot in the new order (the next rebuild's kc.skey)
  int* head_clear;      // arrival heads of the following rebuild, reset to -1 by the scan
  int* rel;             // row-relative bin offsets of the sparse instances (seg_of)
  int* dirty;           // per segment: a sparse row whose relative offsets must be rewritten
  int* scan_bsum;
  int total_cbins;  // segments of all instances (scan length)
  // row pass (only when some instance has a sparse grid)
  bool fill;
  const int* fill_inst;  // row workgroup -> instance
  const int* fill_base;  // row workgroup -> first local row (kFillRows rows per workgroup)
  int fill_blocks;
};
void launch_rebuild(const RebuildArgs& a, int n_blocks, hipStream_t stream);
// Rows of a sparse grid per row workgroup of the rebuild (kFillRows / 4 per wavefront: a
// wavefront writes its rows one after another, so it gets few of them).
constexpr int kFillRows = 64;
int scan_tiles_for(int n);
// Scans of at most this many tiles take the fused path (and leave the tiles' sums in bsum).
constexpr int kScanFusedTileLimit = 4096;
void launch_exclusive_scan(const int* in, int n, int* out, int* bsum, int* head_clear,
                           hipStream_t stream, bool tile_sums_ready = false);

void launch_init_ids(const InstanceDesc* inst, const int* block_inst, int* id, int n_blocks,
                     hipStream_t stream);
struct UploadArgs {
  const InstanceDesc* inst;
  int ii, block_begin;
  const int* id;
  int first, count;  // the host arrays hold cells first .. first + count - 1
  const float* pos;
  const float* pol;      // may be null
  const uint8_t* type;   // may be null
  float4* x;             // current generation
  float2* p;
  float4* trig;
  float4* x2;            // the other generation (inactive cells are kept in both)
  float2* p2;
  float4* trig2;
};
void launch_upload(const UploadArgs& a, int n_blocks, hipStream_t stream);

// set_agent_count: all slots into the other generation, the target instance in index order
struct ReorderArgs {
  const InstanceDesc* inst;
  const int* block_inst;
  int target;
  const int* id_old;
  int* id_new;
  const float4* x_old;
  float4* x_new;
  const float2* pol_old;  // may be null
  float2* pol_new;
  const float4* trig_old;
  float4* trig_new;
  const float4* ov_old;   // may be null
  float4* ov_new;
};
void launch_reorder_identity(const ReorderArgs& a, int n_blocks, hipStream_t stream);

// set_capacity: cells [from, to) of instance ii back to zero state in both generations
struct ClearArgs {
  const InstanceDesc* inst;
  int ii, from, to;
  float4* x[2];
  float2* p[2];
  float4* trig[2];
  float4* ov[2];
};
void launch_clear_cells(const ClearArgs& a, hipStream_t stream);
void launch_download(const InstanceDesc* inst, int ii, int block_begin, int n_blocks,
                     const int* id, const float4* x, const float2* p, const float4* ov,
                     float* pos, float* pol, float* old_vel, hipStream_t stream);

// trig[s] = polar_trig(pol[s]) for all slots of the batch
// Grid queries (kernels_query.hip) on instance 0 of a batch.
enum QueryKind { kQueryNearest = 0, kQueryRandomInBins = 1, kQueryRandomNeighbour = 2 };
struct QueryArgs {
  const InstanceDesc* inst;
  const float4* xs;     // current positions (slot order)
  const int* off;       // segment offsets of the grid (bin_ranges)
  const int* rel;       // row-relative bin offsets of a sparse grid (bin_ranges)
  const int* id;        // slot -> agent index
  int n;                // queries
  const float* points;  // n * 3
  const float* uniforms;
  int radius_bins;
  float radius;
  int32_t* out;
};
void launch_query(int kind, const QueryArgs& a, hipStream_t stream);
// The allocation-free nearest query stages its points and results in a fixed device array
// (kernels_query.hip): at most kNoAllocMaxQueries queries per call.
constexpr int kNoAllocMaxQueries = 65536;
constexpr size_t kNoAllocOutOffset = (size_t)kNoAllocMaxQueries * 3 * sizeof(float);
constexpr size_t kNoAllocScratchBytes = kNoAllocOutOffset + (size_t)kNoAllocMaxQueries * sizeof(int32_t);
// The current device's address of that array (hipGetSymbolAddress, resolved once per device);
// returns a hipError_t value, 0 on success.
int noalloc_scratch(void** out);

// Generic wall forces, link forces and division (kernels_extra.hip).
// inv[slot_begin + agent] = slot over all slots
void launch_inverse_ids(const InstanceDesc* inst, const int* block_inst, const int* id, int* inv,
                        int n_blocks, hipStream_t stream);
// per active cell: d = W + d for the instance's walls, then its link contributions (rows over
// agent space slot_begin + agent, ent = (other endpoint, +1 for the link's a end, -1 for its b
// end) in link order; rows null: no links)
void launch_extra(const InstanceDesc* inst, const int* block_inst, const int* id, const int* inv,
                  const int* rows, const int2* ent, const float4* xs, float4* d, int n_blocks,
                  hipStream_t stream);
void launch_divide(const InstanceDesc* inst, int ii, const int* inv, const int* parents,
                   const float* dirs, int count, int first, float offset, float4* x_cur,
                   float4* x_other, hipStream_t stream);

void launch_polar_trig(const float2* pol, float4* trig, int total_slots, hipStream_t stream);
void launch_copy_probe(const float4* src, float4* dst, int64_t n, hipStream_t stream);

}  // namespace tecs_amd
