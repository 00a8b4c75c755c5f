// batch.hpp - the device state of one or more simulations and their step schedule.
//
// A Batch owns the device state of one or more independent simulations that share the pair
// finder, integrator, COM fixer and field set; each step runs one launch per kernel for all of
// them. tecs_sim is a batch of one; tecs_batch exposes the general form; the server keeps one
// batch per compatible group of requests.
//
// Step schedule (no HIP allocation, memset or host synchronisation inside a step):
//   Heun, NoComFixer:     [rebuild] sweep(pred: d1, x~) -> sweep(corr: x, old_vel)
//   Heun, GlobalComFixer: [rebuild] sweep(delta) -> com -> pred -> sweep(delta) -> com -> corr
//   Euler, NoComFixer:    [rebuild] sweep(euler: x', old_vel) (ping-pong position buffers)
//   Euler, GlobalComFixer:[rebuild] sweep(delta) -> com -> euler update
// [rebuild] follows the reference schedule (binned_all_pairs.hpp:77-84).
#pragma once

#include "../../include/tecs_amd.h"
#include "io_snapshot.hpp"
#include "tecs_internal.hpp"

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

namespace tecs_amd {

extern thread_local std::string g_err;

inline int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}


#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(e_ == hipErrorOutOfMemory ? TECS_ERR_OOM : TECS_ERR_DEVICE,          \
                  std::string(#expr) + " failed: " + hipGetErrorString(e_));           \
  } while (0)

#define TRY(expr)            \
  do {                       \
    const int rc_ = (expr);  \
    if (rc_ != TECS_OK) return rc_; \
  } while (0)

enum TimingClass { kTimeSweepFull = 0, kTimeSweepStale = 1, kTimeRebuild = 2, kTimeClasses = 3 };


// The terms of a force in the order they apply: term_order up to its first TECS_TERM_END, or
// PiecewiseLinear, bending, alignment when it is all zeros. pos[t] = place of term t, -1 if
// it is not applied.
inline void term_places(const tecs_force_desc& f, int pos[4]) {
  pos[0] = pos[1] = pos[2] = pos[3] = -1;
  if (f.term_order[0] == TECS_TERM_END && f.term_order[1] == TECS_TERM_END &&
      f.term_order[2] == TECS_TERM_END) {
    for (int t = 1; t <= 3; ++t) pos[t] = t - 1;
    return;
  }
  for (int k = 0; k < 3 && f.term_order[k] != TECS_TERM_END; ++k) pos[f.term_order[k]] = k;
}

inline bool force_has_drag(const tecs_force_desc& f) {
  if (f.drag_by_type) {
    for (int k = 0; k < f.n_types * f.n_types; ++k)
      if (f.drag_table[k] != 0.0f) return true;
    return false;
  }
  return f.drag != 0.0f;
}

inline int validate_force(const tecs_force_desc& f, const tecs_sim_desc& d) {
  const uint32_t fields = d.fields;
  if (f.n_types < 0 || f.n_types > TECS_MAX_TYPES) return fail(TECS_ERR_INVALID, "bad n_types");
  if (f.n_types > 0 && !(fields & TECS_FIELD_TYPE))
    return fail(TECS_ERR_INVALID, "type table needs TECS_FIELD_TYPE");
  bool seen[4] = {false, false, false, false};
  bool ended = false;
  for (int k = 0; k < 3; ++k) {
    const int t = f.term_order[k];
    if (t < TECS_TERM_END || t > TECS_TERM_ALIGNMENT) return fail(TECS_ERR_INVALID, "bad term in term_order");
    if (t == TECS_TERM_END) {
      ended = true;
      continue;
    }
    if (ended) return fail(TECS_ERR_INVALID, "term_order continues after TECS_TERM_END");
    if (seen[t]) return fail(TECS_ERR_INVALID, "a term appears twice in term_order");
    seen[t] = true;
  }
  int pos[4];
  term_places(f, pos);
  if (((pos[TECS_TERM_BENDING] >= 0 && f.bending_strength != 0.0f) ||
       (pos[TECS_TERM_ALIGNMENT] >= 0 && f.alignment_strength != 0.0f)) &&
      !(fields & TECS_FIELD_POLARITY))
    return fail(TECS_ERR_INVALID, "polarity forces need TECS_FIELD_POLARITY");
  if (f.drag_by_type != 0 && f.drag_by_type != 1) return fail(TECS_ERR_INVALID, "drag_by_type must be 0 or 1");
  if (f.drag_by_type && f.n_types == 0) return fail(TECS_ERR_INVALID, "drag by type needs a type table");
  if (!std::isfinite(f.drag)) return fail(TECS_ERR_INVALID, "drag must be finite");
  for (int k = 0; k < TECS_MAX_TYPES * TECS_MAX_TYPES; ++k)
    if (!std::isfinite(f.drag_table[k])) return fail(TECS_ERR_INVALID, "drag table entries must be finite");
  if (force_has_drag(f)) {
    if (!d.track_old_velocities)
      return fail(TECS_ERR_INVALID, "drag needs the old velocities (track_old_velocities = 1)");
    if (is_symmetric(d.pair_finder))
      return fail(TECS_ERR_UNSUPPORTED, "drag with a symmetric pair finder is not implemented");
  }
  return TECS_OK;
}

inline int validate_desc(const tecs_sim_desc* d) {
  if (!d) return fail(TECS_ERR_INVALID, "null descriptor");
  if (d->capacity < d->agent_count) return fail(TECS_ERR_INVALID, "capacity < agent_count");
  if (d->agent_count > (1u << 30)) return fail(TECS_ERR_INVALID, "agent_count too large");
  if (d->pair_finder < 0 || d->pair_finder > kNaiveDelaunay)
    return fail(TECS_ERR_INVALID, "bad pair_finder");
  if (d->pair_finder == kNaiveDelaunay && d->capacity > (uint32_t)kDelaunayMaxCells)
    return fail(TECS_ERR_INVALID, "NaiveDelaunay takes at most 512 cells (capacity)");
  if (d->integrator != TECS_EULER && d->integrator != TECS_HEUN)
    return fail(TECS_ERR_INVALID, "bad integrator");
  if (d->com_fixer != TECS_NO_COM_FIXER && d->com_fixer != TECS_GLOBAL_COM_FIXER)
    return fail(TECS_ERR_INVALID, "bad com_fixer");
  TRY(validate_force(d->force, *d));
  if (!(d->cutoff_distance > 0.0f) || !std::isfinite(d->cutoff_distance))
    return fail(TECS_ERR_INVALID, "cutoff_distance must be finite and > 0");
  if (is_binned(d->pair_finder)) {
    const tecs_pair_finder_settings& s = d->pair_finder_settings;
    if (!(s.bin_size_scale > 0.0f)) return fail(TECS_ERR_INVALID, "bin_size_scale must be > 0");
    const float bin = s.bin_size_scale * d->cutoff_distance;
    double bins = 1.0;
    for (int k = 0; k < 3; ++k) {
      if (!(s.max_bounds[k] > s.min_bounds[k]))
        return fail(TECS_ERR_INVALID, "max_bounds must exceed min_bounds");
      bins *= std::ceil((double)((s.max_bounds[k] - s.min_bounds[k]) / bin));
    }
    if (bins > 1.0e9) return fail(TECS_ERR_INVALID, "grid has more than 1e9 bins");
    if (s.rebuild_every_n < 0) return fail(TECS_ERR_INVALID, "rebuild_every_n must be >= 0");
  }
  return TECS_OK;
}

// Device memory kept for reuse by batches that come and go: the server re-forms a group's
// batch whenever its membership changes, and the new batch asks for buffers of about the
// sizes the old one returns. Blocks are handed out by size class (multiples of 1/8 of the
// request's power of two, at least 64 KiB, so at most 12.5% is wasted) and returned on
// release; hipFree, which synchronises the whole device, runs only when the cache exceeds its
// limit or the pool is destroyed. Callers return a block only after the stream that used it
// is idle.
class DevicePool {
 public:
  explicit DevicePool(size_t max_cached_bytes) : limit_(max_cached_bytes) {}
  ~DevicePool() { trim(0); }
  DevicePool(const DevicePool&) = delete;
  DevicePool& operator=(const DevicePool&) = delete;

  static size_t size_class(size_t bytes) {
    size_t top = 1;
    while (top <= bytes / 2) top <<= 1;  // largest power of two <= bytes
    const size_t g = std::max<size_t>(65536, top / 8);
    return (std::max<size_t>(bytes, 1) + g - 1) / g * g;
  }
  hipError_t get(size_t bytes, void** p) {
    const size_t c = size_class(bytes);
    auto it = free_.find(c);
    if (it != free_.end()) {
      *p = it->second;
      cached_ -= c;
      free_.erase(it);
      return hipSuccess;
    }
    hipError_t e = hipMalloc(p, c);
    if (e == hipErrorOutOfMemory && cached_ > 0) {
      (void)hipGetLastError();
      trim(0);
      e = hipMalloc(p, c);
    }
    return e;
  }
  void put(void* p, size_t cls) {
    free_.emplace(cls, p);
    cached_ += cls;
    if (cached_ > limit_) trim(limit_);
  }
  void trim(size_t keep) {
    while (cached_ > keep && !free_.empty()) {
      auto it = std::prev(free_.end());  // largest first
      hipFree(it->second);
      cached_ -= it->first;
      free_.erase(it);
    }
  }
  size_t cached_bytes() const { return cached_; }

 private:
  std::multimap<size_t, void*> free_;
  size_t cached_ = 0;
  size_t limit_;
};

struct Batch {
  // homogeneous configuration
  int finder = 0, integrator = 0, com_fixer = 0, rebuild_every_n = 0;
  bool has_pol = false, has_type = false, track_ov = true;
  std::vector<tecs_sim_desc> descs;
  std::vector<InstanceDesc> h_inst;
  std::vector<int> h_block_inst;
  int n_inst = 0, n_blocks = 0, total_slots = 0, max_n = 0;
  // cells the host staging buffers, the pair-export rows and the snapshot slots hold: the
  // largest member slot range, which bounds every agent count set_capacity can allow
  int stage_n = 0;
  int total_cbins = 0;         // counting-sort segments of all instances
  bool sparse_rows = false;    // some instance has a sparse grid (counts per row of bins)
  int total_rel = 0;           // row-relative bin offsets of all sparse instances
  std::vector<int> h_fill_inst, h_fill_base;  // row workgroups: instance, first row
  hipStream_t stream = nullptr;
  bool own_stream = false;
  bool inst_dirty = false;
  int device = 0;  // the HIP device current at creation (all state lives there)

  // device state
  InstanceDesc* d_inst = nullptr;
  int* d_block_inst = nullptr;
  float4* x[2] = {nullptr, nullptr};
  float2* pol[2] = {nullptr, nullptr};
  int* id[2] = {nullptr, nullptr};
  float4* ov[2] = {nullptr, nullptr};
  float4* xp = nullptr;
  float2* polp = nullptr;
  float4* d1 = nullptr;
  float2* d1p = nullptr;
  float4* d3 = nullptr;
  float2* d3p = nullptr;
  unsigned char* mover8 = nullptr;  // binned symmetric finders: the predictor's mover flags
  float4* q = nullptr;   // binned symmetric finders: contributions scattered by movers (zero
  float2* qp = nullptr;  // between evaluations; the consuming update clears what it adds)
  int* keys = nullptr;
  int* tmp = nullptr;
  int* tkey = nullptr;
  int* tidb = nullptr;
  int* cnt = nullptr;
  int* coff[2] = {nullptr, nullptr};   // segment offsets of the order held in id[b]
  int* rel = nullptr;                  // row-relative bin offsets (sparse grids only)
  int* dirty = nullptr;                // per segment: a sparse row to rewrite at the next rebuild
  int* skey[2] = {nullptr, nullptr};   // key of every slot of the order held in id[b]
  int* head[2] = {nullptr, nullptr};   // incremental rebuild: arrival lists per segment
  int* nxt = nullptr;
  unsigned long long* movers = nullptr;  // incremental rebuild: marks of cells that changed segment
  int* d_fill_inst = nullptr;
  int* d_fill_base = nullptr;
  int* bsum = nullptr;
  unsigned long long* block_pairs = nullptr;
  float4* com_partial = nullptr;
  float4* com = nullptr;
  int* nonfinite = nullptr;  // per member: 1 once a step ended with a non-finite position or
                             // polarity (set by the corrector / Euler kernels, never cleared by them)
  float* st_pos = nullptr;
  float* st_pol = nullptr;
  float* st_ov = nullptr;
  uint8_t* st_type = nullptr;
  int* d_row_count = nullptr;
  int64_t* d_row_ptr = nullptr;
  int32_t* d_cols = nullptr;
  int64_t cols_cap = 0;
  uint32_t* dl_cnt = nullptr;  // NaiveDelaunay: counters of the work items found (SweepArgs)
  // generic wall forces (in the instance table) and link forces (kernels_extra.hip)
  std::vector<std::vector<int32_t>> member_links;  // per member: a0, b0, a1, b1, ...
  std::vector<float> member_link_k;
  int* d_link_rows = nullptr;   // incident links per cell: rows over agent space (slot_begin + agent)
  int2* d_link_ent = nullptr;   // (other endpoint, +1 a end / -1 b end), link order per cell
  int64_t link_ent_cap = 0;
  int n_links = 0;              // incidences (two per link with distinct ends)
  int* inv = nullptr;     // slot of every agent (links, division)
  int* st_idx = nullptr;  // staging of agent indices (division)
  int dl_row = 0, dl_splits = 1;
  DevicePool* pool = nullptr;                          // optional source of device memory
  std::vector<std::pair<void*, size_t>> allocations;  // (block, size class; 0 = hipMalloc)

  int variant = TECS_SWEEP_AUTO;  // requested sweep variant
  int tuned = -1;                 // concrete variant picked by the autotuner (AUTO)
  // Autotuning per size class (AUTO, a binned finder whose team sweep is exact, no COM fixer,
  // ordered pairs): the members are grouped by cell count in powers of four and each class
  // gets the variant that sweeps its chunks fastest; the classes that picked the same variant
  // form a group swept by one launch over its chunks. One group: `tuned`, no chunk lists.
  static constexpr int kMaxClasses = 16;
  std::vector<int> member_class;  // size class of every member (filled by the autotuner)
  int class_variant[kMaxClasses];  // the variant a class picked (-1: no member in it)
  double class_tune_ms[kMaxClasses][6];
  struct SweepGroup {
    int variant;
    int n_chunks;
    int* chunks;  // device list of the group's 256-slot chunks
  };
  std::vector<SweepGroup> groups;
  bool team_ok = false;           // the team sweep reproduces this batch exactly
  float4* trig[2] = {nullptr, nullptr};  // polar_trig of pol[b] (kept current by every writer)
  float4* trigp = nullptr;                // polar_trig of polp
  unsigned long long* tune_pairs = nullptr;  // scratch counters for autotuning sweeps
  double tune_ms[6] = {0, 0, 0, 0, 0, 0};
  int xc = 0, ic = 0;  // current position/polarity/old-velocity buffer, current id buffer
  int step_count = 0;
  bool grid_built = false;
  // the last step's final sweep stored the keys of x[xc] and counted them (count_key), so the
  // rebuild that starts the next step skips k_bin_keys
  bool keys_ready = false;
  // the next rebuild derives the new order from the last one: a grid exists and the state has
  // only moved by steps since (an upload or a member copy clears this)
  bool inc_ok = false;
  // the last maybe_rebuild built the grid: the evaluation that follows sees a fresh grid
  bool rebuilt_now = false;
  tecs_stats stats{};

  bool timing = false;
  struct Pending {
    hipEvent_t a, b;
    int cls;
  };
  std::vector<hipEvent_t> ev_pool;
  std::vector<Pending> pending;
  double t_ms[kTimeClasses] = {0, 0, 0};
  int64_t t_launches[kTimeClasses] = {0, 0, 0};

  // hipGraph replay of step sequences: the host-side state (buffer parities, rebuild phase)
  // repeats with period graph_period(); a sequence of that many steps captured from a state
  // can be replayed whenever the batch is in that state again. Kernels read dt from the
  // device-side instance table, so a dt change does not invalidate a graph.
  int graph_mode = TECS_GRAPHS_AUTO;
  struct HostState {
    int xc, ic, step_count;
    bool keys_ready, inc_ok;
    bool operator==(const HostState& o) const {
      return xc == o.xc && ic == o.ic && step_count == o.step_count &&
             keys_ready == o.keys_ready && inc_ok == o.inc_ok;
    }
  };
  struct StepGraph {
    HostState start;
    hipGraphExec_t exec;
    tecs_stats delta;  // counters advanced by one replay
  };
  std::vector<StepGraph> graphs;
  int64_t graph_launches = 0;

  // snapshot hand-off (io_snapshot.hip): gather on `stream`, copy on `side`, write elsewhere
  static constexpr int kSnapshotSlots = 3;
  SnapshotSlot snap[kSnapshotSlots];
  std::mutex snap_mu;
  std::condition_variable snap_cv;
  hipStream_t side = nullptr;
  std::vector<unsigned> write_step_count;  // Simulation::current_step per member

  HostState host_state() const {
    return HostState{xc, ic, step_count, keys_ready, inc_ok};
  }
  int graph_period() const { return 2 * std::max(1, is_binned(finder) ? rebuild_every_n : 1); }

  // Executable graphs that ran once and are not kept (a non-periodic capture); destroyed with
  // the others once the stream is idle.
  std::vector<hipGraphExec_t> retired;

  // Destroys every executable graph. A launched graph may still be running on the stream, so
  // the stream is drained first.
  void drop_graphs() {
    if ((!graphs.empty() || !retired.empty()) && stream) hipStreamSynchronize(stream);
    for (auto& g : graphs) hipGraphExecDestroy(g.exec);
    graphs.clear();
    for (hipGraphExec_t e : retired) hipGraphExecDestroy(e);
    retired.clear();
  }

  ~Batch() { release(); }

  void release() {
    if (stream) hipStreamSynchronize(stream);
    {
      // snapshots in flight still reference the slots: wait for the writer to finish them
      std::unique_lock<std::mutex> g(snap_mu);
      snap_cv.wait(g, [&] {
        for (auto& sl : snap)
          if (sl.busy) return false;
        return true;
      });
    }
    for (auto& sl : snap) {
      if (sl.h_pos) hipHostFree(sl.h_pos);
      if (sl.h_pol) hipHostFree(sl.h_pol);
      if (sl.gathered) hipEventDestroy(sl.gathered);
      if (sl.copied) hipEventDestroy(sl.copied);
      sl = SnapshotSlot{};
    }
    if (side) hipStreamDestroy(side);
    side = nullptr;
    drop_graphs();
    for (auto& p : pending) {
      hipEventDestroy(p.a);
      hipEventDestroy(p.b);
    }
    pending.clear();
    for (hipEvent_t e : ev_pool) hipEventDestroy(e);
    ev_pool.clear();
    for (auto& a : allocations) {
      if (pool && a.second)
        pool->put(a.first, a.second);
      else
        hipFree(a.first);
    }
    allocations.clear();
    if (own_stream && stream) hipStreamDestroy(stream);
    stream = nullptr;
  }

  template <typename T>
  int alloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    void* v = nullptr;
    if (pool) {
      HIP_TRY(pool->get(count * sizeof(T), &v));
      allocations.emplace_back(v, DevicePool::size_class(count * sizeof(T)));
    } else {
      HIP_TRY(hipMalloc(&v, count * sizeof(T)));
      allocations.emplace_back(v, 0);
    }
    *p = static_cast<T*>(v);
    return TECS_OK;
  }

  int init(const tecs_sim_desc* ds, int count, void* user_stream) {
    if (count <= 0) return fail(TECS_ERR_INVALID, "batch needs at least one simulation");
    for (int k = 0; k < count; ++k) TRY(validate_desc(&ds[k]));
    const tecs_sim_desc& d0 = ds[0];
    finder = d0.pair_finder;
    integrator = d0.integrator;
    com_fixer = d0.com_fixer;
    rebuild_every_n = d0.pair_finder_settings.rebuild_every_n;
    has_pol = (d0.fields & TECS_FIELD_POLARITY) != 0;
    has_type = (d0.fields & TECS_FIELD_TYPE) != 0;
    track_ov = d0.track_old_velocities != 0;
    for (int k = 1; k < count; ++k) {
      const tecs_sim_desc& d = ds[k];
      if (d.pair_finder != finder || d.integrator != integrator || d.com_fixer != com_fixer ||
          d.fields != d0.fields || (d.track_old_velocities != 0) != track_ov ||
          (is_binned(finder) && d.pair_finder_settings.rebuild_every_n != rebuild_every_n))
        return fail(TECS_ERR_INVALID,
                    "batched simulations must share pair finder, integrator, com fixer, fields, "
                    "old-velocity tracking and rebuild period");
    }
    descs.assign(ds, ds + count);
    n_inst = count;
    write_step_count.assign(count, 0u);

    // layout: slots padded per instance to whole workgroups; segments concatenated (+1 pad segment)
    int64_t slot = 0, blocks = 0;
    h_inst.resize(count);
    for (int k = 0; k < count; ++k) {
      const tecs_sim_desc& d = ds[k];
      InstanceDesc& in = h_inst[k];
      std::memset(&in, 0, sizeof(in));
      in.slot_begin = (int)slot;
      in.n = (int)d.agent_count;
      in.cap = (int)d.capacity;
      // slots for the whole capacity: set_agent_count moves within it
      const int nb = std::max(1, (int)((d.capacity + kBlock - 1) / kBlock));
      in.padded = nb * kBlock;
      in.block_begin = (int)blocks;
      in.n_blocks = nb;
      const float cutoff = d.cutoff_distance;
      in.cutoff2 = cutoff * cutoff;  // binned_all_pairs.hpp:33
      const tecs_pair_finder_settings& ps = d.pair_finder_settings;
      in.gabriel_factor = ps.gabriel_coefficient * ps.gabriel_coefficient * 0.25f;
      in.gabriel_c2 = ps.gabriel_coefficient * ps.gabriel_coefficient;
      if (is_binned(finder)) {
        // binned_all_pairs.hpp:36-38, grid.hpp:201-215
        in.bin_size = ps.bin_size_scale * cutoff;
        in.inv_bin = 1.0f / in.bin_size;
        {
          int e = 0;
          const float m = std::frexp(in.bin_size, &e);  // bin = m * 2^e, m in [0.5, 1)
          in.bin_pow2 = (m == 0.5f && std::isfinite(in.inv_bin) && in.inv_bin * in.bin_size == 1.0f) ? 1 : 0;
        }
        in.search_radius = (int)std::ceil(cutoff * in.inv_bin);
        int bc = 1;
        for (int q = 0; q < 3; ++q) {
          in.minb[q] = ps.min_bounds[q];
          in.ext[q] = (int)std::ceil((ps.max_bounds[q] - ps.min_bounds[q]) / in.bin_size);
          bc *= in.ext[q];
        }
        in.bin_count = bc;
      } else {
        in.bin_size = 1.0f;
        in.inv_bin = 1.0f;
        in.bin_count = 0;
      }
      // sparse grids (many bins per cell, e.g. a sheet of cells on a sphere) count per row of
      // bins, so the scan does not run over every empty bin and a rebuild rewrites the bin
      // offsets of the rows whose members changed bins only (seg_of); dense grids count per bin
      in.rel_begin = -1;
      if (is_binned(finder) && in.n > 0 && (double)in.bin_count / (double)in.n >= 8.0) {
        in.rel_begin = 0;
        // division of a key (< 2^31) by ext[0] as a multiply: with l = ceil(log2 e), the
        // multiplier ceil(2^(31 + l) / e) fits 32 bits and the quotient is exact
        const uint32_t e = (uint32_t)in.ext[0];
        int l = 0;
        while ((1u << l) < e) ++l;
        in.row_mul = e == 1u ? 0u : (uint32_t)((((uint64_t)1 << (31 + l)) + e - 1) / e);
        in.row_shift = e == 1u ? 0 : l - 1;
      }
      in.cbin_count = (in.rel_begin >= 0 ? in.ext[1] * in.ext[2] : in.bin_count) + 1;
      in.dt = 0.0f;
      load_force(in.force, d.force);
      slot += in.padded;
      blocks += nb;
      max_n = std::max(max_n, in.cap);
      stage_n = std::max(stage_n, in.padded);
      if (slot > (int64_t)1 << 30)
        return fail(TECS_ERR_INVALID, "batch too large for 32-bit slot/bin indices");
    }
    total_slots = (int)slot;
    n_blocks = (int)blocks;
    sparse_rows = false;
    for (int k = 0; k < count; ++k) sparse_rows = sparse_rows || h_inst[k].rel_begin >= 0;
    {
      // segment and row-relative offset ranges of the instances, each starting on a multiple of
      // 64 entries (gap entries keep count 0, so the scan carries the next instance's start)
      int64_t cb = 0, rb = 0;
      for (int k = 0; k < count; ++k) {
        InstanceDesc& in = h_inst[k];
        in.cbin_begin = (int)cb;
        cb += ((in.cbin_count + 63) / 64) * 64;
        if (in.rel_begin >= 0) {
          in.rel_begin = (int)rb;
          rb += (((int64_t)(in.cbin_count - 1) * (in.ext[0] + 1) + 63) / 64) * 64;
          // row workgroups of kFillRows rows each
          for (int q = 0; q < in.cbin_count - 1; q += kFillRows) {
            h_fill_inst.push_back(k);
            h_fill_base.push_back(q);
          }
        }
      }
      if (cb > (int64_t)2000000000 || rb > (int64_t)2000000000)
        return fail(TECS_ERR_INVALID, "batch too large for 32-bit bin indices");
      total_cbins = (int)cb;
      total_rel = (int)rb;
    }
    h_block_inst.resize(n_blocks);
    for (int k = 0; k < count; ++k)
      for (int b = 0; b < h_inst[k].n_blocks; ++b) h_block_inst[h_inst[k].block_begin + b] = k;

    HIP_TRY(hipGetDevice(&device));
    if (user_stream) {
      stream = static_cast<hipStream_t>(user_stream);
    } else {
      HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
      own_stream = true;
    }

    const size_t S = (size_t)total_slots;
    TRY(alloc(&d_inst, n_inst));
    TRY(alloc(&d_block_inst, n_blocks));
    // positions carry one slot past the last: the all-pairs walk loads two candidates at a
    // time and may read that slot (its result unused)
    for (int b = 0; b < 2; ++b) {
      TRY(alloc(&x[b], S + 1));
      TRY(alloc(&id[b], S));
      if (has_pol) TRY(alloc(&pol[b], S));
      if (track_ov) TRY(alloc(&ov[b], S));
    }
    TRY(alloc(&d1, S));
    if (has_pol) TRY(alloc(&d1p, S));
    if (integrator == TECS_HEUN) {
      TRY(alloc(&xp, S + 1));
      if (has_pol) TRY(alloc(&polp, S));
    }
    bool any_drag = false;
    for (int k = 0; k < count; ++k) any_drag = any_drag || force_has_drag(ds[k].force);
    // the Heun corrector's own delta buffer: every step path that evaluates into deltas before
    // the update (COM fixer, delta-only finders, drag; walls and links allocate it when set)
    if (integrator == TECS_HEUN && (com_fixer == TECS_GLOBAL_COM_FIXER || delta_only(finder) || any_drag)) {
      TRY(alloc(&d3, S));
      if (has_pol) TRY(alloc(&d3p, S));
    }
    if (is_symmetric(finder) && is_binned(finder)) {
      TRY(alloc(&q, S));
      HIP_TRY(hipMemsetAsync(q, 0, sizeof(float4) * S, stream));
      TRY(alloc(&mover8, S));
      if (has_pol) {
        TRY(alloc(&qp, S));
        HIP_TRY(hipMemsetAsync(qp, 0, sizeof(float2) * S, stream));
      }
    }
    if (is_binned(finder)) {
      TRY(alloc(&keys, S));
      TRY(alloc(&tmp, S));
      TRY(alloc(&tkey, S));
      TRY(alloc(&tidb, S));
      TRY(alloc(&cnt, (size_t)total_cbins));
      TRY(alloc(&nxt, S));
      TRY(alloc(&movers, S / 64));
      HIP_TRY(hipMemsetAsync(movers, 0, sizeof(unsigned long long) * (S / 64), stream));
      for (int b = 0; b < 2; ++b) {
        TRY(alloc(&coff[b], (size_t)total_cbins + 1));
        TRY(alloc(&skey[b], S));
        TRY(alloc(&head[b], (size_t)total_cbins));
        HIP_TRY(hipMemsetAsync(coff[b], 0, sizeof(int) * ((size_t)total_cbins + 1), stream));
        HIP_TRY(hipMemsetAsync(head[b], 0xff, sizeof(int) * (size_t)total_cbins, stream));
      }
      if (sparse_rows) {
        TRY(alloc(&rel, (size_t)total_rel));
        TRY(alloc(&dirty, (size_t)total_cbins));
        TRY(alloc(&d_fill_inst, h_fill_inst.size()));
        TRY(alloc(&d_fill_base, h_fill_base.size()));
        HIP_TRY(hipMemcpyAsync(d_fill_inst, h_fill_inst.data(), sizeof(int) * h_fill_inst.size(),
                               hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemcpyAsync(d_fill_base, h_fill_base.data(), sizeof(int) * h_fill_base.size(),
                               hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemsetAsync(rel, 0, sizeof(int) * (size_t)total_rel, stream));
        HIP_TRY(hipMemsetAsync(dirty, 0, sizeof(int) * (size_t)total_cbins, stream));
      }
      TRY(alloc(&bsum, (size_t)scan_tiles_for(total_cbins)));
      HIP_TRY(hipMemsetAsync(cnt, 0, sizeof(int) * (size_t)total_cbins, stream));
    }
    // team sweep conditions: a binned finder, search radius 1 and, for Gabriel, c <= 0.95
    // (kernels_team.hip)
    team_ok = is_binned(finder);
    for (int k = 0; k < count; ++k) {
      if (h_inst[k].search_radius != 1) team_ok = false;
      const float c = ds[k].pair_finder_settings.gabriel_coefficient;
      if (is_gabriel(finder) && !(c >= 0.0f && c <= 0.95f)) team_ok = false;
    }
    if (finder == kNaiveDelaunay) {
      // a row of counters per slot of the largest member; the cell pass leaves them at zero
      for (int k = 0; k < count; ++k) dl_row = std::max(dl_row, h_inst[k].padded);
      dl_splits = std::max(1, dl_row / 128);
      const size_t cnt_len = (size_t)n_inst * dl_row * dl_row;
      TRY(alloc(&dl_cnt, cnt_len));
      HIP_TRY(hipMemsetAsync(dl_cnt, 0, sizeof(uint32_t) * cnt_len, stream));
    }
    TRY(alloc(&block_pairs, (size_t)n_blocks * kMaxTeamLanes));
    TRY(alloc(&tune_pairs, (size_t)n_blocks * kMaxTeamLanes));
    if (has_pol) {
      TRY(alloc(&trig[0], S));
      TRY(alloc(&trig[1], S));
      if (integrator == TECS_HEUN) TRY(alloc(&trigp, S));
    }
    TRY(alloc(&com_partial, (size_t)n_blocks * kMaxTeamLanes));
    TRY(alloc(&com, n_inst));
    TRY(alloc(&nonfinite, n_inst));
    HIP_TRY(hipMemsetAsync(nonfinite, 0, sizeof(int) * n_inst, stream));
    HIP_TRY(hipMemsetAsync(block_pairs, 0, sizeof(unsigned long long) * n_blocks * kMaxTeamLanes, stream));
    HIP_TRY(hipMemsetAsync(com, 0, sizeof(float4) * n_inst, stream));
    for (int b = 0; b < 2; ++b) {
      HIP_TRY(hipMemsetAsync(x[b], 0, sizeof(float4) * S, stream));
      if (has_pol) HIP_TRY(hipMemsetAsync(pol[b], 0, sizeof(float2) * S, stream));
      if (track_ov) HIP_TRY(hipMemsetAsync(ov[b], 0, sizeof(float4) * S, stream));
      if (has_pol) launch_polar_trig(pol[b], trig[b], total_slots, stream);
    }
    HIP_TRY(hipGetLastError());
    const size_t M = (size_t)std::max(stage_n, 1);
    TRY(alloc(&st_pos, 3 * M));
    TRY(alloc(&st_ov, 3 * M));
    if (has_pol) TRY(alloc(&st_pol, 2 * M));
    if (has_type) TRY(alloc(&st_type, M));
    TRY(alloc(&d_row_count, M));
    TRY(alloc(&d_row_ptr, M + 1));
    HIP_TRY(hipMemcpyAsync(d_inst, h_inst.data(), sizeof(InstanceDesc) * n_inst,
                           hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(d_block_inst, h_block_inst.data(), sizeof(int) * n_blocks,
                           hipMemcpyHostToDevice, stream));
    for (int b = 0; b < 2; ++b) launch_init_ids(d_inst, d_block_inst, id[b], n_blocks, stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(stream));
    return TECS_OK;
  }

  // ---------------------------------------------------------------------- timing
  int ev_get(hipEvent_t* e) {
    if (!ev_pool.empty()) {
      *e = ev_pool.back();
      ev_pool.pop_back();
      return TECS_OK;
    }
    HIP_TRY(hipEventCreate(e));
    return TECS_OK;
  }
  int time_begin(Pending* p, int cls) {
    if (!timing) return TECS_OK;
    TRY(ev_get(&p->a));
    TRY(ev_get(&p->b));
    p->cls = cls;
    HIP_TRY(hipEventRecord(p->a, stream));
    return TECS_OK;
  }
  int time_end(Pending* p) {
    if (!timing) return TECS_OK;
    HIP_TRY(hipEventRecord(p->b, stream));
    pending.push_back(*p);
    if (pending.size() > 4096) TRY(fold_timing());
    return TECS_OK;
  }
  int fold_timing() {
    if (pending.empty()) return TECS_OK;
    HIP_TRY(hipEventSynchronize(pending.back().b));
    for (auto& p : pending) {
      float ms = 0.0f;
      HIP_TRY(hipEventElapsedTime(&ms, p.a, p.b));
      t_ms[p.cls] += ms;
      t_launches[p.cls]++;
      ev_pool.push_back(p.a);
      ev_pool.push_back(p.b);
    }
    pending.clear();
    return TECS_OK;
  }

  // ---------------------------------------------------------------------- helpers
  int sync_inst() {
    if (!inst_dirty) return TECS_OK;
    HIP_TRY(hipMemcpyAsync(d_inst, h_inst.data(), sizeof(InstanceDesc) * n_inst,
                           hipMemcpyHostToDevice, stream));
    inst_dirty = false;
    return TECS_OK;
  }
  static void load_force(ForceParams& f, const tecs_force_desc& d) {
    int pos[4];
    term_places(d, pos);
    f.h = d.homeostatic_radius;
    f.krep = d.repulsion_scale;
    f.kadh = d.adhesion_scale;
    f.n_types = d.n_types;
    std::memcpy(f.table, d.adhesion_table, sizeof(f.table));
    // a term that is not listed is not applied
    f.bend = pos[TECS_TERM_BENDING] >= 0 ? d.bending_strength : 0.0f;
    f.align = pos[TECS_TERM_ALIGNMENT] >= 0 ? d.alignment_strength : 0.0f;
    const bool linear = pos[TECS_TERM_PIECEWISE_LINEAR] >= 0;
    f.flags = linear ? kForceLinear : 0;
    if (linear && f.bend != 0.0f && pos[TECS_TERM_BENDING] < pos[TECS_TERM_PIECEWISE_LINEAR])
      f.flags |= kForceBendFirst;
    if (f.bend != 0.0f && f.align != 0.0f && pos[TECS_TERM_ALIGNMENT] < pos[TECS_TERM_BENDING])
      f.flags |= kForceAlignFirst;
    if (force_has_drag(d)) f.flags |= kForceDrag;
    f.drag = d.drag;
    f.drag_by_type = d.drag_by_type;
    std::memcpy(f.drag_table, d.drag_table, sizeof(f.drag_table));
  }
  // The force of member k for the following steps. The kernels read it from the device
  // instance table, which is refreshed before the next launch (captured graphs stay valid).
  int set_force(int k, const tecs_force_desc* f) {
    TRY(check_index(k));
    if (!f) return fail(TECS_ERR_INVALID, "null force");
    TRY(validate_force(*f, descs[k]));
    // drag takes the separate-update path, whose Heun corrector evaluates into d3
    if (force_has_drag(*f)) TRY(ensure_d3());
    const bool drag_before = has_drag();
    descs[k].force = *f;
    load_force(h_inst[k].force, *f);
    inst_dirty = true;
    if (has_drag() != drag_before) {
      // the step takes another path (drag needs the separate update): captured graphs are
      // stale; the stream is drained first, a graph may still be running
      drop_graphs();
    }
    return TECS_OK;
  }
  void set_dt(int k, double dt) {
    const float f = (float)dt;  // VectorN * Scalar narrows dt (vector.hpp:263-267)
    if (h_inst[k].dt != f || std::signbit(h_inst[k].dt) != std::signbit(f)) {
      h_inst[k].dt = f;
      inst_dirty = true;
    }
  }

  // keep_ov false: the step that follows writes every active cell's old velocity before any
  // kernel reads one (no member has drag), so the old velocities are not moved: the new
  // generation's active slots hold stale values until that write, inactive slots keep theirs
  // in both generations.
  int rebuild(bool keep_ov = true) {
    if (!is_binned(finder)) return TECS_OK;
    Pending p{};
    TRY(time_begin(&p, kTimeRebuild));
    RebuildArgs r{};
    r.inst = d_inst;
    r.block_inst = d_block_inst;
    r.x_old = x[xc];
    r.pol_old = pol[xc];
    r.trig_old = trig_of(pol[xc]);
    r.id_old = id[ic];
    r.ov_old = keep_ov ? ov[xc] : nullptr;
    r.x_new = x[xc ^ 1];
    r.pol_new = pol[xc ^ 1];
    r.trig_new = trig_of(pol[xc ^ 1]);
    r.id_new = id[ic ^ 1];
    r.ov_new = keep_ov ? ov[xc ^ 1] : nullptr;
    r.kc = key_count();
    r.tmp = tmp;
    r.tkey = tkey;
    r.tid = tidb;
    r.incremental = inc_ok;
    r.coff_old = coff[ic];
    r.coff = coff[ic ^ 1];
    r.skey_new = skey[ic ^ 1];
    r.head_clear = head[ic ^ 1];
    r.rel = rel;
    r.dirty = dirty;
    r.keys_ready = keys_ready;
    r.scan_bsum = bsum;
    r.total_cbins = total_cbins;
    r.fill = sparse_rows;
    r.fill_inst = d_fill_inst;
    r.fill_base = d_fill_base;
    r.fill_blocks = (int)h_fill_inst.size();
    launch_rebuild(r, n_blocks, stream);
    HIP_TRY(hipGetLastError());
    keys_ready = false;
    TRY(time_end(&p));
    xc ^= 1;
    ic ^= 1;
    grid_built = true;
    inc_ok = true;
    stats.rebuilds++;
    return TECS_OK;
  }

  // The reference rebuild schedule for a full evaluation (binned_all_pairs.hpp:77-84).
  int maybe_rebuild(bool is_full_step, bool keep_ov = true) {
    rebuilt_now = false;
    if (!is_binned(finder)) return TECS_OK;
    if (is_full_step) {
      if (step_count == 0 || step_count >= rebuild_every_n) {
        TRY(rebuild(keep_ov));
        step_count = 0;
        rebuilt_now = true;
      }
      step_count++;
    }
    if (!grid_built)
      return fail(TECS_ERR_STATE, "binned pair finder evaluated before any full step built a grid");
    return TECS_OK;
  }

  SweepArgs base_args(const float4* xs, const float2* ps) const {
    SweepArgs a{};
    a.inst = d_inst;
    a.block_inst = d_block_inst;
    a.xs = xs;
    a.ps = ps;
    a.trig = trig_of(ps);
    a.off = coff[ic];
    a.rel = rel;
    a.id = id[ic];
    a.block_pairs = block_pairs;
    a.dl_cnt = dl_cnt;
    a.dl_row = dl_row;
    a.dl_splits = dl_splits;
    a.n_inst = n_inst;
    return a;
  }

  // The concrete sweep: TECS_SWEEP_THREAD or TECS_SWEEP_TEAM4/8/16.
  // Every launch of a force-mode sweep runs the thread variant (per size class when tuned so).
  bool thread_sweeps_only() const {
    const int v = concrete_variant();
    if (!groups.empty() && variant == TECS_SWEEP_AUTO && v == tuned) {
      for (const SweepGroup& g : groups)
        if (g.variant != TECS_SWEEP_THREAD) return false;
      return true;
    }
    return v == TECS_SWEEP_THREAD;
  }
  int concrete_variant() const {
    if (!team_ok) return TECS_SWEEP_THREAD;
    switch (variant) {
      case TECS_SWEEP_THREAD: return TECS_SWEEP_THREAD;
      case TECS_SWEEP_TEAM: return TECS_SWEEP_TEAM8;
      case TECS_SWEEP_TEAM4:
      case TECS_SWEEP_TEAM8:
      case TECS_SWEEP_TEAM16: return variant;
      default: return tuned >= 0 ? tuned : TECS_SWEEP_TEAM8;
    }
  }
  static int lanes_of(int v) {
    return v == TECS_SWEEP_TEAM4 ? 4 : v == TECS_SWEEP_TEAM8 ? 8 : v == TECS_SWEEP_TEAM16 ? 16 : 1;
  }
  int partials_per_block() const { return lanes_of(concrete_variant()); }

  // nb: the chunks the launch covers (a.chunks lists them when nb < n_blocks)
  int launch_variant(int v, int mode, const SweepArgs& a, int nb = -1) {
    if (nb < 0) nb = n_blocks;
    if (v == TECS_SWEEP_THREAD) {
      if (launch_sweep(finder, has_type, has_pol, mode, a, nb, stream) != 0)
        return fail(TECS_ERR_STATE, "sweep mode not available for this pair finder");
    } else if (launch_sweep_team(lanes_of(v), is_gabriel(finder), is_symmetric(finder), has_type,
                                 has_pol, mode, a, nb, stream) != 0) {
      return fail(TECS_ERR_STATE, "sweep mode not available for this pair finder");
    }
    return TECS_OK;
  }

  // Size class of a member: its cell count in powers of four (1k-4k, 4k-16k, ...).
  static int size_class_of(int n) {
    int c = 0;
    for (int v = std::max(n, 1); v >= 4 && c < kMaxClasses - 1; v >>= 2) ++c;
    return c;
  }

  // The next full evaluation rebuilds (maybe_rebuild), so the sweep that ends this step can
  // count the keys of the positions it writes.
  bool emit_keys() const { return is_binned(finder) && step_count >= rebuild_every_n; }

  // What count_key writes for the positions of the current order: counts for a full rebuild,
  // moves and arrivals for an incremental one.
  KeyCount key_count() const {
    KeyCount k{};
    k.keys = keys;
    k.cnt = cnt;
    if (inc_ok) {
      k.skey = skey[ic];
      k.head = head[ic];
      k.nxt = nxt;
      k.movers = movers;
      k.dirty = dirty;
      // the scan's tile sums move with the counts when the scan takes the fused path, which
      // leaves them in bsum (with more tiles the top-level scan overwrites them)
      if (scan_tiles_for(total_cbins) <= kScanFusedTileLimit) k.tsum = bsum;
    }
    return k;
  }

  // Positions are about to change outside a step: the next rebuild starts from scratch, with
  // zero counts (cnt holds the last rebuild's counts, and possibly moves of this order) and
  // empty arrival lists.
  int drop_keys() {
    if (is_binned(finder) && (keys_ready || inc_ok)) {
      HIP_TRY(hipMemsetAsync(cnt, 0, sizeof(int) * (size_t)total_cbins, stream));
      if (keys_ready) HIP_TRY(hipMemsetAsync(head[ic], 0xff, sizeof(int) * (size_t)total_cbins, stream));
      // marks of an incremental count that no rebuild will link
      if (inc_ok) HIP_TRY(hipMemsetAsync(movers, 0, sizeof(unsigned long long) * (size_t)(total_slots / 64), stream));
    }
    keys_ready = false;
    inc_ok = false;
    return TECS_OK;
  }

  // the trig buffer that belongs to a polarity buffer
  float4* trig_of(const float2* ps) const {
    if (!has_pol || !ps) return nullptr;
    if (ps == pol[0]) return trig[0];
    if (ps == pol[1]) return trig[1];
    return trigp;
  }

  // Autotuner: time every exact sweep variant on the live state (the first sweep of a step,
  // into buffers the step rewrites, with scratch counters, so neither the state nor the
  // statistics change) and keep the fastest of 5 timed launches each. Runs once, after the
  // first grid exists.
  int autotune() {
    if (tuned >= 0) return TECS_OK;
    if (!team_ok) {
      const std::vector<std::vector<int>> class_chunks = chunks_by_class();
      for (int c = 0; c < kMaxClasses; ++c) class_variant[c] = class_chunks[c].empty() ? -1 : TECS_SWEEP_THREAD;
      tuned = TECS_SWEEP_THREAD;
      return TECS_OK;
    }
    const int cands[4] = {TECS_SWEEP_TEAM4, TECS_SWEEP_TEAM8, TECS_SWEEP_TEAM16, TECS_SWEEP_THREAD};
    struct Events {  // destroyed on every return path
      hipEvent_t e0 = nullptr, e1 = nullptr;
      ~Events() {
        if (e0) hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
      }
    } ev;
    HIP_TRY(hipEventCreate(&ev.e0));
    HIP_TRY(hipEventCreate(&ev.e1));
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    double best = 1e30;
    int best_v = TECS_SWEEP_TEAM8;
    int rc = TECS_OK;
    // time the sweep the steps start with, writing only buffers that step overwrites before
    // reading them (Heun: d1 and the predicted state; Euler: the other state buffer)
    const bool com_on = com_fixer == TECS_GLOBAL_COM_FIXER;
    int mode = kModeDelta;
    SweepArgs t = base_args(x[xc], pol[xc]);
    t.o_pos = d1;
    t.o_pol = d1p;
    t.block_pairs = tune_pairs;
    // symmetric finders: a Delta sweep as on a fresh grid (no movers, so nothing goes to q)
    if (delta_only(finder)) {
    } else if (!com_on && integrator == TECS_HEUN) {
      mode = kModePred;
      t.o2_pos = xp;
      t.o2_pol = polp;
      t.o_trig = trigp;
    } else if (!com_on) {
      mode = kModeEuler;
      t.o_pos = x[xc ^ 1];
      t.o_pol = pol[xc ^ 1];
      t.o_trig = trig_of(pol[xc ^ 1]);
      t.old_vel = track_ov ? ov[xc ^ 1] : nullptr;
    }
    // the fastest of 5 timed launches of variant v over nb chunks (the first launch warms
    // the caches)
    auto time_variant = [&](int v, const SweepArgs& args, int nb) -> float {
      float ms_min = 1e30f;
      for (int rep = 0; rep < 6 && rc == TECS_OK; ++rep) {
        hipEventRecord(e0, stream);
        rc = launch_variant(v, mode, args, nb);
        if (rc != TECS_OK) break;
        hipEventRecord(e1, stream);
        hipError_t e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) rc = fail(TECS_ERR_DEVICE, std::string("autotune: ") + hipGetErrorString(e));
        if (rep > 0) ms_min = std::min(ms_min, ms);
      }
      return ms_min;
    };
    // size classes of the members, and the chunks of each class
    const std::vector<std::vector<int>> class_chunks = chunks_by_class();
    int n_classes = 0;
    for (int c = 0; c < kMaxClasses; ++c) {
      class_variant[c] = -1;
      for (int v = 0; v < 6; ++v) class_tune_ms[c][v] = 0.0;
      n_classes += class_chunks[c].empty() ? 0 : 1;
    }
    const bool per_class = n_classes > 1 && !com_on && !is_symmetric(finder);
    for (int v = 0; v < 6; ++v) tune_ms[v] = 0.0;
    // one variant for the whole batch: each timed over all chunks in one launch
    for (int v : cands) {
      const float ms = time_variant(v, t, n_blocks);
      tune_ms[v] = ms;
      if (ms < best) {
        best = ms;
        best_v = v;
      }
    }
    for (int c = 0; c < kMaxClasses; ++c) class_variant[c] = class_chunks[c].empty() ? -1 : best_v;
    if (per_class && rc == TECS_OK) {
      // Then class by class (largest first), another variant for the class is kept when the
      // whole sweep - one launch per group of classes, one after another as the steps run
      // them - gets faster by more than 2%. Timing a class alone would favour variants that
      // fill an otherwise idle GPU with its few cells, which they do not in the full sweep.
      int* d_lists = nullptr;
      TRY(alloc(&d_lists, (size_t)n_blocks));
      auto time_assignment = [&](const int* cv) -> float {
        std::vector<int> all;
        std::vector<std::pair<int, int>> grp;  // (variant, chunks)
        for (int v : cands) {
          const size_t before = all.size();
          for (int c = 0; c < kMaxClasses; ++c)
            if (cv[c] == v) all.insert(all.end(), class_chunks[c].begin(), class_chunks[c].end());
          if (all.size() > before) grp.push_back({v, (int)(all.size() - before)});
        }
        if (hipMemcpyAsync(d_lists, all.data(), sizeof(int) * all.size(), hipMemcpyHostToDevice, stream) != hipSuccess) {
          rc = fail(TECS_ERR_DEVICE, "autotune: list upload");
          return 1e30f;
        }
        float ms_min = 1e30f;
        for (int rep = 0; rep < 6 && rc == TECS_OK; ++rep) {
          hipEventRecord(e0, stream);
          int off = 0;
          for (const auto& gv : grp) {
            SweepArgs tg = t;
            tg.chunks = d_lists + off;
            rc = launch_variant(gv.first, mode, tg, gv.second);
            off += gv.second;
            if (rc != TECS_OK) break;
          }
          hipEventRecord(e1, stream);
          hipError_t e = hipEventSynchronize(e1);
          float ms = 0.0f;
          if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
          if (e != hipSuccess) rc = fail(TECS_ERR_DEVICE, std::string("autotune: ") + hipGetErrorString(e));
          if (rep > 0) ms_min = std::min(ms_min, ms);
        }
        return ms_min;
      };
      std::vector<int> by_size;
      for (int c = 0; c < kMaxClasses; ++c)
        if (!class_chunks[c].empty()) by_size.push_back(c);
      std::sort(by_size.begin(), by_size.end(),
                [&](int a, int b) { return class_chunks[a].size() > class_chunks[b].size(); });
      double current = best;
      for (int c : by_size) {
        for (int v : cands) {
          if (rc != TECS_OK) break;
          if (v == class_variant[c]) {
            class_tune_ms[c][v] = current;
            continue;
          }
          int cv[kMaxClasses];
          for (int q = 0; q < kMaxClasses; ++q) cv[q] = class_variant[q];
          cv[c] = v;
          const float ms = time_assignment(cv);
          class_tune_ms[c][v] = ms;
          if (ms < current * 0.98) {
            current = ms;
            class_variant[c] = v;
          }
        }
      }
      free_one(d_lists);
    }
    if (rc != TECS_OK) return rc;
    TRY(build_groups(class_chunks));
    HIP_TRY(hipStreamSynchronize(stream));
    return TECS_OK;
  }

  // Members by size class: their classes (member_class) and each class's chunks.
  std::vector<std::vector<int>> chunks_by_class() {
    member_class.assign(n_inst, 0);
    std::vector<std::vector<int>> class_chunks(kMaxClasses);
    for (int k = 0; k < n_inst; ++k) {
      member_class[k] = size_class_of(h_inst[k].n);
      for (int b = 0; b < h_inst[k].n_blocks; ++b) class_chunks[member_class[k]].push_back(h_inst[k].block_begin + b);
    }
    return class_chunks;
  }

  // The launch groups of class_variant: the classes that use the same variant, swept by one
  // launch over a device list of their chunks (one group: no lists). `tuned` becomes the
  // variant of the largest group.
  int build_groups(const std::vector<std::vector<int>>& class_chunks) {
    const int order[4] = {TECS_SWEEP_TEAM4, TECS_SWEEP_TEAM8, TECS_SWEEP_TEAM16, TECS_SWEEP_THREAD};
    groups.clear();
    std::vector<std::vector<int>> lists;
    int largest = -1, best_v = TECS_SWEEP_THREAD;
    for (int v : order) {
      std::vector<int> ch;
      for (int c = 0; c < kMaxClasses; ++c)
        if (class_variant[c] == v) ch.insert(ch.end(), class_chunks[c].begin(), class_chunks[c].end());
      if (ch.empty()) continue;
      if ((int)ch.size() > largest) {
        largest = (int)ch.size();
        best_v = v;
      }
      groups.push_back(SweepGroup{v, (int)ch.size(), nullptr});
      lists.push_back(std::move(ch));
    }
    if (groups.size() > 1) {
      for (size_t g = 0; g < groups.size(); ++g) {
        TRY(alloc(&groups[g].chunks, lists[g].size()));
        HIP_TRY(hipMemcpyAsync(groups[g].chunks, lists[g].data(), sizeof(int) * lists[g].size(),
                               hipMemcpyHostToDevice, stream));
      }
    } else {
      groups.clear();
    }
    tuned = best_v;
    return TECS_OK;
  }

  // A variant per size class chosen by the caller (AUTO only): class_variant_in[c] for every
  // class that has members (THREAD or, when the team sweep is exact here, TEAM4/8/16).
  int set_class_variants(const int32_t* class_variant_in) {
    const std::vector<std::vector<int>> class_chunks = chunks_by_class();
    for (int c = 0; c < kMaxClasses; ++c) {
      if (class_chunks[c].empty()) continue;
      const int v = class_variant_in[c];
      const bool team = v == TECS_SWEEP_TEAM4 || v == TECS_SWEEP_TEAM8 || v == TECS_SWEEP_TEAM16;
      if (v != TECS_SWEEP_THREAD && !(team && team_ok))
        return fail(TECS_ERR_INVALID, "class variant must be THREAD or, where exact, TEAM4/8/16");
    }
    if (com_fixer == TECS_GLOBAL_COM_FIXER || is_symmetric(finder))
      return fail(TECS_ERR_UNSUPPORTED, "variants per size class need ordered pairs and no COM fixer");
    HIP_TRY(hipStreamSynchronize(stream));
    drop_graphs();
    for (int c = 0; c < kMaxClasses; ++c) {
      class_variant[c] = class_chunks[c].empty() ? -1 : class_variant_in[c];
      for (int v = 0; v < 6; ++v) class_tune_ms[c][v] = 0.0;
    }
    variant = TECS_SWEEP_AUTO;
    TRY(build_groups(class_chunks));
    HIP_TRY(hipStreamSynchronize(stream));
    return TECS_OK;
  }

  // The buffers a sweep of this mode reads and writes besides the state: a missing one is a
  // host-side error, not a launch.
  int check_sweep_args(int mode, const SweepArgs& a) const {
    bool ok = a.inst && a.block_inst && a.xs && a.id && a.block_pairs && (!has_pol || a.ps);
    if (is_binned(finder)) ok = ok && a.off;
    switch (mode) {
      case kModeDeltaDrag: ok = ok && a.ov_in; [[fallthrough]];
      case kModeDelta: ok = ok && a.o_pos && (!has_pol || a.o_pol); break;
      case kModePred:
        ok = ok && a.o_pos && a.o2_pos && (!has_pol || (a.o_pol && a.o2_pol && a.o_trig));
        break;
      case kModeCorr:
        ok = ok && a.o2_pos && a.d1_pos && (!has_pol || (a.o2_pol && a.d1_pol && a.o_trig));
        break;
      case kModeEuler: ok = ok && a.o_pos && (!has_pol || (a.o_pol && a.o_trig)); break;
      case kModeCount: ok = ok && a.row_count; break;
      case kModeEmit: ok = ok && a.row_ptr && a.cols; break;
      default: ok = false;
    }
    if (finder == kNaiveDelaunay) ok = ok && a.dl_cnt;
    if (!ok) return fail(TECS_ERR_STATE, "sweep launched without one of its buffers");
    return TECS_OK;
  }

  int sweep(int mode, const SweepArgs& a_in, int cls) {
    Pending p{};
    TRY(time_begin(&p, cls));
    // the symmetric team sweeps evaluate forces only; the pair export walks one thread per cell
    int v = is_symmetric(finder) && mode != kModeDelta ? TECS_SWEEP_THREAD : concrete_variant();
    SweepArgs a = a_in;
    if (mode == kModeDelta && has_drag()) {
      // the drag terms: one thread per cell sums them in visiting order, reading the old
      // velocities of the current slot order
      mode = kModeDeltaDrag;
      v = TECS_SWEEP_THREAD;
      a.ov_in = ov[xc];
    }
    TRY(check_sweep_args(mode, a));
    // a symmetric Delta sweep with the predictor fused (o2_pos set): only the thread sweep
    // fuses it, so every class takes that variant
    if (mode == kModeDelta && is_symmetric(finder) && a.o2_pos) v = TECS_SWEEP_THREAD;
    const bool force_mode = mode == kModeDelta || mode == kModePred || mode == kModeCorr || mode == kModeEuler;
    if (!groups.empty() && variant == TECS_SWEEP_AUTO && force_mode && v == tuned &&
        !(mode == kModeDelta && is_symmetric(finder) && a.o2_pos)) {
      // per size class: one launch per group of classes, each over its chunks
      for (const SweepGroup& g : groups) {
        SweepArgs ag = a;
        ag.chunks = g.chunks;
        TRY(launch_variant(g.variant, mode, ag, g.n_chunks));
      }
    } else {
      TRY(launch_variant(v, mode, a));
    }
    HIP_TRY(hipGetLastError());
    TRY(time_end(&p));
    if (mode != kModeCount && mode != kModeEmit) stats.evaluations++;
    return TECS_OK;
  }

  // COM fix (when on) and the integrator update of a Delta-mode evaluation whose position
  // deltas are F. The ordered sweeps wrote per-workgroup COM partials themselves; for the
  // symmetric ones a pass computes them from F (+ q, the movers' contributions, when the
  // evaluation saw a stale grid: with_q), and the update adds q and clears it.
  // movers: the predictor flags the cells outside their bin of the last rebuild for the
  // symmetric evaluation that follows on that stale grid (mover8)
  int com_and_update(int kind, const float4* F = nullptr, bool with_q = false, bool movers = false) {
    const bool sym = separate_update();
    if (com_fixer == TECS_GLOBAL_COM_FIXER) {
      if (sym) launch_com_partials(d_inst, d_block_inst, F, with_q ? q : nullptr, com_partial, n_blocks, stream);
      launch_com_finalize(d_inst, n_inst, com_partial, sym ? 1 : partials_per_block(), com, stream);
    }
    UpdateArgs u{};
    u.inst = d_inst;
    u.block_inst = d_block_inst;
    u.com = com;
    u.x = x[xc];
    u.pol = pol[xc];
    u.d1_pos = d1;
    u.d1_pol = d1p;
    u.d3_pos = d3;
    u.d3_pol = d3p;
    u.xp = xp;
    u.polp = polp;
    u.trig = trig_of(pol[xc]);
    u.trigp = trigp;
    u.old_vel = track_ov ? ov[xc] : nullptr;
    u.q_pos = with_q ? q : nullptr;
    u.q_pol = with_q ? qp : nullptr;
    // the update that ends a step counts the keys of the new positions when the next step
    // rebuilds, as the fused sweep epilogues do
    const bool emit = kind != kUpdatePred && emit_keys();
    if (emit) u.kc = key_count();
    if (kind != kUpdatePred) u.nonfinite = nonfinite;
    if (kind == kUpdatePred && movers && mover8) {
      u.mover_key = skey[ic];
      u.mover8 = mover8;
    }
    launch_update(kind, has_pol, u, n_blocks, stream);
    HIP_TRY(hipGetLastError());
    if (kind != kUpdatePred) keys_ready = emit;
    return TECS_OK;
  }

  // A symmetric evaluation on a stale grid (positions other than those the grid was built
  // from) hands the sweep the rebuild keys and the movers' buffer.
  void sym_grid_args(SweepArgs* a, bool stale) const {
    if (!stale) return;
    a->skey = skey[ic];
    a->q_pos = q;
    a->q_pol = qp;
  }

  // Members with walls or links: the forces beyond the pair finder join every evaluation.
  bool has_extra() const {
    if (n_links > 0) return true;
    for (const auto& in : h_inst)
      if (in.n_walls > 0) return true;
    return false;
  }
  // Some member's force has drag: its evaluations read the old velocities of every partner,
  // so none may update them in place (the fused Heun corrector and Euler epilogues do), and the
  // thread sweep, which sums the drag terms in visiting order, evaluates them.
  bool has_drag() const {
    for (const auto& in : h_inst)
      if (in.force.flags & kForceDrag) return true;
    return false;
  }
  // The steps take the separate-update path (below).
  bool separate_update() const { return delta_only(finder) || has_extra() || has_drag(); }

  // The generic wall forces (before the pair forces in the reference's order, so W + F) and the
  // link forces (after them) of an evaluation at stage positions xs into its deltas F. The
  // agent -> slot map of the links is refreshed when `order_changed` (the first evaluation of a
  // step: the slot order changes only at rebuilds).
  int apply_extra(const float4* xs, float4* F, bool order_changed) {
    if (n_links > 0 && order_changed)
      launch_inverse_ids(d_inst, d_block_inst, id[ic], inv, n_blocks, stream);
    launch_extra(d_inst, d_block_inst, id[ic], inv, n_links > 0 ? d_link_rows : nullptr,
                 d_link_ent, xs, F, n_blocks, stream);
    HIP_TRY(hipGetLastError());
    return TECS_OK;
  }

  // A step with a symmetric finder, NaiveDelaunay, walls or links: each evaluation writes its
  // deltas (a symmetric finder's movers' part into q on a stale grid), the wall and link
  // forces join them, then the COM fix and the update run as their own kernels.
  // Heun: evaluate -> predictor -> evaluate (stale grid) -> corrector.
  int step_delta_only() {
    if (integrator == TECS_HEUN && (!d3 || (has_pol && !d3p)))
      return fail(TECS_ERR_STATE, "separate-update step without the corrector's delta buffer");
    const bool sym_binned = is_symmetric(finder) && is_binned(finder);
    const bool stale1 = sym_binned && !rebuilt_now;
    SweepArgs a = base_args(x[xc], pol[xc]);
    a.o_pos = d1;
    a.o_pol = d1p;
    sym_grid_args(&a, stale1);
    // the predictor fused into a fresh symmetric thread sweep: nothing else adds to the delta
    // (no movers on a fresh grid, no COM fix, no walls or links)
    const bool fuse_pred = integrator == TECS_HEUN && sym_binned && !stale1 &&
                           com_fixer != TECS_GLOBAL_COM_FIXER && !has_extra() && !has_drag() &&
                           thread_sweeps_only();
    if (fuse_pred) {
      a.o2_pos = xp;
      a.o2_pol = polp;
      a.o_trig = trigp;
      a.mover8_out = mover8;
    }
    TRY(sweep(kModeDelta, a, kTimeSweepFull));
    if (has_extra()) TRY(apply_extra(x[xc], d1, true));
    if (integrator == TECS_HEUN) {
      if (!fuse_pred) TRY(com_and_update(kUpdatePred, d1, stale1, sym_binned));
      SweepArgs b = base_args(xp, polp);
      b.o_pos = d3;
      b.o_pol = d3p;
      sym_grid_args(&b, sym_binned);
      if (sym_binned) b.mover8 = mover8;
      TRY(sweep(kModeDelta, b, kTimeSweepStale));
      if (has_extra()) TRY(apply_extra(xp, d3, false));
      TRY(com_and_update(kUpdateCorr, d3, sym_binned));
    } else {
      TRY(com_and_update(kUpdateEuler, d1, stale1));
    }
    stats.steps++;
    return TECS_OK;
  }

  // One integrator step (Euler::integrate euler.hpp:81-85, Heun::integrate heun.hpp:201-208).
  int step_once() {
    // every path below ends in a write of each active cell's old velocity (the fused Heun
    // corrector or Euler epilogue, or the update kernel); only drag reads them before that
    TRY(maybe_rebuild(true, has_drag()));
    if (separate_update()) return step_delta_only();
    const bool com_on = com_fixer == TECS_GLOBAL_COM_FIXER;
    if (integrator == TECS_HEUN) {
      if (!com_on) {
        SweepArgs a = base_args(x[xc], pol[xc]);
        a.o_pos = d1;
        a.o_pol = d1p;
        a.o2_pos = xp;
        a.o2_pol = polp;
        a.o_trig = trigp;
        TRY(sweep(kModePred, a, kTimeSweepFull));
        SweepArgs b = base_args(xp, polp);
        b.o2_pos = x[xc];
        b.o2_pol = pol[xc];
        b.o_trig = trig_of(pol[xc]);
        const bool emit = emit_keys();
        if (emit) {
          b.kc = key_count();
        }
        b.d1_pos = d1;
        b.d1_pol = d1p;
        b.old_vel = track_ov ? ov[xc] : nullptr;
        b.nonfinite = nonfinite;
        TRY(sweep(kModeCorr, b, kTimeSweepStale));
        keys_ready = emit;
      } else {
        SweepArgs a = base_args(x[xc], pol[xc]);
        a.o_pos = d1;
        a.o_pol = d1p;
        a.com_partial = com_partial;
        TRY(sweep(kModeDelta, a, kTimeSweepFull));
        TRY(com_and_update(kUpdatePred));
        SweepArgs b = base_args(xp, polp);
        b.o_pos = d3;
        b.o_pol = d3p;
        b.com_partial = com_partial;
        TRY(sweep(kModeDelta, b, kTimeSweepStale));
        TRY(com_and_update(kUpdateCorr));
      }
    } else {
      if (!com_on) {
        SweepArgs a = base_args(x[xc], pol[xc]);
        a.o_pos = x[xc ^ 1];
        a.o_pol = pol[xc ^ 1];
        a.o_trig = trig_of(pol[xc ^ 1]);
        const bool emit = emit_keys();
        if (emit) {
          a.kc = key_count();
        }
        a.old_vel = track_ov ? ov[xc ^ 1] : nullptr;
        a.nonfinite = nonfinite;
        TRY(sweep(kModeEuler, a, kTimeSweepFull));
        xc ^= 1;
        keys_ready = emit;
      } else {
        SweepArgs a = base_args(x[xc], pol[xc]);
        a.o_pos = d1;
        a.o_pol = d1p;
        a.com_partial = com_partial;
        TRY(sweep(kModeDelta, a, kTimeSweepFull));
        TRY(com_and_update(kUpdateEuler));
      }
    }
    stats.steps++;
    return TECS_OK;
  }

  // Capture graph_period() steps from the current state (the host state advances as if they
  // ran; the caller launches the graph once to make the device catch up).
  int capture_graph(StepGraph* out) {
    const HostState start = host_state();
    const tecs_stats before = stats;
    HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    int rc = TECS_OK;
    const int P = graph_period();
    for (int k = 0; k < P && rc == TECS_OK; ++k) rc = step_once();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(stream, &g);
    if (rc != TECS_OK) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(TECS_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ei != hipSuccess) return fail(TECS_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    out->start = start;
    out->exec = exec;
    out->delta.pair_interactions = 0;
    out->delta.evaluations = stats.evaluations - before.evaluations;
    out->delta.rebuilds = stats.rebuilds - before.rebuilds;
    out->delta.steps = stats.steps - before.steps;
    if (!(host_state() == start)) {  // not periodic from here: run it once, do not keep it
      // (destroyed by drop_graphs once the stream is idle, not while the launch is in flight)
      retired.push_back(exec);
      HIP_TRY(hipGraphLaunch(exec, stream));
      out->exec = nullptr;
    }
    return TECS_OK;
  }

  int take_step(const double* dts, int n_steps) {
    if (n_steps < 0) return fail(TECS_ERR_INVALID, "n_steps must be >= 0");
    for (int k = 0; k < n_inst; ++k) set_dt(k, dts[k]);
    TRY(sync_inst());
    int done = 0;
    if (variant == TECS_SWEEP_AUTO && tuned < 0 && n_steps > 0) {
      if (!grid_built && is_binned(finder)) {
        TRY(step_once());
        ++done;
      }
      TRY(autotune());
    }
    const int P = graph_period();
    // AUTO captures as soon as a call covers one period: a captured sequence is kept per
    // starting state and replayed by every later call, so the capture cost is paid once
    const bool graphs_on = graph_mode == TECS_GRAPHS_ON ||
                           (graph_mode == TECS_GRAPHS_AUTO && n_steps >= P);
    if (graphs_on && !timing) {
      // the first step after creation or a manual rebuild may start a new rebuild phase
      while (done < n_steps && (step_count == 0 && is_binned(finder))) {
        TRY(step_once());
        ++done;
      }
      while (n_steps - done >= P) {
        const HostState st = host_state();
        StepGraph* g = nullptr;
        for (auto& c : graphs)
          if (c.start == st) g = &c;
        if (!g) {
          StepGraph fresh{};
          TRY(capture_graph(&fresh));
          done += P;
          if (!fresh.exec) continue;  // ran once eagerly through the graph
          graphs.push_back(fresh);
          HIP_TRY(hipGraphLaunch(graphs.back().exec, stream));
          ++graph_launches;
          continue;
        }
        HIP_TRY(hipGraphLaunch(g->exec, stream));
        ++graph_launches;
        stats.evaluations += g->delta.evaluations;
        stats.rebuilds += g->delta.rebuilds;
        stats.steps += g->delta.steps;
        done += P;
      }
    }
    for (; done < n_steps; ++done) TRY(step_once());
    return TECS_OK;
  }

  int check_index(int k) const {
    if (k < 0 || k >= n_inst) return fail(TECS_ERR_INVALID, "simulation index out of range");
    return TECS_OK;
  }

  int upload(int k, const float* pos, const float* p, const uint8_t* type) {
    TRY(check_index(k));
    return upload_range(k, 0, descs[k].agent_count, pos, p, type);
  }

  // Cells first .. first + count - 1 of member k (active or not) from host arrays.
  int upload_range(int k, uint32_t first, uint32_t count, const float* pos, const float* p,
                   const uint8_t* type) {
    TRY(check_index(k));
    if (!pos && count > 0) return fail(TECS_ERR_INVALID, "positions are required");
    const tecs_sim_desc& d = descs[k];
    if ((uint64_t)first + count > d.capacity)
      return fail(TECS_ERR_INVALID, "cells beyond the capacity");
    const size_t n = count;
    if (n > (size_t)stage_n) return fail(TECS_ERR_INVALID, "upload larger than the staging buffers");
    if (has_type && type && d.force.n_types > 0)
      for (size_t i = 0; i < n; ++i)
        if (type[i] >= (unsigned)d.force.n_types)
          return fail(TECS_ERR_INVALID, "cell type outside the adhesion table");
    TRY(drop_keys());
    if (n == 0) return TECS_OK;
    HIP_TRY(hipMemcpyAsync(st_pos, pos, sizeof(float) * 3 * n, hipMemcpyHostToDevice, stream));
    const bool up_pol = has_pol && p;
    const bool up_type = has_type && type;
    if (up_pol)
      HIP_TRY(hipMemcpyAsync(st_pol, p, sizeof(float) * 2 * n, hipMemcpyHostToDevice, stream));
    if (up_type) HIP_TRY(hipMemcpyAsync(st_type, type, n, hipMemcpyHostToDevice, stream));
    const InstanceDesc& in = h_inst[k];
    UploadArgs u{};
    u.inst = d_inst;
    u.ii = k;
    u.block_begin = in.block_begin;
    u.id = id[ic];
    u.first = (int)first;
    u.count = (int)count;
    u.pos = st_pos;
    u.pol = up_pol ? st_pol : nullptr;
    u.type = up_type ? st_type : nullptr;
    u.x = x[xc];
    u.p = pol[xc];
    u.trig = trig_of(pol[xc]);
    u.x2 = x[xc ^ 1];
    u.p2 = pol[xc ^ 1];
    u.trig2 = trig_of(pol[xc ^ 1]);
    launch_upload(u, in.n_blocks, stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(stream));
    return TECS_OK;
  }

  // Simulation::set_agent_count (simulation.hpp:285-301) for member k. Every member's cells move
  // into the other buffer generation, member k's into index order (so its first n slots hold
  // cells 0..n-1, active or newly active, and the rest keep their state in index order), and
  // the next full step rebuilds every grid (pair_finder.set_agent_count resets the counter,
  // binned_all_pairs.hpp:60-63).
  int set_agent_count(int k, uint32_t n) {
    TRY(check_index(k));
    if (n > descs[k].capacity) return fail(TECS_ERR_INVALID, "agent_count above the capacity");
    if (n_inst > 1 && is_binned(finder) && rebuild_every_n > 0)
      return fail(TECS_ERR_UNSUPPORTED,
                  "set_agent_count in a batch that shares a rebuild period > 0 would move every "
                  "member's rebuild");
    TRY(drop_keys());
    ReorderArgs r{};
    r.inst = d_inst;
    r.block_inst = d_block_inst;
    r.target = k;
    r.id_old = id[ic];
    r.id_new = id[ic ^ 1];
    r.x_old = x[xc];
    r.x_new = x[xc ^ 1];
    if (has_pol) {
      r.pol_old = pol[xc];
      r.pol_new = pol[xc ^ 1];
      r.trig_old = trig[xc];
      r.trig_new = trig[xc ^ 1];
    }
    if (track_ov) {
      r.ov_old = ov[xc];
      r.ov_new = ov[xc ^ 1];
    }
    launch_reorder_identity(r, n_blocks, stream);
    HIP_TRY(hipGetLastError());
    xc ^= 1;
    ic ^= 1;
    // both generations hold member k's inactive cells (in index order)
    const InstanceDesc& in = h_inst[k];
    const size_t b0 = (size_t)in.slot_begin + n, len = (size_t)in.padded - n;
    if (len > 0) {
      HIP_TRY(hipMemcpyAsync(x[xc ^ 1] + b0, x[xc] + b0, sizeof(float4) * len, hipMemcpyDeviceToDevice, stream));
      HIP_TRY(hipMemcpyAsync(id[ic ^ 1] + b0, id[ic] + b0, sizeof(int) * len, hipMemcpyDeviceToDevice, stream));
      if (has_pol) {
        HIP_TRY(hipMemcpyAsync(pol[xc ^ 1] + b0, pol[xc] + b0, sizeof(float2) * len, hipMemcpyDeviceToDevice, stream));
        HIP_TRY(hipMemcpyAsync(trig[xc ^ 1] + b0, trig[xc] + b0, sizeof(float4) * len, hipMemcpyDeviceToDevice, stream));
      }
      if (track_ov)
        HIP_TRY(hipMemcpyAsync(ov[xc ^ 1] + b0, ov[xc] + b0, sizeof(float4) * len, hipMemcpyDeviceToDevice, stream));
    }
    h_inst[k].n = (int)n;
    descs[k].agent_count = n;
    inst_dirty = true;
    TRY(sync_inst());
    step_count = 0;
    grid_built = false;
    rebuilt_now = false;
    HIP_TRY(hipStreamSynchronize(stream));
    return TECS_OK;
  }

  // Grid queries (grid.hpp:413-564) for member 0 on the grid of the last rebuild; points and
  // uniforms are host arrays of n queries, out receives n agent indices.
  int query(int kind, int n, const float* points, int radius_bins, float radius,
            const float* uniforms, int32_t* out) {
    if (n < 0 || (n > 0 && (!points || !out))) return fail(TECS_ERR_INVALID, "bad query arrays");
    if (kind != kQueryNearest && n > 0 && !uniforms) return fail(TECS_ERR_INVALID, "uniforms required");
    if (!is_binned(finder)) return fail(TECS_ERR_UNSUPPORTED, "grid queries need a binned pair finder");
    if (!grid_built) return fail(TECS_ERR_STATE, "no grid yet: queries follow a full step or rebuild");
    if (radius_bins < 0) return fail(TECS_ERR_INVALID, "negative search radius");
    if (n == 0) return TECS_OK;
    void* buf = nullptr;
    const size_t bytes = sizeof(float) * 4 * (size_t)n + sizeof(int32_t) * (size_t)n;
    HIP_TRY(hipMalloc(&buf, bytes));
    float* d_points = static_cast<float*>(buf);
    float* d_u = d_points + 3 * (size_t)n;
    int32_t* d_out = reinterpret_cast<int32_t*>(d_u + n);
    int rc = TECS_OK;
    hipError_t e = hipMemcpyAsync(d_points, points, sizeof(float) * 3 * n, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess && uniforms)
      e = hipMemcpyAsync(d_u, uniforms, sizeof(float) * n, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) {
      QueryArgs q{};
      q.inst = d_inst;
      q.xs = x[xc];
      q.off = coff[ic];
      q.rel = rel;
      q.id = id[ic];
      q.n = n;
      q.points = d_points;
      q.uniforms = d_u;
      q.radius_bins = radius_bins;
      q.radius = radius;
      q.out = d_out;
      launch_query(kind, q, stream);
      e = hipGetLastError();
    }
    if (e == hipSuccess)
      e = hipMemcpyAsync(out, d_out, sizeof(int32_t) * n, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) rc = fail(TECS_ERR_DEVICE, std::string("grid query: ") + hipGetErrorString(e));
    hipFree(buf);
    return rc;
  }

  // get_nearest_point_index for member 0 like query(kQueryNearest, ...), but with the points and
  // results staged in the fixed device array of kernels_query.hip instead of a buffer allocated
  // per call: no device allocation, at most kNoAllocMaxQueries points, the simulation's device
  // current.
  int nearest_noalloc(int n, const float* points, int radius_bins, int32_t* out) {
    if (n < 0 || (n > 0 && (!points || !out))) return fail(TECS_ERR_INVALID, "bad query arrays");
    if (n > kNoAllocMaxQueries)
      return fail(TECS_ERR_CAPACITY, "at most " + std::to_string(kNoAllocMaxQueries) +
                                         " points per allocation-free query");
    if (!is_binned(finder)) return fail(TECS_ERR_UNSUPPORTED, "grid queries need a binned pair finder");
    if (!grid_built) return fail(TECS_ERR_STATE, "no grid yet: queries follow a full step or rebuild");
    if (radius_bins < 0) return fail(TECS_ERR_INVALID, "negative search radius");
    if (n == 0) return TECS_OK;
    int cur = -1;
    HIP_TRY(hipGetDevice(&cur));
    if (cur != device) return fail(TECS_ERR_STATE, "the simulation's device is not the current device");
    void* base = nullptr;
    const int se = noalloc_scratch(&base);
    if (se != 0)
      return fail(TECS_ERR_DEVICE, std::string("scratch address: ") + hipGetErrorString((hipError_t)se));
    float* d_points = static_cast<float*>(base);
    int32_t* d_out = reinterpret_cast<int32_t*>(static_cast<unsigned char*>(base) + kNoAllocOutOffset);
    HIP_TRY(hipMemcpyAsync(d_points, points, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice, stream));
    QueryArgs q{};
    q.inst = d_inst;
    q.xs = x[xc];
    q.off = coff[ic];
      q.rel = rel;
    q.id = id[ic];
    q.n = n;
    q.points = d_points;
    q.uniforms = nullptr;
    q.radius_bins = radius_bins;
    q.out = d_out;
    launch_query(kQueryNearest, q, stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, d_out, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return TECS_OK;
  }

  // Simulation::set_capacity (simulation.hpp:266-274) for member k, within its allocation.
  int set_capacity(int k, uint32_t cap) {
    TRY(check_index(k));
    if (cap < descs[k].agent_count) return fail(TECS_ERR_INVALID, "capacity below agent_count");
    InstanceDesc& in = h_inst[k];
    if (cap > (uint32_t)in.padded)
      return fail(TECS_ERR_UNSUPPORTED, "capacity beyond the member's allocation in this batch");
    if ((int)cap < in.cap) {
      ClearArgs c{};
      c.inst = d_inst;
      c.ii = k;
      c.from = (int)cap;
      c.to = in.cap;
      for (int g = 0; g < 2; ++g) {
        c.x[g] = x[g];
        c.p[g] = has_pol ? pol[g] : nullptr;
        c.trig[g] = has_pol ? trig[g] : nullptr;
        c.ov[g] = track_ov ? ov[g] : nullptr;
      }
      launch_clear_cells(c, stream);
      HIP_TRY(hipGetLastError());
    }
    in.cap = (int)cap;
    descs[k].capacity = cap;
    inst_dirty = true;
    TRY(sync_inst());
    HIP_TRY(hipStreamSynchronize(stream));
    return TECS_OK;
  }

  // Release one block of `allocations` (the stream must be idle).
  void free_one(void* v) {
    for (size_t q = 0; q < allocations.size(); ++q)
      if (allocations[q].first == v) {
        if (pool && allocations[q].second)
          pool->put(v, allocations[q].second);
        else
          hipFree(v);
        allocations.erase(allocations.begin() + (long)q);
        return;
      }
  }

  // The Heun corrector's delta buffer, which the separate-update path needs (allocated when a
  // member first gets walls or links).
  int ensure_d3() {
    if (integrator != TECS_HEUN || d3) return TECS_OK;
    TRY(alloc(&d3, (size_t)total_slots));
    if (has_pol) TRY(alloc(&d3p, (size_t)total_slots));
    return TECS_OK;
  }

  // Generic wall forces of member k (PlaneN normalisation, types/plane.hpp:24-31, on the host
  // in FP32). Captured step graphs are dropped: the step path depends on whether any member
  // has walls or links.
  int set_walls(int k, const tecs_wall* walls, int count) {
    TRY(check_index(k));
    if (count < 0 || count > kMaxWalls) return fail(TECS_ERR_INVALID, "bad wall count");
    if (count > 0 && !walls) return fail(TECS_ERR_INVALID, "null walls");
    InstanceDesc& in = h_inst[k];
    float w[kMaxWalls][5];
    for (int q = 0; q < count; ++q) {
      const float* c = walls[q].plane;
      float sum = 0.0f;
      for (int j = 0; j < 3; ++j) sum += c[j] * c[j];
      if (!(sum > 0.0f) || !std::isfinite(sum)) return fail(TECS_ERR_INVALID, "wall normal is zero");
      const float inv_norm = 1.0f / std::sqrt(sum);
      for (int j = 0; j < 4; ++j) w[q][j] = c[j] * inv_norm;
      w[q][4] = walls[q].strength;
    }
    TRY(ensure_d3());
    HIP_TRY(hipStreamSynchronize(stream));
    drop_graphs();
    in.n_walls = count;
    std::memcpy(in.wall, w, sizeof(float) * 5 * (size_t)count);
    inst_dirty = true;
    return TECS_OK;
  }

  // Links of member k (agent indices below its capacity) and their strength.
  int set_links(int k, const int32_t* links, int64_t count, float strength) {
    TRY(check_index(k));
    if (count < 0 || count > (int64_t)1 << 30) return fail(TECS_ERR_INVALID, "bad link count");
    if (count > 0 && !links) return fail(TECS_ERR_INVALID, "null links");
    for (int64_t q = 0; q < 2 * count; ++q)
      if (links[q] < 0 || (uint32_t)links[q] >= descs[k].capacity)
        return fail(TECS_ERR_INVALID, "link endpoint beyond the capacity");
    TRY(ensure_d3());
    HIP_TRY(hipStreamSynchronize(stream));
    drop_graphs();
    member_links.resize(n_inst);
    member_link_k.resize(n_inst, 0.0f);
    member_links[k].assign(links, links + 2 * count);
    member_link_k[k] = strength;
    h_inst[k].link_k = strength;
    inst_dirty = true;
    // every cell's incident links in link order (a counting sort of the endpoints)
    std::vector<int> rows((size_t)total_slots + 1, 0);
    for (int m = 0; m < n_inst; ++m) {
      const int base = h_inst[m].slot_begin;
      const auto& L = member_links[m];
      for (size_t q = 0; q + 1 < L.size(); q += 2) {
        if (L[q] == L[q + 1]) continue;  // base.hpp:207-208
        rows[(size_t)base + L[q] + 1]++;
        rows[(size_t)base + L[q + 1] + 1]++;
      }
    }
    for (size_t a = 0; a < (size_t)total_slots; ++a) rows[a + 1] += rows[a];
    std::vector<int2> ent((size_t)rows[(size_t)total_slots]);
    {
      std::vector<int> fill(rows.begin(), rows.end() - 1);
      for (int m = 0; m < n_inst; ++m) {
        const int base = h_inst[m].slot_begin;
        const auto& L = member_links[m];
        for (size_t q = 0; q + 1 < L.size(); q += 2) {
          const int a = L[q], b = L[q + 1];
          if (a == b) continue;
          ent[(size_t)fill[(size_t)base + a]++] = make_int2(b, 1);
          ent[(size_t)fill[(size_t)base + b]++] = make_int2(a, -1);
        }
      }
    }
    if ((int64_t)ent.size() > link_ent_cap) {
      if (d_link_ent) free_one(d_link_ent);
      d_link_ent = nullptr;
      link_ent_cap = 0;
      TRY(alloc(&d_link_ent, ent.size()));
      link_ent_cap = (int64_t)ent.size();
    }
    if (!d_link_rows) TRY(alloc(&d_link_rows, (size_t)total_slots + 1));
    if (!inv) TRY(alloc(&inv, (size_t)total_slots));
    TRY(sync_inst());
    HIP_TRY(hipMemcpyAsync(d_link_rows, rows.data(), sizeof(int) * rows.size(), hipMemcpyHostToDevice, stream));
    if (!ent.empty())
      HIP_TRY(hipMemcpyAsync(d_link_ent, ent.data(), sizeof(int2) * ent.size(), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    n_links = (int)ent.size();
    return TECS_OK;
  }

  // Cell division of member k (the UPDATE_FUNC `proliferate` of examples/growth_w_wall.cpp:
  // 52-64): the daughters' directions are formed on the host from the caller's draws, the
  // daughters written on the device next to their parents' current positions.
  int divide(int k, float rate, float offset, const double* u, uint32_t* new_count) {
    TRY(check_index(k));
    const uint32_t n = descs[k].agent_count;
    if (n > 0 && !u) return fail(TECS_ERR_INVALID, "null uniforms");
    std::vector<int> parents;
    std::vector<float> dirs;
    const double pi_f = (double)3.14159265358979323846f;  // Kokkos::numbers::pi_v<float>
    for (uint32_t i = 0; i < n; ++i) {
      const double* v = &u[3 * (size_t)i];
      if (v[0] > (double)rate) continue;  // the reference returns when drand > rate
      // Polarity(acos(2 drand - 1), drand * 2 * pi) narrows both angles to float
      const float theta = (float)std::acos(2.0 * v[1] - 1);
      const float phi = (float)(v[2] * 2.0 * pi_f);
      const float st = (float)std::sin((double)theta), ct = (float)std::cos((double)theta);
      const float sp = (float)std::sin((double)phi), cp = (float)std::cos((double)phi);
      parents.push_back((int)i);
      dirs.push_back(st * cp);  // to_vector3, polarity.hpp:111-118
      dirs.push_back(st * sp);
      dirs.push_back(ct);
    }
    const uint64_t total = (uint64_t)n + parents.size();
    if (total > descs[k].capacity) return fail(TECS_ERR_CAPACITY, "new cells beyond the capacity");
    if (new_count) *new_count = (uint32_t)total;
    if (parents.empty()) return TECS_OK;
    if (!inv) TRY(alloc(&inv, (size_t)total_slots));
    if (!st_idx) TRY(alloc(&st_idx, (size_t)std::max(stage_n, 1)));
    TRY(sync_inst());
    HIP_TRY(hipMemcpyAsync(st_idx, parents.data(), sizeof(int) * parents.size(), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(st_pos, dirs.data(), sizeof(float) * dirs.size(), hipMemcpyHostToDevice, stream));
    launch_inverse_ids(d_inst, d_block_inst, id[ic], inv, n_blocks, stream);
    launch_divide(d_inst, k, inv, st_idx, st_pos, (int)parents.size(), (int)n, offset, x[xc],
                  x[xc ^ 1], stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(stream));
    return TECS_OK;
  }

  // The state of every slot of `old`'s single member, carried into this batch (created for the
  // same member with a larger capacity): both buffer generations get the cells of the current
  // one, so its inactive cells stay in index order in both. The grid is rebuilt at the next full
  // step.
  int adopt_member(Batch& old) {
    if (n_inst != 1 || old.n_inst != 1) return fail(TECS_ERR_INVALID, "adopt_member: single members only");
    HIP_TRY(hipStreamSynchronize(old.stream));
    const size_t len = (size_t)old.h_inst[0].padded;
    for (int g = 0; g < 2; ++g) {
      HIP_TRY(hipMemcpyAsync(x[g], old.x[old.xc], sizeof(float4) * len, hipMemcpyDeviceToDevice, stream));
      HIP_TRY(hipMemcpyAsync(id[g], old.id[old.ic], sizeof(int) * len, hipMemcpyDeviceToDevice, stream));
      if (has_pol) {
        HIP_TRY(hipMemcpyAsync(pol[g], old.pol[old.xc], sizeof(float2) * len, hipMemcpyDeviceToDevice, stream));
        HIP_TRY(hipMemcpyAsync(trig[g], old.trig[old.xc], sizeof(float4) * len, hipMemcpyDeviceToDevice, stream));
      }
      if (track_ov && old.track_ov)
        HIP_TRY(hipMemcpyAsync(ov[g], old.ov[old.xc], sizeof(float4) * len, hipMemcpyDeviceToDevice, stream));
    }
    // the forces beyond the pair finder
    if (old.h_inst[0].n_walls > 0) {
      TRY(ensure_d3());
      h_inst[0].n_walls = old.h_inst[0].n_walls;
      std::memcpy(h_inst[0].wall, old.h_inst[0].wall, sizeof(h_inst[0].wall));
      inst_dirty = true;
      TRY(sync_inst());
    }
    if (!old.member_links.empty() && !old.member_links[0].empty())
      TRY(set_links(0, old.member_links[0].data(), (int64_t)(old.member_links[0].size() / 2),
                    old.member_link_k[0]));
    HIP_TRY(hipMemcpyAsync(nonfinite, old.nonfinite, sizeof(int), hipMemcpyDeviceToDevice, stream));
    variant = old.variant;
    graph_mode = old.graph_mode;
    timing = old.timing;
    stats = old.stats;
    write_step_count = old.write_step_count;
    // the per-block pair counters carry the statistics: move their total into block 0
    {
      tecs_stats st{};
      TRY(old.get_stats(&st));
      stats.pair_interactions = 0;
      const unsigned long long total = st.pair_interactions;
      HIP_TRY(hipMemcpyAsync(block_pairs, &total, sizeof(total), hipMemcpyHostToDevice, stream));
    }
    HIP_TRY(hipStreamSynchronize(stream));
    return TECS_OK;
  }

  int download(int k, float* pos, float* p, float* old_vel) {
    TRY(check_index(k));
    const size_t n = descs[k].agent_count;
    if (n > (size_t)stage_n) return fail(TECS_ERR_STATE, "agent count beyond the staging buffers");
    const InstanceDesc& in = h_inst[k];
    const bool want_pol = p && has_pol;
    const bool want_ov = old_vel != nullptr;
    if (want_ov && !track_ov) {
      std::memset(old_vel, 0, sizeof(float) * 3 * n);
    }
    launch_download(d_inst, k, in.block_begin, in.n_blocks, id[ic], x[xc], pol[xc], ov[xc],
                    pos ? st_pos : nullptr, want_pol ? st_pol : nullptr,
                    (want_ov && track_ov) ? st_ov : nullptr, stream);
    HIP_TRY(hipGetLastError());
    if (pos) HIP_TRY(hipMemcpyAsync(pos, st_pos, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, stream));
    if (want_pol)
      HIP_TRY(hipMemcpyAsync(p, st_pol, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, stream));
    if (want_ov && track_ov)
      HIP_TRY(hipMemcpyAsync(old_vel, st_ov, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return TECS_OK;
  }

  // One evaluation into the d1 scratch buffers, then out in agent order.
  int evaluate(bool is_full_step, float* dpos, float* dpol) {
    TRY(sync_inst());
    TRY(maybe_rebuild(is_full_step));
    const bool sym_stale = is_symmetric(finder) && is_binned(finder) && !rebuilt_now;
    SweepArgs a = base_args(x[xc], pol[xc]);
    a.o_pos = d1;
    a.o_pol = d1p;
    sym_grid_args(&a, sym_stale);
    TRY(sweep(kModeDelta, a, is_full_step ? kTimeSweepFull : kTimeSweepStale));
    if (sym_stale) launch_fold_q(d1, d1p, q, qp, total_slots, stream);
    const InstanceDesc& in = h_inst[0];
    launch_download(d_inst, 0, in.block_begin, in.n_blocks, id[ic], d1, d1p, nullptr,
                    dpos ? st_pos : nullptr, (dpol && has_pol) ? st_pol : nullptr, nullptr, stream);
    HIP_TRY(hipGetLastError());
    const size_t n = descs[0].agent_count;
    if (dpos) HIP_TRY(hipMemcpyAsync(dpos, st_pos, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, stream));
    if (dpol && has_pol)
      HIP_TRY(hipMemcpyAsync(dpol, st_pol, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return TECS_OK;
  }

  int pairs(bool is_full_step, int64_t* row_ptr, int32_t* cols, int64_t cols_capacity,
            int64_t* n_pairs) {
    const size_t n = descs[0].agent_count;
    if (n > (size_t)stage_n) return fail(TECS_ERR_STATE, "agent count beyond the staging buffers");
    if (!cols) {
      // size query: build the grid a due full evaluation would see, advance nothing
      if (is_binned(finder) && is_full_step && (step_count == 0 || step_count >= rebuild_every_n)) {
        TRY(rebuild());
        stats.rebuilds--;
      }
      if (is_binned(finder) && !grid_built)
        return fail(TECS_ERR_STATE, "binned pair finder evaluated before any full step built a grid");
    } else {
      TRY(maybe_rebuild(is_full_step));
    }
    SweepArgs a = base_args(x[xc], pol[xc]);
    a.row_count = d_row_count;
    TRY(sweep(kModeCount, a, kTimeSweepFull));
    std::vector<int> counts(n);
    if (n) HIP_TRY(hipMemcpyAsync(counts.data(), d_row_count, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    std::vector<int64_t> rp(n + 1);
    rp[0] = 0;
    for (size_t i = 0; i < n; ++i) rp[i + 1] = rp[i] + counts[i];
    const int64_t total = rp[n];
    if (n_pairs) *n_pairs = total;
    if (row_ptr) std::memcpy(row_ptr, rp.data(), sizeof(int64_t) * (n + 1));
    if (!cols) return TECS_OK;
    if (cols_capacity < total) return fail(TECS_ERR_CAPACITY, "cols buffer smaller than n_pairs");
    if (total == 0) return TECS_OK;  // nothing to emit (the counting sweep left every counter clear)
    if (total > cols_cap) {
      void* v = nullptr;
      HIP_TRY(hipMalloc(&v, sizeof(int32_t) * (size_t)std::max<int64_t>(total, 1)));
      bool replaced = false;
      for (auto& q : allocations)
        if (q.first == d_cols && d_cols) {
          if (pool && q.second)
            pool->put(q.first, q.second);
          else
            hipFree(q.first);
          q = {v, 0};
          replaced = true;
        }
      if (!replaced) allocations.emplace_back(v, 0);
      d_cols = static_cast<int32_t*>(v);
      cols_cap = total;
    }
    HIP_TRY(hipMemcpyAsync(d_row_ptr, rp.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, stream));
    SweepArgs b = base_args(x[xc], pol[xc]);
    b.row_ptr = d_row_ptr;
    b.cols = d_cols;
    TRY(sweep(kModeEmit, b, kTimeSweepFull));
    if (total) HIP_TRY(hipMemcpyAsync(cols, d_cols, sizeof(int32_t) * total, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return TECS_OK;
  }

  // Simulation::write(time) (simulation.hpp:607-611) for member k, off the critical path.
  int write_snapshot(int k, Writer* w, double time) {
    TRY(check_index(k));
    const InstanceDesc& in = h_inst[k];
    const size_t n = descs[k].agent_count;
    const size_t cap = (size_t)std::max(stage_n, 1);
    if (n > cap) return fail(TECS_ERR_STATE, "agent count beyond the snapshot buffers");
    if (!side) HIP_TRY(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    SnapshotSlot* sl = nullptr;
    {
      std::unique_lock<std::mutex> g(snap_mu);
      snap_cv.wait(g, [&] {
        for (auto& c : snap)
          if (!c.busy) return true;
        return false;
      });
      for (auto& c : snap)
        if (!c.busy) {
          sl = &c;
          break;
        }
      sl->busy = true;
    }
    if (!sl->h_pos) {
      sl->mu = &snap_mu;
      sl->cv = &snap_cv;
      int rc = alloc(&sl->d_pos, 3 * cap);
      if (rc == TECS_OK && has_pol) rc = alloc(&sl->d_pol, 2 * cap);
      hipError_t e = hipSuccess;
      if (rc == TECS_OK) e = hipHostMalloc((void**)&sl->h_pos, sizeof(float) * 3 * cap, hipHostMallocDefault);
      if (rc == TECS_OK && e == hipSuccess && has_pol)
        e = hipHostMalloc((void**)&sl->h_pol, sizeof(float) * 2 * cap, hipHostMallocDefault);
      if (rc == TECS_OK && e == hipSuccess) e = hipEventCreateWithFlags(&sl->gathered, hipEventDisableTiming);
      if (rc == TECS_OK && e == hipSuccess) e = hipEventCreateWithFlags(&sl->copied, hipEventDisableTiming);
      if (rc != TECS_OK || e != hipSuccess) {
        sl->release();
        return rc != TECS_OK ? rc : fail(TECS_ERR_OOM, std::string("snapshot buffers: ") + hipGetErrorString(e));
      }
    }
    sl->n = n;
    sl->has_pol = has_pol;
    sl->time = time;
    sl->step = write_step_count[k]++;
    sl->links.clear();
    if (k < (int)member_links.size())
      sl->links.assign(member_links[k].begin(), member_links[k].end());
    launch_download(d_inst, k, in.block_begin, in.n_blocks, id[ic], x[xc], pol[xc], nullptr,
                    sl->d_pos, has_pol ? sl->d_pol : nullptr, nullptr, stream);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(sl->gathered, stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(side, sl->gathered, 0);
    if (e == hipSuccess && n)
      e = hipMemcpyAsync(sl->h_pos, sl->d_pos, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, side);
    if (e == hipSuccess && n && has_pol)
      e = hipMemcpyAsync(sl->h_pol, sl->d_pol, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, side);
    if (e == hipSuccess) e = hipEventRecord(sl->copied, side);
    if (e != hipSuccess) {
      sl->release();
      return fail(TECS_ERR_DEVICE, std::string("snapshot hand-off: ") + hipGetErrorString(e));
    }
    w->enqueue(sl);
    return TECS_OK;
  }

  // Move member ks of src (its current cell state: positions, types, polarities, ids, old
  // velocities) into member kd of this batch, which must not have stepped yet. The grid is not
  // carried: this batch rebuilds it at its first full step, which the member's own schedule does
  // at that step when rebuild_every_n is 0, or when its steps so far are a multiple of a period
  // > 0 (the server re-forms such groups only then).
  int copy_member_from(Batch& src, int ks, int kd) {
    TRY(check_index(kd));
    TRY(src.check_index(ks));
    const InstanceDesc& a = src.h_inst[ks];
    const InstanceDesc& b = h_inst[kd];
    if (a.n != b.n) return fail(TECS_ERR_INVALID, "member sizes differ");
    TRY(drop_keys());
    HIP_TRY(hipStreamSynchronize(src.stream));
    const size_t n = (size_t)a.n;
    HIP_TRY(hipMemcpyAsync(x[xc] + b.slot_begin, src.x[src.xc] + a.slot_begin, sizeof(float4) * n,
                           hipMemcpyDeviceToDevice, stream));
    HIP_TRY(hipMemcpyAsync(id[ic] + b.slot_begin, src.id[src.ic] + a.slot_begin, sizeof(int) * n,
                           hipMemcpyDeviceToDevice, stream));
    if (has_pol) {
      HIP_TRY(hipMemcpyAsync(pol[xc] + b.slot_begin, src.pol[src.xc] + a.slot_begin,
                             sizeof(float2) * n, hipMemcpyDeviceToDevice, stream));
      HIP_TRY(hipMemcpyAsync(trig[xc] + b.slot_begin, src.trig[src.xc] + a.slot_begin,
                             sizeof(float4) * n, hipMemcpyDeviceToDevice, stream));
    }
    if (track_ov && src.track_ov)
      HIP_TRY(hipMemcpyAsync(ov[xc] + b.slot_begin, src.ov[src.xc] + a.slot_begin,
                             sizeof(float4) * n, hipMemcpyDeviceToDevice, stream));
    write_step_count[kd] = src.write_step_count[ks];
    return TECS_OK;
  }

  // Device bytes a member of this description adds to a batch (state, scratch and grid).
  static size_t estimate_bytes(const tecs_sim_desc& d) {
    const size_t slots = ((size_t)d.capacity + kBlock - 1) / kBlock * kBlock + kBlock;
    const bool pol = (d.fields & TECS_FIELD_POLARITY) != 0;
    size_t per_slot = 2 * 16 + 2 * 4 + 16 + 16 * 2 /* x, id, d1, ov */ + 8 * 4 /* keys.. */;
    if (pol) per_slot += 2 * 8 + 8 + 2 * 16 /* pol, d1 pol, trig */;
    if (d.integrator == TECS_HEUN) per_slot += 16 + (pol ? 8 : 0);
    if (d.integrator == TECS_HEUN && (d.com_fixer == TECS_GLOBAL_COM_FIXER || delta_only(d.pair_finder)))
      per_slot += 16 + (pol ? 8 : 0);
    if (is_symmetric(d.pair_finder) && is_binned(d.pair_finder)) per_slot += 16 + (pol ? 8 : 0);
    size_t bins = 0;
    if (is_binned(d.pair_finder)) {
      const tecs_pair_finder_settings& s = d.pair_finder_settings;
      const float bin = s.bin_size_scale * d.cutoff_distance;
      double b = 1.0;
      for (int k = 0; k < 3; ++k) b *= std::ceil((double)((s.max_bounds[k] - s.min_bounds[k]) / bin));
      bins = (size_t)b + 64;
    }
    size_t dl = 0;  // NaiveDelaunay counters, one per (cell, partner) slot pair
    if (d.pair_finder == kNaiveDelaunay) dl = (slots - kBlock) * (slots - kBlock) * 4;
    return slots * per_slot + bins * 28 + dl + 2 * 4096;
  }

  // Which members' state became non-finite at the end of some step since creation (or the last
  // clear): flags[k] = 1 for member k. Synchronous.
  int get_nonfinite(int32_t* flags) {
    if (!flags) return fail(TECS_ERR_INVALID, "null flags");
    HIP_TRY(hipMemcpyAsync(flags, nonfinite, sizeof(int) * n_inst, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return TECS_OK;
  }
  int clear_nonfinite() {
    HIP_TRY(hipMemsetAsync(nonfinite, 0, sizeof(int) * n_inst, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return TECS_OK;
  }

  int get_stats(tecs_stats* out) {
    std::vector<unsigned long long> bp((size_t)n_blocks * kMaxTeamLanes);
    HIP_TRY(hipMemcpyAsync(bp.data(), block_pairs, sizeof(unsigned long long) * bp.size(),
                           hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    tecs_stats s = stats;
    s.pair_interactions = 0;
    for (auto v : bp) s.pair_interactions += v;
    *out = s;
    return TECS_OK;
  }

  int reset_stats() {
    HIP_TRY(hipMemsetAsync(block_pairs, 0, sizeof(unsigned long long) * n_blocks * kMaxTeamLanes, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    stats = tecs_stats{};
    return TECS_OK;
  }
};


}  // namespace tecs_amd
