// tecs_api.hip - the C ABI (include/tecs_amd.h) over the device engine (batch.hpp).
#include "batch.hpp"
#include "server.hpp"

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace tecs_amd {
thread_local std::string g_err;
}  // namespace tecs_amd

using tecs_amd::Batch;
using tecs_amd::fail;

// A simulation is a batch of one; the batch is replaced when set_capacity grows past its
// allocation.
struct tecs_sim {
  std::unique_ptr<Batch> b;
};
struct tecs_batch {
  Batch b;
};
struct tecs_writer {
  tecs_amd::Writer w;
};
struct tecs_server {
  tecs_amd::Server s;
  tecs_server(uint64_t budget, int chunk) : s(budget, chunk) {}
};

extern "C" {

void tecs_default_pair_finder_settings(tecs_pair_finder_settings* s) {
  if (!s) return;
  for (int k = 0; k < 3; ++k) {
    s->min_bounds[k] = -20.0f;
    s->max_bounds[k] = 20.0f;
  }
  s->bin_size_scale = 1.0f;
  s->rebuild_every_n = 0;
  s->gabriel_coefficient = 0.8f;
}

void tecs_default_sim_desc(tecs_sim_desc* d) {
  if (!d) return;
  std::memset(d, 0, sizeof(*d));
  d->agent_count = 0;
  d->capacity = 0;
  d->fields = 0;
  d->cutoff_distance = 1.0f;
  d->pair_finder = TECS_BINNED_ALL_PAIRS;
  tecs_default_pair_finder_settings(&d->pair_finder_settings);
  d->integrator = TECS_HEUN;
  d->com_fixer = TECS_NO_COM_FIXER;
  d->force.homeostatic_radius = 0.8f;
  d->force.repulsion_scale = 2.0f;
  d->force.adhesion_scale = 1.0f;
  d->force.n_types = 0;
  d->force.bending_strength = 0.0f;
  d->force.alignment_strength = 0.0f;
  d->track_old_velocities = 1;
}

int tecs_abi_version(void) { return TECS_AMD_ABI_VERSION; }

const char* tecs_last_error(void) { return tecs_amd::g_err.c_str(); }

int tecs_device_count(int32_t* count) {
  if (!count) return fail(TECS_ERR_INVALID, "null count");
  int c = 0;
  const hipError_t e = hipGetDeviceCount(&c);
  *count = (e == hipSuccess) ? c : 0;
  return TECS_OK;
}

#define GUARD(h)                                                        \
  do {                                                                  \
    if (!(h)) return fail(TECS_ERR_INVALID, "null handle");             \
  } while (0)

int tecs_sim_create(const tecs_sim_desc* desc, void* stream, tecs_sim** out) {
  if (!out) return fail(TECS_ERR_INVALID, "null output handle");
  *out = nullptr;
  tecs_sim* s = new (std::nothrow) tecs_sim();
  if (!s) return fail(TECS_ERR_OOM, "host allocation failed");
  s->b.reset(new (std::nothrow) Batch());
  if (!s->b) {
    delete s;
    return fail(TECS_ERR_OOM, "host allocation failed");
  }
  const int rc = s->b->init(desc, 1, stream);
  if (rc != TECS_OK) {
    delete s;
    return rc;
  }
  *out = s;
  return TECS_OK;
}

int tecs_sim_destroy(tecs_sim* sim) {
  delete sim;
  return TECS_OK;
}

int tecs_sim_upload(tecs_sim* sim, const float* positions, const float* polarities,
                    const uint8_t* types) {
  GUARD(sim);
  return sim->b->upload(0, positions, polarities, types);
}

int tecs_sim_download(tecs_sim* sim, float* positions, float* polarities, float* old_velocities) {
  GUARD(sim);
  return sim->b->download(0, positions, polarities, old_velocities);
}

int tecs_sim_upload_range(tecs_sim* sim, uint32_t first, uint32_t count, const float* positions,
                          const float* polarities, const uint8_t* types) {
  GUARD(sim);
  return sim->b->upload_range(0, first, count, positions, polarities, types);
}

int tecs_sim_set_agent_count(tecs_sim* sim, uint32_t agent_count) {
  GUARD(sim);
  return sim->b->set_agent_count(0, agent_count);
}

int tecs_sim_set_capacity(tecs_sim* sim, uint32_t capacity) {
  GUARD(sim);
  Batch& old = *sim->b;
  if (capacity < old.descs[0].agent_count) return fail(TECS_ERR_INVALID, "capacity below agent_count");
  if (capacity <= (uint32_t)old.h_inst[0].padded) return old.set_capacity(0, capacity);
  // past the allocation: a new batch for the larger capacity takes over the state
  tecs_sim_desc d = old.descs[0];
  d.capacity = capacity;
  std::unique_ptr<Batch> nb(new (std::nothrow) Batch());
  if (!nb) return fail(TECS_ERR_OOM, "host allocation failed");
  TRY(nb->init(&d, 1, old.own_stream ? nullptr : old.stream));
  TRY(nb->adopt_member(old));
  sim->b = std::move(nb);
  return TECS_OK;
}

int tecs_sim_get_population(tecs_sim* sim, uint32_t* agent_count, uint32_t* capacity) {
  GUARD(sim);
  if (agent_count) *agent_count = sim->b->descs[0].agent_count;
  if (capacity) *capacity = sim->b->descs[0].capacity;
  return TECS_OK;
}

int tecs_sim_nearest(tecs_sim* sim, int32_t n, const float* points, int32_t search_radius,
                     int32_t* out) {
  GUARD(sim);
  return sim->b->query(tecs_amd::kQueryNearest, n, points, search_radius, 0.0f, nullptr, out);
}

int tecs_sim_nearest_noalloc(tecs_sim* sim, int32_t n, const float* points, int32_t search_radius,
                             int32_t* out) {
  GUARD(sim);
  return sim->b->nearest_noalloc(n, points, search_radius, out);
}

int tecs_sim_random_in_bins(tecs_sim* sim, int32_t n, const float* points, int32_t radius_bins,
                            const float* uniforms, int32_t* out) {
  GUARD(sim);
  return sim->b->query(tecs_amd::kQueryRandomInBins, n, points, radius_bins, 0.0f, uniforms, out);
}

int tecs_sim_random_neighbour(tecs_sim* sim, int32_t n, const float* points, float radius,
                              const float* uniforms, int32_t* out) {
  GUARD(sim);
  if (!(radius >= 0.0f)) return fail(TECS_ERR_INVALID, "radius must be >= 0");
  // grid.hpp:514: radius_bins = max(1, int(ceil(search_radius / bin_size)))
  const float bin = sim->b->h_inst[0].bin_size;
  const int rb = std::max(1, (int)std::ceil(radius / bin));
  return sim->b->query(tecs_amd::kQueryRandomNeighbour, n, points, rb, radius, uniforms, out);
}

int tecs_sim_set_walls(tecs_sim* sim, const tecs_wall* walls, int32_t count) {
  GUARD(sim);
  return sim->b->set_walls(0, walls, count);
}

int tecs_sim_set_links(tecs_sim* sim, const int32_t* links, int64_t count, float strength) {
  GUARD(sim);
  return sim->b->set_links(0, links, count, strength);
}

int tecs_sim_divide(tecs_sim* sim, float rate, float offset, const double* uniforms,
                    uint32_t* new_count) {
  GUARD(sim);
  return sim->b->divide(0, rate, offset, uniforms, new_count);
}

int tecs_sim_take_step(tecs_sim* sim, double dt, int32_t n_steps) {
  GUARD(sim);
  return sim->b->take_step(&dt, n_steps);
}

int tecs_sim_set_force(tecs_sim* sim, const tecs_force_desc* force) {
  GUARD(sim);
  return sim->b->set_force(0, force);
}

int tecs_sim_evaluate_force(tecs_sim* sim, int32_t is_full_step, float* dpos, float* dpol) {
  GUARD(sim);
  return sim->b->evaluate(is_full_step != 0, dpos, dpol);
}

int tecs_sim_pairs(tecs_sim* sim, int32_t is_full_step, int64_t* row_ptr, int32_t* cols,
                   int64_t cols_capacity, int64_t* n_pairs) {
  GUARD(sim);
  return sim->b->pairs(is_full_step != 0, row_ptr, cols, cols_capacity, n_pairs);
}

int tecs_sim_rebuild(tecs_sim* sim) {
  GUARD(sim);
  int rc = sim->b->rebuild();
  if (rc == TECS_OK) sim->b->step_count = 1;  // as a due full step: reset to 0, then ++
  return rc;
}

int tecs_sim_synchronize(tecs_sim* sim) {
  GUARD(sim);
  const hipError_t e = hipStreamSynchronize(sim->b->stream);
  if (e != hipSuccess) return fail(TECS_ERR_DEVICE, hipGetErrorString(e));
  return TECS_OK;
}

int tecs_sim_get_nonfinite(tecs_sim* sim, int32_t* flag) {
  GUARD(sim);
  return sim->b->get_nonfinite(flag);
}

int tecs_batch_get_nonfinite(tecs_batch* batch, int32_t* flags) {
  GUARD(batch);
  return batch->b.get_nonfinite(flags);
}

int tecs_sim_get_stats(tecs_sim* sim, tecs_stats* stats) {
  GUARD(sim);
  if (!stats) return fail(TECS_ERR_INVALID, "null stats");
  return sim->b->get_stats(stats);
}

int tecs_sim_reset_stats(tecs_sim* sim) {
  GUARD(sim);
  return sim->b->reset_stats();
}

int tecs_sim_set_sweep_variant(tecs_sim* sim, int32_t variant) {
  GUARD(sim);
  if (variant < TECS_SWEEP_AUTO || variant > TECS_SWEEP_TEAM16)
    return fail(TECS_ERR_INVALID, "bad sweep variant");
  sim->b->variant = variant;
  sim->b->drop_graphs();
  return TECS_OK;
}

int tecs_sim_get_sweep_variant(tecs_sim* sim, int32_t* variant) {
  GUARD(sim);
  if (!variant) return fail(TECS_ERR_INVALID, "null variant");
  *variant = sim->b->concrete_variant();
  return TECS_OK;
}

int tecs_sim_get_autotune(tecs_sim* sim, double* ms_per_variant) {
  GUARD(sim);
  if (!ms_per_variant) return fail(TECS_ERR_INVALID, "null output");
  for (int v = 0; v < 6; ++v) ms_per_variant[v] = sim->b->tune_ms[v];
  return TECS_OK;
}

int tecs_sim_set_graph_mode(tecs_sim* sim, int32_t mode) {
  GUARD(sim);
  if (mode < TECS_GRAPHS_AUTO || mode > TECS_GRAPHS_ON) return fail(TECS_ERR_INVALID, "bad graph mode");
  sim->b->graph_mode = mode;
  return TECS_OK;
}

int tecs_sim_get_graph_launches(tecs_sim* sim, int64_t* launches) {
  GUARD(sim);
  if (!launches) return fail(TECS_ERR_INVALID, "null output");
  *launches = sim->b->graph_launches;
  return TECS_OK;
}

int tecs_sim_set_kernel_timing(tecs_sim* sim, int32_t enabled) {
  GUARD(sim);
  sim->b->timing = enabled != 0;
  return TECS_OK;
}

int tecs_sim_get_kernel_timing(tecs_sim* sim, int32_t kernel_class, double* ms,
                               int64_t* launches) {
  GUARD(sim);
  if (kernel_class < 0 || kernel_class >= tecs_amd::kTimeClasses)
    return fail(TECS_ERR_INVALID, "bad kernel class");
  const int rc = sim->b->fold_timing();
  if (rc) return rc;
  if (ms) *ms = sim->b->t_ms[kernel_class];
  if (launches) *launches = sim->b->t_launches[kernel_class];
  return TECS_OK;
}

// ------------------------------------------------------------------------------ batch
int tecs_batch_create(const tecs_sim_desc* descs, int32_t count, void* stream, tecs_batch** out) {
  if (!out || !descs) return fail(TECS_ERR_INVALID, "null argument");
  *out = nullptr;
  tecs_batch* b = new (std::nothrow) tecs_batch();
  if (!b) return fail(TECS_ERR_OOM, "host allocation failed");
  const int rc = b->b.init(descs, count, stream);
  if (rc != TECS_OK) {
    delete b;
    return rc;
  }
  *out = b;
  return TECS_OK;
}

int tecs_batch_destroy(tecs_batch* batch) {
  delete batch;
  return TECS_OK;
}

int tecs_batch_upload(tecs_batch* batch, int32_t index, const float* positions,
                      const float* polarities, const uint8_t* types) {
  GUARD(batch);
  return batch->b.upload(index, positions, polarities, types);
}

int tecs_batch_download(tecs_batch* batch, int32_t index, float* positions, float* polarities,
                        float* old_velocities) {
  GUARD(batch);
  return batch->b.download(index, positions, polarities, old_velocities);
}

int tecs_batch_take_step(tecs_batch* batch, const double* dts, int32_t n_steps) {
  GUARD(batch);
  if (!dts) return fail(TECS_ERR_INVALID, "null dt array");
  return batch->b.take_step(dts, n_steps);
}

int tecs_batch_set_force(tecs_batch* batch, int32_t index, const tecs_force_desc* force) {
  GUARD(batch);
  return batch->b.set_force(index, force);
}

int tecs_batch_upload_range(tecs_batch* batch, int32_t index, uint32_t first, uint32_t count,
                            const float* positions, const float* polarities, const uint8_t* types) {
  GUARD(batch);
  return batch->b.upload_range(index, first, count, positions, polarities, types);
}

int tecs_batch_set_agent_count(tecs_batch* batch, int32_t index, uint32_t agent_count) {
  GUARD(batch);
  return batch->b.set_agent_count(index, agent_count);
}

int tecs_batch_set_walls(tecs_batch* batch, int32_t index, const tecs_wall* walls, int32_t count) {
  GUARD(batch);
  return batch->b.set_walls(index, walls, count);
}

int tecs_batch_set_links(tecs_batch* batch, int32_t index, const int32_t* links, int64_t count,
                         float strength) {
  GUARD(batch);
  return batch->b.set_links(index, links, count, strength);
}

int tecs_batch_divide(tecs_batch* batch, int32_t index, float rate, float offset,
                      const double* uniforms, uint32_t* new_count) {
  GUARD(batch);
  return batch->b.divide(index, rate, offset, uniforms, new_count);
}

int tecs_batch_set_capacity(tecs_batch* batch, int32_t index, uint32_t capacity) {
  GUARD(batch);
  return batch->b.set_capacity(index, capacity);
}

int tecs_batch_synchronize(tecs_batch* batch) {
  GUARD(batch);
  const hipError_t e = hipStreamSynchronize(batch->b.stream);
  if (e != hipSuccess) return fail(TECS_ERR_DEVICE, hipGetErrorString(e));
  return TECS_OK;
}

int tecs_batch_get_stats(tecs_batch* batch, tecs_stats* stats) {
  GUARD(batch);
  if (!stats) return fail(TECS_ERR_INVALID, "null stats");
  return batch->b.get_stats(stats);
}

int tecs_batch_reset_stats(tecs_batch* batch) {
  GUARD(batch);
  return batch->b.reset_stats();
}

int tecs_batch_set_sweep_variant(tecs_batch* batch, int32_t variant) {
  GUARD(batch);
  if (variant < TECS_SWEEP_AUTO || variant > TECS_SWEEP_TEAM16)
    return fail(TECS_ERR_INVALID, "bad sweep variant");
  batch->b.variant = variant;
  batch->b.drop_graphs();
  return TECS_OK;
}

int tecs_batch_get_sweep_variant(tecs_batch* batch, int32_t* variant) {
  GUARD(batch);
  if (!variant) return fail(TECS_ERR_INVALID, "null variant");
  *variant = batch->b.concrete_variant();
  return TECS_OK;
}

int tecs_batch_get_autotune(tecs_batch* batch, double* ms_per_variant) {
  GUARD(batch);
  if (!ms_per_variant) return fail(TECS_ERR_INVALID, "null output");
  for (int v = 0; v < 6; ++v) ms_per_variant[v] = batch->b.tune_ms[v];
  return TECS_OK;
}

int tecs_batch_get_size_classes(tecs_batch* batch, int32_t* member_class, int32_t* class_variant) {
  GUARD(batch);
  if (!member_class || !class_variant) return fail(TECS_ERR_INVALID, "null output");
  Batch& b = batch->b;
  if (b.member_class.empty()) return fail(TECS_ERR_STATE, "the autotuner has not run");
  static_assert(Batch::kMaxClasses == TECS_MAX_SIZE_CLASSES, "class count of the ABI");
  for (int k = 0; k < b.n_inst; ++k) member_class[k] = b.member_class[k];
  for (int c = 0; c < Batch::kMaxClasses; ++c) class_variant[c] = b.class_variant[c];
  return TECS_OK;
}

int tecs_batch_set_size_class_variants(tecs_batch* batch, const int32_t* class_variant) {
  GUARD(batch);
  if (!class_variant) return fail(TECS_ERR_INVALID, "null class variants");
  if (!tecs_amd::is_binned(batch->b.finder)) return fail(TECS_ERR_UNSUPPORTED, "variants per size class need a binned finder");
  return batch->b.set_class_variants(class_variant);
}

int tecs_batch_set_graph_mode(tecs_batch* batch, int32_t mode) {
  GUARD(batch);
  if (mode < TECS_GRAPHS_AUTO || mode > TECS_GRAPHS_ON) return fail(TECS_ERR_INVALID, "bad graph mode");
  batch->b.graph_mode = mode;
  return TECS_OK;
}

int tecs_batch_get_graph_launches(tecs_batch* batch, int64_t* launches) {
  GUARD(batch);
  if (!launches) return fail(TECS_ERR_INVALID, "null output");
  *launches = batch->b.graph_launches;
  return TECS_OK;
}

int tecs_batch_set_kernel_timing(tecs_batch* batch, int32_t enabled) {
  GUARD(batch);
  batch->b.timing = enabled != 0;
  return TECS_OK;
}

int tecs_batch_get_kernel_timing(tecs_batch* batch, int32_t kernel_class, double* ms,
                                 int64_t* launches) {
  GUARD(batch);
  if (kernel_class < 0 || kernel_class >= tecs_amd::kTimeClasses)
    return fail(TECS_ERR_INVALID, "bad kernel class");
  const int rc = batch->b.fold_timing();
  if (rc) return rc;
  if (ms) *ms = batch->b.t_ms[kernel_class];
  if (launches) *launches = batch->b.t_launches[kernel_class];
  return TECS_OK;
}

// ---------------------------------------------------------------------------- snapshots
int tecs_writer_open(const char* path, int32_t write_xmf, tecs_writer** out) {
  if (!path || !out) return fail(TECS_ERR_INVALID, "null argument");
  *out = nullptr;
  tecs_writer* w = new (std::nothrow) tecs_writer();
  if (!w) return fail(TECS_ERR_OOM, "host allocation failed");
  std::string err;
  if (w->w.open(path, write_xmf != 0, &err) != 0) {
    delete w;
    return fail(err.find("libhdf5") != std::string::npos ? TECS_ERR_UNSUPPORTED : TECS_ERR_INVALID, err);
  }
  *out = w;
  return TECS_OK;
}

int tecs_writer_close(tecs_writer* w) {
  GUARD(w);
  std::string err;
  const int rc = w->w.close(&err);
  delete w;
  if (rc != 0) return fail(TECS_ERR_INVALID, err);
  return TECS_OK;
}

int tecs_writer_flush(tecs_writer* w) {
  GUARD(w);
  std::string err;
  if (w->w.flush(&err) != 0) return fail(TECS_ERR_INVALID, err);
  return TECS_OK;
}

int tecs_writer_write(tecs_writer* w, double time, uint32_t step, uint32_t n, const float* positions,
                      const float* polarities) {
  GUARD(w);
  if (n && !positions) return fail(TECS_ERR_INVALID, "positions required");
  std::string err;
  if (w->w.write_step(time, step, n, positions, polarities, &err) != 0) return fail(TECS_ERR_INVALID, err);
  return TECS_OK;
}

int tecs_writer_write_static(tecs_writer* w, const char* label, uint32_t n, const uint8_t* data) {
  GUARD(w);
  if (!label || (n && !data)) return fail(TECS_ERR_INVALID, "null argument");
  std::string err;
  if (w->w.write_static_u8(label, n, data, &err) != 0) return fail(TECS_ERR_STATE, err);
  return TECS_OK;
}

int tecs_sim_write(tecs_sim* sim, tecs_writer* w, double time) {
  GUARD(sim);
  GUARD(w);
  return sim->b->write_snapshot(0, &w->w, time);
}

int tecs_batch_write(tecs_batch* batch, int32_t index, tecs_writer* w, double time) {
  GUARD(batch);
  GUARD(w);
  return batch->b.write_snapshot(index, &w->w, time);
}

int tecs_read_dataset(const char* h5_path, const char* dataset, float* out, int64_t capacity,
                      int64_t* dims) {
  if (!h5_path || !dataset || !dims) return fail(TECS_ERR_INVALID, "null argument");
  std::string err;
  const int rc = tecs_amd::read_dataset(h5_path, dataset, out, capacity, dims, &err);
  if (rc == -2) return fail(TECS_ERR_CAPACITY, err);
  if (rc != 0) return fail(err.find("libhdf5") != std::string::npos ? TECS_ERR_UNSUPPORTED : TECS_ERR_INVALID, err);
  return TECS_OK;
}

int tecs_read_dataset_typed(const char* h5_path, const char* dataset, int32_t dtype, void* out,
                            int64_t capacity, int64_t* dims) {
  if (!h5_path || !dataset || !dims) return fail(TECS_ERR_INVALID, "null argument");
  std::string err;
  const int rc = tecs_amd::read_dataset_typed(h5_path, dataset, dtype, out, capacity, dims, &err);
  if (rc == -2) return fail(TECS_ERR_CAPACITY, err);
  if (rc != 0) return fail(err.find("libhdf5") != std::string::npos ? TECS_ERR_UNSUPPORTED : TECS_ERR_INVALID, err);
  return TECS_OK;
}

// ------------------------------------------------------------------------------ server
int tecs_server_create(uint64_t hbm_budget_bytes, int32_t max_chunk_steps, tecs_server** out) {
  if (!out) return fail(TECS_ERR_INVALID, "null output handle");
  *out = new (std::nothrow) tecs_server(hbm_budget_bytes, max_chunk_steps);
  if (!*out) return fail(TECS_ERR_OOM, "host allocation failed");
  return TECS_OK;
}

int tecs_server_destroy(tecs_server* server) {
  delete server;
  return TECS_OK;
}

int tecs_server_submit(tecs_server* server, const tecs_sim_desc* desc, const float* positions,
                       const float* polarities, const uint8_t* types, double dt,
                       int32_t n_steps, int64_t* request_id) {
  GUARD(server);
  if (!request_id) return fail(TECS_ERR_INVALID, "null request id");
  return server->s.submit(desc, positions, polarities, types, dt, n_steps, request_id);
}

int tecs_server_advance(tecs_server* server, int32_t* running, int32_t* pending) {
  GUARD(server);
  return server->s.advance(running, pending);
}

int tecs_server_request_info(tecs_server* server, int64_t request_id, tecs_request_info* info) {
  GUARD(server);
  if (!info) return fail(TECS_ERR_INVALID, "null info");
  return server->s.info(request_id, info);
}

int tecs_server_result(tecs_server* server, int64_t request_id, float* positions, float* polarities) {
  GUARD(server);
  return server->s.result(request_id, positions, polarities);
}

int tecs_server_cancel(tecs_server* server, int64_t request_id) {
  GUARD(server);
  return server->s.cancel(request_id);
}

int tecs_server_get_stats(tecs_server* server, tecs_server_stats* stats) {
  GUARD(server);
  if (!stats) return fail(TECS_ERR_INVALID, "null stats");
  return server->s.stats(stats);
}

int tecs_copy_probe(int64_t bytes, int32_t repeats, double* mean_ms) {
  if (bytes < 16 || repeats < 1 || !mean_ms) return fail(TECS_ERR_INVALID, "bad copy probe request");
  const int64_t n = bytes / 16;
  void *src = nullptr, *dst = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  int rc = TECS_OK;
  float ms = 0.0f;
  auto check = [&](hipError_t e) {
    if (e != hipSuccess && rc == TECS_OK) rc = fail(TECS_ERR_DEVICE, hipGetErrorString(e));
    return e == hipSuccess;
  };
  if (check(hipMalloc(&src, n * 16)) && check(hipMalloc(&dst, n * 16)) &&
      check(hipMemset(src, 0, n * 16)) && check(hipEventCreate(&a)) && check(hipEventCreate(&b))) {
    tecs_amd::launch_copy_probe((const float4*)src, (float4*)dst, n, nullptr);  // warm-up
    check(hipEventRecord(a, nullptr));
    for (int r = 0; r < repeats; ++r)
      tecs_amd::launch_copy_probe((const float4*)src, (float4*)dst, n, nullptr);
    check(hipEventRecord(b, nullptr));
    if (check(hipEventSynchronize(b)) && check(hipEventElapsedTime(&ms, a, b)))
      *mean_ms = ms / repeats;
  }
  if (a) hipEventDestroy(a);
  if (b) hipEventDestroy(b);
  if (src) hipFree(src);
  if (dst) hipFree(dst);
  return rc;
}

}  // extern "C"
