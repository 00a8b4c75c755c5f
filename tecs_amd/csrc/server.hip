// server.hip - a continuous-batching serving loop for simulation requests.
//
// Requests (a simulation description, its initial state, dt and a number of steps) arrive at
// any time. Requests that share what a batch must share (pair finder, integrator, COM fixer,
// fields, old-velocity tracking) form a group; each group owns a stream and one batch holding
// its running members. tecs_server_advance:
//   1. admission: pending requests are admitted in arrival order while their estimated device
//      bytes fit the HBM budget (a request that does not fit waits; later, smaller requests may
//      pass it) and join their group's waiting list;
//   2. every group whose stream is idle is serviced: the chunk that just ended is accounted,
//      finished members are downloaded and leave (their bytes return to the budget), the batch
//      is re-formed once if members left or joined (running members' state is copied device
//      to device, new members are uploaded), and the next chunk is enqueued: the smallest
//      remaining step count of the members, at most max_chunk steps;
//   3. if no group was idle, the call waits until one is.
// Groups therefore advance independently: a group of small simulations does not wait for the
// chunk of a group of large ones, and their kernels overlap on the device. The state of a
// running request stays on the device between chunks, and batch buffers come from a device
// pool, so a re-form costs no hipMalloc/hipFree in the steady state. The autotuner's choice is
// kept per group and re-used while the group's cell count stays within a factor of two.
// A re-formed batch rebuilds its grid at its first full step. Requests that rebuild every step
// (rebuild_every_n = 0) do so anyway; requests with a period R > 0 are grouped by R, and their
// group is re-formed only where every continuing member is at a rebuild point of its own
// schedule (its steps so far a multiple of R: the reference rebuilds at steps 1, R + 1, ...,
// binned_all_pairs.hpp:77-84). A member that finishes between two such points stays in the batch
// (its result already taken) until the next one, and joiners wait for it. So regrouping changes
// nothing in a request's arithmetic: its result is bitwise the result of the same simulation run
// alone.
// Failure handling: the kernels that end a step flag members whose new state is not finite;
// after each chunk such requests fail (state TECS_REQUEST_FAILED) and leave their group at its
// next re-form, and the others go on unaffected. tecs_server_cancel drops a pending request at
// once and a running one at the end of its group's chunk. This serving layer has no
// counterpart in the reference.
#include "server.hpp"

#include <chrono>
#include <thread>

namespace tecs_amd {

namespace {
double now_s(const std::chrono::steady_clock::time_point& t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

int period_of(const tecs_sim_desc& d) {
  return is_binned(d.pair_finder) ? d.pair_finder_settings.rebuild_every_n : 0;
}

bool same_group(const tecs_sim_desc& a, const tecs_sim_desc& b) {
  return a.pair_finder == b.pair_finder && a.integrator == b.integrator &&
         a.com_fixer == b.com_fixer && a.fields == b.fields &&
         (a.track_old_velocities != 0) == (b.track_old_velocities != 0) &&
         period_of(a) == period_of(b);
}
}  // namespace

Server::Server(uint64_t budget_bytes, int max_chunk)
    : budget(budget_bytes),
      chunk(std::max(1, max_chunk)),
      pool(budget_bytes / 4),
      t0(std::chrono::steady_clock::now()) {}

Server::~Server() {
  for (auto& g : groups) {
    delete g.batch;  // synchronises the group's stream first
    if (g.stream) hipStreamDestroy(g.stream);
  }
}

int Server::submit(const tecs_sim_desc* d, const float* pos, const float* pol, const uint8_t* types,
                   double dt, int32_t n_steps, int64_t* id_out) {
  int rc = validate_desc(d);
  if (rc) return rc;
  if (n_steps < 0) return fail(TECS_ERR_INVALID, "n_steps must be >= 0");
  if (!pos && d->agent_count) return fail(TECS_ERR_INVALID, "positions required");
  auto r = std::make_unique<Request>();
  r->id = next_id++;
  r->desc = *d;
  r->dt = dt;
  r->n_steps = n_steps;
  r->bytes = Batch::estimate_bytes(*d);
  if (r->bytes > budget) return fail(TECS_ERR_CAPACITY, "request exceeds the HBM budget alone");
  const size_t n = d->agent_count;
  r->pos.assign(pos, pos + 3 * n);
  if ((d->fields & TECS_FIELD_POLARITY) && pol) r->pol.assign(pol, pol + 2 * n);
  if ((d->fields & TECS_FIELD_TYPE) && types) r->types.assign(types, types + n);
  r->t_submit = now_s(t0);
  *id_out = r->id;
  pending.push_back(r.get());
  requests.emplace(r->id, std::move(r));
  return TECS_OK;
}

// Replace the group's batch by one holding `members` (running ones carry their device state).
// Called only while the group's stream is idle.
int Server::regroup(Group& g, std::vector<Request*> members) {
  const auto t_start = std::chrono::steady_clock::now();
  std::unique_ptr<Batch> nb;
  if (!members.empty()) {
    std::vector<tecs_sim_desc> ds;
    int64_t cells = 0;
    for (Request* r : members) {
      ds.push_back(r->desc);
      cells += (int64_t)r->desc.agent_count;
    }
    nb = std::make_unique<Batch>();
    nb->pool = &pool;
    int rc = nb->init(ds.data(), (int)ds.size(), g.stream);
    if (rc) return rc;
    if (g.tuned >= 0 && cells <= 2 * g.tuned_cells && 2 * cells >= g.tuned_cells) nb->tuned = g.tuned;
    for (size_t k = 0; k < members.size(); ++k) {
      Request* r = members[k];
      if (r->slot >= 0) {
        rc = nb->copy_member_from(*g.batch, r->slot, (int)k);
      } else {
        rc = nb->upload((int)k, r->pos.data(), r->pol.empty() ? nullptr : r->pol.data(),
                        r->types.empty() ? nullptr : r->types.data());
        r->t_admit = now_s(t0);
      }
      if (rc) return rc;
    }
    for (size_t k = 0; k < members.size(); ++k) members[k]->slot = (int)k;
    rebuilds++;
  }
  if (g.batch) {
    tecs_stats st{};
    if (g.batch->get_stats(&st) == TECS_OK) retired_pairs += st.pair_interactions;
    delete g.batch;
  }
  g.batch = nb.release();
  g.members = std::move(members);
  g.retired = false;
  g.phase = 0;
  regroup_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  return TECS_OK;
}

// The chunk in flight on the group's (now idle) stream has ended: account it, retire members.
int Server::finish_chunk(Group& g) {
  g.busy = false;
  if (g.period > 0) g.phase = (g.phase + g.chunk_steps) % g.period;
  std::vector<int32_t> bad(g.members.size(), 0);
  if (g.batch && !g.members.empty()) TRY(g.batch->get_nonfinite(bad.data()));
  if (g.batch && g.batch->tuned >= 0 && g.tuned != g.batch->tuned) {
    g.tuned = g.batch->tuned;
    g.tuned_cells = 0;
    for (Request* r : g.members) g.tuned_cells += (int64_t)r->desc.agent_count;
  }
  for (Request* r : g.members) {
    if (r->done) continue;
    r->steps_done += g.chunk_steps;
    steps_run += (uint64_t)g.chunk_steps;
    if (bad[(size_t)r->slot] || r->cancel) {
      // leaves without a result: a step ended with a non-finite state, or it was cancelled
      r->failed = bad[(size_t)r->slot] != 0;
      r->cancelled = !r->failed;
      r->t_done = now_s(t0);
      r->done = true;
      used -= r->bytes;
      (r->failed ? failed : cancelled)++;
      g.retired = true;
      continue;
    }
    if (r->steps_done < r->n_steps) continue;
    const size_t n = r->desc.agent_count;
    r->pos.assign(3 * n, 0.0f);
    if (r->desc.fields & TECS_FIELD_POLARITY) r->pol.assign(2 * n, 0.0f);
    int rc = g.batch->download(r->slot, r->pos.data(), r->pol.empty() ? nullptr : r->pol.data(), nullptr);
    if (rc) return rc;
    r->t_done = now_s(t0);
    r->done = true;
    used -= r->bytes;
    completed++;
    g.retired = true;
  }
  g.chunk_steps = 0;
  return TECS_OK;
}

// Re-form the (idle) group if its membership changed, then enqueue its next chunk.
int Server::start_chunk(Group& g) {
  bool live = false;  // members that continue through a re-form
  for (Request* r : g.members) live = live || !r->done;
  // a re-form rebuilds the grid: with a rebuild period, continuing members must be at a
  // rebuild point of their schedule (every member of a batch has the same phase)
  const bool can_reform = g.period == 0 || g.phase == 0 || !live;
  const bool want_reform = g.retired || !g.joining.empty();
  if (want_reform && can_reform) {
    std::vector<Request*> members;
    for (Request* r : g.members) {
      if (r->done)
        r->slot = -1;
      else
        members.push_back(r);
    }
    for (Request* r : g.joining) members.push_back(r);
    g.joining.clear();
    TRY(regroup(g, members));
  }
  if (g.members.empty()) return TECS_OK;
  live = false;
  for (Request* r : g.members) live = live || !r->done;
  if (!live) return TECS_OK;  // only finished members, waiting for a rebuild point: none left to run
  int steps = chunk;
  for (Request* r : g.members)
    if (!r->done) steps = std::min(steps, r->n_steps - r->steps_done);
  // a re-form is waiting: stop at the next rebuild point
  if (want_reform && !can_reform) steps = std::min(steps, g.period - g.phase);
  std::vector<double> dts;
  for (Request* r : g.members) dts.push_back(r->dt);
  if (steps > 0) TRY(g.batch->take_step(dts.data(), steps));
  g.chunk_steps = steps;
  g.busy = true;
  return TECS_OK;
}

int Server::advance(int32_t* running, int32_t* waiting) {
  // 1. admission
  for (auto it = pending.begin(); it != pending.end();) {
    Request* r = *it;
    if (used + r->bytes > budget) {
      ++it;
      continue;
    }
    used += r->bytes;
    Group* g = nullptr;
    for (auto& c : groups)
      if (same_group(c.key, r->desc)) g = &c;
    if (!g) {
      groups.emplace_back();
      g = &groups.back();
      g->key = r->desc;
      g->period = period_of(r->desc);
      HIP_TRY(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    }
    g->joining.push_back(r);
    it = pending.erase(it);
  }
  // 2./3. service idle groups; wait until at least one was idle
  for (;;) {
    bool serviced = false, any_busy = false;
    for (auto& g : groups) {
      if (g.busy) {
        const hipError_t q = hipStreamQuery(g.stream);
        if (q == hipErrorNotReady) {
          any_busy = true;
          continue;
        }
        if (q != hipSuccess) return fail(TECS_ERR_DEVICE, hipGetErrorString(q));
        TRY(finish_chunk(g));
        serviced = true;
      }
      const bool had_work = g.retired || !g.joining.empty() || !g.members.empty();
      TRY(start_chunk(g));
      serviced = serviced || had_work;
      any_busy = any_busy || g.busy;
    }
    if (serviced || !any_busy) break;
    std::this_thread::sleep_for(std::chrono::microseconds(50));
  }
  if (running) {
    int32_t n_run = 0;
    for (auto& g : groups) n_run += g.busy ? (int32_t)g.members.size() : 0;
    *running = n_run;
  }
  if (waiting) *waiting = (int32_t)pending.size();
  return TECS_OK;
}

int Server::info(int64_t id, tecs_request_info* out) const {
  auto it = requests.find(id);
  if (it == requests.end()) return fail(TECS_ERR_INVALID, "unknown request");
  const Request& r = *it->second;
  out->state = r.failed      ? TECS_REQUEST_FAILED
               : r.cancelled ? TECS_REQUEST_CANCELLED
               : r.done      ? TECS_REQUEST_DONE
               : (r.slot >= 0 ? TECS_REQUEST_RUNNING : TECS_REQUEST_PENDING);
  out->steps_done = r.steps_done;
  out->t_submit = r.t_submit;
  out->t_admit = r.t_admit;
  out->t_done = r.t_done;
  out->hbm_bytes = r.bytes;
  return TECS_OK;
}

int Server::result(int64_t id, float* pos, float* pol) {
  auto it = requests.find(id);
  if (it == requests.end()) return fail(TECS_ERR_INVALID, "unknown request");
  Request& r = *it->second;
  if (r.failed) return fail(TECS_ERR_STATE, "request failed: a step ended with a non-finite state");
  if (r.cancelled) return fail(TECS_ERR_STATE, "request was cancelled");
  if (!r.done) return fail(TECS_ERR_STATE, "request not finished");
  if (pos) std::memcpy(pos, r.pos.data(), sizeof(float) * r.pos.size());
  if (pol && !r.pol.empty()) std::memcpy(pol, r.pol.data(), sizeof(float) * r.pol.size());
  return TECS_OK;
}

int Server::cancel(int64_t id) {
  auto it = requests.find(id);
  if (it == requests.end()) return fail(TECS_ERR_INVALID, "unknown request");
  Request* r = it->second.get();
  if (r->done) return fail(TECS_ERR_STATE, "request already finished");
  auto drop = [&](bool charged) {
    r->cancelled = true;
    r->done = true;
    r->t_done = now_s(t0);
    if (charged) used -= r->bytes;
    cancelled++;
  };
  for (auto p = pending.begin(); p != pending.end(); ++p)
    if (*p == r) {
      pending.erase(p);
      drop(false);
      return TECS_OK;
    }
  for (auto& g : groups)
    for (auto j = g.joining.begin(); j != g.joining.end(); ++j)
      if (*j == r) {
        g.joining.erase(j);
        drop(true);
        return TECS_OK;
      }
  r->cancel = true;  // running: leaves at the end of its group's chunk
  return TECS_OK;
}

int Server::stats(tecs_server_stats* out) {
  out->completed = completed;
  out->failed = failed;
  out->cancelled = cancelled;
  out->pending = pending.size();
  out->running = 0;
  uint64_t pairs = retired_pairs;
  for (auto& g : groups) {
    for (Request* r : g.members) out->running += r->done ? 0 : 1;
    out->running += g.joining.size();
    if (g.batch) {
      tecs_stats st{};
      TRY(g.batch->get_stats(&st));  // waits for the group's chunk in flight
      pairs += st.pair_interactions;
    }
  }
  out->pair_interactions = pairs;
  out->member_steps = steps_run;
  out->regroups = rebuilds;
  out->hbm_bytes_in_use = used;
  out->hbm_budget = budget;
  out->pool_cached_bytes = pool.cached_bytes();
  out->regroup_seconds = regroup_s;
  return TECS_OK;
}

}  // namespace tecs_amd
