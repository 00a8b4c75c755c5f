// server.hip - a continuous-batching serving loop for simulation requests.
//
// Requests (a simulation description, its initial state, dt and a number of steps) arrive at
// any time. Each scheduling round (tecs_server_advance):
//   1. admission: pending requests are admitted in arrival order while their estimated device
//      bytes fit the HBM budget (a request that does not fit waits; later, smaller requests may
//      pass it);
//   2. regrouping: requests are grouped by what a batch must share (pair finder, integrator, COM
//      fixer, fields, old-velocity tracking); a group whose membership changed gets a new batch,
//      the running members' state is copied into it device to device and new members are
//      uploaded;
//   3. stepping: every group advances by the same number of steps, the smallest remaining step
//      count of its members capped at max_chunk; the groups run concurrently on their own
//      streams;
//   4. retirement: finished members are downloaded into their request and leave their group;
//      their bytes return to the budget.
// The state of a running request stays on the device between rounds. Every request rebuilds its
// grid at every full step (rebuild_every_n = 0 is required), so regrouping changes nothing in
// its arithmetic: a request's result is bitwise the result of the same simulation run alone.
// This serving layer has no counterpart in the reference.
#include "server.hpp"

#include <chrono>

namespace tecs_amd {

namespace {
double now_s(const std::chrono::steady_clock::time_point& t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

bool same_group(const tecs_sim_desc& a, const tecs_sim_desc& b) {
  return a.pair_finder == b.pair_finder && a.integrator == b.integrator &&
         a.com_fixer == b.com_fixer && a.fields == b.fields &&
         (a.track_old_velocities != 0) == (b.track_old_velocities != 0);
}
}  // namespace

Server::Server(uint64_t budget_bytes, int max_chunk)
    : budget(budget_bytes), chunk(std::max(1, max_chunk)), t0(std::chrono::steady_clock::now()) {}

Server::~Server() {
  for (auto& g : groups) delete g.batch;
}

int Server::submit(const tecs_sim_desc* d, const float* pos, const float* pol, const uint8_t* types,
                   double dt, int32_t n_steps, int64_t* id_out) {
  int rc = validate_desc(d);
  if (rc) return rc;
  if (is_binned(d->pair_finder) && d->pair_finder_settings.rebuild_every_n != 0)
    return fail(TECS_ERR_UNSUPPORTED, "served requests must rebuild the grid every step");
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

int Server::regroup(Group& g, std::vector<Request*> members) {
  // members: the group's requests for the next batch, running ones first (they own state)
  std::unique_ptr<Batch> nb;
  if (!members.empty()) {
    std::vector<tecs_sim_desc> ds;
    for (Request* r : members) ds.push_back(r->desc);
    nb = std::make_unique<Batch>();
    int rc = nb->init(ds.data(), (int)ds.size(), nullptr);
    if (rc) return rc;
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
  return TECS_OK;
}

int Server::advance(int32_t* running, int32_t* waiting) {
  // 1. admission
  std::vector<Request*> admitted;
  for (auto it = pending.begin(); it != pending.end();) {
    Request* r = *it;
    if (used + r->bytes <= budget) {
      used += r->bytes;
      admitted.push_back(r);
      it = pending.erase(it);
    } else {
      ++it;
    }
  }
  // 2. regroup groups that gained members
  for (Request* r : admitted) {
    Group* g = nullptr;
    for (auto& c : groups)
      if (same_group(c.key, r->desc)) g = &c;
    if (!g) {
      groups.push_back(Group{r->desc, nullptr, {}, {}});
      g = &groups.back();
    }
    g->joining.push_back(r);
  }
  for (auto& g : groups) {
    if (g.joining.empty()) continue;
    std::vector<Request*> members = g.members;
    for (Request* r : g.joining) members.push_back(r);
    g.joining.clear();
    int rc = regroup(g, members);
    if (rc) return rc;
  }
  // 3. stepping: all groups enqueue, then wait
  int n_run = 0;
  for (auto& g : groups) {
    if (g.members.empty()) continue;
    int steps = chunk;
    for (Request* r : g.members) steps = std::min(steps, r->n_steps - r->steps_done);
    std::vector<double> dts;
    for (Request* r : g.members) dts.push_back(r->dt);
    if (steps > 0) {
      int rc = g.batch->take_step(dts.data(), steps);
      if (rc) return rc;
    }
    g.pending_steps = steps;
    n_run += (int)g.members.size();
  }
  for (auto& g : groups) {
    if (g.members.empty()) continue;
    const hipError_t e = hipStreamSynchronize(g.batch->stream);
    if (e != hipSuccess) return fail(TECS_ERR_DEVICE, hipGetErrorString(e));
    for (Request* r : g.members) r->steps_done += g.pending_steps;
    steps_run += (uint64_t)g.pending_steps * g.members.size();
  }
  // 4. retirement
  for (auto& g : groups) {
    if (g.members.empty()) continue;
    std::vector<Request*> keep;
    bool changed = false;
    for (Request* r : g.members) {
      if (r->steps_done < r->n_steps) {
        keep.push_back(r);
        continue;
      }
      const size_t n = r->desc.agent_count;
      r->pos.assign(3 * n, 0.0f);
      if (r->desc.fields & TECS_FIELD_POLARITY) r->pol.assign(2 * n, 0.0f);
      int rc = g.batch->download(r->slot, r->pos.data(), r->pol.empty() ? nullptr : r->pol.data(), nullptr);
      if (rc) return rc;
      r->t_done = now_s(t0);
      r->done = true;
      r->slot = -1;
      used -= r->bytes;
      completed++;
      changed = true;
    }
    if (changed) {
      int rc = regroup(g, keep);
      if (rc) return rc;
    }
  }
  if (running) *running = n_run;
  if (waiting) *waiting = (int32_t)pending.size();
  return TECS_OK;
}

int Server::info(int64_t id, tecs_request_info* out) const {
  auto it = requests.find(id);
  if (it == requests.end()) return fail(TECS_ERR_INVALID, "unknown request");
  const Request& r = *it->second;
  out->state = r.done ? TECS_REQUEST_DONE : (r.slot >= 0 ? TECS_REQUEST_RUNNING : TECS_REQUEST_PENDING);
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
  if (!r.done) return fail(TECS_ERR_STATE, "request not finished");
  if (pos) std::memcpy(pos, r.pos.data(), sizeof(float) * r.pos.size());
  if (pol && !r.pol.empty()) std::memcpy(pol, r.pol.data(), sizeof(float) * r.pol.size());
  return TECS_OK;
}

int Server::stats(tecs_server_stats* out) {
  out->completed = completed;
  out->pending = pending.size();
  out->running = 0;
  uint64_t pairs = retired_pairs;
  for (auto& g : groups) {
    out->running += g.members.size();
    if (g.batch) {
      tecs_stats st{};
      int rc = g.batch->get_stats(&st);
      if (rc) return rc;
      pairs += st.pair_interactions;
    }
  }
  out->pair_interactions = pairs;
  out->member_steps = steps_run;
  out->regroups = rebuilds;
  out->hbm_bytes_in_use = used;
  out->hbm_budget = budget;
  return TECS_OK;
}

}  // namespace tecs_amd
