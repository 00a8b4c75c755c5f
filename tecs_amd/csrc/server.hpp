// server.hpp - continuous-batching serving loop over batches (see server.hip).
#pragma once

#include <chrono>
#include <list>
#include <map>
#include <memory>
#include <vector>

#include "batch.hpp"

namespace tecs_amd {

struct Request {
  int64_t id = 0;
  tecs_sim_desc desc{};
  double dt = 0.0;
  int32_t n_steps = 0;
  int32_t steps_done = 0;
  size_t bytes = 0;
  std::vector<float> pos, pol;  // initial state, then the result
  std::vector<uint8_t> types;
  int slot = -1;  // member index in its group's batch while running
  bool done = false;       // finished, failed or cancelled: no longer advanced
  bool failed = false;     // a step ended with a non-finite state
  bool cancelled = false;
  bool cancel = false;     // cancellation asked while running: leaves at the chunk's end
  double t_submit = 0.0, t_admit = 0.0, t_done = 0.0;
};

class Server {
 public:
  Server(uint64_t budget_bytes, int max_chunk);
  ~Server();
  int submit(const tecs_sim_desc* d, const float* pos, const float* pol, const uint8_t* types,
             double dt, int32_t n_steps, int64_t* id_out);
  int advance(int32_t* running, int32_t* waiting);
  int info(int64_t id, tecs_request_info* out) const;
  int cancel(int64_t id);
  int result(int64_t id, float* pos, float* pol);
  int stats(tecs_server_stats* out);

 private:
  struct Group {
    tecs_sim_desc key;
    hipStream_t stream = nullptr;     // the group's batches run here, one after another
    Batch* batch = nullptr;
    std::vector<Request*> members;    // members of `batch`, in member-index order
    std::vector<Request*> joining;    // admitted, waiting for the next re-form
    bool busy = false;                // a chunk is in flight on `stream`
    bool retired = false;             // some members finished since the last re-form
    int chunk_steps = 0;              // steps of the chunk in flight
    int tuned = -1;                   // sweep variant the autotuner picked for this group
    int64_t tuned_cells = 0;          // the group's cell count when it was tuned
    int period = 0;                   // rebuild_every_n of the group's binned members (0: every step)
    int phase = 0;                    // steps since the batch was formed, modulo period
  };
  int regroup(Group& g, std::vector<Request*> members);
  int finish_chunk(Group& g);
  int start_chunk(Group& g);

  uint64_t budget;
  uint64_t used = 0;
  int chunk;
  DevicePool pool;
  std::chrono::steady_clock::time_point t0;
  int64_t next_id = 1;
  std::map<int64_t, std::unique_ptr<Request>> requests;
  std::list<Request*> pending;
  std::list<Group> groups;  // stable addresses
  uint64_t retired_pairs = 0, steps_run = 0, rebuilds = 0, completed = 0, failed = 0, cancelled = 0;
  double regroup_s = 0.0;
};

}  // namespace tecs_amd
