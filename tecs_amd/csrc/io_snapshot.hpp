// io_snapshot.hpp - HDF5 writer with an asynchronous GPU hand-off (see io_snapshot.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/tecs_amd.h"

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <fstream>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace tecs_amd {

// One snapshot in flight: device staging in agent order, pinned host copies, the event that
// marks the end of the device-to-host copy. Owned by a simulation; busy while queued.
struct SnapshotSlot {
  size_t n = 0;
  bool has_pol = false;
  float* d_pos = nullptr;
  float* d_pol = nullptr;
  float* h_pos = nullptr;  // pinned
  float* h_pol = nullptr;  // pinned
  hipEvent_t gathered = nullptr;
  hipEvent_t copied = nullptr;
  double time = 0.0;
  unsigned step = 0;
  std::vector<uint32_t> links;  // the member's links (a, b pairs) at the time of the write
  bool busy = false;
  std::mutex* mu = nullptr;               // the owner's slot mutex
  std::condition_variable* cv = nullptr;  // signalled when a slot frees
  void release();
};

class Writer {
 public:
  int open(const std::string& path, bool xmf, std::string* err);
  int close(std::string* err);
  // synchronous writes (host data)
  // links: n_links (a, b) pairs of uint32, or null
  int write_step(double time, unsigned step, size_t n, const float* pos, const float* pol,
                 std::string* err, const uint32_t* links = nullptr, size_t n_links = 0);
  int write_static_u8(const std::string& label, size_t n, const uint8_t* data, std::string* err);
  // asynchronous: the slot's copy is in flight; the writer thread waits for it and writes
  void enqueue(SnapshotSlot* slot);
  int flush(std::string* err);

  size_t buffer_threshold = 65536;  // HDF5_Writer::Settings::buffer_treshold

 private:
  void run();
  int64_t file = -1;
  std::string filename;
  bool write_xmf = true;
  std::ofstream xmf_file;
  std::string xmf_buffer, xmf_static;
  bool static_written = false;
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<SnapshotSlot*> queue;
  int in_flight = 0;
  bool stopping = false;
  std::string first_error;
};

int read_dataset(const std::string& h5_path, const std::string& name, float* out, int64_t capacity,
                 int64_t* dims, std::string* err);
int read_dataset_typed(const std::string& h5_path, const std::string& name, int dtype, void* out,
                       int64_t capacity, int64_t* dims, std::string* err);

}  // namespace tecs_amd
