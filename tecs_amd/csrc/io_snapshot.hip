// io_snapshot.hip - HDF5 snapshots of the simulation state, written off the critical path.
//
// File layout of HDF5_Writer (include/io/hdf5_writer.hpp): <path>.h5 (truncated on open) with
// one group "t<step>" per write holding one dataset per field, named by the field label
// ("position" [n x 3] float, "polarity" [n x 2] float; write_single :120-129, write :306-321),
// a "static" group (write_static :335-349) and the XDMF companion <path>.xmf (:131-304).
//
// The hand-off (tecs_sim_write): on the simulation's stream one kernel gathers the state into
// agent order in a device staging slot; a side stream (default priority) waits for it and copies
// the slot into pinned host memory; a writer thread waits for that copy and writes the group.
// The simulation's stream only carries the gather kernel, so steps continue while the copy and
// the file write run. A small pool of slots bounds the snapshots in flight.
//
// libhdf5 is loaded with dlopen when a writer is opened (TECS_HDF5_LIB, else libhdf5.so, else
// /opt/conda/lib/libhdf5.so), so the engine library does not depend on it otherwise.
#include "io_snapshot.hpp"

#include <dlfcn.h>

#include <cstdio>
#include <cstring>
#include <filesystem>

namespace tecs_amd {
namespace {

using hid_t = int64_t;
using herr_t = int;
using hsize_t = unsigned long long;
constexpr hid_t kH5PDefault = 0;
constexpr hid_t kH5SAll = 0;
constexpr unsigned kH5FAccTrunc = 0x0002u;
constexpr unsigned kH5FAccRdonly = 0x0000u;

struct H5Api {
  void* handle = nullptr;
  herr_t (*open)() = nullptr;
  hid_t (*fcreate)(const char*, unsigned, hid_t, hid_t) = nullptr;
  hid_t (*fopen)(const char*, unsigned, hid_t) = nullptr;
  herr_t (*fclose)(hid_t) = nullptr;
  herr_t (*fflush)(hid_t, int) = nullptr;
  hid_t (*gcreate)(hid_t, const char*, hid_t, hid_t, hid_t) = nullptr;
  herr_t (*gclose)(hid_t) = nullptr;
  hid_t (*screate_simple)(int, const hsize_t*, const hsize_t*) = nullptr;
  herr_t (*sclose)(hid_t) = nullptr;
  int (*sget_dims)(hid_t, hsize_t*, hsize_t*) = nullptr;
  int (*sget_ndims)(hid_t) = nullptr;
  hid_t (*dcreate)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t) = nullptr;
  hid_t (*dopen)(hid_t, const char*, hid_t) = nullptr;
  hid_t (*dget_space)(hid_t) = nullptr;
  herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*) = nullptr;
  herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*) = nullptr;
  herr_t (*dclose)(hid_t) = nullptr;
  hid_t* native_float = nullptr;
  hid_t* native_uchar = nullptr;
  hid_t* native_uint = nullptr;
  hid_t* native_double = nullptr;
  hid_t* native_int = nullptr;
  hid_t* native_llong = nullptr;
  std::string error;

  bool load() {
    if (handle) return true;
    const char* env = std::getenv("TECS_HDF5_LIB");
    const char* cands[] = {env, "libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so"};
    for (const char* c : cands) {
      if (!c) continue;
      handle = dlopen(c, RTLD_NOW | RTLD_LOCAL);
      if (handle) break;
    }
    if (!handle) {
      error = "libhdf5 not found (set TECS_HDF5_LIB)";
      return false;
    }
    bool ok = true;
    auto sym = [&](auto& fn, const char* name) {
      fn = reinterpret_cast<std::remove_reference_t<decltype(fn)>>(dlsym(handle, name));
      if (!fn) {
        ok = false;
        error = std::string("libhdf5 lacks ") + name;
      }
    };
    sym(open, "H5open");
    sym(fcreate, "H5Fcreate");
    sym(fopen, "H5Fopen");
    sym(fclose, "H5Fclose");
    sym(fflush, "H5Fflush");
    sym(gcreate, "H5Gcreate2");
    sym(gclose, "H5Gclose");
    sym(screate_simple, "H5Screate_simple");
    sym(sclose, "H5Sclose");
    sym(sget_dims, "H5Sget_simple_extent_dims");
    sym(sget_ndims, "H5Sget_simple_extent_ndims");
    sym(dcreate, "H5Dcreate2");
    sym(dopen, "H5Dopen2");
    sym(dget_space, "H5Dget_space");
    sym(dwrite, "H5Dwrite");
    sym(dread, "H5Dread");
    sym(dclose, "H5Dclose");
    sym(native_float, "H5T_NATIVE_FLOAT_g");
    sym(native_uchar, "H5T_NATIVE_UCHAR_g");
    sym(native_uint, "H5T_NATIVE_UINT_g");
    sym(native_double, "H5T_NATIVE_DOUBLE_g");
    sym(native_int, "H5T_NATIVE_INT_g");
    sym(native_llong, "H5T_NATIVE_LLONG_g");
    if (ok && open() < 0) {
      ok = false;
      error = "H5open failed";
    }
    if (!ok) {
      dlclose(handle);
      handle = nullptr;
    }
    return ok;
  }
};

H5Api& h5() {
  static H5Api api;
  return api;
}
std::mutex& h5_mutex() {
  static std::mutex m;
  return m;
}

}  // namespace

// -------------------------------------------------------------------------------- Writer
int Writer::open(const std::string& path, bool xmf, std::string* err) {
  std::lock_guard<std::mutex> g(h5_mutex());
  if (!h5().load()) {
    *err = h5().error;
    return -1;
  }
  std::filesystem::path p(path);
  std::error_code ec;
  if (p.has_parent_path()) std::filesystem::create_directories(p.parent_path(), ec);
  filename = p.filename().string();
  file = h5().fcreate((path + ".h5").c_str(), kH5FAccTrunc, kH5PDefault, kH5PDefault);
  if (file < 0) {
    *err = "cannot create " + path + ".h5";
    return -1;
  }
  write_xmf = xmf;
  if (write_xmf) {
    // init_xmf, hdf5_writer.hpp:131-141
    xmf_file.open(path + ".xmf", std::ios::out | std::ios::trunc);
    if (xmf_file.is_open())
      xmf_buffer +=
          "<?xml version=\"1.0\"?>\n<Xdmf Version=\"3.0\">\n\t<Domain>\n"
          "\t\t<Grid Name=\"Agents\" GridType=\"Collection\" CollectionType=\"Temporal\">\n";
  }
  worker = std::thread([this] { run(); });
  return 0;
}

int Writer::close(std::string* err) {
  int rc = flush(err);
  {
    std::lock_guard<std::mutex> g(mu);
    stopping = true;
  }
  cv.notify_all();
  if (worker.joinable()) worker.join();
  std::lock_guard<std::mutex> g(h5_mutex());
  if (file >= 0) {
    h5().fclose(file);
    file = -1;
  }
  // finalize_xmf, hdf5_writer.hpp:146-153
  if (xmf_file.is_open()) {
    xmf_buffer += "\t\t</Grid>\n\t</Domain>\n</Xdmf>\n";
    xmf_file.write(xmf_buffer.data(), (std::streamsize)xmf_buffer.size());
    xmf_file.close();
  }
  return rc;
}

static bool write_dataset(hid_t group, const char* label, hsize_t n, hsize_t dims, hid_t type,
                          const void* data) {
  if (n == 0) return true;  // write_single skips empty views (hdf5_writer.hpp:121-122)
  const hsize_t shape[2] = {n, dims};
  const hid_t space = h5().screate_simple(dims > 0 ? 2 : 1, shape, nullptr);
  if (space < 0) return false;
  const hid_t ds = h5().dcreate(group, label, type, space, kH5PDefault, kH5PDefault, kH5PDefault);
  bool ok = ds >= 0 && h5().dwrite(ds, type, kH5SAll, kH5SAll, kH5PDefault, data) >= 0;
  if (ds >= 0) h5().dclose(ds);
  h5().sclose(space);
  return ok;
}

// HDF5_Writer::write, hdf5_writer.hpp:306-321, for the fields position (+ polarity) and the
// links (Simulation::write passes its links view last, simulation.hpp:607-611; a Link is two
// uint32, types/link.hpp:30-84).
int Writer::write_step(double time, unsigned step, size_t n, const float* pos, const float* pol,
                       std::string* err, const uint32_t* links, size_t n_links) {
  std::lock_guard<std::mutex> g(h5_mutex());
  const std::string gname = "t" + std::to_string(step);
  const hid_t group = h5().gcreate(file, gname.c_str(), kH5PDefault, kH5PDefault, kH5PDefault);
  if (group < 0) {
    *err = "cannot create group " + gname;
    return -1;
  }
  bool ok = write_dataset(group, "position", n, 3, *h5().native_float, pos);
  if (ok && pol) ok = write_dataset(group, "polarity", n, 2, *h5().native_float, pol);
  if (ok && links) ok = write_dataset(group, "links", n_links, 2, *h5().native_uint, links);
  h5().gclose(group);
  h5().fflush(file, 0);
  if (!ok) {
    *err = "dataset write failed in " + gname;
    return -1;
  }
  if (!write_xmf || !xmf_file.is_open()) return 0;
  // write_xmf_grid_start, hdf5_writer.hpp:163-197: with links, the links are the topology
  const bool has_links = links && n_links > 0;
  xmf_buffer += "\t\t\t<Grid Name=\"t" + std::to_string(step) + "\" GridType=\"Uniform\">\n";
  xmf_buffer += "\t\t\t\t<Time Value=\"" + std::to_string(time);
  if (!has_links)
    xmf_buffer += "\"/>\n\t\t\t\t<Topology TopologyType=\"Polyvertex\" NumberOfElements=\"" +
                  std::to_string(n);
  xmf_buffer += "\"/>\n\t\t\t\t<Geometry GeometryType=\"XYZ\">\n\t\t\t\t\t<DataItem Format=\"HDF\" "
                "Dimensions=\"" + std::to_string(n) + " 3\">\n\t\t\t\t\t\t" + filename + ".h5:t" +
                std::to_string(step) + "/position\n\t\t\t\t\t</DataItem>\n\t\t\t\t</Geometry>\n";
  // write_xmf_grid_item, hdf5_writer.hpp:205-266, for the polarity (rank 2, 2 components)
  if (pol && n > 0) {
    xmf_buffer += "\t\t\t\t<Attribute Name=\"polarity\" AttributeType=\"Vector\" Center=\"Node\">\n"
                  "\t\t\t\t\t<DataItem Format=\"HDF\" Dimensions=\"" + std::to_string(n) +
                  " 2 \">\n\t\t\t\t\t\t" + filename + ".h5:" + gname +
                  "/polarity\n\t\t\t\t\t</DataItem>\n\t\t\t\t</Attribute>\n";
  }
  // write_xmf_grid_item for the links view: a polyline topology (hdf5_writer.hpp:241-266)
  if (has_links) {
    xmf_buffer += "\t\t\t\t<Topology TopologyType=\"Polyline\" NumberOfElements=\"" +
                  std::to_string(n_links) + "\">\n\t\t\t\t\t<DataItem Format=\"HDF\" "
                  "DataType=\"UInt32\" Dimensions=\"" + std::to_string(n_links) + " 2 \">\n\t\t\t\t\t\t" +
                  filename + ".h5:" + gname + "/links\n\t\t\t\t\t</DataItem>\n\t\t\t\t</Topology>\n";
  }
  // write_xmf_grid_end, hdf5_writer.hpp:274-285
  xmf_buffer += xmf_static;
  xmf_buffer += "\t\t\t</Grid>\n";
  if (xmf_buffer.size() > buffer_threshold) {
    xmf_file.write(xmf_buffer.data(), (std::streamsize)xmf_buffer.size());
    xmf_buffer.clear();
  }
  return 0;
}

// HDF5_Writer::write_static, hdf5_writer.hpp:335-349, for one byte-valued field.
int Writer::write_static_u8(const std::string& label, size_t n, const uint8_t* data,
                            std::string* err) {
  std::lock_guard<std::mutex> g(h5_mutex());
  if (static_written) {
    // hdf5_writer.hpp:337-338
    *err = "Static data has already been written. You can only write to static data once.";
    return -1;
  }
  const hid_t group = h5().gcreate(file, "static", kH5PDefault, kH5PDefault, kH5PDefault);
  if (group < 0) {
    *err = "cannot create group static";
    return -1;
  }
  const bool ok = write_dataset(group, label.c_str(), n, 0, *h5().native_uchar, data);
  h5().gclose(group);
  h5().fflush(file, 0);
  if (!ok) {
    *err = "static dataset write failed";
    return -1;
  }
  static_written = true;
  if (write_xmf && xmf_file.is_open() && n > 0)
    xmf_static += "\t\t\t\t<Attribute Name=\"" + label +
                  "\" AttributeType=\"Scalar\" Center=\"Node\">\n\t\t\t\t\t<DataItem Format=\"HDF\" "
                  "Dimensions=\"" + std::to_string(n) + " \">\n\t\t\t\t\t\t" + filename +
                  ".h5:static/" + label + "\n\t\t\t\t\t</DataItem>\n\t\t\t\t</Attribute>\n";
  return 0;
}

void Writer::enqueue(SnapshotSlot* slot) {
  {
    std::lock_guard<std::mutex> g(mu);
    queue.push_back(slot);
    ++in_flight;
  }
  cv.notify_all();
}

void Writer::run() {
  for (;;) {
    SnapshotSlot* slot = nullptr;
    {
      std::unique_lock<std::mutex> g(mu);
      cv.wait(g, [&] { return stopping || !queue.empty(); });
      if (queue.empty()) return;
      slot = queue.front();
      queue.pop_front();
    }
    std::string err;
    const hipError_t e = hipEventSynchronize(slot->copied);
    int rc = 0;
    if (e != hipSuccess) {
      err = std::string("snapshot copy failed: ") + hipGetErrorString(e);
      rc = -1;
    } else {
      rc = write_step(slot->time, slot->step, slot->n, slot->h_pos, slot->has_pol ? slot->h_pol : nullptr, &err,
                      slot->links.empty() ? nullptr : slot->links.data(), slot->links.size() / 2);
    }
    {
      std::lock_guard<std::mutex> g(mu);
      if (rc != 0 && first_error.empty()) first_error = err;
      --in_flight;
    }
    slot->release();
    cv.notify_all();
  }
}

int Writer::flush(std::string* err) {
  std::unique_lock<std::mutex> g(mu);
  cv.wait(g, [&] { return in_flight == 0; });
  if (!first_error.empty()) {
    *err = first_error;
    first_error.clear();
    return -1;
  }
  return 0;
}

// ----------------------------------------------------------------------------- SnapshotSlot
void SnapshotSlot::release() {
  {
    // notify while holding the owner's mutex: once it is released the owner may see every
    // slot free, return from its teardown wait and destroy the condition variable
    std::lock_guard<std::mutex> g(*mu);
    busy = false;
    cv->notify_all();
  }
}

// --------------------------------------------------------------------------------- reading
// HDF5_Reader::read_dataset, include/io/hdf5_reader.hpp:66-71 (get_dataset_dimensions, :43-53,
// with out null): a dataset of rank 1 or 2 read into the memory type `dtype` (HDF5 converts
// from the file type on read, as HighFive does for the View's element type).
int read_dataset_typed(const std::string& h5_path, const std::string& name, int dtype, void* out,
                       int64_t capacity, int64_t* dims, std::string* err) {
  std::lock_guard<std::mutex> g(h5_mutex());
  if (!h5().load()) {
    *err = h5().error;
    return -1;
  }
  hid_t mem = -1;
  switch (dtype) {
    case TECS_DTYPE_FLOAT32: mem = *h5().native_float; break;
    case TECS_DTYPE_FLOAT64: mem = *h5().native_double; break;
    case TECS_DTYPE_INT32: mem = *h5().native_int; break;
    case TECS_DTYPE_UINT32: mem = *h5().native_uint; break;
    case TECS_DTYPE_UINT8: mem = *h5().native_uchar; break;
    case TECS_DTYPE_INT64: mem = *h5().native_llong; break;
    default:
      *err = "unknown element type";
      return -3;
  }
  const hid_t f = h5().fopen(h5_path.c_str(), kH5FAccRdonly, kH5PDefault);
  if (f < 0) {
    *err = "cannot open " + h5_path;
    return -1;
  }
  int rc = 0;
  const hid_t ds = h5().dopen(f, name.c_str(), kH5PDefault);
  if (ds < 0) {
    *err = "no dataset " + name;
    rc = -1;
  } else {
    const hid_t sp = h5().dget_space(ds);
    const int rank = h5().sget_ndims(sp);
    if (rank < 0 || rank > 2) {
      *err = "dataset " + name + " has rank " + std::to_string(rank) + " (1 or 2 supported)";
      rc = -3;
    } else {
      hsize_t shape[2] = {0, 1};
      h5().sget_dims(sp, shape, nullptr);
      dims[0] = rank >= 1 ? (int64_t)shape[0] : 1;
      dims[1] = rank >= 2 ? (int64_t)shape[1] : 1;
    }
    h5().sclose(sp);
    const int64_t total = rc == 0 ? dims[0] * dims[1] : 0;
    if (rc == 0 && out) {
      if (total > capacity) {
        *err = "buffer too small";
        rc = -2;
      } else if (total > 0 && h5().dread(ds, mem, kH5SAll, kH5SAll, kH5PDefault, out) < 0) {
        *err = "read failed";
        rc = -1;
      }
    }
    h5().dclose(ds);
  }
  h5().fclose(f);
  return rc;
}

int read_dataset(const std::string& h5_path, const std::string& name, float* out, int64_t capacity,
                 int64_t* dims, std::string* err) {
  return read_dataset_typed(h5_path, name, TECS_DTYPE_FLOAT32, out, capacity, dims, err);
}

}  // namespace tecs_amd
