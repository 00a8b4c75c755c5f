// adapter_check.cpp - drives the C++ adapter (include/tecs_amd.hpp) the way a TECS program
// drives Simulation<Config>.
//
//   adapter_check host
//       host-side checks that need no GPU: the Settings defaults are the reference's, the
//       force builders fill tecs_force_desc in the documented order, bad compositions throw,
//       and on a machine without a HIP device creating a Simulation throws tecs_amd::Error.
//   adapter_check epithelium <in.bin> <out.bin> <steps> <dt> [<snapshots>]
//       BinnedGabriel + Heun with polarity, force PiecewiseLinear(0.8, 2, 1) + Bending() +
//       Alignment(); in.bin = uint32 n, float32 bounds[6], positions[n*3], polarities[n*2];
//       out.bin = positions[n*3], polarities[n*2], uint64 pair interactions. With <snapshots>,
//       Simulation::write at the start and the end into <snapshots>.h5 (groups t0, t1).
//   adapter_check sorting <in.bin> <out.bin> <steps> <dt> <tau_ab>
//       BinnedAllPairs + Heun with types, DifferentialAdhesion(0.8, 2, {{1, tau}, {tau, 1}});
//       in.bin = uint32 n, float32 bounds[6], positions[n*3], uint8 types[n]; out.bin =
//       positions[n*3], uint64 pair interactions.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <memory>
#include <type_traits>

#include "tecs_amd.hpp"

using namespace tecs_amd;

struct EpitheliumConfig : DefaultSimulationConfig {
  TECS_CONFIG_PAIR_FINDER(pair_finders::BinnedGabriel)
  TECS_CONFIG_INTEGRATOR(integrators::Heun)
  TECS_CONFIG_FIELDS(fields::polarity)
};

struct SortingConfig : DefaultSimulationConfig {
  TECS_CONFIG_PAIR_FINDER(pair_finders::BinnedAllPairs)
  TECS_CONFIG_INTEGRATOR(integrators::Heun)
  TECS_CONFIG_FIELDS(fields::type)
};

struct DelaunayConfig : DefaultSimulationConfig {
  TECS_CONFIG_PAIR_FINDER(pair_finders::NaiveDelaunay)
};

// every member of the adapter compiles for these configurations
template class tecs_amd::Simulation<EpitheliumConfig>;
template class tecs_amd::Simulation<SortingConfig>;
template class tecs_amd::Simulation<DelaunayConfig>;

#define EXPECT(c)                                                   \
  do {                                                              \
    if (!(c)) {                                                     \
      std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
      return 1;                                                     \
    }                                                               \
  } while (0)

static int host_checks() {
  pair_finders::BinnedGabriel::Settings g;
  EXPECT(g.min_bounds[0] == -20.0f && g.max_bounds[2] == 20.0f);
  EXPECT(g.bin_size_scale == 1.0f && g.rebuild_every_n == 0 && g.gabriel_coefficient == 0.8f);
  pair_finders::NaiveGabriel::Settings ng;
  EXPECT(ng.gabriel_coefficient == 0.8f);
  pair_finders::BinnedGabrielSymmetric::Settings gs{0.9f};  // coefficient first, as in the reference
  EXPECT(gs.gabriel_coefficient == 0.9f && gs.min_bounds[1] == -20.0f && gs.bin_size_scale == 1.0f);
  EXPECT((std::is_same<pair_finders::BinnedAllPairsSymmetric::Settings,
                       pair_finders::BinnedAllPairs::Settings>::value));
  EXPECT(pair_finders::NaiveGabrielSymmetric::kind == TECS_NAIVE_GABRIEL_SYMMETRIC);
  EXPECT(pair_finders::NaiveDelaunay::kind == TECS_NAIVE_DELAUNAY);
  EXPECT((std::is_same<DefaultSimulationConfig::pair_finder_t, pair_finders::NaiveAllPairs>::value));
  EXPECT((std::is_same<DefaultSimulationConfig::integrator_t, integrators::Heun>::value));

  const auto f = forces::PiecewiseLinear(0.7f, 0.8f, 1.0f) + forces::Bending(0.5f) + forces::Alignment();
  EXPECT(f.desc.homeostatic_radius == 0.7f && f.desc.repulsion_scale == 0.8f);
  EXPECT(f.desc.adhesion_scale == 1.0f && f.desc.n_types == 0);
  EXPECT(f.desc.bending_strength == 0.5f && f.desc.alignment_strength == 1.0f);
  EXPECT(f.descriptor().term_order[0] == TECS_TERM_PIECEWISE_LINEAR &&
         f.descriptor().term_order[1] == TECS_TERM_BENDING && f.descriptor().term_order[2] == TECS_TERM_ALIGNMENT);
  // the composition order is kept: bending first, then PiecewiseLinear (kernel_fuser.hpp:40-55)
  const auto g2 = forces::Bending() + forces::RelaxForce(0.8f);
  EXPECT(g2.desc.homeostatic_radius == 0.8f && g2.desc.repulsion_scale == 2.0f && g2.desc.bending_strength == 1.0f);
  EXPECT(g2.descriptor().term_order[0] == TECS_TERM_BENDING &&
         g2.descriptor().term_order[1] == TECS_TERM_PIECEWISE_LINEAR && g2.descriptor().term_order[2] == TECS_TERM_END);
  const auto g3 = forces::Alignment(0.25f) + forces::Drag(1.0f) + forces::RelaxForce(0.8f);
  EXPECT(g3.descriptor().term_order[0] == TECS_TERM_ALIGNMENT &&
         g3.descriptor().term_order[1] == TECS_TERM_PIECEWISE_LINEAR && g3.desc.drag == 1.0f &&
         g3.desc.drag_by_type == 0 && g3.desc.alignment_strength == 0.25f);
  const auto g4 = forces::Drag(2.0f);  // drag alone: a zero-scale PiecewiseLinear carries it
  EXPECT(g4.descriptor().term_order[0] == TECS_TERM_PIECEWISE_LINEAR && g4.descriptor().repulsion_scale == 0.0f &&
         g4.descriptor().drag == 2.0f);
  const auto g5 = forces::DifferentialAdhesion(0.8f, 2.0f, {{1.0f, 0.3f}, {0.3f, 1.0f}}) +
                  forces::DragByType({{1.0f, 0.0f}, {0.0f, 1.0f}});
  EXPECT(g5.desc.drag_by_type == 1 && g5.desc.n_types == 2 && g5.desc.drag_table[1] == 0.0f &&
         g5.desc.drag_table[3] == 1.0f);
  const auto t = forces::DifferentialAdhesion(0.8f, 2.0f, {{1.0f, 0.3f}, {0.3f, 1.0f}});
  EXPECT(t.desc.n_types == 2 && t.desc.adhesion_table[1] == 0.3f && t.desc.adhesion_table[3] == 1.0f);
  bool threw = false;
  try {
    (void)(forces::PiecewiseLinear(1, 1, 1) + forces::RelaxForce(0.8f));  // the same term twice
  } catch (const Error& e) {
    threw = e.code() == TECS_ERR_INVALID;
  }
  EXPECT(threw);
  threw = false;
  try {
    (void)forces::DifferentialAdhesion(0.8f, 2.0f, {{1.0f, 0.3f}});
  } catch (const Error&) {
    threw = true;
  }
  EXPECT(threw);

  int32_t devices = 0;
  EXPECT(tecs_device_count(&devices) == TECS_OK || devices == 0);
  if (devices == 0) {
    threw = false;
    try {
      Simulation<EpitheliumConfig> sim(100, 1.0f);
    } catch (const Error& e) {
      threw = e.code() != TECS_OK && std::string(e.what()).size() > 0;
    }
    EXPECT(threw);
  }
  std::printf("host checks passed (devices=%d)\n", devices);
  return 0;
}

template <class T>
static void read_vec(std::ifstream& in, std::vector<T>& v) {
  in.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
}

static int run_epithelium(const char* in_path, const char* out_path, int steps, double dt,
                          const char* h5_path) {
  std::ifstream in(in_path, std::ios::binary);
  uint32_t n = 0;
  float b[6];
  in.read(reinterpret_cast<char*>(&n), 4);
  in.read(reinterpret_cast<char*>(b), sizeof(b));
  std::vector<std::array<float, 3>> x(n);
  std::vector<std::array<float, 2>> p(n);
  read_vec(in, x);
  read_vec(in, p);
  if (!in) throw Error(TECS_ERR_INVALID, "short input file");
  pair_finders::BinnedGabriel::Settings s;
  for (int k = 0; k < 3; ++k) {
    s.min_bounds[k] = b[k];
    s.max_bounds[k] = b[3 + k];
  }
  Simulation<EpitheliumConfig> sim(n, 1.0f, s);
  sim.set_state(x, p);
  const auto force = forces::PiecewiseLinear(0.8f, 2.0f, 1.0f) + forces::Bending() + forces::Alignment();
  std::unique_ptr<Writer> writer;
  if (h5_path) {
    writer.reset(new Writer(h5_path));
    sim.write(*writer, 0.0);
  }
  for (int i = 0; i < steps; ++i) sim.take_step(dt, force);
  if (writer) {
    sim.write(*writer, steps * dt);
    writer->flush();
  }
  sim.synchronize();
  const auto xo = sim.positions();
  const auto po = sim.polarities();
  const uint64_t pairs = sim.stats().pair_interactions;
  std::ofstream out(out_path, std::ios::binary);
  out.write(reinterpret_cast<const char*>(xo.data()), (std::streamsize)(xo.size() * 12));
  out.write(reinterpret_cast<const char*>(po.data()), (std::streamsize)(po.size() * 8));
  out.write(reinterpret_cast<const char*>(&pairs), 8);
  std::printf("epithelium n=%u steps=%d pairs=%llu\n", n, steps, (unsigned long long)pairs);
  return 0;
}

static int run_sorting(const char* in_path, const char* out_path, int steps, double dt, float tau) {
  std::ifstream in(in_path, std::ios::binary);
  uint32_t n = 0;
  float b[6];
  in.read(reinterpret_cast<char*>(&n), 4);
  in.read(reinterpret_cast<char*>(b), sizeof(b));
  std::vector<std::array<float, 3>> x(n);
  std::vector<uint8_t> t(n);
  read_vec(in, x);
  read_vec(in, t);
  if (!in) throw Error(TECS_ERR_INVALID, "short input file");
  pair_finders::BinnedAllPairs::Settings s;
  for (int k = 0; k < 3; ++k) {
    s.min_bounds[k] = b[k];
    s.max_bounds[k] = b[3 + k];
  }
  Simulation<SortingConfig> sim(n, 1.0f, s);
  sim.set_state(x, {}, t);
  const auto force = forces::DifferentialAdhesion(0.8f, 2.0f, {{1.0f, tau}, {tau, 1.0f}});
  for (int i = 0; i < steps; ++i) sim.take_step(dt, force);
  sim.synchronize();
  const auto xo = sim.positions();
  const uint64_t pairs = sim.stats().pair_interactions;
  std::ofstream out(out_path, std::ios::binary);
  out.write(reinterpret_cast<const char*>(xo.data()), (std::streamsize)(xo.size() * 12));
  out.write(reinterpret_cast<const char*>(&pairs), 8);
  std::printf("sorting n=%u steps=%d pairs=%llu\n", n, steps, (unsigned long long)pairs);
  return 0;
}

int main(int argc, char** argv) {
  try {
    const std::string mode = argc > 1 ? argv[1] : "host";
    if (mode == "host") return host_checks();
    if (mode == "epithelium" && (argc == 6 || argc == 7))
      return run_epithelium(argv[2], argv[3], std::atoi(argv[4]), std::atof(argv[5]),
                            argc == 7 ? argv[6] : nullptr);
    if (mode == "sorting" && argc == 7)
      return run_sorting(argv[2], argv[3], std::atoi(argv[4]), std::atof(argv[5]),
                         (float)std::atof(argv[6]));
    std::fprintf(stderr, "usage: adapter_check host | epithelium in out steps dt | sorting in out steps dt tau\n");
    return 2;
  } catch (const Error& e) {
    std::fprintf(stderr, "tecs_amd::Error %d: %s\n", e.code(), e.what());
    return 3;
  }
}
