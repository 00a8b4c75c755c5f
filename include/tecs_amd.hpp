/* tecs_amd.hpp - the reference's C++ simulation API for the per-step hot path, over the C ABI
 * of tecs_amd.h (header only; link with -ltecs_amd).
 *
 * A TECS program selects its components in a simulation config and drives a Simulation:
 *
 *   struct EpitheliumConfig : tecs_amd::DefaultSimulationConfig {      // simulation_config.hpp:344-379
 *     TECS_CONFIG_PAIR_FINDER(tecs_amd::pair_finders::BinnedGabriel)
 *     TECS_CONFIG_INTEGRATOR(tecs_amd::integrators::Heun)
 *     TECS_CONFIG_FIELDS(tecs_amd::fields::polarity)
 *   };
 *   tecs_amd::Simulation<EpitheliumConfig> sim(n_cells, r_max, settings);
 *   sim.set_state(positions, polarities);
 *   auto layer_force = forces::PiecewiseLinear(0.7f, 0.8f, 1.0f) + forces::Bending();
 *   for (...) { sim.take_step(dt, layer_force); sim.write(writer, t); }
 *
 * The names follow the reference: pair_finders::{NaiveAllPairs, BinnedAllPairs, NaiveGabriel,
 * BinnedGabriel} and their *Symmetric variants with their Settings (binned_all_pairs.hpp:20-25,
 * binned_gabriel.hpp:20-26, naive_gabriel.hpp:22-24, binned_gabriel_symmetric.hpp:46-52),
 * integrators::{Euler, Heun}, com_fixers::{NoComFixer,
 * GlobalComFixer}, Simulation<Config>::take_step(dt, force) (simulation.hpp:548-556) and
 * write(time) (simulation.hpp:607-611). A force is not a lambda here: the engine evaluates the
 * laws of tecs_force_desc, which the forces:: builders compose in the reference's statement
 * order (PiecewiseLinear predefined.hpp:107-144, a type table as in sorting.cpp:43-58,
 * bending_force polarity.hpp:181-198, bidirectional_polarization_force polarity.hpp:165-168).
 * Errors throw tecs_amd::Error (a std::runtime_error), as the reference's writer does
 * (hdf5_writer.hpp:337-338); the engine's message is the exception text. */
#ifndef TECS_AMD_HPP
#define TECS_AMD_HPP

#include <array>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "tecs_amd.h"

namespace tecs_amd {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int rc) {
  if (rc != TECS_OK) {
    const char* msg = tecs_last_error();
    throw Error(rc, msg && *msg ? msg : "tecs_amd call failed");
  }
}

/* ------------------------------------------------------------------ components */
namespace pair_finders {
// binned_all_pairs.hpp:14-25
struct BinnedAllPairs {
  static constexpr int32_t kind = TECS_BINNED_ALL_PAIRS;
  struct Settings {
    std::array<float, 3> min_bounds{{-20.0f, -20.0f, -20.0f}};
    std::array<float, 3> max_bounds{{20.0f, 20.0f, 20.0f}};
    float bin_size_scale = 1.0f;
    unsigned rebuild_every_n = 0;
  };
};
// naive_all_pairs.hpp: no settings
struct NaiveAllPairs {
  static constexpr int32_t kind = TECS_NAIVE_ALL_PAIRS;
  struct Settings {};
};
// binned_gabriel.hpp:20-26
struct BinnedGabriel {
  static constexpr int32_t kind = TECS_BINNED_GABRIEL;
  struct Settings {
    std::array<float, 3> min_bounds{{-20.0f, -20.0f, -20.0f}};
    std::array<float, 3> max_bounds{{20.0f, 20.0f, 20.0f}};
    float bin_size_scale = 1.0f;
    unsigned rebuild_every_n = 0;
    float gabriel_coefficient = 0.8f;
  };
};
// naive_gabriel.hpp:22-24
struct NaiveGabriel {
  static constexpr int32_t kind = TECS_NAIVE_GABRIEL;
  struct Settings {
    float gabriel_coefficient = 0.8f;
  };
};
// The symmetric finders: each unordered pair once (j < i), delta(i) += v, delta(j) -= v.
// binned_all_pairs_symmetric.hpp:43-48 (the same members as BinnedAllPairs)
struct BinnedAllPairsSymmetric {
  static constexpr int32_t kind = TECS_BINNED_ALL_PAIRS_SYMMETRIC;
  using Settings = BinnedAllPairs::Settings;
};
// naive_all_pairs_symmetric.hpp:37
struct NaiveAllPairsSymmetric {
  static constexpr int32_t kind = TECS_NAIVE_ALL_PAIRS_SYMMETRIC;
  using Settings = NaiveAllPairs::Settings;
};
// binned_gabriel_symmetric.hpp:46-52 (its own member order)
struct BinnedGabrielSymmetric {
  static constexpr int32_t kind = TECS_BINNED_GABRIEL_SYMMETRIC;
  struct Settings {
    float gabriel_coefficient = 0.8f;
    std::array<float, 3> min_bounds{{-20.0f, -20.0f, -20.0f}};
    std::array<float, 3> max_bounds{{20.0f, 20.0f, 20.0f}};
    float bin_size_scale = 1.0f;
    unsigned rebuild_every_n = 0;
  };
};
// naive_gabriel_symmetric.hpp:39-41
struct NaiveGabrielSymmetric {
  static constexpr int32_t kind = TECS_NAIVE_GABRIEL_SYMMETRIC;
  using Settings = NaiveGabriel::Settings;
};
// naive_delaunay.hpp:13-16: no settings; at most 512 cells (the capacity)
struct NaiveDelaunay {
  static constexpr int32_t kind = TECS_NAIVE_DELAUNAY;
  using Settings = NaiveAllPairs::Settings;
};
}  // namespace pair_finders

namespace integrators {
struct Euler {  // euler.hpp:37-86
  static constexpr int32_t kind = TECS_EULER;
};
struct Heun {  // heun.hpp:43-208
  static constexpr int32_t kind = TECS_HEUN;
};
}  // namespace integrators

namespace com_fixers {
struct NoComFixer {  // com_fixers.hpp:11-20
  static constexpr int32_t kind = TECS_NO_COM_FIXER;
};
struct GlobalComFixer {  // com_fixers.hpp:22-45 (divides by the capacity)
  static constexpr int32_t kind = TECS_GLOBAL_COM_FIXER;
};
}  // namespace com_fixers

// Optional fields besides the position (CONFIG_FIELDS, simulation_config.hpp:314-325).
namespace fields {
constexpr uint32_t position = 0;
constexpr uint32_t polarity = TECS_FIELD_POLARITY;
constexpr uint32_t type = TECS_FIELD_TYPE;
}  // namespace fields

// The defaults of DefaultSimulationConfig (simulation_config.hpp:344-362): NaiveAllPairs,
// NoComFixer, Heun, positions only.
struct DefaultSimulationConfig {
  using pair_finder_t = pair_finders::NaiveAllPairs;
  using com_fixer_t = com_fixers::NoComFixer;
  using integrator_t = integrators::Heun;
  static constexpr uint32_t field_mask = fields::position;
};

#define TECS_CONFIG_PAIR_FINDER(T) using pair_finder_t = T;
#define TECS_CONFIG_INTEGRATOR(T) using integrator_t = T;
#define TECS_CONFIG_COM_FIXER(T) using com_fixer_t = T;
#define TECS_CONFIG_FIELDS(mask) static constexpr uint32_t field_mask = (mask);

/* ------------------------------------------------------------------ forces */
namespace forces {

// A pairwise force of the engine: the terms of tecs_force_desc. `a + b` applies a's terms, then
// b's, per ordered pair, as take_step(dt, a, b) runs its fused forces in argument order
// (kernel_fuser.hpp:40-55); the order is kept in tecs_force_desc::term_order, so the sums are
// the reference's for every composition. Each term may appear once.
struct PairwiseForce {
  tecs_force_desc desc{};
  int terms[3] = {TECS_TERM_END, TECS_TERM_END, TECS_TERM_END};
  int n_terms = 0;
  bool has_drag = false;

  bool has(int t) const {
    for (int k = 0; k < n_terms; ++k)
      if (terms[k] == t) return true;
    return false;
  }
  PairwiseForce operator+(const PairwiseForce& o) const {
    PairwiseForce r = *this;
    for (int k = 0; k < o.n_terms; ++k) {
      const int t = o.terms[k];
      if (r.has(t)) throw Error(TECS_ERR_INVALID, "a force term appears twice in the composition");
      r.terms[r.n_terms++] = t;
      if (t == TECS_TERM_PIECEWISE_LINEAR) {
        r.desc.homeostatic_radius = o.desc.homeostatic_radius;
        r.desc.repulsion_scale = o.desc.repulsion_scale;
        r.desc.adhesion_scale = o.desc.adhesion_scale;
        if (o.desc.n_types > 0) {
          if (r.desc.drag_by_type && r.desc.n_types != o.desc.n_types)
            throw Error(TECS_ERR_INVALID, "drag and adhesion tables of different sizes");
          r.desc.n_types = o.desc.n_types;
          std::memcpy(r.desc.adhesion_table, o.desc.adhesion_table, sizeof(r.desc.adhesion_table));
        }
      } else if (t == TECS_TERM_BENDING) {
        r.desc.bending_strength = o.desc.bending_strength;
      } else {
        r.desc.alignment_strength = o.desc.alignment_strength;
      }
    }
    if (o.has_drag) {
      if (r.has_drag) throw Error(TECS_ERR_INVALID, "two drag terms");
      r.has_drag = true;
      r.desc.drag = o.desc.drag;
      r.desc.drag_by_type = o.desc.drag_by_type;
      std::memcpy(r.desc.drag_table, o.desc.drag_table, sizeof(r.desc.drag_table));
      if (o.desc.drag_by_type) {
        if (r.desc.n_types > 0 && r.desc.n_types != o.desc.n_types)
          throw Error(TECS_ERR_INVALID, "drag and adhesion tables of different sizes");
        r.desc.n_types = o.desc.n_types;
      }
    }
    for (int k = 0; k < 3; ++k) r.desc.term_order[k] = k < r.n_terms ? r.terms[k] : TECS_TERM_END;
    return r;
  }
  // the descriptor as tecs_sim_create / tecs_sim_set_force take it
  tecs_force_desc descriptor() const {
    tecs_force_desc d = desc;
    for (int k = 0; k < 3; ++k) d.term_order[k] = k < n_terms ? terms[k] : TECS_TERM_END;
    if (n_terms == 0) {
      // an all-zero term_order means the default terms; a force without terms (drag alone) is
      // a PiecewiseLinear with zero scales, whose +-0 contributions leave every sum unchanged
      d.term_order[0] = TECS_TERM_PIECEWISE_LINEAR;
      d.homeostatic_radius = d.repulsion_scale = d.adhesion_scale = 0.0f;
    }
    return d;
  }
};

inline PairwiseForce single_term(int t) {
  PairwiseForce f;
  f.terms[0] = t;
  f.n_terms = 1;
  f.desc.term_order[0] = t;
  return f;
}

// position.delta += PiecewiseLinear(disp, d, homeostatic_radius, repulsion, adhesion)
inline PairwiseForce PiecewiseLinear(float homeostatic_radius, float repulsion_scale,
                                     float adhesion_scale) {
  PairwiseForce f = single_term(TECS_TERM_PIECEWISE_LINEAR);
  f.desc.homeostatic_radius = homeostatic_radius;
  f.desc.repulsion_scale = repulsion_scale;
  f.desc.adhesion_scale = adhesion_scale;
  return f;
}

// RelaxForce(h) (initializers/relax_force.hpp:27-29) = PiecewiseLinear(h, 2, 1).
inline PairwiseForce RelaxForce(float homeostatic_radius) {
  return PiecewiseLinear(homeostatic_radius, 2.0f, 1.0f);
}

// PiecewiseLinear with the adhesion taken from table[type_i][type_j] (differential adhesion,
// the pattern of sorting.cpp:43-58); needs fields::type.
inline PairwiseForce DifferentialAdhesion(float homeostatic_radius, float repulsion_scale,
                                          const std::vector<std::vector<float>>& table) {
  const size_t n = table.size();
  if (n == 0 || n > TECS_MAX_TYPES) throw Error(TECS_ERR_INVALID, "adhesion table size");
  PairwiseForce f = PiecewiseLinear(homeostatic_radius, repulsion_scale, 0.0f);
  f.desc.n_types = (int32_t)n;
  for (size_t i = 0; i < n; ++i) {
    if (table[i].size() != n) throw Error(TECS_ERR_INVALID, "adhesion table must be square");
    for (size_t j = 0; j < n; ++j) f.desc.adhesion_table[i * n + j] = table[i][j];
  }
  return f;
}

// position.delta += s * b.vector, polarity.delta += s * b.polarity with
// b = polarity_i.bending_force(disp, polarity_j, d); needs fields::polarity.
inline PairwiseForce Bending(float strength = 1.0f) {
  PairwiseForce f = single_term(TECS_TERM_BENDING);
  f.desc.bending_strength = strength;
  return f;
}

// polarity.delta += s * polarity_i.bidirectional_polarization_force(polarity_j).
inline PairwiseForce Alignment(float strength = 1.0f) {
  PairwiseForce f = single_term(TECS_TERM_ALIGNMENT);
  f.desc.alignment_strength = strength;
  return f;
}

// drag += value for every pair (examples/intercalation.cpp:62, turing.cpp:68): the finder adds
// sum(drag * old_velocity_j) / fmax(sum(drag), FLT_EPSILON) to the position delta
// (binned_all_pairs.hpp:168-180). Needs an ordered pair finder.
inline PairwiseForce Drag(float value = 1.0f) {
  PairwiseForce f;
  f.has_drag = true;
  f.desc.drag = value;
  return f;
}

// drag += table[type_i][type_j] (the type-dependent drag of examples/tribolium/main.cpp:113-115);
// needs fields::type and, with an adhesion table, one of the same size.
inline PairwiseForce DragByType(const std::vector<std::vector<float>>& table) {
  const size_t n = table.size();
  if (n == 0 || n > TECS_MAX_TYPES) throw Error(TECS_ERR_INVALID, "drag table size");
  PairwiseForce f;
  f.has_drag = true;
  f.desc.drag_by_type = 1;
  f.desc.n_types = (int32_t)n;
  for (size_t i = 0; i < n; ++i) {
    if (table[i].size() != n) throw Error(TECS_ERR_INVALID, "drag table must be square");
    for (size_t j = 0; j < n; ++j) f.desc.drag_table[i * n + j] = table[i][j];
  }
  return f;
}

}  // namespace forces

/* ------------------------------------------------------------------ snapshot writer */
// HDF5_Writer (hdf5_writer.hpp): groups t<k>/<field>, a "static" group, an XDMF companion.
class Writer {
 public:
  explicit Writer(const std::string& path, bool write_xmf = true) {
    check(tecs_writer_open(path.c_str(), write_xmf ? 1 : 0, &h_));
  }
  ~Writer() {
    if (h_) tecs_writer_close(h_);
  }
  Writer(const Writer&) = delete;
  Writer& operator=(const Writer&) = delete;
  void flush() { check(tecs_writer_flush(h_)); }
  void write_static(const std::string& label, const std::vector<uint8_t>& data) {
    check(tecs_writer_write_static(h_, label.c_str(), (uint32_t)data.size(), data.data()));
  }
  tecs_writer* handle() const { return h_; }

 private:
  tecs_writer* h_ = nullptr;
};

/* ------------------------------------------------------------------ simulation */
template <class Config>
class Simulation {
 public:
  using pair_finder_t = typename Config::pair_finder_t;
  using integrator_t = typename Config::integrator_t;
  using com_fixer_t = typename Config::com_fixer_t;
  using Settings = typename pair_finder_t::Settings;
  static constexpr uint32_t field_mask = Config::field_mask;

  // Simulation(agent_count, ..., r_max) (simulation.hpp:162-182); capacity 0 = agent_count.
  Simulation(unsigned agent_count, float r_max, const Settings& settings = Settings{},
             unsigned capacity = 0, void* stream = nullptr) {
    tecs_sim_desc d;
    tecs_default_sim_desc(&d);
    d.agent_count = agent_count;
    d.capacity = capacity ? capacity : agent_count;
    d.fields = field_mask;
    d.cutoff_distance = r_max;
    d.pair_finder = pair_finder_t::kind;
    d.integrator = integrator_t::kind;
    d.com_fixer = com_fixer_t::kind;
    apply_settings(settings, &d.pair_finder_settings);
    check(tecs_sim_create(&d, stream, &h_));
    force_ = d.force;
    n_ = agent_count;
  }
  ~Simulation() {
    if (h_) tecs_sim_destroy(h_);
  }
  Simulation(const Simulation&) = delete;
  Simulation& operator=(const Simulation&) = delete;

  unsigned agent_count() const { return n_; }

  // State in the reference's host layout: VectorN<float, 3> per cell (x, y, z),
  // Polarity_<float> per cell (theta, phi), one type byte per cell.
  void set_state(const std::vector<std::array<float, 3>>& positions,
                 const std::vector<std::array<float, 2>>& polarities = {},
                 const std::vector<uint8_t>& types = {}) {
    if (positions.size() != n_) throw Error(TECS_ERR_INVALID, "positions size");
    if (!polarities.empty() && polarities.size() != n_) throw Error(TECS_ERR_INVALID, "polarities size");
    if (!types.empty() && types.size() != n_) throw Error(TECS_ERR_INVALID, "types size");
    check(tecs_sim_upload(h_, positions.empty() ? nullptr : positions[0].data(),
                          polarities.empty() ? nullptr : polarities[0].data(),
                          types.empty() ? nullptr : types.data()));
  }
  // Writes cells first .. first + positions.size() - 1, active or not (the reference writes into
  // its views at those indices before set_agent_count makes them active).
  void set_state_range(unsigned first, const std::vector<std::array<float, 3>>& positions,
                       const std::vector<std::array<float, 2>>& polarities = {},
                       const std::vector<uint8_t>& types = {}) {
    const size_t k = positions.size();
    if (!polarities.empty() && polarities.size() != k) throw Error(TECS_ERR_INVALID, "polarities size");
    if (!types.empty() && types.size() != k) throw Error(TECS_ERR_INVALID, "types size");
    check(tecs_sim_upload_range(h_, first, (uint32_t)k, k ? positions[0].data() : nullptr,
                                polarities.empty() ? nullptr : polarities[0].data(),
                                types.empty() ? nullptr : types.data()));
  }
  // Simulation::set_agent_count / set_capacity / get_* (simulation.hpp:257-301)
  void set_agent_count(unsigned n) {
    check(tecs_sim_set_agent_count(h_, n));
    n_ = n;
  }
  void set_capacity(unsigned capacity) { check(tecs_sim_set_capacity(h_, capacity)); }
  unsigned get_agent_count() const { return n_; }
  unsigned get_capacity() const {
    uint32_t n = 0, c = 0;
    check(tecs_sim_get_population(h_, &n, &c));
    return c;
  }

  // The generic wall forces (GENERIC_FORCE of examples/growth_w_wall.cpp:44-48): planes
  // a x + b y + c z + d = 0 as PlaneN (types/plane.hpp) with a strength each.
  void set_walls(const std::vector<tecs_wall>& walls) {
    check(tecs_sim_set_walls(h_, walls.empty() ? nullptr : walls.data(), (int32_t)walls.size()));
  }
  // The links (types/link.hpp) and the LINK_FORCE of examples/intercalation.cpp.
  void set_links(const std::vector<std::array<int32_t, 2>>& links, float strength) {
    check(tecs_sim_set_links(h_, links.empty() ? nullptr : links[0].data(), (int64_t)links.size(),
                             strength));
  }
  // The UPDATE_FUNC `proliferate` of examples/growth_w_wall.cpp:52-64 with three draws per
  // active cell; returns the count for set_agent_count.
  unsigned divide(float rate, float offset, const std::vector<std::array<double, 3>>& uniforms) {
    if (uniforms.size() != n_) throw Error(TECS_ERR_INVALID, "uniforms size");
    uint32_t out = 0;
    check(tecs_sim_divide(h_, rate, offset, uniforms.empty() ? nullptr : uniforms[0].data(), &out));
    return out;
  }

  // Grid::get_nearest_point_index (grid.hpp:413-441) for a batch of points on the simulation's
  // grid; -1 where none of the searched bins holds a cell.
  std::vector<int32_t> get_nearest_point_indices(const std::vector<std::array<float, 3>>& points,
                                                 int search_radius = 1) {
    std::vector<int32_t> out(points.size());
    check(tecs_sim_nearest(h_, (int32_t)points.size(), points.empty() ? nullptr : points[0].data(),
                           search_radius, out.data()));
    return out;
  }

  std::vector<std::array<float, 3>> positions() {
    std::vector<std::array<float, 3>> x(n_);
    check(tecs_sim_download(h_, x.empty() ? nullptr : x[0].data(), nullptr, nullptr));
    return x;
  }
  std::vector<std::array<float, 2>> polarities() {
    std::vector<std::array<float, 2>> p(n_);
    check(tecs_sim_download(h_, nullptr, p.empty() ? nullptr : p[0].data(), nullptr));
    return p;
  }
  std::vector<std::array<float, 3>> old_velocities() {
    std::vector<std::array<float, 3>> v(n_);
    check(tecs_sim_download(h_, nullptr, nullptr, v.empty() ? nullptr : v[0].data()));
    return v;
  }

  // Simulation::take_step(dt, force) (simulation.hpp:548-556). Asynchronous.
  void take_step(double dt, const forces::PairwiseForce& force, int n_steps = 1) {
    const tecs_force_desc d = force.descriptor();
    if (std::memcmp(&d, &force_, sizeof(force_)) != 0) {
      check(tecs_sim_set_force(h_, &d));
      force_ = d;
    }
    check(tecs_sim_take_step(h_, dt, n_steps));
  }

  // Simulation::write(time) (simulation.hpp:607-611): t<k>/position and t<k>/polarity,
  // handed off asynchronously.
  void write(Writer& writer, double time) { check(tecs_sim_write(h_, writer.handle(), time)); }

  // The ordered pairs the force would be called with (CSR over cells, visiting order).
  std::pair<std::vector<int64_t>, std::vector<int32_t>> pairs(bool is_full_step) {
    std::vector<int64_t> rp(n_ + 1);
    int64_t total = 0;
    check(tecs_sim_pairs(h_, is_full_step ? 1 : 0, rp.data(), nullptr, 0, &total));
    std::vector<int32_t> cols((size_t)total);
    check(tecs_sim_pairs(h_, is_full_step ? 1 : 0, rp.data(), cols.data(), total, &total));
    return {rp, cols};
  }

  void synchronize() { check(tecs_sim_synchronize(h_)); }
  tecs_stats stats() {
    tecs_stats s;
    check(tecs_sim_get_stats(h_, &s));
    return s;
  }
  tecs_sim* handle() const { return h_; }

 private:
  static void apply_settings(const pair_finders::BinnedAllPairs::Settings& s,
                             tecs_pair_finder_settings* o) {
    for (int k = 0; k < 3; ++k) {
      o->min_bounds[k] = s.min_bounds[k];
      o->max_bounds[k] = s.max_bounds[k];
    }
    o->bin_size_scale = s.bin_size_scale;
    o->rebuild_every_n = (int32_t)s.rebuild_every_n;
  }
  static void apply_settings(const pair_finders::BinnedGabriel::Settings& s,
                             tecs_pair_finder_settings* o) {
    for (int k = 0; k < 3; ++k) {
      o->min_bounds[k] = s.min_bounds[k];
      o->max_bounds[k] = s.max_bounds[k];
    }
    o->bin_size_scale = s.bin_size_scale;
    o->rebuild_every_n = (int32_t)s.rebuild_every_n;
    o->gabriel_coefficient = s.gabriel_coefficient;
  }
  static void apply_settings(const pair_finders::BinnedGabrielSymmetric::Settings& s,
                             tecs_pair_finder_settings* o) {
    for (int k = 0; k < 3; ++k) {
      o->min_bounds[k] = s.min_bounds[k];
      o->max_bounds[k] = s.max_bounds[k];
    }
    o->bin_size_scale = s.bin_size_scale;
    o->rebuild_every_n = (int32_t)s.rebuild_every_n;
    o->gabriel_coefficient = s.gabriel_coefficient;
  }
  static void apply_settings(const pair_finders::NaiveGabriel::Settings& s,
                             tecs_pair_finder_settings* o) {
    o->gabriel_coefficient = s.gabriel_coefficient;
  }
  static void apply_settings(const pair_finders::NaiveAllPairs::Settings&, tecs_pair_finder_settings*) {}

  tecs_sim* h_ = nullptr;
  tecs_force_desc force_{};
  unsigned n_ = 0;
};

}  // namespace tecs_amd

#endif /* TECS_AMD_HPP */
