// tecs_oracle.cpp - CPU restatement of the TECS per-step hot path.
//
// TEST INFRASTRUCTURE ONLY. This file is the parity oracle and the CPU baseline: only the
// tests, the smoke check and the benchmark's CPU-baseline leg load it. The product never
// links, calls or falls back to it.
//
// It restates, operation by operation, the reference path:
//   include/utils/grid.hpp                    Grid::rebuild + bin math           (Grid below)
//   Kokkos::BinSort (Serial backend order)    stable counting sort               (Grid::rebuild)
//   include/pair_finders/binned_all_pairs.hpp BinnedAllPairs::evaluate_force      (eval_binned_all_pairs)
//   include/pair_finders/naive_all_pairs.hpp  NaiveAllPairs::evaluate_force       (eval_naive_all_pairs)
//   include/pair_finders/binned_gabriel.hpp   BinnedGabriel::evaluate_force       (eval_binned_gabriel)
//   include/pair_finders/naive_gabriel.hpp    NaiveGabriel::evaluate_force        (eval_naive_gabriel)
//   include/pair_finders/naive_delaunay.hpp   NaiveDelaunay, 3-D                  (naive_delaunay_cell)
//   include/pair_finders/*_symmetric.hpp      the four symmetric finders          (the same walks
//                                             with j < i, then evaluate_symmetric)
//   include/forces/predefined.hpp:107-144     PiecewiseLinear                     (piecewise_linear)
//   include/types/polarity.hpp:111-198        Polarity_ math                      (pol_*)
//   include/types/vector.hpp:312-343          dot / length_squared / distance_to_squared
//   include/integrators/euler.hpp:49-85       Euler                               (step_euler)
//   include/integrators/heun.hpp:143-208      Heun                                (step_heun)
//   include/com_fixers/com_fixers.hpp:11-45   No/GlobalComFixer                   (com_fix)
//
// Every floating-point expression keeps the reference's operand order and is compiled with
// -ffp-contract=off (no fused multiply-add), so results are those of IEEE binary32 evaluation
// of the reference expressions. Pairs of a cell are visited in the reference's serial order
// (neighbour bins x-fastest, members ascending in agent index, which is the order Kokkos
// BinSort produces on the Serial backend); cells are processed in parallel with OpenMP, one
// owner per cell, so the results do not depend on the thread count.
//
// The Kokkos team reductions of the Gabriel and naive finders and the parallel_reduce of the
// GlobalComFixer have a backend-defined summation order; this restatement sums sequentially
// (parity of those sums is to a tolerance, not bitwise). The symmetric finders accumulate
// with atomics; here they apply the pair results in the Serial backend's order (cells i
// ascending, each cell's pairs in visiting order, delta(i) += v then delta(j) -= v).
//
// A double-precision evaluator (orc_forces_f64) recomputes the forces of a given pair set in
// FP64 for the force-accuracy checks.

#include "../include/tecs_amd.h"

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// ------------------------------------------------------------------ vector helpers (vector.hpp)
struct V3 {
  float c[3];
};
struct P2 {
  float c[2];
};

// VectorN::dot / length_squared (vector.hpp:312-324): result{} then += a_i*b_i in index order.
inline float dot3(const float* a, const float* b) {
  float r = 0.0f;
  for (int i = 0; i < 3; ++i) r += a[i] * b[i];
  return r;
}
// VectorN::distance_to_squared (vector.hpp:335-343).
inline float dist2(const float* a, const float* b) {
  float r = 0.0f;
  for (int i = 0; i < 3; ++i) {
    const float d = a[i] - b[i];
    r += d * d;
  }
  return r;
}

// --------------------------------------------------------------- force laws (predefined.hpp)
// PiecewiseLinearPotential / PiecewiseLinear, predefined.hpp:107-144.
inline void piecewise_linear(const float* disp, float d, float h, float krep, float kadh,
                             float* out) {
  if (d == 0.0f) {
    out[0] = out[1] = out[2] = 0.0f;
    return;
  }
  const float F = krep * std::fmax(h - d, 0.0f) - kadh * std::fmax(d - h, 0.0f);
  for (int c = 0; c < 3; ++c) out[c] = disp[c] * F / d;
}

// ---------------------------------------------------------------- polarity (polarity.hpp)
// to_vector3, polarity.hpp:111-118.
inline void pol_to_vec3(const float* p, float* v) {
  v[0] = std::sin(p[0]) * std::cos(p[1]);
  v[1] = std::sin(p[0]) * std::sin(p[1]);
  v[2] = std::cos(p[0]);
}
// Polarity_(vector, distance), polarity.hpp:62-66.
inline void pol_from_vec3(const float* v, float d, float* p) {
  p[0] = std::acos(std::fmin(1.0f, std::fmax(-1.0f, v[2] / d)));
  p[1] = std::atan2(v[1], v[0]);
}
// dot, polarity.hpp:131-136.
inline float pol_dot(const float* a, const float* b) {
  return std::sin(a[0]) * std::sin(b[0]) * std::cos(a[1] - b[1]) + std::cos(a[0]) * std::cos(b[0]);
}
// unidirectional_polarization_force, polarity.hpp:145-157.
inline void pol_unidirectional(const float* self, const float* other, float* r) {
  r[0] = std::cos(self[0]) * std::sin(other[0]) * std::cos(self[1] - other[1]) -
         std::sin(self[0]) * std::cos(other[0]);
  r[1] = 0.0f;
  const float sin_theta = std::sin(self[0]);
  if (std::fabs(sin_theta) > FLT_EPSILON)
    r[1] = -std::sin(other[0]) * std::sin(self[1] - other[1]) / sin_theta;
}
// bidirectional_polarization_force, polarity.hpp:165-168.
inline void pol_bidirectional(const float* self, const float* other, float* r) {
  float u[2];
  pol_unidirectional(self, other, u);
  const float s = pol_dot(self, other);
  r[0] = s * u[0];
  r[1] = s * u[1];
}
// Kokkos::pow(float, int) promotes to double; the quotient is rounded to float when it
// multiplies the float vector (operator*(Scalar, VectorN), vector.hpp:293-296).
inline float pow2_ratio(float a, float b) {
  return (float)(std::pow((double)a, 2.0) / std::pow((double)b, 2.0));
}
// bending_force, polarity.hpp:181-198.
inline void pol_bending(const float* self, const float* disp, const float* other, float d,
                        float* vec, float* pol) {
  float tv[3], ov[3];
  pol_to_vec3(self, tv);
  const float prod_i = dot3(tv, disp) / d;
  pol_to_vec3(other, ov);
  const float prod_j = dot3(ov, disp) / d;
  const float a = -prod_i / d;
  const float b = pow2_ratio(prod_i, d);
  const float c = -prod_j / d;
  const float e = pow2_ratio(prod_j, d);
  for (int k = 0; k < 3; ++k) vec[k] = ((a * tv[k] + b * disp[k]) + c * ov[k]) + e * disp[k];
  float q[2], u[2];
  pol_from_vec3(disp, d, q);
  pol_unidirectional(self, q, u);
  pol[0] = -prod_i * u[0];
  pol[1] = -prod_i * u[1];
}

// ------------------------------------------------------------------------ the fused force
struct Force {
  float h, krep, kadh;
  int n_types;
  float table[TECS_MAX_TYPES * TECS_MAX_TYPES];
  float bend, align;
  int order[3];  // TECS_TERM_* in the order the fused forces run, TECS_TERM_END after the last
  float drag;
  int drag_by_type;
  float drag_table[TECS_MAX_TYPES * TECS_MAX_TYPES];
};

// The force of a descriptor: the terms in term_order (all zeros = PiecewiseLinear, bending,
// alignment), as take_step(dt, f1, f2, ...) fuses f1, f2, ... in argument order
// (kernel_fuser.hpp:40-55).
void load_force(Force& o, const tecs_force_desc& f) {
  o.h = f.homeostatic_radius;
  o.krep = f.repulsion_scale;
  o.kadh = f.adhesion_scale;
  o.n_types = f.n_types;
  std::memcpy(o.table, f.adhesion_table, sizeof(o.table));
  o.bend = f.bending_strength;
  o.align = f.alignment_strength;
  const bool dflt = f.term_order[0] == 0 && f.term_order[1] == 0 && f.term_order[2] == 0;
  for (int k = 0; k < 3; ++k) o.order[k] = dflt ? k + 1 : f.term_order[k];
  o.drag = f.drag;
  o.drag_by_type = f.drag_by_type;
  std::memcpy(o.drag_table, f.drag_table, sizeof(o.drag_table));
}

struct Acc {
  float pos[3];
  float pol[2];
};

// One invocation of the fused pairwise functor: the forces in argument order, each adding its
// terms to the pack (kernel_fuser.hpp:40-55); the config forces are listed in tecs_amd.h.
// Returns the pair's drag (each force adds to the zeroed pairwise drag,
// binned_all_pairs.hpp:155-166; here one drag value per pair).
inline float apply_force(const Force& f, const float* disp, float d, const float* pol_i,
                         const float* pol_j, int type_i, int type_j, Acc& acc) {
  for (int k = 0; k < 3 && f.order[k] != TECS_TERM_END; ++k) {
    const int term = f.order[k];
    if (term == TECS_TERM_PIECEWISE_LINEAR) {
      float kadh = f.kadh;
      if (f.n_types > 0) kadh = f.table[type_i * f.n_types + type_j];
      float v[3];
      piecewise_linear(disp, d, f.h, f.krep, kadh, v);
      for (int c = 0; c < 3; ++c) acc.pos[c] += v[c];
    } else if (term == TECS_TERM_BENDING && f.bend != 0.0f) {
      float bv[3], bp[2];
      pol_bending(pol_i, disp, pol_j, d, bv, bp);
      for (int c = 0; c < 3; ++c) acc.pos[c] += f.bend * bv[c];
      for (int c = 0; c < 2; ++c) acc.pol[c] += f.bend * bp[c];
    } else if (term == TECS_TERM_ALIGNMENT && f.align != 0.0f) {
      float ap[2];
      pol_bidirectional(pol_i, pol_j, ap);
      for (int c = 0; c < 2; ++c) acc.pol[c] += f.align * ap[c];
    }
  }
  float drag = 0.0f;
  drag += f.drag_by_type ? f.drag_table[type_i * f.n_types + type_j] : f.drag;
  return drag;
}

// ------------------------------------------------------------------------------- the grid
// acceleration::Grid, grid.hpp:29-401, with Kokkos::BinSort's Serial-backend order.
struct Grid {
  float minb[3] = {0, 0, 0}, maxb[3] = {0, 0, 0};
  float bin_size = 0.0f;
  int ext[3] = {0, 0, 0};
  int bin_count = 0;
  bool built = false;
  std::vector<int> particle_bins, perm, offsets;

  // grid.hpp:71-85 (constructor), :200-215 (extents, count).
  void init(const float* mn, const float* mx, float bs) {
    for (int d = 0; d < 3; ++d) {
      minb[d] = mn[d];
      maxb[d] = mx[d];
    }
    bin_size = bs;
    for (int d = 0; d < 3; ++d) ext[d] = (int)std::ceil((maxb[d] - minb[d]) / bin_size);
    int r = 1;
    for (int d = 0; d < 3; ++d) r *= ext[d];
    bin_count = r;
  }
  // calc_bin_coords_from_point, grid.hpp:222-230 (division, truncation toward zero, clamp).
  void coords(const float* p, int* c) const {
    for (int d = 0; d < 3; ++d)
      c[d] = std::max(0, std::min((int)((p[d] - minb[d]) / bin_size), ext[d] - 1));
  }
  // flatten_bin_index, grid.hpp:243-245.
  int flatten(const int* c) const { return c[0] + c[1] * ext[0] + c[2] * (ext[0] * ext[1]); }
  // rebuild, grid.hpp:389-401: keys, then BinSort(keys, 0, n, BinOp1D(bin_count, 0, bin_count))
  // whose BinOp1D has bin_count + 1 bins; counting sort, ascending index within a bin.
  void rebuild(const std::vector<float>& pos, unsigned n) {
    particle_bins.assign(n, 0);
    for (unsigned i = 0; i < n; ++i) {
      int c[3];
      coords(&pos[3 * (size_t)i], c);
      particle_bins[i] = flatten(c);
    }
    offsets.assign((size_t)bin_count + 2, 0);
    for (unsigned i = 0; i < n; ++i) offsets[(size_t)particle_bins[i] + 1]++;
    for (size_t b = 0; b < (size_t)bin_count + 1; ++b) offsets[b + 1] += offsets[b];
    perm.assign(n, 0);
    std::vector<int> fill(offsets.begin(), offsets.end() - 1);
    for (unsigned i = 0; i < n; ++i) perm[(size_t)fill[particle_bins[i]]++] = (int)i;
    built = true;
  }
};

struct Sim {
  tecs_sim_desc desc;
  unsigned n = 0, cap = 0;
  bool has_pol = false, has_type = false;
  Force force;
  // pair finder state (binned_all_pairs.hpp:27-54)
  float cutoff = 0, cutoff2 = 0, bin_size = 0, inv_bin = 0, gabriel_factor = 0, gabriel_c2 = 0;
  int search_radius = 0;
  int step_count = 0;
  Grid grid;
  // stage 0 (the simulation storage), aux stages, old velocities (base.hpp:58-59)
  std::vector<float> pos, pol;
  std::vector<uint8_t> type;
  std::vector<float> d1pos, d1pol, x2pos, x2pol, d3pos, d3pol, old_vel;
  // generic wall forces (normalised plane a, b, c, d and strength) and link forces
  std::vector<std::array<float, 5>> walls;
  std::vector<int32_t> links;  // a0, b0, a1, b1, ...
  float link_strength = 0.0f;
  tecs_stats stats{};
};

// Per-cell accumulation context shared by the finders.
struct Ctx {
  const Sim* s;
  const float* pos;   // stage positions (in_view_pack.first())
  const float* pol;   // stage polarities or nullptr
  const uint8_t* type;
  float* out_pos;     // out_view_pack
  float* out_pol;
  const float* old_vel;
  std::vector<std::vector<int>>* emit;  // pair export (nullptr = force evaluation)
  uint64_t* pair_count;
};

inline void visit_pair(const Ctx& x, unsigned i, unsigned j, const float* disp, float d2,
                       Acc& acc, float& total_drag, float* total_vel, uint64_t& local_pairs,
                       std::vector<int>* row) {
  const float d = std::sqrt(d2);
  local_pairs++;
  if (row) {
    row->push_back((int)j);
    return;
  }
  static const float zero2[2] = {0.0f, 0.0f};
  const float* pi = x.pol ? &x.pol[2 * (size_t)i] : zero2;
  const float* pj = x.pol ? &x.pol[2 * (size_t)j] : zero2;
  const int ti = x.type ? x.type[i] : 0;
  const int tj = x.type ? x.type[j] : 0;
  const float pairwise_drag = apply_force(x.s->force, disp, d, pi, pj, ti, tj, acc);
  // binned_all_pairs.hpp:168-169
  total_drag += pairwise_drag;
  for (int c = 0; c < 3; ++c) total_vel[c] += pairwise_drag * x.old_vel[3 * (size_t)j + c];
}

// Commit, binned_all_pairs.hpp:174-180: out(i) += total; out.pos(i) += vel / fmax(drag, eps).
inline void commit(const Ctx& x, unsigned i, const Acc& acc, float total_drag,
                   const float* total_vel) {
  for (int c = 0; c < 3; ++c) x.out_pos[3 * (size_t)i + c] += acc.pos[c];
  if (x.out_pol)
    for (int c = 0; c < 2; ++c) x.out_pol[2 * (size_t)i + c] += acc.pol[c];
  const float den = std::fmax(total_drag, FLT_EPSILON);
  for (int c = 0; c < 3; ++c) x.out_pos[3 * (size_t)i + c] += total_vel[c] / den;
}

// BinnedAllPairs::evaluate_force, binned_all_pairs.hpp:86-182 (per cell i). With `sym`:
// BinnedAllPairsSymmetric, binned_all_pairs_symmetric.hpp:151-201 (i's bin by
// calc_bin_coords_from_point, partners j < i only).
void binned_all_pairs_cell(const Ctx& x, unsigned i, uint64_t& local_pairs,
                           std::vector<int>* row, bool sym) {
  const Sim& s = *x.s;
  const Grid& g = s.grid;
  const float* p_i = &x.pos[3 * (size_t)i];
  const int side = 2 * s.search_radius + 1;
  const int task_count = side * side * side;
  int bc[3];
  if (sym)
    g.coords(p_i, bc);  // binned_all_pairs_symmetric.hpp:154
  else
    for (int d = 0; d < 3; ++d)
      bc[d] = std::max(0, std::min((int)((p_i[d] - g.minb[d]) * s.inv_bin), g.ext[d] - 1));
  Acc acc{};
  float total_drag = 0.0f;
  float total_vel[3] = {0.0f, 0.0f, 0.0f};
  for (int t = 0; t < task_count; ++t) {
    int temp = t;
    int nb[3];
    for (int d = 0; d < 3; ++d) {
      nb[d] = bc[d] + (temp % side) - s.search_radius;
      temp /= side;
    }
    bool out = false;
    for (int d = 0; d < 3 && !out; ++d) out = (nb[d] < 0 || nb[d] >= g.ext[d]);
    if (out) continue;
    const int bi = nb[0] + nb[1] * g.ext[0] + nb[2] * (g.ext[0] * g.ext[1]);
    const int start = g.offsets[bi], end = g.offsets[(size_t)bi + 1];
    for (int idx = start; idx < end; ++idx) {
      const unsigned j = (unsigned)g.perm[idx];
      if (sym ? j >= i : j == i) continue;
      const float* p_j = &x.pos[3 * (size_t)j];
      float disp[3];
      for (int c = 0; c < 3; ++c) disp[c] = p_i[c] - p_j[c];
      const float d2 = dot3(disp, disp);
      if (d2 >= s.cutoff2) continue;
      visit_pair(x, i, j, disp, d2, acc, total_drag, total_vel, local_pairs, row);
    }
  }
  if (!row) commit(x, i, acc, total_drag, total_vel);
}

// NaiveAllPairs::evaluate_force, naive_all_pairs.hpp:41-103 (sequential j order). With `sym`:
// NaiveAllPairsSymmetric, naive_all_pairs_symmetric.hpp:100-143 (j < i).
void naive_all_pairs_cell(const Ctx& x, unsigned i, uint64_t& local_pairs,
                          std::vector<int>* row, bool sym) {
  const Sim& s = *x.s;
  const float* p_i = &x.pos[3 * (size_t)i];
  Acc acc{};
  float total_drag = 0.0f;
  float total_vel[3] = {0.0f, 0.0f, 0.0f};
  for (unsigned j = 0; j < (sym ? i : s.n); ++j) {
    if (i == j) continue;
    const float* p_j = &x.pos[3 * (size_t)j];
    float disp[3];
    for (int c = 0; c < 3; ++c) disp[c] = p_i[c] - p_j[c];
    const float d2 = dot3(disp, disp);
    if (d2 >= s.cutoff2) continue;
    visit_pair(x, i, j, disp, d2, acc, total_drag, total_vel, local_pairs, row);
  }
  if (!row) commit(x, i, acc, total_drag, total_vel);
}

// BinnedGabriel::evaluate_force, binned_gabriel.hpp:95-361 (tasks and members in order). With
// `sym`: BinnedGabrielSymmetric, binned_gabriel_symmetric.hpp:159-287 (partners j < i; the
// blockers k are all members but i and j, as before).
void binned_gabriel_cell(const Ctx& x, unsigned i, uint64_t& local_pairs,
                         std::vector<int>* row, bool sym) {
  const Sim& s = *x.s;
  const Grid& g = s.grid;
  const float* p_i = &x.pos[3 * (size_t)i];
  const int side = 2 * s.search_radius + 1;
  const int task_count = side * side * side;
  int bc[3];
  g.coords(p_i, bc);  // binned_gabriel.hpp:101
  Acc acc{};
  float total_drag = 0.0f;
  float total_vel[3] = {0.0f, 0.0f, 0.0f};
  for (int t = 0; t < task_count; ++t) {
    // linear_index_to_offset, grid.hpp:317-322
    const int temp = t / side;
    int ni[3] = {bc[0] + (t % side) - s.search_radius, bc[1] + (temp % side) - s.search_radius,
                 bc[2] + (temp / side) - s.search_radius};
    if (ni[0] < 0 || ni[0] >= g.ext[0] || ni[1] < 0 || ni[1] >= g.ext[1] || ni[2] < 0 ||
        ni[2] >= g.ext[2])
      continue;
    const int b = g.flatten(ni);
    const int start = g.offsets[b], end = g.offsets[(size_t)b + 1];
    for (int idx = start; idx < end; ++idx) {
      const unsigned j = (unsigned)g.perm[idx];
      if (sym ? j >= i : j == i) continue;
      const float* p_j = &x.pos[3 * (size_t)j];
      float disp[3];
      for (int c = 0; c < 3; ++c) disp[c] = p_i[c] - p_j[c];
      const float d2 = dot3(disp, disp);
      if (d2 >= s.cutoff2) continue;
      float mid[3];
      for (int c = 0; c < 3; ++c) mid[c] = p_i[c] - disp[c] * 0.5f;
      const float rho2 = d2 * s.gabriel_factor;
      const float rho = std::sqrt(rho2);
      int min_bin[3], span[3];
      int task_count_2 = 1;
      for (int d = 0; d < 3; ++d) {
        min_bin[d] = std::max(
            0, std::min((int)std::floor((mid[d] - rho - g.minb[d]) * s.inv_bin), g.ext[d] - 1));
        const int max_b = std::max(
            0, std::min((int)std::floor((mid[d] + rho - g.minb[d]) * s.inv_bin), g.ext[d] - 1));
        span[d] = max_b - min_bin[d] + 1;
        task_count_2 *= span[d];
      }
      bool blocked = false;
      for (int t2 = 0; t2 < task_count_2 && !blocked; ++t2) {
        const int tmp = t2 / span[0];
        const int b0 = min_bin[0] + (t2 % span[0]);
        const int b1 = min_bin[1] + (tmp % span[1]);
        const int b2 = min_bin[2] + (tmp / span[1]);
        const int bb = b0 + b1 * g.ext[0] + b2 * (g.ext[0] * g.ext[1]);
        const int s2 = g.offsets[bb], e2 = g.offsets[(size_t)bb + 1];
        for (int idx2 = s2; idx2 < e2; ++idx2) {
          const unsigned k = (unsigned)g.perm[idx2];
          if (k == i || k == j) continue;
          if (dist2(&x.pos[3 * (size_t)k], mid) < rho2) {
            blocked = true;
            break;
          }
        }
      }
      if (blocked) continue;
      visit_pair(x, i, j, disp, d2, acc, total_drag, total_vel, local_pairs, row);
    }
  }
  if (!row) commit(x, i, acc, total_drag, total_vel);
}

// NaiveGabriel::evaluate_force, naive_gabriel.hpp:47-119. With `sym`: NaiveGabrielSymmetric,
// naive_gabriel_symmetric.hpp:107-168 (j < i; rho^2 = (d2 * 0.25) * c^2 as before).
void naive_gabriel_cell(const Ctx& x, unsigned i, uint64_t& local_pairs,
                        std::vector<int>* row, bool sym) {
  const Sim& s = *x.s;
  const float* p_i = &x.pos[3 * (size_t)i];
  Acc acc{};
  float total_drag = 0.0f;
  float total_vel[3] = {0.0f, 0.0f, 0.0f};
  for (unsigned j = 0; j < (sym ? i : s.n); ++j) {
    if (i == j) continue;
    const float* p_j = &x.pos[3 * (size_t)j];
    float disp[3];
    for (int c = 0; c < 3; ++c) disp[c] = p_i[c] - p_j[c];
    const float d2 = dot3(disp, disp);
    if (d2 >= s.cutoff2) continue;
    float mid[3];
    for (int c = 0; c < 3; ++c) mid[c] = p_i[c] - disp[c] * 0.5f;
    const float rho2 = d2 * 0.25f * s.gabriel_c2;
    bool blocked = false;
    for (unsigned k = 0; k < s.n && !blocked; ++k) {
      if (k == i || k == j) continue;
      float off[3];
      for (int c = 0; c < 3; ++c) off[c] = x.pos[3 * (size_t)k + c] - mid[c];
      if (dot3(off, off) < rho2) blocked = true;
    }
    if (blocked) continue;
    visit_pair(x, i, j, disp, d2, acc, total_drag, total_vel, local_pairs, row);
  }
  if (!row) commit(x, i, acc, total_drag, total_vel);
}

// NaiveDelaunay::evaluate_force, 3-D branch (naive_delaunay.hpp:157-293), the work items
// (i, j, k) of cell i in index order: j over the other cells, k over the cells other than i
// and j (k shifted past min(i, j), then past max(i, j); a k that lands on agent_count does
// nothing). An item calls the functor for (i, j) when d(i, j) < r and some l > k, l not in
// {i, j, k}, spans a sphere that no other cell lies in or on (a box prefilter, then
// d^2 <= r^2). Each item stops at its first such l. The sphere is the reference's: Cramer's
// rule on the edges e = p - p_i with right-hand sides |p|^2 - |p_i|^2 (|denominator| <= eps
// skips l) gives the circumcentre C itself, to which p_i is then added, and the radius^2 is
// |C|^2; for p_i != 0 this is not the circumsphere and the pairs are not the Delaunay graph.
void naive_delaunay_cell(const Ctx& x, unsigned i, uint64_t& local_pairs, std::vector<int>* row) {
  const Sim& s = *x.s;
  const unsigned n = s.n;
  const float* P = x.pos;
  auto pt = [&](unsigned a) { return &P[3 * (size_t)a]; };
  auto len2v = [](const float* v) { return dot3(v, v); };
  Acc acc{};
  float total_drag = 0.0f;
  float total_vel[3] = {0.0f, 0.0f, 0.0f};
  if (n < 2) {
    if (!row) commit(x, i, acc, total_drag, total_vel);
    return;
  }
  for (unsigned jj = 0; jj + 1 < n; ++jj) {
    const unsigned j = jj + (jj >= i ? 1u : 0u);
    const float* pi = pt(i);
    const float* pj = pt(j);
    float disp[3];
    for (int c = 0; c < 3; ++c) disp[c] = pi[c] - pj[c];
    const float d2 = dot3(disp, disp);
    if (d2 >= s.cutoff2) continue;
    const unsigned lo = i < j ? i : j, hi = i < j ? j : i;
    for (unsigned kk = 0; kk < n; ++kk) {
      unsigned k = kk;
      k += k >= lo ? 1u : 0u;
      k += k >= hi ? 1u : 0u;
      if (k >= n) continue;
      const float* pk = pt(k);
      bool found = false;
      for (unsigned l = 0; l < n && !found; ++l) {
        if (l == i || l == j || l == k || l <= k) continue;
        const float* pl = pt(l);
        float e1[3], e2[3], e3[3];
        for (int c = 0; c < 3; ++c) {
          e1[c] = pj[c] - pi[c];
          e2[c] = pk[c] - pi[c];
          e3[c] = pl[c] - pi[c];
        }
        const float den = 2.0f * (e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) -
                                  e1[1] * (e2[0] * e3[2] - e2[2] * e3[0]) +
                                  e1[2] * (e2[0] * e3[1] - e2[1] * e3[0]));
        if (std::fabs(den) <= FLT_EPSILON) continue;
        const float b1 = len2v(pj) - len2v(pi);
        const float b2 = len2v(pk) - len2v(pi);
        const float b3 = len2v(pl) - len2v(pi);
        const float cc[3] = {
            pi[0] + (b1 * (e2[1] * e3[2] - e2[2] * e3[1]) - e1[1] * (b2 * e3[2] - e2[2] * b3) +
                     e1[2] * (b2 * e3[1] - e2[1] * b3)) / den,
            pi[1] + (e1[0] * (b2 * e3[2] - e2[2] * b3) - b1 * (e2[0] * e3[2] - e2[2] * e3[0]) +
                     e1[2] * (e2[0] * b3 - b2 * e3[0])) / den,
            pi[2] + (e1[0] * (e2[1] * b3 - b2 * e3[1]) - e1[1] * (e2[0] * b3 - b2 * e3[0]) +
                     b1 * (e2[0] * e3[1] - e2[1] * e3[0])) / den};
        const float rad2 = dist2(pi, cc);
        const float rad = std::sqrt(rad2);
        bool valid = true;
        for (unsigned m = 0; m < n && valid; ++m) {
          if (m == i || m == j || m == k || m == l) continue;
          const float* pm = pt(m);
          if (std::fabs(pm[0] - cc[0]) > rad || std::fabs(pm[1] - cc[1]) > rad ||
              std::fabs(pm[2] - cc[2]) > rad)
            continue;
          if (dist2(pm, cc) <= rad2) valid = false;
        }
        found = valid;
      }
      if (!found) continue;
      // each work item evaluates the functor on a zeroed pack and adds that pack to out(i)
      // (naive_delaunay.hpp:258-275); the items' adds are atomics, here in (j, k) order
      Acc v{};
      visit_pair(x, i, j, disp, d2, v, total_drag, total_vel, local_pairs, row);
      for (int c = 0; c < 3; ++c) acc.pos[c] += v.pos[c];
      for (int c = 0; c < 2; ++c) acc.pol[c] += v.pol[c];
    }
  }
  if (!row) commit(x, i, acc, total_drag, total_vel);
}

bool finder_is_binned(const Sim& s) {
  const int f = s.desc.pair_finder;
  return f == TECS_BINNED_ALL_PAIRS || f == TECS_BINNED_GABRIEL ||
         f == TECS_BINNED_ALL_PAIRS_SYMMETRIC || f == TECS_BINNED_GABRIEL_SYMMETRIC;
}
bool finder_is_symmetric(const Sim& s) {
  return s.desc.pair_finder >= TECS_NAIVE_ALL_PAIRS_SYMMETRIC &&
         s.desc.pair_finder <= TECS_BINNED_GABRIEL_SYMMETRIC;
}

// The rebuild schedule of the binned finders, binned_all_pairs.hpp:77-84.
void maybe_rebuild(Sim& s, const float* pos, bool is_full_step) {
  if (!finder_is_binned(s) || !is_full_step) return;
  if (s.step_count == 0 || s.step_count >= s.desc.pair_finder_settings.rebuild_every_n) {
    std::vector<float> p(pos, pos + 3 * (size_t)s.n);
    s.grid.rebuild(p, s.n);
    s.stats.rebuilds++;
    s.step_count = 0;
  }
  s.step_count++;
}

// The force part of the symmetric finders (binned_all_pairs_symmetric.hpp:176-218 and the
// same lines of the other three): for every pair (i, j < i) the functor runs once on a zeroed
// accumulator pack; atomic_add(out(i), v) and atomic_add(out(j), -v) for every field, then the
// drag pass adds total_velocity / fmax(total_drag, eps), which is +0 because no force term
// writes drag. Pair results are computed in parallel and applied in the Serial order.
void apply_symmetric(const Ctx& x, const std::vector<std::vector<int>>& rows) {
  const Sim& s = *x.s;
  const long n = (long)s.n;
  std::vector<int64_t> rp((size_t)n + 1, 0);
  for (long i = 0; i < n; ++i) rp[i + 1] = rp[i] + (int64_t)rows[i].size();
  std::vector<Acc> vals((size_t)rp[n]);
  static const float zero2[2] = {0.0f, 0.0f};
#pragma omp parallel for schedule(dynamic, 256)
  for (long i = 0; i < n; ++i) {
    const float* p_i = &x.pos[3 * (size_t)i];
    for (size_t e = 0; e < rows[i].size(); ++e) {
      const int j = rows[i][e];
      const float* p_j = &x.pos[3 * (size_t)j];
      float disp[3];
      for (int c = 0; c < 3; ++c) disp[c] = p_i[c] - p_j[c];
      const float d = std::sqrt(dot3(disp, disp));
      Acc v{};
      apply_force(s.force, disp, d, x.pol ? &x.pol[2 * (size_t)i] : zero2,
                  x.pol ? &x.pol[2 * (size_t)j] : zero2, x.type ? x.type[i] : 0,
                  x.type ? x.type[j] : 0, v);
      vals[(size_t)rp[i] + e] = v;
    }
  }
  for (long i = 0; i < n; ++i) {
    for (size_t e = 0; e < rows[i].size(); ++e) {
      const size_t j = (size_t)rows[i][e];
      const Acc& v = vals[(size_t)rp[i] + e];
      for (int c = 0; c < 3; ++c) x.out_pos[3 * (size_t)i + c] += v.pos[c];
      for (int c = 0; c < 3; ++c) x.out_pos[3 * j + c] += -v.pos[c];
      if (x.out_pol) {
        for (int c = 0; c < 2; ++c) x.out_pol[2 * (size_t)i + c] += v.pol[c];
        for (int c = 0; c < 2; ++c) x.out_pol[2 * j + c] += -v.pol[c];
      }
    }
  }
  for (long i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) x.out_pos[3 * (size_t)i + c] += 0.0f;
}

// PairFinder::evaluate_force for any finder: in = (pos, pol), out += deltas.
int evaluate(Sim& s, const float* pos, const float* pol, float* out_pos, float* out_pol,
             bool is_full_step, std::vector<std::vector<int>>* emit) {
  maybe_rebuild(s, pos, is_full_step);
  if (finder_is_binned(s) && !s.grid.built)
    return fail(TECS_ERR_STATE, "binned finder evaluated before any full step built a grid");
  Ctx x{&s, pos, s.has_pol ? pol : nullptr, s.has_type ? s.type.data() : nullptr,
        out_pos, s.has_pol ? out_pol : nullptr, s.old_vel.data(), emit, nullptr};
  const bool sym = finder_is_symmetric(s);
  // a symmetric force evaluation first collects each cell's pairs, then applies them
  std::vector<std::vector<int>> sym_rows;
  std::vector<std::vector<int>>* rows = emit;
  if (sym && !emit) {
    sym_rows.assign(s.n, {});
    rows = &sym_rows;
    x.emit = rows;
  }
  uint64_t total_pairs = 0;
  const long n = (long)s.n;
  const int finder = sym ? s.desc.pair_finder - TECS_NAIVE_ALL_PAIRS_SYMMETRIC : s.desc.pair_finder;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : total_pairs)
  for (long ii = 0; ii < n; ++ii) {
    const unsigned i = (unsigned)ii;
    uint64_t lp = 0;
    std::vector<int>* row = rows ? &(*rows)[i] : nullptr;
    switch (finder) {
      case TECS_NAIVE_DELAUNAY: naive_delaunay_cell(x, i, lp, row); break;
      case TECS_BINNED_ALL_PAIRS: binned_all_pairs_cell(x, i, lp, row, sym); break;
      case TECS_NAIVE_ALL_PAIRS: naive_all_pairs_cell(x, i, lp, row, sym); break;
      case TECS_BINNED_GABRIEL: binned_gabriel_cell(x, i, lp, row, sym); break;
      default: naive_gabriel_cell(x, i, lp, row, sym); break;
    }
    total_pairs += lp;
  }
  if (sym && !emit) apply_symmetric(x, sym_rows);
  if (!emit) {
    s.stats.pair_interactions += total_pairs;
    s.stats.evaluations++;
  }
  return TECS_OK;
}

// ComFixer::fix, com_fixers.hpp:11-45. The sum runs over the view extent (the capacity) and
// divides by it; the reduction order of Kokkos is backend-defined, here it is sequential.
void com_fix(const Sim& s, const std::vector<float>& dpos, float* com) {
  com[0] = com[1] = com[2] = 0.0f;
  if (s.desc.com_fixer != TECS_GLOBAL_COM_FIXER) return;
  const size_t n = s.cap;
  if (n == 0) return;
  float t[3] = {0.0f, 0.0f, 0.0f};
  for (size_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) t[c] += dpos[3 * i + c];
  for (int c = 0; c < 3; ++c) com[c] = t[c] / (float)n;
}

// Euler::integrate + apply_euler, euler.hpp:49-85.
// The GENERIC_FORCE of examples/growth_w_wall.cpp:44-48 (and growth_w_squeeze.cpp:37-52) for
// every wall: sd = PlaneN::signed_distance(p) (types/plane.hpp:36-42: d + a x + b y + c z in
// that order); sd < 0 adds normal * (-sd * k) to the cell's delta, wall after wall.
void wall_forces(const Sim& s, const float* pos, float* out) {
  if (s.walls.empty()) return;
  const long n = (long)s.n;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    const float* p = &pos[3 * i];
    for (const auto& w : s.walls) {
      float sd = w[3];
      for (int c = 0; c < 3; ++c) sd += w[c] * p[c];
      if (sd < 0.0f)
        for (int c = 0; c < 3; ++c) out[3 * i + c] += w[c] * (-sd * w[4]);
    }
  }
}

// The LINK_FORCE of examples/intercalation.cpp and sorting_prot.cpp, applied as base.hpp:193-230
// applies link forces (a zeroed pack per endpoint, atomically added to its delta; here in link
// order): u = p_b - p_a, d = |u|, delta_a += (k * u) / d, delta_b += (-k * u) / d. Links with
// a == b are skipped (base.hpp:207-208), and so are links with an endpoint beyond the active
// count (their stage positions are not maintained).
void link_forces(const Sim& s, const float* pos, float* out) {
  const size_t nl = s.links.size() / 2;
  const float k = s.link_strength;
  for (size_t q = 0; q < nl; ++q) {
    const unsigned a = (unsigned)s.links[2 * q], b = (unsigned)s.links[2 * q + 1];
    if (a == b || a >= s.n || b >= s.n) continue;
    float u[3];
    for (int c = 0; c < 3; ++c) u[c] = pos[3 * (size_t)b + c] - pos[3 * (size_t)a + c];
    const float d = std::sqrt(dot3(u, u));
    for (int c = 0; c < 3; ++c) out[3 * (size_t)a + c] += 0.0f + (k * u[c]) / d;
    for (int c = 0; c < 3; ++c) out[3 * (size_t)b + c] += 0.0f + (-k * u[c]) / d;
  }
}

// Simulation::take_step's forces in the order fuse_forces groups them (kernel_fuser.hpp:
// 148-156): generic (walls), pairwise (the pair finder), link.
int step_forces(Sim& s, const float* pos, const float* pol, float* out_pos, float* out_pol,
                bool is_full_step) {
  wall_forces(s, pos, out_pos);
  const int rc = evaluate(s, pos, pol, out_pos, out_pol, is_full_step, nullptr);
  if (rc) return rc;
  link_forces(s, pos, out_pos);
  return TECS_OK;
}

int step_euler(Sim& s, double dt) {
  int rc = step_forces(s, s.pos.data(), s.pol.data(), s.d1pos.data(), s.d1pol.data(), true);
  if (rc) return rc;
  float com[3];
  com_fix(s, s.d1pos, com);
  const float dtf = (float)dt;  // VectorN * Scalar narrows the double dt (vector.hpp:263-267)
  const long n = (long)s.n;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    float* d = &s.d1pos[3 * i];
    for (int c = 0; c < 3; ++c) d[c] -= com[c];
    for (int c = 0; c < 3; ++c) s.old_vel[3 * i + c] = d[c];
    for (int c = 0; c < 3; ++c) {
      s.pos[3 * i + c] += d[c] * dtf;
      d[c] = 0.0f;
    }
    if (s.has_pol)
      for (int c = 0; c < 2; ++c) {
        s.pol[2 * i + c] += s.d1pol[2 * i + c] * dtf;
        s.d1pol[2 * i + c] = 0.0f;
      }
  }
  s.stats.steps++;
  return TECS_OK;
}

// Heun::integrate, heun.hpp:143-208.
int step_heun(Sim& s, double dt) {
  const float dtf = (float)dt;
  const long n = (long)s.n;
  int rc = step_forces(s, s.pos.data(), s.pol.data(), s.d1pos.data(), s.d1pol.data(), true);
  if (rc) return rc;
  float com[3];
  com_fix(s, s.d1pos, com);
  // apply_euler_predictor, heun.hpp:146-156
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    for (int c = 0; c < 3; ++c) s.d1pos[3 * i + c] -= com[c];
    for (int c = 0; c < 3; ++c) s.x2pos[3 * i + c] = s.pos[3 * i + c] + s.d1pos[3 * i + c] * dtf;
    if (s.has_pol)
      for (int c = 0; c < 2; ++c)
        s.x2pol[2 * i + c] = s.pol[2 * i + c] + s.d1pol[2 * i + c] * dtf;
  }
  rc = step_forces(s, s.x2pos.data(), s.x2pol.data(), s.d3pos.data(), s.d3pol.data(), false);
  if (rc) return rc;
  com_fix(s, s.d3pos, com);
  // apply_heun_corrector, heun.hpp:170-186
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    for (int c = 0; c < 3; ++c) s.d3pos[3 * i + c] -= com[c];
    for (int c = 0; c < 3; ++c)
      s.old_vel[3 * i + c] = (s.d1pos[3 * i + c] + s.d3pos[3 * i + c]) * 0.5f;
    for (int c = 0; c < 3; ++c) {
      s.pos[3 * i + c] += (s.d1pos[3 * i + c] + s.d3pos[3 * i + c]) * 0.5f * dtf;
      s.d1pos[3 * i + c] = 0.0f;
      s.d3pos[3 * i + c] = 0.0f;
    }
    if (s.has_pol)
      for (int c = 0; c < 2; ++c) {
        s.pol[2 * i + c] += (s.d1pol[2 * i + c] + s.d3pol[2 * i + c]) * 0.5f * dtf;
        s.d1pol[2 * i + c] = 0.0f;
        s.d3pol[2 * i + c] = 0.0f;
      }
  }
  s.stats.steps++;
  return TECS_OK;
}

int validate(const tecs_sim_desc* d) {
  if (!d) return fail(TECS_ERR_INVALID, "null descriptor");
  if (d->capacity < d->agent_count) return fail(TECS_ERR_INVALID, "capacity < agent_count");
  if (d->pair_finder < 0 || d->pair_finder > TECS_NAIVE_DELAUNAY)
    return fail(TECS_ERR_INVALID, "bad pair_finder");
  if (d->integrator != TECS_EULER && d->integrator != TECS_HEUN)
    return fail(TECS_ERR_INVALID, "bad integrator");
  if (d->com_fixer != TECS_NO_COM_FIXER && d->com_fixer != TECS_GLOBAL_COM_FIXER)
    return fail(TECS_ERR_INVALID, "bad com_fixer");
  const tecs_force_desc& f = d->force;
  if (f.n_types < 0 || f.n_types > TECS_MAX_TYPES) return fail(TECS_ERR_INVALID, "bad n_types");
  if (f.n_types > 0 && !(d->fields & TECS_FIELD_TYPE))
    return fail(TECS_ERR_INVALID, "type table needs TECS_FIELD_TYPE");
  if ((f.bending_strength != 0.0f || f.alignment_strength != 0.0f) &&
      !(d->fields & TECS_FIELD_POLARITY))
    return fail(TECS_ERR_INVALID, "polarity forces need TECS_FIELD_POLARITY");
  if (!(d->cutoff_distance > 0.0f)) return fail(TECS_ERR_INVALID, "cutoff must be > 0");
  for (int k = 0; k < 3; ++k)
    if (f.term_order[k] < TECS_TERM_END || f.term_order[k] > TECS_TERM_ALIGNMENT)
      return fail(TECS_ERR_INVALID, "bad term in term_order");
  if (f.drag_by_type && f.n_types == 0) return fail(TECS_ERR_INVALID, "drag by type needs a type table");
  bool drag = f.drag != 0.0f;
  for (int k = 0; f.drag_by_type && k < f.n_types * f.n_types; ++k) drag = drag || f.drag_table[k] != 0.0f;
  if (drag && (d->pair_finder >= TECS_NAIVE_ALL_PAIRS_SYMMETRIC && d->pair_finder <= TECS_BINNED_GABRIEL_SYMMETRIC))
    return fail(TECS_ERR_UNSUPPORTED, "drag with a symmetric pair finder is not restated");
  return TECS_OK;
}

bool term_on(const Force& f, int term) {
  for (int k = 0; k < 3 && f.order[k] != TECS_TERM_END; ++k)
    if (f.order[k] == term) return true;
  return false;
}

// Double-precision value of the fused force of one pair (i as self, j as other) at the current
// stage-0 state, and per component the sum of the absolute values of its terms (the terms the
// force applies; their order does not matter in this evaluation).
void pair_force_f64(const Sim& s, long i, long j, double* vpos, double* vpol, double* apos,
                    double* apol) {
  const Force& f = s.force;
  const bool use_linear = term_on(f, TECS_TERM_PIECEWISE_LINEAR);
  const bool use_bend = term_on(f, TECS_TERM_BENDING) && f.bend != 0.0f;
  const bool use_align = term_on(f, TECS_TERM_ALIGNMENT) && f.align != 0.0f;
  for (int c = 0; c < 3; ++c) vpos[c] = apos[c] = 0.0;
  for (int c = 0; c < 2; ++c) vpol[c] = apol[c] = 0.0;
  const double pi[3] = {s.pos[3 * i], s.pos[3 * i + 1], s.pos[3 * i + 2]};
  double disp[3];
  for (int c = 0; c < 3; ++c) disp[c] = pi[c] - (double)s.pos[3 * j + c];
  const double d = std::sqrt(disp[0] * disp[0] + disp[1] * disp[1] + disp[2] * disp[2]);
  double kadh = f.kadh;
  if (f.n_types > 0) kadh = f.table[s.type[i] * f.n_types + s.type[j]];
  if (d != 0.0 && use_linear) {
    const double F = (double)f.krep * std::fmax((double)f.h - d, 0.0) -
                     kadh * std::fmax(d - (double)f.h, 0.0);
    for (int c = 0; c < 3; ++c) {
      const double v = disp[c] * F / d;
      vpos[c] += v;
      apos[c] += std::fabs(v);
    }
  }
  if (!s.has_pol) return;
  const double li[2] = {s.pol[2 * i], s.pol[2 * i + 1]};
  const double lj[2] = {s.pol[2 * j], s.pol[2 * j + 1]};
  auto to_vec = [](const double* p, double* v) {
    v[0] = std::sin(p[0]) * std::cos(p[1]);
    v[1] = std::sin(p[0]) * std::sin(p[1]);
    v[2] = std::cos(p[0]);
  };
  auto uni = [](const double* a, const double* b, double* r) {
    r[0] = std::cos(a[0]) * std::sin(b[0]) * std::cos(a[1] - b[1]) - std::sin(a[0]) * std::cos(b[0]);
    r[1] = 0.0;
    const double st = std::sin(a[0]);
    // same guard as the FP32 path (polarity.hpp:154), evaluated on the FP32 sin
    if (std::fabs(std::sin((float)a[0])) > FLT_EPSILON)
      r[1] = -std::sin(b[0]) * std::sin(a[1] - b[1]) / st;
  };
  if (use_bend) {
    double tv[3], ov[3];
    to_vec(li, tv);
    to_vec(lj, ov);
    const double prod_i = (tv[0] * disp[0] + tv[1] * disp[1] + tv[2] * disp[2]) / d;
    const double prod_j = (ov[0] * disp[0] + ov[1] * disp[1] + ov[2] * disp[2]) / d;
    for (int c = 0; c < 3; ++c) {
      const double v = f.bend * (-prod_i / d * tv[c] + prod_i * prod_i / (d * d) * disp[c] -
                                 prod_j / d * ov[c] + prod_j * prod_j / (d * d) * disp[c]);
      vpos[c] += v;
      apos[c] += std::fabs(-prod_i / d * tv[c]) + std::fabs(prod_i * prod_i / (d * d) * disp[c]) +
                 std::fabs(prod_j / d * ov[c]) + std::fabs(prod_j * prod_j / (d * d) * disp[c]);
    }
    double q[2], u[2];
    q[0] = std::acos(std::fmin(1.0, std::fmax(-1.0, disp[2] / d)));
    q[1] = std::atan2(disp[1], disp[0]);
    uni(li, q, u);
    for (int c = 0; c < 2; ++c) {
      const double v = f.bend * (-prod_i * u[c]);
      vpol[c] += v;
      apol[c] += std::fabs(v);
    }
  }
  if (use_align) {
    double u[2];
    uni(li, lj, u);
    const double dd =
        std::sin(li[0]) * std::sin(lj[0]) * std::cos(li[1] - lj[1]) + std::cos(li[0]) * std::cos(lj[0]);
    for (int c = 0; c < 2; ++c) {
      const double v = f.align * dd * u[c];
      vpol[c] += v;
      apol[c] += std::fabs(v);
    }
  }
}

}  // namespace

// =============================================================================== C ABI
extern "C" {

typedef struct orc_sim orc_sim;

const char* orc_last_error(void) { return g_err.c_str(); }

int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

int orc_create(const tecs_sim_desc* desc, orc_sim** out) {
  int rc = validate(desc);
  if (rc) return rc;
  Sim* s = new Sim();
  s->desc = *desc;
  s->n = desc->agent_count;
  s->cap = desc->capacity;
  s->has_pol = (desc->fields & TECS_FIELD_POLARITY) != 0;
  s->has_type = (desc->fields & TECS_FIELD_TYPE) != 0;
  load_force(s->force, desc->force);
  // pair finder constructors, binned_all_pairs.hpp:27-39, binned_gabriel.hpp:28-42,
  // naive_gabriel.hpp:386-392
  const tecs_pair_finder_settings& ps = desc->pair_finder_settings;
  s->cutoff = desc->cutoff_distance;
  s->cutoff2 = s->cutoff * s->cutoff;
  s->bin_size = ps.bin_size_scale * s->cutoff;
  s->inv_bin = 1.0f / s->bin_size;
  s->search_radius = (int)std::ceil(s->cutoff * s->inv_bin);
  s->gabriel_factor = ps.gabriel_coefficient * ps.gabriel_coefficient * 0.25f;
  s->gabriel_c2 = ps.gabriel_coefficient * ps.gabriel_coefficient;
  if (finder_is_binned(*s)) s->grid.init(ps.min_bounds, ps.max_bounds, s->bin_size);
  const size_t c = s->cap;
  s->pos.assign(3 * c, 0.0f);
  s->pol.assign(2 * c, 0.0f);
  s->type.assign(c, 0);
  s->d1pos.assign(3 * c, 0.0f);
  s->d1pol.assign(2 * c, 0.0f);
  s->x2pos.assign(3 * c, 0.0f);
  s->x2pol.assign(2 * c, 0.0f);
  s->d3pos.assign(3 * c, 0.0f);
  s->d3pol.assign(2 * c, 0.0f);
  s->old_vel.assign(3 * c, 0.0f);
  *out = reinterpret_cast<orc_sim*>(s);
  return TECS_OK;
}

int orc_destroy(orc_sim* h) {
  delete reinterpret_cast<Sim*>(h);
  return TECS_OK;
}

int orc_upload(orc_sim* h, const float* pos, const float* pol, const uint8_t* type) {
  Sim& s = *reinterpret_cast<Sim*>(h);
  if (!pos) return fail(TECS_ERR_INVALID, "positions required");
  std::memcpy(s.pos.data(), pos, sizeof(float) * 3 * s.n);
  if (s.has_pol && pol) std::memcpy(s.pol.data(), pol, sizeof(float) * 2 * s.n);
  if (s.has_type && type) std::memcpy(s.type.data(), type, s.n);
  if (s.has_type && type && s.force.n_types > 0)
    for (unsigned i = 0; i < s.n; ++i)
      if (type[i] >= (unsigned)s.force.n_types) return fail(TECS_ERR_INVALID, "type out of range");
  return TECS_OK;
}

// Simulation::take_step(dt, forces...) takes its forces with every call: the force of the
// following steps.
int orc_set_force(orc_sim* h, const tecs_force_desc* f) {
  if (!h || !f) return fail(TECS_ERR_INVALID, "null argument");
  Sim& s = *reinterpret_cast<Sim*>(h);
  tecs_sim_desc d = s.desc;
  d.force = *f;
  const int rc = validate(&d);
  if (rc != TECS_OK) return rc;
  s.desc.force = *f;
  load_force(s.force, *f);
  return TECS_OK;
}

// Writes into the views at indices first .. first + count - 1 (active or not).
int orc_upload_range(orc_sim* h, unsigned first, unsigned count, const float* pos, const float* pol,
                     const uint8_t* type) {
  Sim& s = *reinterpret_cast<Sim*>(h);
  if ((uint64_t)first + count > s.cap) return fail(TECS_ERR_INVALID, "cells beyond the capacity");
  if (count && !pos) return fail(TECS_ERR_INVALID, "positions required");
  if (s.has_type && type && s.force.n_types > 0)
    for (unsigned i = 0; i < count; ++i)
      if (type[i] >= (unsigned)s.force.n_types) return fail(TECS_ERR_INVALID, "type out of range");
  std::memcpy(&s.pos[3 * (size_t)first], pos, sizeof(float) * 3 * count);
  if (s.has_pol && pol) std::memcpy(&s.pol[2 * (size_t)first], pol, sizeof(float) * 2 * count);
  if (s.has_type && type) std::memcpy(&s.type[first], type, count);
  return TECS_OK;
}

// Grid queries on the grid of the last rebuild and the current positions (grid.hpp:413-564):
// kind 0 get_nearest_point_index(point, radius_bins); 1 get_random_point_index_in_neighbourhood_bins
// (point, radius_bins, draw); 2 get_random_point_index_in_neighbourhood(point, radius, draw), whose
// bin radius is max(1, int(ceil(radius / bin_size))). The draw picks candidate
// min(int(u * total), total - 1) of the total in visiting order (the reference's rng.rand /
// urand from its Kokkos pool).
int orc_query(orc_sim* h, int kind, int n, const float* points, int radius_bins, float radius,
              const float* uniforms, int32_t* out) {
  Sim& s = *reinterpret_cast<Sim*>(h);
  if (!finder_is_binned(s)) return fail(TECS_ERR_UNSUPPORTED, "grid queries need a binned finder");
  if (!s.grid.built) return fail(TECS_ERR_STATE, "no grid yet");
  const Grid& g = s.grid;
  if (kind == 2) radius_bins = std::max(1, (int)std::ceil(radius / g.bin_size));
  const float r2 = radius * radius;
  for (int q = 0; q < n; ++q) {
    const float* p = &points[3 * (size_t)q];
    int c[3];
    g.coords(p, c);  // calc_bin_coords_from_point
    const int side = 2 * radius_bins + 1;
    const int tasks = side * side * side;
    // the members of the searched bins in visiting order (linear_index_to_offset, grid.hpp:308-322)
    std::vector<int> members;
    for (int t = 0; t < tasks; ++t) {
      const int tmp = t / side;
      const int ni[3] = {c[0] + (t % side) - radius_bins, c[1] + (tmp % side) - radius_bins,
                         c[2] + (tmp / side) - radius_bins};
      if (ni[0] < 0 || ni[0] >= g.ext[0] || ni[1] < 0 || ni[1] >= g.ext[1] || ni[2] < 0 ||
          ni[2] >= g.ext[2])
        continue;
      const int b = g.flatten(ni);
      for (int idx = g.offsets[b]; idx < g.offsets[(size_t)b + 1]; ++idx) members.push_back(g.perm[idx]);
    }
    int result = -1;
    if (kind == 0) {
      float best = FLT_MAX;
      for (int j : members) {
        const float d2 = dist2(p, &s.pos[3 * (size_t)j]);
        if (d2 < best) {
          best = d2;
          result = j;
        }
      }
    } else {
      std::vector<int> cand;
      for (int j : members)
        if (kind == 1 || dist2(p, &s.pos[3 * (size_t)j]) <= r2) cand.push_back(j);
      if (!cand.empty()) {
        const int total = (int)cand.size();
        const int target = std::min((int)(uniforms[q] * (float)total), total - 1);
        result = cand[(size_t)target];
      }
    }
    out[q] = result;
  }
  return TECS_OK;
}

// Simulation::set_agent_count, simulation.hpp:285-301: the views keep their contents, the pair
// finder's rebuild counter restarts (binned_all_pairs.hpp:60-63).
int orc_set_agent_count(orc_sim* h, unsigned n) {
  Sim& s = *reinterpret_cast<Sim*>(h);
  if (n > s.cap) return fail(TECS_ERR_INVALID, "agent_count above the capacity");
  s.n = n;
  s.desc.agent_count = n;
  s.step_count = 0;
  s.grid.built = false;
  return TECS_OK;
}

// Simulation::set_capacity, simulation.hpp:266-274: Kokkos::resize of every view keeps the
// first min(old, new) entries and zero-fills the rest.
int orc_set_capacity(orc_sim* h, unsigned cap) {
  Sim& s = *reinterpret_cast<Sim*>(h);
  if (cap < s.n) return fail(TECS_ERR_INVALID, "capacity below agent_count");
  const size_t c = cap;
  for (auto* v : {&s.pos, &s.d1pos, &s.x2pos, &s.d3pos, &s.old_vel}) {
    v->resize(std::min<size_t>(v->size(), 3 * c));
    v->resize(3 * c, 0.0f);
  }
  for (auto* v : {&s.pol, &s.d1pol, &s.x2pol, &s.d3pol}) {
    v->resize(std::min<size_t>(v->size(), 2 * c));
    v->resize(2 * c, 0.0f);
  }
  s.type.resize(std::min<size_t>(s.type.size(), c));
  s.type.resize(c, 0);
  s.cap = cap;
  s.desc.capacity = cap;
  return TECS_OK;
}

// Walls for the generic force: five floats each (a, b, c, d, strength); the plane is
// normalised as PlaneN's constructor does (types/plane.hpp:24-31): inv = 1 / sqrt(a^2 + b^2 +
// c^2), every coefficient times inv.
int orc_set_walls(orc_sim* h, const float* walls, int count) {
  Sim& s = *reinterpret_cast<Sim*>(h);
  if (count < 0 || count > TECS_MAX_WALLS) return fail(TECS_ERR_INVALID, "bad wall count");
  std::vector<std::array<float, 5>> ws;
  for (int q = 0; q < count; ++q) {
    const float* w = &walls[5 * q];
    float sum = 0.0f;
    for (int c = 0; c < 3; ++c) sum += w[c] * w[c];
    if (!(sum > 0.0f)) return fail(TECS_ERR_INVALID, "wall normal is zero");
    const float inv = 1.0f / std::sqrt(sum);
    ws.push_back({w[0] * inv, w[1] * inv, w[2] * inv, w[3] * inv, w[4]});
  }
  s.walls = ws;
  return TECS_OK;
}

int orc_set_links(orc_sim* h, const int32_t* links, int64_t count, float strength) {
  Sim& s = *reinterpret_cast<Sim*>(h);
  if (count < 0) return fail(TECS_ERR_INVALID, "bad link count");
  for (int64_t q = 0; q < 2 * count; ++q)
    if (links[q] < 0 || (unsigned)links[q] >= s.cap) return fail(TECS_ERR_INVALID, "link endpoint beyond the capacity");
  s.links.assign(links, links + 2 * count);
  s.link_strength = strength;
  return TECS_OK;
}

// The UPDATE_FUNC `proliferate` of examples/growth_w_wall.cpp:52-64, run by Simulation::init
// (simulation.hpp:520-529) over the active cells: cell i divides when u0 <= rate (the reference
// returns when drand > rate); the new cell takes the next index (the reference's atomic counter;
// here in the order of the parents), at p_i + offset * to_vector3(theta, phi) with theta =
// float(acos(2 u1 - 1)), phi = float(u2 * 2 * double(pi_f)) (drand is double, Polarity narrows;
// polarity.hpp:44, :111-118). The sines and cosines are the double functions rounded to float
// (within an ulp of a float sin / cos). Only positions of the new cells are written.
int orc_divide(orc_sim* h, float rate, float offset, const double* u, unsigned* new_count) {
  Sim& s = *reinterpret_cast<Sim*>(h);
  unsigned m = s.n;
  for (unsigned i = 0; i < s.n; ++i) m += u[3 * (size_t)i] <= (double)rate;
  if (m > s.cap) return fail(TECS_ERR_CAPACITY, "new cells beyond the capacity");
  const double pi_f = (double)3.14159265358979323846f;
  m = s.n;
  for (unsigned i = 0; i < s.n; ++i) {
    const double* v = &u[3 * (size_t)i];
    if (v[0] > (double)rate) continue;
    const float theta = (float)std::acos(2.0 * v[1] - 1);
    const float phi = (float)(v[2] * 2.0 * pi_f);
    const float st = (float)std::sin((double)theta), ct = (float)std::cos((double)theta);
    const float sp = (float)std::sin((double)phi), cp = (float)std::cos((double)phi);
    const float dir[3] = {st * cp, st * sp, ct};
    for (int c = 0; c < 3; ++c) s.pos[3 * (size_t)m + c] = s.pos[3 * (size_t)i + c] + offset * dir[c];
    ++m;
  }
  *new_count = m;
  return TECS_OK;
}

int orc_download(orc_sim* h, float* pos, float* pol, float* old_vel) {
  Sim& s = *reinterpret_cast<Sim*>(h);
  if (pos) std::memcpy(pos, s.pos.data(), sizeof(float) * 3 * s.n);
  if (pol && s.has_pol) std::memcpy(pol, s.pol.data(), sizeof(float) * 2 * s.n);
  if (old_vel) std::memcpy(old_vel, s.old_vel.data(), sizeof(float) * 3 * s.n);
  return TECS_OK;
}

int orc_take_step(orc_sim* h, double dt, int n_steps) {
  Sim& s = *reinterpret_cast<Sim*>(h);
  for (int k = 0; k < n_steps; ++k) {
    const int rc = s.desc.integrator == TECS_EULER ? step_euler(s, dt) : step_heun(s, dt);
    if (rc) return rc;
  }
  return TECS_OK;
}

int orc_rebuild(orc_sim* h) {
  Sim& s = *reinterpret_cast<Sim*>(h);
  if (!finder_is_binned(s)) return TECS_OK;
  s.grid.rebuild(s.pos, s.n);
  s.stats.rebuilds++;
  s.step_count = 1;  // as a due full step: reset to 0, then ++
  return TECS_OK;
}

int orc_evaluate_force(orc_sim* h, int is_full_step, float* dpos, float* dpol) {
  Sim& s = *reinterpret_cast<Sim*>(h);
  std::vector<float> op(3 * (size_t)s.cap, 0.0f), ol(2 * (size_t)s.cap, 0.0f);
  int rc = evaluate(s, s.pos.data(), s.pol.data(), op.data(), ol.data(), is_full_step != 0, nullptr);
  if (rc) return rc;
  if (dpos) std::memcpy(dpos, op.data(), sizeof(float) * 3 * s.n);
  if (dpol && s.has_pol) std::memcpy(dpol, ol.data(), sizeof(float) * 2 * s.n);
  return TECS_OK;
}

int orc_pairs(orc_sim* h, int is_full_step, int64_t* row_ptr, int32_t* cols, int64_t cap,
              int64_t* n_pairs) {
  Sim& s = *reinterpret_cast<Sim*>(h);
  std::vector<std::vector<int>> rows(s.n);
  if (!cols) {
    // size query: evaluate without touching the rebuild schedule
    Sim tmp_state = s;
    int rc = evaluate(tmp_state, tmp_state.pos.data(), tmp_state.pol.data(), nullptr, nullptr,
                      is_full_step != 0, &rows);
    if (rc) return rc;
  } else {
    int rc = evaluate(s, s.pos.data(), s.pol.data(), nullptr, nullptr, is_full_step != 0, &rows);
    if (rc) return rc;
  }
  int64_t total = 0;
  for (unsigned i = 0; i < s.n; ++i) total += (int64_t)rows[i].size();
  if (n_pairs) *n_pairs = total;
  if (!cols) return TECS_OK;
  if (cap < total) return fail(TECS_ERR_CAPACITY, "cols buffer too small");
  int64_t at = 0;
  for (unsigned i = 0; i < s.n; ++i) {
    if (row_ptr) row_ptr[i] = at;
    for (int j : rows[i]) cols[at++] = j;
  }
  if (row_ptr) row_ptr[s.n] = at;
  return TECS_OK;
}

int orc_get_stats(orc_sim* h, tecs_stats* st) {
  *st = reinterpret_cast<Sim*>(h)->stats;
  return TECS_OK;
}

// Double-precision evaluation of the fused force on a given pair set (CSR over i in the
// visiting order) at the current stage-0 state. Writes the FP64 deltas and, per component, the
// sum of absolute pair contributions (the scale the FP32 error is measured against).
// symmetric = 0: each listed pair (i, j) adds its value to i only (the ordered finders).
// symmetric = 1: each listed pair adds its value v to i and -v to j (the symmetric finders).
int orc_forces_f64(orc_sim* h, int symmetric, const int64_t* row_ptr, const int32_t* cols,
                   double* dpos, double* dpol, double* scale_pos, double* scale_pol) {
  Sim& s = *reinterpret_cast<Sim*>(h);
  const long n = (long)s.n;
  std::vector<double> ap(3 * (size_t)n, 0.0), al(2 * (size_t)n, 0.0), sp(3 * (size_t)n, 0.0),
      sl(2 * (size_t)n, 0.0);
  if (!symmetric) {
#pragma omp parallel for schedule(dynamic, 256)
    for (long i = 0; i < n; ++i) {
      for (int64_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
        double vp[3], vl[2], xp[3], xl[2];
        pair_force_f64(s, i, cols[e], vp, vl, xp, xl);
        for (int c = 0; c < 3; ++c) {
          ap[3 * i + c] += vp[c];
          sp[3 * i + c] += xp[c];
        }
        for (int c = 0; c < 2; ++c) {
          al[2 * i + c] += vl[c];
          sl[2 * i + c] += xl[c];
        }
      }
    }
  } else {
    for (long i = 0; i < n; ++i) {
      for (int64_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
        const long j = cols[e];
        double vp[3], vl[2], xp[3], xl[2];
        pair_force_f64(s, i, j, vp, vl, xp, xl);
        for (int c = 0; c < 3; ++c) {
          ap[3 * i + c] += vp[c];
          ap[3 * j + c] -= vp[c];
          sp[3 * i + c] += xp[c];
          sp[3 * j + c] += xp[c];
        }
        for (int c = 0; c < 2; ++c) {
          al[2 * i + c] += vl[c];
          al[2 * j + c] -= vl[c];
          sl[2 * i + c] += xl[c];
          sl[2 * j + c] += xl[c];
        }
      }
    }
  }
  if (dpos) std::memcpy(dpos, ap.data(), sizeof(double) * ap.size());
  if (dpol) std::memcpy(dpol, al.data(), sizeof(double) * al.size());
  if (scale_pos) std::memcpy(scale_pos, sp.data(), sizeof(double) * sp.size());
  if (scale_pol) std::memcpy(scale_pol, sl.data(), sizeof(double) * sl.size());
  return TECS_OK;
}

// ------------------------------------------------------------------ unit-test hooks
// Direct access to the restated vector / polarity / COM arithmetic for known-answer tests
// (tests/test_vector.cpp:91-105, tests/test_polarity.cpp:28-96 of the reference).
//   op 0: dot(a, b)                       a = in[0..3), b = in[3..6)         -> out[0]
//   op 1: length_squared(a)                                                   -> out[0]
//   op 2: distance_to_squared(a, b)                                           -> out[0]
int orc_vector_op(int op, const float* in, float* out) {
  switch (op) {
    case 0: out[0] = dot3(in, in + 3); return TECS_OK;
    case 1: out[0] = dot3(in, in); return TECS_OK;
    case 2: out[0] = dist2(in, in + 3); return TECS_OK;
    default: return fail(TECS_ERR_INVALID, "bad vector op");
  }
}
//   op 0: Polarity_(vector, distance)   in = v[3], d                         -> out[0..2)
//   op 1: Polarity_(vector)             in = v[3] (length computed)          -> out[0..2)
//   op 2: to_vector3(p)                 in = p[2]                            -> out[0..3)
//   op 3: p.dot(q)                      in = p[2], q[2]                      -> out[0]
//   op 4: p.unidirectional(q)           in = p[2], q[2]                      -> out[0..2)
//   op 5: p.bidirectional(q)            in = p[2], q[2]                      -> out[0..2)
//   op 6: p.bending_force(disp, q, d)   in = p[2], disp[3], q[2], d          -> vec[3], pol[2]
int orc_polarity_op(int op, const float* in, float* out) {
  switch (op) {
    case 0: pol_from_vec3(in, in[3], out); return TECS_OK;
    case 1: {
      const float len = std::sqrt(dot3(in, in));
      pol_from_vec3(in, len, out);
      return TECS_OK;
    }
    case 2: pol_to_vec3(in, out); return TECS_OK;
    case 3: out[0] = pol_dot(in, in + 2); return TECS_OK;
    case 4: pol_unidirectional(in, in + 2, out); return TECS_OK;
    case 5: pol_bidirectional(in, in + 2, out); return TECS_OK;
    case 6: pol_bending(in, in + 2, in + 5, in[7], out, out + 3); return TECS_OK;
    default: return fail(TECS_ERR_INVALID, "bad polarity op");
  }
}
// GlobalComFixer::fix on a host array of `capacity` position deltas (com_fixers.hpp:35-43).
int orc_com_fix(const float* deltas, unsigned capacity, float* com) {
  float t[3] = {0.0f, 0.0f, 0.0f};
  if (capacity == 0) {
    com[0] = com[1] = com[2] = 0.0f;
    return TECS_OK;
  }
  for (size_t i = 0; i < capacity; ++i)
    for (int c = 0; c < 3; ++c) t[c] += deltas[3 * i + c];
  for (int c = 0; c < 3; ++c) com[c] = t[c] / (float)capacity;
  return TECS_OK;
}

}  // extern "C"
