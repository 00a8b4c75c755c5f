"""Seeded synthetic inputs and descriptors for the benchmark workloads.

Inputs come from numpy's Philox counter-based generator seeded with 2807 (the reference's
default seed, include/simulation.hpp:89) plus a per-workload stream offset, computed in float64
and rounded to float32 once. Both the GPU engine and the CPU oracle consume the same host
arrays. Bounds of the binned finders are the data bounding box +- 2 r_max (the pattern of
examples/branching.cpp:48-55).

  relaxation  N = 10,000 uniform in a ball of radius 0.5*cbrt(N); PiecewiseLinear(0.8, 2, 1)
              (RelaxForce, initializers/relax_force.hpp:27-29); BinnedAllPairs; Euler dt 0.1
  epithelium  N = 200,000 on a sphere of radius sqrt(N/4pi)*0.8, radial jitter N(0, 0.05);
              polarity = outward normal (theta, phi) + N(0, 0.1); PiecewiseLinear(0.8, 2, 1) +
              bending x1 + bidirectional alignment x1 (examples/epithelium.cpp:30-36,
              examples/polarization.cpp:78-84); BinnedGabriel c = 0.8; Heun dt 0.05
  sorting     N = 4,000,000 uniform in a cube at 1.5 cells per unit volume, types A/B p = 0.5;
              PiecewiseLinear(0.8, 2, tau) with tau_AA = tau_BB = 1, tau_AB = 0.3
              (examples/sorting.cpp:43-58); BinnedAllPairs; Heun dt 0.05
  spheroid    N = 100 uniform in a ball of radius 3; PiecewiseLinear(0.5) x 3; NaiveDelaunay;
              Heun dt 0.001 (examples/delaunay_spheroid.cpp)
  sweep       many independent sorting/relaxation instances of log-uniform size (batched)
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import abi

SEED = 2807
R_MAX = 1.0


def rng(stream: int = 0) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(key=SEED + stream))


def _unit_vectors(g, n):
    v = g.standard_normal((n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v


def ball(n: int, radius: float, g) -> np.ndarray:
    r = radius * np.cbrt(g.random(n))
    return (_unit_vectors(g, n) * r[:, None]).astype(np.float32)


def cube(n: int, edge: float, g) -> np.ndarray:
    return ((g.random((n, 3)) - 0.5) * edge).astype(np.float32)


def shell(n: int, radius: float, jitter: float, g):
    """Positions on a jittered sphere and polarities = outward normal + N(0, 0.1) noise."""
    u = _unit_vectors(g, n)
    r = radius + jitter * g.standard_normal(n)
    pos = (u * r[:, None]).astype(np.float32)
    theta = np.arccos(np.clip(u[:, 2], -1.0, 1.0)) + 0.1 * g.standard_normal(n)
    phi = np.arctan2(u[:, 1], u[:, 0]) + 0.1 * g.standard_normal(n)
    pol = np.stack([theta, phi], axis=1).astype(np.float32)
    return pos, pol


def bounds(pos: np.ndarray, margin: float = 2.0 * R_MAX):
    lo = pos.min(axis=0).astype(np.float64) - margin
    hi = pos.max(axis=0).astype(np.float64) + margin
    return tuple(float(x) for x in lo), tuple(float(x) for x in hi)


@dataclass
class Workload:
    name: str
    desc: abi.SimDesc
    positions: np.ndarray
    polarities: np.ndarray | None
    types: np.ndarray | None
    dt: float
    steps: int

    @property
    def n(self) -> int:
        return int(self.desc.agent_count)


def relaxation(n: int = 10_000, seed_stream: int = 0, finder: int = abi.BINNED_ALL_PAIRS) -> Workload:
    g = rng(seed_stream)
    pos = ball(n, 0.5 * n ** (1.0 / 3.0), g)
    lo, hi = bounds(pos)
    d = abi.make_desc(n, pair_finder=finder, min_bounds=lo, max_bounds=hi, integrator=abi.EULER)
    return Workload("relaxation", d, pos, None, None, 0.1, 1000)


def epithelium(n: int = 200_000, seed_stream: int = 1, finder: int = abi.BINNED_GABRIEL) -> Workload:
    g = rng(seed_stream)
    pos, pol = shell(n, math.sqrt(n / (4.0 * math.pi)) * 0.8, 0.05, g)
    lo, hi = bounds(pos)
    d = abi.make_desc(
        n,
        pair_finder=finder,
        min_bounds=lo,
        max_bounds=hi,
        integrator=abi.HEUN,
        polarity=True,
        bending_strength=1.0,
        alignment_strength=1.0,
    )
    return Workload("epithelium", d, pos, pol, None, 0.05, 2000)


SORTING_TABLE = [[1.0, 0.3], [0.3, 1.0]]


def sorting(n: int = 4_000_000, seed_stream: int = 2, tau_ab: float = 0.3,
            finder: int = abi.BINNED_ALL_PAIRS, density: float = 1.5) -> Workload:
    g = rng(seed_stream)
    pos = cube(n, (n / density) ** (1.0 / 3.0), g)
    types = (g.random(n) < 0.5).astype(np.uint8)
    lo, hi = bounds(pos)
    d = abi.make_desc(
        n,
        pair_finder=finder,
        min_bounds=lo,
        max_bounds=hi,
        integrator=abi.HEUN,
        types=True,
        adhesion_table=[[1.0, tau_ab], [tau_ab, 1.0]],
    )
    return Workload("sorting", d, pos, None, types, 0.05, 500)


def spheroid(n: int = 100, seed_stream: int = 4, finder: int = abi.NAIVE_DELAUNAY) -> Workload:
    """examples/delaunay_spheroid.cpp:19-53: n cells uniform in a ball of radius 3, r_max 1,
    NaiveDelaunay, Heun dt 0.001 for 300 steps, PiecewiseLinear(0.5) (scales 1, 1,
    forces/predefined.hpp:109-143) times 3, the factor folded into both scales."""
    g = rng(seed_stream)
    pos = ball(n, 3.0, g)
    d = abi.make_desc(n, pair_finder=finder, integrator=abi.HEUN, homeostatic_radius=0.5,
                      repulsion_scale=3.0, adhesion_scale=3.0)
    return Workload("spheroid", d, pos, None, None, 0.001, 300)


def sweep(count: int = 256, n_min: int = 1_000, n_max: int = 100_000, seed_stream: int = 3):
    """Batched parameter sweep: instance k uses the relaxation geometry (ball, one cell type)
    when k is even and the sorting geometry (cube, A/B types) when k is odd; tau_AB runs
    linearly over [0.1, 1.0]; sizes are log-uniform in [n_min, n_max]."""
    g = rng(seed_stream)
    sizes = np.exp(g.uniform(math.log(n_min), math.log(n_max), count)).astype(np.int64)
    taus = np.linspace(0.1, 1.0, count)
    out = []
    for k in range(count):
        n = int(sizes[k])
        gk = rng(1000 + seed_stream * 100_000 + k)
        if k % 2 == 0:
            pos = ball(n, 0.5 * n ** (1.0 / 3.0), gk)
            types = np.zeros(n, np.uint8)
        else:
            pos = cube(n, (n / 1.5) ** (1.0 / 3.0), gk)
            types = (gk.random(n) < 0.5).astype(np.uint8)
        lo, hi = bounds(pos)
        t = float(taus[k])
        d = abi.make_desc(
            n,
            pair_finder=abi.BINNED_ALL_PAIRS,
            min_bounds=lo,
            max_bounds=hi,
            integrator=abi.HEUN,
            types=True,
            adhesion_table=[[1.0, t], [t, 1.0]],
        )
        out.append(Workload(f"sweep[{k}]", d, pos, None, types, 0.05, 500))
    return out


@dataclass
class ServeRequest:
    arrival: float  # seconds after the start
    model: str
    workload: Workload
    n_steps: int


def serve_requests(count: int = 64, rate: float = 8.0, n_min: int = 2_000, n_max: int = 500_000,
                   steps_min: int = 100, steps_max: int = 5_000, seed_stream: int = 4):
    """Served requests: Poisson arrivals at `rate` per second, the model drawn uniformly from
    relaxation / epithelium / sorting, cell counts log-uniform in [n_min, n_max] and step counts
    uniform in [steps_min, steps_max]."""
    g = rng(seed_stream)
    gaps = g.exponential(1.0 / rate, count)
    arrivals = np.cumsum(gaps) - gaps[0]
    models = g.integers(0, 3, count)
    sizes = np.exp(g.uniform(math.log(n_min), math.log(n_max), count)).astype(np.int64)
    steps = g.integers(steps_min, steps_max + 1, count)
    out = []
    for k in range(count):
        n = int(sizes[k])
        name = ("relaxation", "epithelium", "sorting")[int(models[k])]
        stream = 10_000 + seed_stream * 1000 + k
        w = {"relaxation": relaxation, "epithelium": epithelium, "sorting": sorting}[name](n, seed_stream=stream)
        out.append(ServeRequest(float(arrivals[k]), name, w, int(steps[k])))
    return out
