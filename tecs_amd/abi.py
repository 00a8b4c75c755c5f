"""ctypes mirror of the plain-C descriptors in include/tecs_amd.h (data layout only).

Nothing here loads the engine library; see tecs_amd._lib for that.
"""
from __future__ import annotations

import ctypes as C

ABI_VERSION = 3

# status codes
OK = 0
ERR_INVALID = -1
ERR_DEVICE = -2
ERR_OOM = -3
ERR_UNSUPPORTED = -4
ERR_STATE = -5
ERR_CAPACITY = -6

# pair finders (include/pair_finders/*.hpp)
NAIVE_ALL_PAIRS = 0
BINNED_ALL_PAIRS = 1
NAIVE_GABRIEL = 2
BINNED_GABRIEL = 3
NAIVE_ALL_PAIRS_SYMMETRIC = 4
BINNED_ALL_PAIRS_SYMMETRIC = 5
NAIVE_GABRIEL_SYMMETRIC = 6
BINNED_GABRIEL_SYMMETRIC = 7
FINDER_NAMES = {
    NAIVE_ALL_PAIRS: "NaiveAllPairs",
    BINNED_ALL_PAIRS: "BinnedAllPairs",
    NAIVE_GABRIEL: "NaiveGabriel",
    BINNED_GABRIEL: "BinnedGabriel",
    NAIVE_ALL_PAIRS_SYMMETRIC: "NaiveAllPairsSymmetric",
    BINNED_ALL_PAIRS_SYMMETRIC: "BinnedAllPairsSymmetric",
    NAIVE_GABRIEL_SYMMETRIC: "NaiveGabrielSymmetric",
    BINNED_GABRIEL_SYMMETRIC: "BinnedGabrielSymmetric",
}
NAIVE_DELAUNAY = 8
MAX_WALLS = 8
FINDER_NAMES[NAIVE_DELAUNAY] = "NaiveDelaunay"
SYMMETRIC_FINDERS = (NAIVE_ALL_PAIRS_SYMMETRIC, BINNED_ALL_PAIRS_SYMMETRIC, NAIVE_GABRIEL_SYMMETRIC,
                     BINNED_GABRIEL_SYMMETRIC)
BINNED_FINDERS = (BINNED_ALL_PAIRS, BINNED_GABRIEL, BINNED_ALL_PAIRS_SYMMETRIC, BINNED_GABRIEL_SYMMETRIC)

EULER = 0
HEUN = 1
NO_COM_FIXER = 0
GLOBAL_COM_FIXER = 1

FIELD_POLARITY = 1
FIELD_TYPE = 2
MAX_TYPES = 4

SWEEP_AUTO = 0
SWEEP_THREAD = 1
SWEEP_TEAM = 2
SWEEP_TEAM4 = 3
SWEEP_TEAM8 = 4
SWEEP_TEAM16 = 5
SWEEP_NAMES = {1: "thread", 3: "team4", 4: "team8", 5: "team16"}
MAX_SIZE_CLASSES = 16  # TECS_MAX_SIZE_CLASSES

GRAPHS_AUTO = 0
GRAPHS_OFF = 1
GRAPHS_ON = 2

KT_SWEEP_FULL = 0
KT_SWEEP_STALE = 1
KT_REBUILD = 2


# terms of the fused pairwise force (tecs_force_desc.term_order)
TERM_END = 0
TERM_PIECEWISE_LINEAR = 1
TERM_BENDING = 2
TERM_ALIGNMENT = 3


class PairFinderSettings(C.Structure):
    _fields_ = [
        ("min_bounds", C.c_float * 3),
        ("max_bounds", C.c_float * 3),
        ("bin_size_scale", C.c_float),
        ("rebuild_every_n", C.c_int32),
        ("gabriel_coefficient", C.c_float),
    ]


class ForceDesc(C.Structure):
    _fields_ = [
        ("homeostatic_radius", C.c_float),
        ("repulsion_scale", C.c_float),
        ("adhesion_scale", C.c_float),
        ("n_types", C.c_int32),
        ("adhesion_table", C.c_float * (MAX_TYPES * MAX_TYPES)),
        ("bending_strength", C.c_float),
        ("alignment_strength", C.c_float),
        ("term_order", C.c_int32 * 3),
        ("drag", C.c_float),
        ("drag_by_type", C.c_int32),
        ("drag_table", C.c_float * (MAX_TYPES * MAX_TYPES)),
    ]


class SimDesc(C.Structure):
    _fields_ = [
        ("agent_count", C.c_uint32),
        ("capacity", C.c_uint32),
        ("fields", C.c_uint32),
        ("cutoff_distance", C.c_float),
        ("pair_finder", C.c_int32),
        ("pair_finder_settings", PairFinderSettings),
        ("integrator", C.c_int32),
        ("com_fixer", C.c_int32),
        ("force", ForceDesc),
        ("track_old_velocities", C.c_int32),
    ]


class Wall(C.Structure):
    _fields_ = [("plane", C.c_float * 4), ("strength", C.c_float)]


def make_walls(rows):
    """A ctypes array of tecs_wall from rows (a, b, c, d, strength)."""
    rows = [tuple(float(v) for v in r) for r in rows]
    arr = (Wall * max(len(rows), 1))()
    for q, r in enumerate(rows):
        for j in range(4):
            arr[q].plane[j] = r[j]
        arr[q].strength = r[4]
    return arr, len(rows)


class Stats(C.Structure):
    _fields_ = [
        ("pair_interactions", C.c_uint64),
        ("evaluations", C.c_uint64),
        ("rebuilds", C.c_uint64),
        ("steps", C.c_uint64),
    ]


def make_desc(
    agent_count: int,
    *,
    capacity: int | None = None,
    cutoff_distance: float = 1.0,
    pair_finder: int = BINNED_ALL_PAIRS,
    min_bounds=(-20.0, -20.0, -20.0),
    max_bounds=(20.0, 20.0, 20.0),
    bin_size_scale: float = 1.0,
    rebuild_every_n: int = 0,
    gabriel_coefficient: float = 0.8,
    integrator: int = HEUN,
    com_fixer: int = NO_COM_FIXER,
    polarity: bool = False,
    types: bool = False,
    homeostatic_radius: float = 0.8,
    repulsion_scale: float = 2.0,
    adhesion_scale: float = 1.0,
    adhesion_table=None,
    bending_strength: float = 0.0,
    alignment_strength: float = 0.0,
    term_order=None,
    drag: float = 0.0,
    drag_table=None,
    track_old_velocities: bool = True,
) -> SimDesc:
    """A descriptor with the reference defaults (Settings structs of the finders, RelaxForce's
    PiecewiseLinear(0.8, 2, 1) from include/initializers/relax_force.hpp:27-29)."""
    d = SimDesc()
    d.agent_count = int(agent_count)
    d.capacity = int(agent_count if capacity is None else capacity)
    d.fields = (FIELD_POLARITY if polarity else 0) | (FIELD_TYPE if types else 0)
    d.cutoff_distance = float(cutoff_distance)
    d.pair_finder = int(pair_finder)
    ps = d.pair_finder_settings
    for k in range(3):
        ps.min_bounds[k] = float(min_bounds[k])
        ps.max_bounds[k] = float(max_bounds[k])
    ps.bin_size_scale = float(bin_size_scale)
    ps.rebuild_every_n = int(rebuild_every_n)
    ps.gabriel_coefficient = float(gabriel_coefficient)
    d.integrator = int(integrator)
    d.com_fixer = int(com_fixer)
    f = d.force
    f.homeostatic_radius = float(homeostatic_radius)
    f.repulsion_scale = float(repulsion_scale)
    f.adhesion_scale = float(adhesion_scale)
    if adhesion_table is not None:
        import numpy as np

        t = np.asarray(adhesion_table, dtype=np.float32)
        n = t.shape[0]
        if t.shape != (n, n) or not 1 <= n <= MAX_TYPES:
            raise ValueError("adhesion_table must be square with 1..4 types")
        f.n_types = n
        for a in range(n):
            for b in range(n):
                f.adhesion_table[a * n + b] = float(t[a, b])
    f.bending_strength = float(bending_strength)
    f.alignment_strength = float(alignment_strength)
    set_terms(f, term_order=term_order, drag=drag, drag_table=drag_table)
    d.track_old_velocities = 1 if track_old_velocities else 0
    return d


def set_terms(f: ForceDesc, term_order=None, drag: float = 0.0, drag_table=None) -> ForceDesc:
    """The order of the force terms (a sequence of TERM_* ids, None = the default
    PiecewiseLinear, bending, alignment) and the drag (a constant, or a type-pair table)."""
    order = list(term_order or [])
    if len(order) > 3:
        raise ValueError("at most three terms")
    for k in range(3):
        f.term_order[k] = int(order[k]) if k < len(order) else TERM_END
    f.drag = float(drag)
    f.drag_by_type = 0
    for k in range(MAX_TYPES * MAX_TYPES):
        f.drag_table[k] = 0.0
    if drag_table is not None:
        import numpy as np

        t = np.asarray(drag_table, dtype=np.float32)
        n = t.shape[0]
        if t.shape != (n, n) or n != f.n_types:
            raise ValueError("drag_table must be square with one row per cell type of the adhesion table")
        f.drag_by_type = 1
        for a in range(n):
            for b in range(n):
                f.drag_table[a * n + b] = float(t[a, b])
    return f


def copy_desc(d: SimDesc) -> SimDesc:
    out = SimDesc()
    C.memmove(C.addressof(out), C.addressof(d), C.sizeof(SimDesc))
    return out


REQUEST_PENDING = 0
REQUEST_RUNNING = 1
REQUEST_DONE = 2
REQUEST_FAILED = 3
REQUEST_CANCELLED = 4


class RequestInfo(C.Structure):
    _fields_ = [
        ("state", C.c_int32),
        ("steps_done", C.c_int32),
        ("t_submit", C.c_double),
        ("t_admit", C.c_double),
        ("t_done", C.c_double),
        ("hbm_bytes", C.c_uint64),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("completed", C.c_uint64),
        ("pending", C.c_uint64),
        ("running", C.c_uint64),
        ("failed", C.c_uint64),
        ("cancelled", C.c_uint64),
        ("pair_interactions", C.c_uint64),
        ("member_steps", C.c_uint64),
        ("regroups", C.c_uint64),
        ("hbm_bytes_in_use", C.c_uint64),
        ("hbm_budget", C.c_uint64),
        ("pool_cached_bytes", C.c_uint64),
        ("regroup_seconds", C.c_double),
    ]
