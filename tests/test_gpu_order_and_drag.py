"""Force composition order and drag on the GPU, through the C ABI.

Bars:
  * composition order: on three-cell systems the GPU's composed delta of cell 0 is bitwise the
    float32 fold, pair by pair in visiting order and term by term in the composition's order,
    of the GPU's own single-term values (the oracle's rule, pinned there); on the epithelium the
    composed forces are within the 1e-5 bar of FP64 for every order, and two orders give
    different GPU sums (the order is applied, not ignored);
  * drag (PiecewiseLinear forces, whose GPU arithmetic is the oracle's operation for
    operation): positions and old velocities bitwise the oracle's after 20 steps for the binned
    and naive all-pairs finders, BinnedGabriel and NaiveDelaunay, Euler and Heun, a constant
    drag and a type-pair drag table; with GlobalComFixer within 1e-5 (its reduction order
    differs); drag switched on between steps, and a batch whose members differ in drag, bitwise.
"""
from __future__ import annotations

import numpy as np
import pytest

from tecs_amd import GpuBatch, GpuSimulation, abi, workloads as W

pytestmark = pytest.mark.gpu

PL, BEND, ALIGN = abi.TERM_PIECEWISE_LINEAR, abi.TERM_BENDING, abi.TERM_ALIGNMENT


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _desc(n, order):
    d = abi.make_desc(n, pair_finder=abi.NAIVE_ALL_PAIRS, integrator=abi.EULER, polarity=True,
                      bending_strength=0.7, alignment_strength=1.3)
    abi.set_terms(d.force, term_order=order)
    return d


def _gpu_eval(order, pos, pol):
    g = GpuSimulation(_desc(len(pos), order))
    g.upload(pos, pol)
    out = g.evaluate_force(True)
    g.close()
    return out


@pytest.mark.parametrize("order", [(PL, BEND, ALIGN), (BEND, PL, ALIGN), (ALIGN, BEND, PL), (BEND, ALIGN, PL)],
                         ids=lambda o: "-".join(str(t) for t in o))
def test_gpu_composition_is_the_ordered_fold(gpu, order):
    rng = W.rng(301)
    print(f"case order {order}", flush=True)
    for _ in range(12):
        pos = np.zeros((3, 3), np.float32)
        for j in (1, 2):
            u = rng.standard_normal(3)
            pos[j] = u / np.linalg.norm(u) * rng.uniform(0.3, 0.95)
        pol = np.stack([rng.uniform(0.3, 2.8, 3), rng.uniform(-3, 3, 3)], axis=1).astype(np.float32)
        f, p = _gpu_eval(order, pos, pol)
        acc_f = np.zeros(3, np.float32)
        acc_p = np.zeros(2, np.float32)
        for j in (1, 2):
            for t in order:
                sf, sp = _gpu_eval((t,), pos[[0, j]], pol[[0, j]])
                acc_f = acc_f + sf[0]
                acc_p = acc_p + sp[0]
        assert np.array_equal(_bits(f[0]), _bits(acc_f)) and np.array_equal(_bits(p[0]), _bits(acc_p)), order


def test_gpu_orders_on_the_epithelium(gpu, oracle_lib):
    w = W.epithelium(4000)
    res = {}
    for order in [(PL, BEND, ALIGN), (BEND, PL, ALIGN), (ALIGN, PL, BEND)]:
        print(f"case epithelium order {order}", flush=True)
        d = abi.copy_desc(w.desc)
        abi.set_terms(d.force, term_order=order)
        g = GpuSimulation(d)
        o = oracle_lib.OracleSim(d)
        for s in (g, o):
            s.upload(w.positions, w.polarities)
        rg, cg = g.pairs(True)
        ro, co = o.pairs(True)
        assert np.array_equal(rg, ro) and np.array_equal(cg, co)
        fg, pg = g.evaluate_force(False)
        f64, p64, s64, sp64 = o.forces_f64(ro, co)
        assert oracle_lib.force_excess(fg, f64, s64) <= 0.0
        assert oracle_lib.force_excess(pg, p64, sp64) <= 0.0
        res[order] = (fg, pg)
    a, b, c = res.values()
    assert not np.array_equal(_bits(a[0]), _bits(b[0]))  # bending before / after PiecewiseLinear
    assert not np.array_equal(_bits(a[1]), _bits(c[1]))  # alignment before / after bending


def _drag_case(name, finder, integrator, com=False):
    if name == "relaxation":
        w = W.relaxation(3000, finder=finder)
        d = abi.copy_desc(w.desc)
        abi.set_terms(d.force, drag=1.0)
    elif name == "sorting":
        w = W.sorting(12000, finder=finder)
        d = abi.copy_desc(w.desc)
        abi.set_terms(d.force, drag_table=[[1.0, 0.0], [0.5, 2.0]])
    else:  # NaiveDelaunay spheroid
        w = W.spheroid(100)
        d = abi.copy_desc(w.desc)
        abi.set_terms(d.force, drag=1.0)
    d.integrator = integrator
    if com:
        d.com_fixer = abi.GLOBAL_COM_FIXER
    return w, d


DRAG_CASES = [
    ("relaxation", abi.BINNED_ALL_PAIRS, abi.HEUN),
    ("relaxation", abi.BINNED_ALL_PAIRS, abi.EULER),
    ("relaxation", abi.NAIVE_ALL_PAIRS, abi.EULER),
    ("relaxation", abi.BINNED_GABRIEL, abi.HEUN),
    ("sorting", abi.BINNED_ALL_PAIRS, abi.HEUN),
    ("spheroid", abi.NAIVE_DELAUNAY, abi.HEUN),
]


@pytest.mark.parametrize("name,finder,integrator", DRAG_CASES,
                         ids=[f"{n}-{abi.FINDER_NAMES[f]}-{'Heun' if i else 'Euler'}" for n, f, i in DRAG_CASES])
def test_drag_bitwise_against_the_oracle(gpu, oracle_lib, name, finder, integrator):
    w, d = _drag_case(name, finder, integrator)
    print(f"case drag {name} {abi.FINDER_NAMES[finder]} integrator={integrator}", flush=True)
    g = GpuSimulation(d)
    o = oracle_lib.OracleSim(d)
    for s in (g, o):
        s.upload(w.positions, None, w.types)
        s.take_step(w.dt, 20)
    xg, _, vg = g.download()
    xo, _, vo = o.download()
    assert np.array_equal(_bits(xg), _bits(xo))
    assert np.array_equal(_bits(vg), _bits(vo))
    # drag changes the trajectory (the term is applied)
    n = abi.copy_desc(d)
    abi.set_terms(n.force)
    r = oracle_lib.OracleSim(n)
    r.upload(w.positions, None, w.types)
    r.take_step(w.dt, 20)
    assert not np.array_equal(_bits(r.download()[0]), _bits(xo))


def test_drag_with_the_com_fixer(gpu, oracle_lib):
    w, d = _drag_case("relaxation", abi.BINNED_ALL_PAIRS, abi.HEUN, com=True)
    print("case drag + GlobalComFixer", flush=True)
    g = GpuSimulation(d)
    o = oracle_lib.OracleSim(d)
    for s in (g, o):
        s.upload(w.positions)
        s.take_step(w.dt, 20)
    assert np.abs(g.download()[0] - o.download()[0]).max() <= 1e-5


def test_drag_switched_on_between_steps(gpu, oracle_lib):
    w = W.relaxation(3000)
    d = abi.copy_desc(w.desc)
    d.integrator = abi.HEUN
    print("case drag switched on after captured graphs", flush=True)
    g = GpuSimulation(d)
    o = oracle_lib.OracleSim(d)
    for s in (g, o):
        s.upload(w.positions)
        s.take_step(w.dt, 12)  # replays captured graphs
    f = abi.ForceDesc()
    abi.C.memmove(abi.C.addressof(f), abi.C.addressof(d.force), abi.C.sizeof(abi.ForceDesc))
    abi.set_terms(f, drag=1.0)
    for s in (g, o):
        s.set_force(f)
        s.take_step(w.dt, 12)
    assert np.array_equal(_bits(g.download()[0]), _bits(o.download()[0]))


def test_batch_members_with_and_without_drag(gpu, oracle_lib):
    ws = W.sweep(count=4, n_min=1000, n_max=6000)
    descs = []
    for k, w in enumerate(ws):
        d = abi.copy_desc(w.desc)
        if k % 2:
            abi.set_terms(d.force, drag=1.0)
        descs.append(d)
    print("case batch with drag on two of four members", flush=True)
    b = GpuBatch(descs)
    for k, w in enumerate(ws):
        b.upload(k, w.positions, None, w.types)
    b.take_step([w.dt for w in ws], 10)
    for k, w in enumerate(ws):
        o = oracle_lib.OracleSim(descs[k])
        o.upload(w.positions, None, w.types)
        o.take_step(w.dt, 10)
        assert np.array_equal(_bits(b.download(k)[0]), _bits(o.download()[0])), k


@pytest.mark.parametrize("integrator", [abi.HEUN, abi.EULER], ids=["Heun", "Euler"])
@pytest.mark.parametrize("variant", [abi.SWEEP_TEAM4, abi.SWEEP_TEAM8, abi.SWEEP_TEAM16], ids=["team4", "team8", "team16"])
def test_team_sweeps_sum_like_the_thread_sweep(gpu, variant, integrator):
    """The team sweeps add each pair's terms to the running sums in the composition's order,
    like the thread sweep, and update a cell with the integrator's arithmetic (by component
    where the team has a lane per component), so their polarity-model forces, trajectories and
    old velocities are bitwise the thread sweep's (both read the same sines and cosines)."""
    w = W.epithelium(4000)
    d = abi.copy_desc(w.desc)
    d.integrator = integrator
    print(f"case epithelium team variant {variant} vs thread, integrator {integrator}", flush=True)
    out = {}
    for v in (abi.SWEEP_THREAD, variant):
        g = GpuSimulation(d)
        g.set_sweep_variant(v)
        g.upload(w.positions, w.polarities)
        f = g.evaluate_force(True)
        g.take_step(w.dt, 10)
        out[v] = (f, g.download())
        g.close()
    (fa, pa), (xa, qa, va) = out[abi.SWEEP_THREAD]
    (fb, pb), (xb, qb, vb) = out[variant]
    assert np.array_equal(_bits(fa), _bits(fb)) and np.array_equal(_bits(pa), _bits(pb))
    assert np.array_equal(_bits(xa), _bits(xb)) and np.array_equal(_bits(qa), _bits(qb))
    assert np.array_equal(_bits(va), _bits(vb)), "old velocities"
    assert np.isfinite(xb).all()
