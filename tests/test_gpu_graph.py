"""Graph replay of the electrostatics (mbgpu_set_graphs) against launching kernel by kernel.

Once an evaluation has set the solver's iteration count, a context captures the next
evaluation's electrostatics (the start, count + 2 iterations, the final pass) as one HIP
graph and replays it while the parameters, the cell and the buffers stay the same.  Every
replay must give exactly what launching the same kernels one by one gives:

* several configurations in a row through engine.Slot (the electrostatics on their own
  stream beside MB-pol and the water terms), with the program's own eager work and
  synchronizations between the replays: energies, forces, J, dipole iterations and the
  per-site electrostatics arrays bit-identical, for the Ewald sum over k, the particle mesh
  and an isolated cluster, with both solvers;
* a graph holding fewer iterations than the solve needs (MBGPU_GRAPH_BATCH): the solve
  continues kernel by kernel and the final pass runs again, still bit-identical;
* a change of parameters captures a new graph;
* the complete evaluation in one call (mbgpu_evaluate_dev, engine.Slot's default: MB-pol,
  electrostatics, water terms and heat flux enqueued together and, from the second call on,
  one captured graph) against the parts called one by one, bit for bit, with graphs and
  without, and its fall-back to the separate entries when a graph holds too few iterations.
"""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DETAILS = ("atomic_energy", "heat_w", "s_ew1p", "vesp", "chge", "f_sites")


@pytest.fixture(scope="module")
def mods(pkg):
    m = lambda n: importlib.import_module(f"dlpoly2-mbpol-heat_amd.{n}")
    return m("engine"), m("waterbox")


def configurations(pkg, wb, kind):
    if kind == "cluster":
        base = wb.make_cluster(30, seed=8)
        e, w = pkg.elec_params.for_cluster(base.L), pkg.water_params.for_cluster(base.L)
    else:
        base = wb.box_64()
        e, w = pkg.elec_params.for_box(base.L), pkg.water_params.for_box(base.L)
    boxes = [base] + [wb.displaced(base, 0.03, seed=70 + i) for i in range(4)]
    if kind == "cluster":
        for b in boxes:
            b.imcon = 0
    return boxes, e, w


def same(ra, sa, rb, sb):
    for k in ("eng", "vir", "eng_elec", "dipole_iterations"):
        assert ra[k] == rb[k], k
    for k in ("stress", "jtot", "jkin", "jpot"):
        assert np.array_equal(ra[k], rb[k]), k
    assert np.array_equal(sa.forces(), sb.forces())
    da, db = sa.ctx.elec_details(), sb.ctx.elec_details()
    for k in DETAILS:
        assert np.array_equal(da[k], db[k]), k


@pytest.mark.parametrize("kind", ["ksum", "mesh", "cluster"])
@pytest.mark.parametrize("solver", [0, 1])
def test_replays_equal_kernel_by_kernel(pkg, mods, kind, solver):
    import torch
    engine, wb = mods
    boxes, e, w = configurations(pkg, wb, kind)
    recip = ("spme", 6, (32, 32, 32)) if kind == "mesh" else None
    eager = engine.Slot(boxes[0], solver=solver, elec=e, water=w, recip=recip, fused=False)
    eager.ctx.set_graphs(False)
    graph = engine.Slot(boxes[0], solver=solver, elec=e, water=w, recip=recip, fused=False)
    print(f"graph replays: {kind}, solver {solver}, {len(boxes)} configurations", flush=True)
    scratch = torch.zeros(1 << 16, dtype=torch.float64, device="cuda")
    for rep in range(2):
        for b in boxes:
            eager.load(b)
            graph.load(b)
            ra = eager.evaluate()
            scratch.add_(1.0)                  # eager work of the program between the replays
            torch.cuda.synchronize()
            rb = graph.evaluate()
            same(ra, eager, rb, graph)
    caps, reps = graph.ctx.graph_counts()
    assert eager.ctx.graph_counts() == (0, 0)
    # the first evaluation ran kernel by kernel; the others replayed (a new iteration count
    # captures again)
    assert caps >= 1 and reps == 2 * len(boxes) - 1
    print(f"{kind}, solver {solver}: {caps} capture(s), {reps} replays", flush=True)


def test_short_graph_continues_kernel_by_kernel(pkg, mods, monkeypatch):
    engine, wb = mods
    boxes, e, w = configurations(pkg, wb, "ksum")
    eager = engine.Slot(boxes[0], solver=1, elec=e, water=w, fused=False)
    eager.ctx.set_graphs(False)
    monkeypatch.setenv("MBGPU_GRAPH_BATCH", "3")
    graph = engine.Slot(boxes[0], solver=1, elec=e, water=w, fused=False)
    print("graph of 3 solver iterations, continued kernel by kernel", flush=True)
    for b in boxes[:3]:
        eager.load(b)
        graph.load(b)
        ra, rb = eager.evaluate(), graph.evaluate()
        assert ra["dipole_iterations"] > 3
        same(ra, eager, rb, graph)
    assert graph.ctx.graph_counts()[1] == 2


def test_new_parameters_capture_again(pkg, mods):
    engine, wb = mods
    boxes, e, w = configurations(pkg, wb, "ksum")
    e2 = pkg.elec_params.ElecParams(rcut=5.5, alpha=0.6, kmax=(10, 10, 10))
    eager = engine.Slot(boxes[0], solver=1, elec=e, water=w, fused=False)
    eager.ctx.set_graphs(False)
    graph = engine.Slot(boxes[0], solver=1, elec=e, water=w, fused=False)
    print("graph replay across a change of electrostatics parameters", flush=True)
    for params in (e, e2, e):
        eager.set_params(params, w)
        graph.set_params(params, w)
        for _ in range(3):
            ra, rb = eager.evaluate(), graph.evaluate()
            same(ra, eager, rb, graph)
    # three parameter sets: one capture each (more if a set's first solve outgrew the batch
    # carried over from the previous set), every evaluation after the first a replay
    caps, reps = graph.ctx.graph_counts()
    assert 3 <= caps <= 6 and reps == 8
    with pytest.raises(pkg.MbgpuError, match="graphs"):
        pkg.mbgpu._check(pkg.mbgpu.load_library().mbgpu_set_graphs(graph.ctx._h, 2), "mbgpu_set_graphs")


@pytest.mark.parametrize("kind", ["ksum", "mesh", "cluster"])
@pytest.mark.parametrize("solver", [0, 1])
def test_complete_evaluation_in_one_call(pkg, mods, kind, solver, monkeypatch):
    import torch
    engine, wb = mods
    boxes, e, w = configurations(pkg, wb, kind)
    recip = ("spme", 6, (32, 32, 32)) if kind == "mesh" else None
    parts = engine.Slot(boxes[0], solver=solver, elec=e, water=w, recip=recip, fused=False)
    parts.ctx.set_graphs(False)
    one = engine.Slot(boxes[0], solver=solver, elec=e, water=w, recip=recip)
    print(f"one-call evaluation: {kind}, solver {solver}", flush=True)
    scratch = torch.zeros(1 << 16, dtype=torch.float64, device="cuda")
    for rep in range(2):
        for b in boxes:
            parts.load(b)
            one.load(b)
            ra = parts.evaluate()
            scratch.add_(1.0)
            torch.cuda.synchronize()
            rb = one.evaluate()
            same(ra, parts, rb, one)
    caps, reps = one.ctx.evaluate_graph_counts()
    # the first call sizes the buffers through the separate entries; every later one replays
    assert caps >= 1 and reps == 2 * len(boxes) - 1
    print(f"{kind}, solver {solver}: {caps} capture(s), {reps} replays of the complete evaluation", flush=True)
    # the same single call without graphs (one enqueue, one read-back)
    monkeypatch.setenv("MBGPU_GRAPHS", "0")
    plain = engine.Slot(boxes[0], solver=solver, elec=e, water=w, recip=recip)
    for b in boxes[:3]:
        parts.load(b)
        plain.load(b)
        ra, rb = parts.evaluate(), plain.evaluate()
        same(ra, parts, rb, plain)
    assert plain.ctx.evaluate_graph_counts() == (0, 0)


def test_one_call_falls_back_when_the_graph_is_short(pkg, mods, monkeypatch):
    engine, wb = mods
    boxes, e, w = configurations(pkg, wb, "ksum")
    parts = engine.Slot(boxes[0], solver=1, elec=e, water=w, fused=False)
    parts.ctx.set_graphs(False)
    monkeypatch.setenv("MBGPU_GRAPH_BATCH", "3")
    one = engine.Slot(boxes[0], solver=1, elec=e, water=w)
    print("one-call evaluation with a 3-iteration graph: the separate entries finish", flush=True)
    for b in boxes[:3]:
        parts.load(b)
        one.load(b)
        ra, rb = parts.evaluate(), one.evaluate()
        assert ra["dipole_iterations"] > 3
        same(ra, parts, rb, one)
    assert one.ctx.evaluate_graph_counts()[1] == 2


def test_one_call_fallback_starts_from_the_callers_dipoles(pkg, mods, monkeypatch):
    import torch
    engine, wb = mods
    boxes, e, w = configurations(pkg, wb, "ksum")
    monkeypatch.setenv("MBGPU_GRAPH_BATCH", "3")
    parts = engine.Slot(boxes[0], solver=1, elec=e, water=w, fused=False)
    parts.ctx.set_graphs(False)
    one = engine.Slot(boxes[0], solver=1, elec=e, water=w)
    # start dipoles: the converged dipoles of another configuration
    d0 = torch.zeros(3 * boxes[0].natms, dtype=torch.float64, device="cuda")
    b4 = boxes[4]
    parts.load(b4)
    torch.cuda.synchronize()
    parts.ctx.polar_electrostatics_dev(b4.imcon, b4.cell, parts.eparams, parts.x.data_ptr(), parts.y.data_ptr(),
                                       parts.z.data_ptr(), d0.data_ptr())
    print("one-call evaluation from given start dipoles, falling back after a 3-iteration graph", flush=True)
    for bx in boxes[:3]:
        parts.load(bx)
        one.load(bx)
        da, db = d0.clone(), d0.clone()
        torch.cuda.synchronize()
        ee, _, _, ita = parts.ctx.polar_electrostatics_dev(bx.imcon, bx.cell, parts.eparams, parts.x.data_ptr(),
                                                           parts.y.data_ptr(), parts.z.data_ptr(), da.data_ptr())
        o = one.ctx.evaluate_dev(bx.imcon, bx.cell, 3 * bx.nw, one.lst.data_ptr(), one.eparams, one.wparams,
                                 one.x.data_ptr(), one.y.data_ptr(), one.z.data_ptr(), db.data_ptr(),
                                 one.v[0].data_ptr(), one.v[1].data_ptr(), one.v[2].data_ptr(), one.ek.data_ptr())
        torch.cuda.synchronize()
        assert ita > 3 and o[49] == 0.0
        assert o[11] == ee and int(o[22]) == ita
        assert torch.equal(da, db)
