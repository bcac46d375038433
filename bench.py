#!/usr/bin/env python3
"""Benchmark: MB-pol force evaluations per second at 512 waters on one MI355X.

One evaluation = one full force evaluation of a 512-water box (L = 24.84 A, 8^3 lattice,
seeded synthetic configuration) with positions, velocities and per-atom kinetic energies
resident in HBM before the timed region:
  * MB-pol 2-body + 3-body forces with the per-atom energy / force-matrix decomposition
    (lists, polynomials, per-atom accumulation);
  * the polarizable Ewald electrostatics (M sites, DMS charges, induced dipoles solved
    from zero start values every step to the reference's stopping test, tolerance 1e-12 D,
    by conjugate gradients (default) or the reference's SOR with omega = 0.3; reciprocal
    space kmax = 16, alpha = 0.3 1/A, rcut = 9 A), with their per-atom energies,
    reciprocal stresses and force-matrix contractions;
  * the Partridge-Schwenke monomer (bonds + coupling terms), the TT-damped dispersion
    within 9 A and the long-range dispersion correction, with their per-atom terms;
  * the Green-Kubo heat flux over all of them.
The read-back of the energy, virial and stress scalars is inside the step.

Legs (all timed over --steps steps after --warmup):
  * plain FP64, one configuration at a time (one step = one evaluation): through the four
    separate entries in the reference's order, with the time split per part and the roofline
    of the dominant kernel, then (that context released) through engine.Slot (one library
    call per evaluation, graph replays);
  * batched FP64: --batch configurations (different seeds) evaluated concurrently, one
    library context, HIP stream and host thread each (one step = one evaluation of each);
  * batched mixed: the same with the 2B/3B polynomials in FP32, with its force, energy,
    virial and heat-flux error against the FP64 leg on the same configuration.
`value` is the fastest leg whose RMS force error against FP64 is <= 1e-5 kcal/mol/A.

The legs after them: NVE on the device (with the heat-current autocorrelation), a ring
polymer of 8 beads of 256 waters, serving of PIMD beads and isolated clusters with p50/p99
latency (FP64, then FP32 polynomials), the 2048-water box by the k-sum and the particle mesh,
and the 2048-water relaxation + NVE workflow.  Prints one
JSON line: metric, value, unit, timing and configuration, a `roofline` object for the
dominant kernel (achieved vs peak FLOP/s, HBM traffic per launch) and a `cpu_baseline`
object (the CPU path timed on this host).
"""
import argparse
import importlib
import json
import os
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "force evals/s @512 H2O, mixed vs FP64 (same GPU, ΔF≤1e-5) + p99 serve latency"
FLOP_DIMER = 31600      # algorithmic FLOP per 2B dimer (value + gradient)
FLOP_TRIMER = 113100    # algorithmic FLOP per 3B trimer (value + gradient)
PEAK_FP64 = 78.6        # TFLOP/s, MI355X FP64 (vector and matrix peaks are equal)
FLOP_SFAC = 12         # FLOP per (k vector, site) of the Ewald structure factors: the six multiply-adds of
                       # mu_a cos(k.r) and mu_a sin(k.r) into S(k); the phase recurrence is not counted
# HBM bytes per launch of k_el_sfac<false> at this workload, from rocprofv3 --pmc FETCH_SIZE /
# WRITE_SIZE (KiB per dispatch) scaled by the same counters on a plain 512 MiB copy (FETCH x 2,
# WRITE x 1); profiles/k_el_sfac_bench512_traffic.md
SFAC_TRAFFIC_BYTES = (4190.5 * 2.0 + 7486.5 * 1.0) * 1024


def _cpu_eval(seed_and_budget):
    """One or more complete 512-water evaluations by the oracle's C restatement on this
    process's core; returns (evaluations, seconds)."""
    seed, budget = seed_and_budget
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import numpy as np
    import oracle_lib

    pkg = importlib.import_module("dlpoly2-mbpol-heat_amd")
    box = pkg.waterbox.make_waterbox(512, 24.84, 8, seed)
    ek = 0.5 * box.mass * (box.vel ** 2).sum(1)
    ep = pkg.elec_params.for_box(box.L)
    wp = pkg.water_params.for_box(box.L)
    nr = box.nreal
    n, t0 = 0, time.perf_counter()
    while True:
        r = oracle_lib.mbpol_forces(box, compat_inplace=1, lists=False)
        el = oracle_lib.elec_forces(box, ep)
        wt = oracle_lib.water_terms(box, wp)
        hw = r["heat_w"].copy()
        hw[:nr] += el["heat_w"] + wt["heat_w"]
        e = r["atomic_energy"].copy()
        e[:nr] += el["atomic_energy"] + wt["atomic_energy"]
        sv = np.zeros((box.natms, 9))
        sv[:nr] = el["s_ew1p"]
        oracle_lib.heat_flux_tensor(box, hw.reshape(-1), box.vel, ek, e, sv.reshape(-1).copy())
        n += 1
        dt = time.perf_counter() - t0
        if dt > budget:
            return n, dt


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cores):
    """The oracle's C restatement of the whole evaluation (MB-pol all-pairs / all-triples
    scan as mbpol.f90, Ewald electrostatics with SOR from zero dipoles as forces.f, water
    terms, dense decomposition and heat flux), FP64, on this host's cores: one evaluation
    on one core, then one evaluation per core on `cores` worker processes at once (each a
    different configuration, as the GPU's batched leg).  Runs in a child process of the
    benchmark that never touches the GPU, so that its workers are plain forks."""
    import multiprocessing as mp
    n1, t1 = _cpu_eval((2, 0.0))
    out = {"value": n1 / t1, "unit": "force evals/s", "cores": 1, "kind": "port",
           "cpu_model": cpu_model(), "sample": f"{n1} complete 512-water evaluation(s) on one core, {t1:.1f} s"}
    if cores > 1:
        t0 = time.perf_counter()
        with mp.get_context("fork").Pool(cores) as pool:
            res = pool.map(_cpu_eval, [(100 + i, 0.0) for i in range(cores)])
        wall = time.perf_counter() - t0
        nall = sum(r[0] for r in res)
        out["all_cores"] = {"value": nall / wall, "unit": "force evals/s", "cores": cores,
                            "sample": f"{nall} complete 512-water evaluations, one per worker process at once, "
                                      f"{wall:.1f} s wall"}
    out["what"] = ("the oracle's C restatement of the reference's algorithms (MB-pol all-pairs / all-triples scan, "
                   "plain Ewald with SOR from zero dipoles, water terms, dense force-matrix heat flux), FP64; the "
                   "reference itself is not run on this host")
    return out


def large_box_leg(pkg, torch, reps):
    """The 2048-water box (L = 39.43 A, alpha 0.3 1/A, rcut 9 A): one complete force
    evaluation with the heat flux (engine.Slot: MB-pol, electrostatics with the CG solver
    from zero dipoles, water terms, J) with the Ewald sum over k at kmax 24 (33 530 k
    vectors), and with the particle-mesh reciprocal space (order 6 on 48^3); the mesh's
    error against the k-sum in forces, energy and heat current."""
    import numpy as np
    engine = importlib.import_module("dlpoly2-mbpol-heat_amd.engine")
    box = pkg.waterbox.box_2048()
    out, res = {}, {}
    for name, recip in (("ewald_kmax24", None), ("spme_order6_48", ("spme", 6, (48, 48, 48)))):
        slot = engine.Slot(box, recip=recip)
        for _ in range(2):
            slot.evaluate()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = slot.evaluate()
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / reps
        res[name] = (r, slot.forces())
        out[name] = {"ms_per_eval": ms, "evals_per_s": 1e3 / ms, "dipole_iterations": r["dipole_iterations"]}
        del slot
    (re_, fe), (rs, fs) = res["ewald_kmax24"], res["spme_order6_48"]
    out["spme_vs_ewald"] = {
        "rel_force": float(np.abs(fs - fe).max() / np.abs(fe).max()),
        "rms_dF_kcal_mol_A": float(np.sqrt(((fs - fe) ** 2).mean()) / 418.4),
        "rel_energy_of_elec": float(abs(rs["eng"] - re_["eng"]) / abs(re_["eng_elec"])),
        "rel_J": float(np.abs(rs["jtot"] - re_["jtot"]).max() / np.abs(re_["jtot"]).max()),
        "speedup": out["ewald_kmax24"]["ms_per_eval"] / out["spme_order6_48"]["ms_per_eval"]}
    out["what"] = ("2048 H2O, complete evaluation with the heat flux, host-timed over " + str(reps) +
                   " evaluations after 2 warm-up, reciprocal space by the k-sum and by the particle mesh")
    return out


def large_md_leg(pkg, args):
    """The large-box workflow (md.large_box_run): 2048 waters, particle mesh order 6 on 48^3,
    --large-md-relax FP64 relaxation steps, --large-md-steps mixed-precision NVE steps at
    0.2 fs with J every step, an FP64 shadow evaluation every 100 steps, the J
    autocorrelation on the device."""
    md = importlib.import_module("dlpoly2-mbpol-heat_amd.md")
    box = pkg.waterbox.box_2048()

    def progress(stage, step, d):
        print(f"large-box {stage} step {step}: " + ", ".join(f"{k} {v:.3g}" for k, v in d.items()
                                                               if isinstance(v, float)), file=sys.stderr, flush=True)

    r = md.large_box_run(box, relax_steps=args.large_md_relax, nve_steps=args.large_md_steps, shadow_every=100,
                         progress=progress)
    keep = ("waters", "relax_steps", "nve_steps", "shadow_every", "dt_fs", "recip", "relax_ms_per_step",
            "relax_temperature_last", "nve_ms_per_step", "drift_per_ps_over_KE0", "max_abs_dE_over_KE0",
            "dipole_iterations_median", "max_rms_dF_kcal_mol_A", "max_rel_dJ", "acf_ms")
    out = {k: r[k] for k in keep}
    out["acf_lags"] = int(len(r["acf"]))
    out["shadow_last"] = r["shadow"][-1]
    out["what"] = ("2048 H2O (13^3 lattice, L = 39.43 A), FP64 relaxation with Berendsen scaling to 300 K, then NVE "
                   "with the 2B/3B polynomials in FP32 and the heat current every step; every 100 steps the same "
                   "configuration evaluated in FP64 (forces, energy, J compared); J autocorrelation on the device")
    return out


def serve_leg(pkg, args, precision=0):
    """Config-3 style serving mix through serve.Server; returns latency and throughput.
    precision 0: FP64; 3: FP32 polynomials (the mixed-precision path)."""
    import numpy as np
    serve = importlib.import_module("dlpoly2-mbpol-heat_amd.serve")
    wb, ep, wp = pkg.waterbox, pkg.elec_params, pkg.water_params
    rng = np.random.default_rng(9)
    base = wb.box_256()
    beads = [wb.displaced(base, 0.05, seed=500 + i) for i in range(32)]
    sizes, clusters = [], []
    if args.serve_mix == "mixed":
        sizes = [int(n) for n in np.exp(rng.uniform(np.log(3), np.log(1024), 24))]
        clusters = [wb.make_cluster(n, seed=600 + i) for i, n in enumerate(sizes)]
    jobs = []
    for i in range(args.serve_requests):
        if clusters and rng.random() < 0.5:
            c = clusters[int(rng.integers(len(clusters)))]
            jobs.append((c, ep.for_cluster(c.L), wp.for_cluster(c.L), "cluster"))
        else:
            b = beads[i % len(beads)]
            jobs.append((b, ep.for_box(b.L), wp.for_box(b.L), "bead"))
    srv = serve.Server(workers=max(1, args.batch), precision=precision, solver=1 if args.solver == "cg" else 0)
    # warm the pool, untimed: one request per worker at once, so every worker holds a slot,
    # then one request of every cluster size (slots of new sizes are still created and
    # evicted during the timed run, as other workers use the same size at once)
    warm = [srv.submit(*j) for j in jobs[:max(1, args.batch)]]
    warm += [srv.submit(c, ep.for_cluster(c.L), wp.for_cluster(c.L), "cluster") for c in clusters]
    for r in warm:
        r.done.wait()
    reqs, wall = serve.run_poisson(srv, jobs, args.serve_rate, seed=3)
    srv.close()
    out = serve.latency_summary(reqs)
    for tag in ("bead", "cluster"):
        sub = [r for r in reqs if r.tag == tag]
        s = serve.latency_summary(sub) if sub else {}
        if "p50_ms" in s:
            out[f"{tag}_p50_ms"], out[f"{tag}_p99_ms"] = s["p50_ms"], s["p99_ms"]
    out.update(arrival_rate=args.serve_rate, wall_s=wall, evals_per_s=len(reqs) / wall, workers=max(1, args.batch),
               slots_created=srv.slots_created, slots_evicted=srv.slots_evicted,
               mix=args.serve_mix, cluster_sizes=sizes, precision="fp64" if precision == 0 else "fp32 polynomials",
               what="requests drawn from 32 PIMD beads x 256 H2O (L=19.72 A, N(0,0.05 A) displacements)"
                    + (" and, half of them, isolated clusters of n in [3,1024] waters, n log-uniform (random "
                       "sequential addition at 0.997 g/cm^3, O-O >= 2.6 A; isolated: imcon 0, direct sums, no images)"
                       if clusters else "")
                    + ", Poisson arrivals, continuous batching over the workers' slots; latency = arrival to "
                      "forces on the host")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-only", type=int, default=0, metavar="CORES",
                    help="only run the CPU baseline on CORES worker processes and print it (used by the benchmark "
                         "itself, in a child process)")
    ap.add_argument("--solver", choices=["sor", "cg"], default="cg",
                    help="induced-dipole solver: the reference's SOR or conjugate gradients to the same test")
    ap.add_argument("--batch", type=int, default=16,
                    help="configurations evaluated concurrently in the batched legs (0 skips them)")
    ap.add_argument("--nve-steps", type=int, default=1000,
                    help="steps of the NVE leg (512 waters, dt 0.2 fs, heat current every step; 0 skips it)")
    ap.add_argument("--pimd-steps", type=int, default=20,
                    help="ring-polymer steps (8 beads x 256 H2O, 0.2 fs) of the PIMD leg; 0 skips it")
    ap.add_argument("--serve-requests", type=int, default=10000,
                    help="requests of the serving leg (Poisson arrivals at --serve-rate; 0 skips it)")
    ap.add_argument("--hw-queues", type=int, default=0,
                    help="GPU_MAX_HW_QUEUES for the run (0: HIP's default, 4)")
    ap.add_argument("--no-overlap", action="store_true",
                    help="evaluate the electrostatics after MB-pol and the water terms instead of beside them")
    ap.add_argument("--only-large", action="store_true",
                    help="run the 2048-water leg alone and print its record (for profiling)")
    ap.add_argument("--large-reps", type=int, default=5,
                    help="evaluations per reciprocal method of the 2048-water leg (0: skip)")
    ap.add_argument("--large-md-steps", type=int, default=1000,
                    help="mixed-precision NVE steps of the large-box workflow (2048 waters, mesh 48^3, order 6, "
                         "FP64 shadow evaluation every 100 steps, J autocorrelation on the device); the full "
                         "workflow is 10000; 0 skips it")
    ap.add_argument("--large-md-relax", type=int, default=500,
                    help="FP64 relaxation steps (Berendsen scaling to 300 K) before the large-box NVE run")
    ap.add_argument("--only-large-md", action="store_true",
                    help="run the large-box workflow alone and print its record")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--no-serve-mixed", action="store_true",
                    help="serve in FP64 only (by default the stream is served again with FP32 polynomials)")
    ap.add_argument("--serve-mix", choices=["beads", "mixed"], default="mixed",
                    help="beads: PIMD beads of 256 waters only (same-size slots); mixed: beads and isolated "
                         "clusters of 3-1024 waters, half and half")
    ap.add_argument("--copy-calibration", type=int, metavar="MIB", default=0,
                    help="only run a plain device copy of MIB MiB five times (scale reference for the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "launches": 5}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every run of this library "
              "uses a single GPU; how the path would shard is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    if args.cpu_only:
        # child process of a benchmark run: the CPU baseline alone, never touching the GPU
        print(json.dumps(cpu_baseline(args.cpu_only)), flush=True)
        return
    if args.hw_queues:
        # HIP's default is 4 hardware queues, which the concurrent legs measured fastest with
        # (16 configurations: 1060 evaluations/s against 925 with 8 queues and 940 with 16);
        # must be set before the HIP runtime starts
        os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
    import numpy as np
    import torch

    pkg = importlib.import_module("dlpoly2-mbpol-heat_amd")
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.only_large:
        print(json.dumps({"large_box": large_box_leg(pkg, torch, max(1, args.large_reps))}), flush=True)
        return
    if args.only_large_md:
        print(json.dumps({"large_box_md": large_md_leg(pkg, args)}), flush=True)
        return
    box = pkg.waterbox.box_512()
    ctx = pkg.Context(box.nw, box.natms)
    ep = pkg.elec_params.for_box(box.L)
    eparams = pkg.ElecParams.from_params(ep, distribute_msite=1, solver=1 if args.solver == "cg" else 0)
    wparams = pkg.WaterParams.from_params(pkg.water_params.for_box(box.L))
    T = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    x, y, z = T(box.x), T(box.y), T(box.z)
    lst = T(box.listttm2, torch.int32)
    dip = torch.zeros(3 * box.natms, dtype=torch.float64, device=dev)
    fsite = torch.zeros(3 * box.natms, dtype=torch.float64, device=dev)
    fw = torch.zeros(3, box.natms, dtype=torch.float64, device=dev)
    vx, vy, vz = T(box.vel[:box.nreal, 0]), T(box.vel[:box.nreal, 1]), T(box.vel[:box.nreal, 2])
    ek = T(0.5 * box.mass[:box.nreal] * (box.vel[:box.nreal] ** 2).sum(1))
    torch.cuda.synchronize()

    parts = {"mbpol": 0.0, "elec": 0.0, "water": 0.0, "heat": 0.0}
    iters = []

    def step(timed=False):
        t0 = time.perf_counter()
        e, v, s = ctx.mbpol_forces_dev(box.imcon, box.cell, lst.data_ptr(), x.data_ptr(), y.data_ptr(),
                                       z.data_ptr())
        t1 = time.perf_counter()
        dip.zero_()                               # zero start dipoles (torch's stream) ...
        torch.cuda.current_stream().synchronize()  # ... complete before the library's stream reads them
        ee, ve, se, it = ctx.polar_electrostatics_dev(box.imcon, box.cell, eparams, x.data_ptr(), y.data_ptr(),
                                                      z.data_ptr(), dip.data_ptr(), fsite.data_ptr())
        t2 = time.perf_counter()
        ew, vw, sw = ctx.water_terms_dev(box.imcon, box.cell, wparams, x.data_ptr(), y.data_ptr(), z.data_ptr(),
                                         fw[0].data_ptr(), fw[1].data_ptr(), fw[2].data_ptr())
        t3 = time.perf_counter()
        jk, jp, jt = ctx.heat_flux_dev(vx.data_ptr(), vy.data_ptr(), vz.data_ptr(), ek.data_ptr())
        t4 = time.perf_counter()
        if timed:
            parts["mbpol"] += t1 - t0
            parts["elec"] += t2 - t1
            parts["water"] += t3 - t2
            parts["heat"] += t4 - t3
            iters.append(it)
        return e + ee + ew.sum(), jt

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = 1e3 * dt / args.steps

    d, t = ctx.lists()
    nd, nt = len(d), len(t)
    kvec = int(ctx.elec_details()["k_vectors"])
    reps = max(5, min(args.steps, 20))
    _, ms_k3 = ctx.time_forces(reps)
    ms_sf, flop_sf_lib = ctx.time_elec(reps)
    tf_k3 = nt * FLOP_TRIMER / (ms_k3 * 1e-3) / 1e12
    # algorithmic work of one launch: FLOP_SFAC per (k vector, polarizable site)
    flop_sf = FLOP_SFAC * kvec * box.nreal
    achieved = flop_sf / (ms_sf * 1e-3) / 1e12
    n_sfac = (int(np.median(iters)) + 2) if iters else None
    single = args.steps / dt

    # the same configuration through engine.Slot: one library call per evaluation
    # (mbgpu_evaluate_dev, graph replays, electrostatics beside MB-pol and the water terms),
    # after the separate entries' context is released: with HIP's 4 hardware queues a
    # context whose two streams are created while another context's exist measured 1.72 ms
    # instead of 1.48 ms per evaluation (how the runtime maps streams to queues)
    ctx.close()
    engine = importlib.import_module("dlpoly2-mbpol-heat_amd.engine")
    one = engine.Slot(box, solver=1 if args.solver == "cg" else 0)
    for _ in range(max(2, args.warmup)):
        one.evaluate()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        one.evaluate()
    torch.cuda.synchronize()
    single_one_call = args.steps / (time.perf_counter() - t0)
    one.close()
    one.ctx.close()
    del one

    # batched legs: B independent 512-water configurations (different seeds), one context,
    # stream and host thread each, evaluated concurrently; FP64, then FP32 polynomials
    batched = {}
    if args.batch > 0:
        engine = importlib.import_module("dlpoly2-mbpol-heat_amd.engine")
        solver = 1 if args.solver == "cg" else 0
        boxes = [pkg.waterbox.make_waterbox(512, 24.84, 8, 100 + i) for i in range(args.batch)]
        for name, prec in (("fp64", 0), ("mixed", 3)):
            slots = [engine.Slot(b, precision=prec, solver=solver, overlap=not args.no_overlap) for b in boxes]
            tb, res = engine.run_concurrently(slots, args.steps, warmup=args.warmup)
            batched[name] = dict(evals_per_s=args.batch * args.steps / tb, ms_per_step=1e3 * tb / args.steps,
                                 forces=slots[0].forces(), result=res[0][-1], nw=boxes[0].nw)
            for sl in slots:             # release the contexts (their streams and memory) now
                sl.close()
                sl.ctx.close()
            del slots
        f64, fmx = batched["fp64"]["forces"] / 418.4, batched["mixed"]["forces"] / 418.4
        df = fmx - f64
        r64, rmx = batched["fp64"]["result"], batched["mixed"]["result"]
        batched["mixed"]["rms_dF"] = float(np.sqrt((df ** 2).mean()))
        batched["mixed"]["max_dF"] = float(np.abs(df).max())
        batched["mixed"]["dE_per_molecule"] = float((rmx["eng"] - r64["eng"]) / 418.4 / boxes[0].nw)
        batched["mixed"]["rel_dJ"] = float(np.abs(rmx["jtot"] - r64["jtot"]).max() / np.abs(r64["jtot"]).max())
        batched["mixed"]["rel_dvir"] = float(abs(rmx["vir"] - r64["vir"]) / abs(r64["vir"]))
    # NVE leg: velocity Verlet on the device, a complete force evaluation and the heat flux
    # every step, dipoles carried between steps (configuration of the plain FP64 path)
    nve_out = None
    if args.nve_steps > 0:
        md = importlib.import_module("dlpoly2-mbpol-heat_amd.md")
        engine = importlib.import_module("dlpoly2-mbpol-heat_amd.engine")
        nve = md.NVE(engine.Slot(box, solver=1 if args.solver == "cg" else 0, dedicated_queues=True), 0.2,
                     extrapolate=2)
        e0, k0 = nve.pot + nve.kin, nve.kin
        tot, its, jser = [], [], []
        torch.cuda.synchronize()
        t0n = time.perf_counter()
        for _ in range(args.nve_steps):
            r = nve.step()
            tot.append(r["tot"])
            its.append(r["iterations"])
            jser.append(r["jtot"])
        torch.cuda.synchronize()
        tn = time.perf_counter() - t0n
        # autocorrelation of the run's J_tot series on the device (its cost; a 0.2 ps run gives
        # no converged conductivity)
        ntau = max(1, min(200, args.nve_steps // 2))
        t0g = time.perf_counter()
        gk = md.green_kubo(nve.slot.ctx, np.array(jser), 0.2e-3, box.L ** 3, 300.0, ntau)
        tg = time.perf_counter() - t0g
        nve.slot.close()                 # release the context before the next leg creates its own
        nve.slot.ctx.close()
        tot = np.array(tot)
        slope = np.polyfit(np.arange(1, len(tot) + 1) * 0.2e-3, tot - e0, 1)[0] if len(tot) > 1 else 0.0
        nve_out = {"steps": args.nve_steps, "dt_fs": 0.2, "ms_per_step": 1e3 * tn / args.nve_steps,
                   "max_abs_dE_over_KE0": float(np.abs(tot - e0).max() / k0),
                   "drift_per_ps_over_KE0": float(slope / k0), "dipole_iterations_median": float(np.median(its)),
                   "heat_acf": {"lags": ntau, "ms": 1e3 * tg, "what": "autocorrelation of the run's J_tot on the device"},
                   "what": "512 H2O from the 300 K start, heat current every step, induced dipoles started from "
                           "the parabola through the last three steps' (md.NVE extrapolate=2)"}
    # PIMD leg: a ring polymer of 8 beads of 256 waters (300 K), every bead evaluated with the
    # complete force field on its own context (overlapping), springs and normal modes on the
    # device, velocity Verlet of the ring-polymer Hamiltonian
    pimd_out = None
    if args.pimd_steps > 0:
        ring = importlib.import_module("dlpoly2-mbpol-heat_amd.ring")
        base = pkg.waterbox.box_256()
        beads = [pkg.waterbox.displaced(base, 0.02, seed=700 + k) for k in range(8)]
        rp = ring.RingPolymer(beads, 300.0, solver=1 if args.solver == "cg" else 0)
        h0, k0 = rp.energy(), rp.kinetic()
        torch.cuda.synchronize()
        t0p = time.perf_counter()
        hs = [rp.step(0.2) for _ in range(args.pimd_steps)]
        torch.cuda.synchronize()
        tp = time.perf_counter() - t0p
        rp.close()
        pimd_out = {"beads": 8, "waters": 256, "steps": args.pimd_steps, "dt_fs": 0.2,
                    "ms_per_step": 1e3 * tp / args.pimd_steps,
                    "bead_evals_per_s": 8 * args.pimd_steps / tp,
                    "max_abs_dH_over_K0": float(np.max(np.abs(np.array(hs) - h0)) / k0),
                    "what": "ring polymer at 300 K, normal modes and springs on the device, beads overlapping"}
    # serving leg: the mixed stream of PIMD beads (256 waters, L = 19.72 A, one configuration
    # displaced by N(0, 0.05 A) per bead) and isolated clusters (n log-uniform in [3, 1024]),
    # Poisson arrivals, continuous batching over --batch workers
    serving = serving_mixed = None
    if args.serve_requests > 0:
        serving = serve_leg(pkg, args)
        if not args.no_serve_mixed:
            serving_mixed = serve_leg(pkg, args, precision=3)
    large_out = large_box_leg(pkg, torch, args.large_reps) if args.large_reps > 0 else None
    large_md = large_md_leg(pkg, args) if args.large_md_steps > 0 else None
    # headline: the fastest leg whose forces are within 1e-5 kcal/mol/A RMS of FP64
    lead = ("fp64 single", single, ms, "f64")
    if single_one_call > single:
        lead = ("fp64 single (one call)", single_one_call, 1e3 / single_one_call, "f64")
    if "fp64" in batched and batched["fp64"]["evals_per_s"] > single:
        lead = (f"fp64 batched x{args.batch}", batched["fp64"]["evals_per_s"], batched["fp64"]["ms_per_step"], "f64")
    if "mixed" in batched and batched["mixed"]["rms_dF"] <= 1e-5 and batched["mixed"]["evals_per_s"] > lead[1]:
        lead = (f"mixed batched x{args.batch}", batched["mixed"]["evals_per_s"], batched["mixed"]["ms_per_step"],
                "f64+f32")
    legs = {"fp64_single_evals_per_s": single, "fp64_single_one_call_evals_per_s": single_one_call}
    for name, b in batched.items():
        legs[f"{name}_batched_evals_per_s"] = b["evals_per_s"]
        legs[f"{name}_batched_speedup_vs_fp64_single"] = b["evals_per_s"] / single
    if serving is not None:
        legs["serving"] = serving
    if serving_mixed is not None:
        legs["serving_mixed"] = serving_mixed
    if nve_out is not None:
        legs["nve"] = nve_out
    if pimd_out is not None:
        legs["pimd"] = pimd_out
    if large_out is not None:
        legs["large_box"] = large_out
    if large_md is not None:
        legs["large_box_md"] = large_md
    if "mixed" in batched:
        m = batched["mixed"]
        legs["mixed_vs_fp64"] = {"rms_dF_kcal_mol_A": m["rms_dF"], "max_dF_kcal_mol_A": m["max_dF"],
                                 "dE_kcal_mol_per_molecule": m["dE_per_molecule"], "rel_dJ": m["rel_dJ"],
                                 "rel_dvirial": m["rel_dvir"], "meets_1e-5": m["rms_dF"] <= 1e-5,
                                 "what": "2B and 3B polynomials in FP32, everything else FP64"}
    out = {
        "metric": METRIC,
        "value": lead[1],
        "unit": "force evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": lead[2],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": lead[3],
        "data": "synthetic (seeded 512-water lattice boxes, rigid-geometry perturbations, 300 K velocities)",
        "config": {"workload": "MB-pol 2B+3B + polarizable Ewald electrostatics (DMS charges, induced dipoles "
                               "from zero to 1e-12 D, kmax 16, rcut 9 A) + PS monomer + TT dispersion and "
                               "long-range tail + per-atom energy/heat-flux decomposition + Green-Kubo heat flux, "
                               "512 H2O L=24.84 A",
                   "headline_leg": lead[0], "batch": args.batch, **legs,
                   "waters": box.nw, "dimers": nd, "trimers": nt, "k_vectors": kvec, "dipole_iterations":
                   int(np.median(iters)) if iters else None, "dipole_solver": args.solver, "parallelism": "single GPU",
                   "single_ms_mbpol": 1e3 * parts["mbpol"] / args.steps,
                   "single_ms_electrostatics": 1e3 * parts["elec"] / args.steps,
                   "single_ms_water_terms": 1e3 * parts["water"] / args.steps,
                   "single_ms_heat_flux": 1e3 * parts["heat"] / args.steps,
                   "k_pip3b_ms": ms_k3, "k_pip3b_tflops": tf_k3, "k_pip3b_frac": tf_k3 / PEAK_FP64,
                   "sfac_launches_per_step": n_sfac},
        "roofline": {"bound": "fp64-valu", "engine": "FP64 vector ALU (no dense contraction, so no MFMA; peak "
                                                     "78.6 TFLOP/s)",
                     "kernel": "k_el_sfac<false> (dipole structure factors, launched once per solver iteration)",
                     "achieved": achieved, "peak": PEAK_FP64, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64,
                     "traffic": SFAC_TRAFFIC_BYTES,
                     "traffic_source": "copied from profiles/k_el_sfac_bench512_traffic.md (rocprofv3 --pmc FETCH_SIZE "
                                       "and WRITE_SIZE of this workload and build, scaled by a plain copy), bytes "
                                       "per launch; not measured in this run",
                     "kernel_ms": ms_sf, "flop_per_launch": flop_sf,
                     "flop_per_unit": FLOP_SFAC, "units_per_launch": f"{kvec} k vectors x {box.nreal} sites",
                     "frac_with_recurrence": flop_sf_lib / (ms_sf * 1e-3) / 1e12 / PEAK_FP64},
    }
    if not args.no_cpu:
        # in a child process that never touches the GPU (its all-core leg forks workers)
        ncpu = min(16, len(os.sched_getaffinity(0)))
        print(f"cpu baseline: oracle restatement, 1 core then {ncpu} worker processes", file=sys.stderr, flush=True)
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-only", str(ncpu)], capture_output=True,
                           text=True, timeout=900)
        lines = [x for x in r.stdout.splitlines() if x.startswith("{")]
        if r.returncode != 0 or not lines:
            raise RuntimeError(f"cpu baseline failed ({r.returncode}): {r.stderr[-2000:]}")
        out["cpu_baseline"] = json.loads(lines[-1])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
