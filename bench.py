#!/usr/bin/env python3
"""Benchmark: MB-pol force evaluations per second at 512 waters on one MI355X.

One step = one evaluation of the MB-pol 2-body + 3-body forces with the per-atom energy /
heat-flux decomposition and the Green-Kubo heat flux, on the 512-water box (L = 24.84 A,
8^3 lattice, seeded synthetic configuration), with positions, velocities and per-atom
kinetic energies resident in HBM before the timed region.  The step includes the
neighbour / dimer / trimer list builds, the polynomial kernels, the per-atom accumulation
and the read-back of the energy, virial and stress scalars.

Prints one JSON line (see DESIGN.md for the field definitions).
"""
import argparse
import ctypes
import importlib
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "force evals/s @512 H2O, mixed vs FP64 (same GPU, ΔF≤1e-5) + p99 serve latency"
FLOP_DIMER = 31600      # algorithmic FLOP per 2B dimer (value + gradient)
FLOP_TRIMER = 113100    # algorithmic FLOP per 3B trimer (value + gradient)
PEAK_FP64 = 78.6        # TFLOP/s, MI355X vector / matrix FP64


def cpu_baseline(box):
    """The CPU restatement of mbpol_forces + compute_heat_flux2 on this host, one core.
    When the reference's own compiled leaves are present (oracle/_ref) the polynomial
    leaves are the reference's; the driver loop is the oracle's restatement."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import numpy as np
    import oracle_lib

    lib = oracle_lib.oracle()
    ref = oracle_lib.reference_leaves()
    kind = "port"
    if ref is not None:
        lib.ora_set_leaves(ctypes.cast(ref.mbpol_2b_poly_, ctypes.c_void_p),
                           ctypes.cast(ref.mbpol_3b_poly_, ctypes.c_void_p))
        kind = "reference"
    ek = 0.5 * box.mass * (box.vel ** 2).sum(1)
    n, t0 = 0, time.perf_counter()
    while True:
        r = oracle_lib.mbpol_forces(box, compat_inplace=1, lists=False)
        oracle_lib.heat_flux_tensor(box, r["heat_w"], box.vel, ek, r["atomic_energy"])
        n += 1
        dt = time.perf_counter() - t0
        if dt > 10.0 or n >= 8:
            break
    lib.ora_set_leaves(None, None)
    what = "reference C++ leaves (oracle/_ref) under the oracle's mbpol_forces loop" if kind == "reference" \
        else "oracle C restatement (extended-precision polynomial sums)"
    return {"value": n / dt, "unit": "force evals/s", "cores": 1, "kind": kind,
            "sample": f"{n} full 512-water MB-pol 2B+3B force + decomposition + heat-flux evaluations "
                      f"(all-pairs / all-triples scan as mbpol.f90), {what}, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (multi-GPU sharding is "
              "described in DESIGN.md but not built)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    pkg = importlib.import_module("dlpoly2-mbpol-heat_amd")
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    box = pkg.waterbox.box_512()
    ctx = pkg.Context(box.nw, box.natms)
    T = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    x, y, z = T(box.x), T(box.y), T(box.z)
    lst = T(box.listttm2, torch.int32)
    vx, vy, vz = T(box.vel[:box.nreal, 0]), T(box.vel[:box.nreal, 1]), T(box.vel[:box.nreal, 2])
    ek = T(0.5 * box.mass[:box.nreal] * (box.vel[:box.nreal] ** 2).sum(1))
    torch.cuda.synchronize()

    def step():
        e, v, s = ctx.mbpol_forces_dev(box.imcon, box.cell, lst.data_ptr(), x.data_ptr(), y.data_ptr(),
                                       z.data_ptr())
        jk, jp, jt = ctx.heat_flux_dev(vx.data_ptr(), vy.data_ptr(), vz.data_ptr(), ek.data_ptr())
        return e, jt

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = 1e3 * dt / args.steps

    d, t = ctx.lists()
    nd, nt = len(d), len(t)
    _, ms_k3 = ctx.time_forces(max(5, min(args.steps, 20)))
    achieved = nt * FLOP_TRIMER / (ms_k3 * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": args.steps / dt,
        "unit": "force evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded 512-water lattice box, rigid-geometry perturbations, 300 K velocities)",
        "config": {"workload": "MB-pol 2B+3B forces + per-atom energy/heat-flux decomposition + Green-Kubo "
                               "heat flux, FP64, 512 H2O L=24.84 A (electrostatics not yet in the step)",
                   "waters": box.nw, "dimers": nd, "trimers": nt, "parallelism": "single GPU"},
        "roofline": {"bound": "valu-fp64", "kernel": "k_pip3b<double>", "achieved": achieved, "peak": PEAK_FP64,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64, "traffic": None,
                     "kernel_ms": ms_k3, "flop_per_launch": nt * FLOP_TRIMER},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(box)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
