"""Pseudo-time iteration throughput (residual + LU-SGS update) on one MI355X.

Workload: cavitating bump channel, 128x64x64 cells, k-omega SST, AUSM+-up,
MLP limiter, Merkle cavitation, FP64.  One step = one cav3d_timeinteg
(time step, boundary fill, residual, forward/backward LU-SGS sweeps,
update with EOS, eddy viscosity, norms) over all cells, with the state
resident in HBM.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# compulsory HBM bytes per cell and iteration of the FP64 SST path: state,
# properties, metrics, residual and LU-SGS work arrays read/written once
B_ALG_FP64_SST = 720.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", default="128x64x64")
    ap.add_argument("--cpu-baseline", type=int, default=1, help="1: time the oracle on the host")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the library keeps every block of a case on a single device "
              "and has no multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import importlib
    pk = importlib.import_module("3dcode_amd")
    ni, nj, nk = (int(x) for x in args.size.split("x"))
    t0 = time.time()
    case = pk.cases.config2(ni, nj, nk)
    case.params["cav_clip"] = 1
    print(f"# case {case.name} built in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    s = pk.Solver(case)
    for ib, st in enumerate(case.init):
        s.put_state(ib, *st)
    for it in range(args.warmup):
        s.timeinteg(1, it + 1, 0.0, want_time=False)
    s.synchronize()
    t1 = time.perf_counter()
    for it in range(args.steps):
        s.timeinteg(1, args.warmup + it + 1, 0.0, want_time=False)
    s.synchronize()
    t2 = time.perf_counter()
    ms = (t2 - t1) * 1e3 / args.steps
    ncell = case.ncell
    value = ncell / (ms * 1e-3) / 1e6
    l2, ijk, rm = s.residual()
    if not np.all(np.isfinite(l2)):
        raise RuntimeError("non-finite residual norms")

    # per-phase device time over a few profiled iterations (HIP events on the library's stream)
    s.profile(True)
    nprof = min(10, max(args.steps, 1))
    acc = np.zeros(8)
    for it in range(nprof):
        s.timeinteg(1, 1, 0.0, want_time=False)
        acc += s.profile(True)
    s.profile(False)
    phase = acc / nprof
    s.close()

    # dominant kernel: the forward LU-SGS sweep, one launch over all cells.  Its
    # algorithmic bytes are the forward-sweep pass of the compulsory-traffic
    # model: read pv (n) and mu_t, read the metrics (G words), write dcv (n).
    npv = case.npv
    G = 11 if npv == 9 else 10
    b_fwd = 8.0 * (2 * npv + 1) + 8.0 * G
    t_fwd = phase[6] * 1e-3
    achieved = b_fwd * ncell / t_fwd / 1e9
    it_achieved = B_ALG_FP64_SST * ncell / (phase[5] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "kernel": "k_sweep_tile<forward> (LU-SGS forward sweep, one launch over all cells)",
                "bytes_per_cell": b_fwd, "launch_ms": round(phase[6], 4),
                "iteration": {"bytes_per_cell": B_ALG_FP64_SST, "achieved": round(it_achieved, 2),
                              "frac": round(it_achieved / HBM_PEAK_GBS, 5)},
                "phase_ms": {k: round(v, 4) for k, v in zip(
                    ["timestep", "boundary", "rhs", "lusgs", "update_eddy", "total", "sweep_fwd", "sweep_bwd"], phase)}}

    cpu = None
    if args.cpu_baseline:
        from oracle import oracle as orc
        threads = min(16, os.cpu_count() or 1)
        orc.set_threads(threads)
        o = orc.Oracle(case)
        for ib, st in enumerate(case.init):
            o.put_state(ib, *st)
        c0 = time.perf_counter()
        o.timeinteg(1, 1, 0.0)
        c1 = time.perf_counter()
        cpu = {"value": round(ncell / (c1 - c0) / 1e6, 5), "unit": "Mcell-iter/s", "cores": threads,
               "kind": "port", "sample": f"one iteration of the same {args.size} case, FP64 C restatement, OpenMP"}

    out = {
        "metric": "Mcell-iter/s (pseudo-time iteration: residual + LU-SGS, FP64)",
        "value": round(value, 4), "unit": "Mcell-iter/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"cavitating bump channel {args.size} cells, k-omega SST, AUSM+-up, MLP, Merkle "
                               f"(species clip), 1 block", "cells": ncell},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
