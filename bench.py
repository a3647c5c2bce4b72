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


TRAFFIC_FILE = "profiles/hbm_traffic_bump128x64x64_fp64.txt"
TRAFFIC_SIZE = "128x64x64"


TRAFFIC_ITERS = 5          # iterations in the counter runs: 1 warm-up + 2 timed + 2 profiled
SETUP_KERNELS = ("k_dv_to_dc", "k_aos_to_soa", "k_cell_from_aos", "k_eos_cells")


def measured_traffic(size, kernel=None):
    """HBM bytes per launch of `kernel`, or per iteration over all kernels when
    kernel is None (FETCH_SIZE + WRITE_SIZE, each scaled by the copy-kernel
    calibration), from the committed rocprofv3 summary of the default
    workload; None for other workloads or without the summary."""
    path = os.path.join(ROOT, TRAFFIC_FILE)
    if size != TRAFFIC_SIZE or not os.path.exists(path):
        return None
    import re
    total = 0.0
    for line in open(path):
        # kernel names may hold spaces ("k_face<1, false, double>")
        m = re.match(r"kernel (.+) dispatches (\d+) fetch_bytes (\S+) write_bytes (\S+)\s*$", line)
        if m:
            name, n = m.group(1), int(m.group(2))
            b = float(m.group(3)) + float(m.group(4))
            if kernel is not None and name == kernel:
                return int(b)
            if name not in SETUP_KERNELS:
                total += b * n / TRAFFIC_ITERS
    return None if kernel is not None else int(total)


def pmc_summary(args, dfetch, dwrite, dcfetch, dcwrite):
    """Per-kernel HBM traffic per dispatch from rocprofv3 --pmc databases."""
    import glob
    import re
    import sqlite3
    from collections import defaultdict

    def rows(folder, counter):
        db = glob.glob(os.path.join(folder, "**", "*results.db"), recursive=True)[0]
        c = sqlite3.connect(db)
        out = defaultdict(list)
        for name, value in c.execute("select kernel_name, value from counters_collection where counter_name = ?",
                                     (counter,)):
            name = re.sub(r"\(.*", "", name.replace("(anonymous namespace)::", "")).replace("void ", "")
            out[name].append(value)
        return out

    cf, cw = rows(dcfetch, "FETCH_SIZE"), rows(dcwrite, "WRITE_SIZE")
    copy_bytes = 256 * (1 << 20)
    kf = np.mean(cf["__amd_rocclr_copyBuffer"]) * 1024.0
    kw = np.mean(cw["__amd_rocclr_copyBuffer"]) * 1024.0
    ff, fw = copy_bytes / kf, copy_bytes / kw
    bf, bw = rows(dfetch, "FETCH_SIZE"), rows(dwrite, "WRITE_SIZE")
    cmd = "rocprofv3 --pmc {c} -d <dir> -o run -- python3 bench.py {a}"
    lines = [
        "# HBM traffic per dispatch, FP64 config-2 workload (cavitating bump 128x64x64, SST, AUSM+-up, MLP, Merkle)",
        "# from four rocprofv3 counter runs on one MI355X (run from /tmp with TMPDIR=/tmp, R = repository root):",
        "#   " + cmd.format(c="FETCH_SIZE", a="--steps 2 --warmup 1 --cpu-baseline 0 --legs 0"),
        "#   " + cmd.format(c="WRITE_SIZE", a="--steps 2 --warmup 1 --cpu-baseline 0 --legs 0"),
        "#   " + cmd.format(c="FETCH_SIZE", a="--copy-probe 256"),
        "#   " + cmd.format(c="WRITE_SIZE", a="--copy-probe 256"),
        "# summarised by: python3 bench.py --pmc-summary <fetch dir> <write dir> <copy fetch dir> <copy write dir>",
        "# the bench runs make %d iterations: per-iteration traffic = sum over kernels of bytes x dispatches / %d"
        % (TRAFFIC_ITERS, TRAFFIC_ITERS),
        "# calibration: a plain copy of %d B reads FETCH_SIZE %.1f KiB and WRITE_SIZE %.1f KiB per dispatch;"
        % (copy_bytes, kf / 1024, kw / 1024),
        "# each counter below is scaled by (bytes copied) / (counter bytes): fetch x %.4f, write x %.4f" % (ff, fw),
        "# kernel <name> dispatches <n> fetch_bytes <B> write_bytes <B>  (mean per dispatch, scaled)",
    ]
    names = sorted(set(bf) | set(bw), key=lambda k: -(np.mean(bf.get(k, [0])) * ff + np.mean(bw.get(k, [0])) * fw))
    for k in names:
        if k.startswith("__amd") or k.startswith("elementwise"):
            continue
        vf = np.mean(bf.get(k, [0.0])) * 1024.0 * ff
        vw = np.mean(bw.get(k, [0.0])) * 1024.0 * fw
        lines.append("kernel %s dispatches %d fetch_bytes %.0f write_bytes %.0f"
                     % (k, max(len(bf.get(k, [])), len(bw.get(k, []))), vf, vw))
    out = os.path.join(ROOT, TRAFFIC_FILE)
    open(out, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))


def grid_metrics_ms(pk, case, n):
    """Best wall time of building the case's metrics on the device from its
    vertex coordinates (context creation excluded)."""
    best = None
    for _ in range(n):
        s = pk.Solver(case)
        s.synchronize()
        x = [np.ascontiguousarray(np.asarray(b.X, dtype=np.float64).transpose(2, 1, 0, 3)).ravel() for b in case.blocks]
        L = pk.solver.lib()
        t = time.perf_counter()
        for ib, xb in enumerate(x):
            pk.solver._check(L.cav3d_set_grid(s.h, ib, pk.solver._p(xb)))
        pk.solver._check(L.cav3d_build_metrics(s.h))
        s.synchronize()
        ms = (time.perf_counter() - t) * 1e3
        best = ms if best is None else min(best, ms)
        s.close()
    return round(best, 3)


FIELDS = ["p", "u", "v", "w", "T", "y1", "y2", "k", "omega"]


def interior_fields(pk, s, case):
    b = case.blocks[0]
    pv = s.get_state(0)[0].reshape(b.kmax + 5, b.jmax + 5, b.imax + 5, case.npv)
    return pv[3:b.kmax + 2, 3:b.jmax + 2, 3:b.imax + 2]


def mixed_precision_run(pk, case, args, ms64):
    """The same warm-up and timed iterations with CAV3D_MIXED (reconstruction,
    face EOS, inviscid and viscous fluxes in FP32; sources, D^-1, sweeps,
    update and norms in FP64), and the per-field difference of the final
    state from the FP64 run's, relative to each field's largest magnitude."""
    out = {}
    states = {}
    for prec in (0, 1):
        s = pk.Solver(case, precision=prec)
        for ib, st in enumerate(case.init):
            s.put_state(ib, *st)
        n = args.warmup + args.steps
        s.synchronize()
        t1 = t2 = None
        for it in range(n):
            if it == args.warmup:
                s.synchronize()
                t1 = time.perf_counter()
            s.timeinteg(1, it + 1, 0.0, want_time=False)
        s.synchronize()
        t2 = time.perf_counter()
        states[prec] = (interior_fields(pk, s, case), s.residual()[0])
        if prec == 1:
            out["ms_per_step"] = round((t2 - t1) * 1e3 / max(args.steps, 1), 4)
        s.close()
    (a, l64), (b, l32) = states[0], states[1]
    out["value"] = round(case.ncell / (out["ms_per_step"] * 1e-3) / 1e6, 4)
    out["unit"] = "Mcell-iter/s"
    out["speedup_vs_fp64"] = round(ms64 / out["ms_per_step"], 4)
    out["iterations_compared"] = args.warmup + args.steps
    err = {}
    for n in range(case.npv):
        scale = float(np.max(np.abs(a[..., n])))
        err[FIELDS[n]] = float(np.max(np.abs(a[..., n] - b[..., n])) / scale) if scale > 0 else 0.0
    out["field_rel_diff_vs_fp64"] = {k: float(f"{v:.3e}") for k, v in err.items()}
    out["residual_l2_rel_diff"] = float(f"{float(np.max(np.abs(l32 - l64) / np.maximum(np.abs(l64), 1e-300))):.3e}")
    return out


def host_cpu():
    """CPU model, logical CPUs this process may use, and whether SMT is on."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    smt = None
    try:
        smt = open("/sys/devices/system/cpu/smt/active").read().strip() == "1"
    except OSError:
        pass
    return {"model": model, "logical_cpus_usable": len(os.sched_getaffinity(0)), "smt_active": smt}


def cpu_baseline(case, size):
    """The oracle (FP64 C restatement, OpenMP over cells and hyperplanes) on
    the same case on the host: one iteration as a warm-up, then the best of
    several timed iterations, with one thread and with every usable core
    (at most 16, the host share of one GPU)."""
    from oracle import oracle as orc
    usable = len(os.sched_getaffinity(0))
    out = {"unit": "Mcell-iter/s", "kind": "port", "host": host_cpu(), "runs": {}}
    for threads, reps in ((min(16, usable), 5), (1, 2)):
        orc.set_threads(threads)
        o = orc.Oracle(case)
        for ib, st in enumerate(case.init):
            o.put_state(ib, *st)
        times = []
        for it in range(reps + 1):
            c0 = time.perf_counter()
            o.timeinteg(1, it + 1, 0.0)
            times.append(time.perf_counter() - c0)
        o.close()
        best = min(times[1:])
        out["runs"][str(threads)] = {"value": round(case.ncell / best / 1e6, 5), "best_s": round(best, 4),
                                     "timed_iterations": reps}
    top = str(min(16, usable))
    out["value"] = out["runs"][top]["value"]
    out["cores"] = int(top)
    out["sample"] = (f"the same {size} case, FP64 C restatement of the step; one warm-up iteration then the "
                     f"best of 5 ({top} threads) and of 2 (1 thread) timed iterations")
    return out


def adjoint_cost(pk, s, case, reps=3):
    """Device time of one discrete-adjoint product (cav3d_rhs_vjp) against one
    evaluation of the residual it differentiates (cav3d_calrhs), on the
    current state, from HIP events on the library's stream (best of reps)."""
    rng = np.random.default_rng(0)
    rb = [rng.standard_normal(case.npv * (b.imax - 1) * (b.jmax - 1) * (b.kmax - 1)) for b in case.blocks]
    s.profile(True)
    s.rhs_vjp(rb)                 # the first call allocates the work arrays
    vjp, rhs = [], []
    for _ in range(reps):
        s.rhs_vjp(rb)
        vjp.append(s.profile(True)[8])
        s.calrhs(0)
        rhs.append(s.profile(True)[9])
    s.profile(False)
    v, r = min(vjp), min(rhs)
    return {"vjp_ms": round(v, 4), "rhs_ms": round(r, 4), "adjoint_over_primal": round(v / r, 3),
            "vjp_mcell_per_s": round(case.ncell / (v * 1e-3) / 1e6, 3), "arithmetic": "f64"}


def alpha_field(pk, case, pv):
    """Vapour volume fraction rho y1 / rho_v of the interior cells (the vof
    field of src/postvariable_module.f90:282-295) from the host EOS."""
    a = pv.reshape(-1, case.npv)
    dv, _ = pk.eos_host(a[:, 0] + case.params["pref"], a[:, 4], a[:, 5], a[:, 6])
    with np.errstate(invalid="ignore", divide="ignore"):
        al = np.where(dv[:, 3] > 0, dv[:, 0] * a[:, 5] / dv[:, 3], 0.0)
    return al


def mixed_study(pk, args):
    """Mixed precision against FP64 on the 128x64x64 cavitating bump over
    `iters` iterations from the same state: residual histories, the
    per-field difference of the final states (p, u, v, w, T, alpha, k,
    omega) relative to each field's largest magnitude, the integrated vapour
    volume, and the wall time each mode needs to reach the residual level
    both reach (the higher of the two final l2 of du).  With --oracle-iters
    M (M <= iters), the FP64 and the mixed states after M iterations are
    also compared with the oracle's (the FP64 C restatement of the step) per
    field, alpha and vapour volume included."""
    if args.oracle_iters > args.iters:
        sys.exit("--oracle-iters must not exceed --iters")
    ni, nj, nk = (int(x) for x in args.size.split("x"))
    case = pk.cases.config2(ni, nj, nk)
    case.params["cav_clip"] = 1
    p = case.params
    b = case.blocks[0]
    inner = (slice(3, b.kmax + 2), slice(3, b.jmax + 2), slice(3, b.imax + 2))
    runs = {}
    for prec, name in ((0, "fp64"), (1, "mixed")):
        s = pk.Solver(case, precision=prec)
        s.set_grid()
        for ib, st in enumerate(case.init):
            s.put_state(ib, *st)
        hist, times, snap = [], [], None
        s.synchronize()
        t0 = time.perf_counter()
        for it in range(args.iters):
            s.timeinteg(1, it + 1, 0.0, want_time=False)
            if (it + 1) % 10 == 0:
                l2 = s.residual()[0]
                times.append(time.perf_counter() - t0)
                hist.append([it + 1, float(l2[0]), float(l2[1])])
            if (it + 1) % 100 == 0:
                print(f"# {name} iteration {it + 1}: l2 dp {hist[-1][1]:.3e} du {hist[-1][2]:.3e}", file=sys.stderr,
                      flush=True)
            if it + 1 == args.oracle_iters:
                ts = time.perf_counter()
                snap = s.get_state(0)[0].copy()
                t0 += time.perf_counter() - ts          # the copy is not part of the run
        pv = s.get_state(0)[0]
        out = s.integrals(p["rhoref"], p["uref"], 0.01, 0.0)
        s.close()
        runs[name] = {"hist": hist, "times": times, "pv": pv, "vap": float(out[0]), "snap": snap}
    P64 = runs["fp64"]["pv"].reshape(b.kmax + 5, b.jmax + 5, b.imax + 5, case.npv)[inner].reshape(-1, case.npv)
    P32 = runs["mixed"]["pv"].reshape(b.kmax + 5, b.jmax + 5, b.imax + 5, case.npv)[inner].reshape(-1, case.npv)
    err = {}
    for n, f in enumerate(FIELDS):
        if f in ("y1", "y2"):
            continue
        sc = float(np.max(np.abs(P64[:, n])))
        err[f] = float(np.max(np.abs(P64[:, n] - P32[:, n])) / sc) if sc > 0 else 0.0
    a64, a32 = alpha_field(pk, case, P64.ravel()), alpha_field(pk, case, P32.ravel())
    err["alpha"] = float(np.max(np.abs(a64 - a32)) / max(float(np.max(np.abs(a64))), 1e-300))
    target = max(runs["fp64"]["hist"][-1][2], runs["mixed"]["hist"][-1][2])
    def time_to(name):
        for (it, _, du), t in zip(runs[name]["hist"], runs[name]["times"]):
            if du <= target:
                return it, t
        return None, None
    i64, t64 = time_to("fp64")
    i32, t32 = time_to("mixed")
    res = {"iterations": args.iters, "field_max_rel_diff": {k: float(f"{v:.3e}") for k, v in err.items()},
           "vapour_volume": {"fp64": runs["fp64"]["vap"], "mixed": runs["mixed"]["vap"],
                             "rel_diff": abs(runs["mixed"]["vap"] - runs["fp64"]["vap"]) / abs(runs["fp64"]["vap"])},
           "matched_residual": {"l2_du_target": target, "fp64": {"iterations": i64, "seconds": t64},
                                "mixed": {"iterations": i32, "seconds": t32}},
           "history_every_10": {"fp64": runs["fp64"]["hist"], "mixed": runs["mixed"]["hist"]}}
    if args.oracle_iters:
        from oracle import oracle as orc
        orc.set_threads(min(16, len(os.sched_getaffinity(0))))
        o = orc.Oracle(case)
        for ib, st in enumerate(case.init):
            o.put_state(ib, *st)
        t0 = time.perf_counter()
        for it in range(args.oracle_iters):
            o.setqres()
            o.timeinteg(1, it + 1, 0.0)
            if (it + 1) % 50 == 0:
                print(f"# oracle iteration {it + 1} ({time.perf_counter() - t0:.0f} s)", file=sys.stderr, flush=True)
        po = o.get_state(0)[0]
        vo = float(o.integrals(p["rhoref"], p["uref"], 0.01, 0.0)[0])
        Po = po.reshape(b.kmax + 5, b.jmax + 5, b.imax + 5, case.npv)[inner].reshape(-1, case.npv)
        ao = alpha_field(pk, case, Po.ravel())
        res["vs_oracle_after"] = args.oracle_iters
        for name in ("fp64", "mixed"):
            Pg = runs[name]["snap"].reshape(b.kmax + 5, b.jmax + 5, b.imax + 5, case.npv)[inner].reshape(-1, case.npv)
            d = {f: float(f"{float(np.max(np.abs(Pg[:, n] - Po[:, n])) / max(float(np.max(np.abs(Po[:, n]))), 1e-300)):.3e}")
                 for n, f in enumerate(FIELDS[:case.npv]) if f != "y2"}
            ag = alpha_field(pk, case, Pg.ravel())
            d["alpha"] = float(f"{float(np.max(np.abs(ag - ao)) / max(float(np.max(np.abs(ao))), 1e-300)):.3e}")
            res[f"{name}_vs_oracle_max_rel_diff"] = d
        res["oracle_vapour_volume"] = vo
        if args.oracle_iters == args.iters:
            for name in ("fp64", "mixed"):
                res["vapour_volume"][f"{name}_vs_oracle_rel_diff"] = abs(runs[name]["vap"] - vo) / abs(vo)
    print(json.dumps(res), flush=True)


def naca_run(pk, args):
    """The NACA 0015 C-grid workload: 4 blocks of 128x160x96 cells (7.9 M)
    resident on the GPU, metrics built on the device from the vertices.
    For FP64 and for the mixed mode from the same initial state: `warmup`
    untimed and `steps` timed iterations (Mcell-iter/s), then more iterations
    up to `iters` in total with the residual norms every 25, the vapour
    volume, lift and drag coefficients every 250 and at the end."""
    ni, nj, nk = (int(x) for x in args.size.split("x"))
    t0 = time.time()
    case = pk.cases.config3(ni, nj, nk, cfl=args.cfl)
    print(f"# case {case.name} ({case.ncell} cells) built in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    p = case.params
    res = {}
    for prec, name in ((0, "fp64"), (1, "mixed")):
        s = pk.Solver(case, precision=prec, metrics="device")
        for ib, st in enumerate(case.init):
            s.put_state(ib, *st)
        first = None
        hist, ihist = [], []
        n = max(args.iters, args.warmup + args.steps)
        t1 = None
        for it in range(n):
            if it == args.warmup:
                s.synchronize()
                t1 = time.perf_counter()
            s.timeinteg(1, it + 1, 0.0, want_time=False)
            if it == args.warmup + args.steps - 1:
                s.synchronize()
                ms = (time.perf_counter() - t1) * 1e3 / args.steps
            if it == 0 or (it + 1) % 25 == 0 or it == n - 1:
                l2 = s.residual()[0]
                first = l2 if first is None else first
                hist.append([it + 1, float(l2[0]), float(l2[1])])
                print(f"# {name} iteration {it + 1}: l2 dp {l2[0]:.3e} du {l2[1]:.3e}", file=sys.stderr, flush=True)
                if not np.all(np.isfinite(l2)):
                    break
            if (it + 1) % 250 == 0:
                q = s.integrals(p["rhoref"], p["uref"], p["area"], np.radians(p["aoa"]))
                ihist.append([it + 1, float(q[0]), float(q[8]), float(q[9])])
        out = s.integrals(p["rhoref"], p["uref"], p["area"], np.radians(p["aoa"]))
        s.close()
        last = hist[-1]
        res[name] = {"ms_per_step": round(ms, 3), "value": round(case.ncell / (ms * 1e-3) / 1e6, 3),
                     "iterations": last[0], "l2_drop_p": last[1] / float(first[0]), "l2_drop_u": last[2] / float(first[1]),
                     "vapour_volume": float(out[0]), "cl": float(out[8]), "cd": float(out[9]), "history": hist,
                     "integrals_every_250": ihist}
    a, b = res["fp64"], res["mixed"]
    rel = {k: abs(b[k] - a[k]) / max(abs(a[k]), 1e-300) for k in ("vapour_volume", "cl", "cd")}
    print(json.dumps({"metric": "Mcell-iter/s (pseudo-time iteration, NACA 0015 C-grid)", "unit": "Mcell-iter/s",
                      "config": {"workload": f"cavitating NACA 0015 C-grid, 4 blocks of {args.size} cells, k-omega "
                                             f"SST, AUSM+-up, MLP, Merkle, AoA 8 deg, sigma 1.0, CFL {args.cfl}",
                                 "cells": case.ncell},
                      "fp64": a, "mixed": b, "mixed_vs_fp64_rel_diff": rel,
                      "speedup_mixed": round(a["ms_per_step"] / b["ms_per_step"], 4)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", default=None, help="cells per block, NIxNJxNK (default 128x64x64; 128x160x96 "
                                                 "with --config 3)")
    ap.add_argument("--config", type=int, default=2, choices=[2, 3],
                    help="2: the cavitating bump channel (the default line); 3: the NACA 0015 C-grid run")
    ap.add_argument("--cfl", type=float, default=50.0, help="CFL of the --config 3 run")
    ap.add_argument("--iters", type=int, default=0, help="total iterations of the --config 3 run / --mixed-study")
    ap.add_argument("--mixed-study", action="store_true",
                    help="run only the mixed-vs-FP64 study on the config-2 workload for --iters iterations")
    ap.add_argument("--oracle-iters", type=int, default=0, help="with --mixed-study: compare FP64 with the oracle "
                                                                  "after this many iterations")
    ap.add_argument("--cpu-baseline", type=int, default=1, help="1: time the oracle on the host")
    ap.add_argument("--legs", type=int, default=1,
                    help="0: only the timed iterations and their phase profile (no grid-metrics, mixed-precision "
                         "or adjoint leg), for counter runs whose dispatch counts must be the iterations' alone")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="run only a plain device-to-device copy of MIB MiB (5 times): the calibration "
                         "kernel for the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--grid-metrics", type=int, default=0, metavar="N",
                    help="run only the device grid-metric build (face vectors, volumes, wall distance) N times")
    ap.add_argument("--pmc-summary", nargs=4, metavar=("FETCH", "WRITE", "COPY_FETCH", "COPY_WRITE"),
                    help="write the HBM-traffic summary from four rocprofv3 --pmc output folders "
                         "(bench FETCH_SIZE, bench WRITE_SIZE, copy probe FETCH_SIZE, copy probe WRITE_SIZE)")
    args = ap.parse_args()
    if args.size is None:
        args.size = "128x160x96" if args.config == 3 else "128x64x64"
    if args.pmc_summary:
        pmc_summary(args, *args.pmc_summary)
        return
    if args.copy_probe:
        import torch
        n = args.copy_probe * (1 << 20) // 8
        a = torch.arange(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_mib": args.copy_probe, "bytes_each_way": n * 8}), flush=True)
        return
    if args.grid_metrics:
        import importlib
        pk = importlib.import_module("3dcode_amd")
        ni, nj, nk = (int(x) for x in args.size.split("x"))
        case = pk.cases.config2(ni, nj, nk)
        ms = grid_metrics_ms(pk, case, args.grid_metrics)
        print(json.dumps({"grid_metrics_ms": ms, "cells": case.ncell, "repeats": args.grid_metrics}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the library keeps every block of a case on a single device "
              "and has no multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import importlib
    pk = importlib.import_module("3dcode_amd")
    if args.config == 3:
        naca_run(pk, args)
        return
    if args.mixed_study:
        mixed_study(pk, args)
        return
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
        acc += s.profile(True)[:8]
    s.profile(False)
    phase = acc / nprof

    # dominant kernel: the forward LU-SGS sweep, one launch over all cells.  Its
    # algorithmic bytes are the forward-sweep pass of the compulsory-traffic
    # model: read pv (n) and mu_t, read the metrics (G words), write dcv (n).
    npv = case.npv
    G = 11 if npv == 9 else 10
    b_fwd = 8.0 * (2 * npv + 1) + 8.0 * G
    t_fwd = phase[6] * 1e-3
    achieved = b_fwd * ncell / t_fwd / 1e9
    it_achieved = B_ALG_FP64_SST * ncell / (phase[5] * 1e-3) / 1e9
    traffic = measured_traffic(args.size, "k_sweep_tile<true>")
    it_traffic = measured_traffic(args.size)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "traffic_source": TRAFFIC_FILE if traffic is not None else None,
                "kernel": "k_sweep_tile<forward> (LU-SGS forward sweep, one launch over all cells)",
                "bytes_per_cell": b_fwd, "launch_ms": round(phase[6], 4),
                "iteration": {"bytes_per_cell": B_ALG_FP64_SST, "achieved": round(it_achieved, 2),
                              "frac": round(it_achieved / HBM_PEAK_GBS, 5),
                              "traffic": it_traffic,
                              "measured_gbs": (round(it_traffic / (phase[5] * 1e-3) / 1e9, 1)
                                               if it_traffic is not None else None)},
                "phase_ms": {k: round(v, 4) for k, v in zip(
                    ["timestep", "boundary", "rhs", "lusgs", "update_eddy", "total", "sweep_fwd", "sweep_bwd"], phase)}}

    grid_ms = grid_metrics_ms(pk, case, 3) if args.legs else None
    mixed = mixed_precision_run(pk, case, args, ms) if args.legs else None

    cpu = cpu_baseline(case, args.size) if args.cpu_baseline else None

    # the discrete adjoint's cost on the stepped state the timed run left
    # (warmup + steps + profiled iterations), its ghost layers refilled
    adjoint = None
    if args.legs:
        try:
            s.setbc()
            adjoint = adjoint_cost(pk, s, case)
            adjoint["state"] = f"after {args.warmup + args.steps + nprof} pseudo-time iterations"
        except pk.Cav3dError as e:
            adjoint = {"error": str(e)[:300]}
    s.close()

    out = {
        "metric": "Mcell-iter/s (pseudo-time iteration: residual + LU-SGS, FP64)",
        "value": round(value, 4), "unit": "Mcell-iter/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"cavitating bump channel {args.size} cells, k-omega SST, AUSM+-up, MLP, Merkle "
                               f"(species clip), 1 block", "cells": ncell},
        "roofline": roofline, "cpu_baseline": cpu,
        "grid_metrics_ms": grid_ms,
        "mixed": mixed,
        "adjoint": adjoint,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
