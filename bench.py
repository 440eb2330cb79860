#!/usr/bin/env python3
"""EVP1D subcycle-loop benchmark on one MI355X.

One step = one model time step's EVP solve: ndte subcycles of
(stress; stepu; ghost update) over the whole 1D point list, inputs resident in
HBM.  Default workload: 1280x1520 grid, ndte = 1000 (the operational-size
single-GPU configuration).  Prints ONE JSON line:

  value      ice-points x subcycles per second (na * ndte * steps / time)
  roofline   dominant kernel: algorithmic bytes per launch / average launch
             duration (HIP events on the library's stream) vs 8 TB/s
  cpu_baseline  the reference's own OpenMP v2b benchmark (flang build under
             oracle/_ref) timed on this host's cores on a bounded sample
             (same grid, fewer subcycles); the oracle's OpenMP C
             restatement if the reference build is absent.

Algorithmic bytes per subcycle (reference data model, f64 = 8 B,
int32/logical = 4 B): stress 4*na + 340*nT, stepu 4*na + 244*nU.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "cice-evp1d_amd"))

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(na: int, nT: int, nU: int) -> dict:
    stress = 4 * na + 340 * nT
    stepu = 4 * na + 244 * nU
    return {"stress": stress, "stepu": stepu, "subcycle": stress + stepu}


def make_state(workload: str):
    import numpy as np
    import evp1d
    from evp1d import layout, synth
    w = dict(synth.WORKLOADS[workload])
    ndte = w.pop("ndte")
    t = time.time()
    g = synth.make_grid(**w)
    ix = evp1d.build_index(g.tmask, g.nghost)
    s = layout.pack(g, ix)
    nT = int((s["skipT"] == 0).sum())
    nU = int((s["skipU"] == 0).sum())
    log(f"[bench] {workload}: {w['nx']}x{w['ny']} na={s['na']} navel={s['navel']} nT={nT} nU={nU} "
        f"ndte={ndte} (generated in {time.time() - t:.1f}s)")
    return w, ndte, g, ix, s, nT, nU


def cpu_baseline(workload: str, s: dict, ndte_sample: int, threads: int) -> dict | None:
    """Reference OpenMP v2b on the same grid for ndte_sample subcycles."""
    sys.path.insert(0, str(REPO))
    from oracle import oracle as O
    st = O.State1D(s["na"], s["navel"])
    for k in st.f:
        st.f[k][:] = s["fields"][k]
    for k in st.idx:
        st.idx[k][:] = s["idx"][k]
    st.skipT[:] = s["skipT"]
    st.skipU[:] = s["skipU"]
    exe = O.REFDIR / "evp_v2b_omp"
    if exe.exists():
        with tempfile.TemporaryDirectory() as d:
            O.write_inputs(d, st, ndte_sample)
            t = time.time()
            r = O.run_reference(Path(d), "evp_v2b_omp", threads=threads, timeout=600,
                                env_extra={"OMP_PROC_BIND": "close", "OMP_PLACES": "cores"})
            wall = time.time() - t
        m = re.search(r"Benchmark took:\s+([0-9.]+)\s+seconds", r.stdout)
        secs = float(m.group(1)) if m else wall
        kind = "reference"
        what = (f"reference src/evp_v2b.F90 (flang -O2 -fopenmp), timer 'bench' around the "
                f"subcycle loop, {workload} grid, {ndte_sample} subcycles")
    else:
        lib = O.OracleLib(omp=True)
        p = lib.params(ndte_sample, **st.scalars)
        os.environ["OMP_NUM_THREADS"] = str(threads)
        t = time.time()
        lib.run(p, st, ndte_sample)
        secs = time.time() - t
        kind = "port"
        what = f"oracle C restatement (gcc -O3 -fopenmp), {workload} grid, {ndte_sample} subcycles"
    return {"value": s["na"] * ndte_sample / secs, "unit": "ice-pts*subcycles/s", "cores": threads,
            "kind": kind, "sample": what, "seconds": round(secs, 3)}


def calibrate_copy(mib: int) -> int:
    """Plain device-to-device copy of known size through torch (plumbing only):
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on this command gives the factor
    between the counters and the bytes a kernel really moves."""
    import torch
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(20):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes_each_way": n * 8, "launches": 20}))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="evp_c1")
    ap.add_argument("--mode", default="auto", choices=("auto", "graph", "persistent", "launch"))
    ap.add_argument("--fma", action="store_true", help="time libevp1d_fma.so instead of the exact build")
    ap.add_argument("--kernel-reps", type=int, default=200, help="launches per individual-kernel timing")
    ap.add_argument("--cpu-subcycles", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--autotune", type=int, default=0, metavar="REPS",
                    help="run evp1d_autotune with REPS launches per configuration before timing")
    ap.add_argument("--tuning", default=None, metavar="SB,SX,UB,UX",
                    help="fixed launch tuning: stress block, stress xcd order, stepu block, stepu xcd order")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB 20 times (the reference kernel for "
                         "scaling rocprofv3 FETCH_SIZE/WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is a design "
                                   "note (DESIGN.md), not a measured path", "n_gpus": args.gpus}))
        return 2
    import evp1d
    w, ndte, g, ix, s, nT, nU = make_state(args.workload)
    na = s["na"]
    p = evp1d.set_evp_parameters(ndte, fma=args.fma)
    eng = evp1d.Engine(na, s["navel"], p, fma=args.fma)
    eng.set_mode(args.mode)
    eng.load_state(s["fields"], s["idx"], s["skipT"], s["skipU"])
    if args.tuning:
        eng.set_tuning(*[int(x) for x in args.tuning.split(",")])
    if args.autotune:
        t = eng.autotune(args.autotune)
        log(f"[bench] autotune: {t}")
    tuning = eng.get_tuning()

    for k in range(args.warmup):
        eng.run(ndte)
        eng.synchronize()
        log(f"[bench] warmup step {k + 1}/{args.warmup} done")
    eng.synchronize()
    t0 = time.perf_counter()
    eng.timer_start()
    for k in range(args.steps):
        eng.run(ndte)
    ev_ms = eng.timer_stop()
    eng.synchronize()
    wall = time.perf_counter() - t0
    log(f"[bench] timed {args.steps} steps: wall {wall:.3f}s, events {ev_ms / 1e3:.3f}s")
    secs = max(wall, ev_ms / 1e3)
    value = na * ndte * args.steps / secs

    # per-kernel average launch durations (HIP events, the library's stream)
    B = algorithmic_bytes(na, nT, nU)
    kernels = {}
    for which in ("stress", "stepu"):
        eng.synchronize()
        eng.timer_start()
        eng.run_individual(which, args.kernel_reps)
        ms = eng.timer_stop()
        avg_us = ms * 1e3 / args.kernel_reps
        kernels[which] = {"avg_us": round(avg_us, 3), "bytes": B[which],
                          "GBps": round(B[which] / (avg_us * 1e-6) / 1e9, 1)}
        log(f"[bench] {which}: {avg_us:.2f} us/launch, {kernels[which]['GBps']} GB/s algorithmic")
    step_us = secs / (args.steps * ndte) * 1e6
    if args.mode == "persistent":
        dom = {"name": "k_persistent", "bytes": B["subcycle"] * ndte, "avg_us": secs / args.steps * 1e6}
    else:
        name = max(kernels, key=lambda k: kernels[k]["avg_us"])
        dom = {"name": "k_" + name, "bytes": B[name], "avg_us": kernels[name]["avg_us"]}
    achieved = dom["bytes"] / (dom["avg_us"] * 1e-6) / 1e9
    traffic = None
    tf = REPO / "profiles" / f"traffic_{args.workload}.json"
    if tf.exists():
        traffic = json.loads(tf.read_text()).get(dom["name"])
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom["name"],
                "bytes_per_launch": dom["bytes"]}
    whole = B["subcycle"] / (step_us * 1e-6) / 1e9
    eng.close()

    cpu = None
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        try:
            cpu = cpu_baseline(args.workload, s, args.cpu_subcycles, threads)
            log(f"[bench] cpu baseline: {cpu}")
        except Exception as e:  # reported, never fatal for the GPU figure
            log(f"[bench] cpu baseline failed: {e!r}")
            cpu = {"value": None, "error": repr(e)[:200]}

    out = {
        "metric": "ice-pts*subcycles/s + % HBM roofline (bit-exact FP64)",
        "value": value, "unit": "ice-pts*subcycles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": secs / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{args.workload} {w['nx']}x{w['ny']} ndte={ndte}", "nx": w["nx"],
                   "ny": w["ny"], "seed": w["seed"], "na": na, "navel": s["navel"], "nT": nT, "nU": nU,
                   "ndte": ndte, "mode": args.mode, "build": "fma" if args.fma else "exact", "tuning": tuning,
                   "us_per_subcycle": round(step_us, 3), "loop_GBps_algorithmic": round(whole, 1)},
        "kernels": kernels,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
