#!/usr/bin/env python3
"""EVP1D subcycle-loop benchmark on one MI355X.

One step = one model time step's EVP solve: ndte subcycles of
(stress; stepu; ghost update) over the whole 1D point list, inputs resident in
HBM.  Default workload: the 1280x1520 grid with ndte = 1000 (the operational-
size single-GPU configuration), library default form (fused subcycles).
Prints ONE JSON line:

  value      ice-points x subcycles per second (na * ndte * steps / time),
             exact (bit-identical) FP64 build
  roofline   dominant kernel (k_pair: two whole subcycles per launch, or
             k_fused: one, whichever form the tuner picked): algorithmic
             bytes per launch / average launch duration (HIP events on the
             library's stream) vs 8 TB/s; traffic from rocprofv3 counters
  adjoint    time of one gradient evaluation (hand-written adjoint kernels;
             whole trajectory stored, and binomial checkpointing with 32
             stored states) over the time of one forward step, with central
             finite-difference and tangent-linear dot-product checks
  cpu_baseline  the reference's own OpenMP v2b benchmark (flang build under
             oracle/_ref) on this host's cores on a bounded sample (same
             grid, fewer subcycles); the oracle's OpenMP C restatement if the
             reference build is absent.

Other measurements (one JSON line each):
  --workload evp_c2 [--mode graph|persistent]   2560x3040 grid, loop forms
  --ensemble 64      64 members in one batched handle vs 64 handles (member
                     after member, and lockstep)
  --api              dyn_evp1d_run end to end through the Fortran CICE driver,
                     2D <-> 1D copies on the device vs in host loops
  --fma              the contraction-on build
  --calibrate-copy MIB   plain copy for scaling rocprofv3 FETCH/WRITE_SIZE

Algorithmic bytes per subcycle (reference data model, f64 = 8 B,
int32/logical = 4 B): stress 4*na + 340*nT, stepu 4*na + 244*nU.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "cice-evp1d_amd"))

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(na: int, nT: int, nU: int) -> dict:
    stress = 4 * na + 340 * nT
    stepu = 4 * na + 244 * nU
    return {"stress": stress, "stepu": stepu, "subcycle": stress + stepu}


def make_state(nx: int, ny: int, seed: int, **kw):
    import evp1d
    from evp1d import layout, synth
    g = synth.make_grid(nx, ny, seed, **kw)
    ix = evp1d.build_index(g.tmask, g.nghost)
    s = layout.pack(g, ix)
    s["nT"] = int((s["skipT"] == 0).sum())
    s["nU"] = int((s["skipU"] == 0).sum())
    return s


def workload_state(workload: str):
    from evp1d import synth
    w = dict(synth.WORKLOADS[workload])
    ndte = w.pop("ndte")
    t = time.time()
    s = make_state(**w)
    log(f"[bench] {workload}: {w['nx']}x{w['ny']} na={s['na']} navel={s['navel']} nT={s['nT']} "
        f"nU={s['nU']} ndte={ndte} (generated in {time.time() - t:.1f}s)")
    return w, ndte, s


def new_engine(s: dict, ndte: int, mode: str = "auto", fma: bool = False):
    import evp1d
    p = evp1d.set_evp_parameters(ndte, fma=fma)
    eng = evp1d.Engine(s["na"], s["navel"], p, fma=fma)
    eng.set_mode(mode)
    eng.load_state(s["fields"], s["idx"], s["skipT"], s["skipU"])
    return eng


DYNAMIC = ("stressp_1", "stressp_2", "stressp_3", "stressp_4", "stressm_1", "stressm_2", "stressm_3",
           "stressm_4", "stress12_1", "stress12_2", "stress12_3", "stress12_4", "uvel", "vvel",
           "strintxU", "strintyU", "Cb", "str1", "str2", "str3", "str4", "str5", "str6", "str7", "str8")


def reset_state(eng, s: dict, overrides: dict | None = None) -> None:
    """Put the initial dynamic state (and optionally perturbed parameters) back."""
    for n in DYNAMIC:
        eng.set_field(n, s["fields"][n])
    for n, a in (overrides or {}).items():
        eng.set_field(n, a)


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
    # the reference's performance flag set (f2: OpenMP, -O3, vector ISA, contraction on;
    # compiler_oneapi.mk:18) as flang has it, else its OpenMP build without contraction
    exe = next((O.REFDIR / b for b in ("evp_v2b_f2", "evp_v2b_omp") if (O.REFDIR / b).exists()), None)
    if exe is not None:
        with tempfile.TemporaryDirectory() as d:
            O.write_inputs(d, st, ndte_sample)
            t = time.time()
            r = O.run_reference(Path(d), exe.name, threads=threads, timeout=600,
                                env_extra={"OMP_PROC_BIND": "close", "OMP_PLACES": "cores"})
            wall = time.time() - t
        m = re.search(r"Benchmark took:\s+([0-9.]+)\s+seconds", r.stdout)
        secs = float(m.group(1)) if m else wall
        kind = "reference"
        flags = ("flang -O3 -fopenmp -march=x86-64-v3 (contraction on)" if exe.name == "evp_v2b_f2"
                 else "flang -O2 -fopenmp -ffp-contract=off")
        what = (f"reference src/evp_v2b.F90 ({flags}), timer 'bench' around the "
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
            "kind": kind, "sample": what, "seconds": round(secs, 3), "host": host_cpu()}


def api_measure(workload: str, nrun: int) -> dict:
    """dyn_evp1d_run end to end through the Fortran CICE driver API (2D fields
    in, ndte subcycles, 2D fields out), with the 2D <-> 1D copies on the device
    (ldevpack) and in host loops."""
    import re
    import subprocess
    import tempfile
    from evp1d import layout, synth
    w = dict(synth.WORKLOADS[workload])
    ndte = w.pop("ndte")
    g = synth.make_grid(**w)
    exe = REPO / "cice-evp1d_amd" / "lib" / "evp1d_fortran" / "evp1d_api_driver"
    out = {"workload": workload, "ndte": ndte, "nrun": nrun}
    with tempfile.TemporaryDirectory() as d:
        with open(Path(d) / "api_in.bin", "wb") as fh:
            fh.write(np.array([w["nx"], w["ny"], g.nghost, ndte, nrun], dtype="<i4").tobytes())
            for m in (g.tmask, g.icetmask, g.iceumask):
                fh.write(np.asfortranarray(m.astype("<i4")).tobytes(order="F"))
            for n in ("dxT", "dyT", "uarear", "HTE", "HTN"):
                fh.write(g.fields[n].tobytes(order="F"))
            for n in layout.DYN_2D:
                fh.write(g.fields[n].tobytes(order="F"))
        # lautotune: the first call tunes the fused form (evp1d_autotune_fused),
        # later calls run the tuned form
        for flag, tune, key in ((".true.", ".true.", "device_pack_tuned_ms"), (".true.", ".false.", "device_pack_ms"),
                                (".false.", ".false.", "host_pack_ms")):
            (Path(d) / "evp_1d.nml").write_text(f"&evp1d_nml\n  ldevpack = {flag}\n  lautotune = {tune}\n/\n")
            r = subprocess.run([str(exe)], cwd=d, capture_output=True, text=True, timeout=600)
            if r.returncode != 0:
                raise RuntimeError(r.stdout[-500:] + r.stderr[-500:])
            ms = [float(x) for x in re.findall(r"dyn_evp1d_run\s+\d+:\s+([0-9.]+) ms", r.stdout)]
            out[key] = min(ms[1:]) if len(ms) > 1 else ms[0]   # the first call also builds the graph
            log(f"[bench] api ldevpack={flag} lautotune={tune}: per call {ms} ms")
    return out


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


def _probe_lib():
    import ctypes
    path = REPO / "cice-evp1d_amd" / "lib" / "libevp1d_probe.so"
    if not path.exists():
        raise ImportError(f"{path} is not built: run __graft_entry__.build()")
    L = ctypes.CDLL(str(path))
    L.evp1d_probe_stream.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64,
                                     ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
    L.evp1d_probe_stream.restype = ctypes.c_int
    L.evp1d_probe_stream_hints.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
    L.evp1d_probe_stream_hints.restype = ctypes.c_int
    return L


def probe_rate(nr: int, nw: int, mib: int = 512, reps: int = 20, hints: int = 0) -> float:
    """GB/s (read + written) of the streaming probe at nr read and nw written
    streams of mib MiB, one point per thread; hints = 1: non-temporal loads
    and stores."""
    import ctypes
    n = mib * (1 << 20) // 16 // 1024 * 1024
    us = ctypes.c_double()
    rc = _probe_lib().evp1d_probe_stream_hints(nr, nw, 1, hints, n, reps, ctypes.byref(us))
    if rc:
        raise RuntimeError(f"evp1d_probe_stream_hints({nr}, {nw}, 1, {hints}) returned {rc}")
    return (nr + nw) * n * 16 / us.value / 1e3


def stream_probe(mib: int, reps: int = 20) -> int:
    """HBM ceiling at several read/write mixes: lib/libevp1d_probe.so streams
    nr double2 arrays of mib MiB each in and nw out (independent coalesced
    loads), timed with HIP events; prints one JSON line."""
    import ctypes
    L = _probe_lib()
    n = mib * (1 << 20) // 16 // 1024 * 1024
    rows = []
    for hints in (0, 1):
        for per in (1, 4):
            for nr, nw in ((1, 0), (1, 1), (2, 2), (2, 1), (3, 1), (4, 1), (5, 1), (6, 1)):
                us = ctypes.c_double()
                rc = L.evp1d_probe_stream_hints(nr, nw, per, hints, n, reps, ctypes.byref(us))
                if rc:
                    raise RuntimeError(f"evp1d_probe_stream_hints({nr}, {nw}, {per}, {hints}) returned {rc}")
                b = (nr + nw) * n * 16
                rows.append({"read_streams": nr, "write_streams": nw, "points_per_thread": per, "hints": hints,
                             "read_fraction": round(nr / (nr + nw), 3), "bytes": b, "us": round(us.value, 2),
                             "GB_s": round(b / us.value / 1e3, 1)})
                log(f"[probe] {nr} in {nw} out, {per} per thread, hints {hints}: {us.value:.1f} us, "
                    f"{b / us.value / 1e3:.1f} GB/s")
    print(json.dumps({"stream_probe_mib_per_stream": mib, "launches": reps, "mixes": rows}))
    return 0


def time_steps(eng, ndte: int, steps: int, warmup: int) -> float:
    for k in range(warmup):
        eng.run(ndte)
        eng.synchronize()
        log(f"[bench] warmup step {k + 1}/{warmup} done")
    eng.synchronize()
    t0 = time.perf_counter()
    eng.timer_start()
    for _ in range(steps):
        eng.run(ndte)
    ev_ms = eng.timer_stop()
    eng.synchronize()
    wall = time.perf_counter() - t0
    log(f"[bench] timed {steps} steps: wall {wall:.3f}s, events {ev_ms / 1e3:.3f}s")
    return max(wall, ev_ms / 1e3)


OUT_FIELDS = ("stressp_1", "stressp_2", "stressp_3", "stressp_4", "stressm_1", "stressm_2", "stressm_3",
              "stressm_4", "stress12_1", "stress12_2", "stress12_3", "stress12_4", "uvel", "vvel",
              "strintxU", "strintyU", "Cb")


def output_fields(eng, s: dict) -> dict:
    """The fields after one ndte-subcycle run from the initial state (the
    engine is left in that state)."""
    reset_state(eng, s)
    eng.run(eng.get_params().ndte)
    eng.synchronize()
    return {n: eng.get_field(n) for n in OUT_FIELDS}


def fma_measure(s: dict, ndte: int, steps: int, warmup: int, tune_nsub: int, exact: dict) -> dict:
    """The contraction-on build (libevp1d_fma.so) on the same workload, tuned
    on its own (evp1d_autotune_fused), timed like the exact build; its
    largest norm-wise difference max|d| / max|exact| over the output fields
    after one run from the initial state."""
    eng = new_engine(s, ndte, "auto", fma=True)
    tuned = eng.autotune_fused(tune_nsub) if tune_nsub > 0 else None
    log(f"[bench] fma: autotune_fused {tuned}")
    secs = time_steps(eng, ndte, steps, warmup)
    got = output_fields(eng, s)
    eng.close()
    err = 0.0
    for n, a in exact.items():
        r = np.max(np.abs(a))
        d = np.max(np.abs(got[n] - a))
        err = max(err, d / r if r > 0 else d)
    out = {"value": s["na"] * ndte * steps / secs, "unit": "ice-pts*subcycles/s",
           "us_per_subcycle": round(secs / (steps * ndte) * 1e6, 3), "ms_per_step": secs / steps * 1e3,
           "max_normwise_diff_vs_exact": err, "tiling": tuned}
    log(f"[bench] fma: {out['us_per_subcycle']} us/subcycle, max norm-wise diff vs exact {err:.2e}")
    return out


def host_cpu() -> dict:
    """Model, sockets and NUMA nodes of this host (lscpu)."""
    import subprocess
    info = {"logical_cpus": os.cpu_count()}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except Exception:
        return info
    for key, name in (("Model name", "model"), ("Socket(s)", "sockets"), ("NUMA node(s)", "numa_nodes"),
                      ("Core(s) per socket", "cores_per_socket"), ("Thread(s) per core", "threads_per_core")):
        m = re.search(rf"^{re.escape(key)}:\s*(.+)$", txt, re.M)
        if m:
            v = m.group(1).strip()
            info[name] = int(v) if v.isdigit() else v
    return info


def adjoint_measure(s: dict, ndte: int, t_fwd_step: float, nsnap: int, ndir: int, ndot: int,
                    fused_cfg: dict | None = None) -> dict:
    """Gradient of L = 1/2 sum_{skipU=0} |u_N - u*|^2 w.r.t. strength,
    cdn_ocn, forcex, forcey; u* from a run with strength x 1.2.  fused_cfg:
    the forward run's tiling and stress storage, used by the trajectory's
    forward sweep too."""
    eng = new_engine(s, ndte)
    if fused_cfg:
        eng.set_tiling(fused_cfg["tile_w"], fused_cfg["tile_h"], fused_cfg["block"], fused_cfg["xcd"])
        eng.set_fused_inplace(fused_cfg["inplace"])
        eng.set_fused_depth(fused_cfg.get("depth", 1))
        eng.set_tile_order(fused_cfg.get("order", 0))
    reset_state(eng, s, {"strength": 1.2 * s["fields"]["strength"]})
    eng.run(ndte)
    ut = eng.get_field("uvel")[:s["na"]]
    vt = eng.get_field("vvel")[:s["na"]]
    reset_state(eng, s, {"strength": s["fields"]["strength"]})
    eng.run(ndte)                      # warm the forward graph and kernels
    # the gradient's forward sweep runs one subcycle per launch (it stores
    # every subcycle): the one-subcycle fused forward on the same tiling,
    # for the ratio against the same kernels
    t_fwd1 = None
    if fused_cfg and fused_cfg.get("depth", 1) == 2:
        eng.set_fused_depth(1)
        reset_state(eng, s)
        eng.run(ndte)
        eng.synchronize()
        t0 = time.perf_counter()
        eng.run(ndte)
        eng.synchronize()
        t_fwd1 = time.perf_counter() - t0
        eng.set_fused_depth(2)
        eng.run(ndte)
    timings = {}
    grads = {}
    for ns in sorted({min(nsnap, ndte), ndte}):
        if ns >= ndte:   # first call allocates the trajectory storage (kept by the handle)
            reset_state(eng, s)
            eng.adjoint_run(ndte, ut, vt, nsnap=ns)
        reset_state(eng, s)
        eng.synchronize()
        t0 = time.perf_counter()
        grads[ns] = eng.adjoint_run(ndte, ut, vt, nsnap=ns)
        timings[ns] = time.perf_counter() - t0
        log(f"[bench] adjoint ({ns} stored states{', whole trajectory' if ns >= ndte else ''}): "
            f"{timings[ns]:.3f}s, forward step {t_fwd_step:.3f}s, ratio {timings[ns] / t_fwd_step:.2f}")
    g = grads[ndte]
    same = all(np.array_equal(grads[ns][k], g[k]) for ns in grads
               for k in ("strength", "cdn_ocn", "forcex", "forcey"))
    t_adj = timings[ndte]
    rng = np.random.default_rng(1234)
    f = s["fields"]

    def direction():
        return {"strength": rng.standard_normal(s["na"]) * f["strength"],
                "cdn_ocn": np.full(s["na"], rng.standard_normal()) * f["cdn_ocn"],
                "forcexU": rng.standard_normal(s["na"]) * np.abs(f["forcexU"]),
                "forceyU": rng.standard_normal(s["na"]) * np.abs(f["forceyU"])}

    def dot(d):
        return (np.dot(g["strength"], d["strength"]) + np.dot(g["cdn_ocn"], d["cdn_ocn"]) +
                np.dot(g["forcex"], d["forcexU"]) + np.dot(g["forcey"], d["forceyU"]))

    h = 1e-6
    fd_err = 0.0
    for k in range(ndir):
        d = direction()
        vals = []
        for sign in (1.0, -1.0):
            reset_state(eng, s, {n: f[n] + sign * h * d[n] for n in d})
            eng.run(ndte)
            vals.append(eng.loss(ut, vt))
        fd = (vals[0] - vals[1]) / (2 * h)
        ad = dot(d)
        fd_err = max(fd_err, abs(fd - ad) / max(abs(fd), abs(ad)))
    dot_err = 0.0
    for k in range(ndot):
        d = direction()
        reset_state(eng, s, {n: f[n] for n in d})
        _, dl = eng.tangent_run(ndte, ut, vt, d["strength"], d["cdn_ocn"], d["forcexU"], d["forceyU"])
        ad = dot(d)
        dot_err = max(dot_err, abs(dl - ad) / max(abs(dl), abs(ad)))
    eng.close()
    log(f"[bench] adjoint checks: FD max rel err {fd_err:.2e} ({ndir} directions, eps 1e-6 relative), "
        f"tangent-linear dot product max rel err {dot_err:.2e} ({ndot} directions)")
    out = {"t_adj_over_t_fwd": round(t_adj / t_fwd_step, 3), "t_adj_s": round(t_adj, 4),
           "t_fwd_s": round(t_fwd_step, 4)}
    if t_fwd1:
        out.update({"t_fwd_one_subcycle_per_launch_s": round(t_fwd1, 4),
                    "t_adj_over_t_fwd_one_subcycle_per_launch": round(t_adj / t_fwd1, 3)})
    return {**out, "stored_states": ndte, "loss": g["loss"],
            "checkpointed": {"stored_states": min(nsnap, ndte),
                             "t_adj_over_t_fwd": round(timings[min(nsnap, ndte)] / t_fwd_step, 3)},
            "gradients_identical": same,
            "fd_directions": ndir, "fd_max_rel_err": fd_err, "dot_directions": ndot,
            "dot_max_rel_err": dot_err,
            "grad_cdn_ocn_scalar": float(np.sum(g["cdn_ocn"]))}


def _ensemble_grid(k: int):
    """Synthetic 2D state of ensemble member k (worker processes: numpy only)."""
    from evp1d import synth
    m = synth.ensemble_member(k)
    return synth.make_grid(m["nx"], m["ny"], m["seed"], force_perturb=0.02)


def ensemble_states(nmem: int) -> list:
    """The members' 1D states; the 2D fields are generated in worker
    processes (at most 16), the lists and packing in this one."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    import evp1d
    from evp1d import layout
    out = []
    workers = max(1, min(16, os.cpu_count() or 1))
    with ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
        for g in ex.map(_ensemble_grid, range(nmem)):
            s = layout.pack(g, evp1d.build_index(g.tmask, g.nghost))
            s["nT"] = int((s["skipT"] == 0).sum())
            s["nU"] = int((s["skipU"] == 0).sum())
            out.append(s)
    return out


def ensemble_measure(nmem: int, steps: int, autotune: int, tune_fused: int = 0, tune_members: bool = True) -> dict:
    """64-member ensemble: one batched handle against member after member
    and lockstep launches.  tune_fused > 0: the batch first picks its fused
    form with evp1d_autotune_fused (tile ladder, stress storage, one or two
    subcycles per launch) on its own data; every single member does the same
    (tune_members) or takes the batch's choice."""
    import evp1d
    t = time.time()
    members = ensemble_states(nmem)
    sizes = [(s["na"], s["navel"]) for s in members]
    tot = sum(a for a, _ in sizes)
    log(f"[bench] ensemble: {nmem} members, na from {min(a for a, _ in sizes)} to "
        f"{max(a for a, _ in sizes)}, total {tot} (generated in {time.time() - t:.1f}s)")
    ndte = 1000
    p = evp1d.set_evp_parameters(ndte)
    batch = evp1d.Engine.batch(sizes, p)
    for k, s in enumerate(members):
        batch.load_member(k, s["fields"], s["idx"], s["skipT"], s["skipU"])
    tuning = batch.autotune(autotune) if autotune else batch.get_tuning()
    fused_batch = batch.autotune_fused(tune_fused) if tune_fused > 0 else None
    if fused_batch:
        log(f"[bench] ensemble: batch autotune_fused: {fused_batch}")
    t_batch = time_steps(batch, ndte, steps, 1) / steps
    # the whole ensemble's run as one cooperative launch (the persistent form
    # on the same tiles), results identical
    t_batch_pers = None
    if batch.get_fused_depth()["active"]:
        batch.set_mode("persistent")
        t_batch_pers = time_steps(batch, ndte, steps, 1) / steps
        batch.set_mode("auto")
    batch.close()
    singles = [new_engine(s, ndte) for s in members]
    depths = []
    if tune_fused > 0 and tune_members:
        for e in singles:
            depths.append(e.autotune_fused(tune_fused)["depth"])
        log(f"[bench] ensemble: members tuned, {depths.count(2)} of {nmem} on two subcycles per launch")
    elif fused_batch:
        for e in singles:
            e.set_tiling(fused_batch["tile_w"], fused_batch["tile_h"], fused_batch["block"], fused_batch["xcd"])
            e.set_fused_inplace(fused_batch["inplace"])
            e.set_fused_depth(fused_batch["depth"])
            e.set_tile_order(fused_batch["order"])
            depths.append(e.get_fused_depth()["depth"] if e.get_fused_depth()["active"] else 1)
    for e in singles:
        e.run(ndte)
    for e in singles:
        e.synchronize()
    # member after member, each its own ndte-subcycle graph
    t0 = time.perf_counter()
    for _ in range(steps):
        for e in singles:
            e.run(ndte)
            e.synchronize()
    t_seq = (time.perf_counter() - t0) / steps
    # lockstep: every subcycle launches each member in turn (64 launches per subcycle)
    for e in singles:
        e.run(1)
    for e in singles:
        e.synchronize()
    t0 = time.perf_counter()
    for _ in range(ndte):
        for e in singles:
            e.run(1)
    for e in singles:
        e.synchronize()
    t_lock = time.perf_counter() - t0
    # member after member in the unfused form (the reference's two kernels,
    # stress then stepu, as one graph per member)
    for e in singles:
        e.set_mode("graph")
        e.run(ndte)
    for e in singles:
        e.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for e in singles:
            e.run(ndte)
            e.synchronize()
    t_seq_graph = (time.perf_counter() - t0) / steps
    for e in singles:
        e.close()
    out = {"members": nmem, "total_points": tot, "ndte": ndte,
           "batched_s": round(t_batch, 4), "sequential_s": round(t_seq, 4), "lockstep_s": round(t_lock, 4),
           "sequential_unfused_s": round(t_seq_graph, 4), "speedup_vs_sequential_unfused": round(t_seq_graph / t_batch, 3),
           "batched_pts_sub_per_s": tot * ndte / t_batch, "sequential_pts_sub_per_s": tot * ndte / t_seq,
           "lockstep_pts_sub_per_s": tot * ndte / t_lock,
           "speedup": round(t_seq / t_batch, 3), "speedup_vs_lockstep": round(t_lock / t_batch, 3),
           "batched_persistent_s": round(t_batch_pers, 4) if t_batch_pers else None,
           "persistent_speedup_vs_lockstep": round(t_lock / t_batch_pers, 3) if t_batch_pers else None,
           "tuning": tuning, "fused_batch": fused_batch,
           "members_tuned": "each member (evp1d_autotune_fused)" if (tune_fused > 0 and tune_members) else
                            "the batch's choice" if fused_batch else "library default",
           "members_on_pair_form": depths.count(2) if depths else None}
    if t_batch_pers:
        log(f"[bench] ensemble: batched as one cooperative launch {t_batch_pers:.3f}s")
    log(f"[bench] ensemble: batched {t_batch:.3f}s, member after member {t_seq:.3f}s (x{t_seq / t_batch:.2f}), "
        f"lockstep {t_lock:.3f}s (x{t_lock / t_batch:.2f}), member after member unfused {t_seq_graph:.3f}s "
        f"(x{t_seq_graph / t_batch:.2f})")
    return out


def v0_measure(workload: str, steps: int, warmup: int, kernel_reps: int, form: str = "fused",
               hints: int | None = None) -> dict:
    """The v0 2D block path (evp2d_*) on the workload's grid as one block:
    CICE's (nx_block, ny_block) arrays, per-block cell lists.  Algorithmic
    bytes per subcycle in the reference's data model: stress 360 B per T cell
    (indxti/indxtj 8, own uvel/vvel 16, the ten metric/strength arrays 80,
    12 stresses in 96 and out 96, str 64), stepu 248 B per U cell
    (indxui/indxuj 8, uvel/vvel 16, 14 forcing/metric arrays 112, str 64,
    uvel/vvel 16, strintx/y 16, taubx/y 16); the reference's clearing of the
    whole str scratch in every stress call is not counted.  form: "fused"
    (one tile kernel per subcycle, the default) or "plain" (--mode graph);
    hints: the fused form's cache hints (None: the library's choice by size)."""
    import evp1d
    from evp1d import layout, synth
    w = dict(synth.WORKLOADS[workload])
    ndte = w.pop("ndte")
    g = synth.make_grid(**w)
    blk = layout.Blocks2D(g)
    nT, nU = int(blk.icellt[0]), int(blk.icellu[0])
    eng = evp1d.Engine2D(blk.nxb, blk.nyb, 1, evp1d.set_evp_parameters(ndte))
    eng.set_mode(form)
    if hints is not None:
        eng.set_cache_hints(hints)
    eng.load_state(blk)
    hints_on = eng.get_cache_hints()["active"]
    log(f"[bench] v0 path: {blk.nxb}x{blk.nyb} one block, icellt={nT} icellu={nU}, cache hints {int(hints_on)}")
    for k in range(warmup):
        eng.run(ndte)
        eng.synchronize()
    eng.synchronize()
    t0 = time.perf_counter()
    eng.timer_start()
    for _ in range(steps):
        eng.run(ndte)
    ev_ms = eng.timer_stop()
    wall = time.perf_counter() - t0
    secs = max(wall, ev_ms / 1e3)
    B = {"stress": 360 * nT, "stepu": 248 * nU}
    B["fused"] = B["stress"] + B["stepu"]   # one whole subcycle
    kernels = {}
    for which in ("stress", "stepu", "fused"):
        eng.synchronize()
        eng.timer_start()
        eng.run_individual(which, kernel_reps)
        us = eng.timer_stop() * 1e3 / kernel_reps
        kernels[which] = {"avg_us": round(us, 3), "bytes": B[which], "GBps": round(B[which] / (us * 1e-6) / 1e9, 1)}
        log(f"[bench] v0 {which}: {us:.2f} us/launch, {kernels[which]['GBps']} GB/s algorithmic")
    mode = eng.get_mode()
    eng.close()
    name = "fused" if mode == "fused" else max(("stress", "stepu"), key=lambda k: kernels[k]["avg_us"])
    dom_us = kernels[name]["avg_us"]
    achieved = B[name] / (dom_us * 1e-6) / 1e9
    traffic = None
    tf = REPO / "profiles" / f"traffic_{workload}.json"
    if tf.exists():
        traffic = json.loads(tf.read_text()).get("k2_" + name + ("@hints" if name == "fused" and hints_on else ""))
    step_us = secs / (steps * ndte) * 1e6
    sub_bytes = B["stress"] + B["stepu"]
    return {
        "metric": "ice-pts*subcycles/s + % HBM roofline (bit-exact FP64), v0 2D block path",
        "value": nT * ndte * steps / secs, "unit": "ice-T-cells*subcycles/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": secs / steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{workload} {w['nx']}x{w['ny']} ndte={ndte}, v0 2D block path (one block)",
                   "form": mode, "cache_hints": int(hints_on),
                   "nx_block": blk.nxb, "ny_block": blk.nyb, "icellt": nT, "icellu": nU, "ndte": ndte,
                   "us_per_subcycle": round(step_us, 3),
                   "loop_GBps_algorithmic": round(sub_bytes / (step_us * 1e-6) / 1e9, 1)},
        "kernels": kernels,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": "k2_" + name,
                     "bytes_per_launch": B[name]},
        "cpu_baseline": None,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="evp_c1")
    ap.add_argument("--mode", default="auto", choices=("auto", "graph", "persistent", "launch", "fused"))
    ap.add_argument("--tiling", default=None, metavar="W,H,BLOCK,XCD",
                    help="fused-form tile (grid cells) and XCD-aware tile order")
    ap.add_argument("--fma", action="store_true", help="time libevp1d_fma.so instead of the exact build")
    ap.add_argument("--kernel-reps", type=int, default=200, help="launches per individual-kernel timing")
    ap.add_argument("--cpu-subcycles", type=int, default=1000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-fma", action="store_true", help="skip timing the contraction-on build beside the exact one")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the 64-member batched-vs-sequential comparison in the default run")
    ap.add_argument("--adjoint-snapshots", type=int, default=32)
    ap.add_argument("--fd-directions", type=int, default=50)
    ap.add_argument("--dot-directions", type=int, default=10)
    ap.add_argument("--path", default="1d", choices=("1d", "v0"),
                    help="v0: time the v0 2D block path (evp2d_*) on the workload instead of the 1D engine")
    ap.add_argument("--api", action="store_true",
                    help="only time dyn_evp1d_run through the Fortran driver API, device vs host 2D<->1D copies")
    ap.add_argument("--ensemble", type=int, default=0, metavar="N",
                    help="only measure an N-member ensemble, batched vs sequential")
    ap.add_argument("--autotune", type=int, default=0, metavar="REPS",
                    help="run evp1d_autotune with REPS launches per configuration before timing")
    ap.add_argument("--tune-fused", type=int, default=100, metavar="NSUB",
                    help="before timing the fused form, run evp1d_autotune_fused with NSUB subcycles per "
                         "candidate (tile ladder x stress storage); 0: keep the defaults or --tiling/--inplace")
    ap.add_argument("--inplace", type=int, default=None, choices=(0, 1),
                    help="fused stress storage: 1 in place, 0 two buffers (implies --tune-fused 0)")
    ap.add_argument("--order", type=int, default=None, choices=(0, 1),
                    help="tile order of the fused forms: 0 rows of tiles, 1 Hilbert curve (evp1d_set_tile_order)")
    ap.add_argument("--depth", type=int, default=None, choices=(1, 2),
                    help="subcycles per fused launch: 2 the pair form (implies --tune-fused 0)")
    ap.add_argument("--hints", type=int, default=None, choices=(-1, 0, 1),
                    help="cache hints of the pair form (or, with --path v0, of the 2D fused form): -1 chosen "
                         "by size, 0 off, 1 on (implies --tune-fused 0)")
    ap.add_argument("--tuning", default=None, metavar="SB,SX,UB,UX",
                    help="fixed launch tuning: stress block, stress xcd order, stepu block, stepu xcd order")
    ap.add_argument("--stream-probe", type=int, default=0, metavar="MIB",
                    help="HBM streaming probe at several read/write mixes, MIB per stream")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB 20 times (the reference kernel for "
                         "scaling rocprofv3 FETCH_SIZE/WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.stream_probe:
        return stream_probe(args.stream_probe)
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)
    if args.gpus != 1:
        print(json.dumps({"error": "the engine runs on one GPU; sharding over several GPUs is designed, "
                                   "not implemented", "n_gpus": args.gpus}))
        return 2
    if args.path == "v0":
        form = "plain" if args.mode == "graph" else "fused"
        print(json.dumps(v0_measure(args.workload, args.steps, args.warmup, args.kernel_reps, form, args.hints)),
              flush=True)
        return 0
    if args.api:
        out = api_measure(args.workload, 3)
        out.update({"metric": "dyn_evp1d_run wall time, 2D fields in and out (Fortran CICE driver API)",
                    "unit": "ms", "n_gpus": 1, "dtype": "f64", "data": "synthetic"})
        print(json.dumps(out), flush=True)
        return 0
    if args.ensemble:
        out = ensemble_measure(args.ensemble, args.steps, args.autotune, args.tune_fused)
        out.update({"metric": "ensemble per-member throughput, batched vs sequential", "n_gpus": 1,
                    "dtype": "f64", "data": "synthetic"})
        print(json.dumps(out), flush=True)
        return 0

    w, ndte, s = workload_state(args.workload)
    na, nT, nU = s["na"], s["nT"], s["nU"]
    eng = new_engine(s, ndte, args.mode, args.fma)
    if args.tuning:
        eng.set_tuning(*[int(x) for x in args.tuning.split(",")])
    if args.tiling:
        eng.set_tiling(*[int(x) for x in args.tiling.split(",")])
    if args.inplace is not None:
        eng.set_fused_inplace(args.inplace)
    if args.depth is not None:
        eng.set_fused_depth(args.depth)
    if args.order is not None:
        eng.set_tile_order(args.order)
    if args.hints is not None:
        eng.set_cache_hints(args.hints)
    if args.autotune:
        log(f"[bench] autotune: {eng.autotune(args.autotune)}")
    tiling = eng.get_tiling()
    # the form evp1d_run takes: auto is the fused form whenever the lists tile (and ndte >= 4)
    fused = args.mode == "fused" or (args.mode == "auto" and tiling["ntiles"] > 0 and ndte >= 4)
    tuned_fused = None
    if (fused and args.tune_fused > 0 and not args.tiling and args.inplace is None and args.depth is None
            and args.order is None and args.hints is None):
        t0 = time.perf_counter()
        tuned_fused = eng.autotune_fused(args.tune_fused)
        log(f"[bench] autotune_fused ({time.perf_counter() - t0:.1f}s, setup): {tuned_fused}")
    tuning = eng.get_tuning()
    tiling = dict(eng.get_tiling(), inplace=int(eng.get_fused_inplace()), order=eng.get_tile_order())
    dinfo = eng.get_fused_depth()
    tiling["depth"] = 2 if dinfo["active"] else 1
    if dinfo["active"]:
        tiling.update(inplace=0, pair_block=dinfo["block"])   # the pair form keeps two stress buffers
        tiling["cache_hints"] = int(eng.get_cache_hints()["active"])
    log(f"[bench] mode {args.mode} ({'fused' if fused else 'unfused'}), tiling {tiling}, tuning {tuning}")
    secs = time_steps(eng, ndte, args.steps, args.warmup)
    value = na * ndte * args.steps / secs
    step_us = secs / (args.steps * ndte) * 1e6

    # each kernel alone (HIP events on the library's stream, launched one by one)
    B = algorithmic_bytes(na, nT, nU)
    B["fused"] = B["subcycle"]
    kernels = {}
    for which in ("stress", "stepu") + (("fused",) if fused else ()):
        eng.synchronize()
        eng.timer_start()
        eng.run_individual(which, args.kernel_reps)
        ms = eng.timer_stop()
        avg_us = ms * 1e3 / args.kernel_reps
        kernels[which] = {"avg_us": round(avg_us, 3), "bytes": B[which],
                          "GBps": round(B[which] / (avg_us * 1e-6) / 1e9, 1)}
        log(f"[bench] {which}: {avg_us:.2f} us/launch alone, {kernels[which]['GBps']} GB/s algorithmic")
    exact_out = output_fields(eng, s) if (fused and not args.fma and not args.no_fma) else None
    eng.close()
    # The dominant kernel's average launch duration INSIDE the timed loop:
    # the loop's HIP-event time divided by the number of its launches.  It
    # is an upper bound of the kernel's own duration (each run's state
    # copies and packing, about 1 % of it, are included), so launches x
    # avg_us equals ms_per_step.
    run_us = secs / args.steps * 1e6
    if args.mode == "persistent":   # one cooperative launch per run: the pair form's tiles when it is active
        dom = {"name": "k_pair_persistent" if tiling["depth"] == 2 else "k_persistent",
               "bytes": B["subcycle"] * ndte, "avg_us": run_us, "launches_per_step": 1}
    elif fused and tiling["depth"] == 2:
        nl = ndte / 2.0     # one k_pair launch = two subcycles (an odd ndte adds one one-subcycle launch)
        dom = {"name": "k_pair", "bytes": 2 * B["subcycle"], "avg_us": run_us / nl,
               "subcycles_per_launch": 2, "launches_per_step": nl}
    elif fused:
        dom = {"name": "k_fused", "bytes": B["fused"], "avg_us": run_us / ndte, "launches_per_step": ndte}
    else:   # two kernels per subcycle: the slower one, timed alone
        name = max(("stress", "stepu"), key=lambda k: kernels[k]["avg_us"])
        dom = {"name": "k_" + name, "bytes": B[name], "avg_us": kernels[name]["avg_us"], "launches_per_step": ndte,
               "timed": "alone (run_individual)"}
    achieved = dom["bytes"] / (dom["avg_us"] * 1e-6) / 1e9
    traffic = None
    tf = REPO / "profiles" / f"traffic_{args.workload}.json"
    if tf.exists():
        # per-launch HBM bytes from rocprofv3 FETCH_SIZE/WRITE_SIZE (profiles/); the fused
        # kernel's depend on the stress storage the run used
        tab = json.loads(tf.read_text())
        key = dom["name"]
        if key == "k_fused":
            key += "@inplace" if tiling.get("inplace") else "@twobuf"
        elif key == "k_pair":   # measured per tile shape, order and cache hints (the tuner's choice varies)
            key += f"@{tiling['tile_w']}x{tiling['tile_h']}" + ("@hilbert" if tiling.get("order") == 1 else "")
            key += "@hints" if tiling.get("cache_hints") else ""
        if args.fma:   # counters were taken on the exact build
            key += "@fma"
        traffic = tab.get(key)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom["name"],
                "bytes_per_launch": dom["bytes"], "avg_us": round(dom["avg_us"], 3),
                "avg_us_timed": dom.get("timed", "inside the timed loop (loop event time / launches)"),
                "launches_per_step": dom["launches_per_step"],
                "subcycles_per_launch": dom.get("subcycles_per_launch", 1)}
    if traffic:
        # the bytes the kernel really moves through HBM (counters) over the same duration
        phys = traffic / (dom["avg_us"] * 1e-6) / 1e9
        roofline.update({"achieved_physical": round(phys, 1), "frac_physical": round(phys / HBM_PEAK_GBS, 4),
                         "traffic_over_algorithmic": round(traffic / dom["bytes"], 3)})
        # the box's practical ceiling at the kernel's read/write mix (counters:
        # k_pair 83 % reads, k_fused 76 %, k_stress 56 %, k_stepu 94 %):
        # the streaming probe at the nearest whole-stream mix, measured now
        mix = {"k_pair": (5, 1), "k_fused": (3, 1), "k_stress": (1, 1), "k_stepu": (6, 1)}.get(dom["name"])
        if mix:
            try:
                # a kernel that runs with the cache hints is held to the higher of
                # the plain and the hinted probe
                ceil = probe_rate(*mix)
                roofline["probe"] = {"read_streams": mix[0], "write_streams": mix[1], "GBps": round(ceil, 1)}
                if tiling.get("cache_hints"):
                    ceil_nt = probe_rate(*mix, hints=1)
                    roofline["probe"]["GBps_hints"] = round(ceil_nt, 1)
                    ceil = max(ceil, ceil_nt)
                roofline["probe"]["frac_physical_of_probe"] = round(phys / ceil, 4)
                log(f"[bench] streaming probe {mix[0]}:{mix[1]}: {ceil:.1f} GB/s; {dom['name']} at "
                    f"{phys / ceil:.3f} of it")
            except Exception as e:   # reported, never fatal
                roofline["probe"] = {"error": repr(e)[:200]}
    if dom["name"] == "k_pair":
        roofline["note"] = ("frac counts the reference's bytes of two subcycles; the pair form keeps "
                            "str1..8 and the first subcycle's state on chip, so it moves fewer bytes than "
                            "that count and frac_physical is the kernel's distance to HBM bandwidth")

    fma = None
    if exact_out is not None:
        try:
            fma = fma_measure(s, ndte, args.steps, args.warmup, args.tune_fused, exact_out)
        except Exception as e:   # reported, never fatal for the exact figure
            log(f"[bench] fma measurement failed: {e!r}")
            fma = {"error": repr(e)[:200]}

    adjoint = None
    if not args.no_adjoint and not args.fma:
        try:
            adjoint = adjoint_measure(s, ndte, secs / args.steps, args.adjoint_snapshots,
                                      args.fd_directions, args.dot_directions, tiling if fused else None)
        except Exception as e:   # reported, never fatal for the forward figure
            log(f"[bench] adjoint measurement failed: {e!r}")
            adjoint = {"error": repr(e)[:200]}

    ensemble = None
    if not args.no_ensemble and not args.fma and args.workload == "evp_c1":
        try:
            # the batch tunes; the member-after-member runs take its form
            ensemble = ensemble_measure(64, 1, 0, args.tune_fused, tune_members=False)
        except Exception as e:   # reported, never fatal for the forward figure
            log(f"[bench] ensemble measurement failed: {e!r}")
            ensemble = {"error": repr(e)[:200]}

    cpu = None
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        try:
            cpu = cpu_baseline(args.workload, s, args.cpu_subcycles, threads)
            log(f"[bench] cpu baseline: {cpu}")
        except Exception as e:
            log(f"[bench] cpu baseline failed: {e!r}")
            cpu = {"value": None, "error": repr(e)[:200]}

    out = {
        "metric": "ice-pts*subcycles/s + % HBM roofline (bit-exact FP64); adjoint/forward time ratio",
        "value": value, "unit": "ice-pts*subcycles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": secs / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{args.workload} {w['nx']}x{w['ny']} ndte={ndte}", "nx": w["nx"],
                   "ny": w["ny"], "seed": w["seed"], "na": na, "navel": s["navel"], "nT": nT, "nU": nU,
                   "ndte": ndte, "mode": args.mode, "form": "fused" if fused else dom["name"] if args.mode == "persistent" else "unfused",
                   "tiling": tiling,
                   "tiling_chosen_by": ("evp1d_autotune_fused (setup, untimed)" if tuned_fused else
                                        "--tiling/--inplace/--depth/--order" if (args.tiling or args.inplace is not None
                                                                         or args.depth is not None
                                                                         or args.order is not None)
                                        else "library default"),
                   "build": "fma" if args.fma else "exact", "tuning": tuning,
                   "us_per_subcycle": round(step_us, 3),
                   "loop_GBps_algorithmic": round(B["subcycle"] / (step_us * 1e-6) / 1e9, 1)},
        "kernels": kernels,
        "roofline": roofline,
        "fma": fma,
        "adjoint": adjoint,
        "ensemble": ensemble,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
