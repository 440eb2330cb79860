"""GPU tests of the Fortran host layer (ISO_C_BINDING -> libevp1d.so).

* evp1d_v2_driver reads the reference's own input files and must write
  output files byte-identical to the reference's (golden fixtures, the
  400x300 md5 sums), in every loop mode.
* evp1d_api_driver drives dyn_evp1d_init / dyn_evp1d_run x nrun /
  dyn_evp1d_finalize on a 2D state; its 2D results must equal the oracle run
  through the reference's own 1D construction and packing
  (oracle.index_reference / pack_dyn, integration/ice_dyn_evp1d.F90).

The drivers run as child processes.  Their persistent mode makes a
cooperative launch, which needs the GPU's global wave-sync resource; one
process holds it at a time, so no test module that runs before this one may
launch the persistent kernel in the test process itself.
"""
import json
import subprocess
import sys
import tempfile
from pathlib import Path

import re

import numpy as np
import pytest

from oracle import oracle as O
from gpu_common import GOLD, golden_cases, load_golden

pytestmark = pytest.mark.gpu
REPO = Path(__file__).resolve().parent.parent
FBIN = REPO / "cice-evp1d_amd" / "lib" / "evp1d_fortran"


def _run(exe, cwd, *args, timeout=120):
    print(f"run {exe.name} {' '.join(args)} in {cwd}", flush=True)
    if not exe.exists():
        pytest.fail(f"{exe} is not built (evp1d.build.build_fortran)")
    r = subprocess.run([str(exe), *args], cwd=str(cwd), capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r


@pytest.mark.parametrize("mode", ["graph", "persistent", "launch", "fused", "fused-inplace", "fused-pair", "tuned"])
@pytest.mark.parametrize("case", golden_cases())
def test_v2_driver_reproduces_reference_files(gpu, case, mode):
    st, ndte, want_stress, want_stepu = load_golden(case)
    with tempfile.TemporaryDirectory() as d:
        O.write_inputs(d, st, ndte)
        r = _run(FBIN / "evp1d_v2_driver", d, mode)
        assert "Benchmark took:" in r.stdout
        got_stress = (Path(d) / "output_stress_v2g_1d.bin").read_bytes()
        got_stepu = (Path(d) / "output_stepu_v2g_1d.bin").read_bytes()
    assert O.md5(got_stress) == O.md5(want_stress)
    assert O.md5(got_stepu) == O.md5(want_stepu)


@pytest.mark.parametrize("mode", ["graph", "fused-pair"])
def test_v2_driver_gradient_through_fortran_binding(gpu, mode):
    """`evp1d_v2_driver <mode> grad` calls evp1d_adjoint_run through the
    Fortran binding: the output files are still the reference's (the adjoint
    run leaves the forward state), and the gradient file equals the
    library's gradient taken through ctypes on the same inputs, target = the
    input velocities (per-point gradients bit for bit; the loss, a sum in
    slot order, to 1e-13)."""
    from gpu_common import make_engine
    case = "small_64x48"
    st, ndte, want_stress, want_stepu = load_golden(case)
    with tempfile.TemporaryDirectory() as d:
        O.write_inputs(d, st, ndte)
        r = _run(FBIN / "evp1d_v2_driver", d, mode, "grad")
        assert "gradient: loss" in r.stdout
        got_stress = (Path(d) / "output_stress_v2g_1d.bin").read_bytes()
        got_stepu = (Path(d) / "output_stepu_v2g_1d.bin").read_bytes()
        raw = np.fromfile(Path(d) / "output_gradient_v2g_1d.bin", dtype=np.float64)
    assert O.md5(got_stress) == O.md5(want_stress)
    assert O.md5(got_stepu) == O.md5(want_stepu)
    na = st.na
    assert raw.size == 1 + 4 * na
    loss, gf = raw[0], raw[1:].reshape(4, na)
    with make_engine(st, ndte, mode) as eng:
        want = eng.adjoint_run(ndte, st.f["uvel"][:na], st.f["vvel"][:na])
    for k, key in enumerate(("strength", "cdn_ocn", "forcex", "forcey")):
        assert np.array_equal(gf[k], want[key]), key
    assert np.any(gf[0] != 0.0)
    assert abs(loss - want["loss"]) <= 1e-13 * abs(want["loss"])


def test_v2_driver_evp_c0_md5(gpu):
    from evp1d import synth
    import evp1d
    rec = json.loads((GOLD / "golden.json").read_text())["evp_c0"]
    w = synth.WORKLOADS["evp_c0"]
    g = synth.make_grid(w["nx"], w["ny"], w["seed"])
    st = O.pack_dyn(g, evp1d.build_index(g.tmask, 1))
    with tempfile.TemporaryDirectory() as d:
        O.write_inputs(d, st, w["ndte"])
        _run(FBIN / "evp1d_v2_driver", d)
        assert O.md5((Path(d) / "output_stress_v2g_1d.bin").read_bytes()) == rec["output_stress_md5"]
        assert O.md5((Path(d) / "output_stepu_v2g_1d.bin").read_bytes()) == rec["output_stepu_md5"]


def test_v2_driver_rejects_short_input(gpu):
    """The reference ignores iostat and runs on garbage; this driver stops."""
    st, ndte, _, _ = load_golden("tiny_24x20")
    with tempfile.TemporaryDirectory() as d:
        O.write_inputs(d, st, ndte)
        p = Path(d) / "input_double_1d_v2.bin"
        p.write_bytes(p.read_bytes()[:-800])
        r = subprocess.run([str(FBIN / "evp1d_v2_driver")], cwd=d, capture_output=True, text=True, timeout=60)
        assert r.returncode != 0 and "short" in (r.stdout + r.stderr)


DYN = ("stressp_1", "stressp_2", "stressp_3", "stressp_4", "stressm_1", "stressm_2", "stressm_3",
       "stressm_4", "stress12_1", "stress12_2", "stress12_3", "stress12_4", "strength", "cdn_ocn",
       "aiu", "uocn", "vocn", "waterxU", "wateryU", "forcexU", "forceyU", "umassdti", "fmU",
       "strintxU", "strintyU", "Tbu", "Cb", "uvel", "vvel")
OUT = ("stressp_1", "stressp_2", "stressp_3", "stressp_4", "stressm_1", "stressm_2", "stressm_3",
       "stressm_4", "stress12_1", "stress12_2", "stress12_3", "stress12_4", "strintxU", "strintyU",
       "Cb", "uvel", "vvel")


def _emulate_api(g, ndte, nrun):
    """The reference integration's data flow, with the oracle as kernels."""
    ix = O.index_reference(g.tmask, g.nghost)
    st = O.pack_dyn(g, ix)
    i = ix["indxti"].astype(np.int64) - 1
    j = ix["indxtj"].astype(np.int64) - 1
    na = st.na
    G = {k: g.fields[k].copy() for k in DYN}
    lib = O.OracleLib()
    p = lib.params(ndte)
    for _ in range(nrun):
        for n in DYN:
            st.f[n][:na] = G[n][i, j]
        st.f["uvel"][na:] = 0.0
        st.f["vvel"][na:] = 0.0
        st.f["uvel_init"][:] = st.f["uvel"][:na]
        st.f["vvel_init"][:] = st.f["vvel"][:na]
        lib.run(p, st, ndte)
        for n in OUT:
            G[n][i, j] = st.f[n][:na]
    return G


@pytest.mark.parametrize("devpack,autotune", [(True, False), (False, False), (True, True)])
@pytest.mark.parametrize("nx,ny,seed,ndte,nrun", [(40, 30, 71, 15, 2), (90, 64, 72, 40, 3)])
def test_api_driver_matches_reference_dataflow(gpu, nx, ny, seed, ndte, nrun, devpack, autotune):
    from evp1d import synth
    g = synth.make_grid(nx, ny, seed)
    # a nonzero stress start, so the unpack of every stress matters
    rng = np.random.default_rng(seed)
    for n in DYN[:12]:
        g.fields[n] = np.asfortranarray(np.where(g.icetmask, rng.standard_normal(g.tmask.shape) * 50.0, 0.0))
    want = _emulate_api(g, ndte, nrun)
    with tempfile.TemporaryDirectory() as d:
        with open(Path(d) / "api_in.bin", "wb") as fh:
            fh.write(np.array([nx, ny, g.nghost, ndte, nrun], dtype="<i4").tobytes())
            for m in (g.tmask, g.icetmask, g.iceumask):
                fh.write(np.asfortranarray(m.astype("<i4")).tobytes(order="F"))
            for n in ("dxT", "dyT", "uarear", "HTE", "HTN"):
                fh.write(g.fields[n].tobytes(order="F"))
            for n in DYN:
                fh.write(g.fields[n].tobytes(order="F"))
        # 2D <-> 1D copies on the device (default) or in host loops
        # with lautotune the first run tunes the fused form on its own fields
        (Path(d) / "evp_1d.nml").write_text(f"&evp1d_nml\n  ldevpack = {'.true.' if devpack else '.false.'}\n"
                                            f"  lautotune = {'.true.' if autotune else '.false.'}\n/\n")
        _run(FBIN / "evp1d_api_driver", d)
        raw = np.fromfile(Path(d) / "api_out.bin", dtype="<f8")
        log = (Path(d) / "evp1d.log").read_text()
    assert "nActive" in log
    assert ("autotune_fused" in log) == autotune
    if autotune:   # the choice: tile, workgroup, depth and cache hints
        assert re.search(r"tile \d+x\d+, workgroup \d+, depth [12], cache hints [01]", log), log
    assert "Time to die for evp1d" in log      # dyn_evp1d_finalize closes the log
    shape = g.tmask.shape
    got = raw.reshape((29,) + shape[::-1]).transpose(0, 2, 1)
    for k, n in enumerate(DYN):
        a, b = got[k], want[n]
        assert np.array_equal(a.view(np.uint64), np.asfortranarray(b).view(np.uint64)), n


@pytest.mark.parametrize("mode", ["fused", "plain"])
@pytest.mark.parametrize("case", [c for c in __import__("gpu_common").golden_cases_v0()])
def test_v0_driver_reproduces_reference_files(gpu, case, mode):
    """evp1d_v0_driver on the reference's own v0 input files: both 2D output
    files byte-identical to the reference v0's; with lconvert21d the 1D files
    hold block 1's cells of the same arrays (convert_2d_1d, vars.F90:534-584)."""
    from gpu_common import load_golden_v0
    st, ndte, want_stress, want_stepu = load_golden_v0(case)
    with tempfile.TemporaryDirectory() as d:
        O.write_inputs_v0(d, st, ndte)
        nml = (Path(d) / "kernel.nml").read_text().replace("/\n", " lconvert21d = .true.\n/\n")
        (Path(d) / "kernel.nml").write_text(nml)
        r = _run(FBIN / "evp1d_v0_driver", d, mode)
        assert "Benchmark took:" in r.stdout
        got_stress = (Path(d) / "output_stress_v0_2d.bin").read_bytes()
        got_stepu = (Path(d) / "output_stepu_v0_2d.bin").read_bytes()
        s1 = (Path(d) / "output_stress_v0_1d.bin").read_bytes()
        u1 = (Path(d) / "output_stepu_v0_1d.bin").read_bytes()
    assert O.md5(got_stress) == O.md5(want_stress)
    assert O.md5(got_stepu) == O.md5(want_stepu)
    n = st.nxb * st.nyb * st.nblocks
    sa = np.frombuffer(want_stress, "<f8").reshape(12, n)
    ua = np.frombuffer(want_stepu, "<f8").reshape(4, n)
    nt, nu = int(st.icellt[0]), int(st.icellu[0])
    tc = (st.indxti[:nt, 0] - 1) + (st.indxtj[:nt, 0] - 1) * st.nxb
    uc = (st.indxui[:nu, 0] - 1) + (st.indxuj[:nu, 0] - 1) * st.nxb
    assert s1 == sa[:, tc].tobytes()
    assert u1 == ua[:, uc].tobytes()


def test_v0_driver_evp_c0_md5(gpu):
    from evp1d import synth
    rec = json.loads((GOLD / "golden.json").read_text())["evp_c0_v0"]
    g = synth.make_grid(rec["nx"], rec["ny"], rec["seed"])
    st = O.pack_v0(g, 1)
    with tempfile.TemporaryDirectory() as d:
        O.write_inputs_v0(d, st, rec["ndte"])
        _run(FBIN / "evp1d_v0_driver", d)
        assert O.md5((Path(d) / "output_stress_v0_2d.bin").read_bytes()) == rec["output_stress_md5"]
        assert O.md5((Path(d) / "output_stepu_v0_2d.bin").read_bytes()) == rec["output_stepu_md5"]
