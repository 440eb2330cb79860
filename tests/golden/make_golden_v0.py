"""Write the golden fixtures of the v0 2D block path (run in a container
that has /root/reference; the fixtures are committed, the reference is not).

For each case: synthetic 2D state (evp1d.synth) -> the v0 arrays, cell lists
and, with more than one block, the halo copy list (oracle.pack_v0) -> the
reference's own v0 input files (input_double_2d.bin, input_integer_2d.bin,
the two halo files; oracle.write_inputs_v0) -> the flang-built reference v0
at f0 flags (oracle/_ref/evp_v0_f0, `make -C oracle ref`) -> its two output
files.  Stored per case in v0_<case>.npz: the input files' bytes and the
bytes of output_stress_v0_2d.bin and output_stepu_v0_2d.bin; the 400x300
workload stores md5 sums only (golden.json, key "evp_c0_v0").

    python tests/golden/make_golden_v0.py
"""
import json
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
sys.path.insert(0, str(REPO / "cice-evp1d_amd"))
sys.path.insert(0, str(REPO))

from evp1d import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402

# name: (make_grid kwargs, nblocks, ndte, scalar overrides)
CASES = {
    "v0_small_32x24": (dict(nx=32, ny=24, seed=21), 1, 20, {}),
    "v0_blocks3_32x24": (dict(nx=32, ny=24, seed=21), 3, 20, {}),
    "v0_blocks4_ragged_36x9": (dict(nx=36, ny=9, seed=23, land_frac=0.5, sigma=2.0), 4, 12, {}),
    "v0_noice_20x16": (dict(nx=20, ny=16, seed=5, ice_prob=0.0), 1, 2, {}),
    "v0_scalars_24x20": (dict(nx=24, ny=20, seed=25), 1, 8, dict(capping=0.5, e_factor=0.3, epp2i=0.2)),
}
INPUT_FILES = ("input_double_2d.bin", "input_integer_2d.bin", "base_output_integer_blockindices_ref_v0_2d.bin",
               "base_output_integer_halo_ref_v0_2d.bin", "kernel.nml")


def build(kw, nblocks, scal):
    g = synth.make_grid(**kw)
    st = O.pack_v0(g, nblocks)
    st.scalars.update(scal)
    return g, st


def run_reference_v0(st, ndte):
    with tempfile.TemporaryDirectory() as d:
        O.write_inputs_v0(d, st, ndte)
        exe = O.REFDIR / "evp_v0_f0"
        subprocess.run([str(exe)], cwd=d, capture_output=True, text=True, timeout=7200, check=True)
        files = {f: (Path(d) / f).read_bytes() for f in INPUT_FILES if (Path(d) / f).exists()}
        a = (Path(d) / "output_stress_v0_2d.bin").read_bytes()
        b = (Path(d) / "output_stepu_v0_2d.bin").read_bytes()
    return files, a, b


def save_case(name):
    kw, nblocks, ndte, scal = CASES[name]
    g, st = build(kw, nblocks, scal)
    files, a, b = run_reference_v0(st, ndte)
    arrs = {"in_" + f.replace(".", "_"): np.frombuffer(v, np.uint8) for f, v in files.items()}
    np.savez_compressed(HERE / f"{name}.npz", **arrs, nxb=np.int32(st.nxb), nyb=np.int32(st.nyb),
                        nblocks=np.int32(nblocks), ndte=np.int32(ndte),
                        out_stress=np.frombuffer(a, np.uint8), out_stepu=np.frombuffer(b, np.uint8))
    print(name, st.nxb, st.nyb, nblocks, st.icellt, O.md5(a), O.md5(b), flush=True)


def md5_c0():
    w = dict(synth.WORKLOADS["evp_c0"])
    ndte = w.pop("ndte")
    g, st = build(w, 1, {})
    files, a, b = run_reference_v0(st, ndte)
    rec = dict(nx=w["nx"], ny=w["ny"], seed=w["seed"], ndte=ndte, nblocks=1,
               icellt=int(st.icellt[0]), icellu=int(st.icellu[0]),
               inputs={f: O.md5(v) for f, v in files.items() if f != "kernel.nml"},
               output_stress_md5=O.md5(a), output_stepu_md5=O.md5(b))
    print("evp_c0_v0", rec, flush=True)
    return rec


if __name__ == "__main__":
    for n in CASES:
        save_case(n)
    gj = HERE / "golden.json"
    golden = json.loads(gj.read_text()) if gj.exists() else {}
    golden["evp_c0_v0"] = md5_c0()
    gj.write_text(json.dumps(golden, indent=1, sort_keys=True) + "\n")
