"""Write the golden fixtures of the EVP1D parity tests (run in a container
that has /root/reference; the fixtures are committed, the reference is not).

For each case: synthetic 2D state (evp1d.synth) -> 1D list and packing as the
reference's init does (oracle.index_reference / oracle.pack_dyn) -> the
reference's own input files -> the flang-built reference v2b at f0 flags
(oracle/_ref/evp_v2b_f0, `make -C oracle ref`) -> its two output files.
Stored per case in <case>.npz: the 1D inputs (fields, neighbour lists, skip
flags, scalars, ndte) and the bytes of output_stress_v2b_1d.bin and
output_stepu_v2b_1d.bin.  Large workloads store only md5 sums (golden.json).

    python tests/golden/make_golden.py            # small cases + evp_c0
    python tests/golden/make_golden.py evp_c1     # add the 1280x1520 md5s
"""
import json
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

# name: (make_grid kwargs, ndte, scalar overrides)
SMALL = {
    "tiny_24x20": (dict(nx=24, ny=20, seed=7), 3, {}),
    "small_64x48": (dict(nx=64, ny=48, seed=11), 25, {}),
    "noice_20x16": (dict(nx=20, ny=16, seed=5, ice_prob=0.0), 2, {}),
    "allice_20x16": (dict(nx=20, ny=16, seed=6, land_frac=0.0, ice_prob=1.0), 5, {}),
    "ragged_40x7": (dict(nx=40, ny=7, seed=9, land_frac=0.5, sigma=2.0), 10, {}),
    "scalars_32x24": (dict(nx=32, ny=24, seed=13), 8,
                      dict(capping=0.5, e_factor=0.3, epp2i=0.2, deltaminEVP=2e-9)),
    "long_16x12": (dict(nx=16, ny=12, seed=17), 400, {}),
}


def build_state(kw, scal):
    g = synth.make_grid(**kw)
    ix = O.index_reference(g.tmask, g.nghost)
    st = O.pack_dyn(g, ix)
    st.scalars.update(scal)
    return g, ix, st


def input_md5(st, ndte):
    with tempfile.TemporaryDirectory() as d:
        O.write_inputs(d, st, ndte)
        return {f: O.md5((Path(d) / f).read_bytes()) for f in
                ("input_double_1d_v2.bin", "input_integer_1d.bin", "input_logical_1d.bin")}


def run_case(st, ndte):
    with tempfile.TemporaryDirectory() as d:
        O.write_inputs(d, st, ndte)
        O.run_reference(Path(d), "evp_v2b_f0", timeout=7200)
        a = (Path(d) / "output_stress_v2b_1d.bin").read_bytes()
        b = (Path(d) / "output_stepu_v2b_1d.bin").read_bytes()
    return a, b


def save_small(name):
    kw, ndte, scal = SMALL[name]
    g, ix, st = build_state(kw, scal)
    a, b = run_case(st, ndte)
    arrs = {f"f_{k}": v for k, v in st.f.items()}
    arrs.update({f"i_{k}": v for k, v in st.idx.items()})
    np.savez_compressed(HERE / f"{name}.npz", **arrs, skipT=st.skipT, skipU=st.skipU,
                        scalars=np.array([st.scalars[k] for k in ("capping", "e_factor", "epp2i", "deltaminEVP")]),
                        ndte=np.int32(ndte), na=np.int32(st.na), navel=np.int32(st.navel),
                        out_stress=np.frombuffer(a, np.uint8), out_stepu=np.frombuffer(b, np.uint8),
                        indxtij=ix["indxtij"].astype(np.int64))
    print(name, st.na, st.navel, O.md5(a), O.md5(b), flush=True)


def md5_workload(name):
    w = dict(synth.WORKLOADS[name])
    ndte = w.pop("ndte")
    g, ix, st = build_state(w, {})
    rec = dict(nx=w["nx"], ny=w["ny"], seed=w["seed"], ndte=ndte, na=st.na, navel=st.navel,
               nT=int((st.skipT == 0).sum()), nU=int((st.skipU == 0).sum()),
               inputs=input_md5(st, ndte))
    a, b = run_case(st, ndte)
    rec["output_stress_md5"] = O.md5(a)
    rec["output_stepu_md5"] = O.md5(b)
    print(name, rec, flush=True)
    return rec


if __name__ == "__main__":
    gj = HERE / "golden.json"
    names = sys.argv[1:] or list(SMALL) + ["evp_c0"]
    for n in names:
        if n in SMALL:
            save_small(n)
        else:
            rec = md5_workload(n)
            # re-read before writing: several workloads may run in parallel processes
            golden = json.loads(gj.read_text()) if gj.exists() else {}
            golden[n] = rec
            gj.write_text(json.dumps(golden, indent=1, sort_keys=True) + "\n")
