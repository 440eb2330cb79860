"""Helpers shared by the GPU parity tests (engine <-> oracle state plumbing)."""
from __future__ import annotations

from pathlib import Path

import numpy as np

import evp1d
from oracle import oracle as O

GOLD = Path(__file__).resolve().parent / "golden"
SCALAR_NAMES = ("capping", "e_factor", "epp2i", "deltaminEVP")


def load_golden(name: str):
    """(State1D, ndte, expected stress bytes, expected stepu bytes)."""
    z = np.load(GOLD / f"{name}.npz", allow_pickle=False)
    na, navel = int(z["na"]), int(z["navel"])
    st = O.State1D(na, navel)
    for k in st.f:
        st.f[k][:] = z[f"f_{k}"]
    for k in st.idx:
        st.idx[k][:] = z[f"i_{k}"]
    st.skipT[:] = z["skipT"]
    st.skipU[:] = z["skipU"]
    st.scalars = dict(zip(SCALAR_NAMES, z["scalars"].tolist()))
    return st, int(z["ndte"]), z["out_stress"].tobytes(), z["out_stepu"].tobytes()


def golden_cases():
    """The 1D (v2b) golden cases."""
    return sorted(p.stem for p in GOLD.glob("*.npz") if not p.stem.startswith("v0_"))


def golden_cases_v0():
    """The 2D v0 golden cases (tests/golden/make_golden_v0.py)."""
    return sorted(p.stem for p in GOLD.glob("v0_*.npz"))


def load_golden_v0(case: str):
    """(State2D, ndte, output_stress bytes, output_stepu bytes) of a v0 case:
    the reference's own input files, read back with the oracle's reader."""
    import tempfile
    z = np.load(GOLD / f"{case}.npz")
    with tempfile.TemporaryDirectory() as d:
        for k in z.files:
            if k.startswith("in_"):
                name = k[3:][::-1].replace("_", ".", 1)[::-1]   # in_input_double_2d_bin -> input_double_2d.bin
                (Path(d) / name).write_bytes(z[k].tobytes())
        st = O.read_inputs_v0(Path(d), int(z["nxb"]), int(z["nyb"]), int(z["nblocks"]))
    return st, int(z["ndte"]), z["out_stress"].tobytes(), z["out_stepu"].tobytes()


def set_mode(eng: evp1d.Engine, mode: str) -> None:
    """mode: a loop form ("auto" is the library's default: the pair form where
    the lists tile), "fused" and "fused-inplace" (one subcycle per launch,
    stresses in two buffers or updated in place), "fused-pair" (two
    subcycles per launch on 32x16 tiles), "pair-hints" (the same with the
    cache hints on) or "persistent-pair" (the pair form's launches as one
    cooperative launch)."""
    if mode in ("fused", "fused-inplace"):
        eng.set_fused_depth(1)
        eng.set_fused_inplace(mode == "fused-inplace")
        mode = "fused"
    elif mode in ("fused-pair", "pair-hints", "persistent-pair"):
        eng.set_fused_depth(2)
        eng.set_tiling(32, 16, 512, 1)
        eng.set_cache_hints(1 if mode == "pair-hints" else 0)
        mode = "persistent" if mode == "persistent-pair" else "fused"
    eng.set_mode(mode)


def make_engine(st: O.State1D, ndte: int, mode: str = "auto", fma: bool = False) -> evp1d.Engine:
    p = evp1d.set_evp_parameters(ndte, fma=fma, **st.scalars)
    eng = evp1d.Engine(st.na, st.navel, p, fma=fma)
    set_mode(eng, mode)
    eng.load_state(st.f, st.idx, st.skipT, st.skipU)
    return eng


def fetch_state(eng: evp1d.Engine, like: O.State1D) -> O.State1D:
    out = like.copy()
    for k in out.f:
        out.f[k][:] = eng.get_field(k)
    return out


def oracle_params(st: O.State1D, ndte: int):
    lib = O.OracleLib()
    return lib, lib.params(ndte, **st.scalars)


def bits(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def assert_state_bitexact(got: O.State1D, want: O.State1D, fields=None) -> None:
    """Every value bit-identical; a NaN matches any NaN (IEEE 754 leaves the
    sign and payload of a generated NaN to the hardware: x86 and CDNA differ)."""
    for k in fields or want.f:
        g, w = bits(got.f[k]), bits(want.f[k])
        both_nan = np.isnan(got.f[k]) & np.isnan(want.f[k])
        bad = np.nonzero((g != w) & ~both_nan)[0]
        assert bad.size == 0, (f"{k}: {bad.size} of {w.size} values differ; first at {bad[0]}: "
                               f"got {got.f[k][bad[0]]!r} want {want.f[k][bad[0]]!r}")


def synthetic_state(nx: int, ny: int, seed: int, **kw) -> O.State1D:
    from evp1d import synth
    g = synth.make_grid(nx, ny, seed, **kw)
    ix = evp1d.build_index(g.tmask, g.nghost)
    return O.pack_dyn(g, ix)


def ensemble_member_grid(k: int):
    """Synthetic 2D state of ensemble member k (numpy only: safe to run in a
    worker process that never touches the GPU)."""
    from evp1d import synth
    m = synth.ensemble_member(k)
    return synth.make_grid(m["nx"], m["ny"], m["seed"], force_perturb=0.02)
