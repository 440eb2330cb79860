"""GPU parity of the v0 2D block path (evp2d_*, csrc/evp2d_device.hip).

Bit-exact against the compiled reference v0 (golden output files, 400x300
md5s), against the v0 oracle on every field including taubx/tauby and the
str1..8 scratch, and against the 1D engine on the same grid.
"""
import json
import sys

import numpy as np
import pytest

import evp1d
from evp1d import synth
from oracle import oracle as O
from gpu_common import GOLD, assert_state_bitexact, fetch_state, golden_cases_v0, load_golden_v0, make_engine

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _params(st, ndte):
    p = evp1d.set_evp_parameters(ndte)
    for k in ("capping", "e_factor", "epp2i"):
        setattr(p, k, st.scalars[k])
    return p


MODES2D = ("fused", "fused-hints", "plain")


def _set_mode(eng, mode):
    """"fused" (cache hints chosen by size: off at these sizes), "fused-hints"
    (forced on) or "plain"."""
    eng.set_mode("fused" if mode.startswith("fused") else mode)
    if mode == "fused-hints":
        eng.set_cache_hints(1)


def _engine(st, ndte, mode="fused"):
    eng = evp1d.Engine2D(st.nxb, st.nyb, st.nblocks, _params(st, ndte))
    _set_mode(eng, mode)
    eng.load_state(st)
    return eng


def _fetch(eng, like):
    out = like.copy()
    for n in O.V0_FIELDS:
        out.f[n][:] = eng.get_field(n)
    return out


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


@pytest.mark.parametrize("mode", MODES2D)
@pytest.mark.parametrize("case", golden_cases_v0())
def test_v0_golden_outputs_bitexact(gpu, case, mode):
    _log(f"v0 golden {case} {mode}")
    st, ndte, want_a, want_b = load_golden_v0(case)
    with _engine(st, ndte, mode) as eng:
        eng.run(ndte)
        got = _fetch(eng, st)
    a, b = O.output_bytes_v0(got)
    assert O.md5(a) == O.md5(want_a)
    assert O.md5(b) == O.md5(want_b)


@pytest.mark.parametrize("mode", MODES2D)
def test_v0_c0_md5(gpu, mode):
    rec = json.loads((GOLD / "golden.json").read_text())["evp_c0_v0"]
    g = synth.make_grid(rec["nx"], rec["ny"], rec["seed"])
    st = O.pack_v0(g, 1)
    with _engine(st, rec["ndte"], mode) as eng:
        eng.run(rec["ndte"])
        got = _fetch(eng, st)
    a, b = O.output_bytes_v0(got)
    assert O.md5(a) == rec["output_stress_md5"]
    assert O.md5(b) == rec["output_stepu_md5"]


@pytest.mark.parametrize("mode", MODES2D)
@pytest.mark.parametrize("nblocks,ndte", [(1, 7), (3, 12), (4, 1), (2, 2)])
def test_v0_every_field_matches_the_oracle(gpu, nblocks, ndte, mode):
    """All arrays after the run, taubx/tauby included; with one block also
    the str1..8 scratch (with several the reference keeps only one block of
    it, the engine one per block)."""
    g = synth.make_grid(50, 33, 41 + nblocks)
    st = O.pack_v0(g, nblocks)
    want = st.copy()
    V = O.OracleV0()
    p = V.params(ndte)
    V.run(p, want, ndte)
    _log(f"v0 fields nblocks={nblocks} ndte={ndte} {mode}")
    with _engine(st, ndte, mode) as eng:
        eng.run(ndte)
        for n in O.V0_FIELDS:
            assert _bits_equal(eng.get_field(n), want.f[n]), n
        if nblocks == 1:
            # the oracle's scratch after its last stress call
            scratch = np.zeros(8 * st.nxb * st.nyb)
            one = want.copy()
            # re-run the last subcycle's stress on the state before it
            prev = st.copy()
            V.run(p, prev, ndte - 1)
            sig = [prev.f[n][:, :, 0].reshape(-1, order="F") for n in O.STRESS_FIELDS]
            sig = [np.ascontiguousarray(s) for s in sig]
            import ctypes
            sp = (ctypes.POINTER(ctypes.c_double) * 12)(*[O._dp(s) for s in sig])
            F = {n: np.ascontiguousarray(prev.f[n][:, :, 0].reshape(-1, order="F")) for n in O.V0_FIELDS}
            V.lib.oracle_v0_stress(ctypes.byref(p), st.nxb, st.nyb, int(st.icellt[0]),
                                   O._ip(np.ascontiguousarray(st.indxti[:, 0])),
                                   O._ip(np.ascontiguousarray(st.indxtj[:, 0])), O._dp(F["uvel"]),
                                   O._dp(F["vvel"]), O._dp(F["dxt"]), O._dp(F["dyt"]), O._dp(F["dxhy"]),
                                   O._dp(F["dyhx"]), O._dp(F["cxp"]), O._dp(F["cyp"]), O._dp(F["cxm"]),
                                   O._dp(F["cym"]), O._dp(F["DminTarea"]), O._dp(F["strength"]), sp,
                                   O._dp(scratch))
            planes = scratch.reshape(8, -1)
            for k in range(8):
                got = eng.get_field(f"str{k + 1}").reshape(-1, order="F")
                assert _bits_equal(got, planes[k]), f"str{k + 1}"


@pytest.mark.parametrize("mode", MODES2D)
@pytest.mark.parametrize("nx,ny,seed,ndte", [(64, 48, 51, 30), (120, 40, 52, 15)])
def test_v0_engine_equals_1d_engine(gpu, nx, ny, seed, ndte, mode):
    """The same grid through both device paths gives the same bits."""
    g = synth.make_grid(nx, ny, seed)
    st = O.pack_v0(g, 1)
    ix = evp1d.build_index(g.tmask, g.nghost)
    s1 = O.pack_dyn(g, ix)
    with _engine(st, ndte, mode) as e2:
        e2.run(ndte)
        got2 = _fetch(e2, st)
    with make_engine(s1, ndte) as e1:
        e1.run(ndte)
        got1 = fetch_state(e1, s1)
    i = ix["indxti"].astype(np.int64) - 1
    j = ix["indxtj"].astype(np.int64) - 1
    keep = s1.skipU == 0
    for n in O.STRESS_FIELDS:
        assert _bits_equal(got2.f[n][:, :, 0][i, j], got1.f[n][:s1.na]), n
    for n2, n1 in (("uvel", "uvel"), ("vvel", "vvel"), ("strintx", "strintxU"), ("strinty", "strintyU")):
        assert _bits_equal(got2.f[n2][:, :, 0][i, j][keep], got1.f[n1][:s1.na][keep]), n2


@pytest.mark.parametrize("mode", MODES2D)
def test_v0_single_calls_and_restart(gpu, mode):
    """stress/stepu/halo one by one equal run; a run continues from the end
    state, in either form, odd and even subcycle counts."""
    g = synth.make_grid(40, 30, 53)
    st = O.pack_v0(g, 3)
    want = st.copy()
    V = O.OracleV0()
    V.run(V.params(9), want, 9)
    with _engine(st, 9, mode) as eng:
        for _ in range(4):
            eng.stress()
            eng.stepu()
            eng.halo()
        eng.run(2)
        eng.run(3)
        for n in O.V0_FIELDS:
            assert _bits_equal(eng.get_field(n), want.f[n]), n


def test_v0_parameters_change_between_runs(gpu):
    g = synth.make_grid(36, 28, 54)
    st = O.pack_v0(g, 1)
    want = st.copy()
    V = O.OracleV0()
    p1 = V.params(10)
    p2 = V.params(10, capping=0.5, e_factor=0.3)
    V.run(p1, want, 4)
    V.run(p2, want, 5)
    with _engine(st, 10) as eng:
        eng.run(4)
        q = evp1d.set_evp_parameters(10)
        q.capping, q.e_factor = 0.5, 0.3
        eng.set_params(q)
        eng.run(5)
        for n in ("uvel", "vvel", "stressp_1", "stress12_4", "strintx"):
            assert _bits_equal(eng.get_field(n), want.f[n]), n


def test_v0_bad_inputs_rejected(gpu):
    g = synth.make_grid(20, 16, 55)
    st = O.pack_v0(g, 2)
    with _engine(st, 3) as eng:
        nxb, nyb = st.nxb, st.nyb
        with pytest.raises(evp1d.Evp1dError):
            eng.set_cells(1, [1], [5], [], [])            # T cell on the first column: stress reads i-1
        with pytest.raises(evp1d.Evp1dError):
            eng.set_cells(1, [], [], [nxb], [3])          # U cell on the last column: stepu reads i+1
        with pytest.raises(evp1d.Evp1dError):
            eng.set_cells(1, [4, 4], [5, 5], [], [])      # listed twice
        with pytest.raises(evp1d.Evp1dError):
            eng.set_cells(3, [], [], [], [])              # no block 3
        with pytest.raises(evp1d.Evp1dError):
            eng.set_field("str1", np.zeros(eng.n))         # scratch
        short = np.zeros(eng.n - 1)                         # size, checked by the library itself
        assert eng.L.evp2d_set_field(eng.h, evp1d.FIELD2D_ID["uvel"], evp1d._dp(short), short.size) != 0
        h = st.halo
        src, dst = h["src"].copy(), h["dst"].copy()
        src[:, 1] = dst[:, 0]                              # copy 2 reads what copy 1 writes
        with pytest.raises(evp1d.Evp1dError):
            eng.set_halo(h["ilo"], h["jlo"], h["ihi"], h["jhi"], src, dst)
        with pytest.raises(evp1d.Evp1dError):
            eng.set_halo(h["ilo"], h["jlo"], np.full_like(h["ihi"], nxb), h["jhi"], h["src"], h["dst"])
        for bad in (-2, 2):
            with pytest.raises(evp1d.Evp1dError):
                eng.set_cache_hints(bad)
        assert eng.get_cache_hints() == {"hints": -1, "active": False}
        # the handle still works after the rejections
        want = st.copy()
        V = O.OracleV0()
        V.run(V.params(3), want, 3)
        eng.run(3)
        assert _bits_equal(eng.get_field("uvel"), want.f["uvel"])


def test_v0_fused_one_by_one_and_mode_switch(gpu):
    """Fused subcycles launched one by one (the bench's kernel timing) and a
    switch of form between runs keep the results."""
    g = synth.make_grid(70, 45, 56)
    st = O.pack_v0(g, 2)
    want = st.copy()
    V = O.OracleV0()
    V.run(V.params(11), want, 11)
    with _engine(st, 11, "fused") as eng:
        eng.run_individual("fused", 4)
        eng.set_mode("plain")
        assert eng.get_mode() == "plain"
        eng.run(3)
        eng.set_mode("fused")
        eng.run(4)
        for n in O.V0_FIELDS:
            assert _bits_equal(eng.get_field(n), want.f[n]), n


@pytest.mark.parametrize("mode", MODES2D)
def test_v0_fma_build_within_1e12(gpu, mode):
    """libevp1d_fma.so's 2D path (contraction on): max|d|/max|ref| <= 1e-12
    per array against the exact oracle."""
    g = synth.make_grid(60, 40, 57)
    st = O.pack_v0(g, 2)
    want = st.copy()
    V = O.OracleV0()
    V.run(V.params(25), want, 25)
    eng = evp1d.Engine2D(st.nxb, st.nyb, st.nblocks, _params(st, 25), fma=True)
    with eng:
        _set_mode(eng, mode)
        eng.load_state(st)
        eng.run(25)
        worst = 0.0
        for n in O.V0_FIELDS:
            a, b = eng.get_field(n), want.f[n]
            den = np.max(np.abs(b))
            if den > 0:
                worst = max(worst, float(np.max(np.abs(a - b)) / den))
    assert worst <= 1e-12, worst
