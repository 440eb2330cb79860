"""GPU parity of the HIP EVP1D kernels (bit-exact against the reference).

Golden fixtures come from the flang-built reference v2b at f0 flags
(tests/golden/make_golden.py); random states are checked against the C
oracle, which is itself pinned to the reference (tests/test_oracle.py).
Exact build: every output bit must match.  FMA build: norm-wise relative
error max|d|/max|ref| <= 1e-12 per field.
"""
import json
import sys

import numpy as np
import pytest

import evp1d
from oracle import oracle as O
from gpu_common import (GOLD, assert_state_bitexact, fetch_state, golden_cases, load_golden,
                        make_engine, oracle_params, set_mode, synthetic_state)

pytestmark = pytest.mark.gpu
MODES = ("auto", "launch", "graph", "persistent", "fused", "fused-inplace", "fused-pair", "pair-hints",
         "persistent-pair")


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("case", golden_cases())
def test_golden_outputs_bitexact(gpu, case, mode):
    """Both reference output files, byte for byte, for every golden case."""
    _log(f"golden {case} mode={mode}")
    st, ndte, want_stress, want_stepu = load_golden(case)
    with make_engine(st, ndte, mode) as eng:
        eng.run(ndte)
        got = fetch_state(eng, st)
    a, b = O.output_bytes(got)
    assert O.md5(a) == O.md5(want_stress)
    assert O.md5(b) == O.md5(want_stepu)


@pytest.mark.parametrize("case", ["small_64x48", "scalars_32x24", "ragged_40x7"])
def test_full_state_matches_oracle(gpu, case):
    """Every field (incl. str1..8 at all navel slots, Cb, strint) vs the oracle."""
    st, ndte, _, _ = load_golden(case)
    lib, p = oracle_params(st, ndte)
    want = st.copy()
    lib.run(p, want, ndte)
    with make_engine(st, ndte, "graph") as eng:
        eng.run(ndte)
        got = fetch_state(eng, st)
    assert_state_bitexact(got, want)


def test_single_kernels_match_oracle(gpu):
    """evp1d_stress and evp1d_stepu, one call each (lb=1, ub=na), vs oracle."""
    st = synthetic_state(96, 80, seed=21)
    lib, p = oracle_params(st, 7)
    lib.run(p, st, 1)                     # start from a non-zero stress state
    want = st.copy()
    with make_engine(st, 7) as eng:
        eng.stress()
        lib.stress(p, want)
        assert_state_bitexact(fetch_state(eng, st), want)
        eng.stepu()
        lib.stepu(p, want)
        assert_state_bitexact(fetch_state(eng, st), want)


@pytest.mark.parametrize("mode", MODES)
def test_repeated_runs_and_restart(gpu, mode):
    """run(n1); run(n2) equals one oracle run of n1+n2 (state stays on device)."""
    st = synthetic_state(70, 50, seed=31)
    lib, p = oracle_params(st, 30)
    want = st.copy()
    lib.run(p, want, 30)
    with make_engine(st, 30, mode) as eng:
        eng.run(12)
        eng.run(18)
        got = fetch_state(eng, st)
    assert_state_bitexact(got, want)


def test_halo_copy_list(gpu):
    """Ghost copy list (vars.F90:100-143 in 1D form) vs the oracle's halo."""
    st = synthetic_state(60, 44, seed=41)
    rng = np.random.default_rng(5)
    ghosts = np.arange(st.na + 1, st.navel + 1, dtype=np.int32)
    dst = rng.choice(ghosts, size=ghosts.size // 2, replace=False).astype(np.int32)
    src = rng.integers(0, st.na + 1, size=dst.size).astype(np.int32)   # 0 = fill with zero
    lib, p = oracle_params(st, 9)
    want = st.copy()
    lib.run(p, want, 9, halo=(dst, src))
    for mode in MODES:
        with make_engine(st, 9, mode) as eng:
            eng.set_halo(dst, src)
            if mode in ("fused-pair", "pair-hints", "persistent-pair"):   # the copy list runs between the two subcycles of a launch
                assert eng.get_fused_depth()["active"]
            assert eng.get_cache_hints()["active"] == (mode == "pair-hints")
            eng.run(9)
            got = fetch_state(eng, st)
        assert_state_bitexact(got, want)


@pytest.mark.parametrize("tile,order,nsub", [((31, 19), 1, 8), ((15, 15), 0, 9), ((5, 7), 1, 4)])
def test_halo_copy_list_pair_form(gpu, tile, order, nsub):
    """The pair form with a ghost copy list (the copy made after each
    subcycle, the first of a launch's two included) on several tiles, tile
    orders and subcycle counts: bit-identical to the oracle's halo run."""
    st = synthetic_state(90, 70, seed=43)
    rng = np.random.default_rng(7)
    ghosts = np.arange(st.na + 1, st.navel + 1, dtype=np.int32)
    dst = rng.choice(ghosts, size=2 * ghosts.size // 3, replace=False).astype(np.int32)
    src = rng.integers(0, st.na + 1, size=dst.size).astype(np.int32)   # 0 = fill with zero
    lib, p = oracle_params(st, nsub)
    want = st.copy()
    lib.run(p, want, nsub, halo=(dst, src))
    _log(f"halo pair tile={tile} order={order} nsub={nsub}")
    with make_engine(st, nsub, "fused") as eng:
        eng.set_tiling(tile[0], tile[1], 512, 1)
        eng.set_fused_depth(2)
        eng.set_tile_order(order)
        eng.set_halo(dst, src)
        eng.run(nsub)
        assert eng.get_fused_depth()["active"]
        got = fetch_state(eng, st)
    assert_state_bitexact(got, want)


def test_evp_c0_md5(gpu):
    """The 400x300, ndte=120 workload: both output files' md5 equal the
    compiled reference's (tests/golden/golden.json)."""
    from evp1d import synth
    rec = json.loads((GOLD / "golden.json").read_text())["evp_c0"]
    w = synth.WORKLOADS["evp_c0"]
    st = synthetic_state(w["nx"], w["ny"], w["seed"])
    assert (st.na, st.navel) == (rec["na"], rec["navel"])
    # the regenerated inputs must be the ones the reference ran on
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        O.write_inputs(d, st, w["ndte"])
        from pathlib import Path
        for f, m in rec["inputs"].items():
            assert O.md5((Path(d) / f).read_bytes()) == m, f
    for mode in MODES:
        _log(f"evp_c0 mode={mode}")
        with make_engine(st, w["ndte"], mode) as eng:
            eng.run(w["ndte"])
            got = fetch_state(eng, st)
        a, b = O.output_bytes(got)
        assert O.md5(a) == rec["output_stress_md5"]
        assert O.md5(b) == rec["output_stepu_md5"]


def _normwise(a, b):
    d = np.max(np.abs(a - b)) if a.size else 0.0
    r = np.max(np.abs(b)) if b.size else 0.0
    return d / r if r > 0 else d


@pytest.mark.parametrize("case", ["small_64x48", "long_16x12"])
def test_fma_build_within_1e12(gpu, case):
    """libevp1d_fma.so (contraction on): max|d|/max|ref| <= 1e-12 per field."""
    st, ndte, _, _ = load_golden(case)
    lib, p = oracle_params(st, ndte)
    want = st.copy()
    lib.run(p, want, ndte)
    assert evp1d.build_info(fma=True)["fma"] == 1
    with make_engine(st, ndte, "graph", fma=True) as eng:
        eng.run(ndte)
        got = fetch_state(eng, st)
    worst = max(_normwise(got.f[k], want.f[k]) for k in want.f)
    assert worst <= 1e-12, worst


def test_bad_inputs_rejected(gpu):
    st = synthetic_state(20, 16, seed=3)
    p = evp1d.set_evp_parameters(5)
    with evp1d.Engine(st.na, st.navel, p) as eng:
        bad = st.idx["ee"].copy()
        bad[3] = st.navel + 1
        with pytest.raises(evp1d.Evp1dError) as e:
            eng.set_index("ee", bad)
        assert e.value.code == 2
        with pytest.raises(evp1d.Evp1dError) as e:
            eng.run(3)                       # lists not set yet
        assert e.value.code == 5
        with pytest.raises(evp1d.Evp1dError):
            eng.set_field("uvel", np.zeros(st.na))   # uvel has navel entries
        with pytest.raises(evp1d.Evp1dError):
            eng.set_halo([1], [1])                   # slot 1 is not a ghost
    with pytest.raises(evp1d.Evp1dError):
        evp1d.Engine(10, 5, p)


@pytest.mark.parametrize("mode", MODES)
def test_skip_flags_change_between_runs(gpu, mode):
    """New skip flags (a new model step's ice masks, set_skipMe) re-partition
    the device slots; results stay bit-identical to the oracle."""
    st = synthetic_state(80, 60, seed=51)
    lib, p = oracle_params(st, 20)
    want = st.copy()
    lib.run(p, want, 10)
    rng = np.random.default_rng(9)
    # the ice retreats: more points skipped (activating a U point without ice
    # mass would divide 0/0 in stepu, as in the reference)
    newT = (st.skipT | (rng.random(st.na) < 0.3)).astype(np.int32)
    newU = (st.skipU | (rng.random(st.na) < 0.2)).astype(np.int32)
    want.skipT[:] = newT
    want.skipU[:] = newU
    lib.run(p, want, 10)
    with make_engine(st, 20, mode) as eng:
        eng.run(10)
        eng.set_skip("skipT", newT)
        eng.set_skip("skipU", newU)
        eng.run(10)
        got = fetch_state(eng, want)
    assert_state_bitexact(got, want)


@pytest.mark.parametrize("mode", ["auto", "persistent-pair", "fused"])
def test_skip_flags_change_with_a_copy_list(gpu, mode):
    """New skip flags between runs with a ghost copy list set: the re-layout
    re-plans the pair form's tables, copy-list sources included; results stay
    bit-identical to the oracle's halo run (odd subcycle counts)."""
    st = synthetic_state(70, 52, seed=57)
    rng = np.random.default_rng(19)
    ghosts = np.arange(st.na + 1, st.navel + 1, dtype=np.int32)
    dst = rng.choice(ghosts, size=ghosts.size // 2, replace=False).astype(np.int32)
    src = rng.integers(0, st.na + 1, size=dst.size).astype(np.int32)
    lib, p = oracle_params(st, 9)
    want = st.copy()
    lib.run(p, want, 9, halo=(dst, src))
    newT = (st.skipT | (rng.random(st.na) < 0.3)).astype(np.int32)
    newU = (st.skipU | (rng.random(st.na) < 0.2)).astype(np.int32)
    want.skipT[:] = newT
    want.skipU[:] = newU
    lib.run(p, want, 7, halo=(dst, src))
    _log(f"skip change with copy list mode={mode}")
    with make_engine(st, 9, mode) as eng:
        eng.set_halo(dst, src)
        eng.run(9)
        eng.set_skip("skipT", newT)
        eng.set_skip("skipU", newU)
        eng.run(7)
        if mode != "fused":
            assert eng.get_fused_depth()["active"]
        got = fetch_state(eng, want)
    assert_state_bitexact(got, want)


def test_fields_roundtrip_through_layout(gpu):
    """set_field/get_field return the caller's order under any skip pattern."""
    st = synthetic_state(40, 30, seed=61)
    with make_engine(st, 3) as eng:
        rng = np.random.default_rng(1)
        for name in ("stressm_3", "uvel", "str6", "forcexU", "Cb"):
            a = rng.standard_normal(eng.length(name))
            eng.set_field(name, a)
            eng.set_skip("skipT", (rng.random(st.na) < 0.5).astype(np.int32))
            assert np.array_equal(eng.get_field(name), a), name


def test_batch_members_match_independent_oracle_runs(gpu):
    """evp1d_create_batch: uneven members in one handle; each member's result
    is bit-identical to its own oracle run (member-local inputs and indices)."""
    sizes = [(30, 22, 81), (64, 40, 82), (12, 9, 83), (50, 70, 84), (5, 4, 85)]
    states = [synthetic_state(nx, ny, seed, force_perturb=0.02) for nx, ny, seed in sizes]
    ndte = 17
    lib, p = oracle_params(states[0], ndte)
    wants = []
    for st in states:
        w = st.copy()
        lib.run(p, w, ndte)
        wants.append(w)
    params = evp1d.set_evp_parameters(ndte)
    for mode in MODES:
        with evp1d.Engine.batch([(st.na, st.navel) for st in states], params) as eng:
            set_mode(eng, mode)
            for k, st in enumerate(states):
                eng.load_member(k, st.f, st.idx, st.skipT, st.skipU)
            info = eng.member_info(3)
            assert info["active_offset"] == sum(s.na for s in states[:3])
            assert info["ghost_offset"] == eng.na + sum(s.navel - s.na for s in states[:3])
            eng.run(ndte)
            for k, st in enumerate(states):
                got = st.copy()
                for n in got.f:
                    got.f[n][:] = eng.get_field_member(k, n)
                assert_state_bitexact(got, wants[k])


@pytest.mark.parametrize("mode", ["auto", "fused-pair", "pair-hints", "persistent-pair", "graph"])
def test_batch_with_copy_lists_matches_member_oracle_runs(gpu, mode):
    """A batch whose ghost copy list is the union of member-local lists (each
    member's ghosts copy that member's active points or zero), written in
    the handle's global numbering through member_info: every member
    bit-identical to its own oracle run with its own list, over an odd
    number of subcycles (the pair forms start with one one-subcycle launch
    and fold the copy after each launch's first subcycle into their plan)."""
    sizes = [(40, 30, 61), (64, 40, 62), (9, 7, 63), (50, 44, 64)]
    states = [synthetic_state(nx, ny, seed, force_perturb=0.02) for nx, ny, seed in sizes]
    ndte = 11
    lib, p = oracle_params(states[0], ndte)
    params = evp1d.set_evp_parameters(ndte)
    rng = np.random.default_rng(17)
    local = []
    for st in states:
        ghosts = np.arange(st.na + 1, st.navel + 1, dtype=np.int32)
        dst = rng.choice(ghosts, size=(2 * ghosts.size) // 3, replace=False).astype(np.int32)
        src = rng.integers(0, st.na + 1, size=dst.size).astype(np.int32)   # 0 = fill with zero
        local.append((dst, src))
    _log(f"batch with copy lists mode={mode}")
    with evp1d.Engine.batch([(st.na, st.navel) for st in states], params) as eng:
        set_mode(eng, mode)
        for k, st in enumerate(states):
            eng.load_member(k, st.f, st.idx, st.skipT, st.skipU)
        gd, gs = [], []
        for k, (st, (dst, src)) in enumerate(zip(states, local)):
            info = eng.member_info(k)
            gd.append(info["ghost_offset"] + (dst - st.na))
            gs.append(np.where(src > 0, info["active_offset"] + src, 0))
        eng.set_halo(np.concatenate(gd).astype(np.int32), np.concatenate(gs).astype(np.int32))
        if mode in ("fused-pair", "pair-hints", "persistent-pair"):
            assert eng.get_fused_depth()["active"]
        eng.run(ndte)
        for k, st in enumerate(states):
            want = st.copy()
            lib.run(p, want, ndte, halo=local[k])
            got = st.copy()
            for n in got.f:
                got.f[n][:] = eng.get_field_member(k, n)
            assert_state_bitexact(got, want)


@pytest.mark.parametrize("mode", ["fused", "fused-pair"])
def test_batch_hilbert_order_matches_independent_oracle_runs(gpu, mode):
    """A batch in the Hilbert tile order (members are separate components of
    one frame): every member bit-identical to its own oracle run."""
    sizes = [(30, 22, 91), (64, 40, 92), (12, 9, 93), (50, 70, 94)]
    states = [synthetic_state(nx, ny, seed, force_perturb=0.02) for nx, ny, seed in sizes]
    ndte = 9
    lib, p = oracle_params(states[0], ndte)
    params = evp1d.set_evp_parameters(ndte)
    _log(f"batch hilbert mode={mode}")
    with evp1d.Engine.batch([(st.na, st.navel) for st in states], params) as eng:
        set_mode(eng, mode)
        eng.set_tile_order(1)
        for k, st in enumerate(states):
            eng.load_member(k, st.f, st.idx, st.skipT, st.skipU)
        eng.run(ndte)
        for k, st in enumerate(states):
            want = st.copy()
            lib.run(p, want, ndte)
            got = st.copy()
            for n in got.f:
                got.f[n][:] = eng.get_field_member(k, n)
            assert_state_bitexact(got, want)


def test_autotune_keeps_results_and_state(gpu):
    st = synthetic_state(120, 90, seed=91)
    lib, p = oracle_params(st, 12)
    want = st.copy()
    lib.run(p, want, 12)
    with make_engine(st, 12) as eng:
        eng.run(5)
        t = eng.autotune(reps=3)
        assert t["stress_block"] in (64, 128, 256, 512, 1024) and t["stress_us"] > 0
        eng.run(7)
        got = fetch_state(eng, st)
    assert_state_bitexact(got, want)
    for cfg in [(64, 1, 1024, 0), (512, 0, 128, 1), (1024, 1, 64, 1)]:
        with make_engine(st, 12) as eng:
            eng.set_tuning(*cfg)
            eng.run(12)
            assert_state_bitexact(fetch_state(eng, st), want)


@pytest.mark.parametrize("inplace", [False, True])
@pytest.mark.parametrize("tile", [(32, 16, 256, 1), (1, 1, 256, 0), (5, 7, 512, 1), (64, 8, 512, 0), (200, 2, 256, 1)])
@pytest.mark.parametrize("nsub", [1, 2, 7])
def test_fused_tilings_match_oracle(gpu, tile, nsub, inplace):
    """Any tile shape and XCD order, odd and even subcycle counts (the state
    ping-pongs between two buffers and must end in the handle's own), both
    stress storages (in place, the shared points' copy alternates too)."""
    st = synthetic_state(70, 50, seed=61)
    lib, p = oracle_params(st, nsub)
    want = st.copy()
    lib.run(p, want, nsub)
    _log(f"fused tile={tile} nsub={nsub} inplace={inplace}")
    with make_engine(st, nsub, "fused-inplace" if inplace else "fused") as eng:
        eng.set_tiling(*tile)
        assert eng.get_tiling()["ntiles"] > 0
        eng.run(nsub)
        assert_state_bitexact(fetch_state(eng, st), want)
        eng.run(nsub)                 # a second run from the end state
        lib.run(p, want, nsub)
        assert_state_bitexact(fetch_state(eng, st), want)


@pytest.mark.parametrize("mode", ["fused", "fused-inplace", "fused-pair", "graph"])
@pytest.mark.parametrize("nsub", [1, 2, 7])
def test_hilbert_tile_order_matches_oracle(gpu, mode, nsub):
    """Tiles visited along a Hilbert curve (evp1d_set_tile_order(h, 1)): a
    different slot order, the same results bit for bit; switching back and
    forth between runs keeps the state."""
    st = synthetic_state(90, 70, seed=63)
    lib, p = oracle_params(st, nsub)
    want = st.copy()
    lib.run(p, want, nsub)
    _log(f"hilbert mode={mode} nsub={nsub}")
    with make_engine(st, nsub, mode) as eng:
        eng.set_tile_order(1)
        assert eng.get_tile_order() == 1
        eng.run(nsub)
        assert_state_bitexact(fetch_state(eng, st), want)
        eng.set_tile_order(0)
        eng.run(nsub)
        lib.run(p, want, nsub)
        assert_state_bitexact(fetch_state(eng, st), want)
        eng.set_tile_order(1)
        eng.run(nsub)
        lib.run(p, want, nsub)
        assert_state_bitexact(fetch_state(eng, st), want)


def test_setting_errors_are_reported(gpu):
    """Out-of-range settings fail with EVP1D_EINVAL and change nothing."""
    st = synthetic_state(20, 16, seed=64)
    with make_engine(st, 3) as eng:
        for bad in (-1, 2):
            with pytest.raises(evp1d.Evp1dError):
                eng.set_tile_order(bad)
        assert eng.get_tile_order() == 0
        with pytest.raises(evp1d.Evp1dError):
            eng.set_adjoint_segment(-1)
        assert eng.get_adjoint_segment() == 0
        eng.set_adjoint_segment(7)
        assert eng.get_adjoint_segment() == 7
        for bad in (-2, 2):
            with pytest.raises(evp1d.Evp1dError):
                eng.set_cache_hints(bad)
        assert eng.get_cache_hints()["hints"] == -1


@pytest.mark.parametrize("mode", ["fused", "fused-inplace", "fused-pair"])
def test_fused_sees_fields_set_between_runs(gpu, mode):
    """str1..8 at points stress never writes (read by stepu from memory),
    stresses and ghost velocities may be changed by the caller between runs."""
    st = synthetic_state(50, 40, seed=62)
    lib, p = oracle_params(st, 12)
    want = st.copy()
    lib.run(p, want, 5)
    rng = np.random.default_rng(7)
    stale = np.concatenate([st.skipT != 0, np.ones(st.navel - st.na, bool)])
    for k in ("str2", "str3", "str7"):
        want.f[k][stale] = rng.standard_normal(stale.sum())
    want.f["uvel"][st.na:] = rng.standard_normal(st.navel - st.na) * 0.1
    want.f["stressp_1"] = want.f["stressp_1"] * 0.5
    with make_engine(st, 12, mode) as eng:
        eng.run(5)
        for k in ("str2", "str3", "str7", "uvel", "stressp_1"):
            eng.set_field(k, want.f[k])
        eng.run(7)
        lib.run(p, want, 7)
        assert_state_bitexact(fetch_state(eng, st), want)


def test_fused_stress_storage_switch_between_runs(gpu):
    """Changing the stress storage between runs keeps the state."""
    st = synthetic_state(60, 45, seed=65)
    lib, p = oracle_params(st, 15)
    want = st.copy()
    lib.run(p, want, 15)
    with make_engine(st, 15, "fused") as eng:
        eng.run(5)
        eng.set_fused_inplace(True)
        assert eng.get_fused_inplace()
        eng.run(7)
        eng.set_fused_inplace(False)
        eng.run(3)
        assert_state_bitexact(fetch_state(eng, st), want)


def test_autotune_fused_keeps_results_and_state(gpu):
    """The fused tuner times every candidate on the handle's own state and
    restores it; the run after tuning continues from where it was."""
    st = synthetic_state(120, 90, seed=92)
    lib, p = oracle_params(st, 12)
    want = st.copy()
    lib.run(p, want, 12)
    with make_engine(st, 12, "fused") as eng:
        eng.run(5)
        t = eng.autotune_fused(nsub=4)
        _log(f"autotune_fused: {t}")
        assert t["ntiles"] > 0 and t["us"] > 0 and (t["tile_w"], t["tile_h"], t["block"], t["depth"]) in \
            [(64, 16, 512, 1), (32, 16, 512, 1), (32, 8, 256, 1), (16, 8, 128, 1),
             (32, 16, 512, 2), (64, 8, 512, 2), (16, 16, 256, 2), (32, 8, 256, 2)]
        eng.run(7)
        assert_state_bitexact(fetch_state(eng, st), want)


@pytest.mark.parametrize("tile", [(32, 16), (16, 16), (64, 8), (5, 7), (1, 1), (200, 2)])
@pytest.mark.parametrize("nsub", [1, 2, 3, 8])
def test_pair_form_matches_oracle(gpu, tile, nsub):
    """Two subcycles per launch (each tile evaluates a two-cell border ring
    again), odd counts starting with one one-subcycle launch, any tile shape:
    every output bit as the oracle, over two runs."""
    st = synthetic_state(70, 50, seed=66)
    lib, p = oracle_params(st, nsub)
    want = st.copy()
    lib.run(p, want, nsub)
    _log(f"pair tile={tile} nsub={nsub}")
    with make_engine(st, nsub, "fused") as eng:
        eng.set_fused_depth(2)
        eng.set_tiling(tile[0], tile[1], 256, 1)
        d = eng.get_fused_depth()
        assert d["active"] and d["block"] >= 64, d
        eng.run(nsub)
        assert_state_bitexact(fetch_state(eng, st), want)
        eng.run(nsub)
        lib.run(p, want, nsub)
        assert_state_bitexact(fetch_state(eng, st), want)


def test_pair_form_switches_and_falls_back(gpu):
    """Depth changes between runs keep the state; a tile whose level-2 set
    exceeds 1024 points (64x16 on an open domain) and a ghost copy-list both
    leave runs on the one-subcycle form, with the same results."""
    st = synthetic_state(150, 60, seed=67)
    lib, p = oracle_params(st, 20)
    want = st.copy()
    lib.run(p, want, 20)
    with make_engine(st, 20, "fused-pair") as eng:
        assert eng.get_fused_depth()["active"]
        eng.run(5)
        eng.set_fused_depth(1)
        assert not eng.get_fused_depth()["active"]
        eng.run(4)
        eng.set_fused_depth(2)
        eng.set_tiling(64, 16, 512, 1)
        d = eng.get_fused_depth()
        _log(f"64x16 pair: {d}")
        eng.run(6)
        eng.set_tiling(16, 16, 256, 0)
        assert eng.get_fused_depth()["active"]
        eng.run(5)
        assert_state_bitexact(fetch_state(eng, st), want)


def test_fused_unavailable_without_a_grid(gpu):
    st = synthetic_state(30, 20, seed=63)
    bad = st.copy()
    bad.idx["nw"] = st.idx["nw"].copy()
    bad.idx["nw"][[3, 10]] = bad.idx["nw"][[10, 3]]
    with make_engine(bad, 3, "fused") as eng:
        assert eng.get_tiling()["ntiles"] == 0
        with pytest.raises(evp1d.Evp1dError):
            eng.run(3)


@pytest.mark.parametrize("mode", ["graph", "fused", "fused-pair"])
def test_parameters_change_between_runs(gpu, mode):
    """The captured graph holds the solver scalars: a changed parameter set
    must reach the next run, an unchanged one must not disturb it."""
    st = synthetic_state(40, 30, seed=64)
    lib = O.OracleLib()
    p1 = lib.params(10)
    p2 = lib.params(10, capping=0.5, e_factor=0.3)
    want = st.copy()
    lib.run(p1, want, 6)
    lib.run(p2, want, 6)
    lib.run(p2, want, 6)
    with make_engine(st, 10, mode) as eng:
        eng.run(6)
        q = evp1d.set_evp_parameters(10)
        q.capping, q.e_factor = 0.5, 0.3
        eng.set_params(q)
        eng.run(6)
        eng.set_params(q)          # same values again
        eng.run(6)
        assert_state_bitexact(fetch_state(eng, st), want)
