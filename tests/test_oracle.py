"""CPU tests that pin the oracle (the checker of every GPU test).

1. Golden vectors: the C restatement reproduces, byte for byte, both output
   files of the flang-built reference v2b (tests/golden/*.npz), and the
   400x300 workload's md5 sums (tests/golden/golden.json).
2. Closed forms that hold whatever the code: the zero-stress, zero-velocity
   state (stresses stay 0, stepu is free drift), uniform translation on a
   uniform grid (strain rates vanish, stresses decay by denom1), and the
   parameter derivation of set_evp_parameters.
3. Planted mistakes: a perturbed oracle (one sign, one neighbour index, one
   dropped term) must fail the golden comparison, so the golden vectors do
   discriminate the mistakes a restatement could make.
"""
import ctypes
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from gpu_common import GOLD, golden_cases, load_golden


@pytest.fixture(scope="module")
def lib():
    return O.OracleLib()


@pytest.mark.parametrize("case", golden_cases())
def test_oracle_matches_reference_golden(lib, case):
    st, ndte, want_stress, want_stepu = load_golden(case)
    p = lib.params(ndte, **st.scalars)
    lib.run(p, st, ndte)
    a, b = O.output_bytes(st)
    assert a == want_stress
    assert b == want_stepu


def test_oracle_openmp_build_identical(lib):
    st, ndte, want_stress, want_stepu = load_golden("small_64x48")
    omp = O.OracleLib(omp=True)
    omp.run(omp.params(ndte, **st.scalars), st, ndte)
    assert O.output_bytes(st) == (want_stress, want_stepu)


def test_oracle_evp_c0_md5(lib):
    from evp1d import synth
    rec = json.loads((GOLD / "golden.json").read_text())["evp_c0"]
    w = synth.WORKLOADS["evp_c0"]
    g = synth.make_grid(w["nx"], w["ny"], w["seed"])
    ix = O.index_reference(g.tmask, g.nghost)
    st = O.pack_dyn(g, ix)
    assert (st.na, st.navel) == (rec["na"], rec["navel"])
    lib.run(lib.params(w["ndte"], **st.scalars), st, w["ndte"])
    a, b = O.output_bytes(st)
    assert O.md5(a) == rec["output_stress_md5"]
    assert O.md5(b) == rec["output_stepu_md5"]


def test_set_evp_parameters_closed_form(lib):
    """ice.F90:109-130 for ndte=120: arlx = 0.72*120, arlx1i = 1/arlx, ..."""
    p = lib.params(120)
    assert p.epp2i == 0.25 and p.e_factor == 0.25 and p.capping == 1.0
    assert p.arlx == 2.0 * 0.36 * 120 and p.arlx1i == 1.0 / p.arlx
    assert p.brlx == 120.0 and p.denom1 == 1.0 / (1.0 + p.arlx1i)
    assert p.deltaminEVP == 1e-8 and p.u0 == 5e-5 and p.rhow == 1026.0
    assert p.revp == 0.0 and p.Ktens == 0.0 and p.cosw == 1.0 and p.sinw == 0.0


def _square_state(n=6, uniform_vel=None, seed=0):
    """All-ocean n x n block with a dry rim, uniform 10 km grid."""
    from evp1d import synth
    g = synth.make_grid(n, n, seed, land_frac=0.0, ice_prob=1.0, sigma=0.0)
    for k in ("dxT", "dyT", "HTE", "HTN"):
        g.fields[k][:] = 1.0e4
    g.fields["uarear"][:] = 1.0e-8
    ix = O.index_reference(g.tmask, g.nghost)
    st = O.pack_dyn(g, ix)
    if uniform_vel is not None:
        st.f["uvel"][:] = uniform_vel[0]
        st.f["vvel"][:] = uniform_vel[1]
    return st


def test_zero_state_is_free_drift(lib):
    """sigma = 0, u = v = 0 everywhere: strain rates vanish, Delta = 0, so
    tmpcalc = strength/DminTarea, rep_prs = 0 and the stresses stay exactly 0;
    stepu reduces to u = (cca*cc1 + ccb*cc2)/(cca^2 + ccb^2) with
    cca = brlx*umassdti + vrel, ccb = fm, cc1 = forcex + vrel*uocn,
    cc2 = forcey + vrel*vocn, vrel = aiu*rhow*Cw*|u_ocn|."""
    st = _square_state(8)
    st.f["uvel"][:] = 0.0
    st.f["vvel"][:] = 0.0
    st.f["uvel_init"][:] = 0.0
    st.f["vvel_init"][:] = 0.0
    p = lib.params(10)
    s0 = st.copy()
    lib.run(p, st, 1)
    for n in O.STRESS_FIELDS:
        assert not np.any(st.f[n])
    act = s0.skipU == 0
    f = s0.f
    vrel = f["aiu"] * 1026.0 * f["cdn_ocn"] * np.sqrt(f["uocn"] ** 2 + f["vocn"] ** 2)
    cca = 10.0 * f["umassdti"] + vrel
    ccb = f["fmU"]
    cc1 = f["forcexU"] + vrel * f["waterxU"]
    cc2 = f["forceyU"] + vrel * f["wateryU"]
    u = (cca * cc1 + ccb * cc2) / (cca ** 2 + ccb ** 2)
    v = (cca * cc2 - ccb * cc1) / (cca ** 2 + ccb ** 2)
    na = st.na
    np.testing.assert_allclose(st.f["uvel"][:na][act], u[act], rtol=1e-13, atol=1e-18)
    np.testing.assert_allclose(st.f["vvel"][:na][act], v[act], rtol=1e-13, atol=1e-18)
    assert np.all(st.f["strintxU"][act] == 0.0)


def test_uniform_translation_decays_stress(lib):
    """Uniform u, v on a uniform grid (ghosts included): every strain rate is
    exactly 0, so one stress call maps sigma -> sigma*denom1 (rep_prs*0 = 0)."""
    st = _square_state(8, uniform_vel=(0.3, -0.2))
    rng = np.random.default_rng(3)
    for n in O.STRESS_FIELDS:
        st.f[n][:] = rng.standard_normal(st.na) * 100.0
    p = lib.params(50)
    s0 = st.copy()
    lib.stress(p, st)
    act = s0.skipT == 0
    for n in O.STRESS_FIELDS:
        assert np.array_equal(st.f[n][act], s0.f[n][act] * p.denom1), n


# ---- planted mistakes ------------------------------------------------------
def _perturbed_oracle(tmp_path, sed_expr):
    """Compile a copy of the oracle with one textual mutation."""
    import subprocess
    src = (Path(O.HERE) / "evp1d_oracle.c").read_text()
    old, new = sed_expr
    assert old in src, old
    mut = src.replace(old, new, 1)
    c = tmp_path / "mut.c"
    c.write_text(mut)
    so = tmp_path / "libmut.so"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-o", str(so), str(c), "-lm"],
                   check=True)
    m = O.OracleLib.__new__(O.OracleLib)
    m.lib = ctypes.CDLL(str(so))
    m.lib.oracle_set_evp_parameters.argtypes = [ctypes.c_int32, ctypes.POINTER(O.Params)]
    m.lib.oracle_run.argtypes = O.OracleLib().lib.oracle_run.argtypes
    return m


MUTATIONS = {
    "sign_in_divune": ("tmp_cyp * tmp_uvel_cc - tmp_dyT * tmp_uvel_ee", "tmp_cyp * tmp_uvel_cc + tmp_dyT * tmp_uvel_ee"),
    "ne_for_se_gather": ("tmp_vvel_se = vvel[se[w] - 1];", "tmp_vvel_se = vvel[ne[w] - 1];"),
    "dropped_dyhx_term": ("+ tmp_dyhx * csig12ne;\n    str2", ";\n    str2"),
    "swapped_nw_sw": ("tmp_str2_nw = str2[nw[w] - 1];", "tmp_str2_nw = str2[sw[w] - 1];"),
    "reassociated_sum": ("(str1[w] + tmp_str2_nw + tmp_str3_sse + tmp_str4_sw)",
                         "(str1[w] + (tmp_str2_nw + tmp_str3_sse) + tmp_str4_sw)"),
}


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_golden_vectors_catch_planted_mistake(tmp_path, name):
    mut = _perturbed_oracle(tmp_path, MUTATIONS[name])
    caught = False
    for case in ("small_64x48", "long_16x12"):
        st, ndte, want_stress, want_stepu = load_golden(case)
        mut.run(mut.params(ndte, **st.scalars), st, ndte)
        if O.output_bytes(st) != (want_stress, want_stepu):
            caught = True
    assert caught, f"mutation {name} not detected by the golden vectors"


def test_input_files_round_trip_and_md5(tmp_path):
    """write_inputs writes the reference's three input files and kernel.nml;
    input_md5s (used by the full-size GPU tests to check regenerated inputs)
    gives the md5 of the same bytes, and read_inputs reads them back."""
    st, ndte, _, _ = load_golden("ragged_40x7")
    O.write_inputs(tmp_path, st, ndte)
    for f, m in O.input_md5s(st).items():
        assert O.md5((tmp_path / f).read_bytes()) == m, f
    nml = (tmp_path / "kernel.nml").read_text()
    assert f"navel = {st.navel}" in nml and f"na = {st.na}" in nml and f"ndte = {ndte}" in nml
    back = O.read_inputs(tmp_path, st.na, st.navel)
    for k in ("strength", "uvel", "vvel", "HTNm1", "stress12_4", "Tbu"):
        assert np.array_equal(back.f[k], st.f[k]), k
    for k in O.NBR:
        assert np.array_equal(back.idx[k], st.idx[k]), k
    assert np.array_equal(back.skipT, st.skipT) and np.array_equal(back.skipU, st.skipU)
