"""GPU tests of the gradient path (hand-written adjoint kernels, binomial
checkpointing, forward-mode tangent-linear kernels).

Loss: L = 1/2 sum_{skipU = 0} (u_N - u*)^2 + (v_N - v*)^2 after ndte
subcycles; target (u*, v*) = final state of a run with strength x 1.2.
Checks:
  * the tangent run's values are the exact forward (bit-identical to the
    oracle) and the adjoint run leaves the handle in that same state;
  * dot-product test: the tangent derivative along d equals <grad, d>
    to 1e-10 relative;
  * central finite differences (eps = 1e-6 relative per component) agree
    with <grad, d> to 1e-6 relative;
  * the gradient does not depend on the number of stored states, from one
    stored state (recompute everything) to all of them (no recomputation),
    nor on the checkpointing unit (single subcycles, or segments reversed
    through a stored velocity trajectory), in the graph and pair forms (the
    latter also with the cache hints forced on);
  * the pair-form forward sweep (two stored subcycles per launch, odd and
    tiny nsub) gives the same gradient as one subcycle per launch.
"""
import numpy as np
import pytest

import evp1d
from gpu_common import assert_state_bitexact, fetch_state, make_engine, oracle_params, synthetic_state

pytestmark = pytest.mark.gpu
NDTE = 40


@pytest.fixture(scope="module")
def problem(gpu):
    st = synthetic_state(50, 40, seed=101)
    tgt = st.copy()
    tgt.f["strength"] *= 1.2
    with make_engine(tgt, NDTE) as eng:
        eng.run(NDTE)
        ut = eng.get_field("uvel")[:st.na]
        vt = eng.get_field("vvel")[:st.na]
    return st, ut, vt


def _loss(st, ut, vt, pert=None, h=0.0):
    s = st.copy()
    if pert:
        for k, d in pert.items():
            s.f[k] = s.f[k] + h * d
    with make_engine(s, NDTE) as eng:
        eng.run(NDTE)
        return eng.loss(ut, vt)


def _direction(st, rng):
    return {
        "strength": rng.standard_normal(st.na) * st.f["strength"],
        "cdn_ocn": np.full(st.na, rng.standard_normal()) * st.f["cdn_ocn"],   # one scalar drag coefficient
        "forcexU": rng.standard_normal(st.na) * np.abs(st.f["forcexU"]),
        "forceyU": rng.standard_normal(st.na) * np.abs(st.f["forceyU"]),
    }


def _dot(g, d):
    return (np.dot(g["strength"], d["strength"]) + np.dot(g["cdn_ocn"], d["cdn_ocn"]) +
            np.dot(g["forcex"], d["forcexU"]) + np.dot(g["forcey"], d["forceyU"]))


def test_tangent_and_adjoint_leave_exact_forward_state(problem):
    st, ut, vt = problem
    lib, p = oracle_params(st, NDTE)
    want = st.copy()
    lib.run(p, want, NDTE)
    d = _direction(st, np.random.default_rng(1))
    with make_engine(st, NDTE) as eng:
        eng.tangent_run(NDTE, ut, vt, d["strength"], d["cdn_ocn"], d["forcexU"], d["forceyU"])
        assert_state_bitexact(fetch_state(eng, st), want)
    for nsnap in (6, NDTE):     # checkpointed; every state stored
        with make_engine(st, NDTE) as eng:
            g = eng.adjoint_run(NDTE, ut, vt, nsnap=nsnap)
            assert_state_bitexact(fetch_state(eng, st), want)
        with make_engine(want, NDTE) as eng:
            assert abs(eng.loss(ut, vt) - g["loss"]) <= 1e-15 * max(1.0, g["loss"])


def test_dot_product_tangent_vs_adjoint(problem):
    st, ut, vt = problem
    with make_engine(st, NDTE) as eng:
        g = eng.adjoint_run(NDTE, ut, vt, nsnap=NDTE)
    assert g["loss"] > 0
    rng = np.random.default_rng(2)
    worst = 0.0
    for _ in range(6):
        d = _direction(st, rng)
        with make_engine(st, NDTE) as eng:
            _, dl = eng.tangent_run(NDTE, ut, vt, d["strength"], d["cdn_ocn"], d["forcexU"], d["forceyU"])
        ad = _dot(g, d)
        worst = max(worst, abs(dl - ad) / max(abs(dl), abs(ad), 1e-300))
    print(f"dot-product relative mismatch: {worst:.3e}")
    assert worst <= 1e-10


def test_central_finite_differences(problem):
    st, ut, vt = problem
    with make_engine(st, NDTE) as eng:
        g = eng.adjoint_run(NDTE, ut, vt, nsnap=8)
    rng = np.random.default_rng(3)
    h = 1e-6
    worst = 0.0
    for _ in range(8):
        d = _direction(st, rng)
        fd = (_loss(st, ut, vt, d, h) - _loss(st, ut, vt, d, -h)) / (2 * h)
        ad = _dot(g, d)
        worst = max(worst, abs(fd - ad) / max(abs(fd), abs(ad)))
    print(f"FD relative mismatch: {worst:.3e}")
    assert worst <= 1e-6


@pytest.mark.parametrize("mode", ["graph", "fused-pair", "pair-hints"])
def test_gradient_independent_of_snapshot_count(problem, mode):
    """From one stored state to all of them, and for every checkpointing
    unit: one subcycle (binomial checkpointing with recomputed str1..8),
    automatic (one unit per stored state, each reversed through its stored
    velocity trajectory), and 5 or 3 subcycles (binomial checkpointing over
    more units than stored states).  Same gradient and loss bit for bit; the
    handle ends in the exact forward state."""
    st, ut, vt = problem
    lib, p = oracle_params(st, NDTE)
    want = st.copy()
    lib.run(p, want, NDTE)
    out = []
    # (2, 2) and (3, 4): units of even length with few stored states, so the
    # pair form's sweeps start from stored stresses also when the schedule
    # recomputes from one stored state
    cases = [(1, 1), (3, 1), (7, 1), (1, 0), (3, 0), (7, 0), (3, 5), (2, 3), (2, 2), (3, 4), (NDTE + 5, 0), (0, 0)]
    for nsnap, seg in cases:     # nsnap 0: chosen by the library (all, here)
        with make_engine(st, NDTE, mode) as eng:
            eng.set_adjoint_segment(seg)
            assert eng.get_adjoint_segment() == seg
            out.append(eng.adjoint_run(NDTE, ut, vt, nsnap=nsnap))
            assert_state_bitexact(fetch_state(eng, st), want)
    for (nsnap, seg), g in zip(cases[1:], out[1:]):
        for k in ("strength", "cdn_ocn", "forcex", "forcey"):
            assert np.array_equal(g[k], out[0][k]), (nsnap, seg, k)
        assert g["loss"] == out[0]["loss"], (nsnap, seg)


def test_gradient_same_for_both_stress_storages(problem):
    """The stored trajectory's fused forward sweep with stresses in place
    gives the same gradient, loss and end state as with two buffers."""
    st, ut, vt = problem
    lib, p = oracle_params(st, NDTE)
    want = st.copy()
    lib.run(p, want, NDTE)
    out = []
    for mode in ("fused", "fused-inplace"):
        for nsnap in (NDTE, 4):
            with make_engine(st, NDTE, mode) as eng:
                out.append(eng.adjoint_run(NDTE, ut, vt, nsnap=nsnap))
                assert_state_bitexact(fetch_state(eng, st), want)
    for g in out[1:]:
        for k in ("strength", "cdn_ocn", "forcex", "forcey"):
            assert np.array_equal(g[k], out[0][k]), k
        assert g["loss"] == out[0]["loss"]


def test_gradient_is_zero_off_ice(problem):
    """No stress where skipT, no stepu where skipU: those inputs are unused."""
    st, ut, vt = problem
    with make_engine(st, NDTE) as eng:
        g = eng.adjoint_run(NDTE, ut, vt, nsnap=5)
    assert np.all(g["strength"][st.skipT != 0] == 0.0)
    for k in ("cdn_ocn", "forcex", "forcey"):
        assert np.all(g[k][st.skipU != 0] == 0.0)
    assert np.any(g["strength"] != 0.0) and np.any(g["forcex"] != 0.0)


def test_repeated_gradient_calls_on_one_handle(problem):
    """The stored trajectory's memory is kept and reused by later calls and
    given back by a checkpointed one; every call starts from the state it finds."""
    st, ut, vt = problem
    with make_engine(st, NDTE) as eng:
        ref = eng.adjoint_run(NDTE, ut, vt, nsnap=NDTE)
    with make_engine(st, NDTE) as eng:
        for nsnap in (NDTE, NDTE, 4, NDTE):
            eng.load_state(st.f, st.idx, st.skipT, st.skipU)
            g = eng.adjoint_run(NDTE, ut, vt, nsnap=nsnap)
            for k in ("strength", "cdn_ocn", "forcex", "forcey"):
                assert np.array_equal(g[k], ref[k]), (nsnap, k)
            assert g["loss"] == ref["loss"]


@pytest.mark.parametrize("mode", ["auto", "fused-pair"])
def test_gradient_storage_reused_across_calls(problem, mode):
    """One handle through checkpointed calls of several stored-state counts
    and units with whole-trajectory calls between them: the stored states and
    the trajectory that the handle keeps from earlier calls (never cleared)
    give every call a fresh handle's gradient bit for bit.  Then a changed
    field: the next forward run and gradient use it (the forward sweeps of a
    gradient reuse the fields it packed; later runs pack them again)."""
    st, ut, vt = problem
    keys = ("strength", "cdn_ocn", "forcex", "forcey")
    with make_engine(st, NDTE, mode) as eng:
        ref = eng.adjoint_run(NDTE, ut, vt, nsnap=NDTE)
    st2 = st.copy()
    st2.f["strength"] = st2.f["strength"] * 1.1
    with make_engine(st2, NDTE, mode) as eng:
        ref2 = eng.adjoint_run(NDTE, ut, vt, nsnap=NDTE)
    lib, p = oracle_params(st2, NDTE)
    want2 = st2.copy()
    lib.run(p, want2, NDTE)
    with make_engine(st, NDTE, mode) as eng:
        for nsnap, seg in [(4, 0), (NDTE, 0), (3, 0), (6, 1), (2, 3), (NDTE, 0), (5, 0), (8, 0)]:
            eng.load_state(st.f, st.idx, st.skipT, st.skipU)
            eng.set_adjoint_segment(seg)
            g = eng.adjoint_run(NDTE, ut, vt, nsnap=nsnap)
            for k in keys:
                assert np.array_equal(g[k], ref[k]), (nsnap, seg, k)
            assert g["loss"] == ref["loss"], (nsnap, seg)
        eng.set_adjoint_segment(0)
        eng.load_state(st2.f, st2.idx, st2.skipT, st2.skipU)
        eng.run(NDTE)
        assert_state_bitexact(fetch_state(eng, st2), want2)
        for nsnap in (5, NDTE):
            eng.load_state(st2.f, st2.idx, st2.skipT, st2.skipU)
            g = eng.adjoint_run(NDTE, ut, vt, nsnap=nsnap)
            for k in keys:
                assert np.array_equal(g[k], ref2[k]), (nsnap, k)
            assert_state_bitexact(fetch_state(eng, st2), want2)


def test_gradient_storage_released_and_regained(problem):
    """The handle reports the memory gradient calls keep with it; releasing
    it leaves nothing held (twice is harmless), and the next call allocates
    the same amount again and gives the same gradient bit for bit."""
    st, ut, vt = problem
    keys = ("strength", "cdn_ocn", "forcex", "forcey")
    with make_engine(st, NDTE) as eng:
        assert eng.gradient_storage() == 0
        out = []
        for nsnap in (5, NDTE):
            eng.load_state(st.f, st.idx, st.skipT, st.skipU)
            g1 = eng.adjoint_run(NDTE, ut, vt, nsnap=nsnap)
            held = eng.gradient_storage()
            assert held > 0, nsnap
            eng.release_gradient_storage()
            assert eng.gradient_storage() == 0
            eng.release_gradient_storage()
            eng.load_state(st.f, st.idx, st.skipT, st.skipU)
            g2 = eng.adjoint_run(NDTE, ut, vt, nsnap=nsnap)
            assert eng.gradient_storage() == held, nsnap
            for k in keys:
                assert np.array_equal(g1[k], g2[k]), (nsnap, k)
            out.append(g2)
        for k in keys:
            assert np.array_equal(out[0][k], out[1][k]), k


@pytest.mark.parametrize("nsnap", [NDTE, 5])
def test_batch_gradient_equals_member_gradients(gpu, nsnap):
    """A batched handle's gradient: each member's part is bit-identical to the
    member's own single-domain gradient (members are independent), and the
    loss is the sum of the members' losses."""
    sizes = [(30, 22, 121), (16, 12, 122), (40, 18, 123)]
    states = [synthetic_state(nx, ny, seed) for nx, ny, seed in sizes]
    targets, singles = [], []
    for st in states:
        tgt = st.copy()
        tgt.f["strength"] *= 1.2
        with make_engine(tgt, NDTE) as e:
            e.run(NDTE)
            targets.append((e.get_field("uvel")[:st.na], e.get_field("vvel")[:st.na]))
        with make_engine(st, NDTE) as e:
            singles.append(e.adjoint_run(NDTE, *targets[-1], nsnap=nsnap))
    params = evp1d.set_evp_parameters(NDTE, **states[0].scalars)
    with evp1d.Engine.batch([(st.na, st.navel) for st in states], params) as eng:
        for k, st in enumerate(states):
            eng.load_member(k, st.f, st.idx, st.skipT, st.skipU)
        ut = np.concatenate([t[0] for t in targets])
        vt = np.concatenate([t[1] for t in targets])
        g = eng.adjoint_run(NDTE, ut, vt, nsnap=nsnap)
    off = 0
    for k, st in enumerate(states):
        for key in ("strength", "cdn_ocn", "forcex", "forcey"):
            assert np.array_equal(g[key][off:off + st.na], singles[k][key]), (k, key)
        off += st.na
    want = sum(s["loss"] for s in singles)
    assert abs(g["loss"] - want) <= 1e-12 * want


@pytest.mark.parametrize("nsub", [NDTE, NDTE - 1, 2, 1])
def test_gradient_unchanged_by_the_pair_form(problem, nsub):
    """With two subcycles per fused launch selected, the gradient's forward
    sweep runs the pair form and stores both subcycles of each launch (after
    one single subcycle when nsub is odd).  The gradient and the end state
    are the same bit for bit as with one subcycle per launch, also after a
    pair-form run on the handle, and with the whole trajectory or 3 stored
    states."""
    st, ut, vt = problem
    lib, p = oracle_params(st, NDTE)
    want = st.copy()
    lib.run(p, want, nsub)
    with make_engine(st, NDTE, "fused") as eng:
        eng.set_tiling(32, 16, 512, 1)   # fused-pair's tiling, so the loss sums in the same slot order
        ref = eng.adjoint_run(nsub, ut, vt)
    for nsnap in (0, 3):
        with make_engine(st, NDTE, "fused-pair") as eng:
            assert eng.get_fused_depth()["active"]
            eng.run(3)
            for k, v in st.f.items():       # back to the start state
                eng.set_field(k, v)
            g = eng.adjoint_run(nsub, ut, vt, nsnap=nsnap)
            assert_state_bitexact(fetch_state(eng, st), want)
        for k in ("strength", "cdn_ocn", "forcex", "forcey"):
            assert np.array_equal(g[k], ref[k]), (nsnap, k)
        assert g["loss"] == ref["loss"]
