"""evp1d_loss_noalloc (the loss without allocations, fixed device scratch)
against evp1d_loss through the C ABI: two calls in a row, the second on a
smaller domain, each within 1e-12 relative of evp1d_loss (the two sum the
per-point terms in different orders); sizes beyond the scratch and an
out-of-date layout are rejected."""
import sys

import numpy as np
import pytest

import evp1d
from gpu_common import synthetic_state

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _engine(nx, ny, seed, ndte):
    st = synthetic_state(nx, ny, seed)
    eng = evp1d.Engine(st.na, st.navel, evp1d.set_evp_parameters(ndte))
    eng.load_state(st.f, st.idx, st.skipT, st.skipU)
    return st, eng


def test_loss_noalloc_matches_loss(gpu):
    rng = np.random.default_rng(77)
    for nx, ny, seed in ((160, 120, 501), (48, 36, 502)):    # the second call is smaller
        _log(f"loss_noalloc {nx}x{ny}")
        st, eng = _engine(nx, ny, seed, 12)
        with eng:
            eng.run(12)
            ut = rng.standard_normal(st.na) * 0.1
            vt = rng.standard_normal(st.na) * 0.1
            want = eng.loss(ut, vt)
            got = eng.loss_noalloc(ut, vt)
            assert want > 0
            assert abs(got - want) <= 1e-12 * abs(want), (got, want)
            # repeated call: same bits (fixed reduction order, scratch reused)
            assert eng.loss_noalloc(ut, vt) == got


def test_loss_noalloc_rejects_what_it_cannot_do_without_allocating(gpu):
    st, eng = _engine(40, 30, 503, 4)
    with eng:
        eng.run(4)
        flags = st.skipU.copy()
        flags[: st.na // 2] = 1
        eng.set_skip("skipU", flags)           # the layout is now out of date
        with pytest.raises(evp1d.Evp1dError) as e:
            eng.loss_noalloc(np.zeros(st.na), np.zeros(st.na))
        assert e.value.code == 5
        eng.run(1)
        assert np.isfinite(eng.loss_noalloc(np.zeros(st.na), np.zeros(st.na)))
    big = (1 << 22) + 1
    with evp1d.Engine(big, big, evp1d.set_evp_parameters(4)) as eng:
        with pytest.raises(evp1d.Evp1dError) as e:
            eng.loss_noalloc(np.zeros(big), np.zeros(big))
        assert e.value.code == 1 and "4194304" in str(e.value)
