"""GPU tests of the torch.autograd wrapper of the gradient kernels
(tests/autograd_harness.py): torch.autograd.gradcheck (central differences
of the engine's own forward runs against the adjoint, every input element)
on a small problem, and backward() returning evp1d_adjoint_run's gradients
exactly."""
import numpy as np
import pytest
import torch

from autograd_harness import EvpMisfit, misfit
from gpu_common import make_engine, synthetic_state

pytestmark = pytest.mark.gpu
NDTE = 12


@pytest.fixture(scope="module")
def small(gpu):
    st = synthetic_state(10, 8, seed=111)
    tgt = st.copy()
    tgt.f["strength"] *= 1.2
    with make_engine(tgt, NDTE) as eng:
        eng.run(NDTE)
        ut = eng.get_field("uvel")[:st.na]
        vt = eng.get_field("vvel")[:st.na]
    return st, ut, vt


def _inputs(st, requires_grad=True):
    return [torch.tensor(st.f[n].copy(), dtype=torch.float64, requires_grad=requires_grad)
            for n in ("strength", "cdn_ocn", "forcexU", "forceyU")]


def test_backward_returns_the_adjoint_gradients(small):
    st, ut, vt = small
    x = _inputs(st)
    loss = EvpMisfit.apply(*x, st, NDTE, ut, vt, 0)
    loss.backward()
    with make_engine(st, NDTE) as eng:
        g = eng.adjoint_run(NDTE, ut, vt, nsnap=NDTE)
    assert float(loss.detach()) == g["loss"]
    for t, k in zip(x, ("strength", "cdn_ocn", "forcex", "forcey")):
        assert np.array_equal(t.grad.numpy(), g[k]), k


def test_gradcheck(small):
    """Central differences on every element of all four inputs (torch's
    default eps 1e-6, atol 1e-5, rtol 1e-3), with the inputs scaled to O(1)
    so that eps means the same thing for each."""
    st, ut, vt = small
    base = [st.f[n].copy() for n in ("strength", "cdn_ocn", "forcexU", "forceyU")]
    scale = [np.where(np.abs(b) > 0, np.abs(b), 1.0) for b in base]
    f = misfit(st, NDTE, ut, vt)

    def scaled(*z):   # z = inputs / scale, so every component is O(1)
        return f(*[zi * torch.from_numpy(si) for zi, si in zip(z, scale)])
    z = [torch.tensor(b / s, dtype=torch.float64, requires_grad=True) for b, s in zip(base, scale)]
    print(f"gradcheck over {sum(t.numel() for t in z)} input elements", flush=True)
    assert torch.autograd.gradcheck(scaled, z, eps=1e-6, atol=1e-5, rtol=1e-3)
