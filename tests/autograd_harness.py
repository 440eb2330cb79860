"""torch.autograd wrapper of the hand-written gradient kernels (verification
harness only; the library itself exposes evp1d_adjoint_run through its C ABI).

EvpMisfit.apply(strength, cdn_ocn, forcex, forcey) returns the loss
L = 1/2 sum_{skipU=0} (u_N - u*)^2 + (v_N - v*)^2 after ndte subcycles run by
the GPU engine from a fixed starting state, as a 0-d float64 tensor; its
backward returns the gradients evp1d_adjoint_run computed in the same call.
Tensors are float64 on the CPU: they cross the ctypes boundary as host
arrays, and the engine copies them to the device.
"""
from __future__ import annotations

import numpy as np
import torch

from gpu_common import make_engine

_INPUTS = ("strength", "cdn_ocn", "forcexU", "forceyU")
_GRADS = ("strength", "cdn_ocn", "forcex", "forcey")


class EvpMisfit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, strength, cdn_ocn, forcex, forcey, state, ndte, u_target, v_target, nsnap=0):
        st = state.copy()
        for name, t in zip(_INPUTS, (strength, cdn_ocn, forcex, forcey)):
            st.f[name][:] = t.detach().cpu().numpy()
        with make_engine(st, ndte) as eng:
            g = eng.adjoint_run(ndte, u_target, v_target, nsnap=nsnap)
        ctx.save_for_backward(*[torch.from_numpy(np.ascontiguousarray(g[k])) for k in _GRADS])
        return torch.tensor(g["loss"], dtype=torch.float64)

    @staticmethod
    def backward(ctx, gout):
        grads = [gout * g for g in ctx.saved_tensors]
        return (*grads, None, None, None, None, None)


def misfit(state, ndte, u_target, v_target, nsnap=0):
    """A function of the four input tensors (for torch.autograd.gradcheck)."""
    return lambda s, c, fx, fy: EvpMisfit.apply(s, c, fx, fy, state, ndte, u_target, v_target, nsnap)
