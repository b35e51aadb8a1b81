"""torch.autograd wrapper of the ARD solver and its discrete adjoint.

``recorder_traces(scene, waves, scales)`` returns the recorder traces
[steps][nrec] of ``scene`` (plain form by default, or ``form=A.FORM_BATCHED``;
one MI355X) for the source waveforms ``waves`` [nsrc][steps] and the six
per-face PML damping scales ``scales``; its backward runs ``ard_adjoint``
(checkpointed reverse sweep of the step on the device; for the batched form
on its plain-form twin, the same discrete step) instead of recording the
time loop. Gradients flow to both inputs; a Gaussian source's row of
``waves`` is ignored by the forward.
"""
import copy

import numpy as np
import torch

import ard_scenes as S
import ard_sim as A


def _scene_with(scene, waves):
    sc = copy.copy(scene)
    sc.sources = [S.Source(s.x, s.y, s.z, wave=None if s.wave is None else np.asarray(w, dtype=np.float32))
                  for s, w in zip(scene.sources, waves)]
    return sc


class RecorderTraces(torch.autograd.Function):
    @staticmethod
    def forward(ctx, waves, scales, scene, steps, checkpoint_stride, form):
        w = waves.detach().cpu().numpy().astype(np.float32)
        th = scales.detach().cpu().numpy().astype(np.float32)
        sim = A.Simulation(_scene_with(scene, w), form=form, pml_scale=th)
        sim.update(steps)
        tr = sim.recorders(0, steps)
        ctx.sim, ctx.stride = sim, checkpoint_stride
        ctx.wdtype, ctx.wdev, ctx.sdtype, ctx.sdev = waves.dtype, waves.device, scales.dtype, scales.device
        return torch.as_tensor(tr).to(device=waves.device)

    @staticmethod
    def backward(ctx, g):
        dw, ds = ctx.sim.adjoint(g.detach().cpu().numpy(), ctx.stride)
        ctx.sim.close()
        return (torch.as_tensor(dw).to(dtype=ctx.wdtype, device=ctx.wdev),
                torch.as_tensor(ds).to(dtype=ctx.sdtype, device=ctx.sdev), None, None, None, None)


def recorder_traces(scene, waves, scales, steps=None, checkpoint_stride=0, form=A.FORM_PLAIN):
    """Differentiable recorder traces; see the module docstring."""
    return RecorderTraces.apply(waves, scales, scene, steps or scene.steps, checkpoint_stride, form)
