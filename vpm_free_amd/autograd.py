"""torch.autograd bindings of the library's adjoint entry points.

The gradients come from the hand-written transposed kernels behind vpmx_adjoint_probe_energy and
vpmx_adjoint_rollout (the library runs the forward and the adjoint on its own stream); torch
only carries the host tensors, so the tensors here are CPU float64 tensors.  No counterpart in the reference,
which has no derivative of vpm() (vpm.f90:54-475).
"""
import numpy as np
import torch


def _np(t):
    return t.detach().to(dtype=torch.float64, device="cpu").contiguous().numpy()


class ProbeEnergy(torch.autograd.Function):
    """J = 1/2 sum_k |u(p_k)|^2 for the velocity grid of vpm(WhatToDo=2) of the particles
    (XP: n x 3, QP: n x (neqpm+1)) at the probe points (m x 3); differentiable in XP and
    QP(:, 0:3) (the volume column gets no gradient).  `vpm` is a VPM context whose geometry
    was set by vpm(WhatToDo=0)."""

    @staticmethod
    def forward(ctx, xp, qp, probes, vpm):
        J, dq, dx = vpm.adjoint_probe_energy(_np(xp), _np(qp), _np(probes))
        gq = np.zeros(tuple(qp.shape))
        gq[:, :3] = dq
        ctx.save_for_backward(torch.from_numpy(dx), torch.from_numpy(gq))
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        dx, gq = ctx.saved_tensors
        return g * dx, g * gq, None, None


def probe_energy(vpm, xp, qp, probes):
    return ProbeEnergy.apply(xp, qp, probes, vpm)


class RolloutEnergy(torch.autograd.Function):
    """The probe energy J after `nsteps` RK2 steps (each with its remesh) of the particles;
    differentiable in the starting XP and QP(:, 0:3).  The remesh's threshold mask is held
    fixed, and the backward pass recomputes `ckpt` steps at a time from checkpoints."""

    @staticmethod
    def forward(ctx, xp, qp, probes, vpm, nsteps, dt, uinf, ckpt):
        J, dq, dx = vpm.adjoint_rollout(_np(xp), _np(qp), _np(probes), nsteps, dt, uinf, ckpt)
        gq = np.zeros(tuple(qp.shape))
        gq[:, :3] = dq
        ctx.save_for_backward(torch.from_numpy(dx), torch.from_numpy(gq))
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        dx, gq = ctx.saved_tensors
        return g * dx, g * gq, None, None, None, None, None, None


def rollout_energy(vpm, xp, qp, probes, nsteps, dt, uinf=(0.0, 0.0, 0.0), ckpt=2):
    return RolloutEnergy.apply(xp, qp, probes, vpm, int(nsteps), float(dt), tuple(uinf), int(ckpt))
