"""Differentiable rollouts: a torch.autograd.Function over the hand-written adjoint of the
HIP library (vxg_adjoint_run / vxg_adjoint_backward in include/vxg.h).

    pos = rollout(sim, log_modulus, phase, nsteps)      # (V, 3) float64, final positions
    loss = -(robot_com(sim, pos)[:, 0]).sum()
    loss.backward()                                     # log_modulus.grad, phase.grad

The forward is the mixed-precision step (VXG_MIXED) with per-voxel Young's modulus
E_v = exp(log_modulus_v) and per-voxel actuation phase; the backward recomputes the
rollout segment by segment from checkpoints and runs the reverse kernels.  PyTorch is
the autograd shell only: every step and every derivative runs in the library.
"""
import numpy as np
import torch


class Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, log_modulus, phase, sim, nsteps, every, dt, reset):
        if reset:
            sim.resetTime()
        sim.setModuli(torch.exp(log_modulus.detach().double()).cpu().numpy().astype(np.float32))
        sim.setPhases(phase.detach().cpu().numpy().astype(np.float32))
        sim.adjointRun(nsteps, every, dt)
        ctx.sim = sim
        ctx.device = log_modulus.device
        return torch.from_numpy(sim.positions()).to(log_modulus.device)

    @staticmethod
    def backward(ctx, grad_pos):
        ge, gp = ctx.sim.adjointBackward(grad_pos.detach().double().cpu().numpy())
        return (torch.from_numpy(ge).to(ctx.device), torch.from_numpy(gp).to(ctx.device),
                None, None, None, None, None)


def rollout(sim, log_modulus, phase, nsteps, every=100, dt=None, reset=True):
    """Final voxel positions after nsteps steps from the initial state (reset=True) or the
    current one; differentiable in log_modulus and phase (both (V,) tensors)."""
    return Rollout.apply(log_modulus, phase, sim, nsteps, every, dt, reset)


def robot_com(sim, pos):
    """Mass-weighted centre of each robot from a (V, 3) position tensor (differentiable): one
    bulk read of the voxels' masses and one segmented sum over the robots."""
    R = sim.robotCount()
    first = np.array([sim.robotVoxels(r)[0] for r in range(R)], np.int64)
    count = np.array([sim.robotVoxels(r)[1] for r in range(R)], np.int64)
    robot = np.repeat(np.arange(R), count)
    m = torch.from_numpy(sim.voxelMasses()).to(device=pos.device, dtype=pos.dtype)
    idx = torch.from_numpy(robot).to(pos.device)
    span = torch.from_numpy(np.concatenate([np.arange(f, f + c) for f, c in zip(first, count)])
                            if R else np.zeros(0, np.int64)).to(pos.device)
    w = pos[span] * m[span, None]
    num = torch.zeros((R, 3), dtype=pos.dtype, device=pos.device).index_add_(0, idx, w)
    den = torch.zeros(R, dtype=pos.dtype, device=pos.device).index_add_(0, idx, m[span])
    return num / den[:, None]
