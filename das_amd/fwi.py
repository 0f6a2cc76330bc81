"""Waveform misfit of the acoustic engine and its gradient with respect to
velocity, behind ``torch.autograd.Function``.

The forward model is the reference's acoustic loop (tsw_rtm/wave_solver.py:
157-183 over tsw_rtm/c_engine/wave2d.c:19-82) on the HIP engine, in FP32 or
FP64: the recorded data of every shot in FP32 are bit-identical to
``WaveSolver2D.forward_propagate_acoustic``.  The misfit is

    J(v) = 1/2 sum_shots sum_{t, r} (d_pred - d_obs)^2

and its gradient is the exact discrete adjoint of that loop (sponge,
free-surface mirror, untouched edge cells and injection included), computed
by hand-written HIP kernels backward in time.  The source wavefield is not
stored: it is checkpointed every ceil(sqrt(nt)) steps and recomputed segment
by segment during the adjoint sweep, bit-identical to the first pass.

The reference migrates with the forward operator run on time-reversed data
(tsw_rtm/wave_solver.py:188-242); that image is not a gradient, and this module
has no counterpart there.  ``WaveformMisfit`` is the differentiable operator:

    op = WaveformMisfit(nx, nz, dx, dz, dt, shots, dtype=torch.float64)
    v = torch.tensor(model, device="cuda", dtype=torch.float64, requires_grad=True)
    loss = op(v); loss.backward()   # v.grad = dJ/dv
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Sequence

import numpy as np
import torch

from . import _native as N


@dataclass
class MisfitShot:
    """One shot: source time function (nt), source cell, receiver columns on
    row riz, and observed data (nt, nrec)."""
    src: np.ndarray
    six: int
    siz: int
    rix: np.ndarray
    riz: int
    d_obs: np.ndarray


class WaveformMisfit:
    """Owns the device state of the misfit (checkpoints, residuals, adjoint
    buffers) for a fixed grid and shot set.  Calling it on a velocity tensor
    (nz, nx) on the GPU returns J as a 0-d tensor that autograd differentiates.

    ckpt_interval: 0 = ceil(sqrt(nt)) (default), k > 0 = every k steps,
    -1 = no checkpoints (the adjoint sweep recomputes the whole forward once
    and keeps every level: for tests)."""

    def __init__(self, nx: int, nz: int, dx: float, dz: float, dt: float,
                 shots: Sequence[MisfitShot], n_sponge: int = 30, sponge_alpha: float = 0.0015,
                 free_surface_top: bool = True, dtype=torch.float32, ckpt_interval: int = 0):
        if dtype not in (torch.float32, torch.float64):
            raise ValueError("dtype must be torch.float32 or torch.float64")
        self.nx, self.nz = int(nx), int(nz)
        self.dtype = dtype
        npdt = np.float32 if dtype == torch.float32 else np.float64
        self._npdt = npdt
        keep = []
        arr = (N.FwiShot * len(shots))()
        self.shots = []
        for d, s in zip(arr, shots):
            src = np.ascontiguousarray(np.asarray(s.src).reshape(-1), dtype=npdt)
            rix = np.ascontiguousarray(np.asarray(s.rix).reshape(-1), dtype=np.int32)
            obs = np.ascontiguousarray(s.d_obs, dtype=npdt)
            if obs.shape != (len(src), len(rix)):
                raise ValueError(f"d_obs has shape {obs.shape}, expected {(len(src), len(rix))}")
            if rix.size and (rix.min() < 0 or rix.max() >= nx):
                raise ValueError("receiver_ix out of range")
            keep += [src, rix, obs]
            d.nt, d.src, d.six, d.siz = len(src), N.ptr(src), int(s.six), int(s.siz)
            d.rix, d.nrec, d.riz, d.d_obs = N.ptr(rix), len(rix), int(s.riz), N.ptr(obs)
            self.shots.append((len(src), len(rix)))
        h = N.P()
        alpha32 = float(np.float32(sponge_alpha))
        N.check(N.lib().tsw_fwi_create(self.nx, self.nz, float(dx), float(dz), float(dt),
                                       int(n_sponge), alpha32, 1 if free_surface_top else 0,
                                       8 if dtype == torch.float64 else 4, len(shots), arr,
                                       int(ckpt_interval), ctypes.byref(h)))
        self._h = h
        self._last_vel = None

    def close(self):
        if getattr(self, "_h", None) is not None and N._lib is not None:
            N.lib().tsw_fwi_destroy(self._h)
            self._h = None

    __del__ = close

    def _check(self, vel: torch.Tensor):
        if not vel.is_cuda:
            raise ValueError("velocity must be a GPU tensor")
        if vel.dtype != self.dtype or tuple(vel.shape) != (self.nz, self.nx):
            raise ValueError(f"velocity must be {self.dtype} of shape {(self.nz, self.nx)}")

    def forward_loss(self, vel: torch.Tensor) -> float:
        """J(vel); keeps the checkpoints and residuals for ``gradient``."""
        self._check(vel)
        vel = vel.detach().contiguous()
        loss = N.D()
        stream = torch.cuda.current_stream(vel.device).cuda_stream
        N.check(N.lib().tsw_fwi_forward(self._h, ctypes.c_void_p(vel.data_ptr()),
                                        ctypes.c_void_p(stream), ctypes.byref(loss)))
        self._last_vel = vel.clone()
        return loss.value

    def gradient(self, vel: torch.Tensor, out: torch.Tensor = None) -> torch.Tensor:
        """dJ/dv at vel (runs the forward again if the last one was at
        another velocity)."""
        self._check(vel)
        vel = vel.detach().contiguous()
        if self._last_vel is None or not torch.equal(self._last_vel, vel):
            self.forward_loss(vel)
        g = torch.empty_like(vel) if out is None else out
        stream = torch.cuda.current_stream(vel.device).cuda_stream
        N.check(N.lib().tsw_fwi_backward(self._h, ctypes.c_void_p(vel.data_ptr()),
                                         ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(stream)))
        return g

    def set_blocking(self, config: int):
        """FP32 forward and recompute: 0 = one launch per step, 1..9 = a
        temporally blocked configuration (default 2); results are
        bit-identical either way."""
        N.check(N.lib().tsw_fwi_set_blocking(self._h, int(config)))
        self._last_vel = None

    def set_adjoint_blocking(self, config: int):
        """FP32 adjoint: 0 = one launch per step (or two, see
        set_adjoint_steps), 1..9 = a temporally blocked configuration, the
        shapes of set_blocking (default 2); the gradient is bit-identical
        either way."""
        N.check(N.lib().tsw_fwi_set_adjoint_blocking(self._h, int(config)))

    def set_adjoint_steps(self, steps: int):
        """Per-step adjoint (set_adjoint_blocking(0)): steps per launch, 1
        (default), or 2 (one launch computes two consecutive levels; slower on
        MI355X at 300 x 800).  The gradient is bit-identical either way."""
        N.check(N.lib().tsw_fwi_set_adjoint_steps(self._h, int(steps)))

    def predicted(self, shot: int) -> np.ndarray:
        """Predicted data (nt, nrec) of a shot from the last forward."""
        nt, nrec = self.shots[shot]
        out = np.empty((nt, nrec), self._npdt)
        N.check(N.lib().tsw_fwi_predicted(self._h, int(shot), N.ptr(out)))
        return out

    def set_residual(self, shot: int, res: np.ndarray):
        """Replaces a shot's residual before the adjoint sweep (adjoint tests)."""
        nt, nrec = self.shots[shot]
        r = np.ascontiguousarray(res, dtype=self._npdt)
        if r.shape != (nt, nrec):
            raise ValueError(f"residual must have shape {(nt, nrec)}")
        N.check(N.lib().tsw_fwi_set_residual(self._h, int(shot), N.ptr(r)))

    def adjoint_at_source(self, shot: int) -> np.ndarray:
        """dJ/du^m at the shot's source cell for m = 0..nt-1, from the last
        adjoint sweep (u^m the post-injection level at iteration m)."""
        nt, _ = self.shots[shot]
        out = np.empty(nt, self._npdt)
        N.check(N.lib().tsw_fwi_probe(self._h, int(shot), N.ptr(out)))
        return out

    def __call__(self, vel: torch.Tensor) -> torch.Tensor:
        return MisfitFunction.apply(vel, self)


class MisfitFunction(torch.autograd.Function):
    """J = op(vel) with dJ/dvel from the hand-written adjoint kernels."""

    @staticmethod
    def forward(ctx, vel: torch.Tensor, op: WaveformMisfit):
        v = vel.detach().contiguous()
        loss = op.forward_loss(v)
        ctx.op = op
        ctx.save_for_backward(v)
        return torch.tensor(loss, dtype=vel.dtype, device=vel.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        (v,) = ctx.saved_tensors
        g = ctx.op.gradient(v)  # runs the forward again if another one ran since
        return g * grad_out, None
