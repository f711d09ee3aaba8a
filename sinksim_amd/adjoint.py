"""The adjoint as a torch.autograd.Function: losses of a batch of instances as a differentiable function of their
scale factors theta [count, nparams] (no counterpart in the reference; its calibration is a gradient-free search,
tools/py/calibrate_gpu.py:95-124).

The forward runs ss_batch_gradient (the GPU adjoint kernel: forward with checkpoints, then the reverse sweep when a
gradient is wanted) and keeps d loss / d theta for backward; backward scales it by the incoming gradient. PyTorch is
used only here and in the calibration driver; the kernel takes and returns host arrays.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .batch import Batch, BatchOptions
from .calib import LossSpec, ParamMap


@dataclass
class GradientOptions:
    checkpoint_every: int = 256
    relaxed: bool = True
    t_max: float = 0.0       # 0: the batch's
    max_steps: int = 0       # <= 0: none
    numerics: str = "portable"  # the forward's: "portable" (FP64) or "portable32" (FP32 partials; the reverse stays FP64)


class GradientRunner:
    """A batch of `count` instances of one model (or a fleet) with a parameter map and a loss; every call evaluates all
    instances at once. Keeps the device time and the simulated steps of every call."""

    def __init__(self, models, inst_model, params, spec: LossSpec, t_max: float, options: GradientOptions | None = None,
                 tables=None):
        self.models = list(models)
        self.inst_model = np.asarray(inst_model, np.int32)
        self.count = int(self.inst_model.size)
        self.params = params
        self.spec = spec
        self.options = options or GradientOptions()
        if self.options.numerics not in ("portable", "portable32"):
            raise ValueError(f"numerics must be portable or portable32, not {self.options.numerics!r}")
        self.batch = Batch.fleet(self.models, self.inst_model,
                                 BatchOptions(count=self.count, mixed_precision=self.options.numerics == "portable32",
                                              traced_instances=0, curve_every=0.0, t_max=t_max))
        if tables is not None:
            self.batch.set_connections(*tables)
        self.kernel_ms = 0.0
        self.calls = 0
        self.sim_steps = 0

    @property
    def nparams(self) -> int:
        return (self.params[0] if isinstance(self.params, (list, tuple)) else self.params).count

    def __call__(self, theta: np.ndarray, observations=None, want_grad: bool = True) -> dict:
        o = self.options
        r = self.batch.gradient(self.params, theta, self.spec, observations=observations, checkpoint_every=o.checkpoint_every,
                                relaxed=o.relaxed, t_max=o.t_max, max_steps=o.max_steps, want_grad=want_grad)
        if int(r["mismatches"].sum()) != 0:
            raise RuntimeError(f"{int(r['mismatches'].sum())} recomputed steps differ from the forward trajectory")
        self.kernel_ms += r["ms"]
        self.calls += 1
        self.sim_steps += int(r["steps"].sum())
        return r

    def close(self) -> None:
        self.batch.close()


class FloodLoss(torch.autograd.Function):
    """loss [count] = FloodLoss.apply(theta [count, nparams] float64, runner, observations or None)."""

    @staticmethod
    def forward(ctx, theta, runner, observations):
        th = theta.detach().to("cpu", torch.float64).numpy()
        r = runner(th, observations=observations, want_grad=bool(ctx.needs_input_grad[0]))
        if r["grad"] is not None:
            ctx.save_for_backward(torch.from_numpy(np.ascontiguousarray(r["grad"])))
        ctx.signals = r["signals"]
        return torch.from_numpy(r["loss"]).to(device=theta.device, dtype=theta.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out[:, None] * g.to(device=grad_out.device, dtype=grad_out.dtype), None, None


def flood_loss(theta: torch.Tensor, runner: GradientRunner, observations=None) -> torch.Tensor:
    """Every instance's loss at theta, differentiable with respect to theta."""
    return FloodLoss.apply(theta, runner, observations)


__all__ = ["GradientOptions", "GradientRunner", "FloodLoss", "flood_loss", "ParamMap"]
