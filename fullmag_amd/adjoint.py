"""Differentiable Heun rollouts of the FDM LLG step (torch.autograd).

`heun_rollout(m0, ms_cell, h_ext, plan, dt, steps, checkpoint_every)` runs
`steps` Heun steps of the HIP backend and returns the final magnetisation as
a CUDA float64 tensor of shape (cells, 3).  Its backward pass is the
discrete adjoint of the Heun step (fullmag_fdm_ext_heun_adjoint): gradients
with respect to per-cell Ms and the uniform external field H_ext [A/m].
Memory: checkpoints every `checkpoint_every` steps (device-resident FP64
states); the backward pass recomputes one segment at a time and stores its
states, so peak memory is (steps / K + K) states of 24 bytes per cell.

There is no counterpart in fullmag's CPU or CUDA engines; the forward is
the same fullmag_fdm_backend_step the runner drives
(crates/fullmag-runner/src/native_fdm.rs:370-410).
"""

from __future__ import annotations

from dataclasses import replace
from typing import Optional

import numpy as np

from .native_fdm import FdmPlan, NativeFdmBackend


def _torch():
    import torch

    return torch


def _backend(plan: FdmPlan, m0: np.ndarray, ms_cell: np.ndarray, h_ext) -> NativeFdmBackend:
    p = replace(plan, initial_magnetization=m0, ms_cell=ms_cell, external_field=tuple(float(v) for v in h_ext))
    if p.integrator != "heun":
        raise ValueError("heun_rollout differentiates the Heun integrator only")
    if p.precision not in ("double", "mixed"):
        raise ValueError("heun_rollout needs an FP64 state (precision 'double' or 'mixed')")
    return NativeFdmBackend(p)


class _HeunRollout:
    """Builds the autograd.Function lazily so importing fullmag_amd does not
    import torch."""

    fn = None

    @classmethod
    def get(cls):
        if cls.fn is not None:
            return cls.fn
        torch = _torch()

        class HeunRollout(torch.autograd.Function):
            @staticmethod
            def forward(ctx, ms_cell, h_ext, plan, m0, dt, steps, every):
                n = plan.cell_count
                b = _backend(plan, m0, ms_cell.detach().cpu().numpy().astype(np.float64),
                             h_ext.detach().cpu().numpy().astype(np.float64))
                ckpts = []
                state = torch.empty(3 * n, dtype=torch.float64, device="cuda")
                for s in range(steps):
                    if s % every == 0:
                        c = torch.empty_like(state)
                        b.copy_state_device(c.data_ptr())
                        ckpts.append(c)
                    b.step(dt)
                b.copy_state_device(state.data_ptr())
                ctx.backend, ctx.ckpts, ctx.dt, ctx.steps, ctx.every = b, ckpts, dt, steps, every
                ctx.ms_dtype, ctx.h_dtype = ms_cell.dtype, h_ext.dtype
                ctx.ms_device, ctx.h_device = ms_cell.device, h_ext.device
                return state.view(3, n).t().contiguous()

            @staticmethod
            def backward(ctx, grad_m):
                b, dt, steps, every = ctx.backend, ctx.dt, ctx.steps, ctx.every
                n = b.cell_count
                lam = grad_m.detach().to(torch.float64).t().contiguous().view(-1).cuda()
                lam_next = torch.empty_like(lam)
                gms = torch.zeros(n, dtype=torch.float64, device="cuda")
                gh = np.zeros(3)
                for seg in range(len(ctx.ckpts) - 1, -1, -1):
                    s0 = seg * every
                    cnt = min(every, steps - s0)
                    states = [ctx.ckpts[seg]]
                    b.set_state_device(ctx.ckpts[seg].data_ptr())
                    for _ in range(cnt - 1):
                        b.step(dt)
                        st = torch.empty_like(lam)
                        b.copy_state_device(st.data_ptr())
                        states.append(st)
                    for j in range(cnt - 1, -1, -1):
                        torch.cuda.synchronize()
                        b.heun_adjoint(dt, states[j].data_ptr(), lam.data_ptr(), lam_next.data_ptr(),
                                       gms.data_ptr(), gh)
                        lam, lam_next = lam_next, lam
                b.close()
                ctx.backend = None
                ctx.ckpts = None
                g_ms = gms.to(device=ctx.ms_device, dtype=ctx.ms_dtype)
                g_h = torch.tensor(gh, dtype=ctx.h_dtype, device=ctx.h_device)
                return g_ms, g_h, None, None, None, None, None

        cls.fn = HeunRollout
        return cls.fn


def heun_rollout(m0: np.ndarray, ms_cell, h_ext, plan: FdmPlan, dt: float, steps: int,
                 checkpoint_every: Optional[int] = None):
    """Final magnetisation (cells, 3) after `steps` Heun steps from m0, as a
    differentiable function of ms_cell (cells,) [A/m] and h_ext (3,) [A/m]
    (torch tensors).  `plan` supplies grid, material, precision and demag
    options; its initial_magnetization, ms_cell and external_field are
    replaced."""
    if steps < 1:
        raise ValueError("steps must be >= 1")
    every = int(checkpoint_every or max(1, int(round(np.sqrt(steps)))))
    m0 = np.ascontiguousarray(m0, dtype=np.float64).reshape(plan.cell_count, 3)
    return _HeunRollout.get().apply(ms_cell, h_ext, plan, m0, float(dt), int(steps), every)
