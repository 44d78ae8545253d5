"""Reverse mode of the finite-volume step loop (hydro and MHD).

``rollout`` is a ``torch.autograd.Function``: forward runs the step loop with
the HIP kernels (``amx_fv_evolve_state``) and keeps a checkpoint every
``checkpoint_every`` steps; backward walks the segments from the last one,
recomputes each segment's step inputs from its checkpoint and applies the
hand-written reverse-mode kernels (``amx_fv_adjoint_step``) step by step.
Step sizes are recorded in the forward pass and held constant, as the
reference stop_gradients dt (time_integration.py:632-655); with
``config.fixed_timestep`` they are ``t_end / num_timesteps``
(time_integration.py:667-668). This replaces ``jax.grad`` through
``time_integration`` with ``differentiation_mode=BACKWARDS``
(time_integration.py:786-801), whose checkpointing is equinox's
``checkpointed_while_loop``.

Scope: the finite-volume path with HLL / HLLC, minmod and periodic
boundaries, hydro or MHD. For MHD the reverse step differentiates the
fixed-point applications the forward step made (their number is the
forward's own: the reverse kernel recomputes the step with the forward
kernels), as jax.grad through the reference's while loop does
(_magnetic_field_update.py:171-174).
"""

from __future__ import annotations

import numpy as np

from astronomix_amd import _native
from astronomix_amd._finite_volume import boundary_codes, native_config, require_supported
from astronomix_amd.option_classes.simulation_config import HLL, HLLC, MINMOD


def _function():
    import torch

    class Rollout(torch.autograd.Function):
        @staticmethod
        def forward(fctx, state, ctx, dts, every):
            s = state.detach().contiguous().clone()
            ckpts = [s.clone()]
            for k, dt in enumerate(dts):
                ctx.evolve_state(s, dt)
                if (k + 1) % every == 0 and k + 1 < len(dts):
                    ckpts.append(s.clone())
            fctx.ctx, fctx.dts, fctx.every, fctx.ckpts = ctx, dts, every, ckpts
            return s

        @staticmethod
        def backward(fctx, grad):
            import torch
            ctx, dts, every = fctx.ctx, fctx.dts, fctx.every
            lam = grad.contiguous().clone()
            lam_in = torch.empty_like(lam)
            nseg = (len(dts) + every - 1) // every
            for j in reversed(range(nseg)):
                lo, hi = j * every, min((j + 1) * every, len(dts))
                s = fctx.ckpts[j].clone()
                inputs = []
                for k in range(lo, hi):
                    inputs.append(s.clone())
                    ctx.evolve_state(s, dts[k])
                for k in reversed(range(lo, hi)):
                    ctx.adjoint_step(inputs[k - lo], dts[k], lam, lam_in)
                    lam, lam_in = lam_in, lam
            fctx.ckpts = None
            return lam, None, None, None

    return Rollout


_ROLLOUT = None


def step_sizes(state, ctx, t_end: float, num_steps: int = None, exact_end: bool = True):
    """Device dt tensors of the loop, the same float32 values the device loop
    of amx_fv_time_integration uses: t_end / num_steps (fixed) or the CFL
    step (adaptive), each clamped to t_end - t with t summed in float32 when
    exact_end (time_integration.py:667-678)."""
    import torch
    if num_steps is not None:
        te = np.float32(t_end)
        dt0 = te / np.float32(num_steps)
        t = np.float32(0.0)
        out = []
        for _ in range(num_steps):
            dt = min(dt0, np.float32(te - t)) if exact_end else dt0
            out.append(torch.tensor(dt, dtype=torch.float32, device=state.device))
            t = np.float32(t + dt)
        return out
    s = state.detach().contiguous().clone()
    dts, t = [], torch.zeros((), dtype=torch.float32, device=state.device)
    d = torch.empty((), dtype=torch.float32, device=state.device)
    t_end_t = torch.tensor(float(t_end), dtype=torch.float32, device=state.device)
    # the loop condition t < t_end is evaluated in float32, as the device loop
    # does (time_integration.py:772-777): comparing a float32 t with the double
    # t_end never ends once the clamp has put t on float32(t_end) < t_end
    while bool(t < t_end_t):
        ctx.cfl_dt(s, d)
        dt = torch.minimum(d, t_end_t - t).clone() if exact_end else d.clone()
        ctx.evolve_state(s, dt)
        t = t + dt
        dts.append(dt)
    return dts


def rollout(state, config, params, registered_variables=None, checkpoint_every: int = 50):
    """Differentiable time_integration of a hydro or MHD state (float32 GPU
    tensor): returns the final state; gradients flow to `state` through the
    reverse-mode kernels."""
    global _ROLLOUT
    import torch
    require_supported(config, registered_variables)
    if config.riemann_solver not in (HLL, HLLC) or config.limiter != MINMOD:
        raise NotImplementedError("reverse mode covers the HLL / HLLC fluxes with the minmod limiter")
    if any(boundary_codes(config)):
        raise NotImplementedError("reverse mode covers periodic boundaries only")
    if not isinstance(state, torch.Tensor) or state.device.type != "cuda" or state.dtype != torch.float32:
        raise TypeError("state must be a float32 tensor on the GPU")
    _native.require_gpu()
    if _ROLLOUT is None:
        _ROLLOUT = _function()
    ctx = _native.fv_context(native_config(state.shape, config, params.gamma, params.C_cfl))
    n = int(config.num_timesteps) if config.fixed_timestep else None
    dts = step_sizes(state, ctx, float(params.t_end), n, bool(config.exact_end_time))
    return _ROLLOUT.apply(state, ctx, dts, max(1, int(checkpoint_every)))


__all__ = ["rollout", "step_sizes"]
