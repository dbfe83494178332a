"""Differentiable window probabilities: a ``torch.autograd.Function`` over the
engine's C-ABI.

No counterpart in the reference (its rates are compile-time constants and it
differentiates nothing).  ``forward`` runs the simulate kernel with the rate
vector ``theta`` read from device memory (stream tag 0) and returns the window
probabilities ``P[w, x]`` (observed state ``x``) as a float64 device tensor
normalised on the device (``mbg_window_probs_device``); ``backward`` hands
``dL/dP`` to ``mbg_backward``, which replays the same trajectories and
returns the likelihood-ratio estimate of ``dL/dtheta``.  Both run on torch's
current stream and read and write device tensors through their
``data_ptr``; nothing goes through the host.
"""
from __future__ import annotations

import ctypes

from ._capi import GradInfo, GradParams, SimParams, check, lib

__all__ = ["window_probabilities", "WindowProbabilities", "observed_marginals", "backward", "window_probs_device",
           "ACCUMULATOR"]

# Precision of the score accumulation: per trajectory, then across trajectories.
ACCUMULATOR = "f64 per trajectory, f64 across trajectories"


def _torch():
    import torch
    return torch


def backward(program, grad_probs, rates=None, n_trajectories: int = -1, seed: int = -1, variant: int = 0,
             first_id: int = 0, max_steps: int = 0, block_size: int = 0, stream=None, info: bool = False):
    """dL/dtheta for ``grad_probs`` = dL/dP (device float64 tensor [windows, 2^n_obs]).

    ``rates`` is a device float32 tensor of the parameters (None: the model's
    values).  Returns a device float64 tensor [n_params] (and the launch's
    :class:`GradInfo` when ``info``)."""
    torch = _torch()
    m = program.model.info
    if grad_probs.dtype != torch.float64 or not grad_probs.is_cuda:
        raise TypeError("grad_probs must be a float64 CUDA tensor")
    g = grad_probs.contiguous()
    if g.numel() != m.windows << m.n_observed:
        raise ValueError(f"grad_probs must have {m.windows} x {1 << m.n_observed} entries")
    r = None
    if rates is not None:
        r = rates.detach().to(device=g.device, dtype=torch.float32).contiguous()
        if r.numel() != m.n_params:
            raise ValueError(f"rates must have {m.n_params} entries")
    out = torch.empty(m.n_params, dtype=torch.float64, device=g.device)
    gp = GradParams()
    lib().mbg_grad_params_default(ctypes.byref(gp))
    gp.n_trajectories = n_trajectories
    gp.seed = seed
    gp.variant = variant
    gp.first_id = first_id
    gp.max_steps = max_steps
    gp.block_size = block_size
    s = stream if stream is not None else torch.cuda.current_stream(g.device).cuda_stream
    gi = GradInfo()
    check(lib().mbg_backward(program._h, ctypes.byref(gp), ctypes.c_void_p(r.data_ptr()) if r is not None else None,
                             ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(s),
                             ctypes.byref(gi) if info else None))
    return (out, gi) if info else out


def window_probs_device(program, rates=None, n_trajectories: int = -1, seed: int = -1, max_steps: int = 0,
                        first_id: int = 0):
    """P[w, x] of a run as a device float64 tensor [windows, 2^n_obs], computed and
    normalised on the device (``mbg_window_probs_device``) on torch's current stream;
    ``rates`` is a device float32 tensor of the parameters (None: the model's).  The
    tensors are allocated on that stream, so the caching allocator orders their reuse
    after the kernels that read and write them."""
    torch = _torch()
    m = program.model.info
    dev = rates.device if rates is not None and rates.is_cuda else torch.device("cuda")
    r = None
    if rates is not None:
        r = rates.detach().to(device=dev, dtype=torch.float32).contiguous()
        if r.numel() != m.n_params:
            raise ValueError(f"rates must have {m.n_params} entries")
    P = torch.empty((m.windows, 1 << m.n_observed), dtype=torch.float64, device=dev)
    p = SimParams()
    lib().mbg_sim_params_default(ctypes.byref(p))
    p.n_trajectories = n_trajectories
    p.seed = seed
    p.max_steps = max_steps
    p.first_id = first_id
    s = torch.cuda.current_stream(dev).cuda_stream
    check(lib().mbg_window_probs_device(program._h, ctypes.byref(p),
                                        ctypes.c_void_p(r.data_ptr()) if r is not None else None,
                                        ctypes.c_void_p(P.data_ptr()), ctypes.c_void_p(s)))
    return P


class WindowProbabilities:
    """``torch.autograd.Function`` (built lazily so importing the package does not import torch)."""

    _cls = None

    @classmethod
    def get(cls):
        if cls._cls is None:
            torch = _torch()

            class _F(torch.autograd.Function):
                @staticmethod
                def forward(ctx, theta, program, n_trajectories, seed, max_steps):
                    dev = theta.device if theta.is_cuda else torch.device("cuda")
                    th = theta.detach().to(device=dev, dtype=torch.float32)
                    P = window_probs_device(program, th, n_trajectories=n_trajectories, seed=seed,
                                            max_steps=max_steps)
                    ctx.program = program
                    ctx.n = n_trajectories
                    ctx.seed = seed
                    ctx.max_steps = max_steps
                    ctx.save_for_backward(theta)
                    return P

                @staticmethod
                def backward(ctx, gP):
                    (theta,) = ctx.saved_tensors
                    dev = gP.device if gP.is_cuda else torch.device("cuda")
                    g = backward(ctx.program, gP.to(device=dev, dtype=torch.float64), rates=theta.to(dev),
                                 n_trajectories=ctx.n, seed=ctx.seed, max_steps=ctx.max_steps)
                    return g.to(device=theta.device, dtype=theta.dtype), None, None, None, None

            cls._cls = _F
        return cls._cls


def window_probabilities(program, theta, n_trajectories: int, seed: int = -1, max_steps: int = 0):
    """P[w, x] at rates ``theta`` (1-D tensor in ``Model.param_names`` order), differentiable in theta."""
    return WindowProbabilities.get().apply(theta, program, int(n_trajectories), int(seed), int(max_steps))


def observed_marginals(P, n_obs: int):
    """Per-window probability that each observed node is on: P @ bits, [windows, n_obs]."""
    torch = _torch()
    x = torch.arange(1 << n_obs, device=P.device)
    bits = ((x[:, None] >> torch.arange(n_obs, device=P.device)[None, :]) & 1).to(P.dtype)
    return P @ bits
