"""torch.autograd bindings of the HIP reverse mode.

The reference differentiates rollouts of its step with jax reverse mode
(``production/autograd_cost_benchmark.py:228-330``: ``jax.value_and_grad`` over
``lax.scan(jax.checkpoint(step))``).  Here the step and its adjoint are
hand-written HIP kernels (``pcf_vjp`` in libpcfdns.so); PyTorch only carries
the cotangents:

* ``rollout(solver, steps, u0, u1, u2, g)`` maps an initial velocity/vorticity
  state (empty SBDF3 history) to the state ``steps`` steps later; its backward
  pulls any cotangent of the final (u0, u1, u2, g) back through the steps.
* ``rollout_energy(solver, steps, u0, u1, u2, g)`` returns the final
  perturbation energy ``Epert`` (``channelflow_kmm.py:1302-1308``); forward and
  reverse run in one ``pcf_vjp`` call and backward only scales the result.

Tensors are complex128 CPU tensors in the reference's full-complex layout.
PyTorch's gradient of a real loss with respect to a complex tensor is
dL/dRe + i dL/dIm, the complex conjugate of jax's convention, which the
library speaks at its boundary; the conversion is a conjugation.
"""

from __future__ import annotations

import numpy as np
import torch


def _state(solver, u0, u1, u2, g):
    u = tuple(t.detach().cpu().numpy().astype(np.complex128) for t in (u0, u1, u2))
    return solver._with_zero_history(u, g.detach().cpu().numpy().astype(np.complex128))


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, solver, steps, checkpoint_every, u0, u1, u2, g):
        state = _state(solver, u0, u1, u2, g)
        final = solver.solve(state, int(steps))
        ctx.solver, ctx.steps, ctx.every, ctx.state = solver, int(steps), int(checkpoint_every), state
        return tuple(torch.from_numpy(np.array(a)) for a in (*final.u, final.g))

    @staticmethod
    def backward(ctx, gu0, gu1, gu2, gg):
        def jax_cot(t, like):
            if t is None:
                return np.zeros_like(like)
            return np.conj(t.detach().cpu().numpy().astype(np.complex128))

        s = ctx.state
        cot = (jax_cot(gu0, s.u[0]), jax_cot(gu1, s.u[1]), jax_cot(gu2, s.u[2]), jax_cot(gg, s.g))
        _, grad = ctx.solver.vjp(s, ctx.steps, cotangent=cot, checkpoint_every=ctx.every)
        out = tuple(torch.from_numpy(np.conj(a)) for a in (*grad.u, grad.g))
        return (None, None, None, *out)


class _RolloutEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, solver, steps, checkpoint_every, u0, u1, u2, g):
        state = _state(solver, u0, u1, u2, g)
        value, grad, _ = solver.value_and_grad(state, int(steps), checkpoint_every=int(checkpoint_every))
        ctx.grad = tuple(torch.from_numpy(np.conj(a)) for a in (*grad.u, grad.g))
        return torch.tensor(value, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gout):
        return (None, None, None, *(gout * t for t in ctx.grad))


def rollout(solver, steps: int, u0, u1, u2, g, checkpoint_every: int = 20):
    """Differentiable ``solver.solve`` from a fresh state; returns (u0, u1, u2, g)."""
    return _Rollout.apply(solver, steps, checkpoint_every, u0, u1, u2, g)


def rollout_energy(solver, steps: int, u0, u1, u2, g, checkpoint_every: int = 20):
    """Differentiable final ``Epert`` after ``steps`` steps from a fresh state."""
    return _RolloutEnergy.apply(solver, steps, checkpoint_every, u0, u1, u2, g)
