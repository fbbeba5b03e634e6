"""GPU reverse mode: the HIP adjoint against the oracle's adjoint and finite differences.

Bars:
* FP64 gradients agree with the oracle adjoint (oracle/adjoint.py, itself pinned
  by central differences and per-stage adjoint identities) to 1e-12 relative
  per leaf group, off the Nyquist planes (states keep those planes at zero, so
  no perturbation of a state moves them; the library returns zero there);
* directional derivatives match central differences of the HIP forward to 1e-6
  relative;
* the gradient does not depend on the checkpoint interval (bitwise);
* the mixed-precision gradient stays within 1e-4 relative of the FP64 one.
"""

import numpy as np
import pytest

from jaxfun_amd import PlaneCouetteFluctuationHIP
from jaxfun_amd.ic import seeded_state
from oracle.adjoint import OracleAdjoint, from_jax_convention, jax_convention, zeros_like_state
from tests.helpers import LEAF_NAMES, leaves, oracle_for, to_oracle

pytestmark = pytest.mark.gpu

GRAD_TOL = 1e-12


def make(N, Re=400.0, dt=2e-3, precision="fp64"):
    return PlaneCouetteFluctuationHIP(N=N, Re=Re, dt=dt, precision=precision, time_integrator="SBDF3",
                                      nonlinear_form="rotational")


def _mask(solver, a):
    a = np.array(a)
    if a.ndim == 3:
        a[:, solver.N[1] // 2, :] = 0
        a[:, :, solver.N[2] // 2] = 0
    return a


def grad_errors(solver, got, ref):
    """Per-leaf max |got - ref| / scale on the masked leaves, scale floored at
    1e-6 of the leaf group's largest entry (as tests/helpers.max_rel_errors)."""
    la = [_mask(solver, x) for x in leaves(got)]
    lb = [_mask(solver, x) for x in leaves(ref)]
    groups = [range(0, 4)] + [range(4 + 4 * k, 8 + 4 * k) for k in range(4)]
    out = {}
    for grp in groups:
        gmax = max(float(np.max(np.abs(lb[i]))) for i in grp)
        for i in grp:
            scale = max(float(np.max(np.abs(lb[i]))), 1e-6 * gmax)
            err = float(np.max(np.abs(la[i] - lb[i])))
            out[LEAF_NAMES[i]] = err / scale if scale > 0 else err
    return out


def _axpy(s, d, eps):
    f = lambda a, b: tuple(x + eps * y for x, y in zip(a, b))  # noqa: E731
    return type(s)(u=f(s.u, d.u), g=s.g + eps * d.g, nonlinear_old=s.nonlinear_old,
                   nonlinear_older=s.nonlinear_older, solution_old=s.solution_old,
                   solution_older=s.solution_older, history_steps=s.history_steps,
                   have_old=s.have_old, previous_dt=s.previous_dt)


def _pair(grad, d):
    """Directional derivative Re sum(g d) over the velocity leaves (jax convention)."""
    return float(sum(np.real(np.sum(a * b)) for a, b in zip(list(grad.u) + [grad.g], list(d.u) + [d.g])))


@pytest.mark.parametrize("N,steps,every", [((17, 16, 16), 5, 2), ((24, 16, 32), 4, 4)])
def test_energy_gradient_matches_oracle(N, steps, every):
    """Fresh state: two IMEXRK3 start-up steps then steady steps, with segments."""
    solver = make(N)
    s0 = seeded_state(solver, seed=1, energy=5e-2, K=4)
    value, grad, final = solver.value_and_grad(s0, steps, checkpoint_every=every)
    o = oracle_for(solver)
    adj = OracleAdjoint(o)
    v_ref, mu = adj.value_and_grad(o.reset_history(to_oracle(s0)), steps)
    assert abs(value - v_ref) <= 1e-12 * abs(v_ref)
    err = grad_errors(solver, grad, jax_convention(mu))
    print(N, steps, max(err.values()), flush=True)
    assert max(err.values()) < GRAD_TOL, err
    # the handle ends on the final state, as solve() leaves it
    ref_final = make(N).solve(s0, steps)
    assert all(np.array_equal(a, b) for a, b in zip(final.u, ref_final.u))


def test_history_state_gradient_matches_oracle():
    """State with full SBDF3 history: steady steps only, history leaves get gradients."""
    solver = make((17, 16, 16))
    s = solver.solve(seeded_state(solver, seed=2, energy=5e-2, K=4), 3)
    _, grad, _ = solver.value_and_grad(s, 4, checkpoint_every=3)
    o = oracle_for(solver)
    _, mu = OracleAdjoint(o).value_and_grad(to_oracle(s), 4)
    err = grad_errors(solver, grad, jax_convention(mu))
    print(max(err.values()), flush=True)
    assert max(err.values()) < GRAD_TOL, err
    assert np.max(np.abs(grad.solution_old[0])) > 0


def test_cotangent_vjp_matches_oracle():
    """A general cotangent of the final (u, g): pcf_set_cotangent seed."""
    solver = make((17, 16, 16))
    s0 = seeded_state(solver, seed=3, energy=5e-2, K=4)
    c = seeded_state(solver, seed=4, energy=1.0, K=6)  # Hermitian, masked
    cot = (np.conj(c.u[0]), np.conj(c.u[1]), np.conj(c.u[2]), np.conj(c.g))
    _, grad = solver.vjp(s0, 3, cotangent=cot, checkpoint_every=2)
    o = oracle_for(solver)
    so = o.reset_history(to_oracle(s0))
    s_final = o.solve(so, 3)
    mu_final = zeros_like_state(s_final)
    mu_final = type(mu_final)(**{**mu_final.__dict__, "u": tuple(np.conj(x) for x in cot[:3]),
                                 "g": np.conj(cot[3])})
    ref = jax_convention(OracleAdjoint(o).vjp(so, 3, from_jax_convention(jax_convention(mu_final))))
    err = grad_errors(solver, grad, ref)
    print(max(err.values()), flush=True)
    assert max(err.values()) < GRAD_TOL, err


def test_gradient_matches_central_differences():
    """Directional derivatives of the 33^3 energy after 12 steps vs the HIP forward."""
    N, steps = (33, 32, 32), 12
    solver = make(N, Re=1000.0, dt=5e-3)
    s0 = seeded_state(solver, seed=0, energy=1e-2, K=8)
    value, grad, _ = solver.value_and_grad(s0, steps, checkpoint_every=5)
    fwd = make(N, Re=1000.0, dt=5e-3)
    for seed in (10, 11, 12):
        d = seeded_state(solver, seed=seed, energy=1e-2, K=8)
        eps = 1e-4
        jp = fwd.diagnostics(fwd.solve(_axpy(s0, d, eps), steps))["Epert"]
        jm = fwd.diagnostics(fwd.solve(_axpy(s0, d, -eps), steps))["Epert"]
        fd = (jp - jm) / (2 * eps)
        an = _pair(grad, d)
        print(seed, an, fd, abs(an - fd) / abs(fd), flush=True)
        assert abs(an - fd) <= 1e-6 * abs(fd)


@pytest.mark.parametrize("N,steps,intervals", [((17, 16, 16), 7, (1, 3, 7)), ((33, 32, 32), 24, (24, 2, 5, 5))])
def test_checkpoint_interval_does_not_change_gradient(N, steps, intervals):
    """Bit-identical gradients for every checkpoint interval: the recompute of
    one segment runs concurrently with the reverse sweep of the next on a second
    stream, with two tape sets; a repeated interval replays cached graphs."""
    solver = make(N)
    s0 = seeded_state(solver, seed=5, energy=5e-2, K=4)
    grads = [solver.value_and_grad(s0, steps, checkpoint_every=k)[1] for k in intervals]
    for g in grads[1:]:
        assert all(np.array_equal(a, b) for a, b in zip(leaves(g), leaves(grads[0])))


def test_mixed_precision_gradient_close_to_fp64():
    N = (33, 32, 32)
    s_solver = make(N, Re=1000.0, dt=5e-3)
    s0 = seeded_state(s_solver, seed=0, energy=1e-3, K=8)
    _, g64, _ = s_solver.value_and_grad(s0, 10, checkpoint_every=5)
    _, g32, _ = make(N, Re=1000.0, dt=5e-3, precision="mixed").value_and_grad(s0, 10, checkpoint_every=5)
    err = max(float(np.max(np.abs(a - b))) / float(np.max(np.abs(b)))
              for a, b in zip(list(g32.u) + [g32.g], list(g64.u) + [g64.g]))
    print("mixed vs fp64 gradient", err, flush=True)
    assert 1e-14 < err < 1e-4


def test_torch_autograd_rollout_matches_finite_differences():
    """torch.autograd over rollout(): a loss mixing several final leaves."""
    import torch

    from jaxfun_amd.autograd import rollout, rollout_energy

    N, steps = (17, 16, 16), 5
    solver = make(N)
    s0 = seeded_state(solver, seed=6, energy=5e-2, K=4)
    w = seeded_state(solver, seed=7, energy=1.0, K=4).u[1]

    def loss_np(final):
        return float(np.sum(np.abs(final.u[0]) ** 2) + np.sum(np.real(final.u[1] * np.conj(w)))
                     + np.sum(np.abs(final.g) ** 2))

    xs = [torch.from_numpy(np.array(a)).requires_grad_() for a in (*s0.u, s0.g)]
    out = rollout(solver, steps, *xs, checkpoint_every=2)
    wt = torch.from_numpy(np.array(w))
    loss = (out[0].abs() ** 2).sum() + (out[1] * wt.conj()).real.sum() + (out[3].abs() ** 2).sum()
    assert abs(float(loss) - loss_np(make(N).solve(s0, steps))) <= 1e-12 * abs(float(loss))
    loss.backward()
    fwd = make(N)
    d = seeded_state(solver, seed=8, energy=5e-2, K=4)
    eps = 1e-5
    fd = (loss_np(fwd.solve(_axpy(s0, d, eps), steps)) - loss_np(fwd.solve(_axpy(s0, d, -eps), steps))) / (2 * eps)
    # torch's gradient of a real loss is dL/dRe + i dL/dIm: pairs as Re sum(conj(grad) d)
    an = float(sum(np.real(np.sum(np.conj(x.grad.numpy()) * b)) for x, b in zip(xs, (*d.u, d.g))))
    print("torch rollout", an, fd, flush=True)
    assert abs(an - fd) <= 1e-7 * abs(fd)

    ys = [torch.from_numpy(np.array(a)).requires_grad_() for a in (*s0.u, s0.g)]
    E = rollout_energy(solver, steps, *ys, checkpoint_every=2)
    (3.0 * E).backward()
    _, grad, _ = make(N).value_and_grad(s0, steps, checkpoint_every=2)
    for y, gj in zip(ys, (*grad.u, grad.g)):
        assert np.array_equal(y.grad.numpy(), 3.0 * np.conj(gj))
