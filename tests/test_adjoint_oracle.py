"""Pins of the oracle's hand-written adjoint (oracle/adjoint.py), CPU only.

* central finite differences of the oracle forward along seeded directions,
  the check the reference itself makes for its gradients
  (``tests/couette/test_differentiability_jax.py:81-136``);
* the adjoint identity <A v, w> = <v, A^H w> for every linear stage, so that a
  dropped term or a wrong sign in one stage fails on its own.
"""

from dataclasses import replace

import numpy as np
import pytest

from oracle.adjoint import OracleAdjoint, jax_convention, pairing
from oracle.kmm import PCFOracle

N = (9, 8, 8)


def _oracle(**kw):
    args = dict(N=N, Re=200.0, dt=2e-3, perturbation_amplitude=0.3)
    args.update(kw)
    return PCFOracle(**args)


def _rand_phys(o, rng):
    return rng.standard_normal((o.N0, o.N[1], o.N[2]))


def _rand_state_dir(o, rng, history=False):
    """A Hermitian direction: coefficients of random real fields (history too if asked)."""
    d = o.state_from_physical(tuple(_rand_phys(o, rng) + 0j for _ in range(3)))

    def grp():
        u0 = o.mask(o.tp_forward(_rand_phys(o, rng) + 0j, o.SB, o.MB1))
        g = o.mask(o.tp_forward(_rand_phys(o, rng) + 0j, o.SD, o.MD1))
        return (u0, g, rng.standard_normal(o.ND), rng.standard_normal(o.ND))

    if history:
        d = replace(d, nonlinear_old=grp(), nonlinear_older=grp(), solution_old=grp(), solution_older=grp())
    return d


def _axpy(s, d, eps):
    f = lambda a, b: tuple(x + eps * y for x, y in zip(a, b))  # noqa: E731
    return replace(s, u=f(s.u, d.u), g=s.g + eps * d.g, nonlinear_old=f(s.nonlinear_old, d.nonlinear_old),
                   nonlinear_older=f(s.nonlinear_older, d.nonlinear_older),
                   solution_old=f(s.solution_old, d.solution_old),
                   solution_older=f(s.solution_older, d.solution_older))


def _fd(o, adj, s, d, steps, eps):
    jp = adj.energy(o.solve(_axpy(s, d, eps), steps))
    jm = adj.energy(o.solve(_axpy(s, d, -eps), steps))
    return (jp - jm) / (2 * eps)


@pytest.mark.parametrize("steps", [1, 2, 5])
def test_gradient_matches_central_differences_from_fresh_state(steps):
    """Startup (IMEXRK3) and steady SBDF3 steps from an initial state."""
    o = _oracle()
    adj = OracleAdjoint(o)
    s0 = o.initial_state()
    value, mu = adj.value_and_grad(s0, steps)
    assert np.isclose(value, o.diagnostics(o.solve(s0, steps))["Epert"], rtol=1e-14)
    rng = np.random.default_rng(steps)
    for _ in range(3):
        d = _rand_state_dir(o, rng)
        fd = _fd(o, adj, s0, d, steps, 1e-5)
        an = pairing(mu, d)
        # central differences at eps=1e-5 carry ~1e-10 truncation + rounding error
        assert abs(an - fd) <= 1e-8 * abs(fd), (an, fd)


def test_gradient_with_full_history_includes_history_leaves():
    """Steady steps only: directions also move the SBDF3 history leaves."""
    o = _oracle()
    adj = OracleAdjoint(o)
    s = o.solve(o.initial_state(), 3)
    assert s.history_steps == 2.0
    value, mu = adj.value_and_grad(s, 3)
    rng = np.random.default_rng(11)
    for _ in range(3):
        d = _rand_state_dir(o, rng, history=True)
        fd = _fd(o, adj, s, d, 3, 1e-5)
        an = pairing(mu, d)
        assert abs(an - fd) <= 1e-8 * abs(fd), (an, fd)
    # history leaves carry gradient in the steady regime
    assert max(np.max(np.abs(a)) for a in mu.solution_old) > 0


def test_fresh_state_history_has_zero_gradient():
    """The IMEXRK3 start-up reads none of the (empty) history."""
    o = _oracle()
    adj = OracleAdjoint(o)
    _, mu = adj.value_and_grad(o.initial_state(), 4)
    for grp in (mu.nonlinear_old, mu.nonlinear_older, mu.solution_old, mu.solution_older):
        for a in grp:
            assert np.max(np.abs(a)) == 0.0


def test_jax_convention_directional_derivative():
    """jax's gradient g pairs as Re sum(g d) (examples/pcf_minimal_seed_jax.py:78-94)."""
    o = _oracle()
    adj = OracleAdjoint(o)
    s0 = o.initial_state()
    _, mu = adj.value_and_grad(s0, 2)
    g = jax_convention(mu)
    d = _rand_state_dir(o, np.random.default_rng(5))
    lhs = sum(float(np.real(np.sum(a * b))) for a, b in zip(list(g.u) + [g.g], list(d.u) + [d.g]))
    assert np.isclose(lhs, pairing(mu, d), rtol=1e-13)


def _cplx(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _dot(a, b):
    return float(np.real(np.sum(np.conj(a) * b)))


def test_stage_adjoint_identities():
    o = _oracle()
    adj = OracleAdjoint(o)
    rng = np.random.default_rng(2)
    n0, N1, N2 = o.N0, o.N[1], o.N[2]
    P1, P2 = o.pad[1], o.pad[2]
    # tp_backward (padded and not, all derivative flavours)
    for S, nS in ((o.SB, o.NB), (o.SD, o.ND)):
        for kw in ({}, {"kx": 1}, {"k1": 1}, {"k2": 1}):
            for padded, (p1, p2) in ((True, (P1, P2)), (False, (N1, N2))):
                v = _cplx(rng, (nS, N1, N2))
                w = _cplx(rng, (n0, p1, p2))
                lhs = _dot(w, o.tp_backward(v, S, padded=padded, **kw))
                rhs = _dot(adj.tp_backward_adj(w, S, padded=padded, **kw), v)
                assert np.isclose(lhs, rhs, rtol=1e-12), (kw, padded)
    # tp_forward_D
    v = _cplx(rng, (n0, P1, P2))
    w = _cplx(rng, (o.ND, N1, N2))
    assert np.isclose(_dot(w, o.tp_forward_D(v)), _dot(adj.tp_forward_D_adj(w), v), rtol=1e-12)
    # nonlinear_rhs (H -> Nu, Ng, Nv, Nw); Nv/Nw read Re H(0,0) only
    H = tuple(_cplx(rng, (o.ND, N1, N2)) for _ in range(3))
    mu = (_cplx(rng, (o.NB, N1, N2)), _cplx(rng, (o.ND, N1, N2)), rng.standard_normal(o.ND),
          rng.standard_normal(o.ND))
    out = o.nonlinear_rhs(H)
    lhs = sum(_dot(m, x) for m, x in zip(mu, out))
    rhs = sum(_dot(m, h) for m, h in zip(adj.nonlinear_rhs_adj(mu), H))
    assert np.isclose(lhs, rhs, rtol=1e-12)
    # reconstruct ((u0, g, v, w) -> u)
    x = (_cplx(rng, (o.NB, N1, N2)), _cplx(rng, (o.ND, N1, N2)), rng.standard_normal(o.ND),
         rng.standard_normal(o.ND))
    mu_u = (_cplx(rng, (o.NB, N1, N2)), _cplx(rng, (o.ND, N1, N2)), _cplx(rng, (o.ND, N1, N2)))
    lhs = sum(_dot(m, y) for m, y in zip(mu_u, o.reconstruct(*x)))
    rhs = sum(_dot(m, y) for m, y in zip(adj.reconstruct_adj(mu_u), x))
    assert np.isclose(lhs, rhs, rtol=1e-12)
    # solves, mass and linear operators
    gamma = 0.37 * o.dt
    rhs_in = (_cplx(rng, (o.NB, N1, N2)), _cplx(rng, (o.ND, N1, N2)), rng.standard_normal(o.ND),
              rng.standard_normal(o.ND))
    sol = o._solve(o._fac(gamma), *rhs_in)
    lhs = sum(_dot(m, y) for m, y in zip(x, sol))
    rhs = sum(_dot(m, y) for m, y in zip(adj.solve_adj(gamma, x), rhs_in))
    assert np.isclose(lhs, rhs, rtol=1e-12)
    for fwd, bwd in ((o.mass_rows, adj.mass_adj),
                     (lambda p: (o.Lu_apply(p[0]), o.Lg_apply(p[1]), o.L00 @ p[2], o.L00 @ p[3]), adj.lin_adj)):
        lhs = sum(_dot(m, y) for m, y in zip(rhs_in, fwd(x)))
        rhs = sum(_dot(m, y) for m, y in zip(bwd(rhs_in), x))
        assert np.isclose(lhs, rhs, rtol=1e-12)
