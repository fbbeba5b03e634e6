"""Pins of the oracle's hand-written MHD adjoint (oracle/adjoint_mhd.py), CPU only.

* the adjoint identity <A v, w> = <v, A^H w> for each linear stage of the
  vector-potential path (B = curl A and J = curl B in coefficient space with
  their TD projections, the padded and unpadded backward transforms of B and
  J, the A mass, Laplacian and solve), so a dropped term or a wrong sign in one
  of them fails on its own;
* the vector-Jacobian product of the coupled nonlinear rows (flow rows and
  induction rows) against central differences of the oracle's forward, with
  rotation, shear and a background field, so every cross product, the
  Coriolis term and the EMF are checked;
* the gradient of Epert + Emag after several steps (IMEXRK3 start-up and
  steady SBDF3 steps, from a fresh state and from one with a full history)
  against central differences of the oracle forward along seeded directions in
  every leaf, as the reference checks its gradients
  (``tests/couette/test_differentiability_jax.py:81-136``).
"""

from dataclasses import replace

import numpy as np
import pytest

from oracle.adjoint_mhd import MHDOracleAdjoint, mhd_pairing
from oracle.kmm import OracleState
from oracle.mhd import MHDOracleState, RotatingShearMHDOracle

N = (9, 8, 8)


def _oracle(**kw):
    args = dict(N=N, Re=150.0, Rm=120.0, omega=0.7, shear_rate=1.3, background_b=(0.05, -0.04, 0.2), dt=3e-3,
                perturbation_amplitude=0.3, magnetic_amplitude=0.3)
    args.update(kw)
    return RotatingShearMHDOracle(**args)


def _phys(o, rng):
    return rng.standard_normal((o.N0, o.N[1], o.N[2])) + 0j


def _td(o, rng):
    return o.mask(o.tp_forward(_phys(o, rng), o.SD, o.MD1))


def _dir(o, rng, history=False):
    """A Hermitian direction in every leaf (flow, A and, if asked, the histories)."""
    f = o.state_from_physical(tuple(_phys(o, rng) for _ in range(3)))
    A = tuple(_td(o, rng) for _ in range(3))

    def grp():
        return (o.mask(o.tp_forward(_phys(o, rng), o.SB, o.MB1)), _td(o, rng), rng.standard_normal(o.ND),
                rng.standard_normal(o.ND))

    z = tuple(np.zeros_like(a) for a in A)
    if history:
        f = replace(f, nonlinear_old=grp(), nonlinear_older=grp(), solution_old=grp(), solution_older=grp())
        return MHDOracleState(flow=f, A=A, nonlinear_A_old=tuple(_td(o, rng) for _ in range(3)),
                              nonlinear_A_older=tuple(_td(o, rng) for _ in range(3)),
                              A_old=tuple(_td(o, rng) for _ in range(3)),
                              A_older=tuple(_td(o, rng) for _ in range(3)))
    return MHDOracleState(flow=f, A=A, nonlinear_A_old=z, nonlinear_A_older=z, A_old=z, A_older=z)


def _axpy(s, d, eps):
    t = lambda a, b: tuple(x + eps * y for x, y in zip(a, b))  # noqa: E731
    f, df = s.flow, d.flow
    flow = replace(f, u=t(f.u, df.u), g=f.g + eps * df.g, nonlinear_old=t(f.nonlinear_old, df.nonlinear_old),
                   nonlinear_older=t(f.nonlinear_older, df.nonlinear_older),
                   solution_old=t(f.solution_old, df.solution_old),
                   solution_older=t(f.solution_older, df.solution_older))
    return MHDOracleState(flow=flow, A=t(s.A, d.A), nonlinear_A_old=t(s.nonlinear_A_old, d.nonlinear_A_old),
                          nonlinear_A_older=t(s.nonlinear_A_older, d.nonlinear_A_older),
                          A_old=t(s.A_old, d.A_old), A_older=t(s.A_older, d.A_older))


def _dot(a, b):
    return float(sum(np.real(np.sum(np.conj(x) * y)) for x, y in zip(a, b)))


def test_linear_stage_adjoints():
    o = _oracle()
    adj = MHDOracleAdjoint(o)
    rng = np.random.default_rng(1)
    A = tuple(_td(o, rng) for _ in range(3))
    mB = (_td(o, rng), o.mask(o.tc_forward(_phys(o, rng).real)), o.mask(o.tc_forward(_phys(o, rng).real)))
    assert np.isclose(_dot(o.B_from_A(A), mB), _dot(A, adj.B_from_A_adj(mB)), rtol=1e-12)
    B = o.B_from_A(A)
    mJ = (o.mask(o.tc_forward(_phys(o, rng).real)), _td(o, rng), _td(o, rng))
    assert np.isclose(_dot(o.J_from_B(B), mJ), _dot(B, adj.J_from_B_adj(mJ)), rtol=1e-12)
    for padded in (True, False):
        shp = (o.N0,) + ((o.pad[1], o.pad[2]) if padded else (o.N[1], o.N[2]))
        w = tuple(rng.standard_normal(shp) + 0j for _ in range(3))
        assert np.isclose(_dot(o.backward_B(B, padded), w), _dot(B, adj.backward_B_adj(w, padded)), rtol=1e-12)
        J = o.J_from_B(B)
        assert np.isclose(_dot(o.backward_J(J, padded), w), _dot(J, adj.backward_J_adj(w, padded)), rtol=1e-12)
    r = tuple(_td(o, rng) for _ in range(3))
    assert np.isclose(_dot(o.A_mass(A), r), _dot(A, adj.A_mass_adj(r)), rtol=1e-12)
    assert np.isclose(_dot(o.A_lap(A), r), _dot(A, adj.A_lap_adj(r)), rtol=1e-12)
    g = 0.37 * o.dt
    assert np.isclose(_dot(o.A_solve_gamma(g, r), A), _dot(r, adj.A_solve_adj(g, A)), rtol=1e-11)


def test_nonlinear_vjp_matches_central_differences():
    """<d(N, NA), (mu_n, mu_nA)> from central differences of the forward rows
    equals <(du, dA), VJP(mu_n, mu_nA)>, with Omega, S and B0 all nonzero."""
    o = _oracle()
    adj = MHDOracleAdjoint(o)
    rng = np.random.default_rng(2)
    s = o.solve(o.initial_state(), 2)
    d = _dir(o, rng)
    mu_n = (o.mask(o.tp_forward(_phys(o, rng), o.SB, o.MB1)), _td(o, rng), rng.standard_normal(o.ND),
            rng.standard_normal(o.ND))
    mu_nA = tuple(_td(o, rng) for _ in range(3))

    def rows(st):
        H, HA = o.mhd_convection(st)
        return o.nonlinear_rhs(H), o.nonlinear_A(HA)

    eps = 1e-6
    sp = MHDOracleState(flow=OracleState(u=tuple(a + eps * b for a, b in zip(s.flow.u, d.flow.u)), g=s.flow.g),
                        A=tuple(a + eps * b for a, b in zip(s.A, d.A)))
    sm = MHDOracleState(flow=OracleState(u=tuple(a - eps * b for a, b in zip(s.flow.u, d.flow.u)), g=s.flow.g),
                        A=tuple(a - eps * b for a, b in zip(s.A, d.A)))
    (np_, na_p), (nm_, na_m) = rows(sp), rows(sm)
    fd = (_dot(mu_n, np_) - _dot(mu_n, nm_) + _dot(mu_nA, na_p) - _dot(mu_nA, na_m)) / (2 * eps)
    mu_u, mu_A = adj.mhd_nonlinear_adj(s, mu_n, mu_nA)
    an = _dot(mu_u, d.flow.u) + _dot(mu_A, d.A)
    assert abs(fd - an) <= 1e-8 * abs(an), (fd, an)


@pytest.mark.parametrize("steps,history", [(2, False), (4, False), (3, True)])
def test_gradient_matches_central_differences(steps, history):
    o = _oracle()
    adj = MHDOracleAdjoint(o)
    rng = np.random.default_rng(10 + steps)
    s0 = o.initial_state()
    if history:
        s0 = o.solve(s0, 3)  # history_steps = 2: only steady steps follow
    value, mu = adj.mhd_value_and_grad(s0, steps)
    assert value == pytest.approx(adj.mhd_energy(o.solve(s0, steps)), rel=1e-14)

    def cd(d, eps):
        return (adj.mhd_energy(o.solve(_axpy(s0, d, eps), steps))
                - adj.mhd_energy(o.solve(_axpy(s0, d, -eps), steps))) / (2 * eps)

    for _ in range(3):
        d = _dir(o, rng, history=history)
        # Richardson extrapolation of two central differences cancels their eps^2 term
        # (their difference at eps = 8e-5 and 4e-5 is rounding, about 1e-10 of the value)
        eps = 8e-5
        fd = (4.0 * cd(d, eps / 2) - cd(d, eps)) / 3.0
        an = mhd_pairing(mu, d)
        print(steps, history, fd, an, abs(fd - an) / abs(an))
        assert abs(fd - an) <= 2e-9 * abs(an), (fd, an)


def test_planted_sign_slip_in_the_induction_adjoint_fails():
    """A copy of the adjoint with the sign of the A mass adjoint flipped (the
    map that carries the induction row's cotangent back to A) misses the
    central differences: the check sees the induction path."""
    o = _oracle()
    rng = np.random.default_rng(7)

    class Flipped(MHDOracleAdjoint):
        def A_mass_adj(self, r):
            return tuple(-x for x in super().A_mass_adj(r))

    good, bad = MHDOracleAdjoint(o), Flipped(o)
    s0 = o.initial_state()
    d = _dir(o, rng)
    eps = 5e-6
    fd = (good.mhd_energy(o.solve(_axpy(s0, d, eps), 3)) - good.mhd_energy(o.solve(_axpy(s0, d, -eps), 3))) / (2 * eps)
    assert abs(mhd_pairing(good.mhd_value_and_grad(s0, 3)[1], d) - fd) <= 1e-8 * abs(fd)
    assert abs(mhd_pairing(bad.mhd_value_and_grad(s0, 3)[1], d) - fd) > 1e-4 * abs(fd)
