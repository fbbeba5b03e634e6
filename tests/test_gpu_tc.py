"""GPU parity of the axisymmetric Taylor-Couette step (include/tc_dns.h) against
the oracle restatement ``oracle/tc.py``, which is pinned by the published
golden run and the checks of ``tests/test_tc_oracle.py``.

Bars (FP64): velocity leaves within 1e-12 of the oracle per leaf group, the
pressure within 1e-10 (the saddle-point blocks have condition numbers near
5e5; the library applies their long-double inverses, the oracle pivoted LU),
the nonlinear histories within 1e-12; the golden CNAB2 run reproduces the
published linear and DNS growth rates.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle.tc import CircularCouette as OBase
from oracle.tc import TCOracle, TCState

pytestmark = pytest.mark.gpu

GOLDEN_LZ = 2.0074074463832545


def _pair(Nr=24, Nz=8, dealias=1.5, ti="SBDF3", dt=2e-3, nu=1e-2, Lz=None, Omega2=0.0):
    from jaxfun_amd.tc import AxisymmetricTCDNSHIP, CircularCouette

    solver = AxisymmetricTCDNSHIP(CircularCouette(1.0, 2.0, 1.0, Omega2), nu=nu, Nr=Nr, Nz=Nz, Lz=Lz, dt=dt,
                                  dealias=dealias, time_integrator=ti)
    oracle = TCOracle(OBase(1.0, 2.0, 1.0, Omega2), nu=nu, Nr=Nr, Nz=Nz, Lz=Lz, dt=dt, dealias=dealias,
                      time_integrator=ti)
    return solver, oracle


def _to_oracle(s) -> TCState:
    return TCState(u=s.u, p=s.p, nonlinear_old=s.nonlinear_old, have_old=s.have_old,
                   nonlinear_older=s.nonlinear_older, solution_old=s.solution_old,
                   solution_older=s.solution_older, history_steps=s.history_steps)


def _rel(a_list, b_list, nyq):
    num = max(float(np.max(np.abs(np.delete(a, nyq, axis=0) - np.delete(b, nyq, axis=0)))) for a, b in
              zip(a_list, b_list))
    den = max(float(np.max(np.abs(b))) for b in b_list)
    return num / den if den > 0 else num


def _compare(solver, got, ref, tol_u=1e-12, tol_p=1e-10, tol_n=1e-12):
    nyq = solver.Nz // 2
    eu = _rel(got.u, ref.u, nyq)
    ep = _rel([got.p], [ref.p], nyq)
    en = _rel(got.nonlinear_old, ref.nonlinear_old, nyq) if ref.have_old else 0.0
    print(f"u {eu:.2e} p {ep:.2e} nl {en:.2e}", flush=True)
    assert eu <= tol_u and ep <= tol_p and en <= tol_n, (eu, ep, en)


@pytest.mark.parametrize("ti", ["SBDF3", "CNAB2", "IMEXRK3"])
@pytest.mark.parametrize("dealias", [1.5, 1.0])
def test_steps_match_oracle(ti, dealias):
    solver, oracle = _pair(ti=ti, dealias=dealias)
    s0 = solver.initial_state(amp=5e-2, kz_mode=1)
    print("case", ti, dealias, flush=True)
    for steps in (1, 2, 5):
        got = solver.solve(s0, steps)
        ref = oracle.solve(_to_oracle(s0), steps)
        _compare(solver, got, ref)
        if ti == "SBDF3":
            assert got.history_steps == ref.history_steps
            for g_grp, r_grp in ((got.solution_old, ref.solution_old), (got.nonlinear_older, ref.nonlinear_older)):
                assert _rel(g_grp, r_grp, solver.Nz // 2) <= 1e-12


def test_graph_replay_equals_eager_steps():
    solver, _ = _pair(ti="SBDF3", Nz=16)
    s0 = solver.initial_state(amp=5e-2, kz_mode=1)
    s = solver.solve(s0, 2)  # history bootstrapped
    from jaxfun_amd import _native as nat

    solver.load(s)
    nat.tc_check(solver._lib.tc_step(solver._h, 30))
    eager = solver.fetch(s)
    graph = solver.solve(s, 30)  # 12-step graphs + eager remainder
    for a, b in zip(eager.u + (eager.p,), graph.u + (graph.p,)):
        assert np.array_equal(a, b)


def test_diagnostics_match_oracle():
    solver, oracle = _pair(ti="SBDF3")
    s0 = solver.initial_state(amp=5e-2, kz_mode=1)
    s = solver.solve(s0, 4)
    d = solver.diagnostics(s)
    ref = oracle.diagnostics(_to_oracle(s))
    for key in ("E", "Eth"):
        assert abs(d[key] - ref[key]) <= 1e-12 * abs(ref[key]), (key, d[key], ref[key])
    assert abs(d["div_linf"] - ref["div_linf"]) <= 1e-12 * max(1.0, ref["div_linf"])
    assert d["wall"] == ref["wall"] == 0.0
    # after a step the weak divergence is satisfied to rounding (a residual of
    # cancelling terms): compare it on a field that is not solenoidal
    assert d["continuity_l2"] < 1e-9 and ref["continuity_l2"] < 1e-9
    arg = np.pi * (solver.R - 1.0)
    kz = 2 * np.pi / solver.Lz
    bad = solver.state_from_physical((1e-2 * np.sin(arg) ** 2 * np.cos(kz * solver.Z), 0 * solver.R,
                                      1e-2 * np.sin(arg) * np.sin(2 * kz * solver.Z)))
    d0, ref0 = solver.diagnostics(bad), oracle.diagnostics(_to_oracle(bad))
    assert ref0["continuity_l2"] > 1e-4
    for key in ("continuity_l2", "div_linf", "E"):
        assert abs(d0[key] - ref0[key]) <= 1e-12 * ref0[key], (key, d0[key], ref0[key])


def test_golden_cnab2_growth_rate():
    """production/goldens/taylor_couette_hydro_dns_v1: Re = 100, Nr = 40, Nz = 8,
    CNAB2, dt = 1e-3, 400 steps from the amp = 1e-6 linear eigenmode."""
    solver, _ = _pair(Nr=40, Nz=8, dealias=1.0, ti="CNAB2", dt=1e-3, Lz=GOLDEN_LZ)
    kz_mode = round(3.13 * solver.Lz / (2 * math.pi))
    s0, eig = solver.seed_linear_eigenmode(kz_mode=kz_mode, amp=1e-6)
    assert abs(eig.real - 0.12730355815948557) <= 1e-10
    e0 = solver.energy(s0)
    s1 = solver.solve(s0, 400)
    d1 = solver.diagnostics(s1)
    growth = 0.5 * math.log(d1["E"] / e0) / 0.4
    print("growth", growth, "E0", e0, "E1", d1["E"], flush=True)
    assert abs(growth - 0.1273035583309162) <= 1e-9
    # the reference port writes 0.5 of the eigenvector at +-k; the published
    # energies carry the full vector once, so they are 4x these
    assert abs(4.0 * e0 / 4.951524923572814e-12 - 1.0) <= 1e-9
    assert abs(4.0 * d1["E"] / 5.4823752189594825e-12 - 1.0) <= 1e-9
    assert d1["div_linf"] < 1e-9


def test_fixed_dt_and_integrator_rules():
    solver, _ = _pair(ti="SBDF3")
    with pytest.raises(NotImplementedError):
        solver.set_dt(1e-3)
    solver.set_dt(2e-3)  # unchanged dt is accepted
    cn, oracle = _pair(ti="CNAB2")
    s0 = cn.initial_state(amp=5e-2)
    cn.set_dt(1e-3)
    oracle.dt = 1e-3
    oracle._build_time_operators()
    _compare(cn, cn.solve(s0, 3), oracle.solve(_to_oracle(s0), 3))


def test_runner_scalars_match_oracle():
    """inner_torque and radial_velocity_linf against the oracle's physical
    velocity and a direct evaluation of d(u_theta)/dr at the inner wall."""
    from numpy.polynomial import chebyshev as npcheb

    solver, oracle = _pair(ti="SBDF3")
    s = solver.solve(solver.initial_state(amp=5e-2), 3)
    ref_ur = np.max(np.abs(oracle.velocity_physical(_to_oracle(s))[0]))
    assert abs(solver.radial_velocity_linf(s) - ref_ur) <= 1e-14 * ref_ur
    df = 2.0
    coef = np.zeros(solver.Nr, dtype=complex)
    coef[: solver.n] += s.u[1][0]
    coef[2:] -= s.u[1][0]
    dut = df * npcheb.chebval(-1.0, npcheb.chebder(coef))
    b = solver.base.b
    ref = 2 * np.pi * solver.nu * abs(-2 * b + dut)
    assert abs(solver.inner_torque(s) - ref) <= 1e-13 * ref


def test_cnab2_graph_replay_equals_eager_steps():
    solver, _ = _pair(ti="CNAB2", Nz=16)
    s = solver.solve(solver.initial_state(amp=5e-2, kz_mode=1), 1)  # have_old = 1
    from jaxfun_amd import _native as nat

    solver.load(s)
    nat.tc_check(solver._lib.tc_step(solver._h, 29))
    eager = solver.fetch(s)
    graph = solver.solve(s, 29)
    for a, b in zip(eager.u + (eager.p,) + eager.nonlinear_old, graph.u + (graph.p,) + graph.nonlinear_old):
        assert np.array_equal(a, b)


def test_supercritical_saturation_against_float32_golden():
    """production/goldens/tc_supercritical_saturation: Re = 100, Nr = 64,
    Nz = 32, dealias 1.5, CNAB2, dt = 0.004, 20000 steps (t = 80) from the
    amp = 1e-4 leading eigenmode into the saturated Taylor-vortex state.

    That golden is a float32 GPU run of an earlier reference commit (its
    metadata: production_run_dtype float32); its linear growth rate,
    0.1273169, is 1.3e-5 away from the FP64 value that this build, the
    oracle and the shenfun golden ``taylor_couette_hydro_dns_v1`` agree on to
    1e-13, and its seeded energy is 12.6 % above the one its own seed rule
    gives in FP64, so it pins nothing to FP64 accuracy.  What it does fix is
    the saturated flow's torque on the inner cylinder and its largest radial
    velocity, which this FP64 run reproduces to 2e-4 and 1e-2; the saturation
    itself is checked by the golden's criterion (relative energy change over
    the last 4 time units below 1e-2)."""
    from jaxfun_amd.tc import AxisymmetricTCDNSHIP, CircularCouette

    solver = AxisymmetricTCDNSHIP(CircularCouette(1.0, 2.0, 1.0, 0.0), nu=0.01, Nr=64, Nz=32, Lz=2.007407446944818,
                                  dt=0.004, dealias=1.5, time_integrator="CNAB2")
    s0, _ = solver.seed_linear_eigenmode(kz_mode=1, amp=1e-4)
    s_a = solver.solve(s0, 19000)
    s = solver.solve(s_a, 1000)
    e_a, d = solver.energy(s_a), solver.diagnostics(s)
    torque, urmax = solver.inner_torque(s), solver.radial_velocity_linf(s)
    print("saturated", d["E"], torque, urmax, d["continuity_l2"], flush=True)
    assert abs(d["E"] - e_a) <= 1e-2 * d["E"]
    assert abs(torque / 0.21806299686431885 - 1.0) <= 2e-4
    assert abs(urmax / 0.15969887375831604 - 1.0) <= 1e-2
    assert d["continuity_l2"] < 1e-6 and d["div_linf"] < 1e-3
