"""GPU parity of the 3D conducting-wall Taylor-Couette MHD step
(include/tc_dns.h with mhd = 1, ntheta > 1) against ``oracle.tc.TC3DMRIOracle``,
which is pinned by the published golden run ``taylor_couette_mhd_3d_v1`` and
the checks of ``tests/test_tc_mri_oracle.py``.  Bars as for the other
Taylor-Couette steps (FP64): fields 1e-12, total pressure 1e-10, nonlinear
histories 1e-12, relative to the largest leaf.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle.tc import CircularCouette as OBase
from oracle.tc import TC3DMRIOracle, TCState

pytestmark = pytest.mark.gpu

KEPLER = 0.5**1.5
GOLDEN_LZ = 1.0471975511965976


def _pair(Nr=12, Ntheta=6, Nz=8, dealias=1.5, ti="SBDF3", dt=2e-3, nu=2e-3, eta=3e-3, B0=0.1, Lz=None):
    from jaxfun_amd.tc import CircularCouette, TaylorCouetteMRIDNSHIP

    solver = TaylorCouetteMRIDNSHIP(CircularCouette(1.0, 2.0, 1.0, KEPLER), B0=B0, nu=nu, eta_mag=eta, Nr=Nr,
                                    Ntheta=Ntheta, Nz=Nz, Lz=Lz, dt=dt, dealias=dealias, time_integrator=ti)
    oracle = TC3DMRIOracle(OBase(1.0, 2.0, 1.0, KEPLER), B0=B0, nu=nu, eta_mag=eta, Nr=Nr, Ntheta=Ntheta, Nz=Nz,
                           Lz=Lz, dt=dt, dealias=dealias, time_integrator=ti)
    return solver, oracle


def _to_oracle(s) -> TCState:
    return TCState(u=s.x, p=s.p, nonlinear_old=s.nonlinear_old, have_old=s.have_old,
                   nonlinear_older=s.nonlinear_older, solution_old=s.solution_old,
                   solution_older=s.solution_older, history_steps=s.history_steps)


def _rel(a_list, b_list):
    num = max(float(np.max(np.abs(a - b))) for a, b in zip(a_list, b_list))
    den = max(float(np.max(np.abs(b))) for b in b_list)
    return num / den if den > 0 else num


def _state(solver, amp=3e-2):
    """The m = 1 MRI eigenmode at a nonlinear amplitude plus m = 0 and m = 2 content."""
    s, _ = solver.seed_linear_eigenmode(m=1, kz_mode=1, amp=amp)
    kz = 2 * np.pi / solver.Lz
    arg = np.pi * (solver.R - 1.0)
    zero = 0 * solver.R
    extra = solver.state_from_physical((zero, amp * np.sin(arg) * np.cos(2 * solver.Theta + kz * solver.Z), zero,
                                        zero, amp * solver.R * np.sin(kz * solver.Z),
                                        amp * np.cos(arg) * np.cos(2 * solver.Theta)))
    return type(s)(x=tuple(a + b for a, b in zip(s.x, extra.x)), p=s.p, nonlinear_old=s.nonlinear_old)


@pytest.mark.parametrize("ti", ["SBDF3", "CNAB2", "IMEXRK3"])
@pytest.mark.parametrize("dealias", [1.5, 1.0])
def test_3d_mri_steps_match_oracle(ti, dealias):
    solver, oracle = _pair(ti=ti, dealias=dealias)
    s0 = _state(solver)
    print("case", ti, dealias, flush=True)
    for steps in (1, 2, 5):
        got = solver.solve(s0, steps)
        ref = oracle.solve(_to_oracle(s0), steps)
        ex, ep = _rel(got.x, ref.u), _rel([got.p], [ref.p])
        en = _rel(got.nonlinear_old, ref.nonlinear_old) if ref.have_old else 0.0
        print(f"x {ex:.2e} p {ep:.2e} nl {en:.2e}", flush=True)
        assert ex <= 1e-12 and ep <= 1e-10 and en <= 1e-12, (ex, ep, en)


def test_3d_mri_golden_sbdf3_run():
    """production/goldens/taylor_couette_mhd_3d_v1: Nr = 12, Ntheta = 4, Nz = 4,
    dealias 3/2, SBDF3, dt = 2e-3, three steps from the amp = 1e-7 m = 1 MRI
    eigenmode (divergences: grid maxima that depend on the eigenvector's
    phase, compared in size)."""
    solver, _ = _pair(Nr=12, Ntheta=4, Nz=4, dealias=1.5, ti="SBDF3", dt=2e-3, nu=1e-3, eta=1e-3, Lz=GOLDEN_LZ)
    s0, eig = solver.seed_linear_eigenmode(m=1, kz_mode=1, amp=1e-7)
    assert abs(eig.real - 0.2912937448361517) <= 1e-10
    d0 = solver.diagnostics(s0)
    d1 = solver.diagnostics(solver.solve(s0, 3))
    growth = 0.5 * math.log(d1["E"] / d0["E"]) / 6e-3
    print("growth", growth, d0, d1, flush=True)
    assert abs(d0["Ekin"] / 7.85609816898076e-15 - 1.0) <= 1e-9
    assert abs(d0["Emag"] / 1.1298011285580755e-14 - 1.0) <= 1e-9
    assert abs(d1["Ekin"] / 7.883607407650011e-15 - 1.0) <= 1e-9
    assert abs(d1["Emag"] / 1.1337572869748595e-14 - 1.0) <= 1e-9
    assert abs(growth - 0.29129374483210074) <= 1e-7
    assert abs(d1["divu"] / 1.486388530128091e-08 - 1.0) < 0.05
    assert abs(d1["divb"] / 1.0541502761126154e-08 - 1.0) < 0.05


def test_3d_mri_diagnostics_and_graph_replay():
    solver, oracle = _pair(ti="SBDF3")
    s = solver.solve(_state(solver), 4)
    d = solver.diagnostics(s)
    ref = oracle.diagnostics(_to_oracle(s))
    for key in ("Ekin", "Emag", "E"):
        assert abs(d[key] - ref[key]) <= 1e-12 * abs(ref[key]), (key, d[key], ref[key])
    for key in ("divu", "divb"):
        assert abs(d[key] - ref[key]) <= 1e-12 * max(ref[key], math.sqrt(ref["E"])), (key, d[key], ref[key])
    from jaxfun_amd import _native as nat

    solver.load(s)
    nat.tc_check(solver._lib.tc_step(solver._h, 30))
    eager = solver.fetch(s)
    graph = solver.solve(s, 30)
    for a, b in zip(eager.x + (eager.p,), graph.x + (graph.p,)):
        assert np.array_equal(a, b)


def test_3d_mri_reference_cases():
    """tests/couette/test_taylor_couette_dns_jax.py:334-355,386-405."""
    from jaxfun_amd.tc import CircularCouette, TaylorCouetteMRIDNSHIP

    solver = TaylorCouetteMRIDNSHIP(CircularCouette(1.0, 2.0, 1.0, KEPLER), B0=0.1, nu=1e-3, eta_mag=1e-3, Nr=8,
                                    Ntheta=4, Nz=6, dt=1e-3, dealias=1.0)
    z = solver.step(solver.zero_state())
    d = solver.diagnostics(z)
    assert all(np.max(np.abs(c)) == 0.0 for c in z.x) and np.max(np.abs(z.p)) == 0.0
    assert d["E"] == 0.0 and d["divu"] < 1e-12 and d["divb"] < 1e-12 and d["wall_u"] < 1e-18
    s, eig = solver.seed_linear_eigenmode(m=1, kz_mode=1, amp=1e-8)
    rate, out = solver.growth_rate(s, steps=100)
    d = solver.diagnostics(out)
    assert abs(rate - eig.real) <= 1e-6 * max(1.0, abs(eig.real))
    assert d["divu"] < 1e-7 and d["divb"] < 1e-7
    assert abs(out.p[0, 0, 0]) < 1e-18


def test_3d_mri_production_resolution_matches_oracle():
    """The production tier of taylor_couette_mhd_3d_v1 (Nr = 24, Ntheta = 16,
    Nz = 16, dealias 3/2): the 3D MHD product kernel needs about 125 KB of LDS
    per block here, beyond the 64 KB default."""
    solver, oracle = _pair(Nr=24, Ntheta=16, Nz=16, dealias=1.5, ti="SBDF3", nu=1e-3, eta=1e-3)
    s0 = _state(solver, amp=1e-2)
    for steps in (1, 3):
        got = solver.solve(s0, steps)
        ref = oracle.solve(_to_oracle(s0), steps)
        ex, ep = _rel(got.x, ref.u), _rel([got.p], [ref.p])
        print(f"x {ex:.2e} p {ep:.2e}", flush=True)
        assert ex <= 1e-12 and ep <= 1e-10, (ex, ep)
