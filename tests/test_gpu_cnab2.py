"""GPU parity of the CNAB2 steps (hydrodynamic and the three MHD classes)
against the oracle, with constant and changing dt, and of set_dt.

CNAB2 is Crank-Nicolson for the linear terms and variable-step
Adams-Bashforth-2 for the nonlinear ones; the state it returns keeps the
step's nonlinear rows as the only history level, with have_old = 1 and
previous_dt = dt.  Bars: state fields within 1e-12 relative of the FP64
oracle, the nonlinear history rows within 5e-11 (their rounding sensitivity,
the same bar as the SBDF3 history rows).
"""

import math

import numpy as np
import pytest

from jaxfun_amd import PlaneCouetteFluctuationHIP
from jaxfun_amd.ic import seeded_state

from tests.helpers import oracle_for, to_oracle

pytestmark = pytest.mark.gpu

TOL = 1e-12
HIST_TOL = 5e-11


def make(N, integ="CNAB2", precision="fp64", dt=2e-3):
    return PlaneCouetteFluctuationHIP(N=N, Re=400.0, dt=dt, perturbation_amplitude=0.05, time_integrator=integ,
                                      nonlinear_form="rotational", precision=precision)


def _group_err(xs, ys):
    """Largest per-array error over a group, each scaled by its own maximum
    (floored at 1e-3 of the group's largest entry)."""
    gmax = max(float(np.max(np.abs(np.asarray(y)))) for y in ys)
    out = 0.0
    for x, y in zip(xs, ys):
        x, y = np.asarray(x), np.asarray(y)
        scale = max(float(np.max(np.abs(y))), 1e-3 * gmax, 1e-300)
        out = max(out, float(np.max(np.abs(x - y))) / scale)
    return out


def assert_flow_parity(s, so, where):
    e_state = _group_err(list(s.u) + [s.g], list(so.u) + [so.g])
    e_hist = _group_err(s.nonlinear_old, so.nonlinear_old)
    print(where, e_state, e_hist, flush=True)
    assert e_state < TOL, (where, e_state)
    assert e_hist < HIST_TOL, (where, e_hist)
    assert s.have_old == so.have_old == 1.0
    assert math.isclose(s.previous_dt, so.previous_dt, rel_tol=0, abs_tol=0)


@pytest.mark.parametrize("N", [(17, 16, 16), (24, 16, 32)], ids=["17", "24x16x32"])
def test_cnab2_steps_match_oracle(N):
    solver = make(N)
    oracle = oracle_for(solver)
    s = seeded_state(solver, seed=2, energy=5e-2, K=4)  # no history: the bootstrap step
    so = to_oracle(s)
    done = 0
    for k in (1, 2, 5):
        s = solver.solve(s, k - done)
        so = oracle.solve(so, k - done)
        done = k
        assert_flow_parity(s, so, (N, k))
    # the reference's CNAB2 state: one nonlinear level, nothing else
    assert s.nonlinear_older is None and s.solution_old is None and s.solution_older is None
    assert s.history_steps == 0.0
    d, do = solver.diagnostics(s), oracle.diagnostics(so)
    assert math.isclose(d["Epert"], do["Epert"], rel_tol=1e-12)
    assert d["divL2"] < 1e-12


def test_cnab2_variable_dt_matches_oracle():
    """set_dt between steps and a per-call dt (step(state, dt)): the AB2 ratio
    dt / previous_dt and the Crank-Nicolson systems follow each step's dt."""
    solver = make((17, 16, 16))
    oracle = oracle_for(solver)
    s = seeded_state(solver, seed=5, energy=5e-2, K=4)
    so = to_oracle(s)
    for k, h in enumerate((2e-3, 1e-3, 3e-3, 2.5e-3, 2.5e-3)):
        solver.set_dt(h)
        oracle.set_dt(h)
        s, so = solver.step(s), oracle.step(so)
        assert_flow_parity(s, so, ("set_dt", k, h))
    assert solver.dt == 2.5e-3
    s = solver.step(s, dt=1.7e-3)  # this step only
    assert solver.dt == 2.5e-3
    oracle.set_dt(1.7e-3)
    so = oracle.step(so)
    assert_flow_parity(s, so, ("step dt", 1.7e-3))
    oracle.set_dt(2.5e-3)
    s, so = solver.step(s), oracle.step(so)
    assert_flow_parity(s, so, ("after step dt", 2.5e-3))


def test_cnab2_continues_from_sbdf3_history():
    """A state with a nonlinear level from another integrator (have_old = 1):
    the first CNAB2 step extrapolates from it at once."""
    sb = make((17, 16, 16), "SBDF3")
    s0 = sb.solve(sb.initial_state(), 3)
    c = make((17, 16, 16))
    oracle = oracle_for(c)
    s, so = c.step(s0), oracle.step(to_oracle(s0))
    assert_flow_parity(s, so, "from sbdf3")
    # the same state without its history bootstraps instead (have_old = 0)
    bare = type(s0)(u=s0.u, g=s0.g)
    s1 = c.step(bare)
    so1 = oracle.step(to_oracle(c._device_history(bare)))
    assert_flow_parity(s1, so1, "bare")
    assert _group_err(list(s1.u), list(s.u)) > 1e-8  # the extrapolation mattered


def test_set_dt_per_integrator():
    sb = make((17, 16, 16), "SBDF3")
    sb.set_dt(2e-3)  # unchanged dt is accepted
    with pytest.raises(NotImplementedError):
        sb.set_dt(1e-3)
    with pytest.raises(NotImplementedError):
        sb.step(sb.initial_state(), dt=1e-3)
    # IMEXRK222 refactors its stage systems at the new dt
    a = make((17, 16, 16), "IMEXRK222")
    oracle = oracle_for(a)
    s = seeded_state(a, seed=2, energy=5e-2, K=4)
    so = to_oracle(s)
    s, so = a.solve(s, 2), oracle.solve(so, 2)
    a.set_dt(1.3e-3)
    oracle.set_dt(1.3e-3)
    s, so = a.solve(s, 3), oracle.solve(so, 3)
    e = _group_err(list(s.u) + [s.g], list(so.u) + [so.g])
    print("imexrk222 set_dt", e, flush=True)
    assert e < TOL
    with pytest.raises(ValueError):
        a.set_dt(-1.0)


def test_cnab2_mixed_close_to_fp64():
    N = (33, 32, 32)
    a, b = make(N), make(N, precision="mixed")
    s0 = seeded_state(a, seed=4, energy=1e-3, K=8)
    sa, sb = a.solve(s0, 20), b.solve(s0, 20)
    for x, y in zip(list(sb.u) + [sb.g], list(sa.u) + [sa.g]):
        rel = float(np.max(np.abs(x - y)) / np.max(np.abs(y)))
        print("cnab2 mixed vs fp64", rel, flush=True)
        assert rel < 2e-6


# ---------------------------------------------------------------- MHD
@pytest.mark.parametrize("kind", ["conducting", "shear", "insulating"])
def test_cnab2_mhd_steps_match_oracle(kind):
    """The CNAB2 step of the three MHD classes (flow and vector potential on
    the same Crank-Nicolson / AB2 update, _step_cnab2_mhd) with a dt change."""
    from jaxfun_amd.ic import seeded_mhd_state

    from tests.helpers import insulating_oracle_for, mhd_oracle_for, shear_oracle_for, to_mhd_oracle
    from tests.test_gpu_mhd import make as make_cond
    from tests.test_gpu_mhd import make_insulating, make_shear

    if kind == "conducting":
        solver = make_cond((17, 16, 16), magnetic_amplitude=0.1, background_b=(0.0, 0.0, 0.1),
                           time_integrator="CNAB2")
        oracle = mhd_oracle_for(solver)
        s = solver.initial_state()
    elif kind == "shear":
        solver = make_shear((17, 16, 16), time_integrator="CNAB2")
        oracle = shear_oracle_for(solver)
        s = solver.initial_state()
    else:
        solver = make_insulating((17, 16, 16), time_integrator="CNAB2")
        oracle = insulating_oracle_for(solver)
        s = seeded_mhd_state(solver, seed=3, energy_u=1e-3, energy_b=1e-3, K=4)
    so = to_mhd_oracle(s)
    dt0 = solver.dt
    for k, h in enumerate((dt0, dt0, 0.5 * dt0, 0.5 * dt0, dt0)):
        solver.set_dt(h)
        oracle.set_dt(h)
        s, so = solver.step(s), oracle.step(so)
        assert_flow_parity(s.flow, so.flow, (kind, k))
        e_a = _group_err(s.A, so.A)
        e_na = _group_err(s.nonlinear_A_old, so.nonlinear_A_old)
        print(kind, k, "A", e_a, e_na, flush=True)
        assert e_a < TOL and e_na < HIST_TOL, (kind, k, e_a, e_na)
    assert s.A_old is None and s.nonlinear_A_older is None
    d, do = solver.diagnostics(s), oracle.diagnostics(so)
    for key in ("Epert", "Emag"):
        assert math.isclose(d[key], do[key], rel_tol=1e-12), (key, d[key], do[key])
    assert d["divB_L2"] < 1e-12


def test_imexrk222_mhd_set_dt_matches_oracle():
    """IMEXRK222 under MHD after set_dt: the flow and vector-potential stage
    systems are refactored in place at the new dt."""
    from jaxfun_amd.ic import seeded_mhd_state

    from tests.helpers import insulating_oracle_for, to_mhd_oracle
    from tests.test_gpu_mhd import make_insulating

    solver = make_insulating((17, 16, 16), time_integrator="IMEXRK222")
    oracle = insulating_oracle_for(solver)
    s = seeded_mhd_state(solver, seed=3, energy_u=1e-3, energy_b=1e-3, K=4)
    so = to_mhd_oracle(s)
    s, so = solver.solve(s, 2), oracle.solve(so, 2)
    h = 0.6 * solver.dt
    solver.set_dt(h)
    oracle.set_dt(h)
    s, so = solver.solve(s, 3), oracle.solve(so, 3)
    e_u = _group_err(list(s.flow.u) + [s.flow.g], list(so.flow.u) + [so.flow.g])
    e_a = _group_err(s.A, so.A)
    print("imexrk222 mhd set_dt", e_u, e_a, flush=True)
    assert e_u < TOL and e_a < TOL
