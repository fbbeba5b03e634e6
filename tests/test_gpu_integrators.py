"""GPU parity of the IMEXRK3 and IMEXRK222 hydrodynamic steps against the oracle.

The oracle writes the reference's ARS step literally (explicit L x_1 in the
second stage); the library folds that term into the stage-1 system
((1 - g) dt L x_1 = c1 (M x_1 - rhs_1)), so agreement to 1e-12 also checks
that identity.  History semantics follow the reference: IMEXRK3 passes every
history field through, IMEXRK222 keeps only nonlinear_old, have_old and
previous_dt.
"""

import math

import numpy as np
import pytest

from jaxfun_amd import PlaneCouetteFluctuationHIP
from jaxfun_amd.ic import seeded_state

from tests.helpers import oracle_for, to_oracle

pytestmark = pytest.mark.gpu

TOL = 1e-12


def make(N, integ, precision="fp64"):
    return PlaneCouetteFluctuationHIP(N=N, Re=400.0, dt=2e-3, perturbation_amplitude=0.05, time_integrator=integ,
                                      nonlinear_form="rotational", precision=precision)


def _state_err(s, so):
    errs = []
    for a, b in zip(list(s.u) + [s.g], list(so.u) + [so.g]):
        errs.append(float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300)))
    return max(errs)


@pytest.mark.parametrize("integ", ["IMEXRK3", "IMEXRK222"])
@pytest.mark.parametrize("N", [(17, 16, 16), (24, 16, 32)], ids=["17", "24x16x32"])
def test_steps_match_oracle(integ, N):
    solver = make(N, integ)
    oracle = oracle_for(solver)
    s = seeded_state(solver, seed=2, energy=5e-2, K=4)
    so = to_oracle(s)
    done = 0
    for k in (1, 2, 5):
        s = solver.solve(s, k - done)
        so = oracle.solve(so, k - done)
        done = k
        err = _state_err(s, so)
        print(integ, N, k, err, flush=True)
        assert err < TOL, (integ, N, k, err)
    d, do = solver.diagnostics(s), oracle.diagnostics(so)
    for key in ("Epert", "Etot", "divL2"):
        if key == "divL2":
            assert d[key] < 1e-12
        else:
            assert math.isclose(d[key], do[key], rel_tol=1e-12), (key, d[key], do[key])


def test_history_semantics():
    sb = make((17, 16, 16), "SBDF3")
    s = sb.solve(sb.initial_state(), 3)  # SBDF3 history filled
    a = make((17, 16, 16), "IMEXRK222").step(s)
    assert a.nonlinear_old is s.nonlinear_old and a.have_old == s.have_old and a.previous_dt == s.previous_dt
    assert a.nonlinear_older is None and a.solution_old is None and a.solution_older is None
    assert a.history_steps == 0.0
    b = make((17, 16, 16), "IMEXRK3").step(s)
    assert b.nonlinear_older is s.nonlinear_older and b.solution_older is s.solution_older
    assert b.history_steps == s.history_steps
    # the step itself does not read the history: same result from a fresh copy
    c = make((17, 16, 16), "IMEXRK3").step(type(s)(u=s.u, g=s.g))
    assert all(np.array_equal(x, y) for x, y in zip(list(b.u) + [b.g], list(c.u) + [c.g]))


def test_imexrk222_mixed_close_to_fp64():
    N = (33, 32, 32)
    a, b = make(N, "IMEXRK222"), make(N, "IMEXRK222", precision="mixed")
    s0 = seeded_state(a, seed=4, energy=1e-3, K=8)
    sa, sb = a.solve(s0, 20), b.solve(s0, 20)
    for x, y in zip(list(sb.u) + [sb.g], list(sa.u) + [sa.g]):
        rel = float(np.max(np.abs(x - y)) / np.max(np.abs(y)))
        print("imexrk222 mixed vs fp64", rel, flush=True)
        assert rel < 2e-6


@pytest.mark.parametrize("integ", ["IMEXRK222", "IMEXRK3", "CNAB2"])
def test_sbdf3_only_entry_points_reject_other_integrators(integ):
    solver = make((17, 16, 16), integ)
    s = solver.initial_state()
    with pytest.raises(NotImplementedError):
        solver.value_and_grad(s, 2)
    solver.load(s)
    with pytest.raises(NotImplementedError):
        solver.advance_steady(12, graph_block=12)
    with pytest.raises(NotImplementedError):
        solver.advance_steady(3)
