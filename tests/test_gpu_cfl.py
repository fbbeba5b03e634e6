"""GPU parity of the device field maxima and the CFL numbers built on them
(production/health.py:120-193) against the oracle's unpadded backward
transforms, for the hydrodynamic class and the three MHD classes.

Bars: FP64 maxima within 1e-12 relative of the oracle (a maximum of values
that agree to rounding); mixed precision within 1e-5 (FP32 transforms).
"""

import math

import pytest

from jaxfun_amd import PlaneCouetteFluctuationHIP
from jaxfun_amd.ic import seeded_state

from tests.helpers import oracle_for, to_oracle

pytestmark = pytest.mark.gpu


def _check(solver, cfl, cfl_o, tol):
    assert set(cfl) == set(cfl_o)
    for k, v in cfl_o.items():
        print(k, cfl[k], v, flush=True)
        assert math.isclose(cfl[k], v, rel_tol=tol, abs_tol=1e-300), (k, cfl[k], v)


@pytest.mark.parametrize("N", [(17, 16, 16), (24, 16, 32), (33, 24, 48)], ids=["17", "24x16x32", "33x24x48"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_hydro_cfl_matches_oracle(N, precision):
    solver = PlaneCouetteFluctuationHIP(N=N, Re=400.0, dt=2e-3, perturbation_amplitude=0.05, precision=precision,
                                        time_integrator="SBDF3", nonlinear_form="rotational")
    oracle = oracle_for(solver)
    s = seeded_state(solver, seed=7, energy=5e-2, K=4)
    s = solver.solve(s, 3)
    u_max, b_max = solver.field_maxima(s)
    uo, bo = oracle.field_maxima(to_oracle(s))
    tol = 1e-12 if precision == "fp64" else 1e-5
    for a, b in zip(u_max, uo):
        assert math.isclose(a, b, rel_tol=tol), (u_max, uo)
    assert b_max == (0.0, 0.0, 0.0)
    _check(solver, solver.cfl_scalars(), oracle.cfl(to_oracle(s)), tol)


@pytest.mark.parametrize("kind", ["conducting", "shear", "insulating"])
def test_mhd_cfl_matches_oracle(kind):
    from jaxfun_amd.ic import seeded_mhd_state

    from tests.helpers import insulating_oracle_for, mhd_oracle_for, shear_oracle_for, to_mhd_oracle
    from tests.test_gpu_mhd import make as make_cond
    from tests.test_gpu_mhd import make_insulating, make_shear

    if kind == "conducting":
        solver = make_cond((17, 16, 16), magnetic_amplitude=0.1, background_b=(0.0, 0.0, 0.1))
        oracle = mhd_oracle_for(solver)
        s = solver.initial_state()
    elif kind == "shear":
        solver = make_shear((24, 16, 32))
        oracle = shear_oracle_for(solver)
        s = seeded_mhd_state(solver, seed=3, energy_u=1e-3, energy_b=1e-3, K=4)
    else:
        solver = make_insulating((17, 16, 16))
        oracle = insulating_oracle_for(solver)
        s = seeded_mhd_state(solver, seed=3, energy_u=1e-3, energy_b=1e-3, K=4)
    s = solver.solve(s, 2)
    so = to_mhd_oracle(s)
    u_max, b_max = solver.field_maxima(s)
    uo, bo = oracle.field_maxima(so)
    for a, b in zip(u_max + b_max, uo + bo):
        assert math.isclose(a, b, rel_tol=1e-12), (kind, u_max + b_max, uo + bo)
    _check(solver, solver.cfl_scalars(), oracle.cfl(so), 1e-12)
