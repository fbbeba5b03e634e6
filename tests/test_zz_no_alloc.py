"""pcf_field_maxima_noalloc against pcf_field_maxima, bit for bit, through the
C ABI: two calls in a row on one GPU, the second with a smaller grid (the
shared device array keeps the first call's contents), for a hydrodynamic and
an insulating-wall MHD handle (three and six fields)."""

import ctypes as C

import numpy as np
import pytest

from jaxfun_amd import PlaneCouetteFluctuationHIP, PlaneCouetteMRIShearpyInsulatingHIP
from jaxfun_amd import _native as nat
from jaxfun_amd.ic import seeded_mhd_state, seeded_state

pytestmark = pytest.mark.gpu


def _maxima(solver):
    ref, noalloc = np.full(6, -1.0), np.full(6, -1.0)
    dp = C.POINTER(C.c_double)
    nat.check(solver._lib.pcf_field_maxima(solver._h, ref.ctypes.data_as(dp)))
    nat.check(solver._lib.pcf_field_maxima_noalloc(solver._h, noalloc.ctypes.data_as(dp)))
    return ref, noalloc


def test_noalloc_maxima_equal_allocating_maxima_hydro():
    for N in ((33, 32, 32), (17, 16, 16)):
        solver = PlaneCouetteFluctuationHIP(N=N, Re=400.0, dt=2e-3, time_integrator="SBDF3",
                                            nonlinear_form="rotational")
        solver.load(seeded_state(solver, seed=3, energy=5e-2, K=4))
        solver.advance(3)
        ref, noalloc = _maxima(solver)
        print(N, ref, flush=True)
        assert np.array_equal(ref, noalloc), (N, ref, noalloc)
        assert np.all(ref[:3] > 0) and np.all(ref[3:] == 0)
        solver.close()


def test_noalloc_maxima_equal_allocating_maxima_mhd():
    for N in ((24, 16, 32), (17, 16, 16)):
        solver = PlaneCouetteMRIShearpyInsulatingHIP(N=N, Re=400.0, Rm=300.0, omega=0.4, shear_rate=1.5,
                                                     background_b=(0.0, 0.02, 0.1), dt=2e-3,
                                                     time_integrator="SBDF3", nonlinear_form="rotational")
        solver.load(seeded_mhd_state(solver, seed=3, energy_u=1e-3, energy_b=1e-3, K=4))
        solver.advance(3)
        ref, noalloc = _maxima(solver)
        print(N, ref, flush=True)
        assert np.array_equal(ref, noalloc), (N, ref, noalloc)
        assert np.all(ref > 0)
        solver.close()
