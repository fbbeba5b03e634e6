"""The per-launch stage timer (pcf_profile_stage) returns positive times and
leaves the state untouched: steps after it equal steps without it, bit for bit."""

import ctypes as C

import numpy as np
import pytest

from jaxfun_amd import PlaneCouetteFluctuationHIP
from jaxfun_amd import _native as nat
from jaxfun_amd.ic import seeded_state

pytestmark = pytest.mark.gpu


def test_profile_stage_leaves_state_untouched():
    a = PlaneCouetteFluctuationHIP(N=(17, 16, 16), Re=400.0, dt=2e-3, time_integrator="SBDF3",
                                   nonlinear_form="rotational")
    b = PlaneCouetteFluctuationHIP(N=(17, 16, 16), Re=400.0, dt=2e-3, time_integrator="SBDF3",
                                   nonlinear_form="rotational")
    s = seeded_state(a, seed=1, energy=5e-2, K=4)
    for x in (a, b):
        x.load(s)
        x.advance(3)
    for stage in range(5):
        ms = C.c_double()
        nat.check(a._lib.pcf_profile_stage(a._h, stage, 5, C.byref(ms)))
        print("stage", stage, ms.value, flush=True)
        assert ms.value > 0.0
    with pytest.raises(ValueError):
        nat.check(a._lib.pcf_profile_stage(a._h, 5, 5, C.byref(C.c_double())))
    for x in (a, b):
        x.advance(4)
    sa, sb = a.fetch(), b.fetch()
    for x, y in zip(list(sa.u) + [sa.g] + list(sa.nonlinear_old), list(sb.u) + [sb.g] + list(sb.nonlinear_old)):
        assert np.array_equal(x, y)
