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


def test_profile_reverse_leaves_state_untouched_and_gradient_unchanged():
    """pcf_profile_reverse (reverse steady steps alone on one stream) returns
    positive times, puts the state back, and a pcf_vjp after it gives the same
    gradient as one on a handle that never ran it."""
    kw = dict(N=(17, 16, 16), Re=400.0, dt=2e-3, time_integrator="SBDF3", nonlinear_form="rotational")
    a, b = PlaneCouetteFluctuationHIP(**kw), PlaneCouetteFluctuationHIP(**kw)
    s = seeded_state(a, seed=2, energy=5e-2, K=4)
    for x in (a, b):
        x.load(s)
        x.advance(3)
    out = np.zeros(2)
    nat.check(a._lib.pcf_profile_reverse(a._h, 6, out.ctypes.data_as(C.POINTER(C.c_double))))
    print("reverse ms", out, flush=True)
    assert out[0] > 0.0 and out[1] > 0.0
    with pytest.raises(ValueError):
        nat.check(a._lib.pcf_profile_reverse(a._h, 0, out.ctypes.data_as(C.POINTER(C.c_double))))
    sa, sb = a.fetch(), b.fetch()
    for x, y in zip(list(sa.u) + [sa.g], list(sb.u) + [sb.g]):
        assert np.array_equal(x, y)
    for x in (a, b):
        nat.check(x._lib.pcf_vjp(x._h, 5, 2, nat.SEED_ENERGY))
    ga = np.zeros(np.asarray(sa.u[0]).size * 2)
    gb = np.zeros_like(ga)
    nat.check(a._lib.pcf_get_gradient(a._h, nat.FIELD_ID["U0"], ga.ctypes.data_as(C.POINTER(C.c_double))))
    nat.check(b._lib.pcf_get_gradient(b._h, nat.FIELD_ID["U0"], gb.ctypes.data_as(C.POINTER(C.c_double))))
    assert np.array_equal(ga, gb)
