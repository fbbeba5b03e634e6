"""Device-memory state exchange (pcf_set_field_device / pcf_get_field_device):
states whose leaves are torch CUDA tensors load and fetch without host copies
and step exactly as host states do.

Bar: bitwise equality with the host path for every leaf (the same kernels run
on the same values; only the copies differ).
"""

import numpy as np
import pytest

from jaxfun_amd import PlaneCouetteFluctuationHIP, PlaneCouetteMHDHIP
from jaxfun_amd.ic import seeded_state
from jaxfun_amd.pcf import KMMState, _leaves

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def make(N=(17, 16, 16), **kw):
    args = dict(Re=400.0, dt=1e-3, perturbation_amplitude=0.05, time_integrator="SBDF3",
                nonlinear_form="rotational")
    args.update(kw)
    return PlaneCouetteFluctuationHIP(N=N, **args)


def host(x):
    return x.cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)


def assert_same(a, b):
    la, lb = _leaves(a), _leaves(b)
    assert len(la) == len(lb)
    for x, y in zip(la, lb):
        hx, hy = host(x), host(y)
        assert hx.shape == hy.shape and hx.dtype == hy.dtype
        assert np.array_equal(hx, hy)


def test_fetch_device_matches_fetch():
    solver = make((24, 16, 32))
    st = solver.solve(seeded_state(solver, seed=3, energy=1e-3, K=4), 4)
    d = solver.fetch_device()
    for leaf in _leaves(d):
        assert isinstance(leaf, torch.Tensor) and leaf.is_cuda
        assert leaf.dtype in (torch.complex128, torch.float64)
    assert d.u[0].shape == (20, 16, 32) and d.g.shape == (22, 16, 32) and d.nonlinear_old[2].shape == (22,)
    assert_same(d, st)
    assert (d.history_steps, d.have_old, d.previous_dt) == (st.history_steps, st.have_old, st.previous_dt)


@pytest.mark.parametrize("integrator", ["SBDF3", "CNAB2", "IMEXRK3"])
def test_solve_device_chain_equals_host_chain(integrator):
    a, b = make(time_integrator=integrator), make(time_integrator=integrator)
    s0 = a.initial_state()
    sa = a.solve(a.solve(s0, 2), 5)
    sb = b.solve_device(b.solve_device(s0, 2), 5)
    assert isinstance(sb.u[0], torch.Tensor)
    assert_same(sb, sa)
    assert (sb.history_steps, sb.have_old, sb.previous_dt) == (sa.history_steps, sa.have_old, sa.previous_dt)


def test_device_state_loads_into_another_solver():
    """A device state handed to a second solver is copied device to device
    (loaded u1, u2 checked against their reconstruction as for host states)
    and continues bitwise like the host state."""
    a, b, c = make(), make(), make()
    s0 = seeded_state(a, seed=5, energy=1e-2, K=4)
    sa = a.solve(s0, 6)
    sb = b.solve_device(s0, 6)
    assert_same(c.solve(sb, 4), a.solve(sa, 4))
    # a device state without history restarts the SBDF3 history, like a host one
    bare = KMMState(u=tuple(t.clone() for t in sb.u), g=sb.g.clone())
    ref = KMMState(u=tuple(host(t) for t in sb.u), g=host(sb.g))
    assert_same(make().solve(bare, 3), make().solve(ref, 3))


def test_mhd_device_state_round_trip():
    kw = dict(N=(17, 16, 16), Re=200.0, Rm=200.0, dt=1e-3, perturbation_amplitude=0.05, magnetic_amplitude=0.05,
              time_integrator="SBDF3", nonlinear_form="rotational")
    a, b = PlaneCouetteMHDHIP(**kw), PlaneCouetteMHDHIP(**kw)
    s0 = a.initial_state()
    sa = a.solve(s0, 4)
    sb = b.solve_device(s0, 4)
    assert isinstance(sb.A[0], torch.Tensor)
    assert_same(sb, sa)
    assert_same(PlaneCouetteMHDHIP(**kw).solve(sb, 2), a.solve(sa, 2))
