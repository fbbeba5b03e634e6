"""GPU reverse mode of the conducting-wall MHD step against the oracle's MHD adjoint.

The oracle adjoint (oracle/adjoint_mhd.py) is pinned by central differences
and stage adjoint identities (tests/test_adjoint_mhd_oracle.py).  Bars:
* FP64 gradients of Epert + Emag agree with it leaf by leaf (flow, A and every
  history level) to 1e-11 relative per leaf group, off the Nyquist planes
  (states keep those planes at zero; the library returns zero there), from a
  fresh state (IMEXRK3 start-up then steady steps) and from one with a full
  history, with recomputed checkpoint segments;
* a cotangent seed on every final leaf of u, g and A gives the oracle's VJP;
* directional derivatives match central differences of the HIP forward.
"""

import numpy as np
import pytest

from oracle.adjoint import jax_convention as flow_jax
from oracle.adjoint_mhd import MHDOracleAdjoint
from tests.helpers import MHD_A_GROUPS, shear_oracle_for, to_mhd_oracle

pytestmark = pytest.mark.gpu


def make(N=(17, 16, 16), **kw):
    from jaxfun_amd import PlaneCouetteMRIShearpyHIP

    args = dict(Re=300.0, Rm=250.0, omega=0.5, shear_rate=1.2, background_b=(0.03, -0.02, 0.15), dt=3e-3,
                perturbation_amplitude=0.05, magnetic_amplitude=0.2, time_integrator="SBDF3",
                nonlinear_form="rotational")
    args.update(kw)
    return PlaneCouetteMRIShearpyHIP(N=N, **args)


def _mask(solver, a):
    a = np.array(a)
    if a.ndim == 3:
        a[:, solver.N[1] // 2, :] = 0
        a[:, :, solver.N[2] // 2] = 0
    return a


def _groups(s):
    f = s.flow
    out = [list(f.u) + [f.g]]
    for grp in (f.nonlinear_old, f.nonlinear_older, f.solution_old, f.solution_older):
        out.append(list(grp))
    for name in MHD_A_GROUPS:
        out.append(list(getattr(s, name)))
    return out


def grad_errors(solver, got, ref):
    """Per leaf group: max |got - ref| / scale on the masked leaves, the scale
    floored at 1e-6 of the group's largest entry."""
    out = []
    for ga, gb in zip(_groups(got), _groups(ref)):
        la = [_mask(solver, x) for x in ga]
        lb = [_mask(solver, x) for x in gb]
        gmax = max(float(np.max(np.abs(x))) for x in lb)
        worst = 0.0
        for a, b in zip(la, lb):
            scale = max(float(np.max(np.abs(b))), 1e-6 * gmax)
            if scale > 0:
                worst = max(worst, float(np.max(np.abs(a - b))) / scale)
        out.append(worst)
    return out


def mhd_jax(mu):
    """conj of every complex cotangent (jax's gradient convention)."""
    from dataclasses import replace

    c = lambda grp: tuple(np.conj(x) if np.iscomplexobj(x) else x for x in grp)  # noqa: E731
    return replace(mu, flow=flow_jax(mu.flow), **{name: c(getattr(mu, name)) for name in MHD_A_GROUPS})


@pytest.mark.parametrize("steps,every,history", [(3, 3, False), (6, 2, False), (4, 2, True)])
def test_energy_gradient_matches_oracle(steps, every, history):
    from jaxfun_amd.ic import seeded_mhd_state

    solver = make()
    s0 = seeded_mhd_state(solver, seed=4, energy_u=1e-2, energy_b=1e-2, K=4)
    if history:
        s0 = solver.solve(s0, 3)
    print("case", steps, every, history, flush=True)
    value, grad, final = solver.value_and_grad(s0, steps, checkpoint_every=every)
    o = shear_oracle_for(solver)
    adj = MHDOracleAdjoint(o)
    so = to_mhd_oracle(s0)
    if not history:
        so = o.reset_mhd_history(so)
    v_ref, mu = adj.mhd_value_and_grad(so, steps)
    assert abs(value - v_ref) <= 1e-12 * abs(v_ref), (value, v_ref)
    err = grad_errors(solver, grad, mhd_jax(mu))
    print(steps, every, history, ["%.1e" % e for e in err], flush=True)
    assert max(err) <= 1e-11, err
    # the state after the call is the forward's final state
    ref_final = solver.solve(s0, steps)
    for a, b in zip(list(final.flow.u) + list(final.A), list(ref_final.flow.u) + list(ref_final.A)):
        assert np.array_equal(np.asarray(a), np.asarray(b))


def test_cotangent_seed_matches_oracle_vjp():
    from dataclasses import replace

    from jaxfun_amd.ic import seeded_mhd_state

    solver = make()
    s0 = seeded_mhd_state(solver, seed=6, energy_u=1e-2, energy_b=1e-2, K=4)
    rng = np.random.default_rng(8)
    fin = solver.solve(s0, 4)

    def herm(a):
        # a Hermitian cotangent (the gradient of a real function of a real field)
        a = _mask(solver, a)
        flip = np.conj(np.roll(np.flip(a, axis=(1, 2)), 1, axis=(1, 2)))
        return _mask(solver, 0.5 * (a + flip))

    rnd = lambda x: herm(rng.standard_normal(np.shape(x)) + 1j * rng.standard_normal(np.shape(x)))  # noqa: E731
    cu = tuple(rnd(x) for x in fin.flow.u)
    cg = rnd(fin.flow.g)
    ca = tuple(rnd(x) for x in fin.A)
    final, grad = solver.vjp(s0, 4, cotangent=((*cu, cg), ca), checkpoint_every=2)
    o = shear_oracle_for(solver)
    adj = MHDOracleAdjoint(o)
    so = o.reset_mhd_history(to_mhd_oracle(s0))
    mu_f = adj.zeros(o.solve(so, 4))
    # jax cotangent c pairs as Re sum(c d) = Re sum(conj(conj c) d): the oracle cotangent is conj(c)
    mu_f = replace(mu_f, flow=replace(mu_f.flow, u=tuple(np.conj(x) for x in cu), g=np.conj(cg)),
                   A=tuple(np.conj(x) for x in ca))
    mu = adj.mhd_vjp(so, 4, mu_f)
    err = grad_errors(solver, grad, mhd_jax(mu))
    print("cotangent", ["%.1e" % e for e in err], flush=True)
    assert max(err) <= 1e-11, err


def test_gradient_matches_hip_central_differences():
    """Re sum(grad d) against central differences of the HIP forward along a
    seeded divergence-free direction in u and a direction in A."""
    from dataclasses import replace

    from jaxfun_amd.ic import seeded_mhd_state

    solver = make()
    s0 = seeded_mhd_state(solver, seed=2, energy_u=1e-2, energy_b=1e-2, K=4)
    d = seeded_mhd_state(solver, seed=9, energy_u=1e-2, energy_b=1e-2, K=3)
    steps = 5
    value, grad, _ = solver.value_and_grad(s0, steps, checkpoint_every=2)

    def J(eps):
        f = s0.flow
        flow = replace(f, u=tuple(a + eps * b for a, b in zip(f.u, d.flow.u)), g=f.g + eps * d.flow.g)
        st = replace(s0, flow=flow, A=tuple(a + eps * b for a, b in zip(s0.A, d.A)))
        solver.solve(st, steps)
        from jaxfun_amd.mhd import PlaneCouetteMHDHIP

        dd = PlaneCouetteMHDHIP.diagnostics(solver)
        return dd["Epert"] + dd["Emag"]

    eps = 1e-4
    fd = (J(eps) - J(-eps)) / (2 * eps)
    an = float(sum(np.real(np.sum(a * b)) for a, b in zip(list(grad.flow.u) + [grad.flow.g] + list(grad.A),
                                                           list(d.flow.u) + [d.flow.g] + list(d.A))))
    print("fd", fd, "adjoint", an, flush=True)
    assert abs(fd - an) <= 1e-6 * abs(an)


def test_reverse_mode_scope_is_enforced():
    """Hydrodynamic handles refuse vector-potential gradients; insulating walls
    and mixed precision refuse the MHD reverse mode with NotImplementedError."""
    import ctypes as C

    from jaxfun_amd import PlaneCouetteFluctuationHIP, PlaneCouetteMRIShearpyInsulatingHIP
    from jaxfun_amd import _native as nat

    hyd = PlaneCouetteFluctuationHIP(N=(17, 16, 16), Re=400.0, dt=2e-3, time_integrator="SBDF3",
                                     nonlinear_form="rotational")
    hyd.value_and_grad(hyd.initial_state(), 2, checkpoint_every=2)
    buf = np.zeros(2 * 15 * 16 * 16)
    assert hyd._lib.pcf_get_gradient(hyd._h, nat.FIELD_ID["AX"], buf.ctypes.data_as(C.POINTER(C.c_double))) == -1
    assert b"flow leaves only" in hyd._lib.pcf_last_error()
    ins = PlaneCouetteMRIShearpyInsulatingHIP(N=(17, 16, 16), time_integrator="SBDF3", nonlinear_form="rotational")
    with pytest.raises(NotImplementedError):
        ins.value_and_grad(ins.initial_state(), 2)
    mixed = make(precision="mixed")
    with pytest.raises(NotImplementedError):
        mixed.value_and_grad(mixed.initial_state(), 2)
