"""GPU parity of the rotating-shear diagnostic set (``pcf_mhd_extended_diagnostics``).

The device reductions (nodal-spectral evaluation at the Gauss nodes and both
walls, Parseval sums, DFTs to the unpadded grid for the electric field) are
compared key by key with the oracle's literal restatement of
``PlaneCouetteMRIShearpyJax.diagnostics`` (examples/pcf_mhd_mri_shearpy_jax.py:178-513)
on seeded stepped states, and with the closed forms of the manufactured state
of ``tests/test_mhd_extended_diag_oracle.py``, for conducting and insulating
walls.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from tests.helpers import insulating_oracle_for, shear_oracle_for, to_mhd_oracle

pytestmark = pytest.mark.gpu


def _solver(insulating, N, **kw):
    from jaxfun_amd import PlaneCouetteMRIShearpyHIP, PlaneCouetteMRIShearpyInsulatingHIP

    args = dict(Re=400.0, Rm=300.0, omega=0.4, shear_rate=1.5, background_b=(0.03, 0.02, 0.1), dt=2e-3,
                perturbation_amplitude=0.05, magnetic_amplitude=0.3, time_integrator="SBDF3",
                nonlinear_form="rotational")
    args.update(kw)
    cls = PlaneCouetteMRIShearpyInsulatingHIP if insulating else PlaneCouetteMRIShearpyHIP
    return cls(N=N, **args)


def _assert_keys(d, ref, where):
    scale = max(1.0, max(abs(v) for v in ref.values() if isinstance(v, float) and math.isfinite(v)))
    worst, wkey = 0.0, None
    for key, r in ref.items():
        assert key in d, (where, key)
        v = d[key]
        if isinstance(r, float) and math.isnan(r):
            assert math.isnan(v), (where, key)
            continue
        if isinstance(r, float) and math.isinf(r):
            assert v == r, (where, key)
            continue
        err = abs(v - r)
        if abs(r) > 1e-10 * scale and err / abs(r) > worst:  # keys above rounding level
            worst, wkey = err / abs(r), key
        assert err <= 1e-11 * abs(r) + 1e-14 * scale, (where, key, v, r)
    print(where, "keys", len(ref), "worst relative error", f"{worst:.2e}", wkey, flush=True)


@pytest.mark.parametrize("insulating", [False, True], ids=["conducting", "insulating"])
@pytest.mark.parametrize("N", [(17, 16, 16), (24, 16, 32)], ids=["17x16x16", "24x16x32"])
def test_extended_diagnostics_match_oracle(insulating, N):
    from jaxfun_amd.ic import seeded_mhd_state

    solver = _solver(insulating, N)
    oracle = (insulating_oracle_for if insulating else shear_oracle_for)(solver)
    s = seeded_mhd_state(solver, seed=5, energy_u=1e-3, energy_b=1e-3, K=4)
    s = solver.solve(s, 3)
    print("case", "insulating" if insulating else "conducting", N, flush=True)
    d = solver.diagnostics(s)
    ref = oracle.extended_diagnostics(to_mhd_oracle(s))
    _assert_keys(d, ref, (N, insulating))
    # a second call reuses the device buffers and sees the new state
    s2 = solver.solve(s, 2)
    _assert_keys(solver.diagnostics(s2), oracle.extended_diagnostics(to_mhd_oracle(s2)), (N, insulating, "again"))


@pytest.mark.parametrize("insulating", [False, True], ids=["conducting", "insulating"])
def test_extended_diagnostics_closed_forms(insulating):
    from jaxfun_amd.pcf import KMMState
    from oracle.mhd import RotatingShearMHDOracle
    from oracle.mhd_insulating import InsulatingMHDOracle
    from tests import test_mhd_extended_diag_oracle as M

    cls = InsulatingMHDOracle if insulating else RotatingShearMHDOracle
    o = M._oracle(cls)
    so = M._state(o)
    solver = _solver(insulating, M.N, Re=M.NU_RE, Rm=M.ETA_RM, omega=M.OMEGA, shear_rate=M.S, background_b=M.B0,
                     dt=0.01, perturbation_amplitude=0.0, magnetic_amplitude=0.0)
    s = solver._new_state(KMMState(u=so.flow.u, g=so.flow.g), so.A)
    print("case closed forms", "insulating" if insulating else "conducting", flush=True)
    d = solver.diagnostics(s)
    exp = M._expected(o)
    scale = max(abs(v) for v in exp.values())
    for key, ref in exp.items():
        assert abs(d[key] - ref) <= 1e-12 * max(abs(ref), 1e-3 * scale), (key, d[key], ref)
    assert d["divergence_e_resistive_l2"] < 1e-13
    assert d["mean_b_trace_mismatch_linf"] < 1e-14
    assert np.isfinite(d["alpha_Sh"]) and d["q_shear"] == M.S / M.OMEGA
