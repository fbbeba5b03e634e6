"""GPU runs of the benchmark workloads against the oracle.

* 33x32x32, Re = 400, dt = 0.01, seeded field (seed 0, E = 1e-4, lowest 8
  modes), 100 SBDF3 steps in FP64: every state and history leaf against the
  FP64 oracle run live (1e-12 relative; nonlinear history 5e-11).
* 65x64x64, Re = 1000, dt = 0.005, seeded field (seed 0, E = 1e-3, lowest 12
  modes), 1000 SBDF3 steps in FP64: against the oracle fixture written by
  ``tests/golden/make_pcf_seeded_fixture.py`` -- the low-mode coefficients of
  u0, u1, u2, g, every leaf's l2 norm and the diagnostics at steps 2, 100, 250,
  500 and 1000 (the error growth along the run is printed).
* 97x96x96 and 129x128x128, Re = 2000, dt = 0.002, seeded field (E = 1e-2,
  lowest 12 modes): the mixed-precision step against the HIP FP64 step after
  200 steps (velocity within the stated tolerance, energy, wall shear,
  divergence at round-off).
"""

import base64
import json
import math
from pathlib import Path

import numpy as np
import pytest

from jaxfun_amd import PlaneCouetteFluctuationHIP
from jaxfun_amd.ic import seeded_state

from tests.helpers import LEAF_NAMES, leaves, max_rel_errors, oracle_for, to_oracle

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
HIST_TOL = 5e-11  # nonlinear history rows: see tests/test_gpu_parity.py
FIXTURE = Path(__file__).parent / "golden" / "pcf_seeded_65x64x64_re1000.json"


def make(N, Re, dt, precision="fp64"):
    return PlaneCouetteFluctuationHIP(N=N, Re=Re, dt=dt, precision=precision, time_integrator="SBDF3",
                                      nonlinear_form="rotational", coefficient_path="optimized")


def test_seeded_33_re400_dt001_100_steps_match_oracle():
    N = (33, 32, 32)
    solver = make(N, 400.0, 0.01)
    oracle = oracle_for(solver)
    s0 = seeded_state(solver, seed=0, energy=1e-4, K=8)
    s = solver.solve(s0, 100)
    so = oracle.solve(to_oracle(s0), 100)
    err = max_rel_errors(s, so)
    print("33^3 dt=0.01 after 100 steps:", max(err.values()), flush=True)
    for k, v in err.items():
        assert v < (HIST_TOL if k.startswith("nl_") else FP64_TOL), (k, v)
    d, do = solver.diagnostics(s), oracle.diagnostics(so)
    for k in ("Epert", "Etot", "wall_shear_top", "wall_shear_bot", "mean_shear"):
        assert math.isclose(d[k], do[k], rel_tol=1e-12), (k, d[k], do[k])
    assert d["divL2"] < 1e-14


def _decode(text, shape):
    return np.frombuffer(base64.b64decode(text), dtype="<c16").reshape(shape)


def test_seeded_65_re1000_1000_steps_match_oracle_fixture():
    fx = json.loads(FIXTURE.read_text())
    N = tuple(fx["N"])
    w = fx["workload"]
    solver = make(N, w["Re"], w["dt"])
    s = seeded_state(solver, seed=w["seed"], energy=w["energy"], K=w["K"])
    iy, iz = fx["sample_ky_index"], fx["sample_kz_index"]
    done = 0
    for k in sorted(int(x) for x in fx["records"]):
        s = solver.solve(s, k - done)
        done = k
        rec = fx["records"][str(k)]
        worst_sample = 0.0
        for name, a in zip(("u0", "u1", "u2", "g"), (*s.u, s.g)):
            got = a[:, iy][:, :, iz]
            want = _decode(rec["samples"][name], got.shape)
            worst_sample = max(worst_sample, float(np.max(np.abs(got - want)) / np.max(np.abs(want))))
        worst_norm = {}
        for name, v in zip(LEAF_NAMES, leaves(s)):
            ref = rec["leaves"][name]["l2_norm"]
            worst_norm[name] = abs(float(np.linalg.norm(v)) - ref) / max(ref, 1e-300)
        d = solver.diagnostics(s)
        dE = abs(d["Epert"] - rec["diagnostics"]["Epert"]) / rec["diagnostics"]["Epert"]
        print(f"65^3 step {k}: low-mode coefficients {worst_sample:.2e}, leaf norms "
              f"{max(worst_norm.values()):.2e}, Epert {dE:.2e}", flush=True)
        assert worst_sample < FP64_TOL, (k, worst_sample)
        for name, v in worst_norm.items():
            assert v < (HIST_TOL if name.startswith("nl_") else FP64_TOL), (k, name, v)
        assert dE < FP64_TOL, (k, dE)
        for key in ("wall_shear_top", "wall_shear_bot", "mean_shear"):
            assert abs(d[key] - rec["diagnostics"][key]) < 1e-13, (k, key)
        assert [s.history_steps, s.have_old, s.previous_dt] == rec["history"]


# Mixed precision against the FP64 step at the large sizes: velocity error
# bound after 200 steps at E = 1e-2, Re = 2000 (the strongest nonlinearity of
# the workloads); the FP64 solves and reconstruction keep div u at round-off.
LARGE_TOL = 1e-5


@pytest.mark.parametrize("N", [(97, 96, 96), (129, 128, 128)])
def test_mixed_large_close_to_fp64(N):
    steps = 200
    a, b = make(N, 2000.0, 0.002), make(N, 2000.0, 0.002, precision="mixed")
    s0 = seeded_state(a, seed=0, energy=1e-2, K=12)
    sa, sb = a.solve(s0, steps), b.solve(s0, steps)
    errs = {name: float(np.max(np.abs(x - y)) / np.max(np.abs(y)))
            for name, x, y in zip(("u0", "u1", "u2", "g"), (*sb.u, sb.g), (*sa.u, sa.g))}
    da, db = a.diagnostics(sa), b.diagnostics(sb)
    print(N, "mixed vs fp64 after", steps, "steps:", errs, "Epert", da["Epert"], db["Epert"], flush=True)
    assert max(errs.values()) < LARGE_TOL, errs
    assert abs(db["Epert"] - da["Epert"]) < LARGE_TOL * da["Epert"]
    for k in ("wall_shear_top", "wall_shear_bot"):
        assert abs(db[k] - da[k]) < LARGE_TOL
    assert db["divL2"] < 1e-13 and da["divL2"] < 1e-13
    assert errs["u0"] > 1e-12  # a different arithmetic path really ran
