"""Host logic of the problem-spec runner: spec parsing, scalar rules, health
guards and the checkpoint container and payload (no GPU).

Known answers come from the reference's own tests where they exist
(``tests/production/test_run_problem.py:472-475`` for the saturation rule) and
otherwise from values worked by hand in each test.
"""

import json
import math
import warnings
from pathlib import Path

import numpy as np
import pytest

from jaxfun_amd.pcf import KMMState
from jaxfun_amd.runner import (
    ProductionOracleNotImplementedError,
    checkpoint_payload,
    dedupe_time_rows,
    growth_rate_from_energy,
    kmm_time_integrator,
    materialize_resolution_tier,
    padding_factor,
    raise_on_unhealthy,
    read_checkpoint,
    run_pcf_fluctuation_saturation,
    saturation_passed,
    selected_resolution,
    state_from_checkpoint_payload,
    stationarity_scalars,
    steps_from_spec,
    write_checkpoint,
)

SPEC = {
    "problem_id": "pcf_fluct_re400",
    "nondimensional_groups": {"Re": 400.0, "U_wall": 1.0},
    "domain": {"x": [-1.0, 1.0], "y_period": 4 * math.pi, "z_period": 2 * math.pi},
    "resolution": {"smoke": {"Nx": 9, "Ny": 4, "Nz": 4, "dealias": [1.0, 1.0, 1.0]},
                   "start": {"Nx": 32, "Ny": 64, "Nz": 32}, "production": {"Nx": 64, "Ny": 128, "Nz": 64},
                   "dealias": [1.0, 1.5, 1.5], "family": "C"},
    "time": {"integrator": "IMEXRK222", "dt": 0.01, "final_time": 50.0},
    "initial_condition": {"type": "finite_amplitude_fluctuation", "amplitude": 0.1},
}


def test_resolution_tiers():
    assert selected_resolution(SPEC)["Nx"] == 64  # the production tier by default
    start = materialize_resolution_tier(SPEC, "start")
    assert selected_resolution(start)["Nx"] == 32 and start["resolution"]["dealias"] == [1.0, 1.5, 1.5]
    assert SPEC["resolution"].get("Nx") is None  # the spec itself is not changed
    smoke = materialize_resolution_tier(SPEC, "smoke")
    assert selected_resolution(smoke)["dealias"] == [1.0, 1.0, 1.0]
    with pytest.raises(ValueError, match="not defined"):
        materialize_resolution_tier(SPEC, "huge")


def test_padding_forms():
    assert padding_factor({"dealias": {"z": 1.5, "x": 1.0, "y": 2.0}}) == (1.0, 2.0, 1.5)
    assert padding_factor({"dealias": 1.5}) == (1.5, 1.5, 1.5)
    assert padding_factor({}) == (1.0, 1.0, 1.0)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        assert padding_factor({"dealias": [1.0, 1.5, 1.5]}) == (1.0, 1.5, 1.5)
        assert padding_factor({"dealias": [1.5, 1.5, 1.5]}) == (1.5, 1.5, 1.5)
    assert [x.category for x in w] == [DeprecationWarning]  # only the non-uniform list
    for bad in ({"x": 1.0, "y": 1.5}, [1.0, 1.5], 0.5, [1.0, "a", 1.5]):
        with pytest.raises(ValueError):
            padding_factor({"dealias": bad})


def test_time_integrator_and_steps():
    assert kmm_time_integrator(SPEC) == "IMEXRK222"
    for bad in ("RK4", "Dedalus"):
        with pytest.raises(ProductionOracleNotImplementedError, match="time-stepping integrator"):
            kmm_time_integrator({"time": {"integrator": bad}})
    with pytest.raises(ProductionOracleNotImplementedError, match="adaptive_cfl"):
        kmm_time_integrator({"time": {"integrator": "SBDF3", "adaptive_cfl": {"target": 0.5}}})
    assert steps_from_spec(SPEC) == 5000
    assert steps_from_spec(SPEC, 7) == 7
    with pytest.raises(ValueError):
        steps_from_spec(SPEC, -1)
    with pytest.raises(ValueError, match="integer multiple"):
        steps_from_spec({"time": {"dt": 0.03, "final_time": 0.1}})


def test_growth_rate_rule():
    # 0.5 log(e1/e0) / (steps dt): e1/e0 = e^2 over t = 1 gives 1
    assert math.isclose(growth_rate_from_energy(1.0, math.exp(2.0), 10, 0.1), 1.0, rel_tol=1e-15)
    assert math.isclose(growth_rate_from_energy(2.0, 0.5, 4, 0.5), 0.5 * math.log(0.25) / 2.0, rel_tol=1e-15)
    assert growth_rate_from_energy(1.0, 0.0, 1, 0.1) == -math.inf
    assert math.isnan(growth_rate_from_energy(0.0, 1.0, 1, 0.1))
    assert math.isnan(growth_rate_from_energy(1.0, math.inf, 1, 0.1))
    with pytest.raises(ValueError, match="at least one"):
        growth_rate_from_energy(1.0, 2.0, 0, 0.1)


def test_saturation_rule_reference_cases():
    assert not saturation_passed(float("inf"), threshold=2.0, final_energies=(1.0,))
    assert not saturation_passed(3.0, threshold=2.0, final_energies=(float("inf"),))
    assert not saturation_passed(1.5, threshold=2.0, final_energies=(1.0,))
    assert saturation_passed(3.0, threshold=2.0, final_energies=(1.0, 0.0))
    assert not saturation_passed(2.0, threshold=2.0, final_energies=(1.0,))  # strictly above
    assert not saturation_passed(3.0, threshold=2.0, final_energies=(-1e-30,))
    assert not saturation_passed("x", threshold=2.0, final_energies=())


def test_stationarity_windows():
    rows = [{"t": i, "E": v} for i, v in enumerate([1.0, 2.0, 3.0, 4.0])]
    out = stationarity_scalars(rows, key="E")
    # quarter = 1: previous mean 3, current mean 4, change 1/4
    assert out["stationarity_previous_mean"] == 3.0 and out["stationarity_current_mean"] == 4.0
    assert out["stationarity_relative_change"] == 0.25 and out["stationarity_check_passed"] is False
    flat = [{"E": v} for v in [5.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.02]]
    out = stationarity_scalars(flat, key="E")  # quarter = 2: means 1.0 and 1.01
    assert math.isclose(out["stationarity_relative_change"], 0.01 / 1.01, rel_tol=1e-12)
    assert out["stationarity_check_passed"] is True and out["stationarity_window_samples"] == 8
    few = stationarity_scalars(rows[:3] + [{"t": 9, "E": math.nan}], key="E")
    assert few["stationarity_window_samples"] == 3 and few["stationarity_check_passed"] is None


def test_dedupe_keeps_the_last_row_of_a_time():
    rows = [{"t": 0.0, "a": 1}, {"t": 0.1, "a": 2}, {"t": 0.1, "a": 3}, {"t": 0.2, "a": 4}]
    assert dedupe_time_rows(rows) == [{"t": 0.0, "a": 1}, {"t": 0.1, "a": 3}, {"t": 0.2, "a": 4}]


def test_health_guards():
    good = {"Epert": 1e-3, "Etot": 1.0, "divL2": 1e-12}
    raise_on_unhealthy(good, t=0.1, tstep=10)
    for bad, match in (({"Epert": math.nan}, "nonfinite"), ({"Etot": math.inf}, "nonfinite"),
                       ({"Etot": 2e30}, "runaway"), ({"divL2": 2e-2}, "divergence"),
                       ({"divL2": math.nan}, "divergence")):
        with pytest.raises(FloatingPointError, match=match):
            raise_on_unhealthy({**good, **bad}, t=0.1, tstep=10)


def _state(rng, with_history=True):
    def f(n):
        return rng.standard_normal((n, 4, 6)) + 1j * rng.standard_normal((n, 4, 6))

    u = (f(5), f(7), f(7))
    if not with_history:
        return KMMState(u=u, g=f(7))
    rows = (f(5), f(7), rng.standard_normal(7) + 0j, rng.standard_normal(7) + 0j)
    return KMMState(u=u, g=f(7), nonlinear_old=rows, nonlinear_older=tuple(r * 2 for r in rows),
                    solution_old=(f(5), f(7), rng.standard_normal(7) + 0j, rng.standard_normal(7) + 0j),
                    solution_older=(f(5), f(7), rng.standard_normal(7) + 0j, rng.standard_normal(7) + 0j),
                    history_steps=2.0, have_old=1.0, previous_dt=1e-3)


def _same(a: KMMState, b: KMMState):
    for name in ("u", "nonlinear_old", "nonlinear_older", "solution_old", "solution_older"):
        x, y = getattr(a, name), getattr(b, name)
        assert (x is None) == (y is None), name
        if x is not None:
            assert len(x) == len(y) and all(np.array_equal(p, q) and p.dtype == q.dtype for p, q in zip(x, y))
    assert np.array_equal(a.g, b.g)
    assert (a.history_steps, a.have_old, a.previous_dt) == (b.history_steps, b.have_old, b.previous_dt)


def test_payload_keys_follow_the_reference():
    rng = np.random.default_rng(3)
    full = checkpoint_payload(_state(rng))
    assert set(full) == {"u", "g", "nonlinear_old", "have_old", "previous_dt", "nonlinear_older", "solution_old",
                         "solution_older", "history_steps"}
    assert full["have_old"].dtype == np.float32
    bare = checkpoint_payload(_state(rng, with_history=False))
    assert set(bare) == {"u", "g", "history_steps"}
    with pytest.raises(NotImplementedError):
        state_from_checkpoint_payload(full, "axisymmetric_tc_hydro")


def test_checkpoint_container_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    path = tmp_path / "checkpoints" / "checkpoints.npz"
    path.parent.mkdir()
    states = {}
    for tstep in (12, 24, 36):
        st = _state(rng, with_history=tstep != 24)
        states[tstep] = st
        write_checkpoint(path, {"state": checkpoint_payload(st), "extra": {"a": [np.arange(3), (1.5,)]}},
                         t=tstep * 1e-3, tstep=tstep, attrs={"state_kind": "pcf_fluctuation_saturation"})
    write_checkpoint(path, {"state": checkpoint_payload(states[24])}, t=0.024, tstep=24,
                     attrs={"state_kind": "pcf_fluctuation_saturation"})  # a rewritten step replaces its group
    rec = read_checkpoint(path)
    assert rec.tstep == 24 and rec.t == 0.024  # the last write is the latest step
    assert rec.attrs == {"t": 0.024, "tstep": 24, "state_kind": "pcf_fluctuation_saturation"}
    assert set(rec.fields) == {"state"}
    _same(state_from_checkpoint_payload(rec.fields["state"]), states[24])
    for tstep in (12, 36):
        rec = read_checkpoint(path, step=tstep)
        _same(state_from_checkpoint_payload(rec.fields["state"]), states[tstep])
        assert isinstance(rec.fields["extra"]["a"], list) and isinstance(rec.fields["extra"]["a"][1], tuple)
        assert np.array_equal(rec.fields["extra"]["a"][0], np.arange(3))
    with np.load(path, allow_pickle=False) as f:  # the reference's tree layout
        assert "checkpoints/36/fields/state/u/0" in f.files
        assert "checkpoints/36/fields/state/nonlinear_old/3" in f.files
        assert json.loads(str(f["__meta__"]))["latest_step"] == 24


def test_run_validates_cadence_before_building(tmp_path):
    with pytest.raises(ValueError, match="out_dir is required"):
        run_pcf_fluctuation_saturation(SPEC, steps=2, checkpoint_every=1)
    with pytest.raises(ValueError, match="must be positive"):
        run_pcf_fluctuation_saturation(SPEC, steps=2, out_dir=tmp_path, diagnostics_every=0)


# ---------------------------------------------------------------- vector-potential MHD runner
VP_GROUPS = {"Re": 1000.0, "Rm": 1000.0, "Pm": 1.0, "Omega": 2.0 / 3.0, "S": 1.0, "q_shear": 1.5, "B0": 0.025,
             "nu": 0.001, "eta_mag": 0.001}


def _vp_spec(**groups):
    return {"geometry": "pcf", "physics": "mri", "nondimensional_groups": {**VP_GROUPS, **groups},
            "domain": {"x": [-1.0, 1.0], "y_period": 4.0, "z_period": 1.0},
            "resolution": {"smoke": {"Nx": 9, "Ny": 8, "Nz": 8}, "start": {"Nx": 17, "Ny": 16, "Nz": 16},
                           "production": {"Nx": 33, "Ny": 32, "Nz": 32}},
            "time": {"integrator": "SBDF3", "dt": 0.005, "final_time": 0.06}}


def test_pcf_physics_resolution():
    """``resolve_physics`` (``production/physics.py:408-515``) on the reference's
    ``production/runs/exp_pcf_mri_vector_potential.json`` groups: h = 1, the
    scale |S| h^2 = 1, so nu = 1/Re_h and eta = 1/Rm_h; a Pm, nu or eta that
    disagrees with the groups is rejected."""
    from jaxfun_amd.runner import ProblemSpecError, resolve_pcf_physics

    p = resolve_pcf_physics(_vp_spec())
    assert (p["h"], p["S"], p["Re_h"], p["Rm_h"], p["B0"]) == (1.0, 1.0, 1000.0, 1000.0, 0.025)
    assert p["Omega"] == pytest.approx(2.0 / 3.0) and p["Pm"] == pytest.approx(1.0)
    for bad in ({"Pm": 2.0}, {"nu": 0.002}, {"eta_mag": 0.01}):
        with pytest.raises(ProblemSpecError):
            resolve_pcf_physics(_vp_spec(**bad))
    # without S the shear follows the wall speed: U0 = U_wall, S = U_wall / h
    g = {k: v for k, v in VP_GROUPS.items() if k not in ("S", "nu", "eta_mag", "Pm")}
    spec = _vp_spec()
    spec["nondimensional_groups"] = {**g, "U_wall": 2.0}
    p = resolve_pcf_physics(spec)
    assert p["S"] == 2.0 and p["Re_h"] == 1000.0 and p["nu"] == pytest.approx(2e-3)
    # the imposed field from forcing, as a component vector
    spec = _vp_spec()
    del spec["nondimensional_groups"]["B0"]
    spec["forcing"] = {"B0": [0.0, 0.0, 0.03]}
    assert resolve_pcf_physics(spec)["B0"] == pytest.approx(0.03)


def test_vector_potential_spec_rules():
    from jaxfun_amd.runner import magnetic_bc, vector_potential_divergence_guard, vp_time_integrator

    spec = _vp_spec()
    assert magnetic_bc(spec) == "conducting"
    spec["boundary_conditions"] = {"magnetic": {"type": "insulating"}}
    assert magnetic_bc(spec) == "insulating"
    spec["boundary_conditions"] = {"magnetic": "pseudo_vacuum"}
    assert magnetic_bc(spec) == "pseudo_vacuum"
    assert vector_potential_divergence_guard(spec) == 1e-12
    spec["expected_oracle"] = {"divergence_b_guard_l2": {"smoke": 1e-10, "production": 1e-9}}
    spec["resolution"].update({"Nx": 9, "Ny": 8, "Nz": 8})
    assert vector_potential_divergence_guard(spec) == 1e-10
    spec["resolution"].update({"Nx": 17, "Ny": 16, "Nz": 16})
    with pytest.raises(ValueError, match="start"):
        vector_potential_divergence_guard(spec)
    spec["expected_oracle"] = {"divergence_b_guard_l2": -1.0}
    with pytest.raises(ValueError):
        vector_potential_divergence_guard(spec)
    assert vp_time_integrator(spec) == "SBDF3"
    spec["time"]["integrator"] = "IMEXRK3"
    with pytest.raises(ProductionOracleNotImplementedError):
        vp_time_integrator(spec)


def test_divergence_guard_keys_and_limits():
    """``_is_divergence_key`` / ``_raise_on_divergence_drift``
    (``production/oracles.py:3558-3612``): divergence keys by name, magnetic
    ones against the configured ceiling, the rest against 1e-2."""
    from jaxfun_amd.runner import _is_divergence_key, raise_on_divergence_drift

    for key in ("divL2", "divB_L2", "divergence_b_l2", "divergence_u_l2", "continuity_residual", "diag.divu"):
        assert _is_divergence_key(key), key
    for key in ("divergence_e_l2", "divergence_a_l2", "Epert", "mean_bx_trace"):
        assert not _is_divergence_key(key), key
    raise_on_divergence_drift({"divL2": 1e-3, "divB_L2": 1e-13, "divergence_e_l2": 5.0}, t=0.0, tstep=0,
                              magnetic_limit=1e-12)
    with pytest.raises(FloatingPointError, match="divB_L2"):
        raise_on_divergence_drift({"divB_L2": 1e-11}, t=0.0, tstep=3, magnetic_limit=1e-12)
    with pytest.raises(FloatingPointError, match="divL2"):
        raise_on_divergence_drift({"divL2": 0.05}, t=0.0, tstep=3, magnetic_limit=1e-12)
    with pytest.raises(FloatingPointError):
        raise_on_divergence_drift({"divL2": float("nan")}, t=0.0, tstep=3)


def test_budget_and_correlation_estimators():
    """``energy_budget_residual`` vanishes on rows that satisfy
    d(E/2)/dt = S V <stress> - dissipation exactly (linear in time, so the
    midpoint rule is exact) and sees a wrong-sign stress; ``correlation_time``
    of white noise is near one row spacing and of a slow ramp much longer."""
    from jaxfun_amd.runner import correlation_time, effective_independent_samples, energy_budget_residual

    S, V = 1.5, 8.0
    rows = []
    for k in range(6):
        t = 0.1 * k
        stress, dk, dm = 0.01 + 0.002 * t, 0.03, 0.005 + 0.001 * t
        rows.append({"t": t, "total_stress": stress, "dissipation_kinetic": dk, "dissipation_magnetic": dm})
    e = 1.0
    rows[0]["total_energy"] = e
    for a, b in zip(rows[:-1], rows[1:]):
        rhs = S * V * 0.5 * (a["total_stress"] + b["total_stress"]) - 0.5 * (
            a["dissipation_kinetic"] + a["dissipation_magnetic"] + b["dissipation_kinetic"] + b["dissipation_magnetic"])
        e += 2.0 * rhs * (b["t"] - a["t"])
        b["total_energy"] = e
    assert energy_budget_residual(rows, shear_rate=S, volume=V) < 1e-13
    flipped = [dict(r, total_stress=-r["total_stress"]) for r in rows]
    assert energy_budget_residual(flipped, shear_rate=S, volume=V) > 0.5
    rng = np.random.default_rng(0)
    noise = [{"t": 0.1 * k, "total_stress": float(v)} for k, v in enumerate(rng.standard_normal(400))]
    tau = correlation_time(noise)
    assert 0.05 < tau < 0.2
    ramp = [{"t": 0.1 * k, "total_stress": math.sin(0.01 * k)} for k in range(400)]
    assert correlation_time(ramp) > 10 * tau
    assert correlation_time(noise[:7]) is None
    assert effective_independent_samples(noise, correlation=tau) == pytest.approx(39.9 / tau)


def test_mhd_payload_round_trip(tmp_path):
    """The MHDState payload (``production/oracles.py:3872-3900``) through the
    container: flow leaves under ``flow_*`` with a float32 ``flow_have_old``,
    the A levels as written; the state read back equals the state written."""
    from jaxfun_amd.mhd import MHDState
    from jaxfun_amd.runner import mhd_checkpoint_payload, mhd_state_from_checkpoint_payload

    rng = np.random.default_rng(3)
    c = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)  # noqa: E731
    lv = lambda: (c(5, 8, 8), c(7, 8, 8), c(7), c(7))  # noqa: E731
    flow = KMMState(u=(c(5, 8, 8), c(7, 8, 8), c(7, 8, 8)), g=c(7, 8, 8), nonlinear_old=lv(), nonlinear_older=lv(),
                    solution_old=lv(), solution_older=lv(), history_steps=2.0, have_old=1.0, previous_dt=0.005)
    A = lambda: (c(7, 8, 8), c(9, 8, 8), c(9, 8, 8))  # noqa: E731
    st = MHDState(flow=flow, A=A(), nonlinear_A_old=A(), nonlinear_A_older=A(), A_old=A(), A_older=A())
    p = mhd_checkpoint_payload(st)
    assert p["flow_have_old"].dtype == np.float32
    assert set(p) == {"flow_u", "flow_g", "A", "flow_nonlinear_old", "flow_have_old", "flow_previous_dt",
                      "flow_nonlinear_older", "flow_solution_old", "flow_solution_older", "flow_history_steps",
                      "nonlinear_A_old", "nonlinear_A_older", "A_old", "A_older"}
    write_checkpoint(tmp_path / "m.npz", {"state": p}, t=0.03, tstep=6,
                     attrs={"state_kind": "pcf_vector_potential_mhd_saturation"})
    back = mhd_state_from_checkpoint_payload(read_checkpoint(tmp_path / "m.npz").fields["state"])
    for name in ("u", "nonlinear_old", "nonlinear_older", "solution_old", "solution_older"):
        for x, y in zip(getattr(back.flow, name), getattr(st.flow, name)):
            assert np.array_equal(x, y)
    assert np.array_equal(back.flow.g, st.flow.g)
    assert (back.flow.history_steps, back.flow.have_old, back.flow.previous_dt) == (2.0, 1.0, 0.005)
    for name in ("A", "nonlinear_A_old", "nonlinear_A_older", "A_old", "A_older"):
        for x, y in zip(getattr(back, name), getattr(st, name)):
            assert np.array_equal(x, y)


def test_tc_spec_helpers():
    """Host logic of the Taylor-Couette runner: kz_mode and the physics groups."""
    from jaxfun_amd.runner import _tc_physics, tc_kz_mode

    spec = {"mode": {"axial_wavenumber": 3.13, "azimuthal_wavenumber": 0}, "resolution": {"Nr": 40, "Nz": 8},
            "nondimensional_groups": {"Omega1": 1.0, "R1": 1.0, "R2": 2.0, "Re": 100.0}}
    assert tc_kz_mode(spec, 2.0074074463832545) == 1
    assert _tc_physics(spec) == {"nu": 0.01, "eta": 0.01, "B0": 0.0}
    spec["mode"]["azimuthal_wavenumber"] = 1
    with pytest.raises(NotImplementedError):
        tc_kz_mode(spec, 2.0)
    spec["resolution"]["Ntheta"] = 4
    assert tc_kz_mode(spec, 2.0074074463832545) == 1
    spec["nondimensional_groups"].update(Rm=500.0, B0=0.1)
    assert _tc_physics(spec) == pytest.approx({"nu": 0.01, "eta": 0.002, "B0": 0.1})
