"""GPU tests of the runner path: the reference runner's fluctuation scalars and
physical velocity from the device, a spec-driven run against the oracle, and
exact resume from the checkpoint payload.

Bars: physical fields within 1e-12 relative (max |diff| / max |oracle|) of the
FP64 oracle; energies and rms scalars within 1e-10 relative, wall values
within 1e-13 absolute (as the published diagnostics), the divergence (round-off
of a divergence-free field) within 1e-13 absolute; resumed runs bitwise
equal to uninterrupted ones (``tests/couette/test_pcf_sbdf3_jax.py:170-186``
and ``tests/production/test_pcf_cnab2_restart.py:29-52`` of the reference hold
the same bar).
"""

import math

import numpy as np
import pytest

from jaxfun_amd import PlaneCouetteFluctuationHIP, PlaneCouetteMHDHIP
from jaxfun_amd.ic import seeded_state
from jaxfun_amd.pcf import KMMState
from jaxfun_amd.runner import (
    checkpoint_payload,
    materialize_resolution_tier,
    read_checkpoint,
    run_pcf_fluctuation_saturation,
    state_from_checkpoint_payload,
    write_checkpoint,
)

from tests.helpers import mhd_oracle_for, oracle_for, to_mhd_oracle, to_oracle

pytestmark = pytest.mark.gpu

WALL_KEYS = ("wall_shear_lower", "wall_shear_upper", "wall_velocity_lower", "wall_velocity_upper", "mean_shear")


def make(N, precision="fp64", **kw):
    args = dict(Re=400.0, dt=1e-3, perturbation_amplitude=0.05, time_integrator="SBDF3",
                nonlinear_form="rotational", coefficient_path="optimized")
    args.update(kw)
    return PlaneCouetteFluctuationHIP(N=N, precision=precision, **args)


def assert_scalars(got, want, rel=1e-10):
    assert set(got) == set(want)
    for k, w in want.items():
        if k in WALL_KEYS:
            assert abs(got[k] - w) < 1e-13 * max(1.0, abs(w)), (k, got[k], w)
        elif k == "divergence_l2":  # round-off of a divergence-free field
            assert abs(got[k] - w) < 1e-13, (k, got[k], w)
        else:
            assert math.isclose(got[k], w, rel_tol=rel, abs_tol=1e-300), (k, got[k], w)


def rel_field(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def assert_bitwise(a: KMMState, b: KMMState):
    for name in ("u", "nonlinear_old", "nonlinear_older", "solution_old", "solution_older"):
        x, y = getattr(a, name), getattr(b, name)
        assert (x is None) == (y is None), name
        for p, q in zip(x or (), y or ()):
            assert np.array_equal(np.asarray(p), np.asarray(q)), name
    assert np.array_equal(a.g, b.g)
    assert (a.history_steps, a.have_old, a.previous_dt) == (b.history_steps, b.have_old, b.previous_dt)


@pytest.mark.parametrize("N", [(17, 16, 16), (33, 32, 32), (24, 16, 32)])
def test_fluctuation_scalars_and_velocity_match_oracle(N):
    solver = make(N)
    oracle = oracle_for(solver)
    s = seeded_state(solver, seed=1, energy=1e-3, K=4)
    s = solver.solve(s, 5)
    so = to_oracle(s)
    got, want = solver.fluctuation_scalars(s), oracle.fluctuation_scalars(so)
    print(N, {k: abs(got[k] - want[k]) for k in want}, flush=True)
    assert_scalars(got, want)
    up, uo = solver._backward_velocity(s.u), oracle.backward_velocity(so)
    for a, b in zip(up, uo):
        assert a.shape == b.shape and rel_field(a, np.real(b)) < 1e-12
        assert float(np.max(np.abs(np.imag(b)))) < 1e-13 * float(np.max(np.abs(b)))
    e = solver.velocity_energies(s)
    V = 2.0 * solver.Ly * solver.Lz
    assert math.isclose(e[1], want["streak_rms"] ** 2 * V, rel_tol=1e-10)
    assert math.isclose(e[0] + e[2], want["roll_rms"] ** 2 * V, rel_tol=1e-10)


def test_backward_velocity_of_a_foreign_field():
    """A velocity the solver does not hold is loaded for the transform and gives
    the oracle's physical field; the resident state is unchanged afterwards for
    a state passed explicitly."""
    solver = make((17, 16, 16))
    oracle = oracle_for(solver)
    s = solver.solve(solver.initial_state(), 3)
    foreign = seeded_state(solver, seed=4, energy=1e-2, K=4)
    u = tuple(np.array(c) for c in foreign.u)
    got = solver._backward_velocity(u)
    want = oracle.backward_velocity(to_oracle(foreign))
    for a, b in zip(got, want):
        assert rel_field(a, np.real(b)) < 1e-12
    assert_scalars(solver.fluctuation_scalars(s), oracle.fluctuation_scalars(to_oracle(s)))


def test_mixed_precision_scalars_close_to_fp64():
    a, b = make((33, 32, 32)), make((33, 32, 32), precision="mixed")
    s = seeded_state(a, seed=2, energy=1e-3, K=4)
    fa, fb = a.fluctuation_scalars(a.solve(s, 4)), b.fluctuation_scalars(b.solve(s, 4))
    for k in ("kinetic_energy", "total_kinetic_energy", "streak_rms", "roll_rms"):
        assert math.isclose(fa[k], fb[k], rel_tol=1e-5), (k, fa[k], fb[k])


def test_mhd_physical_fields_match_oracle():
    solver = PlaneCouetteMHDHIP(N=(17, 16, 16), Re=200.0, Rm=200.0, dt=1e-3, perturbation_amplitude=0.05,
                                magnetic_amplitude=0.05, time_integrator="SBDF3", nonlinear_form="rotational")
    oracle = mhd_oracle_for(solver)
    s = solver.solve(solver.initial_state(), 3)
    so = to_mhd_oracle(s)
    phys = solver._physical(s)
    assert phys.shape == (6, 17, 16, 16)
    want = list(oracle.backward_velocity(so.flow)) + list(oracle.backward_B(oracle.B_from_A(so.A), padded=False))
    for f in range(6):
        err = rel_field(phys[f], np.real(want[f]))
        print("mhd field", f, err, flush=True)
        assert err < 1e-12


SPEC = {
    "problem_id": "pcf_fluct_re400",
    "nondimensional_groups": {"Re": 400.0, "U_wall": 1.0},
    "domain": {"x": [-1.0, 1.0], "y_period": 4 * math.pi, "z_period": 2 * math.pi},
    "resolution": {"start": {"Nx": 17, "Ny": 16, "Nz": 16}, "production": {"Nx": 64, "Ny": 128, "Nz": 64},
                   "dealias": {"x": 1.0, "y": 1.5, "z": 1.5}, "family": "C"},
    "time": {"integrator": "SBDF3", "dt": 1e-3, "final_time": 0.024},
    "initial_condition": {"type": "finite_amplitude_fluctuation", "amplitude": 0.1},
}


def test_run_matches_oracle_and_resumes_exactly(tmp_path):
    spec = materialize_resolution_tier(SPEC, "start")
    rows = []
    out = run_pcf_fluctuation_saturation(spec, out_dir=tmp_path / "a", checkpoint_every=12, diagnostics_every=6,
                                         on_row=rows.append)
    assert [r["t"] for r in rows] == pytest.approx([0.006, 0.012, 0.018, 0.024], abs=1e-15)
    series = out["time_series"]
    assert [r["t"] for r in series] == pytest.approx([0.0, 0.006, 0.012, 0.018, 0.024], abs=1e-15)
    assert "growth_rate" in series[-1] and "streak_rms" in series[-1]
    sc = out["scalars"]
    assert sc["stationarity_window_samples"] == 5 and sc["stationarity_key"] == "kinetic_energy"

    solver = PlaneCouetteFluctuationHIP(N=(17, 16, 16), domain=((-1.0, 1.0), (0.0, 4 * math.pi), (0.0, 2 * math.pi)),
                                        Re=400.0, dt=1e-3, perturbation_amplitude=0.1, time_integrator="SBDF3",
                                        nonlinear_form="rotational")
    oracle = oracle_for(solver)
    s0 = oracle.initial_state()
    e0 = oracle.fluctuation_scalars(s0)["kinetic_energy"]
    want = oracle.fluctuation_scalars(oracle.solve(s0, 24))
    assert_scalars({k: sc[k] for k in want}, want)
    growth = 0.5 * math.log(want["kinetic_energy"] / e0) / 0.024
    assert math.isclose(sc["growth_rate"], growth, rel_tol=1e-9)
    assert math.isclose(sc["energy_growth_factor"], want["kinetic_energy"] / e0, rel_tol=1e-10)
    assert sc["saturation_check_passed"] == (sc["energy_growth_factor"] > 2.0)

    path = tmp_path / "a" / "checkpoints" / "checkpoints.npz"
    assert read_checkpoint(path).tstep == 24
    rec = read_checkpoint(path, step=12)
    assert rec.attrs["state_kind"] == "pcf_fluctuation_saturation" and rec.t == pytest.approx(0.012)
    resumed = run_pcf_fluctuation_saturation(spec, out_dir=tmp_path / "b", checkpoint_every=12, diagnostics_every=6,
                                             resume_checkpoint=rec)
    assert_bitwise(resumed["final_state"], out["final_state"])
    assert [r["t"] for r in resumed["time_series"]] == pytest.approx([0.0, 0.018, 0.024], abs=1e-15)
    # the device energy sums end in cross-block atomic adds: one ulp of order noise
    assert math.isclose(resumed["scalars"]["kinetic_energy"], sc["kinetic_energy"], rel_tol=1e-14)
    with pytest.raises(ValueError, match="state_kind"):
        bad = type(rec)(fields=rec.fields, t=rec.t, tstep=rec.tstep, attrs={"state_kind": "other"})
        run_pcf_fluctuation_saturation(spec, resume_checkpoint=bad)


def test_sbdf3_checkpoint_preserves_two_level_history_exactly(tmp_path):
    """``tests/couette/test_pcf_sbdf3_jax.py:170-186`` of the reference on the
    hydrodynamic solver, through the checkpoint file."""
    parent = make((17, 16, 16))
    checkpoint_state = parent.solve(parent.initial_state(), 4)
    path = tmp_path / "c.npz"
    write_checkpoint(path, {"state": checkpoint_payload(checkpoint_state)}, t=0.004, tstep=4,
                     attrs={"state_kind": "pcf_fluctuation_saturation"})
    restored = state_from_checkpoint_payload(read_checkpoint(path).fields["state"])
    assert int(restored.history_steps) == 2
    expected = parent.solve(checkpoint_state, 2)
    resumed = make((17, 16, 16)).solve(restored, 2)
    assert_bitwise(resumed, expected)


def test_cnab2_checkpoint_preserves_variable_step_history_exactly():
    """``tests/production/test_pcf_cnab2_restart.py:29-52`` of the reference on
    the hydrodynamic CNAB2 solver: a restart at a new dt keeps the AB2 history."""
    parent = make((17, 16, 16), time_integrator="CNAB2")
    checkpoint_state = parent.solve(parent.initial_state(), 3)
    restored = state_from_checkpoint_payload(checkpoint_payload(checkpoint_state))
    parent.set_dt(7.0e-4)
    expected = parent.solve(checkpoint_state, 2)
    actual = make((17, 16, 16), time_integrator="CNAB2", dt=7.0e-4).solve(restored, 2)
    assert actual.have_old == 1.0 and actual.previous_dt == 7.0e-4
    assert_bitwise(actual, expected)


def test_solve_with_cadence_matches_solve_and_calls_back():
    """``solve_with_cadence`` (the loop the reference runner drives): the same
    state as one ``solve``, callbacks at every due step in the reference's
    order, and ``should_stop`` ending the run after the block it returns true."""
    import types

    s0 = make((17, 16, 16)).initial_state()
    cad = types.SimpleNamespace(diagnostics_every=5, snapshot_every=None, checkpoint_every=12)
    seen = []
    out = make((17, 16, 16)).solve_with_cadence(
        s0, 30, cad, block_size=8, on_diagnostics=lambda t, k, d: seen.append(("d", k, t, d["Epert"])),
        on_checkpoint=lambda t, k, st: seen.append(("c", k, t)), should_stop=lambda t, k, st: False)
    assert_bitwise(out, make((17, 16, 16)).solve(s0, 30))
    assert [r[1] for r in seen if r[0] == "d"] == [5, 10, 15, 20, 25, 30]
    assert [r[1] for r in seen if r[0] == "c"] == [12, 24]
    assert all(math.isclose(r[2], r[1] * 1e-3, rel_tol=1e-12) for r in seen)
    stops = []
    early = make((17, 16, 16)).solve_with_cadence(s0, 30, cad, block_size=8,
                                                  should_stop=lambda t, k, st: stops.append(k) or k >= 10)
    assert stops == [5, 10]
    assert_bitwise(early, make((17, 16, 16)).solve(s0, 10))


def test_rollout_cache_info_counts_graph_lookups():
    s = make((17, 16, 16))
    s.load(s.initial_state())
    s.advance(2)
    info = s.rollout_cache_info()
    assert (info.live_entries, info.hits, info.misses, info.max_entries) == (0, 0, 0, -1)
    s.advance_steady(24, graph_block=12)  # one lookup per replay: captured, then served
    s.advance_steady(12, graph_block=12)  # same ring phase: served from the cache
    info = s.rollout_cache_info()
    assert (info.misses, info.hits, info.live_entries, info.step_counts) == (1, 2, 1, (12,))
    s.advance_steady(5, graph_block=12)
    info = s.rollout_cache_info()
    assert info.misses == 2 and info.step_counts == (5, 12) and info.generation == 0


VP_SPEC = {
    "problem_id": "exp_pcf_mri_vector_potential_smoke",
    "geometry": "pcf",
    "physics": "mri",
    "nondimensional_groups": {"Re": 1000.0, "Rm": 1000.0, "Pm": 1.0, "Omega": 2.0 / 3.0, "S": 1.0,
                              "q_shear": 1.5, "B0": 0.025, "nu": 0.001, "eta_mag": 0.001},
    "boundary_conditions": {"velocity": {"type": "no_slip_shearbox_walls"}, "magnetic": {"type": "conducting"}},
    "domain": {"x": [-1.0, 1.0], "y_period": 4.0, "z_period": 1.0},
    "resolution": {"start": {"Nx": 17, "Ny": 16, "Nz": 16}, "dealias": {"x": 1.0, "y": 1.5, "z": 1.5},
                   "family": "C"},
    "time": {"integrator": "SBDF3", "dt": 0.005, "final_time": 0.06},
    "initial_condition": {"type": "shearpy_channel_modes", "velocity_amplitude": 0.05, "magnetic_amplitude": 0.05},
    "expected_oracle": {"divergence_b_guard_l2": 1e-12},
}


@pytest.mark.parametrize("wall", ["conducting", "insulating"])
def test_vector_potential_run_matches_oracle_and_resumes_exactly(tmp_path, wall):
    """``_run_pcf_vector_potential_mhd_saturation`` on the HIP rotating-shear
    solvers: the final scalars equal the oracle's after the same steps, every
    cadence row closes the shearing-box energy budget to the level the oracle's
    own trajectory does, and a resume from the step-6 checkpoint is bitwise
    the uninterrupted run."""
    from jaxfun_amd.runner import run_pcf_vector_potential_mhd_saturation
    from oracle.mhd import RotatingShearMHDOracle
    from oracle.mhd_insulating import InsulatingMHDOracle

    spec = materialize_resolution_tier(dict(VP_SPEC, boundary_conditions={"magnetic": {"type": wall}}), "start")
    print("case vector-potential run", wall, flush=True)
    rows = []
    out = run_pcf_vector_potential_mhd_saturation(spec, out_dir=tmp_path / "a", checkpoint_every=6,
                                                  diagnostics_every=1, on_row=rows.append)
    assert len(rows) == 12 and rows[-1]["t"] == pytest.approx(0.06)
    sc = out["scalars"]
    assert sc["magnetic_bc"] == wall and sc["representation"] == "vector_potential"
    cls = InsulatingMHDOracle if wall == "insulating" else RotatingShearMHDOracle
    o = cls(N=(17, 16, 16), domain=((-1.0, 1.0), (0.0, 4.0), (0.0, 1.0)), Re=1000.0, Rm=1000.0, omega=2.0 / 3.0,
            shear_rate=1.0, background_b=(0.0, 0.0, 0.025), dt=0.005, perturbation_amplitude=0.05,
            magnetic_amplitude=0.05)
    from jaxfun_amd.runner import curl_solver_from_spec

    s0 = curl_solver_from_spec(spec).initial_state()
    so = to_mhd_oracle(s0)
    orows = []
    for k in range(12):
        so = o.step(so)
        d = o.extended_diagnostics(so)
        orows.append({"t": 0.005 * (k + 1), "total_energy": d["Epert"] + d["Emag"],
                      "total_stress": d["total_stress"], "dissipation_kinetic": d["dissipation_kinetic"],
                      "dissipation_magnetic": d["dissipation_magnetic"]})
    want = o.extended_diagnostics(so)
    for key, ref in (("kinetic_energy", want["Epert"]), ("magnetic_energy", want["Emag"]),
                     ("maxwell_stress_xy", want["maxwell_stress"]), ("reynolds_stress", want["reynolds_stress"]),
                     ("dissipation_kinetic", want["dissipation_kinetic"]),
                     ("channel_kz1_total_rms", want["channel_kz1_total_rms"])):
        got = sc[key] if key in sc else rows[-1][key]
        assert math.isclose(got, ref, rel_tol=1e-10), (key, got, ref)
    for key in ("electric_total_l2", "divergence_a_l2", "mean_bz", "mag_energy_fluct"):
        assert math.isclose(sc[key], want[key], rel_tol=1e-10, abs_tol=1e-15), (key, sc[key], want[key])
    from jaxfun_amd.runner import energy_budget_residual

    ob = energy_budget_residual(orows, shear_rate=1.0, volume=8.0)
    gb = energy_budget_residual(rows, shear_rate=1.0, volume=8.0)
    print(wall, "energy budget residual", gb, "oracle", ob, flush=True)
    assert math.isclose(gb, ob, rel_tol=1e-6) and gb < 1e-2
    assert sc["energy_budget_residual"] == pytest.approx(gb, rel=1e-12)
    assert sc["divergence_b_l2"] < 1e-12
    rec = read_checkpoint(tmp_path / "a" / "checkpoints" / "checkpoints.npz", step=6)
    assert rec.attrs["state_kind"] == "pcf_vector_potential_mhd_saturation"
    resumed = run_pcf_vector_potential_mhd_saturation(spec, out_dir=tmp_path / "b", checkpoint_every=6,
                                                      diagnostics_every=1, resume_checkpoint=rec)
    a, b = out["final_state"], resumed["final_state"]
    assert_bitwise(a.flow, b.flow)
    for name in ("A", "nonlinear_A_old", "nonlinear_A_older", "A_old", "A_older"):
        for p, q in zip(getattr(a, name), getattr(b, name)):
            assert np.array_equal(np.asarray(p), np.asarray(q)), name
