"""Problem-spec runner and checkpoint wire format for the HIP plane-Couette solvers.

``run_pcf_fluctuation_saturation`` is the counterpart of the reference
runner's ``_run_pcf_fluctuation_saturation`` (``production/oracles.py:1442-1573``):
the solver is built from a problem spec, stepped in blocks with diagnostics
rows at a cadence, checkpointed with the reference's payload
(``_checkpoint_payload``, ``production/oracles.py:3872-3962``) and resumable
exactly from it (``_state_from_checkpoint_payload``, ``:327-357``); the final
scalars carry the keys of ``_pcf_fluctuation_scalars`` (``:4566-4597``) plus the
growth-rate, saturation and stationarity entries (``:1531-1573``, ``:531-548``,
``:3347-3380``, ``:4359-4369``).

Two deliberate differences, both spec knobs: ``numerics.nonlinear_form`` and
``numerics.coefficient_path`` select the solver's form and path (the reference
runner passes neither and so gets its default gradient form; this build runs
the rotational form only, so the runner defaults to it), and the checkpoint
container is a NumPy ``.npz`` archive with the reference's tree layout
(``src/jaxfun/io/__init__.py:50-96,189-244`` write HDF5 through h5py, which is
not part of this image): every dataset sits at
``checkpoints/<step>/fields/<name>/<child>...`` and the tree kinds and
attributes travel in a JSON entry.
"""

from __future__ import annotations

import copy
import json
import math
import warnings
from dataclasses import dataclass
from pathlib import Path
from typing import Any

import numpy as np

from .pcf import KMMState, PlaneCouetteFluctuationHIP, to_host

KMM_INTEGRATORS = frozenset({"IMEXRK222", "IMEXRK3", "SBDF3", "CNAB2"})


class ProductionOracleNotImplementedError(NotImplementedError):
    """A spec the solver family cannot run (``production/oracles.py:30-58``)."""


# ---------------------------------------------------------------- spec parsing
def selected_resolution(spec: dict) -> dict:
    """``_selected_resolution`` (``production/oracles.py:4372-4377``)."""
    resolution = spec["resolution"]
    if any(key in resolution for key in ("N", "Nr", "Nx", "Nz")):
        return resolution
    selected = resolution.get("production", resolution)
    return {**resolution, **selected}


def materialize_resolution_tier(spec: dict, resolution_tier: str | None) -> dict:
    """``_materialize_resolution_tier`` (``production/adapters.py:150-163``): a
    copy of ``spec`` whose resolution has the named tier merged over it."""
    data = copy.deepcopy(spec)
    if resolution_tier is None:
        return data
    resolution = data["resolution"]
    selected = resolution.get(resolution_tier)
    if not isinstance(selected, dict):
        raise ValueError(f"resolution tier {resolution_tier!r} is not defined for {data.get('problem_id')}")
    data["resolution"] = {**resolution, **selected}
    return data


def padding_factor(resolution: dict) -> tuple:
    """The KMM solver's native (x, y, z) padding from ``resolution.dealias``
    (``native_padding_for_solver`` / ``to_native_padding``,
    ``production/axes.py:69-134``): a semantic {x, y, z} map (every axis
    required), a scalar for every axis, or a legacy positional list read in
    (x, y, z) order (non-uniform ones warn as deprecated); each value >= 1."""
    dealias = resolution.get("dealias", 1.0)
    context = "resolution.dealias[pcf_kmm]"

    def finite(value, label):
        try:
            out = float(value)
        except (TypeError, ValueError) as exc:
            raise ValueError(f"{label} padding must be numeric") from exc
        if not out >= 1.0:
            raise ValueError(f"{label} padding must be >= 1.0, got {out}")
        return out

    if isinstance(dealias, dict):
        missing = [axis for axis in "xyz" if axis not in dealias]
        if missing:
            raise ValueError(f"{context}: semantic padding is missing axis/axes {missing}; "
                             "native order is ('x', 'y', 'z')")
        return tuple(finite(dealias[axis], f"{context}.{axis}") for axis in "xyz")
    if isinstance(dealias, (list, tuple)):
        values = tuple(finite(v, context) for v in dealias)
        if len(values) != 3:
            raise ValueError(f"{context}: expected 3 padding values, got {len(values)}")
        if len(set(values)) > 1:
            warnings.warn(f"{context}: positional padding tuple {values} is deprecated; use a semantic "
                          "{'x': .., 'y': .., 'z': ..} map. Interpreting the tuple in canonical "
                          "('x', 'y', 'z') order.", DeprecationWarning, stacklevel=2)
        return values
    return (finite(dealias, context),) * 3


def kmm_time_integrator(spec: dict) -> str:
    """``_pcf_kmm_time_integrator`` (``production/oracles.py:39-58``)."""
    integrator = str(spec["time"]["integrator"])
    if integrator not in KMM_INTEGRATORS:
        raise ProductionOracleNotImplementedError(
            "PCF hydrodynamic KMM requires a time-stepping integrator "
            f"({', '.join(sorted(KMM_INTEGRATORS))}); got time.integrator={integrator!r}")
    if integrator == "SBDF3" and spec.get("time", {}).get("adaptive_cfl") is not None:
        raise ProductionOracleNotImplementedError(
            "PCF hydrodynamic KMM does not support adaptive_cfl with fixed-step SBDF3")
    return integrator


def steps_from_spec(spec: dict, steps: int | None = None) -> int:
    """``_steps_from_spec`` (``production/oracles.py:4641-4653``)."""
    if steps is not None:
        if steps < 0:
            raise ValueError("steps override must be non-negative")
        return int(steps)
    dt = float(spec["time"]["dt"])
    final_time = float(spec["time"]["final_time"])
    n = int(round(final_time / dt))
    if not math.isclose(n * dt, final_time, rel_tol=1e-12, abs_tol=1e-12):
        raise ValueError("final_time must be an integer multiple of dt for DNS parity runs")
    return n


def solver_from_spec(spec: dict, precision: str = "fp64", device: int = 0) -> PlaneCouetteFluctuationHIP:
    """The solver ``_run_pcf_fluctuation_saturation`` constructs
    (``production/oracles.py:1466-1480``), on the GPU."""
    resolution = selected_resolution(spec)
    domain = (tuple(float(v) for v in spec["domain"]["x"]), (0.0, float(spec["domain"]["y_period"])),
              (0.0, float(spec["domain"]["z_period"])))
    groups = spec["nondimensional_groups"]
    numerics = spec.get("numerics", {})
    return PlaneCouetteFluctuationHIP(
        N=(int(resolution.get("Nx", resolution.get("N", 32))), int(resolution.get("Ny", 64)),
           int(resolution.get("Nz", 32))),
        domain=domain, Re=float(groups["Re"]), U_wall=float(groups.get("U_wall", 1.0)),
        dt=float(spec["time"]["dt"]), family=resolution.get("family", "C"),
        padding_factor=padding_factor(resolution),
        perturbation_amplitude=float(spec["initial_condition"].get("amplitude", 0.1)),
        time_integrator=kmm_time_integrator(spec),
        nonlinear_form=numerics.get("nonlinear_form", "rotational"),
        coefficient_path=numerics.get("coefficient_path", "optimized"),
        precision=precision, device=device)


# ---------------------------------------------------------------- checkpoint payload
def checkpoint_payload(state: KMMState) -> dict:
    """``_checkpoint_payload`` (``production/oracles.py:3872-3962``), KMM branch:
    the state leaves as a tree of tuples, ``have_old`` as float32."""
    payload: dict[str, Any] = {"u": tuple(state.u), "g": state.g}
    if state.nonlinear_old is not None:
        payload.update({"nonlinear_old": tuple(state.nonlinear_old),
                        "have_old": np.asarray(state.have_old, dtype=np.float32),
                        "previous_dt": np.asarray(state.previous_dt)})
    if state.nonlinear_older is not None:
        payload["nonlinear_older"] = tuple(state.nonlinear_older)
    if state.solution_old is not None:
        payload["solution_old"] = tuple(state.solution_old)
    if state.solution_older is not None:
        payload["solution_older"] = tuple(state.solution_older)
    payload["history_steps"] = np.asarray(state.history_steps)
    return payload


def state_from_checkpoint_payload(payload: dict, state_kind: str = "pcf_fluctuation_saturation") -> KMMState:
    """``_state_from_checkpoint_payload`` (``production/oracles.py:327-357``)."""
    if state_kind != "pcf_fluctuation_saturation":
        raise ProductionOracleNotImplementedError(f"state kind {state_kind!r} is not built")

    def group(name):
        return None if name not in payload else tuple(np.asarray(v) for v in payload[name])

    return KMMState(u=tuple(np.asarray(v) for v in payload["u"]), g=np.asarray(payload["g"]),
                    nonlinear_old=group("nonlinear_old"), nonlinear_older=group("nonlinear_older"),
                    solution_old=group("solution_old"), solution_older=group("solution_older"),
                    history_steps=float(np.asarray(payload.get("history_steps", 0.0))),
                    have_old=float(np.asarray(payload.get("have_old", 0.0), dtype=np.float32)),
                    previous_dt=float(np.asarray(payload.get("previous_dt", 0.0))))


@dataclass(frozen=True)
class CheckpointRecord:
    """``CheckpointRecord`` of ``src/jaxfun/io/__init__.py``."""

    fields: dict
    t: float
    tstep: int
    attrs: dict


def _flatten(prefix: str, value, arrays: dict, kinds: dict) -> None:
    if isinstance(value, dict):
        kinds[prefix] = {"kind": "dict", "keys": list(value)}
        for key, item in value.items():
            _flatten(f"{prefix}/{key}", item, arrays, kinds)
    elif isinstance(value, (tuple, list)):
        kinds[prefix] = {"kind": "tuple" if isinstance(value, tuple) else "list", "length": len(value)}
        for idx, item in enumerate(value):
            _flatten(f"{prefix}/{idx}", item, arrays, kinds)
    else:
        arrays[prefix] = np.asarray(value)


def _unflatten(prefix: str, arrays: dict, kinds: dict):
    meta = kinds.get(prefix)
    if meta is None:
        return arrays[prefix]
    if meta["kind"] == "dict":
        return {key: _unflatten(f"{prefix}/{key}", arrays, kinds) for key in meta["keys"]}
    items = [_unflatten(f"{prefix}/{i}", arrays, kinds) for i in range(int(meta["length"]))]
    return tuple(items) if meta["kind"] == "tuple" else items


def write_checkpoint(filename, fields: dict, *, t: float, tstep: int, attrs: dict | None = None) -> None:
    """``write_checkpoint`` (``src/jaxfun/io/__init__.py:189-220``) into an
    .npz archive; later steps are added beside earlier ones and the latest
    step is recorded."""
    path = Path(filename)
    arrays, meta = {}, {"steps": {}, "latest_step": int(tstep)}
    if path.exists():
        with np.load(path, allow_pickle=False) as old:
            meta = json.loads(str(old["__meta__"]))
            arrays = {k: old[k] for k in old.files if k != "__meta__"}
        meta["latest_step"] = int(tstep)
    step = str(int(tstep))
    base = f"checkpoints/{step}/fields"
    for key in [k for k in arrays if k.startswith(f"checkpoints/{step}/")]:
        del arrays[key]
    kinds: dict = {}
    for name, value in fields.items():
        _flatten(f"{base}/{name}", value, arrays, kinds)
    meta["steps"][step] = {"t": float(t), "tstep": int(tstep), "names": list(fields), "kinds": kinds,
                           "attrs": {"t": float(t), "tstep": int(tstep),
                                     **{str(k): v for k, v in (attrs or {}).items()}}}
    arrays["__meta__"] = np.asarray(json.dumps(meta))
    tmp = path.with_name(path.name + ".tmp.npz")
    np.savez(tmp, **arrays)
    tmp.replace(path)


def read_checkpoint(filename, *, step: int | None = None) -> CheckpointRecord:
    """``read_checkpoint`` (``src/jaxfun/io/__init__.py:223-244``)."""
    with np.load(Path(filename), allow_pickle=False) as f:
        meta = json.loads(str(f["__meta__"]))
        key = str(int(meta["latest_step"] if step is None else step))
        rec = meta["steps"][key]
        base = f"checkpoints/{key}/fields"
        arrays = {k: f[k] for k in f.files if k.startswith(base + "/")}
        fields = {name: _unflatten(f"{base}/{name}", arrays, rec["kinds"]) for name in rec["names"]}
    return CheckpointRecord(fields=fields, t=float(rec["t"]), tstep=int(rec["tstep"]), attrs=rec["attrs"])


# ---------------------------------------------------------------- scalars
def growth_rate_from_energy(e0: float, e1: float, steps: int, dt: float) -> float:
    """``_growth_rate_from_energy`` (``production/oracles.py:4359-4369``)."""
    elapsed = int(steps) * float(dt)
    if elapsed <= 0.0:
        raise ValueError("growth-rate diagnostics require at least one DNS step")
    e0, e1 = float(e0), float(e1)
    if not (math.isfinite(e0) and math.isfinite(e1)) or e0 <= 0.0 or e1 <= 0.0:
        if math.isfinite(e0) and e0 > 0.0 and e1 == 0.0:
            return -math.inf
        return math.nan
    return 0.5 * math.log(e1 / e0) / elapsed


def saturation_passed(growth, *, threshold: float, final_energies) -> bool:
    """``_saturation_passed`` (``production/oracles.py:531-548``)."""
    try:
        g = float(growth)
    except (TypeError, ValueError):
        return False
    if not math.isfinite(g) or g <= threshold:
        return False
    for e in final_energies:
        try:
            v = float(e)
        except (TypeError, ValueError):
            return False
        if not math.isfinite(v) or v < 0.0:
            return False
    return True


def stationarity_scalars(rows: list, *, key: str, tolerance: float = 5.0e-2) -> dict:
    """``_stationarity_scalars`` (``production/oracles.py:3347-3380``)."""
    values = [float(r[key]) for r in rows if key in r and math.isfinite(float(r[key]))]
    n = len(values)
    if n < 4:
        return {"stationarity_key": key, "stationarity_window_samples": n,
                "stationarity_relative_tolerance": tolerance, "stationarity_relative_change": None,
                "stationarity_check_passed": None}
    q = max(1, n // 4)
    prev, cur = float(np.mean(values[-2 * q:-q])), float(np.mean(values[-q:]))
    change = abs(cur - prev) / max(abs(prev), abs(cur), 1.0e-300)
    return {"stationarity_key": key, "stationarity_window_samples": n, "stationarity_relative_tolerance": tolerance,
            "stationarity_previous_mean": prev, "stationarity_current_mean": cur,
            "stationarity_relative_change": float(change), "stationarity_check_passed": bool(change <= tolerance)}


def dedupe_time_rows(rows: list) -> list:
    """``_dedupe_time_rows`` (``production/oracles.py:3411-3421``)."""
    out, seen = [], set()
    for row in rows:
        t = float(row.get("t", len(out)))
        if t in seen:
            out[-1] = row
            continue
        seen.add(t)
        out.append(row)
    return out


# ---------------------------------------------------------------- health guards
DIVERGENCE_GUARD_LIMIT = 1.0e-2
ENERGY_RUNAWAY_LIMIT = 1.0e30


def raise_on_unhealthy(diag: dict, *, t: float, tstep: int) -> None:
    """The per-block guards of ``_solve_with_optional_checkpoints``'s
    ``should_stop`` (``production/oracles.py:3255-3282``), on the device
    diagnostics: a non-finite state (``_tree_all_finite``, ``:3615-3625``; a
    non-finite coefficient makes the device energy sums non-finite), the
    energy ceiling (``_raise_on_energy_runaway``, ``:3527-3555``) and the
    divergence limit (``_raise_on_divergence_drift``, ``:3558-3598``)."""
    energies = (diag["Epert"], diag["Etot"])
    if not all(math.isfinite(v) for v in energies):
        raise FloatingPointError(f"nonfinite solver state at tstep={int(tstep)} t={float(t):g}")
    if abs(diag["Etot"]) > ENERGY_RUNAWAY_LIMIT:
        raise FloatingPointError(
            f"energy runaway ceiling exceeded at tstep={int(tstep)} "
            f"t={float(t):g}: E={diag['Etot']:g} > {ENERGY_RUNAWAY_LIMIT:g}")
    div = abs(float(diag["divL2"]))
    if not math.isfinite(div) or div > DIVERGENCE_GUARD_LIMIT:
        raise FloatingPointError(f"divergence guard failed at tstep={int(tstep)} t={float(t):g}: "
                                 f"divL2={div:g} > {DIVERGENCE_GUARD_LIMIT:g}")


# ---------------------------------------------------------------- the run
def run_pcf_fluctuation_saturation(spec: dict, *, steps: int | None = None, out_dir=None,
                                   checkpoint_every: int | None = None, diagnostics_every: int | None = None,
                                   resume_checkpoint: CheckpointRecord | None = None, on_row=None,
                                   precision: str = "fp64", solver: PlaneCouetteFluctuationHIP | None = None) -> dict:
    """``_run_pcf_fluctuation_saturation`` (``production/oracles.py:1442-1573``)
    on the HIP solver.  Steps run in blocks between cadence points with the
    state in device memory (torch tensor leaves; host copies only for
    checkpoints and the returned final state); each block ends with the health
    guards.  With ``out_dir`` and
    ``checkpoint_every`` the state payload goes to
    ``out_dir/checkpoints/checkpoints.npz`` at that cadence and at the end
    (attr ``state_kind``, ``:3167-3192``, ``:3321-3326``), and
    ``resume_checkpoint`` continues from one (``_resume_or_initial_state``,
    ``:268-292``).  ``solver`` reuses a solver built by ``solver_from_spec``
    for the same spec."""
    state_kind = "pcf_fluctuation_saturation"
    for name, value in (("checkpoint_every", checkpoint_every), ("diagnostics_every", diagnostics_every)):
        if value is not None and value <= 0:
            raise ValueError(f"{name} must be positive")
    if checkpoint_every is not None and out_dir is None:
        raise ValueError("out_dir is required when checkpoint, snapshot, or profile output is set")
    solver = solver if solver is not None else solver_from_spec(spec, precision=precision)
    state = solver.initial_state()
    initial = solver.fluctuation_scalars(state)
    tstep0, t0 = 0, 0.0
    if resume_checkpoint is not None:
        kind = str(resume_checkpoint.attrs.get("state_kind"))
        if kind != state_kind:
            raise ValueError(f"resume checkpoint state_kind {kind!r} does not match expected {state_kind!r}")
        payload = resume_checkpoint.fields.get("state")
        if not isinstance(payload, dict):
            raise ValueError("resume checkpoint is missing state payload")
        state = state_from_checkpoint_payload(payload, state_kind)
        tstep0, t0 = int(resume_checkpoint.tstep), float(resume_checkpoint.t)
    target = steps_from_spec(spec, steps)
    if tstep0 > target:
        raise ValueError(f"resume checkpoint step {tstep0} is beyond target step {target}")
    n_steps = target - tstep0
    dt = solver.dt
    cadence = [int(c) for c in (checkpoint_every, diagnostics_every) if c]
    rows = []
    ckpt = Path(out_dir) / "checkpoints" / "checkpoints.npz" if checkpoint_every is not None else None
    if ckpt is not None:
        ckpt.parent.mkdir(parents=True, exist_ok=True)

    def save(t, tstep, st):
        write_checkpoint(ckpt, {"state": checkpoint_payload(to_host(st))}, t=t, tstep=tstep,
                         attrs={"state_kind": state_kind})

    done = 0
    while done < n_steps:
        block = n_steps - done
        for c in cadence:
            block = min(block, c - ((tstep0 + done) % c))
        state = solver.solve_device(state, block)  # the state stays in device memory between blocks
        done += block
        tstep = tstep0 + done
        t = t0 + done * dt
        d = solver.diagnostics(state)
        if diagnostics_every and tstep % diagnostics_every == 0:
            row = {"t": float(t), "kinetic_energy": d["Epert"], "total_kinetic_energy": d["Etot"],
                   "divergence_l2": d["divL2"], "mean_shear": d["mean_shear"]}
            rows.append(row)
            if on_row is not None:
                on_row(row)
        raise_on_unhealthy(d, t=t, tstep=tstep)
        if ckpt is not None and tstep % checkpoint_every == 0:
            save(t, tstep, state)
    if ckpt is not None and (n_steps == 0 or target % checkpoint_every != 0):
        save(t0 + n_steps * dt, target, state)
    final = solver.fluctuation_scalars(state)
    growth = growth_rate_from_energy(initial["kinetic_energy"], final["kinetic_energy"], target, dt)
    elapsed = target * float(spec["time"]["dt"])
    egrowth = final["kinetic_energy"] / initial["kinetic_energy"] if initial["kinetic_energy"] > 0.0 else 0.0
    scalars = {**final, "growth_rate": float(growth), "energy_growth_factor": float(egrowth),
               "saturation_check_passed": saturation_passed(
                   egrowth, threshold=2.0,
                   final_energies=(final["kinetic_energy"], final["total_kinetic_energy"], final["divergence_l2"],
                                   final["streak_rms"], final["roll_rms"]))}
    first = {"t": 0.0, "kinetic_energy": initial["kinetic_energy"],
             "total_kinetic_energy": initial["total_kinetic_energy"], "divergence_l2": initial["divergence_l2"],
             "mean_shear": initial["mean_shear"]}
    last = {"t": elapsed, "kinetic_energy": final["kinetic_energy"],
            "total_kinetic_energy": final["total_kinetic_energy"], "divergence_l2": final["divergence_l2"],
            "mean_shear": final["mean_shear"], "growth_rate": float(growth), "energy_growth_factor": float(egrowth),
            "saturation_check_passed": scalars["saturation_check_passed"],
            "wall_shear_lower": final["wall_shear_lower"], "wall_shear_upper": final["wall_shear_upper"],
            "streak_rms": final["streak_rms"], "roll_rms": final["roll_rms"]}
    series = dedupe_time_rows([first, *rows, last])
    scalars.update(stationarity_scalars(series, key="kinetic_energy"))
    return {"scalars": scalars, "time_series": series, "final_state": to_host(state)}


# ================================================================ vector-potential MHD
VP_INTEGRATORS = frozenset({"IMEXRK222", "SBDF3", "CNAB2"})
VP_STATE_KIND = "pcf_vector_potential_mhd_saturation"
MIN_CORRELATION_SAMPLES = 8
# the electromagnetic keys the runner copies into every row and the final
# scalars (``_PCF_ELECTROMAGNETIC_KEYS``, ``production/oracles.py:1890-1918``)
PCF_ELECTROMAGNETIC_KEYS = (
    "electric_ideal_l2", "electric_resistive_l2", "electric_total_l2", "divergence_e_l2",
    "divergence_e_ideal_l2", "divergence_e_resistive_l2", "divergence_a_l2", "divergence_a_wall_linf",
    "electric_ideal_wall_tangential_linf", "electric_resistive_wall_tangential_linf",
    "electric_wall_tangential_linf", "faraday_mean_bx_tendency", "faraday_mean_by_tendency",
    "faraday_mean_bz_tendency", "mean_bx_trace", "mean_by_trace", "mean_bz_trace", "mean_b_trace_mismatch_linf",
    "mean_ex_ideal", "mean_ey_ideal", "mean_ez_ideal", "mean_ex_resistive", "mean_ey_resistive",
    "mean_ez_resistive", "mean_ex_total", "mean_ey_total", "mean_ez_total",
)


class ProblemSpecError(ValueError):
    """An inconsistent or under-specified problem spec (``production/physics.py``)."""


def magnetic_bc(spec: dict) -> str:
    """``_magnetic_bc`` (``production/oracles.py:2320-2322``)."""
    magnetic = spec.get("boundary_conditions", {}).get("magnetic", "conducting")
    return magnetic.get("type", magnetic) if isinstance(magnetic, dict) else magnetic


def vector_potential_divergence_guard(spec: dict) -> float:
    """``_vector_potential_divergence_guard`` (``production/oracles.py:2325-2352``):
    the div-B ceiling, a scalar or one value per resolution tier."""
    configured = spec.get("expected_oracle", {}).get("divergence_b_guard_l2", 1.0e-12)
    if isinstance(configured, dict):
        resolution = selected_resolution(spec)
        tier = "production"
        for candidate in ("smoke", "start", "production"):
            block = spec.get("resolution", {}).get(candidate)
            if isinstance(block, dict) and all(resolution.get(k) == v for k, v in block.items()):
                tier = candidate
                break
        if tier not in configured:
            raise ValueError(f"expected_oracle.divergence_b_guard_l2 is missing the {tier!r} resolution tier")
        configured = configured[tier]
    limit = float(configured)
    if not math.isfinite(limit) or limit <= 0.0:
        raise ValueError("divergence_b_guard_l2 must be finite and positive")
    return limit


def _finite(value, label: str) -> float:
    try:
        out = float(value)
    except (TypeError, ValueError) as exc:
        raise ProblemSpecError(f"{label} must be numeric") from exc
    if not math.isfinite(out):
        raise ProblemSpecError(f"{label} must be finite")
    return out


def _positive(value: float, label: str) -> float:
    if not value > 0.0:
        raise ProblemSpecError(f"{label} must be positive, got {value}")
    return value


def _b0_magnitude(value, label: str) -> float:
    if isinstance(value, (list, tuple)):
        return math.sqrt(sum(_finite(v, label) ** 2 for v in value))
    amplitude = _finite(value, label)
    if amplitude < 0.0:
        raise ProblemSpecError(f"{label} is a field amplitude and must be nonnegative; "
                               "use an explicit component vector for signed direction")
    return amplitude


def _resolve_coefficient(supplied, group, scale: float, coeff_label: str, group_label: str):
    if supplied is None and group is None:
        raise ProblemSpecError(f"must supply {coeff_label} or {group_label}")
    if supplied is not None and group is not None:
        coeff, grp = _positive(supplied, coeff_label), _positive(group, group_label)
        derived = scale / coeff
        if not math.isclose(grp, derived, rel_tol=1.0e-6, abs_tol=1.0e-9):
            raise ProblemSpecError(f"over-specified inconsistent inputs: {group_label}={grp:g} but "
                                   f"{coeff_label}={coeff:g} implies {group_label}={derived:g} "
                                   f"(scale |S| h^2 = {scale:g})")
        return coeff, grp
    if supplied is not None:
        coeff = _positive(supplied, coeff_label)
        return coeff, scale / coeff
    grp = _positive(group, group_label)
    return scale / grp, grp


def resolve_pcf_physics(spec: dict) -> dict:
    """``resolve_physics`` (``production/physics.py:408-515``) for plane
    Couette: half gap h, shear rate S (given, or U_wall / h), rotation Omega,
    the imposed field B0 (``_single_b0``, ``:136-162``), nu / Re_h and
    eta / Rm_h on the scale |S| h^2 with the consistency checks, Pm."""
    groups = spec["nondimensional_groups"]
    x0, x1 = (_finite(v, "domain.x") for v in spec["domain"]["x"])
    h = _positive(0.5 * (x1 - x0), "half-gap h")
    omega = _finite(groups.get("Omega", 0.0), "Omega")
    forcing = spec.get("forcing", {})
    group_b0 = groups.get("B0")
    forcing_b0 = forcing.get("B0") if isinstance(forcing, dict) else None
    if group_b0 is not None and forcing_b0 is not None:
        gb, fb = _b0_magnitude(group_b0, "nondimensional_groups.B0"), _b0_magnitude(forcing_b0, "forcing.B0")
        if not math.isclose(gb, fb, rel_tol=1.0e-9, abs_tol=1.0e-12):
            raise ProblemSpecError(f"imposed field disagrees: nondimensional_groups.B0={gb:g} vs "
                                   f"forcing.B0={fb:g}; state a single value")
        b0 = gb
    elif group_b0 is not None:
        b0 = _b0_magnitude(group_b0, "nondimensional_groups.B0")
    elif forcing_b0 is not None:
        b0 = _b0_magnitude(forcing_b0, "forcing.B0")
    else:
        b0 = 0.0
    if groups.get("S") is not None:
        S = _finite(groups["S"], "S")
        U0 = abs(S) * h
    else:
        U_wall, Re_in, nu_in = groups.get("U_wall"), groups.get("Re"), groups.get("nu")
        if U_wall is not None:
            U0 = abs(_finite(U_wall, "U_wall"))
        elif Re_in is not None and nu_in is not None:
            U0 = _positive(_finite(Re_in, "Re"), "Re") * _finite(nu_in, "nu") / h
        else:
            U0 = 1.0
        S = U0 / h
    scale = U0 * h
    nu_in, Re_in = groups.get("nu"), groups.get("Re", groups.get("Re_h"))
    nu, Re_h = _resolve_coefficient(None if nu_in is None else _finite(nu_in, "nu"),
                                    None if Re_in is None else _finite(Re_in, "Re"), scale, "nu", "Re_h")
    out = {"h": h, "S": S, "Omega": omega, "B0": b0, "nu": nu, "Re_h": Re_h, "U0": U0,
           "eta": None, "Rm_h": None, "Pm": None}
    if spec.get("physics") in {"mhd", "mri"}:
        eta_in, Rm_in = groups.get("eta_mag", groups.get("eta")), groups.get("Rm", groups.get("Rm_h"))
        eta, Rm_h = _resolve_coefficient(None if eta_in is None else _finite(eta_in, "eta"),
                                         None if Rm_in is None else _finite(Rm_in, "Rm"), scale, "eta", "Rm_h")
        Pm = nu / eta
        if groups.get("Pm") is not None:
            pm_in = _finite(groups["Pm"], "Pm")
            if not math.isclose(pm_in, Pm, rel_tol=1.0e-6, abs_tol=1.0e-9):
                raise ProblemSpecError(f"over-specified inconsistent Pm={pm_in:g}; nu/eta={Pm:g}")
        out.update(eta=eta, Rm_h=Rm_h, Pm=Pm)
    return out


def vp_time_integrator(spec: dict) -> str:
    """``_pcf_kmm_time_integrator`` with the vector-potential family's set
    (``production/oracles.py:35,1822-1826``)."""
    integrator = str(spec["time"]["integrator"])
    if integrator not in VP_INTEGRATORS:
        raise ProductionOracleNotImplementedError(
            "PCF vector-potential MHD/MRI requires a time-stepping integrator "
            f"({', '.join(sorted(VP_INTEGRATORS))}); got time.integrator={integrator!r}")
    if integrator == "SBDF3" and spec.get("time", {}).get("adaptive_cfl") is not None:
        raise ProductionOracleNotImplementedError(
            "PCF vector-potential MHD/MRI does not support adaptive_cfl with fixed-step SBDF3")
    return integrator


def curl_solver_from_spec(spec: dict, precision: str = "fp64", device: int = 0):
    """``_curl_solver_from_spec`` (``production/oracles.py:1819-1874``): the
    conducting or insulating rotating-shear solver on the GPU."""
    from .mhd import PlaneCouetteMRIShearpyHIP, PlaneCouetteMRIShearpyInsulatingHIP

    integrator = vp_time_integrator(spec)
    bc = magnetic_bc(spec)
    classes = {"conducting": PlaneCouetteMRIShearpyHIP, "insulating": PlaneCouetteMRIShearpyInsulatingHIP}
    if bc not in classes:
        raise ProductionOracleNotImplementedError(
            "the vector-potential (curl) PCF-MRI family is wired for conducting or insulating walls, "
            f"got {bc!r}; pseudo-vacuum runs use the primitive-b family")
    resolution = selected_resolution(spec)
    physics = resolve_pcf_physics(spec)
    initial = spec["initial_condition"]
    x0, x1 = (float(v) for v in spec["domain"]["x"])
    numerics = spec.get("numerics", {})
    seed_default = "sinusoidal_bz_x" if initial.get("type") == "mri_znf_sinusoidal_bz" else "ax_yz"
    return classes[bc](
        N=(int(resolution.get("Nx", resolution.get("N", 17))), int(resolution.get("Ny", 16)),
           int(resolution.get("Nz", 16))),
        domain=((x0, x1), (0.0, float(spec["domain"]["y_period"])), (0.0, float(spec["domain"]["z_period"]))),
        Re=physics["Re_h"], Rm=physics["Rm_h"] if physics["Rm_h"] is not None else physics["Re_h"],
        omega=physics["Omega"], shear_rate=physics["S"], background_b=(0.0, 0.0, physics["B0"]),
        dt=float(spec["time"]["dt"]), family=resolution.get("family", "C"),
        padding_factor=padding_factor(resolution),
        perturbation_amplitude=float(initial.get("velocity_amplitude", 0.05)),
        magnetic_amplitude=float(initial.get("magnetic_amplitude", 0.0)),
        magnetic_seed=str(initial.get("magnetic_seed", seed_default)),
        solenoidal_velocity_seed=bool(initial.get("solenoidal_velocity_seed", False)),
        time_integrator=integrator,
        nonlinear_form=numerics.get("nonlinear_form", "rotational"),
        coefficient_path=numerics.get("coefficient_path", "optimized"),
        precision=precision, device=device)


def mhd_checkpoint_payload(state) -> dict:
    """``_checkpoint_payload`` (``production/oracles.py:3872-3900``), MHDState
    branch: the flow leaves under ``flow_*`` (``have_old`` float32), A and the
    A history levels present."""
    f = state.flow
    payload: dict[str, Any] = {"flow_u": tuple(f.u), "flow_g": f.g, "A": tuple(state.A)}
    if f.nonlinear_old is not None:
        payload.update({"flow_nonlinear_old": tuple(f.nonlinear_old),
                        "flow_have_old": np.asarray(f.have_old, dtype=np.float32),
                        "flow_previous_dt": np.asarray(f.previous_dt)})
    if f.nonlinear_older is not None:
        payload["flow_nonlinear_older"] = tuple(f.nonlinear_older)
    if f.solution_old is not None:
        payload["flow_solution_old"] = tuple(f.solution_old)
    if f.solution_older is not None:
        payload["flow_solution_older"] = tuple(f.solution_older)
    payload["flow_history_steps"] = np.asarray(f.history_steps)
    for name in ("nonlinear_A_old", "nonlinear_A_older", "A_old", "A_older"):
        if getattr(state, name) is not None:
            payload[name] = tuple(getattr(state, name))
    return payload


def mhd_state_from_checkpoint_payload(payload: dict):
    """``_state_from_checkpoint_payload`` (``production/oracles.py:417-460``),
    vector-potential branch."""
    from .mhd import MHDState

    def group(name):
        return None if name not in payload else tuple(np.asarray(v) for v in payload[name])

    flow = KMMState(u=tuple(np.asarray(v) for v in payload["flow_u"]), g=np.asarray(payload["flow_g"]),
                    nonlinear_old=group("flow_nonlinear_old"), nonlinear_older=group("flow_nonlinear_older"),
                    solution_old=group("flow_solution_old"), solution_older=group("flow_solution_older"),
                    history_steps=float(np.asarray(payload.get("flow_history_steps", 0.0))),
                    have_old=float(np.asarray(payload.get("flow_have_old", 0.0), dtype=np.float32)),
                    previous_dt=float(np.asarray(payload.get("flow_previous_dt", 0.0))))
    return MHDState(flow=flow, A=group("A"), nonlinear_A_old=group("nonlinear_A_old"),
                    nonlinear_A_older=group("nonlinear_A_older"), A_old=group("A_old"), A_older=group("A_older"))


def pcf_curl_scalars(solver, state) -> dict:
    """``_pcf_curl_scalars`` (``production/oracles.py:1921-1952``)."""
    d = solver.diagnostics(state)
    out = {"box_volume": 2.0 * solver.Ly * solver.Lz if solver.domain[0] == (-1.0, 1.0)
           else float((solver.domain[0][1] - solver.domain[0][0]) * solver.Ly * solver.Lz),
           "kinetic_energy": float(d["Epert"]), "magnetic_energy": float(d["Emag"]),
           "magnetic_energy_total": float(d["Emag_total"]), "total_energy": float(d["Epert"]) + float(d["Emag"]),
           "divergence_u_l2": float(d["divL2"]), "divergence_b_l2": float(d["divB_L2"]),
           "maxwell_stress_xy": float(d["maxwell_stress"]), "reynolds_stress": float(d["reynolds_stress"]),
           "total_stress": float(d["total_stress"]), "alpha_Sh": float(d["alpha_Sh"]),
           "channel_kz1_velocity_rms": float(d["channel_kz1_velocity_rms"]),
           "channel_kz1_magnetic_rms": float(d["channel_kz1_magnetic_rms"]),
           "channel_kz1_total_rms": float(d["channel_kz1_total_rms"]),
           "mean_bx": float(d["mean_bx"]), "mean_by": float(d["mean_by"]), "mean_bz": float(d["mean_bz"]),
           "mag_energy_mean": float(d["mag_energy_mean"]), "mag_energy_fluct": float(d["mag_energy_fluct"])}
    out.update({key: float(d[key]) for key in PCF_ELECTROMAGNETIC_KEYS})
    if "alpha" in d:
        out["transport_alpha"] = float(d["alpha"])
        out["alpha_B0"] = float(d["alpha_B0"])
    if "insulating_bc_residual" in d:
        out["insulating_bc_residual"] = float(d["insulating_bc_residual"])
    return out


_DIVERGENCE_KEY_RE = None


def _is_divergence_key(key: str) -> bool:
    """``_is_divergence_key`` (``production/oracles.py:3600-3612``)."""
    import re

    global _DIVERGENCE_KEY_RE
    if _DIVERGENCE_KEY_RE is None:
        _DIVERGENCE_KEY_RE = re.compile(
            r"^(?:divergence(?:[_-](?:u|b))?(?:[_-]?(?:l2|linf|norm|rms|max))?"
            r"|div(?:u|b)(?:[_-]?(?:l2|linf|norm|rms|max))?"
            r"|div(?:[_-]?(?:l2|linf|norm|rms|max))"
            r"|continuity(?:[_-]?(?:residual|l2|linf|norm|rms|max))?)$")
    return _DIVERGENCE_KEY_RE.match(key.rsplit(".", 1)[-1].lower()) is not None


def raise_on_divergence_drift(diag: dict, *, t: float, tstep: int, magnetic_limit: float | None = None) -> None:
    """``_raise_on_divergence_drift`` (``production/oracles.py:3558-3598``): every
    divergence key of ``diag`` below its limit (the magnetic ones below
    ``magnetic_limit`` when given)."""
    offenders = []
    for key, value in diag.items():
        if not _is_divergence_key(str(key)):
            continue
        try:
            magnitude = abs(float(value))
        except (TypeError, ValueError):
            offenders.append(f"{key}=non-numeric")
            continue
        name = str(key).lower().replace("-", "_")
        is_magnetic = name.startswith("divb") or name.startswith("divergence_b")
        limit = float(magnetic_limit) if is_magnetic and magnetic_limit is not None else DIVERGENCE_GUARD_LIMIT
        if not math.isfinite(magnitude) or magnitude > limit:
            offenders.append(f"{key}={magnitude:g} > {limit:g}")
    if offenders:
        raise FloatingPointError(f"divergence guard failed at tstep={int(tstep)} t={float(t):g}: {', '.join(offenders)}")


def correlation_time(rows: list, *, key: str = "total_stress") -> float | None:
    """``correlation_time`` (``production/health.py:304-338``): integrated
    autocorrelation time by the initial-positive-sequence estimator."""
    samples = [(float(r["t"]), float(r[key])) for r in rows if key in r and "t" in r and math.isfinite(float(r[key]))]
    if len(samples) < MIN_CORRELATION_SAMPLES:
        return None
    times = np.asarray([s[0] for s in samples])
    values = np.asarray([s[1] for s in samples])
    dt_row = float(np.median(np.diff(times)))
    if dt_row <= 0.0:
        return None
    centered = values - values.mean()
    variance = float(np.dot(centered, centered))
    if variance <= 0.0:
        return None
    acf = 0.0
    for lag in range(1, centered.size - 1):
        rho = float(np.dot(centered[:-lag], centered[lag:])) / variance
        if rho <= 0.0:
            break
        acf += rho
    return dt_row * (1.0 + 2.0 * acf)


def effective_independent_samples(rows: list, *, correlation: float | None, key: str = "total_stress"):
    """``effective_independent_samples`` (``production/health.py:341-365``)."""
    if correlation is None or not math.isfinite(float(correlation)) or correlation <= 0.0:
        return None
    times = [float(r["t"]) for r in rows if key in r and "t" in r and math.isfinite(float(r[key]))
             and math.isfinite(float(r["t"]))]
    if len(times) < 2:
        return None
    duration = max(times) - min(times)
    return float(duration / float(correlation)) if duration > 0.0 else None


def energy_budget_residual(rows: list, *, shear_rate: float, volume: float) -> float | None:
    """``energy_budget_residual`` (``production/health.py:463-510``): median
    relative closure of d(E/2)/dt = S V <total_stress> - dissipation between
    consecutive rows."""
    required = ("t", "total_energy", "total_stress", "dissipation_kinetic", "dissipation_magnetic")
    usable = [r for r in rows if all(k in r for k in required)]
    if len(usable) < 2:
        return None
    residuals = []
    for a, b in zip(usable[:-1], usable[1:]):
        dtp = float(b["t"]) - float(a["t"])
        if dtp <= 0.0:
            continue
        lhs = 0.5 * (float(b["total_energy"]) - float(a["total_energy"])) / dtp
        stress = 0.5 * (float(a["total_stress"]) + float(b["total_stress"]))
        diss = 0.5 * (float(a["dissipation_kinetic"]) + float(a["dissipation_magnetic"])
                      + float(b["dissipation_kinetic"]) + float(b["dissipation_magnetic"]))
        rhs = float(shear_rate) * float(volume) * stress - diss
        residuals.append(abs(lhs - rhs) / max(abs(lhs), abs(rhs), diss, 1.0e-300))
    return float(np.median(residuals)) if residuals else None


def _vp_row(t: float, d: dict, dt: float) -> dict:
    """The diagnostics row of the vector-potential runner's ``collect``
    (``production/oracles.py:2022-2058``)."""
    row = {"t": float(t), "kinetic_energy": float(d["Epert"]), "magnetic_energy": float(d["Emag"]),
           "total_energy": float(d["Epert"]) + float(d["Emag"]), "divergence_u_l2": float(d["divL2"]),
           "divergence_b_l2": float(d["divB_L2"]), "reynolds_stress": float(d["reynolds_stress"]),
           "maxwell_stress_xy": float(d["maxwell_stress"]), "total_stress": float(d["total_stress"]),
           "alpha_Sh": float(d["alpha_Sh"]), "channel_kz1_velocity_rms": float(d["channel_kz1_velocity_rms"]),
           "channel_kz1_magnetic_rms": float(d["channel_kz1_magnetic_rms"]),
           "channel_kz1_total_rms": float(d["channel_kz1_total_rms"]), "mean_bx": float(d["mean_bx"]),
           "mean_by": float(d["mean_by"]), "mean_bz": float(d["mean_bz"]),
           "mag_energy_mean": float(d["mag_energy_mean"]), "mag_energy_fluct": float(d["mag_energy_fluct"]),
           "dissipation_kinetic": float(d["dissipation_kinetic"]),
           "dissipation_magnetic": float(d["dissipation_magnetic"]), "energy_convention": "integral_abs2",
           "dt": float(dt)}
    if "alpha" in d:
        row["transport_alpha"] = float(d["alpha"])
    if "insulating_bc_residual" in d:
        row["insulating_bc_residual"] = float(d["insulating_bc_residual"])
    row.update({key: float(d[key]) for key in PCF_ELECTROMAGNETIC_KEYS})
    return row


def run_pcf_vector_potential_mhd_saturation(spec: dict, *, steps: int | None = None, out_dir=None,
                                            checkpoint_every: int | None = None,
                                            diagnostics_every: int | None = None,
                                            resume_checkpoint: CheckpointRecord | None = None, on_row=None,
                                            precision: str = "fp64", solver=None) -> dict:
    """``_run_pcf_vector_potential_mhd_saturation`` (``production/oracles.py:1956-2250``),
    fixed-dt path, on the HIP rotating-shear solvers: the conducting or
    insulating solver from the spec, the div-B guard at the start, after every
    block and at the end, cadence rows with the stresses, mean-field split,
    dissipation and electromagnetic keys (device diagnostics), the MHDState
    checkpoint payload with exact resume, and the final scalars (growth rate
    of the total energy, magnetic energy growth factor and saturation check,
    flux drift, stationarity of the total energy, correlation time of the
    total stress and the energy-budget closure).  The state stays in device
    memory between blocks.  Not built: the adaptive-CFL path, quench
    continuation and burn-in windows, snapshots and profiles."""
    from .pcf import to_host

    if spec.get("time", {}).get("adaptive_cfl") is not None:
        raise ProductionOracleNotImplementedError("the adaptive-CFL vector-potential runner is not built")
    bc = magnetic_bc(spec)
    if bc not in {"conducting", "insulating"}:
        raise ProductionOracleNotImplementedError(
            "the vector-potential (curl) PCF-MRI family is wired for conducting or insulating walls, "
            f"got {bc!r}; pseudo-vacuum runs use the primitive-b family")
    x0, x1 = (float(v) for v in spec["domain"]["x"])
    if not (math.isclose(x0, -1.0, rel_tol=0.0, abs_tol=1e-12) and math.isclose(x1, 1.0, rel_tol=0.0, abs_tol=1e-12)):
        raise ProductionOracleNotImplementedError(
            f"PCF vector-potential DNS is wired for the half-gap domain [-1, 1], got [{x0:g}, {x1:g}]")
    guard = vector_potential_divergence_guard(spec)
    for name, value in (("checkpoint_every", checkpoint_every), ("diagnostics_every", diagnostics_every)):
        if value is not None and value <= 0:
            raise ValueError(f"{name} must be positive")
    if checkpoint_every is not None and out_dir is None:
        raise ValueError("out_dir is required when checkpoint, snapshot, or profile output is set")
    solver = solver if solver is not None else curl_solver_from_spec(spec, precision=precision)
    state = solver.initial_state()
    tstep0, t0 = 0, 0.0
    if resume_checkpoint is not None:
        kind = str(resume_checkpoint.attrs.get("state_kind"))
        if kind != VP_STATE_KIND:
            raise ValueError(f"resume checkpoint state_kind {kind!r} does not match expected {VP_STATE_KIND!r}")
        payload = resume_checkpoint.fields.get("state")
        if not isinstance(payload, dict):
            raise ValueError("resume checkpoint is missing state payload")
        state = mhd_state_from_checkpoint_payload(payload)
        tstep0, t0 = int(resume_checkpoint.tstep), float(resume_checkpoint.t)
    first = pcf_curl_scalars(solver, state)
    raise_on_divergence_drift({"divergence_b_l2": first["divergence_b_l2"]}, t=t0, tstep=tstep0,
                              magnetic_limit=guard)
    target = steps_from_spec(spec, steps)
    if tstep0 > target:
        raise ValueError(f"resume checkpoint step {tstep0} is beyond target step {target}")
    n_steps = target - tstep0
    dt = solver.dt
    cadence = [int(c) for c in (checkpoint_every, diagnostics_every) if c]
    rows: list = []
    ckpt = Path(out_dir) / "checkpoints" / "checkpoints.npz" if checkpoint_every is not None else None
    if ckpt is not None:
        ckpt.parent.mkdir(parents=True, exist_ok=True)

    def save(t, tstep, st):
        write_checkpoint(ckpt, {"state": mhd_checkpoint_payload(to_host(st))}, t=t, tstep=tstep,
                         attrs={"state_kind": VP_STATE_KIND})

    done = 0
    while done < n_steps:
        block = n_steps - done
        for c in cadence:
            block = min(block, c - ((tstep0 + done) % c))
        state = solver.solve_device(state, block)
        done += block
        tstep, t = tstep0 + done, t0 + done * dt
        d = solver.diagnostics(state)
        if diagnostics_every and tstep % diagnostics_every == 0:
            row = _vp_row(t, d, dt)
            rows.append(row)
            if on_row is not None:
                on_row(row)
        raise_on_unhealthy(d, t=t, tstep=tstep)
        raise_on_divergence_drift({"divL2": d["divL2"], "divB_L2": d["divB_L2"]}, t=t, tstep=tstep,
                                  magnetic_limit=guard)
        if ckpt is not None and tstep % checkpoint_every == 0:
            save(t, tstep, state)
    if ckpt is not None and (n_steps == 0 or target % checkpoint_every != 0):
        save(t0 + n_steps * dt, target, state)
    final = pcf_curl_scalars(solver, state)
    elapsed = n_steps * float(spec["time"]["dt"])
    growth = growth_rate_from_energy(first["total_energy"], final["total_energy"], n_steps, dt) if n_steps else math.nan
    raise_on_divergence_drift({"divergence_b_l2": final["divergence_b_l2"]}, t=t0 + elapsed, tstep=target,
                              magnetic_limit=guard)
    mgrowth = final["magnetic_energy"] / first["magnetic_energy"] if first["magnetic_energy"] > 0.0 else 0.0
    passed = saturation_passed(mgrowth, threshold=2.0,
                               final_energies=(final["kinetic_energy"], final["magnetic_energy"],
                                               final["total_energy"], final["divergence_u_l2"],
                                               final["divergence_b_l2"]))
    scalars = {**final, **{f"flux_drift_b{a}": final[f"mean_b{a}"] - first[f"mean_b{a}"] for a in "xyz"},
               "growth_rate": float(growth), "magnetic_energy_growth_factor": float(mgrowth),
               "saturation_check_passed": bool(passed), "magnetic_bc": bc, "representation": "vector_potential",
               "divergence_b_guard_l2": float(guard), "energy_convention": "integral_abs2"}
    numeric = lambda v: isinstance(v, (int, float)) and not isinstance(v, bool)  # noqa: E731
    first_row = {"t": float(t0), **{k: v for k, v in first.items() if numeric(v)}}
    last_row = {"t": float(t0) + elapsed, "dt": float(dt), **{k: v for k, v in final.items() if numeric(v)},
                "growth_rate": float(growth), "magnetic_energy_growth_factor": float(mgrowth),
                "saturation_check_passed": bool(passed)}
    series = dedupe_time_rows([first_row, *rows, last_row])
    scalars.update(stationarity_scalars(series, key="total_energy"))
    tau = correlation_time(rows)
    if tau is not None:
        scalars["correlation_time_total_stress"] = float(tau)
        n_ind = effective_independent_samples(rows, correlation=tau)
        if n_ind is not None:
            scalars["effective_independent_samples_total_stress"] = float(n_ind)
    budget = energy_budget_residual(rows, shear_rate=float(solver.shear_rate), volume=scalars["box_volume"])
    if budget is not None:
        scalars["energy_budget_residual"] = float(budget)
    return {"scalars": scalars, "time_series": series, "final_state": to_host(state)}


# ---------------------------------------------------------------------------
# Taylor-Couette DNS problems (production/oracles.py:1237-1330, 2942-3039)
# ---------------------------------------------------------------------------


def tc_kz_mode(spec: dict, Lz: float) -> int:
    """``_kz_mode_from_spec`` (``production/oracles.py:4656-4675``)."""
    mode = spec.get("mode", {})
    m = int(mode.get("azimuthal_wavenumber", 0))
    if m != 0 and selected_resolution(spec).get("Ntheta") is None:
        raise NotImplementedError("non-axisymmetric Taylor-Couette modes require resolution.Ntheta; "
                                  f"got azimuthal_wavenumber={m} for an axisymmetric solver")
    kz_mode = int(round(float(mode["axial_wavenumber"]) * float(Lz) / (2.0 * math.pi)))
    if kz_mode < 1:
        raise ValueError("axial_wavenumber does not map to a positive Fourier mode")
    return kz_mode


def _tc_physics(spec: dict) -> dict:
    """nu, eta and B0 of a Taylor-Couette spec: the explicit groups, else from
    Re / Rm on the inner-cylinder velocity scale Omega1 R1 (R2 - R1)."""
    g = spec["nondimensional_groups"]
    scale = float(g["Omega1"]) * float(g["R1"]) * (float(g["R2"]) - float(g["R1"]))
    nu = float(g["nu"]) if "nu" in g else scale / float(g["Re"])
    eta = float(g["eta_mag"]) if "eta_mag" in g else (scale / float(g["Rm"]) if "Rm" in g else nu)
    b0 = float(g.get("B0", spec.get("forcing", {}).get("B0", 0.0)))
    return {"nu": nu, "eta": eta, "B0": b0}


def run_taylor_couette_dns(spec: dict, *, steps: int | None = None, device: int = 0) -> dict:
    """The Taylor-Couette DNS golden problems on the device: the hydrodynamic
    ``_run_taylor_couette_hydro_dns`` (``production/oracles.py:1237-1330``) and,
    for ``physics == "mhd"`` with conducting walls, ``_run_taylor_couette_mhd_dns``
    (``:2942-3039``): the solver class from the resolution (``Ntheta`` selects
    the 3D class), the linear-eigenmode seed, the run and the scalars and time
    series the reference reports.  Checkpoints and snapshots of the reference's
    runner are not written here."""
    from .tc import (AxisymmetricMRIDNSHIP, AxisymmetricTCDNSHIP, CircularCouette, TaylorCouetteDNSHIP,
                     TaylorCouetteMRIDNSHIP)

    resolution = selected_resolution(spec)
    g = spec["nondimensional_groups"]
    phys = _tc_physics(spec)
    base = CircularCouette(float(g["R1"]), float(g["R2"]), float(g["Omega1"]), float(g["Omega2"]))
    mhd = spec.get("physics") == "mhd"
    if mhd and magnetic_bc(spec) != "conducting":
        raise NotImplementedError("TC MHD DNS is built for conducting walls only, "
                                  f"got {magnetic_bc(spec)!r}")
    ntheta = resolution.get("Ntheta")
    kw = dict(nu=phys["nu"], Nr=int(resolution.get("Nr", resolution.get("N", 40))), Nz=int(resolution.get("Nz", 8)),
              Lz=float(spec["domain"]["z_period"]), dt=float(spec["time"]["dt"]),
              family=resolution.get("family", "C"), dealias=float(resolution.get("dealias", 1.0)),
              time_integrator=str(spec["time"].get("integrator", "CNAB2")), device=device)
    if mhd:
        kw.update(B0=phys["B0"], eta_mag=phys["eta"])
    if ntheta is not None:
        kw["Ntheta"] = int(ntheta)
        cls = TaylorCouetteMRIDNSHIP if mhd else TaylorCouetteDNSHIP
    else:
        cls = AxisymmetricMRIDNSHIP if mhd else AxisymmetricTCDNSHIP
    solver = cls(base, **kw)
    try:
        seed = {"kz_mode": tc_kz_mode(spec, solver.Lz),
                "amp": float(spec["initial_condition"].get("amplitude", 1.0e-7 if mhd else 1.0e-6))}
        if ntheta is not None:
            seed["m"] = int(spec.get("mode", {}).get("azimuthal_wavenumber", 0))
        state, eig = solver.seed_linear_eigenmode(**seed)
        initial = solver.diagnostics(state)
        n = steps_from_spec(spec, steps)
        out = solver.solve(state, n)
        final = solver.diagnostics(out)
    finally:
        solver.close()
    growth = growth_rate_from_energy(initial["E"], final["E"], n, solver.dt)
    elapsed = n * float(spec["time"]["dt"])
    if mhd:
        scalars = {"kinetic_energy": final["Ekin"], "magnetic_energy": final["Emag"], "growth_rate": growth,
                   "growth_rate_linear": eig.real, "divergence_u": final["divu"], "divergence_b": final["divb"],
                   "magnetic_bc": "conducting"}
        series = [{"t": 0.0, "kinetic_energy": initial["Ekin"], "magnetic_energy": initial["Emag"],
                   "growth_rate_linear": eig.real},
                  {"t": elapsed, "kinetic_energy": final["Ekin"], "magnetic_energy": final["Emag"],
                   "growth_rate": growth}]
    else:
        scalars = {"kinetic_energy": final["E"], "growth_rate": growth, "growth_rate_linear": eig.real,
                   "divergence_linf": final["div_linf"], "rayleigh_stable": bool(base.rayleigh_stable())}
        series = [{"t": 0.0, "kinetic_energy": initial["E"], "growth_rate_linear": eig.real},
                  {"t": elapsed, "kinetic_energy": final["E"], "growth_rate": growth}]
    return {"scalars": scalars, "time_series": series}
