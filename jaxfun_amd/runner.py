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
