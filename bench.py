#!/usr/bin/env python3
"""Throughput of the plane-Couette KMM SBDF3 step on one MI355X.

Workload (one "step" = one steady SBDF3 rotational step of the whole field):
65x64x64 (65 Chebyshev points, 64x64 Fourier modes, 3/2 dealiasing), Re=1000,
dt=0.005, seeded divergence-free perturbation with energy 1e-3 in the lowest
12 modes, FP64 throughout.  The two IMEXRK3 startup steps, operator setup and
the warm-up are excluded; the timed steps replay 12-step HIP graphs with the
state resident in HBM.

Also reported on the same line:
  * the mixed-precision step (FP32 transforms and nonlinear term, FP64 solves)
    and its throughput ratio to FP64;
  * per-kernel mean durations from HIP events on the library's stream, and the
    roofline of the dominant kernel;
  * the CPU FP64 oracle timed on this host's cores on a bounded sample;
  * the adjoint/forward cost ratio with its finite-difference check;
  * the insulating-wall MHD step (FP64 and mixed) at the same resolution;
  * FP64 steps/s on the reference's other published configurations (hydro
    97^3 and 128^3, conducting-wall MHD 65^3 and 97^3, the IMEXRK3 integrator
    and the conducting-wall MHD CNAB2 integrator at 65^3) beside its figures.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "DNS steps/s at 65×64×64: mixed vs FP64 on GPU; adjoint/forward cost ratio"
WORKLOAD = dict(N=(65, 64, 64), Re=1000.0, dt=0.005, energy=1e-3, K=12, seed=0)
PUBLISHED_STEPS_PER_S = 1000.0 / 27.552  # reference 65x64x64 SBDF3 FP64, RTX 5090 Laptop (benchmark-results/FINAL_LOCAL_CHECKPOINT.md:58)
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
GRAPH_BLOCK = 12
# the path every published SBDF3 number of the reference uses
SBDF3_ROT = dict(time_integrator="SBDF3", nonlinear_form="rotational", coefficient_path="optimized")


def stage_model(solver, precision: str) -> dict:
    """Algorithmic HBM bytes and flops per launch of each stage kernel.

    Each stage reads its inputs once and writes its outputs once; spectral
    state and history are complex FP64 (16 B); nodal/physical intermediates
    are complex in the compute precision (16 B FP64, 8 B FP32).
    """
    n0, ny, nz = solver.N
    nb, nd = n0 - 4, n0 - 2
    kzs = nz // 2
    m = ny * kzs
    my = int(1.5 * ny)
    V = 16 if precision == "fp64" else 8
    c16 = 16
    # (u0, g, v00, w00) in (the reconstruction of u1, u2 is folded into the
    # contraction: five products, three over u0 and two over g), 6 nodal fields out
    nodal_b = (nb + nd) * m * c16 + 6 * n0 * m * V
    nodal_f = m * n0 * (12 * nb + 8 * nd)
    ffti_b = 6 * n0 * m * V + 6 * n0 * my * kzs * V
    mz = int(1.5 * nz)
    fft_flops = lambda n, count: 5.0 * n * __import__("math").log2(n) * count  # noqa: E731
    ffti_f = fft_flops(my, 6 * n0 * kzs)
    rot_b = 6 * n0 * my * kzs * V + 3 * n0 * my * kzs * V
    rot_f = fft_flops(mz, 3 * n0 * my // 2) * 2 + fft_flops(mz, 3 * n0 * my // 2) + 30 * n0 * my * mz
    fftf_b = 3 * n0 * my * kzs * V + 3 * n0 * m * V
    fftf_f = fft_flops(my, 3 * n0 * kzs)
    rhs_b = 3 * n0 * m * V + (nb + nd) * m * c16
    rhs_f = m * n0 * 16 * nd
    # three solution + three nonlinear levels in, new (u0, g) out, and the
    # per-mode factors (7 values per row of the clamped blocks, 4 of the
    # Dirichlet ones, FP64)
    solve_b = 7 * (nb + nd) * m * c16 + (7 * nb + 4 * nd) * m * 8
    solve_f = m * 2 * (nb + nd) * 40  # O(n) structured sweeps, ~40 flops per row
    return {
        "nodal_backward": (nodal_b, nodal_f),
        "fft_y_inverse": (ffti_b, ffti_f),
        "rotational": (rot_b, rot_f),
        "fft_y_forward": (fftf_b, fftf_f),
        "forward_rhs": (rhs_b, rhs_f),
        "solve": (solve_b, solve_f),
    }


def profile_stages(solver, nsteps: int) -> dict:
    lib = solver._lib
    n = lib.pcf_stage_count()
    out = (C.c_double * n)()
    from jaxfun_amd import _native as nat

    nat.check(lib.pcf_profile_steady(solver._h, int(nsteps), out))
    return {lib.pcf_stage_name(i).decode(): float(out[i]) for i in range(n)}


def run_gpu(precision: str, steps: int, warmup: int, eager: bool = False):
    """Exactly ``warmup`` untimed and ``steps`` timed steady steps after the two
    IMEXRK3 start-up steps.  The timed steps replay 12-step HIP graphs and one
    graph of the remainder, all captured (not run) before the timed region."""
    from jaxfun_amd import PlaneCouetteFluctuationHIP
    from jaxfun_amd.ic import seeded_state

    w = WORKLOAD
    solver = PlaneCouetteFluctuationHIP(N=w["N"], Re=w["Re"], dt=w["dt"], precision=precision, **SBDF3_ROT)
    state = seeded_state(solver, seed=w["seed"], energy=w["energy"], K=w["K"])
    solver.load(state)
    solver.advance(2)  # IMEXRK3 startup, excluded
    block = 0 if eager else GRAPH_BLOCK
    solver.advance_steady(warmup, graph_block=block)
    if block:
        solver.prepare_steady(steps, block)
    solver.synchronize()
    t0 = time.perf_counter()
    solver.advance_steady(steps, graph_block=block)
    solver.synchronize()
    t1 = time.perf_counter()
    stages = profile_stages(solver, 48)
    diag = solver.diagnostics()
    return solver, state, (t1 - t0), stages, diag


def block_protocol(solver, block_steps: int = 25, timed_blocks: int = 10) -> dict:
    """The reference's timing protocol (production/benchmark.py:282-354): warm
    ms per step as the median (and p90) over ``timed_blocks`` blocks of
    ``block_steps`` steps, each block ended by a device synchronisation.  Each
    block replays one cached graph of ``block_steps`` steps; a 25-step block
    moves the rings one phase on (25 = 2 * 12 + 1), so the warm-up runs one
    block per ring phase (12, where the reference runs 2) and every timed block
    hits the cache."""
    import numpy as np

    for _ in range(12):
        solver.advance_steady(block_steps, graph_block=block_steps)
    solver.synchronize()
    per = []
    for _ in range(timed_blocks):
        t0 = time.perf_counter()
        solver.advance_steady(block_steps, graph_block=block_steps)
        solver.synchronize()
        per.append((time.perf_counter() - t0) / block_steps)
    arr = np.asarray(per)
    return {"block_steps": block_steps, "warmup_blocks": 12, "timed_blocks": timed_blocks,
            "ms_per_step_p50": 1e3 * float(np.percentile(arr, 50)),
            "ms_per_step_p90": 1e3 * float(np.percentile(arr, 90)),
            "steps_per_s_p50": 1.0 / float(np.percentile(arr, 50))}


# stage name -> kernel-name fragment (FP64 instance) in the rocprofv3 traffic summary
_KERNEL_OF = {"nodal_backward": "k_nodal_prim_mfma<double>", "fft_y_inverse": "k_fft_y<double, true",
              "rotational": "k_rot_wave<double", "fft_y_forward": "k_fft_y<double, false",
              "forward_rhs": "k_rhs_mfma<double>", "solve": "k_solve_factored"}
TRAFFIC_FILE = ROOT / "profiles" / "traffic_65x64x64.json"


def measured_traffic(stage: str, precision: str):
    """HBM bytes per launch from the committed counter summary (FETCH_SIZE and
    WRITE_SIZE runs, copy-calibrated), or None when no summary covers it."""
    if not TRAFFIC_FILE.exists():
        return None
    data = json.loads(TRAFFIC_FILE.read_text())
    frag = _KERNEL_OF.get(stage, stage)
    if precision != "fp64":
        frag = frag.replace("double", "float")
    for name, rec in data["kernels"].items():
        if frag in name:
            return rec["traffic_bytes"]
    return None


def launch_ms(solver, stage: str, reps: int = 200) -> float | None:
    """Mean duration of one launch of a nonlinear stage: `reps` back-to-back
    launches between two HIP events on the library's stream (pcf_profile_stage);
    None for the solve, whose repeats are not idempotent."""
    lib = solver._lib
    names = [lib.pcf_stage_name(i).decode() for i in range(lib.pcf_stage_count())]
    idx = names.index(stage)
    if idx > 4:
        return None
    from jaxfun_amd import _native as nat

    out = C.c_double()
    nat.check(lib.pcf_profile_stage(solver._h, idx, int(reps), C.byref(out)))
    return out.value


def roofline(solver, precision: str, stages: dict) -> dict:
    """Roofline of the dominant kernel: algorithmic bytes (or flops) per launch
    over its mean launch duration from back-to-back launches (the per-stage
    event time inside a step, which also holds a launch gap, is kept as
    `avg_ms_in_step`)."""
    model = stage_model(solver, precision)
    name = max(stages, key=lambda k: stages[k])
    nbytes, flops = model[name]
    per_launch = launch_ms(solver, name)
    avg_ms = per_launch if per_launch is not None else stages[name]
    sec = avg_ms * 1e-3
    traffic = measured_traffic(name, precision)
    peak_f = (FP64_PEAK_TFLOPS if precision == "fp64" else FP32_PEAK_TFLOPS) * 1e12
    hbm_t = nbytes / (HBM_PEAK_GBS * 1e9)
    cmp_t = flops / peak_f
    if cmp_t > hbm_t:
        ach = flops / sec / 1e12
        return {"kernel": name, "bound": "mfma", "achieved": ach, "peak": peak_f / 1e12,
                "unit": "TFLOP/s", "frac": ach / (peak_f / 1e12), "traffic": traffic,
                "algorithmic_bytes": nbytes, "flops": flops, "avg_ms": avg_ms,
                "avg_ms_in_step": stages[name]}
    ach = nbytes / sec / 1e9
    return {"kernel": name, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS, "traffic": traffic, "algorithmic_bytes": nbytes,
            "flops": flops, "avg_ms": avg_ms, "avg_ms_in_step": stages[name]}


def copy_calibration(mb: int) -> dict:
    """Kernels of known HBM traffic for the FETCH_SIZE / WRITE_SIZE scale: a
    device-to-device copy of ``mb`` MiB through the HIP runtime (its blit copy
    kernel), then, where torch is importable, a fill of mb / 2 MiB (write only,
    16-byte stores) and elementwise products over mb / 2 MiB in float32 and in
    float16: the scale is checked across four kernels."""
    hip = C.CDLL("libamdhip64.so")
    n = mb * (1 << 20)
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), C.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, C.c_size_t(n))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(dst, src, C.c_size_t(n), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    out = {"copy": {"bytes_read": n, "bytes_written": n}}
    try:
        import torch
    except ImportError:
        return out
    half = n // 2
    x = torch.full((half // 4,), 1.5, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.mul(x, 2.0, out=y)
    torch.cuda.synchronize()
    out["fill"] = {"bytes_read": 0, "bytes_written": half}
    out["elementwise"] = {"bytes_read": half, "bytes_written": half}
    del x, y
    # the same product in float16 (a third kernel; torch vectorises it 8 wide, so
    # also 16 bytes per lane)
    x = torch.full((half // 2,), 1.5, dtype=torch.float16, device="cuda")
    y = torch.empty_like(x)
    torch.mul(x, 2.0, out=y)
    torch.cuda.synchronize()
    out["elementwise_8b"] = {"bytes_read": half, "bytes_written": half}
    del x, y
    return out


def _cpu_blocks(state, threads: int, budget_s: float) -> dict:
    """The FP64 oracle on ``threads`` host threads (scipy.fft workers and the
    BLAS pool): 2 start-up + 2 warm-up steps, then the median of 3 blocks of
    steady steps, 10 per block unless the warm-up says that 3 blocks would
    exceed ``budget_s`` (then fewer, at least 2)."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    from oracle.kmm import OracleState, PCFOracle, set_workers

    w = WORKLOAD
    with threadpool_limits(limits=threads):
        set_workers(threads)
        o = PCFOracle(N=w["N"], Re=w["Re"], dt=w["dt"])
        s = o.reset_history(OracleState(u=tuple(np.array(a) for a in state.u), g=np.array(state.g)))
        s = o.solve(s, 3)  # two start-up steps and one warm-up step, excluded
        t0 = time.perf_counter()
        s = o.step(s)  # second warm-up step, timed to size the blocks
        t_step = time.perf_counter() - t0
        per_block = int(max(2, min(10, budget_s / (3 * t_step))))
        rates = []
        for _ in range(3):
            t0 = time.perf_counter()
            s = o.solve(s, per_block)
            rates.append(per_block / (time.perf_counter() - t0))
        set_workers(None)
    return {"value": float(np.median(rates)), "blocks": 3, "steps_per_block": per_block, "threads": threads}


def cpu_baseline(state, budget_s: float) -> dict:
    """FP64 oracle (numpy/scipy port of the reference algorithm) on the host
    cores this process may use, and on one thread."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        cores = os.cpu_count() or 1
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    many = _cpu_blocks(state, cores, budget_s)
    one = _cpu_blocks(state, 1, budget_s)
    return {"value": many["value"], "unit": "steps/s", "cores": cores, "kind": "port",
            "sample": f"median of 3 blocks of {many['steps_per_block']} steady SBDF3 steps of the 65x64x64 "
                      f"workload after 2 start-up + 2 warm-up steps (FP64 oracle, scipy.fft workers and BLAS "
                      f"threads = {cores})",
            "one_thread": {"value": one["value"], "unit": "steps/s", "cores": 1,
                           "sample": f"median of 3 blocks of {one['steps_per_block']} steady steps, 1 thread"}}


def mixed_error(state, steps: int) -> dict:
    """Per-field error of the mixed-precision step against the FP64 step after
    ``steps`` steps (startup included) from the same seeded state:
    max |mixed - fp64| / max |fp64| over each field's coefficients."""
    import numpy as np

    from jaxfun_amd import PlaneCouetteFluctuationHIP

    w = WORKLOAD
    out = {}
    for prec in ("fp64", "mixed"):
        s = PlaneCouetteFluctuationHIP(N=w["N"], Re=w["Re"], dt=w["dt"], precision=prec, **SBDF3_ROT)
        out[prec] = s.solve(state, steps)
        s.close()
    ref, mix = out["fp64"], out["mixed"]
    err = {}
    for name, a, b in (("u0", mix.u[0], ref.u[0]), ("u1", mix.u[1], ref.u[1]),
                       ("u2", mix.u[2], ref.u[2]), ("g", mix.g, ref.g)):
        err[name] = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    err["velocity"] = max(err["u0"], err["u1"], err["u2"])
    err["steps"] = steps
    return err


def pcie_inclusive(state, steps: int) -> dict:
    """``solve(state, steps)`` through the host boundary: state copied in from
    host memory, stepped, and copied back (full-complex reference layout)."""
    from jaxfun_amd import PlaneCouetteFluctuationHIP

    w = WORKLOAD
    s = PlaneCouetteFluctuationHIP(N=w["N"], Re=w["Re"], dt=w["dt"], precision="fp64", **SBDF3_ROT)
    s.solve(state, 2)  # warm the handle (first-touch allocations)
    t0 = time.perf_counter()
    s.solve(state, steps)
    el = time.perf_counter() - t0
    s.close()
    return {"value": steps / el, "unit": "steps/s", "steps": steps,
            "note": "host state in, host state out, startup steps included"}


def state_exchange(state, reps: int = 5) -> dict:
    """One full state exchange (load then fetch of every leaf, history
    included) through host memory (pcf_set_field / pcf_get_field, PCIe) and
    through device memory (pcf_set_field_device / pcf_get_field_device, torch
    tensor leaves), mean over ``reps``, and the state's size."""
    from jaxfun_amd import PlaneCouetteFluctuationHIP

    w = WORKLOAD
    s = PlaneCouetteFluctuationHIP(N=w["N"], Re=w["Re"], dt=w["dt"], precision="fp64", **SBDF3_ROT)
    s.load(state)
    host = s.fetch()
    nbytes = sum(int(a.nbytes) for a in list(host.u) + [host.g] + [x for grp in (
        host.nonlinear_old, host.nonlinear_older, host.solution_old, host.solution_older) for x in grp])
    t0 = time.perf_counter()
    for _ in range(reps):
        s.load(host)
        host = s.fetch()
    el_host = (time.perf_counter() - t0) / reps
    dev = s.fetch_device()
    t0 = time.perf_counter()
    for _ in range(reps):
        s.load(dev)
        dev = s.fetch_device()
    el_dev = (time.perf_counter() - t0) / reps
    s.close()
    return {"state_bytes": nbytes, "host_ms": 1e3 * el_host, "device_ms": 1e3 * el_dev,
            "note": "load + fetch of the whole state (20 leaves); the device path keeps torch CUDA tensors"}


def stage_rooflines(solver, precision: str, stages: dict) -> dict:
    """Every stage kernel against its bound: algorithmic bytes over the
    stage's event time inside the step, the measured HBM traffic per launch
    from the committed counter summary, and for the two matrix-core
    contractions the model flops over the same time against the matrix peak."""
    model = stage_model(solver, precision)
    peak_f = FP64_PEAK_TFLOPS if precision == "fp64" else FP32_PEAK_TFLOPS
    out = {}
    for name, ms in stages.items():
        nbytes, flops = model[name]
        sec = ms * 1e-3
        rec = {"algorithmic_bytes": nbytes, "ms_in_step": ms, "achieved_gbs": nbytes / sec / 1e9,
               "frac_hbm": nbytes / sec / 1e9 / HBM_PEAK_GBS, "traffic": measured_traffic(name, precision)}
        if name in ("nodal_backward", "forward_rhs"):
            rec["mfma_tflops"] = flops / sec / 1e12
            rec["frac_mfma"] = flops / sec / 1e12 / peak_f
        out[name] = rec
    return out


# Mixed precision at the large sizes: 97x96x96 and 129x128x128, Re = 2000,
# dt = 0.002, seeded perturbation energy 1e-2 over the lowest 12 modes, 2000 steps.
LARGE = dict(sizes=((97, 96, 96), (129, 128, 128)), Re=2000.0, dt=0.002, energy=1e-2, K=12, seed=0, steps=2000)


def large_mixed_leg(steps: int) -> list:
    """FP64 and mixed precision from the same seeded state for ``steps`` steps
    (2 IMEXRK3 start-up steps, then steady steps replayed from graphs): steady
    throughput of each, the final energy, wall shear and divergence, and the
    mixed run's max relative error per velocity field against the FP64 run."""
    import numpy as np

    from jaxfun_amd import PlaneCouetteFluctuationHIP
    from jaxfun_amd.ic import seeded_state

    L = LARGE
    out = []
    for N in L["sizes"]:
        rec = {"N": list(N), "Re": L["Re"], "dt": L["dt"], "energy": L["energy"], "K": L["K"], "steps": steps}
        finals = {}
        for prec in ("fp64", "mixed"):
            solver = PlaneCouetteFluctuationHIP(N=N, Re=L["Re"], dt=L["dt"], precision=prec, **SBDF3_ROT)
            s0 = seeded_state(solver, seed=L["seed"], energy=L["energy"], K=L["K"])
            solver.load(s0)
            solver.advance(2)
            solver.prepare_steady(steps - 2, GRAPH_BLOCK)
            solver.synchronize()
            t0 = time.perf_counter()
            solver.advance_steady(steps - 2, graph_block=GRAPH_BLOCK)
            solver.synchronize()
            el = time.perf_counter() - t0
            finals[prec] = solver.fetch()
            d = solver.diagnostics()
            rec[prec] = {"steady_steps_per_s": (steps - 2) / el, "Epert": d["Epert"], "Etot": d["Etot"],
                         "wall_shear_top": d["wall_shear_top"], "wall_shear_bot": d["wall_shear_bot"],
                         "divL2": d["divL2"]}
            solver.close()
        a, b = finals["fp64"], finals["mixed"]
        err = {name: float(np.max(np.abs(x - y)) / np.max(np.abs(y)))
               for name, x, y in zip(("u0", "u1", "u2", "g"), (*b.u, b.g), (*a.u, a.g))}
        err["velocity"] = max(err["u0"], err["u1"], err["u2"])
        err["Epert_rel"] = abs(rec["mixed"]["Epert"] - rec["fp64"]["Epert"]) / rec["fp64"]["Epert"]
        err["wall_shear_abs"] = max(abs(rec["mixed"][k] - rec["fp64"][k]) for k in ("wall_shear_top", "wall_shear_bot"))
        rec["mixed_error_vs_fp64"] = err
        rec["ratio_mixed_vs_fp64"] = rec["mixed"]["steady_steps_per_s"] / rec["fp64"]["steady_steps_per_s"]
        print(f"# large {N}: fp64 {rec['fp64']['steady_steps_per_s']:.0f}, mixed "
              f"{rec['mixed']['steady_steps_per_s']:.0f} steps/s; velocity error {err['velocity']:.2e}",
              file=sys.stderr, flush=True)
        out.append(rec)
    return out


# Adjoint workload: gradient of the final perturbation energy after 400 SBDF3
# steps (IMEXRK3 start-up included) with respect to the seeded initial field,
# checkpoints every 20 steps.
ADJOINT = dict(N=(65, 64, 64), Re=1000.0, dt=0.005, energy=1e-4, K=12, seed=0, steps=400, every=20)


def adjoint_leg(precision: str, reps: int = 3) -> dict:
    """Wall time of value + gradient over ADJOINT['steps'] steps (recompute
    included) against the same forward step sequence, state resident on the GPU."""
    from jaxfun_amd import PlaneCouetteFluctuationHIP
    from jaxfun_amd import _native as nat
    from jaxfun_amd.ic import seeded_state

    a = ADJOINT
    solver = PlaneCouetteFluctuationHIP(N=a["N"], Re=a["Re"], dt=a["dt"], precision=precision, **SBDF3_ROT)
    s0 = seeded_state(solver, seed=a["seed"], energy=a["energy"], K=a["K"])
    lib, h = solver._lib, solver._h

    def fwd():
        solver.load(s0)
        solver.synchronize()
        t0 = time.perf_counter()
        solver.advance(a["steps"])
        solver.synchronize()
        return time.perf_counter() - t0

    def adj():
        solver.load(s0)
        solver.synchronize()
        t0 = time.perf_counter()
        nat.check(lib.pcf_vjp(h, a["steps"], a["every"], nat.SEED_ENERGY))
        return time.perf_counter() - t0

    fwd()
    adj()  # first call allocates the tape and snapshots
    tf = sorted(fwd() for _ in range(reps))[reps // 2]
    ta = sorted(adj() for _ in range(reps))[reps // 2]
    solver.close()
    return {"precision": precision, "forward_ms": 1e3 * tf, "value_and_grad_ms": 1e3 * ta, "ratio": ta / tf}


def reverse_leg(precision: str, nsteps: int = 40, reps: int = 3) -> dict:
    """Time of one reverse steady step, measured by ``pcf_profile_reverse``
    (reverse kernels alone on one stream, no capture, no overlapped
    recompute), beside an eager forward steady step with its tape write, at
    the adjoint workload (state resident)."""
    from jaxfun_amd import PlaneCouetteFluctuationHIP
    from jaxfun_amd import _native as nat
    from jaxfun_amd.ic import seeded_state
    import numpy as np

    a = ADJOINT
    solver = PlaneCouetteFluctuationHIP(N=a["N"], Re=a["Re"], dt=a["dt"], precision=precision, **SBDF3_ROT)
    solver.load(seeded_state(solver, seed=a["seed"], energy=a["energy"], K=a["K"]))
    solver.advance(2)  # the IMEXRK3 start-up: the history is bootstrapped
    out = np.zeros(2)
    runs = []
    for _ in range(reps + 1):  # the first call allocates the tape
        nat.check(solver._lib.pcf_profile_reverse(solver._h, nsteps, out.ctypes.data_as(C.POINTER(C.c_double))))
        runs.append(out.copy())
    solver.close()
    rev = sorted(r[0] for r in runs[1:])[reps // 2]
    fwd = sorted(r[1] for r in runs[1:])[reps // 2]
    return {"precision": precision, "reverse_steady_step_ms": rev, "forward_steady_step_with_tape_ms": fwd,
            "reverse_over_forward": rev / fwd, "steps": nsteps,
            "method": "HIP events on the reverse stream around nsteps-1 reverse steady steps (pcf_profile_reverse)"}


ADJOINT_MHD = dict(N=(65, 64, 64), Re=1000.0, Rm=1000.0, omega=2.0 / 3.0, shear_rate=1.0,
                   background_b=(0.0, 0.0, 0.025), domain=((-1.0, 1.0), (0.0, 4.0), (0.0, 1.0)), dt=0.005,
                   energy_u=1e-3, energy_b=1e-3, K=12, seed=0, steps=100, every=20)


def adjoint_mhd_leg(reps: int = 3) -> dict:
    """Value + gradient of Epert + Emag over ADJOINT_MHD['steps'] steps of the
    conducting-wall rotating-shear MHD step (the reference's ``mhd_mri_pcf``
    autograd case: Re = Rm = 1000, Omega = 2/3, S = 1, B0 = (0, 0, 0.025),
    4 x 1 box), FP64, checkpoints every ADJOINT_MHD['every'] steps, against the
    same forward step sequence, state resident."""
    from jaxfun_amd import PlaneCouetteMRIShearpyHIP
    from jaxfun_amd import _native as nat
    from jaxfun_amd.ic import seeded_mhd_state

    a = ADJOINT_MHD
    solver = PlaneCouetteMRIShearpyHIP(N=a["N"], domain=a["domain"], Re=a["Re"], Rm=a["Rm"], omega=a["omega"],
                                       shear_rate=a["shear_rate"], background_b=a["background_b"], dt=a["dt"],
                                       **SBDF3_ROT)
    s0 = seeded_mhd_state(solver, seed=a["seed"], energy_u=a["energy_u"], energy_b=a["energy_b"], K=a["K"])
    lib, h = solver._lib, solver._h

    def fwd():
        solver.load(s0)
        solver.synchronize()
        t0 = time.perf_counter()
        solver.advance(a["steps"])
        solver.synchronize()
        return time.perf_counter() - t0

    def adj():
        solver.load(s0)
        solver.synchronize()
        t0 = time.perf_counter()
        nat.check(lib.pcf_vjp(h, a["steps"], a["every"], nat.SEED_ENERGY))
        return time.perf_counter() - t0

    fwd()
    adj()
    tf = sorted(fwd() for _ in range(reps))[reps // 2]
    ta = sorted(adj() for _ in range(reps))[reps // 2]
    solver.close()
    return {"workload": f"grad of Epert + Emag after {a['steps']} steps wrt the seeded initial state (flow, A), "
                        "conducting-wall rotating-shear MHD 65x64x64, Re=Rm=1000, Omega=2/3, S=1, B0=(0,0,0.025), "
                        f"dt=0.005, checkpoints every {a['every']}, FP64",
            "forward_ms": 1e3 * tf, "value_and_grad_ms": 1e3 * ta, "ratio": ta / tf}


TC_CASES = (
    # production/examples/taylor_couette_hydro_dns_v1.json with the autograd
    # benchmark's SBDF3 (production/autograd_cost_benchmark.py:64-71,396)
    dict(name="tc_hydro 40x8", Nr=40, Nz=8, dealias=1.0),
    dict(name="48x32 dealiased", Nr=48, Nz=32, dealias=1.5),
    # the 3D class: production/examples/taylor_couette_hydro_3d_v1.json and the
    # constructor defaults of TaylorCouetteDNSJax (Nr=24, Ntheta=8, Nz=16)
    dict(name="tc_hydro_3d 12x4x4", Nr=12, Ntheta=4, Nz=4, dealias=1.5),
    dict(name="3D defaults 24x8x16", Nr=24, Ntheta=8, Nz=16, dealias=1.5),
    # conducting-wall MHD: production/examples/taylor_couette_mhd_dns_v1.json
    # (Re = Rm = 1000, B0 = 0.1, Keplerian Omega2, Lz = pi / 3) with SBDF3
    dict(name="tc_mhd 40x8", Nr=40, Nz=8, dealias=1.0, mhd=True),
    dict(name="MHD 48x32 dealiased", Nr=48, Nz=32, dealias=1.5, mhd=True),
    # the 3D MHD class: production/examples/taylor_couette_mhd_3d_v1.json, its
    # production resolution tier, and 24x8x16 at the same physics
    dict(name="tc_mhd_3d 12x4x4", Nr=12, Ntheta=4, Nz=4, dealias=1.5, mhd=True),
    dict(name="3D MHD 24x8x16", Nr=24, Ntheta=8, Nz=16, dealias=1.5, mhd=True),
    dict(name="tc_mhd_3d production tier 24x16x16", Nr=24, Ntheta=16, Nz=16, dealias=1.5, mhd=True),
)


def tc_leg(steps: int = 2400, cpu_steps: int = 60) -> list:
    """Taylor-Couette SBDF3 steps/s on the GPU (graph replay of 12-step
    blocks, state resident) beside the oracle on the host cores, from the
    golden cases' seeds (Re = 100, R1 = 1, R2 = 2, Omega1 = 1, Omega2 = 0,
    Lz = 2 pi / 3.13, dt = 1e-3, amp 1e-6 leading eigenmode: m = 0 for the
    axisymmetric class, m = 1 for the 3D one), and the growth rate against
    the linear one."""
    from jaxfun_amd.tc import (AxisymmetricMRIDNSHIP, AxisymmetricTCDNSHIP, CircularCouette, TaylorCouetteDNSHIP,
                               TaylorCouetteMRIDNSHIP)
    from oracle.tc import CircularCouette as OBase
    from oracle.tc import TC3DMRIOracle, TC3DOracle, TCMRIOracle, TCOracle, TCState

    out = []
    lz = 2.0 * math.pi / 3.13
    for case in TC_CASES:
        kw = dict(nu=0.01, Nr=case["Nr"], Nz=case["Nz"], Lz=lz, dt=1e-3, dealias=case["dealias"],
                  time_integrator="SBDF3")
        nth = case.get("Ntheta", 1)
        mhd = case.get("mhd", False)
        om2 = 0.5**1.5 if mhd else 0.0
        if mhd and nth > 1:
            kw.update(nu=1e-3, eta_mag=1e-3, B0=0.1, Lz=math.pi / 3.0, dt=2e-3, Ntheta=nth)
            solver = TaylorCouetteMRIDNSHIP(CircularCouette(1.0, 2.0, 1.0, om2), **kw)
            s0, eig = solver.seed_linear_eigenmode(m=1, kz_mode=1, amp=1e-7)
        elif mhd:
            kw.update(nu=1e-3, eta_mag=1e-3, B0=0.1, Lz=math.pi / 3.0, dt=2e-3)
            solver = AxisymmetricMRIDNSHIP(CircularCouette(1.0, 2.0, 1.0, om2), **kw)
            s0, eig = solver.seed_linear_eigenmode(kz_mode=1, amp=1e-7)
        elif nth > 1:
            kw["Ntheta"] = nth
            solver = TaylorCouetteDNSHIP(CircularCouette(1.0, 2.0, 1.0, 0.0), **kw)
            s0, eig = solver.seed_linear_eigenmode(m=1, kz_mode=1, amp=1e-6)
        else:
            solver = AxisymmetricTCDNSHIP(CircularCouette(1.0, 2.0, 1.0, 0.0), **kw)
            s0, eig = solver.seed_linear_eigenmode(kz_mode=1, amp=1e-6)
        s = solver.solve(s0, 2)  # start-up steps
        solver.load(s)
        solver.advance(24)
        t0 = time.perf_counter()
        solver.advance(steps)
        el = time.perf_counter() - t0  # advance returns after a stream synchronisation
        fin = solver.fetch(s)
        e0, e1 = solver.energy(s), solver.energy(fin)
        growth = 0.5 * math.log(e1 / e0) / ((steps + 24) * solver.dt)
        ocls = (TC3DMRIOracle if nth > 1 else TCMRIOracle) if mhd else (TC3DOracle if nth > 1 else TCOracle)
        oracle = ocls(OBase(1.0, 2.0, 1.0, om2), **kw)
        os_ = TCState(u=s.x if mhd else s.u, p=s.p, nonlinear_old=s.nonlinear_old, have_old=s.have_old,
                      nonlinear_older=s.nonlinear_older, solution_old=s.solution_old,
                      solution_older=s.solution_older, history_steps=s.history_steps)
        oracle.solve(os_, 2)
        t1 = time.perf_counter()
        oracle.solve(os_, cpu_steps)
        cpu = cpu_steps / (time.perf_counter() - t1)
        if mhd and nth > 1:
            kind = f"3D conducting-wall Taylor-Couette MHD SBDF3, {case['name']} (Nr={case['Nr']}, Ntheta={nth}, "
        elif mhd:
            kind = f"axisymmetric conducting-wall Taylor-Couette MHD SBDF3, {case['name']} (Nr={case['Nr']}, "
        elif nth > 1:
            kind = f"3D Taylor-Couette SBDF3, {case['name']} (Nr={case['Nr']}, Ntheta={nth}, "
        else:
            kind = f"axisymmetric Taylor-Couette SBDF3, {case['name']} (Nr={case['Nr']}, "
        phys = "Re=Rm=1000, B0=0.1, Keplerian, dt=2e-3" if mhd else "Re=100, dt=1e-3"
        out.append({"workload": kind + f"Nz={case['Nz']}, dealias={case['dealias']}), {phys}, FP64",
                    "steps_per_s": steps / el, "ms_per_step": 1e3 * el / steps, "steps": steps,
                    "growth_rate": growth, "growth_rate_linear": eig.real,
                    "cpu_oracle_steps_per_s": cpu, "cpu_sample": f"{cpu_steps} steady steps, numpy/scipy "
                                                                 "oracle, one process",
                    "gpu_vs_cpu": (steps / el) / cpu})
        solver.close()
    return out


def adjoint_fd_check(n_dirs: int = 8, eps: float = 1e-6) -> dict:
    """FP64 gradient against central differences of the HIP forward along
    seeded divergence-free directions (same energy as the initial field), and
    the mixed-precision gradient against the FP64 one."""
    import numpy as np

    from jaxfun_amd import PlaneCouetteFluctuationHIP
    from jaxfun_amd.ic import seeded_state

    a = ADJOINT
    solver = PlaneCouetteFluctuationHIP(N=a["N"], Re=a["Re"], dt=a["dt"], precision="fp64", **SBDF3_ROT)
    s0 = seeded_state(solver, seed=a["seed"], energy=a["energy"], K=a["K"])
    value, grad, _ = solver.value_and_grad(s0, a["steps"], checkpoint_every=a["every"])
    errs = []
    for k in range(n_dirs):
        d = seeded_state(solver, seed=100 + k, energy=a["energy"], K=a["K"])

        def J(sign):
            s = type(s0)(u=tuple(x + sign * eps * y for x, y in zip(s0.u, d.u)), g=s0.g + sign * eps * d.g)
            return solver.diagnostics(solver.solve(s, a["steps"]))["Epert"]

        fd = (J(1.0) - J(-1.0)) / (2 * eps)
        an = float(sum(np.real(np.sum(x * y)) for x, y in zip((*grad.u, grad.g), (*d.u, d.g))))
        errs.append(abs(an - fd) / abs(fd))
    solver.close()
    mixed = PlaneCouetteFluctuationHIP(N=a["N"], Re=a["Re"], dt=a["dt"], precision="mixed", **SBDF3_ROT)
    _, gm, _ = mixed.value_and_grad(s0, a["steps"], checkpoint_every=a["every"])
    mixed.close()
    gerr = max(float(np.max(np.abs(x - y)) / np.max(np.abs(y))) for x, y in zip((*gm.u, gm.g), (*grad.u, grad.g)))
    return {"directions": n_dirs, "eps": eps, "max_rel_err": max(errs), "value": value,
            "mixed_vs_fp64_grad_rel": gerr}


MHD = dict(N=(65, 64, 64), Re=1000.0, Rm=1000.0, omega=0.0, shear_rate=-1.0, background_b=(0.0, 0.0, 0.1),
           dt=0.005, energy_u=1e-3, energy_b=1e-3, K=12, seed=0, steps=500)


def mhd_leg(steps: int, warmup: int) -> dict:
    """Insulating-wall MHD (B0 = (0, 0, 0.1), Omega = 0, S = -1, Re = Rm = 1000,
    dt = 0.005) at 65x64x64 from a seeded state: steady-step throughput in FP64
    and mixed precision, and the mixed run's per-field deviation from FP64 with
    the magnetic diagnostics after MHD['steps'] steps."""
    from jaxfun_amd import PlaneCouetteMRIShearpyInsulatingHIP
    from jaxfun_amd.ic import seeded_mhd_state

    w = MHD
    out = {"workload": "insulating-wall rotating-shear MHD, 65x64x64, Re=Rm=1000, B0=(0,0,0.1), Omega=0, S=-1, "
                       "dt=0.005, E_u=E_b=1e-3 over the lowest 12 modes"}
    finals = {}
    for prec in ("fp64", "mixed"):
        solver = PlaneCouetteMRIShearpyInsulatingHIP(N=w["N"], Re=w["Re"], Rm=w["Rm"], omega=w["omega"],
                                                     shear_rate=w["shear_rate"], background_b=w["background_b"],
                                                     dt=w["dt"], precision=prec, **SBDF3_ROT)
        s0 = seeded_mhd_state(solver, seed=w["seed"], energy_u=w["energy_u"], energy_b=w["energy_b"], K=w["K"])
        solver.load(s0)
        solver.advance(2)
        solver.advance_steady(warmup, graph_block=GRAPH_BLOCK)
        solver.synchronize()
        t0 = time.perf_counter()
        solver.advance_steady(steps, graph_block=GRAPH_BLOCK)
        solver.synchronize()
        el = time.perf_counter() - t0
        stages = profile_stages(solver, 24)
        out[prec] = {"value": steps / el, "unit": "steps/s", "ms_per_step": 1e3 * el / steps, "stages_ms": stages}
        final = solver.solve(s0, w["steps"])
        d = solver.diagnostics(final)
        out[prec]["diagnostics_after"] = {k: d[k] for k in ("Epert", "Emag", "Emag_total", "divB_L2",
                                                            "insulating_bc_residual")}
        finals[prec] = final
        solver.close()
    import numpy as np

    err = {}
    a, b = finals["fp64"], finals["mixed"]
    for name, x, y in zip(("u0", "u1", "u2", "g", "Ax", "Ay", "Az"), list(b.flow.u) + [b.flow.g] + list(b.A),
                          list(a.flow.u) + [a.flow.g] + list(a.A)):
        err[name] = float(np.max(np.abs(x - y)) / np.max(np.abs(y)))
    err["steps"] = w["steps"]
    out["mixed_error_vs_fp64"] = err
    out["ratio_mixed_vs_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    return out


# The reference's other published FP64 throughputs (RTX 5090 Laptop, Re = Rm =
# 400, dt = 1e-3, analytic initial field of amplitude 0.02 / magnetic 0.005,
# benchmark-results/FINAL_LOCAL_CHECKPOINT.md:58-59,64,69,79-81 and
# CHECKPOINT_PHASES_1_3.md:16-18): ms per step.  The two 65^3 integrator figures
# are earlier checkpoints of the same solver; the MHD CNAB2 one used the
# coefficient (gradient-form) nonlinear path, run here in the rotational form.
PUBLISHED = [("hydro", (97, 96, 96), 92.94), ("hydro", (128, 128, 128), 219.72),
             ("mhd_conducting", (65, 64, 64), 45.790), ("mhd_conducting", (97, 96, 96), 163.28),
             ("hydro_imexrk3", (65, 64, 64), 84.740),  # the rotational IMEXRK3 integrator alone
             ("mhd_cnab2", (65, 64, 64), 49.73)]  # conducting MHD, CNAB2, batched solves
STEPWISE = {"hydro_imexrk3": "IMEXRK3", "mhd_cnab2": "CNAB2"}  # launched step by step (pcf_step)


def published_leg(steps: int = 120, warmup: int = 24) -> list:
    """Steady FP64 steps/s of this build on each published configuration (same
    solver class, N, Re, Rm, dt and initial field), beside the published rate."""
    from jaxfun_amd import PlaneCouetteFluctuationHIP, PlaneCouetteMHDHIP

    out = []
    for kind, N, ms in PUBLISHED:
        integ = STEPWISE.get(kind, "SBDF3")
        if kind.startswith("hydro"):
            solver = PlaneCouetteFluctuationHIP(N=N, Re=400.0, dt=1e-3, perturbation_amplitude=0.02,
                                                time_integrator=integ, nonlinear_form="rotational")
        else:
            solver = PlaneCouetteMHDHIP(N=N, Re=400.0, Rm=400.0, dt=1e-3, perturbation_amplitude=0.02,
                                        magnetic_amplitude=0.005, time_integrator=integ,
                                        nonlinear_form="rotational")
        solver.load(solver.initial_state())
        if kind in STEPWISE:
            solver.advance(warmup)
            solver.synchronize()
            t0 = time.perf_counter()
            solver.advance(steps)
        else:
            solver.advance(2)
            solver.advance_steady(warmup, graph_block=GRAPH_BLOCK)
            solver.synchronize()
            t0 = time.perf_counter()
            solver.advance_steady(steps, graph_block=GRAPH_BLOCK)
        solver.synchronize()
        el = time.perf_counter() - t0
        solver.close()
        value = steps / el
        out.append({"solver": kind, "N": list(N), "dtype": "f64", "value": value, "unit": "steps/s",
                    "published_steps_per_s": 1e3 / ms, "vs_published": value * ms / 1e3})
        print(f"# published config {kind} {N}: {value:.1f} steps/s ({value * ms / 1e3:.0f}x)", file=sys.stderr,
              flush=True)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2400)
    ap.add_argument("--warmup", type=int, default=240)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--error-steps", type=int, default=1000,
                    help="steps of the mixed-vs-FP64 per-field error check (0: skip)")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--large-steps", type=int, default=LARGE["steps"],
                    help="steps of the 97^3 / 129^3 mixed-vs-FP64 runs (0: skip)")
    ap.add_argument("--no-mhd", action="store_true")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the adjoint timing leg (for kernel profiles)")
    ap.add_argument("--tc-only", action="store_true",
                    help="run only the Taylor-Couette leg (for kernel profiles)")
    ap.add_argument("--no-tc", action="store_true")
    ap.add_argument("--mhd-adjoint-only", action="store_true",
                    help="run only the MHD value-and-gradient timing (for kernel profiles)")
    ap.add_argument("--reverse-only", action="store_true",
                    help="run only the reverse-step timing (pcf_profile_reverse; for kernel profiles)")
    ap.add_argument("--no-published", action="store_true",
                    help="skip the FP64 runs on the reference's other published configurations")
    ap.add_argument("--mhd-only", action="store_true",
                    help="run only the insulating-wall MHD leg (for kernel profiles)")
    ap.add_argument("--pcie-steps", type=int, default=240,
                    help="steps of the host-boundary (PCIe-inclusive) solve timing (0: skip)")
    ap.add_argument("--eager", action="store_true",
                    help="launch steps one by one instead of replaying HIP graphs (for counter runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="also run a device-to-device copy of MB megabytes (HBM counter calibration)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(this build has no multi-GPU path)", file=sys.stderr)
        return 2
    if args.adjoint_only:
        print(json.dumps({"adjoint": {p: adjoint_leg(p) for p in ("fp64", "mixed")}}), flush=True)
        return 0
    if args.reverse_only:
        print(json.dumps({"reverse": {p: reverse_leg(p) for p in ("fp64", "mixed")}}), flush=True)
        return 0
    if args.mhd_adjoint_only:
        print(json.dumps({"adjoint_mhd": adjoint_mhd_leg()}), flush=True)
        return 0
    if args.tc_only:
        print(json.dumps({"taylor_couette": tc_leg()}), flush=True)
        return 0
    if args.steps < 1 or args.warmup < 0:
        print("--steps must be >= 1 and --warmup >= 0", file=sys.stderr)
        return 2
    steps, warmup = args.steps, args.warmup
    if args.mhd_only:
        print(json.dumps({"mhd": mhd_leg(min(steps, 1200), warmup)}), flush=True)
        return 0

    solver, state, el, stages, diag = run_gpu("fp64", steps, warmup, args.eager)
    value = steps / el
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": 1e3 * el / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": value / PUBLISHED_STEPS_PER_S,
        "dtype": "f64",
        "data": "synthetic (seeded divergence-free perturbation, seed 0)",
        "config": {"workload": "plane-Couette KMM SBDF3 rotational step, 65x64x64, Re=1000, "
                               "dt=0.005, E=1e-3 over the lowest 12 modes, FP64",
                   "N": list(WORKLOAD["N"]), "Re": WORKLOAD["Re"], "dt": WORKLOAD["dt"],
                   "padding": [1.0, 1.5, 1.5], "graph_block_steps": GRAPH_BLOCK},
        "stages_ms": stages,
        "roofline": roofline(solver, "fp64", stages),
        "stage_rooflines": stage_rooflines(solver, "fp64", stages),
        "protocol": block_protocol(solver),
        "final_diagnostics": diag,
        "adjoint_forward_ratio": None,
    }
    print(f"# fp64: {value:.1f} steps/s ({1e3 * el / steps:.4f} ms/step)", file=sys.stderr, flush=True)
    solver.close()
    if not args.no_mixed:
        msolver, _, mel, mstages, mdiag = run_gpu("mixed", steps, warmup, args.eager)
        mval = steps / mel
        line["mixed"] = {"value": mval, "unit": "steps/s", "ms_per_step": 1e3 * mel / steps,
                         "ratio_vs_fp64": mval / value, "stages_ms": mstages,
                         "roofline": roofline(msolver, "mixed", mstages),
                         "stage_rooflines": stage_rooflines(msolver, "mixed", mstages),
                         "protocol": block_protocol(msolver)}
        line["mixed"]["protocol_ratio_vs_fp64"] = (line["mixed"]["protocol"]["steps_per_s_p50"]
                                                   / line["protocol"]["steps_per_s_p50"])
        print(f"# mixed: {mval:.1f} steps/s", file=sys.stderr, flush=True)
        msolver.close()
        if args.error_steps:
            line["mixed"]["error_vs_fp64"] = mixed_error(state, args.error_steps)
            print(f"# mixed error after {args.error_steps} steps: "
                  f"{line['mixed']['error_vs_fp64']['velocity']:.3e}", file=sys.stderr, flush=True)
    if args.pcie_steps:
        line["pcie_inclusive"] = pcie_inclusive(state, args.pcie_steps)
        line["state_exchange"] = state_exchange(state)
    if not args.no_adjoint:
        adj = {"workload": f"grad of Epert after {ADJOINT['steps']} steps wrt the seeded initial field, "
                           f"65x64x64, Re=1000, dt=0.005, E=1e-4, checkpoints every {ADJOINT['every']}",
               "mixed": adjoint_leg("mixed"), "fp64": adjoint_leg("fp64")}
        adj["fd_check"] = adjoint_fd_check()
        adj["reverse_step"] = {p: reverse_leg(p) for p in ("fp64", "mixed")}
        adj["mhd"] = adjoint_mhd_leg()
        line["adjoint"] = adj
        line["adjoint_forward_ratio"] = adj["mixed"]["ratio"]
        print(f"# adjoint/forward: mixed {adj['mixed']['ratio']:.3f}, fp64 {adj['fp64']['ratio']:.3f}; "
              f"FD max rel err {adj['fd_check']['max_rel_err']:.2e}", file=sys.stderr, flush=True)
    if not args.no_mhd:
        line["mhd"] = mhd_leg(min(steps, 1200), warmup)
        print(f"# mhd (insulating): fp64 {line['mhd']['fp64']['value']:.1f}, mixed "
              f"{line['mhd']['mixed']['value']:.1f} steps/s; mixed error "
              f"{max(v for k, v in line['mhd']['mixed_error_vs_fp64'].items() if k != 'steps'):.2e}",
              file=sys.stderr, flush=True)
    if args.large_steps:
        line["large_mixed"] = {"workload": "plane-Couette SBDF3 rotational, 97x96x96 and 129x128x128, Re=2000, "
                                           "dt=0.002, E=1e-2 over the lowest 12 modes, mixed vs FP64",
                               "runs": large_mixed_leg(args.large_steps)}
    if not args.no_tc:
        line["taylor_couette"] = tc_leg()
        print("# taylor-couette: " + ", ".join(f"{c['steps_per_s']:.0f} steps/s" for c in line["taylor_couette"]),
              file=sys.stderr, flush=True)
    if not args.no_published:
        line["published_configs"] = published_leg()
    if args.copy_calibration:
        line["copy_calibration"] = copy_calibration(args.copy_calibration)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(state, args.cpu_seconds)
        print(f"# cpu: {line['cpu_baseline']['value']:.3f} steps/s", file=sys.stderr, flush=True)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
