"""Write the oracle fixture of the seeded 65x64x64 plane-Couette run (CPU, minutes).

Workload: 65 Shen/Chebyshev points x 64 x 64 Fourier modes, Re = 1000,
dt = 0.005, SBDF3 (two IMEXRK3 start-up steps), rotational form, FP64; the
seeded divergence-free initial field of ``jaxfun_amd.ic.seeded_state`` with
seed 0, energy 1e-3 over the lowest 12 modes per direction.  The FP64 oracle
(``oracle/kmm.py``) is stepped 1000 times; at the recorded steps the fixture
keeps

* the diagnostics of ``PlaneCouetteFluctuationJax.diagnostics``
  (``examples/pcf_fluctuations_jax.py:131-161``);
* per leaf of the state pytree (``examples/channelflow_kmm.py:65-76``): the
  l2 norm, the real sum and the imaginary sum;
* every wall-normal coefficient of u0, u1, u2 and g at the streamwise indices
  |ky| <= 2 and spanwise indices 0 <= kz <= 2 (the modes that hold the
  energy), as little-endian float64 (re, im) pairs in base64.

Usage: python tests/golden/make_pcf_seeded_fixture.py [--steps 1000] [--workers 8]
Writes tests/golden/pcf_seeded_65x64x64_re1000.json.
"""

from __future__ import annotations

import argparse
import base64
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from jaxfun_amd.ic import seeded_state  # noqa: E402
from oracle.kmm import PCFOracle, set_workers  # noqa: E402
from tests.helpers import HostSolver, leaves, to_oracle, LEAF_NAMES  # noqa: E402

N = (65, 64, 64)
WORKLOAD = dict(Re=1000.0, dt=0.005, seed=0, energy=1e-3, K=12)
RECORD = (2, 100, 250, 500, 1000)
SAMPLE_Y = (0, 1, 2, -2, -1)
SAMPLE_Z = (0, 1, 2)
OUT = Path(__file__).resolve().parent / "pcf_seeded_65x64x64_re1000.json"


def sample(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a[:, list(SAMPLE_Y)][:, :, list(SAMPLE_Z)])


def encode(a: np.ndarray) -> str:
    return base64.b64encode(np.ascontiguousarray(a, dtype="<c16").tobytes()).decode()


def record(oracle, s) -> dict:
    ls = leaves(s)
    return {
        "diagnostics": oracle.diagnostics(s),
        "leaves": {name: {"l2_norm": float(np.linalg.norm(np.asarray(v))),
                          "real_sum": float(np.sum(np.real(v))), "imag_sum": float(np.sum(np.imag(v)))}
                   for name, v in zip(LEAF_NAMES, ls)},
        "history": [s.history_steps, s.have_old, s.previous_dt],
        "samples": {name: encode(sample(v)) for name, v in zip(("u0", "u1", "u2", "g"), list(s.u) + [s.g])},
    }


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=RECORD[-1])
    ap.add_argument("--workers", type=int, default=8)
    args = ap.parse_args()
    set_workers(args.workers)
    host = HostSolver(N)
    s0 = seeded_state(host, seed=WORKLOAD["seed"], energy=WORKLOAD["energy"], K=WORKLOAD["K"])
    oracle = PCFOracle(N=N, Re=WORKLOAD["Re"], dt=WORKLOAD["dt"])
    s = to_oracle(s0)
    out = {"N": list(N), "workload": WORKLOAD, "sample_ky_index": list(SAMPLE_Y),
           "sample_kz_index": list(SAMPLE_Z), "records": {}}
    done = 0
    t0 = time.time()
    for k in RECORD:
        if k > args.steps:
            break
        s = oracle.solve(s, k - done)
        done = k
        out["records"][str(k)] = record(oracle, s)
        print(f"step {k}: Epert={out['records'][str(k)]['diagnostics']['Epert']:.15e} "
              f"({time.time() - t0:.0f} s)", flush=True)
    OUT.write_text(json.dumps(out, indent=1))
    print(OUT)


if __name__ == "__main__":
    main()
