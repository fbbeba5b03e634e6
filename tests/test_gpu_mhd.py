"""GPU parity of the MHD steps (conducting walls, rotating shear with conducting
and insulating walls): HIP against the MHD oracles and the reference's
published MHD goldens.

Bars as for the hydrodynamic step: state leaves (flow and vector potential)
within 1e-12 relative per leaf group of the FP64 oracle, nonlinear history rows
within 5e-11 (their rounding sensitivity, see test_gpu_parity.py), published
energies within 1e-10 relative and checksums within 1e-11 (l2) / 1e-9 (sum).
"""

import json
import math
from pathlib import Path

import numpy as np
import pytest

from jaxfun_amd import PlaneCouetteMHDHIP, mhd_state_checksum

from tests.helpers import mhd_max_rel_errors, mhd_oracle_for, to_mhd_oracle

pytestmark = pytest.mark.gpu

GOLDEN = json.loads((Path(__file__).parent / "golden" / "reference_mhd_published.json").read_text())
FP64_TOL = 1e-12
HIST_TOL = 5e-11


def make(N, precision="fp64", **kw):
    c = GOLDEN["config"]
    args = dict(Re=c["Re"], Rm=c["Rm"], dt=c["dt"], perturbation_amplitude=c["velocity_amplitude"],
                magnetic_amplitude=c["magnetic_amplitude"], time_integrator="SBDF3", nonlinear_form="rotational")
    args.update(kw)
    return PlaneCouetteMHDHIP(N=N, precision=precision, **args)


def assert_parity(err, where):
    hist = {k: v for k, v in err.items() if k.startswith(("nl_", "nonlinear_"))}
    rest = {k: v for k, v in err.items() if k not in hist}
    assert max(rest.values()) < FP64_TOL, (where, rest)
    assert max(hist.values()) < HIST_TOL, (where, hist)


@pytest.mark.parametrize("N,kw", [((17, 16, 16), {}), ((24, 16, 32), {}),
                                  ((17, 16, 16), dict(magnetic_amplitude=0.2, Rm=50.0)),
                                  ((17, 16, 16), dict(magnetic_amplitude=0.1, background_b=(0.0, 0.0, 0.1)))],
                         ids=["17", "24x16x32", "strong-field", "background-B0"])
def test_fp64_mhd_steps_match_oracle(N, kw):
    solver = make(N, **kw)
    oracle = mhd_oracle_for(solver)
    s = solver.initial_state()
    so = to_mhd_oracle(s)
    done = 0
    for k in (1, 2, 3, 7):  # two IMEXRK3 start-up steps, then steady SBDF3
        s = solver.solve(s, k - done)
        so = oracle.solve(so, k - done)
        done = k
        err = mhd_max_rel_errors(s, so)
        print(N, kw, k, max(err.values()), flush=True)
        assert_parity(err, (N, k))
    d, do = solver.diagnostics(s), oracle.diagnostics(so)
    for key in ("Epert", "Etot", "Emag"):
        assert math.isclose(d[key], do[key], rel_tol=1e-12), (key, d[key], do[key])
    assert d["divB_L2"] < 1e-13 and do["divB_L2"] < 1e-13


def test_mhd_graph_replay_matches_eager():
    solver = make((17, 16, 16), magnetic_amplitude=0.1)
    s = solver.solve(solver.initial_state(), 3)
    solver.load(s)
    solver.advance_steady(24)
    eager = solver.fetch()
    solver.load(s)
    solver.advance_steady(24, graph_block=12)
    graph = solver.fetch()
    assert max(mhd_max_rel_errors(graph, eager).values()) == 0.0


def test_mhd_mixed_precision_close_to_fp64():
    N = (33, 32, 32)
    a, b = make(N, magnetic_amplitude=0.05), make(N, precision="mixed", magnetic_amplitude=0.05)
    s0 = a.initial_state()
    sa, sb = a.solve(s0, 40), b.solve(s0, 40)
    for x, y in zip(list(sb.flow.u) + list(sb.A), list(sa.flow.u) + list(sa.A)):
        rel = float(np.max(np.abs(x - y)) / np.max(np.abs(y)))
        print("mixed vs fp64", rel, flush=True)
        assert rel < 2e-6


@pytest.mark.parametrize("case", GOLDEN["cases"], ids=lambda c: "x".join(map(str, c["N"])))
def test_published_mhd_golden(case):
    """Reference-published FP64 MHD results after the benchmark's step count."""
    solver = make(tuple(case["N"]))
    s = solver.solve(solver.initial_state(), case["steps"])
    d = solver.diagnostics(s)
    g = case["diagnostics"]
    print(case["N"], {k: (d[k], g[k]) for k in ("Epert", "Emag")}, flush=True)
    for k in ("Epert", "Etot", "Emag"):
        assert math.isclose(d[k], g[k], rel_tol=1e-10), (k, d[k], g[k])
    for k in ("wall_shear_top", "wall_shear_bot", "u_top", "u_bot", "mean_shear"):
        assert abs(d[k] - g[k]) < 1e-13, (k, d[k], g[k])
    assert d["divL2"] < 1e-14 and d["divB_L2"] < 1e-15
    cs = mhd_state_checksum(s)
    gc = case["checksum"]
    assert cs["leaf_count"] == gc["leaf_count"] and cs["element_count"] == gc["element_count"]
    assert math.isclose(cs["l2_norm"], gc["l2_norm"], rel_tol=1e-11)
    assert math.isclose(cs["real_sum"], gc["real_sum"], rel_tol=1e-9)


def test_mhd_reverse_mode_scope():
    """The conducting-wall class differentiates (gradients checked in
    tests/test_gpu_adjoint_mhd.py); a non-SBDF3 integrator is refused."""
    solver = make((17, 16, 16))
    value, grad, _ = solver.value_and_grad(solver.initial_state(), 2)
    assert np.isfinite(value) and all(np.all(np.isfinite(np.asarray(a))) for a in grad.A)
    cn = make((17, 16, 16), time_integrator="CNAB2")
    with pytest.raises(NotImplementedError):
        cn.value_and_grad(cn.initial_state(), 2)


# ---------------------------------------------------------------- insulating walls
def make_insulating(N, precision="fp64", **kw):
    from jaxfun_amd import PlaneCouetteMRIShearpyInsulatingHIP

    args = dict(Re=400.0, Rm=300.0, omega=0.4, shear_rate=1.5, background_b=(0.0, 0.02, 0.1), dt=2e-3,
                perturbation_amplitude=0.05, magnetic_amplitude=0.3, time_integrator="SBDF3",
                nonlinear_form="rotational")
    args.update(kw)
    return PlaneCouetteMRIShearpyInsulatingHIP(N=N, precision=precision, **args)


def test_insulating_initial_state_matches_oracle():
    from tests.helpers import insulating_oracle_for

    solver = make_insulating((17, 16, 16))
    err = mhd_max_rel_errors(solver.initial_state(), insulating_oracle_for(solver).initial_state())
    assert max(err.values()) < 1e-13, err


@pytest.mark.parametrize("N,kw", [((17, 16, 16), {}), ((24, 16, 32), dict(omega=0.0, shear_rate=-1.0)),
                                  ((33, 32, 32), dict(Re=1000.0, Rm=1000.0, omega=0.0, shear_rate=-1.0,
                                                      background_b=(0.0, 0.0, 0.1), dt=5e-3))],
                         ids=["rot-shear", "couette", "config3-like"])
def test_insulating_steps_match_oracle(N, kw):
    from jaxfun_amd.ic import seeded_mhd_state
    from tests.helpers import insulating_oracle_for

    solver = make_insulating(N, **kw)
    oracle = insulating_oracle_for(solver)
    s = seeded_mhd_state(solver, seed=3, energy_u=1e-3, energy_b=1e-3, K=4)
    so = to_mhd_oracle(s)
    assert oracle.insulating_bc_residual(so) < 1e-13  # the seed meets the wall rows
    done = 0
    for k in (1, 2, 3, 6):
        s = solver.solve(s, k - done)
        so = oracle.solve(so, k - done)
        done = k
        err = mhd_max_rel_errors(s, so)
        print(N, kw, k, max(err.values()), flush=True)
        assert_parity(err, (N, k))
    d, do = solver.diagnostics(s), oracle.diagnostics(so)
    for key in ("Epert", "Etot", "Emag", "Emag_total"):
        assert math.isclose(d[key], do[key], rel_tol=1e-12), (key, d[key], do[key])
    assert d["insulating_bc_residual"] < 1e-10 and do["insulating_bc_residual"] < 1e-10
    assert d["divB_L2"] < 1e-12


def test_insulating_graph_replay_and_mixed_precision():
    from jaxfun_amd.ic import seeded_mhd_state

    N = (33, 32, 32)
    kw = dict(Re=1000.0, Rm=1000.0, omega=0.0, shear_rate=-1.0, background_b=(0.0, 0.0, 0.1), dt=5e-3)
    a = make_insulating(N, **kw)
    s0 = seeded_mhd_state(a, seed=0, energy_u=1e-3, energy_b=1e-3, K=8)
    s = a.solve(s0, 3)
    a.load(s)
    a.advance_steady(24)
    eager = a.fetch()
    a.load(s)
    a.advance_steady(24, graph_block=12)
    assert max(mhd_max_rel_errors(a.fetch(), eager).values()) == 0.0
    b = make_insulating(N, precision="mixed", **kw)
    sa, sb = a.solve(s0, 30), b.solve(s0, 30)
    for x, y in zip(list(sb.flow.u) + list(sb.A), list(sa.flow.u) + list(sa.A)):
        rel = float(np.max(np.abs(x - y)) / np.max(np.abs(y)))
        print("insulating mixed vs fp64", rel, flush=True)
        assert rel < 2e-6
    assert b.diagnostics(sb)["insulating_bc_residual"] < 1e-10


# ---------------------------------------------------------------- rotating shear, conducting walls
def make_shear(N, precision="fp64", **kw):
    from jaxfun_amd import PlaneCouetteMRIShearpyHIP

    args = dict(Re=400.0, Rm=300.0, omega=0.4, shear_rate=1.5, background_b=(0.0, 0.02, 0.1), dt=2e-3,
                perturbation_amplitude=0.05, magnetic_amplitude=0.3, time_integrator="SBDF3",
                nonlinear_form="rotational")
    args.update(kw)
    return PlaneCouetteMRIShearpyHIP(N=N, precision=precision, **args)


def test_shear_initial_states_match_oracle():
    from tests.helpers import shear_oracle_for

    solver = make_shear((17, 16, 16))
    o = shear_oracle_for(solver)
    err = mhd_max_rel_errors(solver.initial_state(), o.initial_state())
    assert max(err.values()) < 1e-13, err
    # sinusoidal_bz_x: Ay = -(a / pi)(cos(pi x) + 1), uniform in y and z, projected onto TD
    s = make_shear((17, 16, 16), magnetic_seed="sinusoidal_bz_x").initial_state()
    x = o.mesh()[0]
    ay = np.broadcast_to(-(0.3 / np.pi) * (np.cos(np.pi * x) + 1.0), (o.N0, o.N[1], o.N[2])).astype(complex)
    want = o.mask(o.tp_forward(ay, o.SD, o.MD1))
    assert np.max(np.abs(s.A[1] - want)) <= 1e-13 * np.max(np.abs(want))
    assert np.max(np.abs(s.A[0])) == 0.0 and np.max(np.abs(s.A[2])) == 0.0


@pytest.mark.parametrize("N,seeded", [((17, 16, 16), False), ((24, 16, 32), True)], ids=["17-initial", "24x16x32-seeded"])
def test_shear_steps_match_oracle(N, seeded):
    from jaxfun_amd.ic import seeded_mhd_state
    from tests.helpers import shear_oracle_for

    solver = make_shear(N)
    oracle = shear_oracle_for(solver)
    s = seeded_mhd_state(solver, seed=3, energy_u=1e-3, energy_b=1e-3, K=4) if seeded else solver.initial_state()
    so = to_mhd_oracle(s)
    done = 0
    for k in (1, 2, 3, 6):
        s = solver.solve(s, k - done)
        so = oracle.solve(so, k - done)
        done = k
        err = mhd_max_rel_errors(s, so)
        print(N, seeded, k, max(err.values()), flush=True)
        assert_parity(err, (N, k))
    d, do = solver.diagnostics(s), oracle.diagnostics(so)
    for key in ("Epert", "Etot", "Emag", "Emag_total"):
        assert math.isclose(d[key], do[key], rel_tol=1e-12), (key, d[key], do[key])
    assert d["divB_L2"] < 1e-12 and d["q_shear"] == 1.5 / 0.4


def test_shear_graph_replay_and_mixed_precision():
    N = (33, 32, 32)
    a = make_shear(N)
    s0 = a.initial_state()
    s = a.solve(s0, 3)
    a.load(s)
    a.advance_steady(24)
    eager = a.fetch()
    a.load(s)
    a.advance_steady(24, graph_block=12)
    assert max(mhd_max_rel_errors(a.fetch(), eager).values()) == 0.0
    b = make_shear(N, precision="mixed")
    sa, sb = a.solve(s0, 30), b.solve(s0, 30)
    for x, y in zip(list(sb.flow.u) + list(sb.A), list(sa.flow.u) + list(sa.A)):
        rel = float(np.max(np.abs(x - y)) / np.max(np.abs(y)))
        print("shear mixed vs fp64", rel, flush=True)
        assert rel < 2e-6


# ---------------------------------------------------------------- IMEXRK222 (ARS 2,2,2)
@pytest.mark.parametrize("kind", ["conducting", "shear", "insulating", "insulating-config3"])
def test_imexrk222_mhd_steps_match_oracle(kind):
    """The ARS step of the three MHD classes against the oracles, which write
    the reference's second stage literally (explicit L x_1 and eta-Laplacian
    of A_1); histories pass through on both sides."""
    from jaxfun_amd.ic import seeded_mhd_state
    from tests.helpers import insulating_oracle_for, shear_oracle_for

    if kind == "conducting":
        solver = make((17, 16, 16), magnetic_amplitude=0.1, background_b=(0.0, 0.0, 0.1), time_integrator="IMEXRK222")
        oracle = mhd_oracle_for(solver)
        s = solver.initial_state()
    elif kind == "shear":
        solver = make_shear((17, 16, 16), time_integrator="IMEXRK222")
        oracle = shear_oracle_for(solver)
        s = solver.initial_state()
    elif kind == "insulating":
        solver = make_insulating((17, 16, 16), time_integrator="IMEXRK222")
        oracle = insulating_oracle_for(solver)
        s = seeded_mhd_state(solver, seed=3, energy_u=1e-3, energy_b=1e-3, K=4)
    else:
        solver = make_insulating((33, 32, 32), Re=1000.0, Rm=1000.0, omega=0.0, shear_rate=-1.0,
                                 background_b=(0.0, 0.0, 0.1), dt=5e-3, time_integrator="IMEXRK222")
        oracle = insulating_oracle_for(solver)
        s = seeded_mhd_state(solver, seed=3, energy_u=1e-3, energy_b=1e-3, K=4)
    so = to_mhd_oracle(s)
    done = 0
    for k in (1, 2, 5):
        s = solver.solve(s, k - done)
        so = oracle.solve(so, k - done)
        done = k
        err = mhd_max_rel_errors(s, so)
        print(kind, k, max(err.values()), flush=True)
        assert_parity(err, (kind, k))
    d, do = solver.diagnostics(s), oracle.diagnostics(so)
    for key in ("Epert", "Emag"):
        assert math.isclose(d[key], do[key], rel_tol=1e-12), (key, d[key], do[key])
    assert d["divB_L2"] < 1e-12
    if kind.startswith("insulating"):
        assert d["insulating_bc_residual"] < 1e-10
