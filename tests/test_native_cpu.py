"""Host-side checks of libpcfdns.so that need no GPU.

* the library loads and exports every entry point include/pcf_dns.h declares;
* the closed-form setup matrices agree with the oracle's independent
  brute-force quadrature and transforms;
* configuration errors surface as the reference's exception types;
* the seeded initial condition is divergence-free, Hermitian and at the
  requested energy.
"""

import ctypes as C
import math
import re
from pathlib import Path

import numpy as np
import pytest

from jaxfun_amd import _native as nat
from oracle.kmm import (
    PCFOracle,
    cheb_backward,
    cheb_derivative,
    galerkin_matrix,
    stencil_clamped,
    stencil_dirichlet,
)

ROOT = Path(__file__).resolve().parents[1]


def cfg_for(n0, ny=16, nz=16, **kw):
    c = nat.PcfConfig()
    c.abi_version = nat.ABI_VERSION
    c.n0, c.ny, c.nz = n0, ny, nz
    c.lx[0], c.lx[1] = -1.0, 1.0
    c.ly, c.lz = 4 * math.pi, 2 * math.pi
    c.nu, c.u_wall, c.dt = 1 / 400, 1.0, 1e-3
    c.padding_y = c.padding_z = 1.5
    c.precision = nat.PREC_FP64
    c.time_integrator = nat.TI_SBDF3
    c.nonlinear_form = nat.NL_ROTATIONAL
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def test_library_exports_every_declared_symbol():
    lib = nat.load()
    header = (ROOT / "include" / "pcf_dns.h").read_text()
    declared = set(re.findall(r"^\s*(?:int32_t|int64_t|const char \*)\s*\*?\s*(pcf_\w+)\s*\(", header, re.M))
    assert declared, "no declarations parsed"
    assert declared == set(nat.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.pcf_abi_version() == nat.ABI_VERSION


@pytest.mark.parametrize("n0", [9, 17, 33, 65])
def test_galerkin_blocks_match_quadrature(n0):
    cfg = cfg_for(n0)
    SB, SD = stencil_clamped(n0), stencil_dirichlet(n0)
    ref = {"C0B": galerkin_matrix(SB, SB, 0), "C2B": galerkin_matrix(SB, SB, 2),
           "C4B": galerkin_matrix(SB, SB, 4), "C0D": galerkin_matrix(SD, SD, 0),
           "C2D": galerkin_matrix(SD, SD, 2)}
    for name, R in ref.items():
        M = nat.setup_matrix(cfg, name)
        assert M.shape == R.shape
        # the quadrature sums entries up to ~1e9 in FP64: ~1e-13 relative rounding
        assert np.max(np.abs(M - R)) <= 1e-12 * np.max(np.abs(R)), name


def test_nodal_transforms_match_oracle():
    n0 = 17
    cfg = cfg_for(n0)
    SB, SD = stencil_clamped(n0), stencil_dirichlet(n0)
    VB, VD, VD1, VB1 = (nat.setup_matrix(cfg, k) for k in ("VB", "VD", "VD1", "VB1"))
    rng = np.random.default_rng(0)
    cb, cd = rng.standard_normal(n0 - 4), rng.standard_normal(n0 - 2)
    assert np.allclose(VB @ cb, cheb_backward(SB.T @ cb), atol=1e-13)
    assert np.allclose(VD @ cd, cheb_backward(SD.T @ cd), atol=1e-13)
    assert np.allclose(VD1 @ cd, cheb_backward(cheb_derivative(SD.T @ cd, 1)), atol=1e-12)
    assert np.allclose(VB1 @ cb, cheb_backward(cheb_derivative(SB.T @ cb, 1)), atol=1e-12)


def test_projection_and_nonlinear_rows_match_oracle():
    """FD/FB/A1/A0/AG/R against the oracle's transform + mass-solve route."""
    N = (17, 8, 8)
    cfg = cfg_for(N[0], N[1], N[2])
    o = PCFOracle(N=N, Re=400.0, dt=1e-3)
    rng = np.random.default_rng(3)
    vals = rng.standard_normal((N[0], N[1], N[2])) + 1j * rng.standard_normal((N[0], N[1], N[2]))
    # projection of nodal values along x only (Fourier axes untouched)
    from oracle.kmm import cheb_scalar_product

    L = np.tensordot(o.SD, cheb_scalar_product(vals, N[0]), axes=(1, 0))
    ref_fd = np.linalg.solve(o.MD1, L.reshape(N[0] - 2, -1)).reshape(L.shape)
    FD = nat.setup_matrix(cfg, "FD")
    assert np.allclose(np.tensordot(FD, vals, axes=(1, 0)), ref_fd, atol=1e-13)
    L = np.tensordot(o.SB, cheb_scalar_product(vals, N[0]), axes=(1, 0))
    ref_fb = np.linalg.solve(o.MB1, L.reshape(N[0] - 4, -1)).reshape(L.shape)
    FB = nat.setup_matrix(cfg, "FB")
    assert np.allclose(np.tensordot(FB, vals, axes=(1, 0)), ref_fb, atol=1e-12)
    # Nu / Ng rows: H given as nodal values; oracle projects then differentiates
    H = tuple(rng.standard_normal(vals.shape) + 1j * rng.standard_normal(vals.shape) for _ in range(3))
    H = tuple(o.mask(h) for h in H)
    Hd = tuple(np.tensordot(FD, h, axes=(1, 0)) for h in H)
    Nu, Ng, Nv, Nw = o.nonlinear_rhs(Hd)
    A1, A0, AG = (nat.setup_matrix(cfg, k) for k in ("A1", "A0", "AG"))
    ky, kz = o.K1e, o.K2e
    k2 = o.K1**2 + o.K2**2
    c1 = 1j * ky * H[1] + 1j * kz * H[2]
    cg = 1j * kz * H[1] - 1j * ky * H[2]
    mine_u = o.W * (np.tensordot(A1, c1, axes=(1, 0)) + np.tensordot(A0, k2 * H[0], axes=(1, 0)))
    mine_g = o.W * np.tensordot(AG, cg, axes=(1, 0))
    assert np.max(np.abs(mine_u - Nu)) <= 1e-12 * np.max(np.abs(Nu))
    assert np.max(np.abs(mine_g - Ng)) <= 1e-13 * np.max(np.abs(Ng))
    assert np.allclose(-AG @ np.real(H[1][:, 0, 0]), Nv, atol=1e-13)
    # reconstruction f = P_D d/dx u0
    R = nat.setup_matrix(cfg, "R")
    u0 = rng.standard_normal((N[0] - 4, 1, 1)) + 0j
    a = cheb_derivative(o.to_orth(u0, o.SB), 1)
    w = np.tensordot(o.SD, a * o.h[:, None, None], axes=(1, 0))
    f_ref = np.linalg.solve(o.MD1, w.reshape(N[0] - 2, -1)).reshape(w.shape)
    assert np.allclose(np.tensordot(R, u0, axes=(1, 0)), f_ref, atol=1e-12)
    # the reconstruction folded into the nodal contraction: nodal values of f
    # and of df/dx against the oracle's backward transform of f
    VDR, VD1R = nat.setup_matrix(cfg, "VDR"), nat.setup_matrix(cfg, "VD1R")
    f_orth = o.to_orth(f_ref, o.SD)[:, 0, 0]
    assert np.allclose(VDR @ u0[:, 0, 0], cheb_backward(f_orth), atol=1e-12)
    assert np.allclose(VD1R @ u0[:, 0, 0], cheb_backward(cheb_derivative(f_orth, 1)), atol=1e-11)


def test_mhd_current_matrices_match_oracle_projection():
    """VQ1 / VQ2 = VD P_D D^k S_D^T against the oracle's Galerkin projection
    (composite.py:293-299) of the x-derivatives of Dirichlet fields."""
    from oracle.mhd import MHDOracle

    N = (17, 8, 8)
    cfg = cfg_for(N[0], N[1], N[2])
    o = MHDOracle(N=N, Re=400.0, dt=1e-3)
    rng = np.random.default_rng(5)
    a = rng.standard_normal((N[0] - 2, 1, 1)) + 0j
    VD = nat.setup_matrix(cfg, "VD")
    for k, name in ((1, "VQ1"), (2, "VQ2")):
        Q = nat.setup_matrix(cfg, name)
        ref = np.tensordot(VD, o.project_D(cheb_derivative(o.to_orth(a, o.SD), k)), axes=(1, 0))
        got = np.tensordot(Q, a, axes=(1, 0))
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref)), name


def test_structured_setup_accepts_all_benchmark_sizes():
    for n0 in (9, 17, 33, 65, 97, 128, 129):
        M = nat.setup_matrix(cfg_for(n0), "C4B")
        assert M.shape == (n0 - 4, n0 - 4)


def test_config_errors():
    lib = nat.load()
    bad = [cfg_for(8), cfg_for(17, ny=15), cfg_for(17, dt=0.0), cfg_for(17, nu=-1.0),
           cfg_for(17, precision=7), cfg_for(17, time_integrator=7), cfg_for(17, abi_version=99),
           cfg_for(17, mhd=1, eta=1e-3, time_integrator=nat.TI_IMEXRK3)]
    for c in bad:
        assert lib.pcf_field_size(C.byref(c), 0) < 0
        assert lib.pcf_last_error()
    c = cfg_for(17)
    assert lib.pcf_field_size(C.byref(c), nat.FIELD_ID["U0"]) == 13 * 16 * 16
    assert lib.pcf_field_size(C.byref(c), nat.FIELD_ID["G"]) == 15 * 16 * 16
    assert lib.pcf_field_size(C.byref(c), nat.FIELD_ID["SOL_OLD_V"]) == 15
    with pytest.raises(ValueError):
        nat.check(-1)
    # MHD: eta must be positive, mhd a flag; A leaves are Dirichlet fields
    for c in (cfg_for(17, mhd=1, eta=0.0), cfg_for(17, mhd=2, eta=1e-3)):
        assert lib.pcf_field_size(C.byref(c), 0) < 0
    c = cfg_for(17, mhd=1, eta=1 / 400)
    assert lib.pcf_field_size(C.byref(c), nat.FIELD_ID["AX"]) == 15 * 16 * 16
    assert lib.pcf_field_size(C.byref(c), nat.FIELD_ID["A_OLDER_Z"]) == 15 * 16 * 16


def test_python_boundary_errors_match_reference():
    """Invalid values raise ValueError, valid-but-unbuilt options
    NotImplementedError (examples/channelflow_kmm.py:140-187)."""
    from jaxfun_amd import PlaneCouetteFluctuationHIP as S

    with pytest.raises(ValueError):
        S(nonlinear_form="skew")
    with pytest.raises(ValueError):
        S(coefficient_path="fast")
    with pytest.raises(ValueError):
        S(time_integrator="RK4")
    with pytest.raises(ValueError):
        S(family="X")
    with pytest.raises(NotImplementedError):
        S(family="L")
    with pytest.raises(NotImplementedError):
        S(nonlinear_form="gradient")
    with pytest.raises(NotImplementedError):
        S()  # the reference's defaults: gradient form, transform path, IMEXRK222
    with pytest.raises(NotImplementedError):
        S(nonlinear_form="rotational", family="L")


def _HostSolver(N):
    from tests.helpers import HostSolver

    return HostSolver(N)


def test_seeded_state_properties():
    from jaxfun_amd.ic import perturbation_energy, seeded_state

    N = (17, 16, 16)
    hs = _HostSolver(N)
    s = seeded_state(hs, seed=0, energy=1e-3, K=4)
    o = PCFOracle(N=N, Re=1000.0, dt=5e-3)
    # energy: Epert = int |u|^2 = 2 V E
    V = 2 * hs.Ly * hs.Lz
    assert math.isclose(perturbation_energy(hs, s.u), 2 * V * 1e-3, rel_tol=1e-12)
    assert math.isclose(o.diagnostics(_to_oracle(s))["Epert"], 2 * V * 1e-3, rel_tol=1e-12)
    # divergence-free and Hermitian (real physical fields)
    assert o.diagnostics(_to_oracle(s))["divL2"] < 1e-14
    for a in list(s.u) + [s.g]:
        iy = (-np.arange(N[1])) % N[1]
        iz = (-np.arange(N[2])) % N[2]
        mirror = np.conj(a[:, iy][:, :, iz])
        assert np.max(np.abs(a - mirror)) < 1e-15
        assert np.all(a[:, N[1] // 2, :] == 0) and np.all(a[:, :, N[2] // 2] == 0)
    # deterministic in the seed
    s2 = seeded_state(hs, seed=0, energy=1e-3, K=4)
    assert all(np.array_equal(a, b) for a, b in zip(s.u, s2.u))


def _to_oracle(s):
    from tests.helpers import to_oracle

    return to_oracle(s)


def test_bench_refuses_multi_gpu():
    import subprocess
    import sys

    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr


@pytest.mark.parametrize("integ", ["SBDF3", "CNAB2", "IMEXRK222", "IMEXRK3"])
def test_device_history_follows_reference_ensure_rules(integ):
    """The history a state carries onto the device (_ensure_flow_history,
    examples/channelflow_kmm.py:612-635): SBDF3 restarts unless all four
    levels are present; CNAB2 without nonlinear_old starts at have_old = 0,
    previous_dt = 0 and keeps a present level; the other integrators keep the
    scalars and travel missing levels as zeros."""
    from dataclasses import replace

    from jaxfun_amd.ic import seeded_state
    from jaxfun_amd.pcf import KMMState
    from jaxfun_amd.pcf import PlaneCouetteFluctuationHIP as S

    hs = _HostSolver((17, 16, 16))
    hs.time_integrator = integ
    hs._device_history = S._device_history.__get__(hs)
    s = seeded_state(hs, seed=0, energy=1e-3, K=4)
    nl = tuple(np.ones_like(a) for a in s.nonlinear_old)
    partial = KMMState(u=s.u, g=s.g, nonlinear_old=nl, history_steps=2.0, have_old=1.0, previous_dt=0.5)
    d = hs._device_history(partial)
    assert all(getattr(d, k) is not None for k in ("nonlinear_old", "nonlinear_older", "solution_old",
                                                   "solution_older"))
    if integ == "SBDF3":  # incomplete: restart with zero levels and scalars
        assert d.history_steps == 0.0 and d.have_old == 0.0 and d.previous_dt == 0.0
        assert all(np.all(a == 0) for a in d.nonlinear_old)
    else:  # levels kept, missing ones zero, scalars kept
        assert d.history_steps == 2.0 and d.have_old == 1.0 and d.previous_dt == 0.5
        assert all(np.all(a == 1) for a in d.nonlinear_old)
        assert all(np.all(a == 0) for a in d.solution_old)
    bare = replace(partial, nonlinear_old=None)
    d = hs._device_history(bare)
    if integ == "CNAB2":
        assert d.have_old == 0.0 and d.previous_dt == 0.0 and d.history_steps == 2.0
    elif integ != "SBDF3":
        assert d.have_old == 1.0 and d.previous_dt == 0.5
    assert all(np.all(a == 0) for a in d.nonlinear_old)


def test_product_cfl_arithmetic_matches_oracle():
    """The host arithmetic of solver.cfl_scalars (spacings, rates, totals) on
    the oracle's own maxima equals the oracle's restatement of _cfl_scalars
    (production/health.py:120-153); the maxima themselves are a GPU test."""
    from jaxfun_amd.pcf import PlaneCouetteFluctuationHIP as S
    from oracle.mhd import RotatingShearMHDOracle

    o = RotatingShearMHDOracle(N=(17, 16, 16), Re=300.0, Rm=200.0, omega=1.3, shear_rate=0.7, dt=3e-3,
                               background_b=(0.0, 0.0, 0.05), magnetic_amplitude=0.0)
    so = o.initial_state()
    hs = _HostSolver((17, 16, 16))
    hs.dt, hs.nu, hs.eta, hs.omega, hs.shear_rate = o.dt, o.nu, o.eta, o.omega, o.shear_rate
    hs.field_maxima = lambda state=None: o.field_maxima(so)
    got = S.cfl_scalars.__get__(hs)()
    want = o.cfl(so)
    assert set(got) == set(want)
    for k, v in want.items():
        assert math.isclose(got[k], v, rel_tol=1e-14, abs_tol=1e-300), (k, got[k], v)


def test_mhd_sbdf3_history_keeps_complete_A_levels():
    """_ensure_mhd_history (examples/pcf_mhd_jax.py:217-231) under SBDF3: an
    incomplete flow history restarts the flow alone and a complete A history is
    kept as loaded; a missing A level restarts flow and A together."""
    from dataclasses import replace

    from jaxfun_amd.ic import seeded_state
    from jaxfun_amd.mhd import MHDState, PlaneCouetteMHDHIP as M
    from jaxfun_amd.pcf import KMMState
    from jaxfun_amd.pcf import PlaneCouetteFluctuationHIP as S

    hs = _HostSolver((17, 16, 16))
    hs.time_integrator = "SBDF3"
    for name in ("_device_history",):
        setattr(hs, name, getattr(S, name).__get__(hs))
    for name in ("_device_mhd_history", "_new_state", "_zero_A", "_a_rows"):
        setattr(hs, name, getattr(M, name).__get__(hs))
    s = seeded_state(hs, seed=1, energy=1e-3, K=4)
    flow = KMMState(u=s.u, g=s.g, nonlinear_old=tuple(np.ones_like(a) for a in s.nonlinear_old),
                    history_steps=2.0, have_old=1.0, previous_dt=0.5)
    A = tuple(np.full((hs.nd, 16, 16), 2.0 + 0j) for _ in range(3))
    lvl = {k: tuple(np.full((hs.nd, 16, 16), v + 0j) for _ in range(3))
           for k, v in (("nonlinear_A_old", 3.0), ("nonlinear_A_older", 4.0), ("A_old", 5.0), ("A_older", 6.0))}
    st = MHDState(flow=flow, A=A, **lvl)
    d = hs._device_mhd_history(st)
    assert d.flow.history_steps == 0.0 and d.flow.have_old == 0.0 and d.flow.previous_dt == 0.0
    assert all(np.all(a == 0) for a in d.flow.nonlinear_old)
    for k, v in (("nonlinear_A_old", 3.0), ("nonlinear_A_older", 4.0), ("A_old", 5.0), ("A_older", 6.0)):
        assert all(np.all(a == v) for a in getattr(d, k)), k
    d = hs._device_mhd_history(replace(st, A_older=None))
    assert all(np.all(a == 0) for k in lvl for a in getattr(d, k))
    assert d.flow.history_steps == 0.0


def test_gradient_rejects_unknown_fields():
    """pcf_get_gradient refuses field ids outside the state before it looks at
    the handle; the vector-potential fields are valid ids (their gradient
    exists for MHD handles: a hydrodynamic handle refuses them, which needs a
    GPU handle and is checked in the GPU tests)."""
    lib = nat.load()
    buf = (C.c_double * 8)()
    for f in (-1, 999, len(nat.FIELDS)):
        assert lib.pcf_get_gradient(None, f, buf) == -1  # PCF_E_INVALID
        assert b"no such field" in lib.pcf_last_error()
    for f in (nat.FIELD_ID["U0"], nat.FIELD_ID["AX"], nat.FIELD_ID["A_OLDER_Z"]):
        assert lib.pcf_get_gradient(None, f, buf) == -1  # PCF_E_INVALID
        assert b"null" in lib.pcf_last_error()


def test_taylor_couette_exports_and_pencil_match_oracle():
    """include/tc_dns.h: every declared entry point is exported; the library's
    long-double linear pencil (tc_linear_pencil, host only) equals the
    golden-pinned oracle pencil (oracle/tc.py) for axisymmetric and m = 2
    modes; bad configurations raise the reference's exception types."""
    from oracle.tc import CircularCouette, TCLinearOracle

    lib = nat.load()
    header = (ROOT / "include" / "tc_dns.h").read_text()
    declared = set(re.findall(r"^\s*(?:int32_t|int64_t|const char \*)\s*\*?\s*(tc_\w+)\s*\(", header, re.M))
    assert declared == set(nat.TC_EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name
    cfg = nat.TcConfig()
    cfg.abi_version = nat.TC_ABI_VERSION
    cfg.nr, cfg.nz = 24, 8
    cfg.r1, cfg.r2, cfg.omega1, cfg.omega2 = 1.0, 2.5, 1.0, -0.2
    cfg.nu, cfg.lz, cfg.dt, cfg.dealias = 0.02, 2.0, 1e-3, 1.5
    cfg.ntheta = 1
    n4 = 4 * 22
    dp = C.POINTER(C.c_double)
    for m, kz in ((0, 3.13), (2, 1.7)):
        L = np.empty((n4, n4), dtype=complex)
        M = np.empty((n4, n4), dtype=complex)
        nat.tc_check(lib.tc_linear_pencil(C.byref(cfg), m, kz, L.view(np.float64).ctypes.data_as(dp),
                                          M.view(np.float64).ctypes.data_as(dp)))
        Lo, Mo = TCLinearOracle(CircularCouette(1.0, 2.5, 1.0, -0.2), nu=0.02, N=24).assemble(m, kz)
        assert np.max(np.abs(L - Lo)) <= 1e-13 * np.max(np.abs(Lo))
        assert np.max(np.abs(M - Mo)) <= 1e-13 * np.max(np.abs(Mo))
    # the conducting MRI pencil (tc_mri_pencil) against the oracle's assembly
    from oracle.tc import TCMRILinearOracle

    cfg.b0, cfg.eta_mag = 0.1, 0.015
    n7 = 7 * 22
    for m, kz in ((0, 6.0), (1, 2.5)):
        L = np.empty((n7, n7), dtype=complex)
        M = np.empty((n7, n7), dtype=complex)
        nat.tc_check(lib.tc_mri_pencil(C.byref(cfg), m, kz, L.view(np.float64).ctypes.data_as(dp),
                                       M.view(np.float64).ctypes.data_as(dp)))
        Lo, Mo = TCMRILinearOracle(CircularCouette(1.0, 2.5, 1.0, -0.2), B0=0.1, nu=0.02, eta_mag=0.015,
                                   N=24).assemble(m, kz)
        assert np.max(np.abs(L - Lo)) <= 1e-13 * np.max(np.abs(Lo))
        assert np.max(np.abs(M - Mo)) <= 1e-13 * np.max(np.abs(Mo))
    L = np.empty((4 * 22, 4 * 22), dtype=complex)
    M = np.empty((4 * 22, 4 * 22), dtype=complex)
    cfg.r2 = 0.5
    with pytest.raises(ValueError, match="R1 < R2"):
        nat.tc_check(lib.tc_linear_pencil(C.byref(cfg), 0, 1.0, L.view(np.float64).ctypes.data_as(dp),
                                          M.view(np.float64).ctypes.data_as(dp)))
