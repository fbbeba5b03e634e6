"""Pins of the CPU FP64 oracle (test infrastructure) against the reference's
published results, library routines and closed forms.

Each test is chosen so that a plausible slip (dropped term, wrong sign or
index, transposed operand) fails it; none re-types the formula under test.
"""

import json
import math
from dataclasses import replace
from pathlib import Path

import numpy as np
import pytest
import scipy.fft as sfft
from numpy.polynomial import chebyshev as npcheb

from oracle.kmm import (
    PCFOracle,
    cheb_backward,
    cheb_derivative,
    cheb_forward,
    cheb_nodes,
    cheb_scalar_product,
    fejer_weights,
    fourier_backward,
    fourier_forward,
    galerkin_matrix,
    set_workers,
    stencil_clamped,
    stencil_dirichlet,
)

GOLDEN = json.loads((Path(__file__).parent / "golden" / "reference_sbdf3_published.json").read_text())


# ---------------------------------------------------------------- transforms
def test_chebyshev_transforms_match_vandermonde():
    rng = np.random.default_rng(0)
    for n in (5, 8, 17):
        c = rng.standard_normal(n)
        x = cheb_nodes(n)
        direct = npcheb.chebval(x, c)  # library evaluation of sum c_k T_k(x)
        assert np.allclose(cheb_backward(c), direct, rtol=0, atol=1e-13)
        assert np.allclose(cheb_forward(direct, n), c, rtol=0, atol=1e-13)
        # scalar product (u, T_k)_w by Gauss-Chebyshev quadrature = pi/n sum u(x_j) T_k(x_j)
        V = npcheb.chebvander(x, n - 1)
        assert np.allclose(cheb_scalar_product(direct, n), (np.pi / n) * V.T @ direct, atol=1e-13)
    assert np.all(np.diff(cheb_nodes(9)) > 0)  # ascending nodes


def test_chebyshev_derivative_matches_library():
    rng = np.random.default_rng(1)
    c = rng.standard_normal(12)
    for k in (1, 2, 4):
        ref = npcheb.chebder(c, k)
        ours = cheb_derivative(c.copy(), k)
        assert np.allclose(ours[: len(ref)], ref, atol=1e-11)
        assert np.all(ours[len(ref):] == 0)


def test_fourier_padding_and_truncation():
    rng = np.random.default_rng(2)
    N, M = 16, 24
    c = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    c[N // 2] = 0
    u = fourier_backward(c, 0, M)
    k = np.where(np.arange(N) < N // 2, np.arange(N), np.arange(N) - N)
    t = 2 * np.pi * np.arange(M) / M
    direct = np.exp(1j * np.outer(t, k)) @ c  # brute-force synthesis on the padded grid
    assert np.allclose(u, direct, atol=1e-12)
    assert np.allclose(fourier_forward(u, 0, N), c, atol=1e-13)


def test_fejer_weights_integrate_polynomials():
    n = 17
    x = cheb_nodes(n)
    w = fejer_weights(n)
    for p in range(0, n):
        exact = 0.0 if p % 2 else 2.0 / (p + 1)
        assert abs(np.sum(w * x**p) - exact) < 1e-13


# ---------------------------------------------------------------- Galerkin blocks
def test_dirichlet_blocks_closed_form():
    """(phi_i, phi_j) and (phi_i, phi_j'') for phi_i = T_i - T_{i+2} (weighted)."""
    N = 17
    S = stencil_dirichlet(N)
    C0 = galerkin_matrix(S, S, 0)
    C2 = galerkin_matrix(S, S, 2)
    n = N - 2
    for i in range(n):
        for j in range(n):
            c0 = (np.pi * (2 if i == 0 else 1) / 2 + np.pi / 2) if i == j else (-np.pi / 2 if abs(i - j) == 2 else 0.0)
            assert abs(C0[i, j] - c0) < 1e-12
            if j < i:
                c2 = 0.0
            elif j == i:
                c2 = -2 * np.pi * (i + 1) * (i + 2)
            else:
                c2 = -4 * np.pi * (i + 1) if (j - i) % 2 == 0 else 0.0
            assert abs(C2[i, j] - c2) < 1e-9 * max(1.0, abs(c2))


def test_clamped_basis_satisfies_boundary_conditions():
    S = stencil_clamped(17)
    for i in range(S.shape[0]):
        c = S[i]
        for x in (-1.0, 1.0):
            assert abs(npcheb.chebval(x, c)) < 1e-13
            assert abs(npcheb.chebval(x, npcheb.chebder(c))) < 1e-12


# ---------------------------------------------------------------- solver invariants
def test_laminar_state_is_a_fixed_point():
    o = PCFOracle(N=(17, 16, 16), Re=400.0, dt=1e-2)
    s = o.reset_history(o.state_from_physical(tuple(np.zeros((17, 16, 16), complex) for _ in range(3))))
    s = o.solve(s, 4)
    worst = max(float(np.max(np.abs(a))) for a in list(s.u) + [s.g])
    assert worst <= 2e-13


def test_rotational_term_does_no_pointwise_work():
    """u_tot . (u_tot x w_tot) = 0 at every padded node (reference
    tests/couette/test_pcf_fluctuations_jax.py:139-151, tolerance 2e-13)."""
    o = PCFOracle(N=(17, 16, 16), Re=400.0, dt=1e-3, perturbation_amplitude=0.05)
    s = o.initial_state()
    u0, u1, u2 = s.u
    up = (o.tp_backward(u0, o.SB), o.tp_backward(u1, o.SD), o.tp_backward(u2, o.SD))
    w0 = o.tp_backward(u2, o.SD, k1=1) - o.tp_backward(u1, o.SD, k2=1)
    w1 = o.tp_backward(u0, o.SB, k2=1) - o.tp_backward(u2, o.SD, kx=1)
    w2 = o.tp_backward(u1, o.SD, kx=1) - o.tp_backward(u0, o.SB, k1=1)
    a = (up[0], up[1] + o.xb, up[2])
    b = (w0, w1, w2 + 1.0)
    cross = (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])
    work = sum(np.real(ai * ci) for ai, ci in zip(a, cross))
    assert float(np.max(np.abs(work))) < 2e-13


def test_linear_decay_matches_closed_form():
    """Streamwise-invariant wall-normal vorticity decays at s = -nu (kz^2 + (pi/2)^2)
    (reference production/goldens/pcf_hydro_laminar_v1: eigenvalue_real =
    -0.0034674010999505545 at nu = 1e-3, kz = 1)."""
    o = PCFOracle(N=(17, 4, 4), Re=1000.0, dt=0.5)
    x, y, z = o.mesh()
    amp = 1e-9  # linear regime
    g_phys = amp * np.cos(np.pi * x / 2) * np.cos(z) * np.ones_like(y)
    zero = np.zeros((17, 4, 4))
    s = o.state_from_physical((zero.astype(complex), (g_phys * 0).astype(complex), zero.astype(complex)))
    g = o.mask(o.tp_forward(g_phys.astype(complex), o.SD, o.MD1))
    # u1 = i kz g / k^2 (reconstruction with f = 0) keeps the state consistent
    s = replace(s, g=g)
    s = replace(s, u=o.reconstruct(s.u[0], g, np.zeros(15), np.zeros(15)))
    s = o.reset_history(s)
    # late window: the next even eigenmode (rate -nu (1 + 9 pi^2 / 4)) excited by
    # the startup has decayed by e^-2 before it opens
    s = o.solve(s, 200)
    e1 = float(np.sum(np.abs(s.g) ** 2))
    s = o.solve(s, 100)
    e2 = float(np.sum(np.abs(s.g) ** 2))
    rate = 0.5 * math.log(e2 / e1) / (100 * o.dt)
    assert abs(rate - (-0.0034674010999505545)) < 2e-8


def _consistent(o, s):
    """KMM-consistent initial state: u1, u2 rebuilt from (u0, g) and the mean
    profiles (reference tests/couette/test_pcf_sbdf3_jax.py:61-70)."""
    means = (np.real(s.u[1][:, 0, 0]), np.real(s.u[2][:, 0, 0]))
    return replace(s, u=o.reconstruct(s.u[0], s.g, *means))


def test_sbdf3_is_third_order():
    """Self-convergence (reference tests/couette/test_pcf_sbdf3_jax.py:225-246):
    N=(9,4,4), Re=200, no padding, amplitude 0.01, T=0.04 in 8..64 steps;
    successive-difference orders in [2.7, 3.3]."""
    T = 0.04
    sols = []
    for steps in (8, 16, 32, 64):
        o = PCFOracle(N=(9, 4, 4), Re=200.0, dt=T / steps, padding_factor=(1.0, 1.0, 1.0),
                      perturbation_amplitude=0.01)
        s = o.solve(_consistent(o, o.initial_state()), steps)
        sols.append(np.concatenate([a.ravel() for a in list(s.u) + [s.g]]))
    errs = [float(np.linalg.norm(sols[i] - sols[i + 1])) for i in range(3)]
    orders = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(2.7 <= q <= 3.3 for q in orders), (errs, orders)


def test_sbdf3_agrees_with_small_dt_imexrk3():
    """SBDF3 at dt=T/20 against the IMEXRK3 startup integrator run alone at
    dt=T/160 (reference tests/couette/test_pcf_sbdf3_jax.py:249-265: < 2e-10)."""
    T = 0.01
    o = PCFOracle(N=(9, 4, 4), Re=200.0, dt=T / 20, padding_factor=(1.0, 1.0, 1.0),
                  perturbation_amplitude=0.01)
    a = o.solve(_consistent(o, o.initial_state()), 20)
    r = PCFOracle(N=(9, 4, 4), Re=200.0, dt=T / 160, padding_factor=(1.0, 1.0, 1.0),
                  perturbation_amplitude=0.01)
    b = _consistent(r, r.initial_state())
    for _ in range(160):
        b = r.step_imexrk3(b)
    va = np.concatenate([x.ravel() for x in list(a.u) + [a.g]])
    vb = np.concatenate([x.ravel() for x in list(b.u) + [b.g]])
    assert float(np.linalg.norm(va - vb)) < 2e-10


def _self_convergence_orders(integrator, T=0.04, counts=(8, 16, 32, 64)):
    sols = []
    for steps in counts:
        o = PCFOracle(N=(9, 4, 4), Re=200.0, dt=T / steps, padding_factor=(1.0, 1.0, 1.0),
                      perturbation_amplitude=0.01, time_integrator=integrator)
        s = o.solve(_consistent(o, o.initial_state()), steps)
        sols.append(np.concatenate([a.ravel() for a in list(s.u) + [s.g]]))
    errs = [float(np.linalg.norm(sols[i] - sols[i + 1])) for i in range(len(sols) - 1)]
    return errs, [math.log2(errs[i] / errs[i + 1]) for i in range(len(errs) - 1)]


def test_imexrk222_is_second_order():
    """ARS(2,2,2) self-convergence at the SBDF3 test's setting: orders in [1.8, 2.2]."""
    errs, orders = _self_convergence_orders("IMEXRK222")
    assert all(1.8 <= q <= 2.2 for q in orders), (errs, orders)


def test_imexrk3_alone_is_second_order():
    """The low-storage IMEXRK3 integrator run alone: third order in its explicit
    terms, but the implicit stages (Crank-Nicolson-like, gamma = (a + b) dt / 2)
    make the coupled scheme second order: orders in [1.8, 2.2]."""
    errs, orders = _self_convergence_orders("IMEXRK3")
    assert all(1.8 <= q <= 2.2 for q in orders), (errs, orders)


def test_cnab2_is_second_order():
    errs, orders = _self_convergence_orders("CNAB2")
    assert all(1.8 <= q <= 2.2 for q in orders), (errs, orders)


def test_cnab2_variable_dt_stays_second_order():
    """dt alternating between h and h/2 (ratio 1/2 and 2 in the variable-step
    AB2): the error against a fine SBDF3 run falls by about 4 per halving of h."""
    T = 0.03
    ref_o = PCFOracle(N=(9, 4, 4), Re=200.0, dt=T / 300, padding_factor=(1.0, 1.0, 1.0),
                      perturbation_amplitude=0.01)
    ref = ref_o.solve(_consistent(ref_o, ref_o.initial_state()), 300)
    vref = np.concatenate([x.ravel() for x in list(ref.u) + [ref.g]])
    errs = []
    for n in (10, 20, 40):  # n pairs (h, h/2) with 1.5 h n = T
        h = T / (1.5 * n)
        o = PCFOracle(N=(9, 4, 4), Re=200.0, dt=h, padding_factor=(1.0, 1.0, 1.0), perturbation_amplitude=0.01,
                      time_integrator="CNAB2")
        s = _consistent(o, o.initial_state())
        for _ in range(n):
            o.set_dt(h)
            s = o.step(s)
            o.set_dt(h / 2)
            s = o.step(s)
        v = np.concatenate([x.ravel() for x in list(s.u) + [s.g]])
        errs.append(float(np.linalg.norm(v - vref)))
    orders = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(1.8 <= q <= 2.3 for q in orders), (errs, orders)


def test_imexrk222_converges_to_sbdf3_solution():
    """Two independent integrators agree as dt -> 0: ARS(2,2,2) at dt = T/640
    against SBDF3 at dt = T/80 (errors ~ dt^2 and dt^3 respectively)."""
    T = 0.02
    runs = []
    for integ, steps in (("IMEXRK222", 640), ("SBDF3", 80)):
        o = PCFOracle(N=(9, 4, 4), Re=200.0, dt=T / steps, padding_factor=(1.0, 1.0, 1.0),
                      perturbation_amplitude=0.01, time_integrator=integ)
        s = o.solve(_consistent(o, o.initial_state()), steps)
        runs.append(np.concatenate([x.ravel() for x in list(s.u) + [s.g]]))
    assert float(np.linalg.norm(runs[0] - runs[1])) < 1e-9 * float(np.linalg.norm(runs[1]))


def test_imexrk222_history_semantics():
    """After an IMEXRK222 step only ``nonlinear_old``, ``have_old`` and
    ``previous_dt`` survive (channelflow_kmm.py:1194-1200); IMEXRK3 passes every
    history field through (``replace`` at :935)."""
    o = PCFOracle(N=(9, 8, 8), Re=200.0, dt=1e-3, perturbation_amplitude=0.01)
    s = o.solve(o.initial_state(), 3)  # SBDF3 history filled
    a = PCFOracle(N=(9, 8, 8), Re=200.0, dt=1e-3, perturbation_amplitude=0.01, time_integrator="IMEXRK222").step(s)
    assert a.nonlinear_old is s.nonlinear_old and a.nonlinear_older is None and a.solution_old is None
    assert a.history_steps == 0.0 and a.have_old == s.have_old
    b = PCFOracle(N=(9, 8, 8), Re=200.0, dt=1e-3, perturbation_amplitude=0.01, time_integrator="IMEXRK3").step(s)
    assert b.solution_older is s.solution_older and b.history_steps == s.history_steps


def test_nonlinear_rows_rounding_sensitivity():
    """Nu is a wall-normal derivative of the dealiased product: a 1e-16 relative
    perturbation of the state moves it by ~1e-12 relative already at 33^3.  This
    bounds how closely two FP64 implementations can agree on the stored
    nonlinear history rows (the state itself agrees to ~1e-15)."""
    set_workers(None)
    o = PCFOracle(N=(33, 32, 32), Re=400, dt=1e-3, perturbation_amplitude=0.02)
    s = o.solve(o.initial_state(), 2)
    base = o.nonlinear_rhs(o.convection(s))
    rng = np.random.default_rng(1)
    s2 = replace(s, u=tuple(a * (1 + 1e-16 * rng.standard_normal(a.shape)) for a in s.u))
    pert = o.nonlinear_rhs(o.convection(s2))
    d_nu = float(np.max(np.abs(pert[0] - base[0])) / np.max(np.abs(base[0])))
    d_ng = float(np.max(np.abs(pert[1] - base[1])) / np.max(np.abs(base[1])))
    assert d_nu > 1e-13  # ill-conditioned rows ...
    assert d_ng < 1e-14  # ... next to well-conditioned ones


@pytest.mark.slow
@pytest.mark.parametrize("case", GOLDEN["cases"], ids=lambda c: "x".join(map(str, c["N"])))
def test_oracle_reproduces_published_golden(case):
    """FP64 diagnostics and full-state checksum published by the reference
    (benchmark-results/*.json) after 325/50/20 SBDF3 steps."""
    set_workers(8)
    o = PCFOracle(N=tuple(case["N"]), Re=400.0, dt=1e-3, perturbation_amplitude=0.02)
    s = o.solve(o.initial_state(), case["steps"])
    d = o.diagnostics(s)
    g = case["diagnostics"]
    assert math.isclose(d["Epert"], g["Epert"], rel_tol=1e-10)
    assert math.isclose(d["Etot"], g["Etot"], rel_tol=1e-10)
    assert abs(d["wall_shear_top"] - g["wall_shear_top"]) < 1e-13
    cs = o.checksum(s)
    assert cs["element_count"] == case["checksum"]["element_count"]
    assert math.isclose(cs["l2_norm"], case["checksum"]["l2_norm"], rel_tol=1e-11)


# ---------------------------------------------------------------- CFL numbers
def test_cfl_scalars_known_answer():
    """_cfl_scalars (production/health.py:120-153) on hand-computed numbers:
    per-axis u dt / h and b dt / h, the largest directional sum, max(nu, eta)
    dt / min(h)^2, 2|Omega| dt, |S| dt, and the total over the explicit terms."""
    from oracle.kmm import cfl_scalars

    c = cfl_scalars((1.0, 2.0, 3.0), (0.5, 0.0, 1.0), (0.1, 0.2, 0.3), 0.01, 1e-3, 2e-3, 1.5, -2.0)
    want = {"cfl_advective_x": 0.1, "cfl_alfven_x": 0.05, "cfl_advective_y": 0.1, "cfl_alfven_y": 0.0,
            "cfl_advective_z": 0.1, "cfl_alfven_z": 0.01 / 0.3, "cfl_diffusive": 2e-3 * 0.01 / 0.01,
            "cfl_rotation": 0.03, "cfl_shear_source": 0.02, "cfl_total": 0.15}
    assert set(c) == set(want)
    for k, v in want.items():
        assert math.isclose(c[k], v, rel_tol=1e-14, abs_tol=1e-16), (k, c[k], v)
    # rotation-limited: 2 |Omega| dt above every directional sum
    assert math.isclose(cfl_scalars((1.0, 2.0, 3.0), (0.5, 0.0, 1.0), (0.1, 0.2, 0.3), 0.01, 1e-3, 0.0, 10.0,
                                    0.0)["cfl_total"], 0.2, rel_tol=1e-14)


def test_field_maxima_match_direct_series_sum():
    """The unpadded physical maxima against a direct evaluation of the
    Chebyshev x Fourier series at every grid point, base flow included."""
    o = PCFOracle(N=(9, 4, 6), Re=200.0, dt=1e-3, perturbation_amplitude=0.3)
    s = o.initial_state()
    n0, n1, n2 = o.N
    x = cheb_nodes(n0)
    y = np.arange(n1) * o.L1 / n1
    z = np.arange(n2) * o.L2 / n2
    ky = np.fft.fftfreq(n1, 1.0 / n1) * 2 * np.pi / o.L1
    kz = np.fft.fftfreq(n2, 1.0 / n2) * 2 * np.pi / o.L2
    Ey = np.exp(1j * np.outer(y, ky))  # [j, p]
    Ez = np.exp(1j * np.outer(z, kz))  # [k, q]
    V = npcheb.chebvander(x, n0 - 1)   # [i, b]
    want = []
    for c, S in zip(s.u, (o.SB, o.SD, o.SD)):
        phi = V @ S.T  # [i, a]: the Shen basis at the points
        u = np.einsum("ia,apq,jp,kq->ijk", phi, c, Ey, Ez)
        want.append(u)
    want[1] = want[1] + o.dUb * x[:, None, None]
    got, bmax = o.field_maxima(s)
    for g, w in zip(got, want):
        assert math.isclose(g, float(np.max(np.abs(w))), rel_tol=1e-12), (g, float(np.max(np.abs(w))))
    assert bmax == (0.0, 0.0, 0.0)
    dx = float(np.min(np.diff(np.sort(x))))
    assert o.grid_spacings() == (dx, o.L1 / n1, o.L2 / n2)


@pytest.mark.parametrize("N", [(17, 16, 16), (24, 16, 32)])
def test_transform_coefficient_path_equals_optimized(N):
    """coefficient_path="transform" (channelflow_kmm.py:785-803,831-847) and the
    optimized path (:810-829,855-870) form the same nonlinear rows and the same
    reconstruction: every quadrature on the unpadded grid is exact for these
    polynomial degrees, so the two agree to round-off on arbitrary fields."""
    o = PCFOracle(N=N, Re=500.0, dt=1e-3)
    rng = np.random.default_rng(11)

    def field(rows):
        c = rng.standard_normal((rows,) + N[1:]) + 1j * rng.standard_normal((rows,) + N[1:])
        return o.mask(c)

    H = tuple(field(o.ND) for _ in range(3))
    a, b = o.nonlinear_rhs_transform(H), o.nonlinear_rhs(H)
    for x, y in zip(a, b):
        assert np.max(np.abs(x - y)) < 1e-12 * np.max(np.abs(y))
    u0, g = field(o.NB), field(o.ND)
    v, w = rng.standard_normal(o.ND), rng.standard_normal(o.ND)
    for x, y in zip(o.reconstruct_transform(u0, g, v, w), o.reconstruct(u0, g, v, w)):
        assert np.max(np.abs(x - y)) < 1e-12 * np.max(np.abs(y))


def test_backward_velocity_matches_direct_series_sum():
    """The unpadded physical fluctuation velocity against a direct evaluation of
    the Chebyshev x Fourier series at every grid point."""
    o = PCFOracle(N=(9, 4, 6), Re=200.0, dt=1e-3, perturbation_amplitude=0.3)
    s = o.initial_state()
    n0, n1, n2 = o.N
    x = cheb_nodes(n0)
    ky = np.fft.fftfreq(n1, 1.0 / n1) * 2 * np.pi / o.L1
    kz = np.fft.fftfreq(n2, 1.0 / n2) * 2 * np.pi / o.L2
    Ey = np.exp(1j * np.outer(np.arange(n1) * o.L1 / n1, ky))
    Ez = np.exp(1j * np.outer(np.arange(n2) * o.L2 / n2, kz))
    V = npcheb.chebvander(x, n0 - 1)
    for got, c, S in zip(o.backward_velocity(s), s.u, (o.SB, o.SD, o.SD)):
        want = np.einsum("ia,apq,jp,kq->ijk", V @ S.T, c, Ey, Ez)
        assert np.max(np.abs(got - want)) < 1e-13 * np.max(np.abs(want))


def test_fluctuation_scalars_split_the_perturbation_energy():
    """streak_rms^2 + roll_rms^2 = Epert / V (the two integrals partition the
    energy the published diagnostics hold), and each measures its own
    components: a state with only u_y has no roll, one without u_y no streak."""
    o = PCFOracle(N=(17, 16, 16), Re=400.0, dt=1e-3, perturbation_amplitude=0.05)
    s = o.solve(o.initial_state(), 3)
    V = 2.0 * o.L1 * o.L2
    f = o.fluctuation_scalars(s)
    d = o.diagnostics(s)
    assert f["roll_rms"] > 0.0 and f["streak_rms"] > 0.0
    assert math.isclose((f["streak_rms"] ** 2 + f["roll_rms"] ** 2) * V, d["Epert"], rel_tol=1e-13)
    assert {k: f[k] for k in ("kinetic_energy", "total_kinetic_energy", "divergence_l2", "mean_shear")} == {
        "kinetic_energy": d["Epert"], "total_kinetic_energy": d["Etot"], "divergence_l2": d["divL2"],
        "mean_shear": d["mean_shear"]}
    assert (f["wall_shear_lower"], f["wall_shear_upper"]) == (d["wall_shear_bot"], d["wall_shear_top"])
    assert (f["wall_velocity_lower"], f["wall_velocity_upper"]) == (d["u_bot"], d["u_top"])
    u0, u1, u2 = s.u
    only_streak = replace(s, u=(0 * u0, u1, 0 * u2))
    only_roll = replace(s, u=(u0, 0 * u1, u2))
    fs, fr = o.fluctuation_scalars(only_streak), o.fluctuation_scalars(only_roll)
    assert fs["roll_rms"] == 0.0 and math.isclose(fs["streak_rms"], f["streak_rms"], rel_tol=1e-14)
    assert fr["streak_rms"] == 0.0 and math.isclose(fr["roll_rms"], f["roll_rms"], rel_tol=1e-14)
