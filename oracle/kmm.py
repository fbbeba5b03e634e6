"""CPU FP64 oracle: a numpy restatement of the reference plane-Couette KMM step.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped package imports this
module; it is the checker that the HIP path is compared against (tests,
``__graft_entry__.smoke`` and the ``cpu_baseline`` leg of ``bench.py``).

It restates, in the reference's own full-complex spectral layout
(axis 0 = wall-normal Chebyshev/Shen, axis 1 = streamwise Fourier,
axis 2 = spanwise Fourier), the path

* ``examples/channelflow_kmm.py:111-870,937-1064``   (KMM solver, SBDF3 step)
* ``examples/pcf_fluctuations_jax.py:35-161``        (plane-Couette specialisation)
* ``src/jaxfun/integrators/sbdf3.py:11-62``          (SBDF3 weights)
* ``src/jaxfun/integrators/imex_rk.py:96-105``       (IMEXRK3 startup table)
* ``src/jaxfun/galerkin/Chebyshev.py:159-355``       (nodes, DCT transforms, derivative recurrence)
* ``src/jaxfun/galerkin/Fourier.py:17-224``          (padded FFT, truncation, wavenumbers)
* ``src/jaxfun/galerkin/composite.py:283-371,797-830`` (Shen stencils, mass projections)
* ``src/jaxfun/galerkin/tensorproductspace.py:80-99,253-293,459-754`` (separable transforms)
* ``src/jaxfun/la/tpmatrix.py:1034-1150`` / ``diamatrix.py:1937-1973`` (no-pivot per-mode LU)
* ``src/jaxfun/galerkin/inner.py:222-262`` + ``Chebyshev.py:183-209`` (diagnostic integrals)

Only ``nonlinear_form="rotational"`` with ``coefficient_path="optimized"`` (the
configuration every published SBDF3 number uses) is restated, for the
Chebyshev family.

Pins (``tests/test_oracle.py``): the reference's published FP64
diagnostics and full-state checksums after 325/50/20 SBDF3 steps at
65x64x64/97x96x96/128^3 (``benchmark-results/*.json``), closed-form linear
decay rates, library DCT/FFT identities, brute-force quadrature of every
Galerkin matrix, the zero-self-work identity of the rotational term and
third-order temporal convergence.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np
import scipy.fft as sfft

# --------------------------------------------------------------------------
# 1-D bases
# --------------------------------------------------------------------------

_WORKERS = None


def set_workers(n: int | None) -> None:
    """Thread count handed to scipy.fft (None = scipy default of 1)."""
    global _WORKERS
    _WORKERS = n


def cheb_nodes(n: int) -> np.ndarray:
    """Gauss-Chebyshev nodes in ascending order (``Chebyshev.py:159-181``)."""
    return np.cos(np.pi + (2 * np.arange(n) + 1) * np.pi / (2 * n))


def _sign(n: int, ndim: int) -> np.ndarray:
    s = (-1.0) ** np.arange(n)
    return s.reshape((n,) + (1,) * (ndim - 1))


def cheb_backward(c: np.ndarray, n: int | None = None) -> np.ndarray:
    """Series values at the Gauss nodes along axis 0 (``Chebyshev.py:263-280``)."""
    n = c.shape[0] if n is None else n
    if n > c.shape[0]:
        pad = [(0, n - c.shape[0])] + [(0, 0)] * (c.ndim - 1)
        c = np.pad(c, pad)
    uh = c * _sign(n, c.ndim)
    return 0.5 * uh[0] + n * sfft.idct(uh, n=n, axis=0, workers=_WORKERS)


def cheb_forward(u: np.ndarray, N: int) -> np.ndarray:
    """Discrete Chebyshev transform along axis 0 (``Chebyshev.py:282-301``)."""
    n = u.shape[0]
    uh = sfft.dct(u, n=n, axis=0, workers=_WORKERS)
    uh[0] = uh[0] / 2
    uh = uh * _sign(n, u.ndim) / n
    return uh[:N]


def cheb_scalar_product(u: np.ndarray, N: int) -> np.ndarray:
    """Weighted scalar products (u, T_k) along axis 0 (``Chebyshev.py:303-318``)."""
    n = u.shape[0]
    uh = sfft.dct(u, n=n, axis=0, workers=_WORKERS)
    uh = uh * np.pi * _sign(n, u.ndim) / n / 2
    return uh[:N]


def cheb_derivative(c: np.ndarray, k: int = 1) -> np.ndarray:
    """Chebyshev derivative coefficients along axis 0 (``Chebyshev.py:320-355``)."""
    for _ in range(k):
        N = c.shape[0] - 1
        out = np.zeros_like(c)
        if N == 0:
            c = out
            continue
        # b_{N} = 0, b_{N-1} = 2 N c_N, b_n = b_{n+2} + 2 (n+1) c_{n+1}, b_0 halved
        out[N - 1] = 2 * N * c[N]
        for m in range(N - 2, -1, -1):
            out[m] = 2 * (m + 1) * c[m + 1] + (out[m + 2] if m + 2 <= N else 0)
        out[0] = out[0] / 2
        c = out
    return c


def cheb_norm_squared(n: int) -> np.ndarray:
    """h_k = pi (k=0), pi/2 otherwise (``Chebyshev.py:437-446``)."""
    h = np.full(n, np.pi / 2)
    h[0] = np.pi
    return h


def fejer_weights(n: int) -> np.ndarray:
    """Fejer-I weights for the unweighted measure on [-1,1] (``Chebyshev.py:183-209``)."""
    theta = np.pi + (2 * np.arange(n) + 1) * np.pi / (2 * n)
    modes = np.arange(1, n // 2 + 1)
    corr = np.sum(
        2.0 * np.cos(2.0 * theta[:, None] * modes[None, :]) / (4.0 * modes[None, :] ** 2 - 1.0),
        axis=1,
    )
    return (2.0 / n) * (1.0 - corr)


def stencil_dirichlet(N: int) -> np.ndarray:
    """phi_i = T_i - T_{i+2} (``composite.py:816-817``); shape (N-2, N)."""
    S = np.zeros((N - 2, N))
    i = np.arange(N - 2)
    S[i, i] = 1.0
    S[i, i + 2] = -1.0
    return S


def stencil_clamped(N: int) -> np.ndarray:
    """Clamped (D+N) Chebyshev stencil (``composite.py:818-826``); shape (N-4, N)."""
    S = np.zeros((N - 4, N))
    n = np.arange(N - 4, dtype=float)
    i = np.arange(N - 4)
    S[i, i] = 1.0
    S[i, i + 2] = 2 * (-n - 2) / (n + 3)
    S[i, i + 4] = (n + 1) / (n + 3)
    return S


def fourier_wavenumbers(N: int, eliminate_highest: bool = False) -> np.ndarray:
    """Integer FFT-ordered wavenumbers (``Fourier.py:17-22``)."""
    idx = np.arange(N)
    k = np.where(idx < (N + 1) // 2, idx, idx - N).astype(float)
    if eliminate_highest and N % 2 == 0:
        k[N // 2] = 0
    return k


def fourier_backward(c: np.ndarray, axis: int, n: int) -> np.ndarray:
    """Zero-insert to n then ifft(norm='forward') (``Fourier.py:128-151``)."""
    c = np.moveaxis(c, axis, 0)
    N = c.shape[0]
    if n > N:
        z = np.zeros((n - N,) + c.shape[1:], dtype=complex)
        c = np.concatenate((c[: N // 2], z, c[N // 2 :]), axis=0)
    out = sfft.ifft(c, axis=0, norm="forward", workers=_WORKERS)
    return np.moveaxis(out, 0, axis)


def fourier_forward(u: np.ndarray, axis: int, N: int) -> np.ndarray:
    """fft(norm='forward') with truncation to wavenumbers() (``Fourier.py:170-185``)."""
    u = np.moveaxis(u, axis, 0)
    out = sfft.fft(u, axis=0, norm="forward", workers=_WORKERS)
    if u.shape[0] > N:
        idx = fourier_wavenumbers(N).astype(int)
        out = out[idx]
    return np.moveaxis(out, 0, axis)


# --------------------------------------------------------------------------
# Galerkin matrices by brute-force quadrature (exact for these degrees)
# --------------------------------------------------------------------------


def _cheb_vandermonde(nq: int, nmodes: int) -> tuple[np.ndarray, np.ndarray]:
    theta = (2 * np.arange(nq) + 1) * np.pi / (2 * nq)
    x = np.cos(theta)
    V = np.cos(np.outer(theta, np.arange(nmodes)))  # T_n(x_q)
    return x, V


def galerkin_matrix(S_test: np.ndarray, S_trial: np.ndarray, deriv: int) -> np.ndarray:
    """(psi_i, d^m psi_j)_w with the Chebyshev weight, by Gauss quadrature.

    Q = 2N+8 nodes integrate every product here exactly.  Independent of the
    closed forms used by the HIP library's setup, so comparing the two pins both.
    """
    N = S_trial.shape[1]
    nq = 2 * N + 8
    _, V = _cheb_vandermonde(nq, N)
    w = np.pi / nq
    trial_orth = S_trial.T.copy()  # (N, ntrial) orthogonal coefficients of each trial fn
    trial_orth = cheb_derivative(trial_orth, deriv) if deriv else trial_orth
    vals_trial = V @ trial_orth  # (nq, ntrial)
    vals_test = V @ S_test.T  # (nq, ntest)
    M = w * vals_test.T @ vals_trial
    # Entries that vanish by orthogonality come out as ~1e-17 quadrature noise;
    # the reference assembles them as exact zeros, so restore those zeros.
    M[np.abs(M) < 1e-11 * np.max(np.abs(M))] = 0.0
    return M


# --------------------------------------------------------------------------
# Solver
# --------------------------------------------------------------------------

IMEXRK3_A = (8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0)
IMEXRK3_B = (0.0, -17.0 / 60.0, -5.0 / 12.0)
SBDF3_GAMMA = 6.0 / 11.0  # sbdf3.py:11-12 IMPLICIT_SCALE
# ARS(2,2,2) tableau (src/jaxfun/integrators/imex_rk.py:175-181): a[1][1] = a[2][2] = g,
# a[2][1] = 1 - g, b[1][0] = g, b[2][0] = d, b[2][1] = 1 - d
ARS222_G = (2.0 - 2.0**0.5) / 2.0
ARS222_D = 1.0 - 1.0 / (2.0 * ARS222_G)
INTEGRATORS = ("SBDF3", "IMEXRK3", "IMEXRK222", "CNAB2")


@dataclass
class OracleState:
    """Mirror of ``KMMState`` (``channelflow_kmm.py:50-101``), full-complex layout."""

    u: tuple
    g: np.ndarray
    nonlinear_old: tuple | None = None
    nonlinear_older: tuple | None = None
    solution_old: tuple | None = None
    solution_older: tuple | None = None
    history_steps: float = 0.0
    have_old: float = 0.0
    previous_dt: float = 0.0

    def leaves(self) -> list:
        out = list(self.u) + [self.g]
        for group in (self.nonlinear_old, self.nonlinear_older, self.solution_old, self.solution_older):
            if group is not None:  # None is an empty subtree, as in jax
                out.extend(group)
        out.extend([np.asarray(self.history_steps), np.asarray(self.have_old), np.asarray(self.previous_dt)])
        return out


def _lower_bandwidth(A: np.ndarray) -> int:
    n = A.shape[-1]
    nz = np.any(A != 0, axis=0)
    p = 0
    for off in range(1, n):
        if np.any(np.diagonal(nz, offset=-off)):
            p = off
    return p


def _nopivot_lu(A: np.ndarray) -> tuple[np.ndarray, int]:
    """Batched banded LU without pivoting (``diamatrix.py:1937-1973``).

    Returns the packed factors (L strictly below the diagonal, U on and above)
    and the lower bandwidth p; only the p rows under each pivot are eliminated,
    as in the reference's DIA kernel.
    """
    A = A.copy()
    n = A.shape[-1]
    p = _lower_bandwidth(A)
    for i in range(n - 1):
        hi = min(n, i + p + 1)
        A[:, i + 1 : hi, i] /= A[:, i, i][:, None]
        A[:, i + 1 : hi, i + 1 :] -= A[:, i + 1 : hi, i][:, :, None] * A[:, i, i + 1 :][:, None, :]
    return A, p


def _lu_solve(fac: tuple[np.ndarray, int], b: np.ndarray) -> np.ndarray:
    """Solve with packed no-pivot LU; b (B, n) or (B, n, r) complex.

    Forward substitution over the p-row band of L, then back substitution with
    the dense upper factor, row by row as the reference's compact recurrence
    (``tpmatrix.py:686-799``).  The factors are real, so real and imaginary
    parts ride as separate right-hand-side columns.
    """
    LU, p = fac
    n = LU.shape[-1]
    squeeze = b.ndim == 2
    bb = b[..., None] if squeeze else b
    y = np.concatenate((bb.real, bb.imag), axis=-1).astype(float)
    for i in range(1, n):
        lo = max(0, i - p)
        y[:, i] -= np.matmul(LU[:, i : i + 1, lo:i], y[:, lo:i])[:, 0]
    for i in range(n - 1, -1, -1):
        if i < n - 1:
            y[:, i] -= np.matmul(LU[:, i : i + 1, i + 1 :], y[:, i + 1 :])[:, 0]
        y[:, i] /= LU[:, i, i][:, None]
    r = bb.shape[-1]
    out = y[..., :r] + 1j * y[..., r:]
    return out[..., 0] if squeeze else out


def cfl_scalars(u_max, b_max, spacings, dt, nu, eta, rotation_rate, shear_rate) -> dict:
    """``_cfl_scalars`` (``production/health.py:120-153``): advective and Alfven
    CFL per axis, their largest directional sum, the diffusive number
    max(nu, eta) dt / min(h)^2, the rotation and shear-source rates, and
    ``cfl_total`` over the explicit terms."""
    out = {}
    worst = 0.0
    for axis, label in enumerate("xyz"):
        adv = u_max[axis] * dt / spacings[axis]
        alf = b_max[axis] * dt / spacings[axis]
        out["cfl_advective_" + label] = adv
        out["cfl_alfven_" + label] = alf
        worst = max(worst, adv + alf)
    out["cfl_diffusive"] = max(nu, eta) * dt / min(spacings) ** 2
    out["cfl_rotation"] = 2.0 * abs(rotation_rate) * dt
    out["cfl_shear_source"] = abs(shear_rate) * dt
    out["cfl_total"] = max(worst, out["cfl_rotation"], out["cfl_shear_source"])
    return out


def variable_ab2_extrapolate(current, previous, have_old, dt, previous_dt):
    """``variable_ab2_extrapolate`` (``src/jaxfun/integrators/cnab2.py:211-239``):
    (1 + r/2) current - (r/2) previous with r = dt / previous_dt when a previous
    level exists (``have_old != 0``), else current."""
    if previous is None or float(have_old) == 0.0:
        return tuple(current)
    r = float(dt) / float(previous_dt)
    return tuple((1.0 + 0.5 * r) * c - 0.5 * r * q for c, q in zip(current, previous))


class PCFOracle:
    """Restatement of ``PlaneCouetteFluctuationJax`` + ``KMM`` for SBDF3/rotational.

    Parameters follow ``examples/pcf_fluctuations_jax.py:35-75``.
    """

    def __init__(
        self,
        N=(17, 16, 16),
        domain=((-1.0, 1.0), (0.0, 4.0 * math.pi), (0.0, 2.0 * math.pi)),
        Re: float = 600.0,
        U_wall: float = 1.0,
        dt: float = 0.01,
        padding_factor=(1.0, 1.5, 1.5),
        perturbation_amplitude: float = 0.05,
        time_integrator: str = "SBDF3",
    ) -> None:
        if time_integrator not in INTEGRATORS:
            raise ValueError(f"time_integrator must be one of {INTEGRATORS}")
        self.time_integrator = time_integrator
        self.N = tuple(int(n) for n in N)
        self.domain = domain
        self.Re = float(Re)
        self.U_wall = float(U_wall)
        self.nu = self.U_wall / self.Re
        # slope of the laminar base flow U_b = dUb * x (pcf_fluctuations_jax.py:73-75;
        # -shear_rate in examples/pcf_mhd_mri_shearpy_jax.py:89-92)
        self.dUb = self.U_wall
        self.dt = float(dt)
        self.perturbation_amplitude = float(perturbation_amplitude)
        N0, N1, N2 = self.N
        self.N0 = N0
        self.NB = N0 - 4
        self.ND = N0 - 2
        # padded counts (orthogonal.py:348-359): floor(factor * N)
        self.pad = tuple(max(n, int(f * n)) for n, f in zip(self.N, padding_factor))
        # domain factors (orthogonal.py:471-483)
        L1 = domain[1][1] - domain[1][0]
        L2 = domain[2][1] - domain[2][0]
        self.df1 = 2 * math.pi / L1 if abs(L1 - 2 * math.pi) > 1e-12 else 1.0
        self.df2 = 2 * math.pi / L2 if abs(L2 - 2 * math.pi) > 1e-12 else 1.0
        self.L1, self.L2 = L1, L2
        self.k1 = fourier_wavenumbers(N1)
        self.k2 = fourier_wavenumbers(N2)
        self.K1 = (self.k1 * self.df1)[None, :, None]
        self.K2 = (self.k2 * self.df2)[None, None, :]
        self.K1e = (fourier_wavenumbers(N1, True) * self.df1)[None, :, None]
        self.K2e = (fourier_wavenumbers(N2, True) * self.df2)[None, None, :]
        k2sum = self.K1**2 + self.K2**2
        denom = np.where(k2sum == 0, 1, k2sum)
        self.KoK2 = (self.K1 / denom, self.K2 / denom)
        mask = np.zeros((1, N1, N2), dtype=bool)
        mask |= (np.abs(self.k1) == N1 // 2)[None, :, None]
        mask |= (np.abs(self.k2) == N2 // 2)[None, None, :]
        self.nyq = mask
        # Fourier mass weights 2 pi / df per axis (Fourier.py:227-229 / orthogonal.py:301-307)
        self.W = (2 * math.pi / self.df1) * (2 * math.pi / self.df2)

        self.SB = stencil_clamped(N0)
        self.SD = stencil_dirichlet(N0)
        self.h = cheb_norm_squared(N0)
        self.MB1 = self.SB @ np.diag(self.h) @ self.SB.T
        self.MD1 = self.SD @ np.diag(self.h) @ self.SD.T
        # Galerkin blocks (quadrature); C{m} = (psi_i, psi_j^{(m)})
        self.B_C0 = galerkin_matrix(self.SB, self.SB, 0)
        self.B_C2 = galerkin_matrix(self.SB, self.SB, 2)
        self.B_C4 = galerkin_matrix(self.SB, self.SB, 4)
        self.D_C0 = galerkin_matrix(self.SD, self.SD, 0)
        self.D_C2 = galerkin_matrix(self.SD, self.SD, 2)

        self.x = cheb_nodes(N0)
        self.xb = self.x[:, None, None]
        # flatten k2 per mode
        self.k2flat = np.broadcast_to(k2sum, (1, N1, N2)).reshape(-1)
        self._factors = {}
        self._build_factors()

    # ---------------------------------------------------------------- operators
    def _mass_u(self, k2):
        return self.W * (self.B_C2[None] - k2[:, None, None] * self.B_C0[None])

    def _lin_u(self, k2):
        return self.W * self.nu * (
            self.B_C4[None] - 2 * k2[:, None, None] * self.B_C2[None] + (k2**2)[:, None, None] * self.B_C0[None]
        )

    def _mass_g(self, k2):
        return self.W * np.broadcast_to(self.D_C0, (k2.size,) + self.D_C0.shape)

    def _lin_g(self, k2):
        return self.W * self.nu * (self.D_C2[None] - k2[:, None, None] * self.D_C0[None])

    def _build_factors(self):
        k2 = self.k2flat
        self.Mu = self._mass_u(k2)
        self.Lu = self._lin_u(k2)
        self.Mg = self._mass_g(k2)
        self.Lg = self._lin_g(k2)
        self.M00 = self.D_C0
        self.L00 = self.nu * self.D_C2

        def fac(gamma):
            key = round(gamma, 17)
            if key not in self._factors:
                Su = _nopivot_lu(self.Mu - gamma * self.Lu)
                Sg = _nopivot_lu(self.Mg - gamma * self.Lg)
                S00 = _nopivot_lu((self.M00 - gamma * self.L00)[None])
                self._factors[key] = (Su, Sg, S00)
            return self._factors[key]

        self._fac = fac
        self.steady = fac(self.dt * SBDF3_GAMMA)
        self.startup = tuple(fac((IMEXRK3_A[rk] + IMEXRK3_B[rk]) * self.dt / 2.0) for rk in range(3))
        self.ars = fac(self.dt * ARS222_G) if self.time_integrator == "IMEXRK222" else None

    # ---------------------------------------------------------------- helpers
    def mask(self, c):
        return np.where(self.nyq, 0, c)

    def _modes(self, c):
        n = c.shape[0]
        return np.moveaxis(c, 0, -1).reshape(-1, n)

    def _unmodes(self, m):
        n = m.shape[-1]
        return np.moveaxis(m.reshape(self.N[1], self.N[2], n), -1, 0)

    def _permode(self, Mats, c):
        m = self._modes(c)
        re = np.matmul(Mats, m.real[:, :, None])[:, :, 0]
        im = np.matmul(Mats, m.imag[:, :, None])[:, :, 0]
        return self._unmodes(re + 1j * im)

    @staticmethod
    def _apply0(A, c):
        """A (p, q) applied along axis 0 of c (q, ...)."""
        return np.tensordot(A, c, axes=(1, 0))

    def to_orth(self, c, S):
        return self._apply0(S.T, c)

    def tp_backward(self, c, S, padded=True, kx=0, k1=0, k2=0):
        """TensorProductSpace.backward_primitive, axis order (1, 2, 0)."""
        n1, n2 = (self.pad[1], self.pad[2]) if padded else (self.N[1], self.N[2])
        z = c
        if k1:
            K = self.K1e if k1 % 2 else self.K1
            z = (1j * K) ** k1 * z
        z = fourier_backward(z, 1, n1)
        if k2:
            K = self.K2e if k2 % 2 else self.K2
            z = (1j * K) ** k2 * z
        z = fourier_backward(z, 2, n2)
        a = self.to_orth(z, S) if S is not None else z
        if kx:
            a = cheb_derivative(a, kx)
        return cheb_backward(a, self.N0)

    def tp_forward_D(self, u):
        """TDp.forward with truncation, axis order (0, 2, 1)."""
        L = cheb_scalar_product(u, self.N0)
        L = self._apply0(self.SD, L)
        shp = L.shape
        sol = np.linalg.solve(self.MD1, L.reshape(shp[0], -1)).reshape(shp)
        z = fourier_forward(sol, 2, self.N[2])
        z = fourier_forward(z, 1, self.N[1])
        return z

    def tp_forward(self, u, S, M1):
        """TB/TD.forward on the unpadded grid (axis order 0, 2, 1)."""
        L = cheb_scalar_product(u, self.N0)
        L = self._apply0(S, L)
        shp = L.shape
        sol = np.linalg.solve(M1, L.reshape(shp[0], -1)).reshape(shp)
        z = fourier_forward(sol, 2, self.N[2])
        z = fourier_forward(z, 1, self.N[1])
        return z

    # ---------------------------------------------------------------- state
    def mesh(self):
        y = np.linspace(self.domain[1][0], self.domain[1][1], self.N[1], endpoint=False)
        z = np.linspace(self.domain[2][0], self.domain[2][1], self.N[2], endpoint=False)
        return self.x[:, None, None], y[None, :, None], z[None, None, :]

    def state_from_physical(self, up) -> OracleState:
        """``channelflow_kmm.py:680-686``."""
        u0 = self.mask(self.tp_forward(up[0], self.SB, self.MB1))
        u1 = self.mask(self.tp_forward(up[1], self.SD, self.MD1))
        u2 = self.mask(self.tp_forward(up[2], self.SD, self.MD1))
        g = self.mask(1j * self.K1 * u2 - 1j * self.K2 * u1)
        return self.reset_history(OracleState(u=(u0, u1, u2), g=g))

    def initial_state(self) -> OracleState:
        """Analytic fluctuation IC (``pcf_fluctuations_jax.py:77-97``)."""
        x, y, z = self.mesh()
        amp = self.perturbation_amplitude
        wall = 1.0 - x**2
        Ly, Lz = self.L1, self.L2
        u0 = amp * wall * np.sin(2 * np.pi * y / Ly) * np.cos(2 * np.pi * z / Lz)
        u1 = amp * wall * np.cos(2 * np.pi * y / Ly) * np.sin(2 * np.pi * z / Lz)
        u2 = amp * wall * np.sin(4 * np.pi * y / Ly) * np.cos(4 * np.pi * z / Lz)
        shape = (self.N0, self.N[1], self.N[2])
        return self.state_from_physical(tuple(np.broadcast_to(v, shape).astype(complex) for v in (u0, u1, u2)))

    def primary(self, s: OracleState):
        return (s.u[0], s.g, np.real(s.u[1][:, 0, 0]).copy(), np.real(s.u[2][:, 0, 0]).copy())

    def reset_history(self, s: OracleState) -> OracleState:
        p = self.primary(s)
        z = tuple(np.zeros_like(a) for a in p)
        return replace(s, nonlinear_old=z, nonlinear_older=z, solution_old=z, solution_older=z,
                       history_steps=0.0, have_old=0.0, previous_dt=0.0)

    # ---------------------------------------------------------------- physics
    def convection_physical(self, s: OracleState):
        """Physical -(u_tot x w_tot) and u on the padded grid (``channelflow_kmm.py:754-762``)."""
        u0, u1, u2 = s.u
        SB, SD = self.SB, self.SD
        up = (self.tp_backward(u0, SB), self.tp_backward(u1, SD), self.tp_backward(u2, SD))
        w0 = self.tp_backward(u2, SD, k1=1) - self.tp_backward(u1, SD, k2=1)
        w1 = self.tp_backward(u0, SB, k2=1) - self.tp_backward(u2, SD, kx=1)
        w2 = self.tp_backward(u1, SD, kx=1) - self.tp_backward(u0, SB, k1=1)
        a = (up[0], up[1] + self.dUb * self.xb, up[2])
        b = (w0, w1, w2 + self.dUb)
        cross = (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])
        return tuple(-c for c in cross), up

    def grid_spacings(self):
        """(dx, Ly/ny, Lz/nz) of ``curl_health_scalars`` (``production/health.py:165-172``):
        dx the smallest gap between Chebyshev points."""
        x1d = np.sort(np.unique(self.x))
        return (float(np.diff(x1d).min()), self.L1 / self.N[1], self.L2 / self.N[2])

    def field_maxima(self, s: OracleState):
        """max |.| of the total velocity (``total_velocity_physical``,
        ``pcf_fluctuations_jax.py:114-116``) on the unpadded grid per component,
        and of the magnetic field (none here)."""
        u0, u1, u2 = s.u
        up = (self.tp_backward(u0, self.SB, False), self.tp_backward(u1, self.SD, False),
              self.tp_backward(u2, self.SD, False))
        ut = (up[0], up[1] + self.dUb * self.xb, up[2])
        return tuple(float(np.max(np.abs(c))) for c in ut), (0.0, 0.0, 0.0)

    def cfl(self, s: OracleState, eta: float = 0.0, rotation_rate: float = 0.0, shear_rate: float = 0.0) -> dict:
        u_max, b_max = self.field_maxima(s)
        return cfl_scalars(u_max, b_max, self.grid_spacings(), self.dt, self.nu, eta, rotation_rate, shear_rate)

    def convection(self, s: OracleState):
        """Rotational convection H = -(u_tot x w_tot) (``channelflow_kmm.py:772-782``)."""
        n, _ = self.convection_physical(s)
        return tuple(self.mask(self.tp_forward_D(c)) for c in n)

    def _dcoef(self, c, kx, k1, k2):
        """differentiate_orthogonal_coeffs (``tensorproductspace.py:638-682``)."""
        z = c
        if k1:
            K = self.K1e if k1 % 2 else self.K1
            z = (1j * K) ** k1 * z
        if k2:
            K = self.K2e if k2 % 2 else self.K2
            z = (1j * K) ** k2 * z
        if kx:
            z = cheb_derivative(z, kx)
        return z

    def _sp_from_orth(self, c, S):
        """scalar_product_from_orthogonal (``tensorproductspace.py:730-754``)."""
        w = c * self.h[:, None, None]
        return self._apply0(S, w) * self.W

    def nonlinear_rhs(self, H):
        """``_nonlinear_rhs_coefficient`` (``channelflow_kmm.py:810-829``)."""
        Ho = tuple(self.to_orth(h, self.SD) for h in H)
        Hu = (self._dcoef(Ho[1], 1, 1, 0) + self._dcoef(Ho[2], 1, 0, 1)
              - self._dcoef(Ho[0], 0, 2, 0) - self._dcoef(Ho[0], 0, 0, 2))
        Hg = self._dcoef(Ho[1], 0, 0, 1) - self._dcoef(Ho[2], 0, 1, 0)
        Nu = self._sp_from_orth(Hu, self.SB)
        Ng = self._sp_from_orth(Hg, self.SD)
        Nv = -(self.M00 @ np.real(H[1][:, 0, 0]))
        Nw = -(self.M00 @ np.real(H[2][:, 0, 0]))
        return (Nu, Ng, Nv, Nw)

    def tp_scalar_product(self, u, S):
        """``TensorProductSpace.scalar_product`` on the unpadded grid
        (``tensorproductspace.py:501-526``): the Fourier scalar products
        (forward sums scaled by 2 pi / df, ``Fourier.py:153-167``), then the
        weighted Chebyshev scalar products (``Chebyshev.py:303-318``) mapped
        onto the basis rows of S."""
        z = fourier_forward(u, 2, self.N[2])
        z = fourier_forward(z, 1, self.N[1])
        return self._apply0(S, cheb_scalar_product(z, self.N0)) * self.W

    def nonlinear_rhs_transform(self, H):
        """``_nonlinear_rhs_physical_reference`` (``channelflow_kmm.py:785-803``),
        the ``coefficient_path="transform"`` rows: derivatives of H taken to the
        unpadded quadrature grid, then scalar products with the test bases."""
        SD = self.SD
        Hu = (self.tp_backward(H[1], SD, padded=False, kx=1, k1=1)
              + self.tp_backward(H[2], SD, padded=False, kx=1, k2=1)
              - self.tp_backward(H[0], SD, padded=False, k1=2)
              - self.tp_backward(H[0], SD, padded=False, k2=2))
        Hg = self.tp_backward(H[1], SD, padded=False, k2=1) - self.tp_backward(H[2], SD, padded=False, k1=1)
        Nu = self.tp_scalar_product(Hu, self.SB)
        Ng = self.tp_scalar_product(Hg, self.SD)
        Nv = -(self.M00 @ np.real(H[1][:, 0, 0]))
        Nw = -(self.M00 @ np.real(H[2][:, 0, 0]))
        return (Nu, Ng, Nv, Nw)

    def reconstruct_transform(self, u0, g, v00, w00):
        """``_reconstruct_velocity_physical_reference`` (``channelflow_kmm.py:831-847``):
        f = TD.forward of d(u0)/dx sampled on the unpadded grid."""
        f = self.tp_forward(self.tp_backward(u0, self.SB, padded=False, kx=1), self.SD, self.MD1)
        u1 = 1j * (self.KoK2[0] * f + self.KoK2[1] * g)
        u2 = 1j * (self.KoK2[1] * f - self.KoK2[0] * g)
        u1[:, 0, 0] = np.real(v00)
        u2[:, 0, 0] = np.real(w00)
        return (self.mask(u0), self.mask(u1), self.mask(u2))

    def reconstruct(self, u0, g, v00, w00):
        """``_reconstruct_velocity_coefficient`` (``channelflow_kmm.py:855-870``)."""
        a = cheb_derivative(self.to_orth(u0, self.SB), 1)
        w = self._apply0(self.SD, a * self.h[:, None, None])
        shp = w.shape
        f = np.linalg.solve(self.MD1, w.reshape(shp[0], -1)).reshape(shp)
        u1 = 1j * (self.KoK2[0] * f + self.KoK2[1] * g)
        u2 = 1j * (self.KoK2[1] * f - self.KoK2[0] * g)
        u1[:, 0, 0] = np.real(v00)
        u2[:, 0, 0] = np.real(w00)
        return (self.mask(u0), self.mask(u1), self.mask(u2))

    def _terms(self, c, mats):
        """Apply sum_m s_m(k) (A_m along axis 0): the TPMatrices matvec as a sum of
        separable Kronecker terms (``tpmatrix.py:153-157``); mats = [(A_m, s_m)]."""
        out = 0
        for A, sk in mats:
            out = out + sk * self._apply0(A, c)
        return out

    def Mu_apply(self, c):
        k2 = self.K1**2 + self.K2**2
        return self._terms(c, [(self.B_C2, self.W), (self.B_C0, -self.W * k2)])

    def Lu_apply(self, c):
        k2 = self.K1**2 + self.K2**2
        nu = self.nu * self.W
        return self._terms(c, [(self.B_C4, nu), (self.B_C2, -2 * nu * k2), (self.B_C0, nu * k2**2)])

    def Mg_apply(self, c):
        return self._terms(c, [(self.D_C0, self.W)])

    def Lg_apply(self, c):
        k2 = self.K1**2 + self.K2**2
        nu = self.nu * self.W
        return self._terms(c, [(self.D_C2, nu), (self.D_C0, -nu * k2)])

    def mass_rows(self, p):
        return (self.Mu_apply(p[0]), self.Mg_apply(p[1]), self.M00 @ p[2], self.M00 @ p[3])

    def _solve(self, factors, rhs_u, rhs_g, rhs_v, rhs_w):
        Su, Sg, S00 = factors
        u0 = self._unmodes(_lu_solve(Su, self._modes(self.mask(rhs_u))))
        g = self.mask(self._unmodes(_lu_solve(Sg, self._modes(self.mask(rhs_g)))))
        vw = np.real(_lu_solve(S00, np.stack((rhs_v, rhs_w), axis=-1)[None])[0])
        return u0, g, vw[:, 0], vw[:, 1]

    def step_imexrk3(self, s: OracleState) -> OracleState:
        """``_step_imexrk3`` (``channelflow_kmm.py:872-935``) with startup factors."""
        dt = self.dt
        u_stage, g_stage = s.u, s.g
        prev = (np.zeros_like(s.u[0]), np.zeros_like(s.g), np.zeros(self.ND), np.zeros(self.ND))
        for rk in range(3):
            cur = self.nonlinear_rhs(self.convection(OracleState(u=u_stage, g=g_stage)))
            a, b = IMEXRK3_A[rk], IMEXRK3_B[rk]
            gamma = (a + b) * dt / 2.0
            v00 = np.real(u_stage[1][:, 0, 0])
            w00 = np.real(u_stage[2][:, 0, 0])
            rhs_u = self.Mu_apply(u_stage[0]) + gamma * self.Lu_apply(u_stage[0])
            rhs_g = self.Mg_apply(g_stage) + gamma * self.Lg_apply(g_stage)
            rhs_v = self.M00 @ v00 + gamma * (self.L00 @ v00)
            rhs_w = self.M00 @ w00 + gamma * (self.L00 @ w00)
            rhs_u = rhs_u + dt * (a * cur[0] + b * prev[0])
            rhs_g = rhs_g + dt * (a * cur[1] + b * prev[1])
            rhs_v = rhs_v + dt * (a * cur[2] + b * prev[2])
            rhs_w = rhs_w + dt * (a * cur[3] + b * prev[3])
            u0n, gn, vn, wn = self._solve(self.startup[rk], rhs_u, rhs_g, rhs_v, rhs_w)
            u_stage = self.reconstruct(u0n, gn, vn, wn)
            g_stage = gn
            prev = cur
        return replace(s, u=u_stage, g=g_stage)

    def linear_rows(self, p):
        """L applied to the primary variables (u0, g, v00, w00)."""
        return (self.Lu_apply(p[0]), self.Lg_apply(p[1]), self.L00 @ p[2], self.L00 @ p[3])

    def step_ars222(self, s: OracleState) -> OracleState:
        """The ARS IMEX-RK step with the IMEXRK222 tableau (``channelflow_kmm.py:1123-1201``,
        ``ars_stage_rhs`` of ``src/jaxfun/integrators/coupled.py:13-44``): one
        implicit operator M - g dt L, stage k's right-hand side
        M x_n + dt sum_j b[k+1][j] N_j + dt sum_j a[k+1][j+1] L x_j.  The history
        fields other than ``nonlinear_old`` are dropped, as the reference's
        ``KMMState`` constructor there leaves them at their defaults."""
        dt = self.dt
        a = ((0.0, 0.0, 0.0), (0.0, ARS222_G, 0.0), (0.0, 1.0 - ARS222_G, ARS222_G))
        b = ((0.0, 0.0, 0.0), (ARS222_G, 0.0, 0.0), (ARS222_D, 1.0 - ARS222_D, 0.0))
        base = self.mass_rows(self.primary(s))
        stage = OracleState(u=s.u, g=s.g)
        nl, lin = [], []
        for rk in range(2):
            nl.append(self.nonlinear_rhs(self.convection(stage)))
            if rk > 0:
                lin.append(self.linear_rows(self.primary(stage)))
            rhs = list(base)
            for j in range(rk + 1):
                rhs = [r + dt * b[rk + 1][j] * n for r, n in zip(rhs, nl[j])]
            for j in range(rk):
                rhs = [r + dt * a[rk + 1][j + 1] * ell for r, ell in zip(rhs, lin[j])]
            u0n, gn, vn, wn = self._solve(self.ars, *rhs)
            stage = OracleState(u=self.reconstruct(u0n, gn, vn, wn), g=gn)
        return OracleState(u=stage.u, g=stage.g, nonlinear_old=s.nonlinear_old, have_old=s.have_old,
                           previous_dt=s.previous_dt)

    def set_dt(self, dt: float) -> None:
        """``KMM.set_dt`` (``channelflow_kmm.py:1207-1222``): SBDF3 keeps its dt;
        the other integrators take the new one (factors at the new gammas)."""
        if self.time_integrator == "SBDF3" and dt != self.dt:
            raise NotImplementedError("SBDF3 runs a fixed dt")
        self.dt = float(dt)
        self.startup = tuple(self._fac((IMEXRK3_A[rk] + IMEXRK3_B[rk]) * self.dt / 2.0) for rk in range(3))
        if self.time_integrator == "IMEXRK222":
            self.ars = self._fac(self.dt * ARS222_G)

    def step_cnab2(self, s: OracleState) -> OracleState:
        """``_step_cnab2`` / ``_cnab2_flow_update`` (``channelflow_kmm.py:1071-1121``):
        Crank-Nicolson for L, variable-step Adams-Bashforth-2
        (``variable_ab2_extrapolate``, ``src/jaxfun/integrators/cnab2.py:211-239``)
        for N; the new state keeps only ``nonlinear_old``, ``have_old``,
        ``previous_dt``."""
        dt = self.dt
        p = self.primary(s)
        cur = self.nonlinear_rhs(self.convection(s))
        ext = variable_ab2_extrapolate(cur, s.nonlinear_old, s.have_old, dt, s.previous_dt)
        m, ell = self.mass_rows(p), self.linear_rows(p)
        rhs = [a + 0.5 * dt * b + dt * e for a, b, e in zip(m, ell, ext)]
        u0n, gn, vn, wn = self._solve(self._fac(0.5 * dt), *rhs)
        return OracleState(u=self.reconstruct(u0n, gn, vn, wn), g=gn, nonlinear_old=cur, have_old=1.0,
                           previous_dt=dt)

    def step(self, s: OracleState) -> OracleState:
        """One step of the configured integrator: ``_step_sbdf3``
        (``channelflow_kmm.py:937-1064``, startup then steady), ``_step_imexrk3``
        (``:872-935``, history passed through), the IMEXRK222 ARS step or CNAB2."""
        if self.time_integrator == "IMEXRK3":
            return self.step_imexrk3(s)
        if self.time_integrator == "IMEXRK222":
            return self.step_ars222(s)
        if self.time_integrator == "CNAB2":
            return self.step_cnab2(s)
        cur_p = self.primary(s)
        cur_n = self.nonlinear_rhs(self.convection(s))
        if s.history_steps < 2:
            adv = self.step_imexrk3(s)
            hist = min(2.0, s.history_steps + 1.0)
        else:
            m0 = self.mass_rows(cur_p)
            m1 = self.mass_rows(s.solution_old)
            m2 = self.mass_rows(s.solution_older)
            rhs = tuple(
                (18.0 * a - 9.0 * b + 2.0 * c) / 11.0 + self.dt * ((18.0 * n0 - 18.0 * n1 + 6.0 * n2) / 11.0)
                for a, b, c, n0, n1, n2 in zip(m0, m1, m2, cur_n, s.nonlinear_old, s.nonlinear_older)
            )
            u0n, gn, vn, wn = self._solve(self.steady, *rhs)
            adv = replace(s, u=self.reconstruct(u0n, gn, vn, wn), g=gn)
            hist = 2.0
        return replace(
            adv,
            nonlinear_old=cur_n,
            nonlinear_older=s.nonlinear_old,
            solution_old=cur_p,
            solution_older=s.solution_old,
            history_steps=hist,
            have_old=1.0,
            previous_dt=self.dt,
        )

    def solve(self, s: OracleState, steps: int) -> OracleState:
        for _ in range(int(steps)):
            s = self.step(s)
        return s

    # ---------------------------------------------------------------- diagnostics
    def _weights(self):
        wx = fejer_weights(self.N0)[:, None, None]
        w1 = (2 * np.pi / self.N[1]) / self.df1
        w2 = (2 * np.pi / self.N[2]) / self.df2
        return wx * w1 * w2

    def integrate(self, vals):
        return np.sum(vals * self._weights())

    def diagnostics(self, s: OracleState) -> dict:
        """``PlaneCouetteFluctuationJax.diagnostics`` (``pcf_fluctuations_jax.py:131-161``)."""
        u0, u1, u2 = s.u
        up = (self.tp_backward(u0, self.SB, padded=False), self.tp_backward(u1, self.SD, padded=False),
              self.tp_backward(u2, self.SD, padded=False))
        ut = (up[0], up[1] + self.dUb * self.xb, up[2])
        epert = sum(np.real(self.integrate(np.conj(v) * v)) for v in up)
        etot = sum(np.real(self.integrate(np.conj(v) * v)) for v in ut)
        divu = (self.tp_backward(u0, self.SB, padded=False, kx=1) + self.tp_backward(u1, self.SD, padded=False, k1=1)
                + self.tp_backward(u2, self.SD, padded=False, k2=1))
        divl2 = np.sqrt(np.real(self.integrate(np.conj(divu) * divu)))
        dvdx = self.tp_backward(u1, self.SD, padded=False, kx=1)
        volume = 2.0 * self.L1 * self.L2
        # wall traces: phi_i(+-1)=0; phi_i'(1)=i^2-(i+2)^2, phi_i'(-1)=-(i^2-(i+2)^2)*(-1)^i
        i = np.arange(self.ND)
        d_top = i**2 - (i + 2) ** 2
        d_bot = ((-1.0) ** (i + 1)) * i**2 - ((-1.0) ** (i + 3)) * (i + 2) ** 2
        top = np.real(np.dot(d_top, u1[:, 0, 0])) + self.dUb
        bot = np.real(np.dot(d_bot, u1[:, 0, 0])) + self.dUb
        return {
            "Epert": float(epert),
            "Etot": float(etot),
            "divL2": float(divl2),
            "u_top": self.dUb * 1.0,
            "u_bot": -self.dUb * 1.0,
            "wall_shear_top": float(top),
            "wall_shear_bot": float(bot),
            "mean_shear": float(np.real(self.integrate(dvdx + self.dUb)) / volume),
        }

    def backward_velocity(self, s: OracleState):
        """``_backward_velocity`` (``channelflow_kmm.py:688-697``) on the unpadded grid."""
        u0, u1, u2 = s.u
        return (self.tp_backward(u0, self.SB, padded=False), self.tp_backward(u1, self.SD, padded=False),
                self.tp_backward(u2, self.SD, padded=False))

    def fluctuation_scalars(self, s: OracleState) -> dict:
        """``_pcf_fluctuation_scalars`` (``production/oracles.py:4566-4597``)."""
        d = self.diagnostics(s)
        up = self.backward_velocity(s)
        volume = 2.0 * self.L1 * self.L2
        streak = np.sqrt(np.real(self.integrate(up[1] * np.conj(up[1]))) / volume)
        roll = np.sqrt((np.real(self.integrate(up[0] * np.conj(up[0])))
                        + np.real(self.integrate(up[2] * np.conj(up[2])))) / volume)
        return {"kinetic_energy": d["Epert"], "total_kinetic_energy": d["Etot"], "divergence_l2": d["divL2"],
                "wall_shear_lower": d["wall_shear_bot"], "wall_shear_upper": d["wall_shear_top"],
                "wall_velocity_lower": d["u_bot"], "wall_velocity_upper": d["u_top"],
                "mean_shear": d["mean_shear"], "streak_rms": float(streak), "roll_rms": float(roll)}

    @staticmethod
    def checksum(s: OracleState) -> dict:
        """Full-state checksum (``production/benchmark.py:234-260``)."""
        real_sum = imag_sum = norm2 = 0.0
        count = 0
        for v in s.leaves():
            v = np.asarray(v)
            real_sum += float(np.sum(np.real(v), dtype=np.float64))
            imag_sum += float(np.sum(np.imag(v), dtype=np.float64))
            norm2 += float(np.sum(np.abs(v) ** 2, dtype=np.float64))
            count += int(v.size)
        return {"leaf_count": len(s.leaves()), "element_count": count, "real_sum": real_sum,
                "imag_sum": imag_sum, "l2_norm": math.sqrt(norm2)}
