"""Pin of the oracle's MHD forcing terms by direct-series evaluation (CPU).

The oracle's flow forcing H and induction forcing HA
(``oracle/mhd.py:mhd_convection`` with the rotating-shear hooks) restate
``_mhd_convection`` of ``examples/pcf_mhd_mri_shearpy_jax.py:153-176``:

    n   = -(u_tot x w_tot)                       (channelflow_kmm.py:754-762)
    n  += (-2 Omega u1, +2 Omega u0, 0)          (Coriolis, fluctuation u; :156-160)
    H   = P_TD[ n - j x (b + B0) ]               (:162-170; B0 added in :149-151)
    HA  = P_A [ u_tot x (b + B0) ]               (:172-175)

with u_tot = (u0, u1 + U_b, u2), w_tot = (w0, w1, w2 + dU_b/dx), U_b = -S x
(:86-92), b = curl A and j = curl b.  Here every physical field is evaluated
at every node of the 3/2-padded grid as a direct sum of its series -- basis
polynomials from ``numpy.polynomial.chebyshev`` with the Shen stencils derived
from their boundary conditions, explicit exponentials for the Fourier
directions, derivatives taken term by term -- with no FFT, DCT or stencil of
the oracle.  The products are formed pointwise, then projected by explicit
Fourier sums and Gauss-Chebyshev quadrature Galerkin solves.

With the insulating walls A = (A_x in TD, A_y and A_z in TC), b and j lie in
the Chebyshev space exactly, so the direct series is the exact field.  A state
with Omega, S and all three B0 components nonzero exercises every term: the
planted-mistake tests below show that the comparison fails by many orders of
magnitude when the oracle's Coriolis sign is flipped, its factor 2 dropped, it
acts on the total instead of the fluctuation velocity, or B0 is left out of
either product.  The conducting rotating-shear oracle shares ``mhd_convection``
and ``flow_forcing`` with the insulating one.
"""

from __future__ import annotations

import math

import numpy as np
import numpy.polynomial.chebyshev as npc
import pytest

from oracle.kmm import OracleState
from oracle.mhd import MHDOracleState
from oracle.mhd_insulating import InsulatingMHDOracle

N = (9, 8, 8)
PARAMS = dict(Re=300.0, Rm=250.0, omega=0.37, shear_rate=0.8, background_b=(0.21, -0.13, 0.3), dt=1e-3)
TOL = 1e-12


# ---------------------------------------------------------------- bases from boundary conditions
def _dirichlet_cheb(n0: int) -> np.ndarray:
    """phi_k = T_k + a T_{k+2} with phi(+-1) = 0 (Chebyshev coefficient rows)."""
    S = np.zeros((n0 - 2, n0))
    for k in range(n0 - 2):
        # T_k(1) = T_{k+2}(1) = 1 and parity gives phi(-1) = 0 too
        S[k, k], S[k, k + 2] = 1.0, -1.0
    return S


def _clamped_cheb(n0: int) -> np.ndarray:
    """psi_k = T_k + a T_{k+2} + b T_{k+4} with psi(+-1) = psi'(+-1) = 0.

    T_m(1) = 1 and T_m'(1) = m^2; parity of T_k, T_{k+2}, T_{k+4} carries the
    conditions at x = 1 to x = -1, so a and b solve a 2x2 system."""
    S = np.zeros((n0 - 4, n0))
    for k in range(n0 - 4):
        a, b = np.linalg.solve([[1.0, 1.0], [(k + 2.0) ** 2, (k + 4.0) ** 2]], [-1.0, -float(k * k)])
        S[k, k], S[k, k + 2], S[k, k + 4] = 1.0, a, b
    return S


def _cheb_identity(n0: int) -> np.ndarray:
    return np.eye(n0)


class _Grid:
    """Padded nodes, FFT-ordered wavenumbers and direct-series evaluation."""

    def __init__(self, N, pad, Ly, Lz):
        self.n0, self.n1, self.n2 = N
        self.m1, self.m2 = pad
        self.x = np.cos((2 * np.arange(self.n0) + 1) * np.pi / (2 * self.n0))  # any order
        self.w = np.full(self.n0, np.pi / self.n0)  # Gauss-Chebyshev weights
        self.y = Ly * np.arange(self.m1) / self.m1
        self.z = Lz * np.arange(self.m2) / self.m2

        def k(n, L):
            i = np.arange(n)
            return np.where(i < (n + 1) // 2, i, i - n) * (2 * np.pi / L)

        self.ky, self.kz = k(self.n1, Ly), k(self.n2, Lz)
        self.Ey = np.exp(1j * np.outer(self.y, self.ky))  # (m1, n1)
        self.Ez = np.exp(1j * np.outer(self.z, self.kz))  # (m2, n2)
        self.nyq = np.zeros((self.n1, self.n2), bool)
        self.nyq[self.n1 // 2, :] = True
        self.nyq[:, self.n2 // 2] = True

    def basis_values(self, S, dx):
        """Values of d^dx/dx^dx of every basis function at the nodes: (n0, nbasis)."""
        return np.stack([npc.chebval(self.x, npc.chebder(row, dx) if dx else row) for row in S], axis=1)

    def value(self, c, S, dx=0, dy=0, dz=0):
        """Physical values of d^(dx,dy,dz) of the series c (nbasis, n1, n2) on the padded grid."""
        V = self.basis_values(S, dx)
        fac = ((1j * self.ky[:, None]) ** dy) * ((1j * self.kz[None, :]) ** dz)
        cf = c * fac[None]
        return np.einsum("nk,pa,qb,kab->npq", V, self.Ey, self.Ez, cf, optimize=True)

    def project(self, f, S):
        """Truncating forward transform: Fourier sums over the padded grid, then
        the Galerkin solve onto the basis rows of S by quadrature."""
        fh = np.einsum("npq,pa,qb->nab", f, self.Ey.conj(), self.Ez.conj(), optimize=True) / (self.m1 * self.m2)
        V = self.basis_values(S, 0)
        M = (V * self.w[:, None]).T @ V
        rhs = np.einsum("nk,n,nab->kab", V, self.w, fh)
        c = np.linalg.solve(M, rhs.reshape(M.shape[0], -1)).reshape(rhs.shape)
        return np.where(self.nyq[None], 0, c)


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _hermitian(rng, rows, n1, n2):
    """Random coefficients of a real field: c(-ky,-kz) = conj c(ky,kz), Nyquist zero."""
    c = rng.standard_normal((rows, n1, n2)) + 1j * rng.standard_normal((rows, n1, n2))
    ia, ib = (-np.arange(n1)) % n1, (-np.arange(n2)) % n2
    c = 0.5 * (c + np.conj(c[:, ia][:, :, ib]))
    c[:, n1 // 2, :] = 0
    c[:, :, n2 // 2] = 0
    return c


def _state(seed=3):
    rng = np.random.default_rng(seed)
    n0, n1, n2 = N
    u = (_hermitian(rng, n0 - 4, n1, n2), _hermitian(rng, n0 - 2, n1, n2), _hermitian(rng, n0 - 2, n1, n2))
    # A scaled so that b = curl A and B0 are of one size
    A = tuple(0.05 * _hermitian(rng, r, n1, n2) for r in (n0 - 2, n0, n0))
    return u, A


def _direct_forcing(u, A, omega, shear_rate, b0, grid: _Grid):
    """H and HA of the reference formulas, every field a direct series sum."""
    SB, SD, SC = _clamped_cheb(grid.n0), _dirichlet_cheb(grid.n0), _cheb_identity(grid.n0)
    spaces_u = (SB, SD, SD)
    spaces_A = (SD, SC, SC)
    ux = lambda i, **d: grid.value(u[i], spaces_u[i], **d)  # noqa: E731
    ax = lambda i, **d: grid.value(A[i], spaces_A[i], **d)  # noqa: E731
    up = (ux(0), ux(1), ux(2))
    w = (ux(2, dy=1) - ux(1, dz=1), ux(0, dz=1) - ux(2, dx=1), ux(1, dx=1) - ux(0, dy=1))
    xb = grid.x[:, None, None]
    ut = (up[0], up[1] - shear_rate * xb, up[2])
    wt = (w[0], w[1], w[2] - shear_rate)
    n = tuple(-c for c in _cross(ut, wt))
    n = (n[0] - 2.0 * omega * up[1], n[1] + 2.0 * omega * up[0], n[2])
    # b = curl A
    b = (ax(2, dy=1) - ax(1, dz=1), ax(0, dz=1) - ax(2, dx=1), ax(1, dx=1) - ax(0, dy=1))
    # j = curl b, each derivative of b taken term by term on A's series
    dby_dz = ax(0, dz=2) - ax(2, dx=1, dz=1)
    dbz_dy = ax(1, dx=1, dy=1) - ax(0, dy=2)
    dbx_dz = ax(2, dy=1, dz=1) - ax(1, dz=2)
    dbz_dx = ax(1, dx=2) - ax(0, dx=1, dy=1)
    dbx_dy = ax(2, dy=2) - ax(1, dy=1, dz=1)
    dby_dx = ax(0, dx=1, dz=1) - ax(2, dx=2)
    j = (dbz_dy - dby_dz, dbx_dz - dbz_dx, dby_dx - dbx_dy)
    bt = tuple(bi + b0i for bi, b0i in zip(b, b0))
    lor = _cross(j, bt)
    H = tuple(grid.project(ni - li, SD) for ni, li in zip(n, lor))
    emf = _cross(ut, bt)
    HA = (grid.project(emf[0], SD), grid.project(emf[1], SC), grid.project(emf[2], SC))
    return H, HA


def _mismatch(oracle, u, A, omega, shear_rate, b0):
    grid = _Grid(N, oracle.pad[1:], oracle.L1, oracle.L2)
    s = MHDOracleState(flow=OracleState(u=u, g=np.zeros_like(u[1])), A=A)
    H, HA = oracle.mhd_convection(s)
    Hd, HAd = _direct_forcing(u, A, omega, shear_rate, b0, grid)
    out = {}
    for name, a, bref in (("H", H, Hd), ("HA", HA, HAd)):
        scale = max(float(np.max(np.abs(x))) for x in bref)
        out[name] = max(float(np.max(np.abs(x - y))) for x, y in zip(a, bref)) / scale
    return out


def test_direct_series_bases_match_boundary_conditions():
    """The test's own bases: clamped psi_k(+-1) = psi_k'(+-1) = 0, Dirichlet phi_k(+-1) = 0."""
    n0 = 11
    for S, orders in ((_clamped_cheb(n0), (0, 1)), (_dirichlet_cheb(n0), (0,))):
        for row in S:
            for d in orders:
                vals = npc.chebval(np.array([-1.0, 1.0]), npc.chebder(row, d) if d else row)
                assert np.max(np.abs(vals)) < 1e-13


def test_oracle_mhd_forcing_matches_direct_series():
    """Oracle H and HA at Omega, S and B0 all nonzero equal the direct-series
    evaluation of the reference formulas to 1e-12 relative."""
    u, A = _state()
    oracle = InsulatingMHDOracle(N=N, **PARAMS)
    err = _mismatch(oracle, u, A, PARAMS["omega"], PARAMS["shear_rate"], PARAMS["background_b"])
    assert err["H"] < TOL and err["HA"] < TOL, err


class _FlippedCoriolis(InsulatingMHDOracle):
    def flow_forcing(self, n, up):
        return (n[0] + 2.0 * self.omega * up[1], n[1] - 2.0 * self.omega * up[0], n[2])


class _HalfCoriolis(InsulatingMHDOracle):
    def flow_forcing(self, n, up):
        return (n[0] - self.omega * up[1], n[1] + self.omega * up[0], n[2])


class _TotalVelocityCoriolis(InsulatingMHDOracle):
    def flow_forcing(self, n, up):
        ut1 = up[1] + self.dUb * self.xb
        return (n[0] - 2.0 * self.omega * ut1, n[1] + 2.0 * self.omega * up[0], n[2])


@pytest.mark.parametrize("planted,kw", [
    (_FlippedCoriolis, {}),
    (_HalfCoriolis, {}),
    (_TotalVelocityCoriolis, {}),
    (InsulatingMHDOracle, dict(background_b=(0.0, 0.0, 0.0))),      # B0 dropped everywhere
    (InsulatingMHDOracle, dict(background_b=(0.21, -0.13, -0.3))),  # one B0 sign flipped
    (InsulatingMHDOracle, dict(omega=-0.37)),                       # rotation sign flipped
])
def test_planted_forcing_mistakes_fail_the_pin(planted, kw):
    """A wrong Coriolis sign, scale or velocity, or a dropped/flipped B0, in the
    oracle moves H (and HA for B0) far beyond the pin's tolerance."""
    u, A = _state()
    params = dict(PARAMS)
    params.update(kw)
    err = _mismatch(planted(N=N, **params), u, A, PARAMS["omega"], PARAMS["shear_rate"],
                    PARAMS["background_b"])
    assert err["H"] > 1e3 * TOL, err
    if "background_b" in kw:
        assert err["HA"] > 1e3 * TOL, err


def test_emf_without_b0_only_through_b():
    """HA of the direct series at B0 = 0 differs from B0 != 0 by P[u_tot x B0],
    which is O(1) here: the B0 term of the EMF is not negligible in the pin."""
    u, A = _state()
    oracle = InsulatingMHDOracle(N=N, **PARAMS)
    grid = _Grid(N, oracle.pad[1:], oracle.L1, oracle.L2)
    _, HA1 = _direct_forcing(u, A, PARAMS["omega"], PARAMS["shear_rate"], PARAMS["background_b"], grid)
    _, HA0 = _direct_forcing(u, A, PARAMS["omega"], PARAMS["shear_rate"], (0.0, 0.0, 0.0), grid)
    rel = max(float(np.max(np.abs(a - b))) for a, b in zip(HA1, HA0)) / max(float(np.max(np.abs(a))) for a in HA1)
    assert rel > 1e-2
    assert math.isfinite(rel)
