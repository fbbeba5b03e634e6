"""Taylor-Couette DNS on the MI355X (C ABI include/tc_dns.h).

Host-side mirror of the reference's ``AxisymmetricTCDNSJax``
(``examples/taylor_couette_dns_jax.py:217-1078``), of the full 3D
``TaylorCouetteDNSJax`` (``:1081-1356``) and of their base flow
``CircularCouette`` (``examples/taylor_couette_linear_jax.py:31-89``): the same
constructor arguments, state fields (``AxisymmetricTCState``, ``:114-165``)
and methods.  Every step runs in the HIP library; the host converts states
between the reference's full-complex layout and the library, assembles the
linear-eigenmode seed from the library's pencil with ``scipy.linalg.eig`` (as
the reference does on the host), and evaluates physical fields for
``velocity_physical``.  Without ``libpcfdns.so`` the constructor raises.
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, replace

import numpy as np
import scipy.linalg as sla

from . import _native as nat

_dp = C.POINTER(C.c_double)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


class CircularCouette:
    """Circular-Couette base flow V = a r + b / r between cylinders R1 < R2
    rotating at Omega1, Omega2 (``examples/taylor_couette_linear_jax.py:31-89``)."""

    def __init__(self, R1=1.0, R2=2.0, Omega1=1.0, Omega2=0.0):
        if not (R2 > R1 > 0):
            raise ValueError("require 0 < R1 < R2")
        self.R1, self.R2 = float(R1), float(R2)
        self.Omega1, self.Omega2 = float(Omega1), float(Omega2)
        d2 = self.R2**2 - self.R1**2
        self.a = (self.Omega2 * self.R2**2 - self.Omega1 * self.R1**2) / d2
        self.b = (self.Omega1 - self.Omega2) * self.R1**2 * self.R2**2 / d2
        self.gap = self.R2 - self.R1
        self.eta = self.R1 / self.R2
        self.mu = self.Omega2 / self.Omega1 if self.Omega1 != 0 else math.inf
        self.shear2a = 2 * self.a

    def Omega(self, r):
        return self.a + self.b / np.asarray(r) ** 2

    def V(self, r):
        r = np.asarray(r)
        return self.a * r + self.b / r

    def kappa2(self, r):
        return 4 * self.a * self.Omega(r)

    def rayleigh_stable(self):
        rr = np.linspace(self.R1, self.R2, 200)
        return bool(np.all(self.kappa2(rr) > 0))

    def q_shear(self, r):
        r = np.asarray(r)
        return 2 * self.b / (self.a * r**2 + self.b)


@dataclass(frozen=True)
class AxisymmetricTCState:
    """Coefficient state (``examples/taylor_couette_dns_jax.py:114-165``):
    u = (u_r, u_theta, u_z), p and the histories, each (Nz, Nr - 2) complex in
    the full-complex layout."""

    u: tuple
    p: np.ndarray
    nonlinear_old: tuple
    have_old: float = 0.0
    nonlinear_older: tuple | None = None
    solution_old: tuple | None = None
    solution_older: tuple | None = None
    history_steps: float = 0.0


def _positive_pivot_phase(vec):
    """``examples/taylor_couette_dns_jax.py:99-111``."""
    i = int(np.argmax(np.abs(vec)))
    piv = vec[i]
    mag = float(abs(piv))
    if mag == 0.0:
        return vec
    if piv.real < 0.0 or (abs(piv.real) <= 1.0e-14 * mag and piv.imag < 0.0):
        return -vec
    return vec


class AxisymmetricTCDNSHIP:
    """``AxisymmetricTCDNSJax`` on the GPU (``examples/taylor_couette_dns_jax.py:217-1078``)."""

    def __init__(self, base: CircularCouette, nu: float = 1.0e-2, Nr: int = 32, Nz: int = 16, Lz: float | None = None,
                 dt: float = 2.0e-3, family: str = "C", dealias: float = 1.5, time_integrator: str = "CNAB2",
                 device: int = 0):
        ti = str(time_integrator).upper()
        if ti not in nat.TC_TI:
            raise ValueError("time_integrator must be one of {'CNAB2', 'SBDF3', 'IMEXRK3'}")
        fam = family.upper()
        if fam.startswith("L"):
            raise NotImplementedError("the Legendre family is not built; family='C' runs")
        if not fam.startswith("C"):
            raise ValueError("family must be 'L' or 'C'")
        self.base = base
        self.nu = float(nu)
        self.Nr, self.Nz = int(Nr), int(Nz)
        self.dt = float(dt)
        self.time_integrator = ti
        self.family = fam
        self.dealias = float(dealias)
        self.Lz = float(Lz) if Lz is not None else 2.0 * math.pi / 3.13 * base.gap
        self.Re = base.Omega1 * base.R1 * base.gap / self.nu
        self.n = self.Nr - 2
        self._lib = nat.load()
        self._cfg = self._config(device)
        h = C.c_void_p()
        nat.tc_check(self._lib.tc_create(C.byref(self._cfg), C.byref(h)))
        self._h = h
        # standard radial grid (ascending Gauss-Chebyshev nodes mapped to [R1, R2])
        x = np.cos(np.pi + (2 * np.arange(self.Nr) + 1) * np.pi / (2 * self.Nr))
        self._x = x
        self.r = base.R1 + (x + 1.0) * (base.R2 - base.R1) / 2.0
        z = np.arange(self.Nz) * self.Lz / self.Nz
        self.Z, self.R = np.meshgrid(z, self.r, indexing="ij")

    def _config(self, device: int) -> nat.TcConfig:
        c = nat.TcConfig()
        c.abi_version = nat.TC_ABI_VERSION
        c.nr, c.nz = self.Nr, self.Nz
        c.r1, c.r2 = self.base.R1, self.base.R2
        c.omega1, c.omega2 = self.base.Omega1, self.base.Omega2
        c.nu, c.lz, c.dt = self.nu, self.Lz, self.dt
        c.dealias = self.dealias
        c.time_integrator = nat.TC_TI[self.time_integrator]
        c.device = int(device)
        c.ntheta = 1
        return c

    def close(self):
        if getattr(self, "_h", None):
            self._lib.tc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- states
    @property
    def _shape(self):
        return (self.Nz, self.n)

    def _zeros(self):
        return np.zeros(self._shape, dtype=complex)

    def zero_state(self) -> AxisymmetricTCState:
        u = tuple(self._zeros() for _ in range(3))
        return AxisymmetricTCState(u=u, p=self._zeros(), nonlinear_old=tuple(self._zeros() for _ in range(3)))

    def _vander(self):
        return np.polynomial.chebyshev.chebvander(self._x, self.Nr - 1)  # (q, k)

    def _forward(self, values) -> np.ndarray:
        """TD.forward: axial fft (norm='forward'), then the weighted Galerkin
        projection onto phi_i = T_i - T_{i+2}."""
        c = np.fft.fft(np.asarray(values, dtype=complex), axis=0, norm="forward")
        V = self._vander()
        phi = V[:, : self.n] - V[:, 2:]
        w = np.pi / self.Nr
        rhs = w * (c @ phi)
        M = w * phi.T @ phi
        return np.linalg.solve(M, rhs.T).T

    def _backward(self, coef) -> np.ndarray:
        V = self._vander()
        phi = V[:, : self.n] - V[:, 2:]
        return np.fft.ifft(np.asarray(coef) @ phi.T, axis=0, norm="forward")

    def state_from_physical(self, values) -> AxisymmetricTCState:
        u = tuple(self._forward(v) for v in values)
        return AxisymmetricTCState(u=u, p=self._zeros(), nonlinear_old=tuple(self._zeros() for _ in range(3)))

    def initial_state(self, amp: float = 1.0e-3, kz_mode: int = 1) -> AxisymmetricTCState:
        """Divergence-free streamfunction seed (``examples/taylor_couette_dns_jax.py:651-667``)."""
        R1, d = self.base.R1, self.base.gap
        kz = 2.0 * math.pi * int(kz_mode) / self.Lz
        arg = math.pi * (self.R - R1) / d
        g = np.sin(arg) ** 2
        gp = (2.0 * math.pi / d) * np.sin(arg) * np.cos(arg)
        ur = amp * (kz / self.R) * g * np.sin(kz * self.Z)
        uz = amp * (1.0 / self.R) * gp * np.cos(kz * self.Z)
        ut = amp * np.sin(arg) * np.cos(kz * self.Z)
        return self.state_from_physical((ur, ut, uz))

    def linear_pencil(self, m: int, kz: float):
        """(L, M) of ``TaylorCouetteLinearJax.assemble`` on this radial space."""
        n4 = 4 * self.n
        L = np.empty((n4, n4), dtype=complex)
        M = np.empty((n4, n4), dtype=complex)
        nat.tc_check(self._lib.tc_linear_pencil(C.byref(self._cfg), int(m), float(kz),
                                                L.view(np.float64).ctypes.data_as(_dp),
                                                M.view(np.float64).ctypes.data_as(_dp)))
        return L, M

    def linear_eigs(self, m: int, kz: float, n_return: int = 8, finite_cap: float = 1.0e6):
        """Growth-sorted finite eigenpairs (``src/jaxfun/la/eig.py:145-195``)."""
        w, V = sla.eig(*self.linear_pencil(m, kz))
        keep = np.isfinite(w) & (np.abs(w) < finite_cap)
        w, V = w[keep], V[:, keep]
        order = np.lexsort((-w.imag, -w.real))
        return w[order][:n_return], V[:, order][:, :n_return]

    def seed_linear_eigenmode(self, kz_mode: int = 1, amp: float = 1.0e-6, which: int = 0):
        """Real part of the leading linear eigenmode (``:669-701``)."""
        kz = 2.0 * math.pi * int(kz_mode) / self.Lz
        w, vecs = self.linear_eigs(0, kz, n_return=which + 1)
        vec = _positive_pivot_phase(vecs[:, which])
        n = self.n
        kpos, kneg = int(kz_mode) % self.Nz, (-int(kz_mode)) % self.Nz
        comps = []
        for c in range(3):
            block = vec[c * n:(c + 1) * n] * amp
            arr = self._zeros()
            if kpos == kneg:
                arr[kpos, :n] = np.real(block)
            else:
                arr[kpos, :n] = 0.5 * block
                arr[kneg, :n] = 0.5 * np.conj(block)
            comps.append(arr)
        st = AxisymmetricTCState(u=tuple(comps), p=self._zeros(), nonlinear_old=tuple(self._zeros() for _ in range(3)),
                                 have_old=0.0)
        return st, complex(w[which])

    # ------------------------------------------------------------ device I/O
    def _set(self, name: str, arr):
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.complex128))
        if a.shape != self._shape:
            raise ValueError(f"{name}: expected shape {self._shape}, got {a.shape}")
        nat.tc_check(self._lib.tc_set_field(self._h, nat.TC_FIELD_ID[name], a.view(np.float64).ctypes.data_as(_dp)))

    def _get(self, name: str) -> np.ndarray:
        a = np.empty(self._shape, dtype=np.complex128)
        nat.tc_check(self._lib.tc_get_field(self._h, nat.TC_FIELD_ID[name], a.view(np.float64).ctypes.data_as(_dp)))
        return a

    @staticmethod
    def _ensure_history(s: AxisymmetricTCState) -> AxisymmetricTCState:
        """``:851-863``."""
        z = tuple(np.zeros_like(v) for v in s.u)
        return replace(s, nonlinear_older=z if s.nonlinear_older is None else s.nonlinear_older,
                       solution_old=s.u if s.solution_old is None else s.solution_old,
                       solution_older=s.u if s.solution_older is None else s.solution_older)

    def load(self, s: AxisymmetricTCState) -> None:
        if self.time_integrator == "SBDF3":
            s = self._ensure_history(s)
        for c, name in enumerate(("UR", "UT", "UZ")):
            self._set(name, s.u[c])
        self._set("P", s.p)
        groups = (("NL_OLD", s.nonlinear_old), ("NL_OLDER", s.nonlinear_older), ("SOL_OLD", s.solution_old),
                  ("SOL_OLDER", s.solution_older))
        for prefix, grp in groups:
            if grp is not None:
                for c, sfx in enumerate("RTZ"):
                    self._set(f"{prefix}_{sfx}", grp[c])
        nat.tc_check(self._lib.tc_set_scalars(self._h, float(s.have_old), float(s.history_steps)))

    def fetch(self, like: AxisymmetricTCState) -> AxisymmetricTCState:
        grp = lambda p: tuple(self._get(f"{p}_{c}") for c in "RTZ")  # noqa: E731
        ho, hs = C.c_double(), C.c_double()
        nat.tc_check(self._lib.tc_get_scalars(self._h, C.byref(ho), C.byref(hs)))
        u = tuple(self._get(n) for n in ("UR", "UT", "UZ"))
        if self.time_integrator == "SBDF3":
            return AxisymmetricTCState(u=u, p=self._get("P"), nonlinear_old=grp("NL_OLD"), have_old=ho.value,
                                       nonlinear_older=grp("NL_OLDER"), solution_old=grp("SOL_OLD"),
                                       solution_older=grp("SOL_OLDER"), history_steps=hs.value)
        if self.time_integrator == "CNAB2":
            return replace(like, u=u, p=self._get("P"), nonlinear_old=grp("NL_OLD"), have_old=ho.value)
        return replace(like, u=u, p=self._get("P"))

    # ------------------------------------------------------------- stepping
    def advance(self, steps: int) -> None:
        """Steps on the device state (graph replay of steady SBDF3 blocks)."""
        nat.tc_check(self._lib.tc_step_graph(self._h, int(steps)))

    def step(self, state: AxisymmetricTCState) -> AxisymmetricTCState:
        """``:948-969``."""
        self.load(state)
        nat.tc_check(self._lib.tc_step(self._h, 1))
        return self.fetch(state)

    def solve(self, state: AxisymmetricTCState, steps: int) -> AxisymmetricTCState:
        """``:987-990``."""
        if int(steps) < 0:
            raise ValueError("steps must be non-negative")
        self.load(state)
        self.advance(int(steps))
        return self.fetch(state)

    def set_dt(self, dt: float) -> None:
        """``:971-985``: SBDF3 keeps a fixed dt; the others rebuild their operators."""
        nat.tc_check(self._lib.tc_set_dt(self._h, float(dt)))
        self.dt = float(dt)
        self._cfg.dt = self.dt

    # ------------------------------------------------------------ diagnostics
    def _device_diagnostics(self, state: AxisymmetricTCState | None) -> np.ndarray:
        if state is not None:
            self.load(state)
        out = np.zeros(len(nat.TC_DIAG_KEYS))
        nat.tc_check(self._lib.tc_diagnostics(self._h, _ptr(out)))
        return out

    def diagnostics(self, state: AxisymmetricTCState) -> dict:
        """E, div_linf, continuity_l2, wall, Eth (``:1061-1069``), device reductions."""
        d = self._device_diagnostics(state)
        return dict(zip(nat.TC_DIAG_KEYS, (float(v) for v in d)))

    def energy(self, state: AxisymmetricTCState) -> float:
        return self.diagnostics(state)["E"]

    def divergence_linf(self, state: AxisymmetricTCState) -> float:
        return self.diagnostics(state)["div_linf"]

    def continuity_residual_l2(self, state: AxisymmetricTCState) -> float:
        return self.diagnostics(state)["continuity_l2"]

    def velocity_physical(self, state: AxisymmetricTCState):
        """u on the standard (Nz, Nr) grid (``:1036-1037``), evaluated on the host."""
        return tuple(self._backward(c) for c in state.u)

    def radial_velocity_linf(self, state: AxisymmetricTCState) -> float:
        """max |u_r| on the standard grid (``production/oracles.py:4600-4604``)."""
        return float(np.max(np.abs(self.velocity_physical(state)[0])))

    def inner_torque(self, state: AxisymmetricTCState) -> float:
        """Torque on the inner cylinder (``production/oracles.py:4620-4638``):
        2 pi nu R1^2 |d(u_theta)/dr - u_theta / r + the base shear -2b / R1^2|
        at r = R1, the axial mean.  The Dirichlet modes vanish at the wall and
        T_k'(-1) = (-1)^(k+1) k^2."""
        r1, b = self.base.R1, self.base.b
        df = 2.0 / (self.base.R2 - self.base.R1)
        k = np.arange(self.Nr, dtype=float)
        dT = ((-1.0) ** (k + 1)) * k**2
        dphi = df * (dT[: self.n] - dT[2:])
        mean_dut = np.asarray(state.u[1])[0] @ dphi  # the k = 0 mode is the axial mean
        return float(2.0 * math.pi * self.nu * r1**2 * abs(-2.0 * b / r1**2 + mean_dut))

    def growth_rate(self, state: AxisymmetricTCState, steps: int):
        """``:1071-1078``."""
        e0 = self.energy(state)
        out = self.solve(state, steps)
        e1 = self.energy(out)
        return 0.5 * math.log(e1 / e0) / (int(steps) * self.dt), out


class TaylorCouetteDNSHIP(AxisymmetricTCDNSHIP):
    """``TaylorCouetteDNSJax`` on the GPU (``examples/taylor_couette_dns_jax.py:1081-1356``):
    Fourier(theta) x Fourier(z) x Chebyshev(r), one radial velocity-pressure
    block per (m, kz) mode.  State arrays are (Ntheta, Nz, Nr - 2) complex in
    FFT order along both Fourier axes; the library stores the kz >= 0 half."""

    def __init__(self, base: CircularCouette, nu: float = 1.0e-2, Nr: int = 24, Ntheta: int = 8, Nz: int = 16,
                 Lz: float | None = None, dt: float = 2.0e-3, family: str = "C", dealias: float = 1.5,
                 time_integrator: str = "CNAB2", device: int = 0):
        self.Ntheta = int(Ntheta)
        super().__init__(base, nu=nu, Nr=Nr, Nz=Nz, Lz=Lz, dt=dt, family=family, dealias=dealias,
                         time_integrator=time_integrator, device=device)
        th = np.arange(self.Ntheta) * 2.0 * math.pi / self.Ntheta
        z = np.arange(self.Nz) * self.Lz / self.Nz
        self.Theta, self.Z, self.R = np.meshgrid(th, z, self.r, indexing="ij")

    def _config(self, device: int) -> nat.TcConfig:
        c = super()._config(device)
        c.ntheta = self.Ntheta
        return c

    @property
    def _shape(self):
        return (self.Ntheta, self.Nz, self.n)

    def _forward(self, values) -> np.ndarray:
        c = np.fft.fft2(np.asarray(values, dtype=complex), axes=(0, 1), norm="forward")
        V = self._vander()
        phi = V[:, : self.n] - V[:, 2:]
        w = np.pi / self.Nr
        rhs = w * (c @ phi)
        M = w * phi.T @ phi
        return np.linalg.solve(M, rhs.reshape(-1, self.n).T).T.reshape(self._shape)

    def _backward(self, coef) -> np.ndarray:
        V = self._vander()
        phi = V[:, : self.n] - V[:, 2:]
        return np.fft.ifft2(np.asarray(coef) @ phi.T, axes=(0, 1), norm="forward")

    def _require_m(self, m: int) -> None:
        """``_require_resolved_m`` (``examples/taylor_couette_dns_jax.py:76-84``)."""
        if 2 * abs(int(m)) >= self.Ntheta:
            raise ValueError(f"azimuthal mode |m|={abs(int(m))} is unresolved by Ntheta={self.Ntheta}; "
                             "require 2|m| < Ntheta")

    def initial_state(self, amp: float = 1.0e-3, m: int = 1, kz_mode: int = 1) -> AxisymmetricTCState:
        """``:1264-1275``: amp sin(pi (r - R1) / d) cos(m theta) cos(kz z) in every component."""
        self._require_m(m)
        kz = 2.0 * math.pi * int(kz_mode) / self.Lz
        shape = np.sin(math.pi * (self.R - self.base.R1) / self.base.gap)
        f = amp * shape * np.cos(int(m) * self.Theta) * np.cos(kz * self.Z)
        return self.state_from_physical((f, f, f))

    def seed_linear_eigenmode(self, m: int = 1, kz_mode: int = 1, amp: float = 1.0e-6, which: int = 0):
        """``:1277-1308``: the pivot phase is fixed only for m = 0."""
        self._require_m(m)
        kz = 2.0 * math.pi * int(kz_mode) / self.Lz
        w, vecs = self.linear_eigs(int(m), kz, n_return=which + 1)
        vec = vecs[:, which]
        if int(m) == 0:
            vec = _positive_pivot_phase(vec)
        n = self.n
        mpos, mneg = int(m) % self.Ntheta, (-int(m)) % self.Ntheta
        kpos, kneg = int(kz_mode) % self.Nz, (-int(kz_mode)) % self.Nz
        comps = []
        for c in range(3):
            block = vec[c * n:(c + 1) * n] * amp
            arr = self._zeros()
            if mpos == mneg and kpos == kneg:
                arr[mpos, kpos, :n] = np.real(block)
            else:
                arr[mpos, kpos, :n] = 0.5 * block
                arr[mneg, kneg, :n] = 0.5 * np.conj(block)
            comps.append(arr)
        st = AxisymmetricTCState(u=tuple(comps), p=self._zeros(), nonlinear_old=tuple(self._zeros() for _ in range(3)),
                                 have_old=0.0)
        return st, complex(w[which])

    def velocity_physical(self, state: AxisymmetricTCState):
        """u on the standard (Ntheta, Nz, Nr) grid (``:1323-1324``), evaluated on the host."""
        return tuple(self._backward(c) for c in state.u)

    def inner_torque(self, state: AxisymmetricTCState) -> float:
        """The inner-cylinder torque of the (theta, z)-mean azimuthal velocity (the (0, 0) mode)."""
        r1, b = self.base.R1, self.base.b
        df = 2.0 / (self.base.R2 - self.base.R1)
        k = np.arange(self.Nr, dtype=float)
        dT = ((-1.0) ** (k + 1)) * k**2
        dphi = df * (dT[: self.n] - dT[2:])
        mean_dut = np.asarray(state.u[1])[0, 0] @ dphi
        return float(2.0 * math.pi * self.nu * r1**2 * abs(-2.0 * b / r1**2 + mean_dut))


@dataclass(frozen=True)
class AxisymmetricMRIState:
    """Coefficient state of the axisymmetric MHD step
    (``examples/taylor_couette_dns_jax.py:169-191``): x = (u_r, u_theta, u_z,
    b_r, b_theta, b_z), p the total pressure and the histories, each
    (Nz, Nr - 2) complex in the full-complex layout."""

    x: tuple
    p: np.ndarray
    nonlinear_old: tuple
    have_old: float = 0.0
    nonlinear_older: tuple | None = None
    solution_old: tuple | None = None
    solution_older: tuple | None = None
    history_steps: float = 0.0


def _wall_stencils(N: int, R1: float, R2: float):
    """Radial bases of b_theta and b_z as rows of Chebyshev coefficients
    (``src/jaxfun/galerkin/composite.py:797-871``): the Robin basis of
    d(r b_theta)/dr = 0 (psi + R df psi_x = 0 at each wall) and the Neumann
    basis T_i - i^2/(i+2)^2 T_{i+2}."""
    df = 2.0 / (R2 - R1)
    n = N - 2
    rob = np.zeros((n, N))
    neu = np.zeros((n, N))
    for i in range(n):
        lft = lambda j: (-1.0) ** j * (1.0 - R1 * df * j * j)  # noqa: E731
        rgt = lambda j: 1.0 + R2 * df * j * j  # noqa: E731
        d = np.linalg.solve([[lft(i + 1), lft(i + 2)], [rgt(i + 1), rgt(i + 2)]], [-lft(i), -rgt(i)])
        rob[i, i], rob[i, i + 1], rob[i, i + 2] = 1.0, d[0], d[1]
        neu[i, i], neu[i, i + 2] = 1.0, -(i * i) / (i + 2.0) ** 2
    return rob, neu


class AxisymmetricMRIDNSHIP(AxisymmetricTCDNSHIP):
    """``AxisymmetricMRIDNSJax`` on the GPU (``examples/taylor_couette_dns_jax.py:1359-1903``):
    conducting walls, the 7-field total-pressure system per axial mode."""

    def __init__(self, base: CircularCouette, B0: float = 0.1, nu: float = 1.0e-3, eta_mag: float = 1.0e-3,
                 Nr: int = 24, Nz: int = 16, Lz: float | None = None, dt: float = 2.0e-3, family: str = "C",
                 dealias: float = 1.5, time_integrator: str = "CNAB2", device: int = 0):
        self.B0 = float(B0)
        self.eta_mag = float(eta_mag)
        Lz = float(Lz) if Lz is not None else 2.0 * math.pi / 3.0 * base.gap
        super().__init__(base, nu=nu, Nr=Nr, Nz=Nz, Lz=Lz, dt=dt, family=family, dealias=dealias,
                         time_integrator=time_integrator, device=device)
        self.Rm = base.Omega1 * base.R1 * base.gap / self.eta_mag
        self.Pm = self.nu / self.eta_mag
        self.S = self.B0 * base.gap / self.eta_mag
        self.Jm = 0.5 * (base.R2 - base.R1)
        rob, neu = _wall_stencils(self.Nr, base.R1, base.R2)
        dirichlet = np.zeros((self.n, self.Nr))
        dirichlet[np.arange(self.n), np.arange(self.n)] = 1.0
        dirichlet[np.arange(self.n), np.arange(self.n) + 2] = -1.0
        self._stencils = (dirichlet,) * 4 + (rob, neu)

    def _config(self, device: int) -> nat.TcConfig:
        c = super()._config(device)
        c.mhd = 1
        c.b0 = self.B0
        c.eta_mag = self.eta_mag
        return c

    # ---------------------------------------------------------------- states
    def zero_state(self) -> AxisymmetricMRIState:
        x = tuple(self._zeros() for _ in range(6))
        return AxisymmetricMRIState(x=x, p=self._zeros(), nonlinear_old=tuple(self._zeros() for _ in range(6)))

    def _forward_basis(self, values, S) -> np.ndarray:
        c = np.fft.fft(np.asarray(values, dtype=complex), axis=0, norm="forward")
        V = self._vander()
        psi = V @ S.T
        w = np.pi / self.Nr
        return np.linalg.solve(w * psi.T @ psi, (w * (c @ psi)).T).T

    def _backward_basis(self, coef, S) -> np.ndarray:
        return np.fft.ifft(np.asarray(coef) @ (self._vander() @ S.T).T, axis=0, norm="forward")

    def state_from_physical(self, values) -> AxisymmetricMRIState:
        x = tuple(self._forward_basis(v, S) for v, S in zip(values, self._stencils))
        return AxisymmetricMRIState(x=x, p=self._zeros(), nonlinear_old=tuple(self._zeros() for _ in range(6)))

    def mri_pencil(self, m: int, kz: float):
        """(L, M) of ``TaylorCouetteMRIJax.assemble`` (conducting) on this radial space."""
        n7 = 7 * self.n
        L = np.empty((n7, n7), dtype=complex)
        M = np.empty((n7, n7), dtype=complex)
        nat.tc_check(self._lib.tc_mri_pencil(C.byref(self._cfg), int(m), float(kz),
                                             L.view(np.float64).ctypes.data_as(_dp),
                                             M.view(np.float64).ctypes.data_as(_dp)))
        return L, M

    def linear_eigs(self, m: int, kz: float, n_return: int = 8, finite_cap: float = 1.0e6):
        w, V = sla.eig(*self.mri_pencil(m, kz))
        keep = np.isfinite(w) & (np.abs(w) < finite_cap)
        w, V = w[keep], V[:, keep]
        order = np.lexsort((-w.imag, -w.real))
        return w[order][:n_return], V[:, order][:, :n_return]

    def seed_linear_eigenmode(self, kz_mode: int = 1, amp: float = 1.0e-6, which: int = 0):
        """``:1809-1842``: the conducting MRI eigenmode, fields from blocks (0, 1, 2, 4, 5, 6)."""
        kz = 2.0 * math.pi * int(kz_mode) / self.Lz
        w, vecs = self.linear_eigs(0, kz, n_return=which + 1)
        vec = _positive_pivot_phase(vecs[:, which])
        n = self.n
        kpos, kneg = int(kz_mode) % self.Nz, (-int(kz_mode)) % self.Nz
        comps = []
        for blk in (0, 1, 2, 4, 5, 6):
            block = vec[blk * n:(blk + 1) * n] * amp
            arr = self._zeros()
            if kpos == kneg:
                arr[kpos, :n] = np.real(block)
            else:
                arr[kpos, :n] = 0.5 * block
                arr[kneg, :n] = 0.5 * np.conj(block)
            comps.append(arr)
        st = AxisymmetricMRIState(x=tuple(comps), p=self._zeros(), nonlinear_old=tuple(self._zeros() for _ in range(6)))
        return st, complex(w[which])

    # ------------------------------------------------------------ device I/O
    def _set_c(self, group: str, comp: int, arr):
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.complex128))
        if a.shape != self._shape:
            raise ValueError(f"{group}[{comp}]: expected shape {self._shape}, got {a.shape}")
        nat.tc_check(self._lib.tc_set_component(self._h, nat.TC_GROUP[group], comp,
                                                a.view(np.float64).ctypes.data_as(_dp)))

    def _get_c(self, group: str, comp: int) -> np.ndarray:
        a = np.empty(self._shape, dtype=np.complex128)
        nat.tc_check(self._lib.tc_get_component(self._h, nat.TC_GROUP[group], comp,
                                                a.view(np.float64).ctypes.data_as(_dp)))
        return a

    @staticmethod
    def _ensure_history(s: AxisymmetricMRIState) -> AxisymmetricMRIState:
        """``_ensure_mhd_history`` (``:1705-1718``)."""
        z = tuple(np.zeros_like(v) for v in s.x)
        return replace(s, nonlinear_older=z if s.nonlinear_older is None else s.nonlinear_older,
                       solution_old=s.x if s.solution_old is None else s.solution_old,
                       solution_older=s.x if s.solution_older is None else s.solution_older)

    def load(self, s: AxisymmetricMRIState) -> None:
        if self.time_integrator == "SBDF3":
            s = self._ensure_history(s)
        for c in range(6):
            self._set_c("X", c, s.x[c])
        self._set_c("P", 0, s.p)
        for group, grp in (("NL_OLD", s.nonlinear_old), ("NL_OLDER", s.nonlinear_older),
                           ("SOL_OLD", s.solution_old), ("SOL_OLDER", s.solution_older)):
            if grp is not None:
                for c in range(6):
                    self._set_c(group, c, grp[c])
        nat.tc_check(self._lib.tc_set_scalars(self._h, float(s.have_old), float(s.history_steps)))

    def fetch(self, like: AxisymmetricMRIState) -> AxisymmetricMRIState:
        grp = lambda g: tuple(self._get_c(g, c) for c in range(6))  # noqa: E731
        ho, hs = C.c_double(), C.c_double()
        nat.tc_check(self._lib.tc_get_scalars(self._h, C.byref(ho), C.byref(hs)))
        x, p = grp("X"), self._get_c("P", 0)
        if self.time_integrator == "SBDF3":
            return AxisymmetricMRIState(x=x, p=p, nonlinear_old=grp("NL_OLD"), have_old=ho.value,
                                        nonlinear_older=grp("NL_OLDER"), solution_old=grp("SOL_OLD"),
                                        solution_older=grp("SOL_OLDER"), history_steps=hs.value)
        if self.time_integrator == "CNAB2":
            return replace(like, x=x, p=p, nonlinear_old=grp("NL_OLD"), have_old=ho.value)
        return replace(like, x=x, p=p)

    # ------------------------------------------------------------ diagnostics
    def _device_diagnostics(self, state: AxisymmetricMRIState | None) -> np.ndarray:
        if state is not None:
            self.load(state)
        out = np.zeros(len(nat.TC_MHD_DIAG_KEYS))
        nat.tc_check(self._lib.tc_diagnostics(self._h, _ptr(out)))
        return out

    def diagnostics(self, state: AxisymmetricMRIState) -> dict:
        """Ekin, Emag, E, divu, divb, wall_u (``:1876-1886``), device reductions."""
        d = self._device_diagnostics(state)
        return dict(zip(nat.TC_MHD_DIAG_KEYS, (float(v) for v in d)))

    def fields_physical(self, state: AxisymmetricMRIState):
        """The six fields on the standard (Nz, Nr) grid (``:1844-1848``), on the host."""
        return tuple(self._backward_basis(c, S) for c, S in zip(state.x, self._stencils))

    def energy_parts(self, state: AxisymmetricMRIState):
        d = self.diagnostics(state)
        return d["Ekin"], d["Emag"]

    def energy(self, state: AxisymmetricMRIState) -> float:
        return self.diagnostics(state)["E"]

    def divergence_linf(self, state: AxisymmetricMRIState):
        d = self.diagnostics(state)
        return d["divu"], d["divb"]

    def velocity_physical(self, state: AxisymmetricMRIState):
        return self.fields_physical(state)[:3]

    def inner_torque(self, state):
        raise NotImplementedError("the inner-cylinder torque is a hydrodynamic runner scalar")


class TaylorCouetteMRIDNSHIP(AxisymmetricMRIDNSHIP):
    """``TaylorCouetteMRIDNSJax`` on the GPU (``examples/taylor_couette_dns_jax.py:1906-2245``):
    the conducting-wall MHD step on Fourier(theta) x Fourier(z) x Chebyshev(r),
    one 7-field block per (m, kz).  State arrays are (Ntheta, Nz, Nr - 2)."""

    def __init__(self, base: CircularCouette, B0: float = 0.1, nu: float = 1.0e-3, eta_mag: float = 1.0e-3,
                 Nr: int = 24, Ntheta: int = 8, Nz: int = 16, Lz: float | None = None, dt: float = 2.0e-3,
                 family: str = "C", dealias: float = 1.5, time_integrator: str = "CNAB2", device: int = 0):
        self.Ntheta = int(Ntheta)
        super().__init__(base, B0=B0, nu=nu, eta_mag=eta_mag, Nr=Nr, Nz=Nz, Lz=Lz, dt=dt, family=family,
                         dealias=dealias, time_integrator=time_integrator, device=device)
        th = np.arange(self.Ntheta) * 2.0 * math.pi / self.Ntheta
        z = np.arange(self.Nz) * self.Lz / self.Nz
        self.Theta, self.Z, self.R = np.meshgrid(th, z, self.r, indexing="ij")

    def _config(self, device: int) -> nat.TcConfig:
        c = super()._config(device)
        c.ntheta = self.Ntheta
        return c

    @property
    def _shape(self):
        return (self.Ntheta, self.Nz, self.n)

    def _forward_basis(self, values, S) -> np.ndarray:
        c = np.fft.fft2(np.asarray(values, dtype=complex), axes=(0, 1), norm="forward")
        psi = self._vander() @ S.T
        w = np.pi / self.Nr
        rhs = w * (c @ psi)
        return np.linalg.solve(w * psi.T @ psi, rhs.reshape(-1, self.n).T).T.reshape(self._shape)

    def _backward_basis(self, coef, S) -> np.ndarray:
        return np.fft.ifft2(np.asarray(coef) @ (self._vander() @ S.T).T, axes=(0, 1), norm="forward")

    def seed_linear_eigenmode(self, m: int = 1, kz_mode: int = 1, amp: float = 1.0e-6, which: int = 0):
        """``:2192-2230``: the pivot phase is fixed only for m = 0."""
        if 2 * abs(int(m)) >= self.Ntheta:
            raise ValueError(f"azimuthal mode |m|={abs(int(m))} is unresolved by Ntheta={self.Ntheta}; "
                             "require 2|m| < Ntheta")
        kz = 2.0 * math.pi * int(kz_mode) / self.Lz
        w, vecs = self.linear_eigs(int(m), kz, n_return=which + 1)
        vec = vecs[:, which]
        if int(m) == 0:
            vec = _positive_pivot_phase(vec)
        n = self.n
        mpos, mneg = int(m) % self.Ntheta, (-int(m)) % self.Ntheta
        kpos, kneg = int(kz_mode) % self.Nz, (-int(kz_mode)) % self.Nz
        comps = []
        for blk in (0, 1, 2, 4, 5, 6):
            block = vec[blk * n:(blk + 1) * n] * amp
            arr = self._zeros()
            if mpos == mneg and kpos == kneg:
                arr[mpos, kpos, :n] = np.real(block)
            else:
                arr[mpos, kpos, :n] = 0.5 * block
                arr[mneg, kneg, :n] = 0.5 * np.conj(block)
            comps.append(arr)
        st = AxisymmetricMRIState(x=tuple(comps), p=self._zeros(), nonlinear_old=tuple(self._zeros() for _ in range(6)))
        return st, complex(w[which])
