"""ctypes binding of libpcfdns.so (C ABI declared in include/pcf_dns.h).

The library is built in-tree by ``jaxfun_amd.build`` and loaded from the
package directory.  There is no fallback: if the shared object is missing or
fails to load, every entry point raises.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libpcfdns.so"

ABI_VERSION = 2
WALLS_CONDUCTING = 0
WALLS_INSULATING = 1
PREC_FP64 = 0
PREC_MIXED = 1
TI_IMEXRK3 = 1
TI_IMEXRK222 = 2
TI_SBDF3 = 3
TI_CNAB2 = 4
TIME_INTEGRATORS = {"IMEXRK3": TI_IMEXRK3, "IMEXRK222": TI_IMEXRK222, "SBDF3": TI_SBDF3, "CNAB2": TI_CNAB2}
NL_ROTATIONAL = 1
XDIAG_COUNT = 38
# output order of pcf_mhd_extended_diagnostics
XDIAG_KEYS = (
    "reynolds_stress", "maxwell_stress", "Emag_total", "mean_bx", "mean_by", "mean_bz",
    "mag_energy_fluct_total", "dissipation_kinetic", "dissipation_magnetic",
    "channel_kz1_velocity_rms", "channel_kz1_magnetic_rms", "bmax", "bmax_total",
    "electric_ideal_l2", "electric_resistive_l2", "electric_total_l2",
    "divergence_e_l2", "divergence_e_ideal_l2", "divergence_e_resistive_l2", "divergence_a_l2",
    "divergence_a_wall_linf", "electric_ideal_wall_tangential_linf",
    "electric_resistive_wall_tangential_linf", "electric_wall_tangential_linf",
    "faraday_mean_by_tendency", "faraday_mean_bz_tendency",
    "mean_bx_trace", "mean_by_trace", "mean_bz_trace",
    "mean_ex_ideal", "mean_ey_ideal", "mean_ez_ideal",
    "mean_ex_resistive", "mean_ey_resistive", "mean_ez_resistive",
    "mean_ex_total", "mean_ey_total", "mean_ez_total",
)

# enum pcf_field
FIELDS = (
    "U0", "U1", "U2", "G",
    "NL_OLD_U", "NL_OLD_G", "NL_OLD_V", "NL_OLD_W",
    "NL_OLDER_U", "NL_OLDER_G", "NL_OLDER_V", "NL_OLDER_W",
    "SOL_OLD_U", "SOL_OLD_G", "SOL_OLD_V", "SOL_OLD_W",
    "SOL_OLDER_U", "SOL_OLDER_G", "SOL_OLDER_V", "SOL_OLDER_W",
    "AX", "AY", "AZ",
    "NL_A_OLD_X", "NL_A_OLD_Y", "NL_A_OLD_Z",
    "NL_A_OLDER_X", "NL_A_OLDER_Y", "NL_A_OLDER_Z",
    "A_OLD_X", "A_OLD_Y", "A_OLD_Z",
    "A_OLDER_X", "A_OLDER_Y", "A_OLDER_Z",
)
FIELD_ID = {name: i for i, name in enumerate(FIELDS)}

# pcf_setup_matrix indices
MATRIX_ID = {
    "C0B": 0, "C2B": 1, "C4B": 2, "C0D": 3, "C2D": 4, "VB": 5, "VD": 6, "VD1": 7,
    "A1": 8, "A0": 9, "AG": 10, "R": 11, "VB1": 12, "FB": 13, "FD": 14, "VQ1": 15, "VQ2": 16,
    "VC": 17, "VC1": 18, "VC2": 19, "WQ": 20, "VDR": 21, "VD1R": 22,
}

# Every symbol include/pcf_dns.h declares.
EXPORTS = (
    "pcf_abi_version", "pcf_last_error", "pcf_field_size", "pcf_setup_matrix",
    "pcf_create", "pcf_destroy", "pcf_set_stream", "pcf_synchronize",
    "pcf_set_field", "pcf_get_field", "pcf_set_field_device", "pcf_get_field_device", "pcf_set_history",
    "pcf_get_history",
    "pcf_step", "pcf_set_dt", "pcf_step_steady", "pcf_step_steady_graph", "pcf_graph_cache_info", "pcf_diagnostics", "pcf_mhd_diagnostics",
    "pcf_mhd_extended_diagnostics", "pcf_field_maxima", "pcf_field_maxima_noalloc", "pcf_velocity_energies", "pcf_physical_fields",
    "pcf_kernels_per_step", "pcf_stage_count", "pcf_stage_name", "pcf_profile_steady", "pcf_profile_stage",
    "pcf_profile_reverse",
    "pcf_set_cotangent", "pcf_vjp", "pcf_get_gradient",
)
SEED_ENERGY = 0
SEED_COTANGENT = 1

# include/tc_dns.h (the Taylor-Couette step in the same library)
TC_ABI_VERSION = 3
TC_TI = {"CNAB2": 0, "SBDF3": 1, "IMEXRK3": 2}
TC_FIELDS = ("UR", "UT", "UZ", "P", "NL_OLD_R", "NL_OLD_T", "NL_OLD_Z", "NL_OLDER_R", "NL_OLDER_T", "NL_OLDER_Z",
             "SOL_OLD_R", "SOL_OLD_T", "SOL_OLD_Z", "SOL_OLDER_R", "SOL_OLDER_T", "SOL_OLDER_Z")
TC_FIELD_ID = {name: i for i, name in enumerate(TC_FIELDS)}
TC_DIAG_KEYS = ("E", "div_linf", "continuity_l2", "wall", "Eth")
TC_MHD_DIAG_KEYS = ("Ekin", "Emag", "E", "divu", "divb", "wall_u")
TC_GROUP = {"X": 0, "P": 1, "NL_OLD": 2, "NL_OLDER": 3, "SOL_OLD": 4, "SOL_OLDER": 5}
TC_EXPORTS = (
    "tc_last_error", "tc_create", "tc_destroy", "tc_set_field", "tc_get_field", "tc_set_scalars", "tc_get_scalars",
    "tc_set_component", "tc_get_component", "tc_step", "tc_step_graph", "tc_diagnostics", "tc_linear_pencil",
    "tc_mri_pencil", "tc_set_dt",
)


class TcConfig(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32),
        ("nr", C.c_int32),
        ("nz", C.c_int32),
        ("r1", C.c_double),
        ("r2", C.c_double),
        ("omega1", C.c_double),
        ("omega2", C.c_double),
        ("nu", C.c_double),
        ("lz", C.c_double),
        ("dt", C.c_double),
        ("dealias", C.c_double),
        ("time_integrator", C.c_int32),
        ("device", C.c_int32),
        ("ntheta", C.c_int32),
        ("mhd", C.c_int32),
        ("b0", C.c_double),
        ("eta_mag", C.c_double),
    ]


class PcfConfig(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32),
        ("n0", C.c_int32),
        ("ny", C.c_int32),
        ("nz", C.c_int32),
        ("lx", C.c_double * 2),
        ("ly", C.c_double),
        ("lz", C.c_double),
        ("nu", C.c_double),
        ("u_wall", C.c_double),
        ("dt", C.c_double),
        ("padding_y", C.c_double),
        ("padding_z", C.c_double),
        ("precision", C.c_int32),
        ("time_integrator", C.c_int32),
        ("nonlinear_form", C.c_int32),
        ("device", C.c_int32),
        ("mhd", C.c_int32),
        ("walls", C.c_int32),
        ("eta", C.c_double),
        ("b0", C.c_double * 3),
        ("omega", C.c_double),
    ]


class PcfError(RuntimeError):
    """Raised when a library call returns a negative status."""


_lib = None


def load() -> C.CDLL:
    """Load the shared library (once); raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise PcfError(
            f"{LIB_PATH} is not built; run `python -m jaxfun_amd.build` (hipcc, gfx950)"
        )
    # torch ships its own HIP and HSA runtime libraries under other file names.
    # Loaded before this library, its HSA runtime is the one both HIP runtimes
    # bind to, and device pointers pass between torch and the library; loaded
    # after, it brings up a second HSA runtime that finds no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(str(LIB_PATH))
    P = C.POINTER
    dp = P(C.c_double)
    cfgp = P(PcfConfig)
    h = C.c_void_p
    sig = {
        "pcf_abi_version": ([], C.c_int32),
        "pcf_last_error": ([], C.c_char_p),
        "pcf_field_size": ([cfgp, C.c_int32], C.c_int64),
        "pcf_setup_matrix": ([cfgp, C.c_int32, dp, P(C.c_int32), P(C.c_int32)], C.c_int32),
        "pcf_create": ([cfgp, P(C.c_void_p)], C.c_int32),
        "pcf_destroy": ([h], C.c_int32),
        "pcf_set_stream": ([h, C.c_void_p], C.c_int32),
        "pcf_synchronize": ([h], C.c_int32),
        "pcf_set_field": ([h, C.c_int32, dp], C.c_int32),
        "pcf_get_field": ([h, C.c_int32, dp], C.c_int32),
        "pcf_set_field_device": ([h, C.c_int32, C.c_void_p], C.c_int32),
        "pcf_get_field_device": ([h, C.c_int32, C.c_void_p], C.c_int32),
        "pcf_set_history": ([h, C.c_double, C.c_double, C.c_double], C.c_int32),
        "pcf_get_history": ([h, dp, dp, dp], C.c_int32),
        "pcf_step": ([h, C.c_int32], C.c_int32),
        "pcf_set_dt": ([h, C.c_double], C.c_int32),
        "pcf_step_steady": ([h, C.c_int32], C.c_int32),
        "pcf_step_steady_graph": ([h, C.c_int32, C.c_int32], C.c_int32),
        "pcf_graph_cache_info": ([h, P(C.c_int64), C.c_int32], C.c_int32),
        "pcf_diagnostics": ([h, dp], C.c_int32),
        "pcf_mhd_diagnostics": ([h, dp], C.c_int32),
        "pcf_mhd_extended_diagnostics": ([h, dp], C.c_int32),
        "pcf_field_maxima": ([h, dp], C.c_int32),
        "pcf_field_maxima_noalloc": ([h, dp], C.c_int32),
        "pcf_velocity_energies": ([h, dp], C.c_int32),
        "pcf_physical_fields": ([h, dp], C.c_int32),
        "pcf_kernels_per_step": ([h], C.c_int32),
        "pcf_stage_count": ([], C.c_int32),
        "pcf_stage_name": ([C.c_int32], C.c_char_p),
        "pcf_profile_steady": ([h, C.c_int32, dp], C.c_int32),
        "pcf_profile_stage": ([h, C.c_int32, C.c_int32, dp], C.c_int32),
        "pcf_profile_reverse": ([h, C.c_int32, dp], C.c_int32),
        "pcf_set_cotangent": ([h, C.c_int32, dp], C.c_int32),
        "pcf_vjp": ([h, C.c_int32, C.c_int32, C.c_int32], C.c_int32),
        "pcf_get_gradient": ([h, C.c_int32, dp], C.c_int32),
    }
    tcp = P(TcConfig)
    sig.update({
        "tc_last_error": ([], C.c_char_p),
        "tc_create": ([tcp, P(C.c_void_p)], C.c_int32),
        "tc_destroy": ([h], C.c_int32),
        "tc_set_field": ([h, C.c_int32, dp], C.c_int32),
        "tc_get_field": ([h, C.c_int32, dp], C.c_int32),
        "tc_set_scalars": ([h, C.c_double, C.c_double], C.c_int32),
        "tc_get_scalars": ([h, dp, dp], C.c_int32),
        "tc_step": ([h, C.c_int32], C.c_int32),
        "tc_step_graph": ([h, C.c_int32], C.c_int32),
        "tc_diagnostics": ([h, dp], C.c_int32),
        "tc_linear_pencil": ([tcp, C.c_int32, C.c_double, dp, dp], C.c_int32),
        "tc_mri_pencil": ([tcp, C.c_int32, C.c_double, dp, dp], C.c_int32),
        "tc_set_component": ([h, C.c_int32, C.c_int32, dp], C.c_int32),
        "tc_get_component": ([h, C.c_int32, C.c_int32, dp], C.c_int32),
        "tc_set_dt": ([h, C.c_double], C.c_int32),
    })
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if lib.pcf_abi_version() != ABI_VERSION:
        raise PcfError("libpcfdns ABI version mismatch; rebuild the library")
    _lib = lib
    return lib


def check(status: int) -> int:
    if status < 0:
        msg = load().pcf_last_error().decode(errors="replace")
        if status in (-3, -4):  # option not built / SBDF3's fixed dt
            raise NotImplementedError(msg)
        if status == -1:
            raise ValueError(msg)
        raise PcfError(f"libpcfdns error {status}: {msg}")
    return status


def setup_matrix(cfg: PcfConfig, name: str) -> np.ndarray:
    """Fetch a closed-form setup matrix (host only, no GPU)."""
    lib = load()
    r, c = C.c_int32(), C.c_int32()
    check(lib.pcf_setup_matrix(C.byref(cfg), MATRIX_ID[name], None, C.byref(r), C.byref(c)))
    out = np.empty((r.value, c.value), dtype=np.float64)
    check(lib.pcf_setup_matrix(C.byref(cfg), MATRIX_ID[name],
                               out.ctypes.data_as(C.POINTER(C.c_double)), None, None))
    return out


def tc_check(status: int) -> int:
    """Status of a tc_* call: the reference's exception types, as check()."""
    if status < 0:
        msg = load().tc_last_error().decode(errors="replace")
        if status in (-3, -4):
            raise NotImplementedError(msg)
        if status == -1:
            raise ValueError(msg)
        raise PcfError(f"Taylor-Couette library error {status}: {msg}")
    return status
