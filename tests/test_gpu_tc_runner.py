"""The Taylor-Couette DNS golden problems through ``runner.run_taylor_couette_dns``
(``production/oracles.py:1237-1330,2942-3039``) on the GPU, against the
published scalars of ``production/goldens/taylor_couette_{hydro_dns,hydro_3d,
mhd_dns,mhd_3d}_v1``.  The specs below restate those goldens' problem data.
The shenfun-written goldens (``hydro_dns``, ``mhd_dns``) carry the whole
eigenvector once where the reference's port writes half of it at +k and at
-k, so their energies are 4x the runner's; the 3D ones were written by the
port itself."""

from __future__ import annotations

import pytest

pytestmark = pytest.mark.gpu

LZ_HYDRO, LZ_MHD, KEPLER = 2.0074074463832545, 1.0471975511965976, 0.3535533905932738


def _spec(physics, groups, lz, resolution, time, amp, mode):
    return {"physics": physics, "nondimensional_groups": groups, "domain": {"z_period": lz},
            "resolution": resolution, "time": time, "initial_condition": {"amplitude": amp, "type": "linear_eigenmode"},
            "mode": mode, "boundary_conditions": {"magnetic": {"type": "conducting"} if physics == "mhd" else None}}


HYDRO = {"Omega1": 1.0, "Omega2": 0.0, "R1": 1.0, "R2": 2.0, "Re": 100.0, "nu": 0.01}
MHD = {"B0": 0.1, "Omega1": 1.0, "Omega2": KEPLER, "R1": 1.0, "R2": 2.0, "Re": 1000.0, "Rm": 1000.0,
       "eta_mag": 0.001, "nu": 0.001}

CASES = [
    ("hydro_dns", _spec("hydro", HYDRO, LZ_HYDRO, {"Nr": 40, "Nz": 8, "family": "C"},
                        {"dt": 1e-3, "final_time": 0.4, "integrator": "CNAB2"}, 1e-6,
                        {"axial_wavenumber": 3.13, "azimuthal_wavenumber": 0}),
     4.0, {"kinetic_energy": 5.4823752189594825e-12, "growth_rate": 0.1273035583309162,
           "growth_rate_linear": 0.12730355815948557}),
    ("hydro_3d", _spec("hydro", HYDRO, LZ_HYDRO, {"Nr": 12, "Ntheta": 4, "Nz": 4, "dealias": 1.5, "family": "C"},
                       {"dt": 1e-3, "final_time": 0.003, "integrator": "SBDF3"}, 1e-6,
                       {"axial_wavenumber": 3.13, "azimuthal_wavenumber": 1}),
     1.0, {"kinetic_energy": 7.955617099356648e-12, "growth_rate": 0.08387873243632032,
           "growth_rate_linear": 0.08387873243612082}),
    ("mhd_dns", _spec("mhd", MHD, LZ_MHD, {"Nr": 40, "Nz": 8, "family": "C"},
                      {"dt": 2e-3, "final_time": 0.5, "integrator": "CNAB2"}, 1e-7,
                      {"axial_wavenumber": 6.0, "azimuthal_wavenumber": 0}),
     4.0, {"kinetic_energy": 7.339824415880832e-15, "magnetic_energy": 1.039589006705832e-14,
           "growth_rate": 0.3403972182913979, "growth_rate_linear": 0.3403972051440628}),
    ("mhd_3d", _spec("mhd", MHD, LZ_MHD, {"Nr": 12, "Ntheta": 4, "Nz": 4, "dealias": 1.5, "family": "C"},
                     {"dt": 2e-3, "final_time": 0.006, "integrator": "SBDF3"}, 1e-7,
                     {"axial_wavenumber": 6.0, "azimuthal_wavenumber": 1}),
     1.0, {"kinetic_energy": 7.883607407650011e-15, "magnetic_energy": 1.1337572869748595e-14,
           "growth_rate": 0.29129374483210074, "growth_rate_linear": 0.2912937448361517}),
]


@pytest.mark.parametrize("name,spec,efac,golden", CASES, ids=[c[0] for c in CASES])
def test_runner_reproduces_golden_scalars(name, spec, efac, golden):
    from jaxfun_amd.runner import run_taylor_couette_dns

    out = run_taylor_couette_dns(spec)
    s = out["scalars"]
    print(name, s, flush=True)
    for key in ("kinetic_energy", "magnetic_energy"):
        if key in golden:
            assert abs(efac * s[key] / golden[key] - 1.0) <= 1e-9, (key, s[key])
    assert abs(s["growth_rate"] - golden["growth_rate"]) <= 1e-7
    assert abs(s["growth_rate_linear"] - golden["growth_rate_linear"]) <= 1e-10
    assert out["time_series"][1]["t"] == pytest.approx(spec["time"]["final_time"])
