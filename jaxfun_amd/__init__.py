"""MI355X-native plane-Couette KMM SBDF3 DNS step and axisymmetric Taylor-Couette step (HIP kernels behind a C ABI)."""

from .mhd import (
    MHDState,
    PlaneCouetteMHDHIP,
    PlaneCouetteMRIShearpyHIP,
    PlaneCouetteMRIShearpyInsulatingHIP,
    mhd_state_checksum,
    mhd_state_leaves,
)
from .pcf import KMMState, PlaneCouetteFluctuationHIP, state_checksum, state_leaves
from .tc import AxisymmetricTCDNSHIP, AxisymmetricTCState, CircularCouette

__all__ = ["AxisymmetricTCDNSHIP", "AxisymmetricTCState", "CircularCouette", "KMMState", "MHDState", "PlaneCouetteFluctuationHIP", "PlaneCouetteMHDHIP", "PlaneCouetteMRIShearpyHIP",
           "PlaneCouetteMRIShearpyInsulatingHIP", "mhd_state_checksum", "mhd_state_leaves", "state_checksum",
           "state_leaves"]
