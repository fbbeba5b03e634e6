"""MI355X-native plane-Couette KMM SBDF3 DNS step (HIP kernels behind a C ABI)."""

from .mhd import (
    MHDState,
    PlaneCouetteMHDHIP,
    PlaneCouetteMRIShearpyHIP,
    PlaneCouetteMRIShearpyInsulatingHIP,
    mhd_state_checksum,
    mhd_state_leaves,
)
from .pcf import KMMState, PlaneCouetteFluctuationHIP, state_checksum, state_leaves

__all__ = ["KMMState", "MHDState", "PlaneCouetteFluctuationHIP", "PlaneCouetteMHDHIP", "PlaneCouetteMRIShearpyHIP",
           "PlaneCouetteMRIShearpyInsulatingHIP", "mhd_state_checksum", "mhd_state_leaves", "state_checksum",
           "state_leaves"]
