"""MI355X-native plane-Couette KMM SBDF3 DNS step (HIP kernels behind a C ABI)."""

from .pcf import KMMState, PlaneCouetteFluctuationHIP, state_checksum, state_leaves

__all__ = ["KMMState", "PlaneCouetteFluctuationHIP", "state_checksum", "state_leaves"]
