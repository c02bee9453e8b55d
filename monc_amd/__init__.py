"""MI355X-native TVD advection for MONC: gfx950 HIP kernels behind a C ABI.

`monc_amd.tvd.TvdAdvection` mirrors the tvd_advection component's callback
surface over libmonc_tvd.so; `monc_amd.setup_state` builds synthetic states.
"""
from . import setup_state  # noqa: F401

__all__ = ["setup_state", "tvd"]
