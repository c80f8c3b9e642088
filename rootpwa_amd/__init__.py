"""rootpwa_amd -- MI355X (gfx950) backend of ROOTPWA's unbinned extended log-likelihood.

The compute path is hand-written HIP in rootpwa_amd/csrc, built into
rootpwa_amd/librootpwa_amd.so and reached through the C ABI of
include/rootpwa_amd.h (low level) and include/rootpwa_amd_pwalik.h (the
pwaLikelihood host class).  There is no CPU fallback: without the shared
library the import of the bindings raises.
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
