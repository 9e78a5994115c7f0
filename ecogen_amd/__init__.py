"""MI355X-native explicit finite-volume step of ECOGEN (Cartesian meshes).

Host interface over the gfx950 shared library ``libecogen_amd.so`` (C ABI of
``include/ecogen.h``).  The numerical work runs only in the HIP kernels; this
package has no CPU fallback.
"""
from .capi import Context, EcoError, EcoStatus, lib, make_config, nfields  # noqa: F401

__all__ = ["Context", "EcoError", "EcoStatus", "lib", "make_config", "nfields"]
