"""MI355X (gfx950) backend for dmrgpy's two-site DMRG / TDVP hot path.

* ``dmrgpy_amd.chain.Chain``       -- the session object (drop-in for the
  reference's ``pyitensor.chain.Chain``)
* ``dmrgpy_amd.frontend``          -- ``Spin_Chain`` / ``Fermionic_Chain`` /
  ``Spinful_Fermionic_Chain_Native`` with ``itensor_version="mi355x"``
* ``dmrgpy_amd.engine.Engine``     -- the device engine over libdmrg_hip.so
* ``dmrgpy_amd.build``             -- builds libdmrg_hip.so with hipcc

Importing the package does not touch the GPU; the engine is created on first
use and raises if the native library or the GPU is missing.
"""

__version__ = "0.1.0"
BACKEND = "mi355x"


def __getattr__(name):
    if name in ("Spin_Chain", "Fermionic_Chain", "Spinful_Fermionic_Chain_Native", "Many_Body_Chain"):
        from . import frontend
        return getattr(frontend, name)
    if name == "Chain":
        from .chain import Chain
        return Chain
    raise AttributeError(name)
