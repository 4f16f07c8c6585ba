"""tinyvlm_amd: MI355X (gfx950) HIP backend for the tinyvlm vortex-lattice solver.

The compute path is the C-ABI library ``libtinyvlm_hip.so`` (sources in ``csrc/``, interface in
``include/tinyvlm_hip.h``).  ``capi`` binds it for Python callers; ``build`` compiles it.
"""
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT_DIR = os.path.dirname(PKG_DIR)

from . import capi  # noqa: E402,F401
