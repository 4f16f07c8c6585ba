"""Drop-in for the reference's pybind11 module ``libhbvlm3`` (lib/libhbvlm3.cpp:554-642).

With this directory on ``sys.path`` (for instance ``PYTHONPATH=<repo>/tinyvlm_amd/pymod``), a script's
``from libhbvlm3 import HBVLM`` (scripts/dof3_hbvlm.py:23) resolves to the HIP harmonic-balance
driver, with the same constructor ``HBVLM(backend, [wing_mesh, flap_mesh])``, ``init(harmonics,
scaling)`` and ``run(omega, U, dyn_f, force_t)`` (force_t mutated in place).  The backend name this
build provides is "hip"; any other name raises RuntimeError("Unsupported backend: ..."), as the
reference's create_backend does for a backend that was not compiled in (vlm/src/vlm_backend.cpp:15-27).
"""
import os
import sys

_ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if _ROOT not in sys.path:
    sys.path.insert(0, _ROOT)

from tinyvlm_amd.hb import HBVLM  # noqa: E402

__all__ = ["HBVLM"]
