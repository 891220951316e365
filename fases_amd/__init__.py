"""MI355X-native FASES 3D-solver time-stepping core (WENO5-Z / Godunov / SSP-RK3 FVM
with cPML and ADE impedance boundaries) behind a C ABI (include/fases_core.h)."""
from . import _lib  # noqa: F401

__all__ = ["_lib"]
