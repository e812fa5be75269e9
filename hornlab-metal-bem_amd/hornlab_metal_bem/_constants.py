"""Physical constants of the solver (reference: hornlab_metal_bem/_constants.py:3-5)."""
from __future__ import annotations

SPEED_OF_SOUND: float = 343.0  # m/s
AIR_DENSITY: float = 1.2041  # kg/m^3
REFERENCE_PRESSURE: float = 20e-6  # Pa
