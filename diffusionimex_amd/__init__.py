"""diffusionimex_amd: MI355X (gfx950) HIP implementation of the SPB 6+1-equation semi-implicit IMEX-RK time step.

Drop-in for the stepper path of diffusionIMEX's SPB6_PC_IMEX solver; the C ABI is include/spb6_imex.h and the
Python binding is diffusionimex_amd.spb6.
"""
from .spb6 import Stepper, Config, make_config, load, SPB6Error, STATUS  # noqa: F401
