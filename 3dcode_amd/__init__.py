"""MI355X-native pseudo-time step of the 3dcode cavitating-flow solver.

The product is libcav3d.so (HIP for gfx950, C-ABI in include/cav3d.h);
``solver`` binds it with ctypes and ``cases`` builds the synthetic bump
channel cases in the reference's array layouts.
"""
from . import solver, cases  # noqa: F401
from .solver import Solver, Cav3dError, lib, eos_host  # noqa: F401
