"""fundem_amd: MI355X-native level-set DEM timestep (FunDEM-compatible).

The compute path is the gfx950 HIP library (fundem_amd/lib/libfundem_amd.so)
behind the C-ABI of include/fundem_amd.h, driven by the C++ host API
(fundem_amd/include/fundem/LSSolver.h). Python reaches it through
`Solver` (and `Adjoint`, its reverse-mode gradients), a ctypes binding of
that API; there is no CPU fallback.
"""
from . import generators  # noqa: F401
from .build import CORE_LIB, SOLVER_LIB, build_all  # noqa: F401
from .solver import Adjoint, NativeLibraryMissing, Solver, solver_lib  # noqa: F401

__all__ = ["Solver", "Adjoint", "generators", "build_all", "solver_lib", "NativeLibraryMissing"]
