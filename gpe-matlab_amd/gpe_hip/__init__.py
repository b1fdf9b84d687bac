"""MI355X (gfx950) split-step Fourier solvers with GPE-MATLAB's grid/task API.

    from gpe_hip import grid2d, grid3d, GPEtask
    grid = grid2d(17, 256, 17, 256)
    task = GPEtask(grid, lambda X, Y, Z: 0.5 * (X**2 + Y**2))
    task.Ntotal, task.g = 1e5, 0.06
    phi = task.groundstate_itp(0.02, 1e-6)
    task.init_state = phi
    task.solve_split(0.004, 50, 500)

The solvers run through the C ABI of libgpehip.so (include/gpe_hip.h); the
package raises GPELibraryError when that library is missing.
"""
from ._lib import GPELibraryError, GPENonFiniteError, device_info, load  # noqa: F401
from .grid import grid2d, grid3d  # noqa: F401
from .task import GPEtask, GPEtask2comp, GPEtaskSpinor, SGPEtask  # noqa: F401
from .diagnostics import detect_core  # noqa: F401
from .vtd import BasisVtd  # noqa: F401

__all__ = ["grid2d", "grid3d", "GPEtask", "GPEtask2comp", "GPEtaskSpinor", "SGPEtask", "BasisVtd", "detect_core", "GPELibraryError", "GPENonFiniteError", "device_info", "load"]
