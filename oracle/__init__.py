"""FP64 CPU oracle for the split-step Fourier path of GPE-MATLAB.

TEST INFRASTRUCTURE ONLY.  Nothing in the product path (the HIP library, its C-ABI
and the ``gpe_hip`` host package) imports, links or calls anything here.  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s parity and ``cpu_baseline``
legs use it, and only as the checker / the timed CPU baseline.

The oracle is a numpy/scipy restatement of the reference MATLAB classes (the
reference is MATLAB and cannot run here).  Each function cites the reference
file:line it follows.  Arrays keep MATLAB's logical shape ``[ny, nx(, nz)]``
(numpy ``meshgrid`` 'xy' indexing gives the same shape); the C-ABI receives the
same data flattened in column-major order.

Pinning: every function is checked in ``tests/test_oracle_*.py`` against
closed forms, invariants, brute-force DFTs or a textbook routine (see the test
docstrings).  No reference golden vectors exist (the reference ships no tests or
data), so end-to-end MATLAB trajectories themselves remain pinned only through
those checks.
"""
from .grids import Grid2D, Grid3D, grid2d, grid3d  # noqa: F401
from .gpetask import GPEtask, GPEtask2comp, GPEtaskSpinor, SGPEtask  # noqa: F401
from .diagnostics import detect_core  # noqa: F401
