"""Drop-in for lib/util/detect_core.m on a host array.

``detect_core(phi, rx, ry, refine)`` runs the same device kernels as
``GPEtask.detect_core`` (which reads the resident state instead) after uploading
``phi``.  rx and ry must be meshgrid coordinates, as ``grid.mesh.x`` / ``grid.mesh.y``
or their 2D slices: the kernels take the column coordinates rx(1,:) and the row
coordinates ry(:,1).
"""
import numpy as np

from . import _lib

_ctx = None


def _context():
    # any handle serves as the device / stream / error context of gpe_detect_core_array
    global _ctx
    if _ctx is None:
        from .grid import grid2d
        _ctx = _lib.Device(grid2d(1.0, 8, 1.0, 8))
    return _ctx


def detect_core(phi, rx, ry, refine=0):
    """detect_core.m:1-51.  Returns (resp, resm): (n, 2) arrays with the [x, y] of the
    positively / negatively charged cores, in MATLAB's column-major order."""
    phi = np.asarray(phi)
    if phi.ndim != 2:
        raise ValueError("phi must be a 2D slice")
    rx, ry = np.asarray(rx, dtype=np.float64), np.asarray(ry, dtype=np.float64)
    if rx.shape != phi.shape or ry.shape != phi.shape:
        raise ValueError("rx, ry must have the shape of phi")
    x, y = rx[0, :], ry[:, 0]
    if not (np.array_equal(rx, np.broadcast_to(x, rx.shape)) and np.array_equal(ry, np.broadcast_to(y[:, None], ry.shape))):
        raise ValueError("rx, ry must be meshgrid coordinates (constant along columns / rows)")
    return _context().detect_core(x, y, refine, phi=phi)
