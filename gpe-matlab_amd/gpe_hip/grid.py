"""grid2d / grid3d with the reference's properties (lib/@grid2d, lib/@grid3d).

The grid objects hold the host-side mesh and k tables exactly as the reference
builds them (init.m); the device copies of the k tables are built by the HIP
library from the same (xmax, n) description.  The quadrature helpers below are
host conveniences of the grid API (grid1d/integrate.m etc.); the split-step hot
path never calls them.  fft, ifft and lap run on the GPU through the library's
transform sweeps (gpe_grid_transform), on a device handle the grid creates when it
first needs one.
"""
from types import SimpleNamespace

import numpy as np


def _kvec(x, dx, n):
    # grid2d/init.m:27 -- k = [0:n/2, -(n/2-1):-1] * 2 pi / (x(end)-x(1)+dx), Nyquist +n/2
    idx = np.concatenate([np.arange(0, n // 2 + 1), -np.arange(n // 2 - 1, 0, -1)])
    return idx * 2.0 * np.pi / (x[-1] - x[0] + dx)


class _Grid:
    def integrate(self, fun):
        # grid1d/integrate.m:11
        return np.sum(fun) * self.weight

    def inner(self, lhs, rhs):
        # grid1d/inner.m:12
        return self.integrate(np.conj(lhs) * rhs)

    def norm(self, psi):
        # grid1d/norm.m:11
        return np.sqrt(self.integrate(np.abs(psi) ** 2))

    def normalize(self, psi):
        # grid1d/normalize.m:13-15
        return psi * (1.0 / self.norm(psi))

    def _device(self):
        dev = getattr(self, "_dev", None)
        if dev is None:
            from ._lib import Device
            dev = self._dev = Device(self)
        return dev

    def fft(self, fun):
        """fftn(fun) (grid2d/fft.m:11, grid3d/fft.m:11), on the GPU."""
        return self._device().grid_transform("fft", fun)

    def ifft(self, ifun):
        """ifftn(ifun) (grid2d/ifft.m:11, grid3d/ifft.m:11), on the GPU."""
        return self._device().grid_transform("ifft", ifun)

    def lap(self, phi, omega=0.0):
        """lap(phi, omega) (grid2d/lap.m:5-9, grid3d/lap.m:5-10): ifftn(kk .* fftn(phi)),
        or for omega != 0 the per-axis rotating-frame form, on the GPU."""
        return self._device().grid_transform("lap", phi, omega)


class grid2d(_Grid):
    """grid2d(xmax, nx, ymax, ny) -- lib/@grid2d/grid2d.m:16-31, init.m:5-37."""

    ndims = 2

    def __init__(self, xmax, nx, ymax, ny):
        for n in (nx, ny):
            if n % 2:
                raise ValueError("grid sizes must be even")
        x = np.linspace(-xmax, xmax, nx)
        y = np.linspace(-ymax, ymax, ny)
        self.nx, self.ny, self.nz = int(nx), int(ny), 1
        self.xmax, self.ymax, self.zmax = float(xmax), float(ymax), 0.0
        self.x, self.y = x, y
        self.dx = x[1] - x[0]
        self.dy = y[1] - y[0]
        kx = _kvec(x, self.dx, nx)
        ky = _kvec(y, self.dy, ny)
        self.kx, self.ky = np.meshgrid(kx, ky)
        self.kz = 0
        self.kk = (self.kx ** 2 + self.ky ** 2) / 2
        X, Y = np.meshgrid(x, y)
        self.mesh = SimpleNamespace(x=X, y=Y, z=0, x2=X, y2=Y)
        self.weight = self.dx * self.dy
        self.size = X.shape


class grid3d(_Grid):
    """grid3d(xmax, nx, ymax, ny, zmax, nz) -- lib/@grid3d/grid3d.m:14-29, init.m:5-46."""

    ndims = 3

    def __init__(self, xmax, nx, ymax, ny, zmax, nz):
        for n in (nx, ny, nz):
            if n % 2:
                raise ValueError("grid sizes must be even")
        x = np.linspace(-xmax, xmax, nx)
        y = np.linspace(-ymax, ymax, ny)
        z = np.linspace(-zmax, zmax, nz)
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.xmax, self.ymax, self.zmax = float(xmax), float(ymax), float(zmax)
        self.x, self.y, self.z = x, y, z
        self.dx = x[1] - x[0]
        self.dy = y[1] - y[0]
        self.dz = z[1] - z[0]
        kx = _kvec(x, self.dx, nx)
        ky = _kvec(y, self.dy, ny)
        kz = _kvec(z, self.dz, nz)
        self.kx, self.ky, self.kz = np.meshgrid(kx, ky, kz)
        self.kk = (self.kx ** 2 + self.ky ** 2 + self.kz ** 2) / 2
        X, Y, Z = np.meshgrid(x, y, z)
        X2, Y2 = np.meshgrid(x, y)
        self.mesh = SimpleNamespace(x=X, y=Y, z=Z, x2=X2, y2=Y2)
        self.weight = self.dx * self.dy * self.dz
        self.size = X.shape
