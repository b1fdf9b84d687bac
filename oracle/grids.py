"""Oracle grids: restatement of lib/@grid1d, lib/@grid2d, lib/@grid3d.

Test infrastructure only (see ``oracle/__init__.py``).
"""
import os
from types import SimpleNamespace

import numpy as np
import scipy.fft as sfft

# FFT worker threads for the CPU baseline; the oracle's numerics do not depend on it.
WORKERS = int(os.environ.get("GPE_ORACLE_WORKERS", "0")) or len(os.sched_getaffinity(0))


def _kvec(x, dx, n):
    """Wave vectors with the +N/2 Nyquist entry.

    grid2d/init.m:27 and grid3d/init.m:33-35:
    ``k = [ (0:n/2) -(n/2-1:-1:1) ] * 2*pi/(x(end)-x(1)+dx)``.
    """
    if n % 2:
        raise ValueError("grid sizes must be even (grid2d/init.m:27 assumes n/2 integer)")
    idx = np.concatenate([np.arange(0, n // 2 + 1), -np.arange(n // 2 - 1, 0, -1)])
    return idx * 2.0 * np.pi / (x[-1] - x[0] + dx)


class _GridBase:
    """Shared reductions: grid1d/integrate.m:11, inner.m:12, norm.m:11, normalize.m:13-15."""

    def integrate(self, fun):
        # integrate.m:11  int = sum(fun(:)) * obj.weight
        return np.sum(fun) * self.weight

    def inner(self, lhs, rhs):
        # inner.m:12  y = integrate(conj(lhs).*rhs)
        return self.integrate(np.conj(lhs) * rhs)

    def norm(self, psi):
        # norm.m:11  n = sqrt(integrate(abs(psi).^2))
        return np.sqrt(self.integrate(np.abs(psi) ** 2))

    def normalize(self, psi):
        # normalize.m:13-15  psi = psi / norm(psi)
        return psi * (1.0 / self.norm(psi))

    def fft(self, fun):
        # grid2d/fft.m:11 (fft2), grid3d/fft.m:11 (fftn): unnormalised, e^{-2 pi i jk/N}
        return sfft.fftn(fun, workers=WORKERS)

    def ifft(self, ifun):
        # grid2d/ifft.m:11, grid3d/ifft.m:11: inverse with 1/N
        return sfft.ifftn(ifun, workers=WORKERS)

    # one-axis transforms: grid2d/fftx.m:11 (fft along dim 2 = x), ffty.m:11 (dim 1 = y),
    # grid3d/fftz.m (dim 3); the inverses carry 1/n of their axis
    def fftx(self, fun):
        return sfft.fft(fun, axis=1, workers=WORKERS)

    def ifftx(self, ifun):
        return sfft.ifft(ifun, axis=1, workers=WORKERS)

    def ffty(self, fun):
        return sfft.fft(fun, axis=0, workers=WORKERS)

    def iffty(self, ifun):
        return sfft.ifft(ifun, axis=0, workers=WORKERS)

    def lap(self, phi, omega=0.0):
        # grid2d/lap.m:5-10, grid3d/lap.m:5-11.  omega = 0: ifft(kk .* fft(phi)); else the
        # rotating-frame kinetic operator, one axis at a time:
        #   ifftx((kx^2/2 - kx y omega) fftx) + iffty((ky^2/2 + ky x omega) ffty) [+ ifftz(kz^2/2 fftz)]
        if omega == 0:
            return self.ifft(self.kk * self.fft(phi))
        res = self.ifftx((self.kx ** 2 / 2 - self.kx * self.mesh.y * omega) * self.fftx(phi)) + \
            self.iffty((self.ky ** 2 / 2 + self.ky * self.mesh.x * omega) * self.ffty(phi))
        if self.ndims == 3:
            res = res + self.ifftz((self.kz ** 2 / 2) * self.fftz(phi))
        return res


class Grid2D(_GridBase):
    """grid2d(xmax, nx, ymax, ny) -- grid2d.m:16-31, init.m:5-37."""

    ndims = 2

    def __init__(self, xmax, nx, ymax, ny):
        x = np.linspace(-xmax, xmax, nx)          # init.m:12
        y = np.linspace(-ymax, ymax, ny)          # init.m:13
        self.nx, self.ny, self.nz = nx, ny, 1
        self.xmax, self.ymax, self.zmax = float(xmax), float(ymax), 0.0
        self.dx = x[1] - x[0]                     # init.m:20
        self.dy = y[1] - y[0]                     # init.m:21
        self.x, self.y = x, y
        kx = _kvec(x, self.dx, nx)                # init.m:27
        ky = _kvec(y, self.dy, ny)                # init.m:28
        self.kx1, self.ky1, self.kz1 = kx, ky, np.zeros(1)
        self.kx, self.ky = np.meshgrid(kx, ky)    # init.m:30  -> [ny, nx]
        self.kz = 0
        self.kk = (self.kx ** 2 + self.ky ** 2) / 2   # init.m:31
        X, Y = np.meshgrid(x, y)                  # init.m:34
        self.mesh = SimpleNamespace(x=X, y=Y, z=0, x2=X, y2=Y)
        self.weight = self.dx * self.dy           # init.m:36
        self.size = X.shape                       # init.m:37

    # grid2d.m:45-51: fftz / ifftz are the identity on a 2D grid
    def fftz(self, phi):
        return phi

    def ifftz(self, phi):
        return phi


class Grid3D(_GridBase):
    """grid3d(xmax, nx, ymax, ny, zmax, nz) -- grid3d.m:14-29, init.m:5-46."""

    ndims = 3

    def __init__(self, xmax, nx, ymax, ny, zmax, nz):
        x = np.linspace(-xmax, xmax, nx)          # init.m:13
        y = np.linspace(-ymax, ymax, ny)          # init.m:14
        z = np.linspace(-zmax, zmax, nz)          # init.m:15
        self.nx, self.ny, self.nz = nx, ny, nz
        self.xmax, self.ymax, self.zmax = float(xmax), float(ymax), float(zmax)
        self.dx = x[1] - x[0]                     # init.m:25
        self.dy = y[1] - y[0]                     # init.m:26
        self.dz = z[1] - z[0]                     # init.m:27
        self.x, self.y, self.z = x, y, z
        kx = _kvec(x, self.dx, nx)                # init.m:33
        ky = _kvec(y, self.dy, ny)                # init.m:34
        kz = _kvec(z, self.dz, nz)                # init.m:35
        self.kx1, self.ky1, self.kz1 = kx, ky, kz
        self.kx, self.ky, self.kz = np.meshgrid(kx, ky, kz)   # init.m:37 -> [ny, nx, nz]
        self.kk = (self.kx ** 2 + self.ky ** 2 + self.kz ** 2) / 2   # init.m:38
        X, Y, Z = np.meshgrid(x, y, z)            # init.m:41
        X2, Y2 = np.meshgrid(x, y)                # init.m:42
        self.mesh = SimpleNamespace(x=X, y=Y, z=Z, x2=X2, y2=Y2)
        self.weight = self.dx * self.dy * self.dz  # init.m:45
        self.size = X.shape                        # init.m:46

    def fftz(self, fun):
        return sfft.fft(fun, axis=2, workers=WORKERS)

    def ifftz(self, ifun):
        return sfft.ifft(ifun, axis=2, workers=WORKERS)


def grid2d(xmax, nx, ymax, ny):
    return Grid2D(xmax, nx, ymax, ny)


def grid3d(xmax, nx, ymax, ny, zmax, nz):
    return Grid3D(xmax, nx, ymax, ny, zmax, nz)
