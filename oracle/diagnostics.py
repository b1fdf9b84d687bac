"""Vortex-core detection on a 2D slice (test infrastructure only).

Restates lib/util/detect_core.m:1-66 and the part of MATLAB's ``del2`` it uses
(uniform spacing, two dimensions).  Arrays keep MATLAB's shape [ny, nx]: dimension 0
(rows) is y, dimension 1 (columns) is x; ``find`` and logical indexing enumerate
points in column-major order, reproduced here with ``order="F"``.
"""
import numpy as np
import scipy.linalg


def del2_uniform(f, hx, hy):
    """MATLAB ``del2(f, hx, hy)`` for a 2D array with uniform spacings: per dimension the
    centred second difference divided by 2 on interior points, linear extrapolation of
    those differences onto the first and last points (n > 3), then the sum over the two
    dimensions divided by 2 -- i.e. (f_xx + f_yy)/4.  hx is the spacing along
    dimension 1 (columns), hy along dimension 0 (rows)."""
    f = np.asarray(f, dtype=np.float64)
    v = np.zeros_like(f)
    for axis, h in ((1, hx), (0, hy)):
        g = np.moveaxis(np.zeros_like(f), axis, 0)
        a = np.moveaxis(f, axis, 0)
        n = a.shape[0]
        if n > 2:
            g[1:n - 1] = ((a[2:] - a[1:n - 1]) / h - (a[1:n - 1] - a[:n - 2]) / h) / (2 * h)
        if n > 3:
            g[0] = 2 * g[1] - g[2]
            g[n - 1] = 2 * g[n - 2] - g[n - 3]
        elif n == 3:
            g[0] = g[1]
            g[n - 1] = g[1]
        v = v + np.moveaxis(g, 0, axis)
    return v / 2


def _refine_core(rep, imp, rx, ry, dx, dy, ix, iy):
    """detect_core.m:54-64 (0-based ix, iy): the zero of the bilinear interpolants of
    Re and Im over the cell [ix-1, ix] x [iy-1, iy] as the generalised eigenproblem
    A v = d B v with v = (s, 1); s runs toward row ix-1, d toward column iy-1."""
    A = -np.array([[rep[ix - 1, iy] - rep[ix, iy], rep[ix, iy]],
                   [imp[ix - 1, iy] - imp[ix, iy], imp[ix, iy]]])
    B = np.array([[rep[ix, iy] - rep[ix - 1, iy] - rep[ix, iy - 1] + rep[ix - 1, iy - 1], rep[ix, iy - 1] - rep[ix, iy]],
                  [imp[ix, iy] - imp[ix - 1, iy] - imp[ix, iy - 1] + imp[ix - 1, iy - 1], imp[ix, iy - 1] - imp[ix, iy]]])
    with np.errstate(all="ignore"):
        d, v = scipy.linalg.eig(A, B)           # LAPACK xGGEV, as MATLAB's eig(A, B)
        s = v[0, :] / v[1, :]
    for i in range(2):
        # MATLAB compares complex values by their real parts
        di, si = d[i].real, s[i].real
        if di > 0 and di < 1 and si > 0 and si < 1:
            return [rx[ix, iy] - dy * di, ry[ix, iy] - dx * si]
    return [0.0, 0.0]


def detect_core(phi, rx, ry, refine=0):
    """lib/util/detect_core.m:1-51.  Returns (resp, resm), arrays of shape (n, 2) with
    the [x, y] of positively and negatively charged cores.  With ``refine`` the complex
    case keeps the real parts (the reference's row would turn complex)."""
    phi = np.asarray(phi, dtype=np.complex128)
    n, m = phi.shape                                            # 13
    dx = rx[0, m // 2] - rx[0, m // 2 - 1]                      # 14: rx(1,end/2+1)-rx(1,end/2)
    dy = ry[n // 2, 0] - ry[n // 2 - 1, 0]                      # 15
    an = np.angle(phi)                                          # 16
    rep, imp = phi.real, phi.imag                               # 17-18
    ddphi = np.abs(del2_uniform(np.abs(phi) ** 2, dx, dy))      # 19
    shifts = [(0, 0), (1, 0), (1, 1), (0, 1)]                   # 20
    angs = np.stack([np.roll(an, s, axis=(0, 1)) for s in shifts], axis=2)   # 22-25
    dif = angs - np.roll(angs, 1, axis=2)                       # 26
    res1 = (np.sum(dif > np.pi, axis=2) - np.sum(dif < -np.pi, axis=2)) * (ddphi > 0.01 * np.max(ddphi))   # 27
    out = []
    for sign in (1, -1):                                        # 29-50
        res = res1 > 0 if sign > 0 else res1 < 0
        iy_, ix_ = np.nonzero(res.T)                            # column-major enumeration
        ix_, iy_ = ix_.astype(int), iy_.astype(int)
        if refine > 0:
            pts = [_refine_core(rep, imp, rx, ry, dx, dy, max(ix, 1), max(iy, 1)) for ix, iy in zip(ix_, iy_)]
            out.append(np.array(pts, dtype=np.float64).reshape(-1, 2))
        else:
            ysh = dy if sign > 0 else dx                        # 39 uses dy, 48 uses dx
            out.append(np.column_stack([rx[ix_, iy_] - dx / 2, ry[ix_, iy_] - ysh / 2]).reshape(-1, 2))
    return out[0], out[1]
