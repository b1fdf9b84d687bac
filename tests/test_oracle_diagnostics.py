"""Pins of the oracle's detect_core (lib/util/detect_core.m) and of its MATLAB del2:
closed forms of the discrete Laplacian, fields whose zeros are known exactly (linear
and bilinear fields, where the bilinear interpolation of the reference is exact),
the charge sign and the column-major order of the output."""
import numpy as np
import pytest

from oracle.diagnostics import del2_uniform, detect_core


def mesh(nx, ny, hx, hy, x0=-3.0, y0=-2.0):
    x = x0 + hx * np.arange(nx)
    y = y0 + hy * np.arange(ny)
    X, Y = np.meshgrid(x, y)
    return x, y, X, Y


def test_del2_closed_forms():
    # MATLAB del2 in 2D is (f_xx + f_yy)/4; second differences of cubics are exact, and
    # the linear extrapolation onto the edges is exact when f_xx is linear
    x, y, X, Y = mesh(12, 10, 0.5, 0.3)
    assert np.allclose(del2_uniform(X ** 2 + Y ** 2, 0.5, 0.3), 1.0, rtol=0, atol=1e-12)
    assert np.allclose(del2_uniform(X ** 3 + 2 * Y ** 3, 0.5, 0.3), (6 * X + 12 * Y) / 4, rtol=0, atol=1e-11)
    assert np.allclose(del2_uniform(X ** 2 * Y, 0.5, 0.3), Y / 2, rtol=0, atol=1e-12)
    # hx belongs to the columns: with the spacings swapped x^2 gives 2 (hx/hy)^2 / 4
    assert np.allclose(del2_uniform(X ** 2, 0.5, 0.3), 0.5, rtol=0, atol=1e-12)
    assert not np.allclose(del2_uniform(X ** 2, 0.3, 0.5), 0.5, rtol=0, atol=1e-3)
    # x^4: the centred second difference is 12 x^2 + 2 h^2 inside; at the edges the
    # linear extrapolation from the two inner columns, 12 (2 x_1^2 - x_2^2) + 2 h^2
    d = del2_uniform(X ** 4, 0.5, 0.3)
    assert np.allclose(d[:, 1:-1], 3 * X[:, 1:-1] ** 2 + 0.5 ** 2 / 2, rtol=0, atol=1e-10)
    assert np.allclose(d[:, 0], 3 * (2 * x[1] ** 2 - x[2] ** 2) + 0.5 ** 2 / 2, rtol=0, atol=1e-10)
    assert np.allclose(d[:, -1], 3 * (2 * x[-2] ** 2 - x[-3] ** 2) + 0.5 ** 2 / 2, rtol=0, atol=1e-10)


def _interior(pts, x, y, refined=None):
    # plaquettes at row 0 or column 0 wrap around the array (circshift); keep the others.
    # The refined rows come in the same order as the plaquette centres `pts`.
    keep = (pts[:, 0] > x[0]) & (pts[:, 1] > y[0]) if pts.size else np.zeros(0, bool)
    return (pts if refined is None else refined)[keep]


@pytest.mark.parametrize("charge", [1, -1])
@pytest.mark.parametrize("hx,hy", [(0.25, 0.25), (0.25, 0.4)])
def test_linear_field_core(charge, hx, hy):
    rng = np.random.Generator(np.random.PCG64(3))
    x, y, X, Y = mesh(24, 20, hx, hy)
    for _ in range(5):
        xc = x[5] + rng.uniform(0.05, 0.95) * (x[15] - x[5])
        yc = y[5] + rng.uniform(0.05, 0.95) * (y[12] - y[5])
        phi = (X - xc) + charge * 1j * (Y - yc)
        j = int(np.searchsorted(x, xc))        # x[j-1] < xc < x[j]
        i = int(np.searchsorted(y, yc))
        rp0, rm0 = detect_core(phi, X, Y, 0)
        mine, other = (rp0, rm0) if charge > 0 else (rm0, rp0)
        assert _interior(other, x, y).shape[0] == 0
        got = _interior(mine, x, y)
        ysh = hy if charge > 0 else hx          # detect_core.m:39 / :48
        assert got.shape == (1, 2)
        assert got[0, 0] == x[j] - hx / 2 and got[0, 1] == y[i] - ysh / 2
        # refined: the exact zero, with the reference's use of dy along x and dx along y
        rp, rm = detect_core(phi, X, Y, 1)
        got = _interior(mine, x, y, rp if charge > 0 else rm)
        lam, s = (x[j] - xc) / hx, (y[i] - yc) / hy
        assert np.allclose(got[0], [x[j] - hy * lam, y[i] - hx * s], rtol=0, atol=1e-12)
        if hx == hy:
            assert np.allclose(got[0], [xc, yc], rtol=0, atol=1e-12)


def test_bilinear_field_two_zeros():
    # Re = (s - s1)(l + c), Im = (l - l1)(s + e) in plaquette units s = (y_i - y)/hy,
    # l = (x_j - x)/hx: zeros at (s1, l1) and (-e, -c), of opposite charge
    hx, hy = 0.5, 0.4
    x, y, X, Y = mesh(24, 20, hx, hy, x0=-6.0, y0=-4.0)
    i, j = 7, 8
    s1, l1, c, e = 0.3, 0.6, 7.4, 5.2
    S, L = (y[i] - Y) / hy, (x[j] - X) / hx
    phi = (S - s1) * (L + c) + 1j * (L - l1) * (S + e)
    rp0, rm0 = detect_core(phi, X, Y, 0)
    rp, rm = detect_core(phi, X, Y, 1)
    rp, rm = _interior(rp0, x, y, rp), _interior(rm0, x, y, rm)
    # zero 1 sits in plaquette (i, j); zero 2 at x = x_j + c hx, y = y_i + e hy, whose
    # plaquette corner is (i + 6, j + 8) with local coordinates s = 0.8, l = 0.6
    z1 = [x[j] - hy * l1, y[i] - hx * s1]
    z2 = [x[j + 8] - hy * 0.6, y[i + 6] - hx * 0.8]
    # Jacobian signs: d(Re, Im)/d(s, l) is positive at zero 1, negative at zero 2, and
    # (s, l) reverses the orientation of (x, y)
    assert rm.shape == (1, 2) and rp.shape == (1, 2)
    assert np.allclose(rm[0], z1, rtol=0, atol=1e-12)
    assert np.allclose(rp[0], z2, rtol=0, atol=1e-12)


def _vortex_field(X, Y, cores, width=4.0, xi=0.5):
    # each core z / sqrt(|z|^2 + xi^2) vanishes exactly at (xc, yc); the envelope is real
    phi = np.exp(-(X ** 2 + Y ** 2) / (2 * width ** 2)).astype(np.complex128)
    for xc, yc, q in cores:
        z = (X - xc) + q * 1j * (Y - yc)
        phi *= z / np.sqrt(np.abs(z) ** 2 + xi ** 2)
    return phi


def test_cores_in_column_major_order_and_cells():
    rng = np.random.Generator(np.random.PCG64(9))
    hx = hy = 0.2
    x, y, X, Y = mesh(64, 64, hx, hy, x0=-6.4, y0=-6.4)
    cores = [(rng.uniform(-3, 3), rng.uniform(-3, 3), q) for q in (1, 1, -1, 1, -1, -1, 1)]
    phi = _vortex_field(X, Y, cores)
    dx, dy = x[32] - x[31], y[32] - y[31]      # detect_core.m:14-15 (not exactly 0.2)
    rp, rm = detect_core(phi, X, Y, 0)
    for got, q in ((rp, 1), (rm, -1)):
        exp = []
        for xc, yc, qq in cores:
            if qq != q:
                continue
            j, i = int(np.searchsorted(x, xc)), int(np.searchsorted(y, yc))
            exp.append((j, i, x[j] - dx / 2, y[i] - (dy if q > 0 else dx) / 2))
        exp.sort()                          # find order: column (x) first, then row (y)
        assert got.shape == (len(exp), 2)
        assert np.array_equal(got, np.array([[e[2], e[3]] for e in exp]))
    # refined positions lie near the cores (the core profiles bend the interpolants)
    rp1, rm1 = detect_core(phi, X, Y, 1)
    for got, q in ((rp1, 1), (rm1, -1)):
        want = sorted([(xc, yc) for xc, yc, qq in cores if qq == q])
        assert np.allclose(sorted(map(tuple, got)), want, rtol=0, atol=0.1 * hx)


def test_refined_point_zeroes_the_bilinear_interpolants():
    # noisy fields: every refined core of a plaquette that does not wrap makes both
    # bilinear interpolants vanish (a winding of +-1 forces a zero inside)
    rng = np.random.Generator(np.random.PCG64(21))
    hx, hy = 0.3, 0.3
    x, y, X, Y = mesh(16, 16, hx, hy)
    checked = 0
    for _ in range(200):
        phi = np.exp(-(X ** 2 + Y ** 2) / 8) * ((X - rng.uniform(-1, 1)) + 1j * (Y - rng.uniform(-1, 1)))
        phi = phi + 0.05 * (rng.normal(size=X.shape) + 1j * rng.normal(size=X.shape))
        rp0, rm0 = detect_core(phi, X, Y, 0)
        rp, rm = detect_core(phi, X, Y, 1)
        for px, py in np.concatenate([_interior(rp0, x, y, rp), _interior(rm0, x, y, rm)]):
            j = int(np.searchsorted(x, px))
            i = int(np.searchsorted(y, py))
            lam, s = (x[j] - px) / hy, (y[i] - py) / hx     # as reported (hx = hy here)
            f = (phi[i, j] * (1 - s) * (1 - lam) + phi[i - 1, j] * s * (1 - lam)
                 + phi[i, j - 1] * (1 - s) * lam + phi[i - 1, j - 1] * s * lam)
            assert abs(f) < 1e-12 * np.max(np.abs(phi))
            checked += 1
    assert checked > 100
