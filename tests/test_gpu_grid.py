"""The grid's own transforms on the device (gpe_grid_transform): grid2d/grid3d fft, ifft
and lap (lib/@grid2d/fft.m:11, ifft.m:11, lap.m:5-9; lib/@grid3d/fft.m:11, ifft.m:11,
lap.m:5-10) against numpy's fftn/ifftn, the oracle's lap and a plane-wave closed form.

The device transforms axis by axis with its own Stockham FFT; numpy's pocketfft uses
another factorisation, so the two agree to rounding.  Tolerance: relative L2 1e-13 for
the transforms, 1e-12 for lap (its result is a sum of per-axis terms)."""
import numpy as np
import pytest

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu

# every supported axis length (8 ... 512) appears on some axis
GRIDS = [(2, (8.0, 48, 6.0, 24)), (2, (10.0, 256, 9.0, 96)), (2, (6.0, 512, 5.0, 384)),
         (3, (6.0, 24, 5.0, 32, 4.0, 16)), (3, (6.0, 64, 7.0, 256, 5.0, 12)),
         (3, (5.0, 192, 5.0, 12, 5.0, 8)), (3, (5.0, 128, 4.0, 16, 6.0, 32))]


def _pair(dims, args):
    import gpe_hip
    if dims == 2:
        return gpe_hip.grid2d(*args), oracle.grid2d(*args)
    return gpe_hip.grid3d(*args), oracle.grid3d(*args)


def _field(G, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.standard_normal(G.size) + 1j * rng.standard_normal(G.size)


@pytest.mark.parametrize("dims,args", GRIDS)
def test_grid_fft_ifft_match_numpy(dims, args):
    G, _ = _pair(dims, args)
    phi = _field(G, sum(G.size))
    F = G.fft(phi)
    e_f = rel_l2(F, np.fft.fftn(phi))
    e_i = rel_l2(G.ifft(F), np.fft.ifftn(F))
    e_r = rel_l2(G.ifft(F), phi)
    print(f"grid {G.size}: fft {e_f:.2e}, ifft {e_i:.2e}, round trip {e_r:.2e}", flush=True)
    assert e_f < 1e-13 and e_i < 1e-13 and e_r < 1e-13


@pytest.mark.parametrize("dims,args", GRIDS)
@pytest.mark.parametrize("omega", [0.0, 0.35])
def test_grid_lap_matches_oracle(dims, args, omega):
    G, O = _pair(dims, args)
    phi = _field(G, 7 + sum(G.size))
    e = rel_l2(G.lap(phi, omega), O.lap(phi, omega))
    print(f"grid {G.size} lap omega={omega}: {e:.2e}", flush=True)
    assert e < 1e-12


def test_grid_lap_plane_wave_closed_form():
    """lap(e^{i k.r}) = |k|^2/2 e^{i k.r} for k on the grid (omega = 0), and with omega
    the rotating-frame operator adds omega (x k_y - y k_x) e^{i k.r}."""
    import gpe_hip
    G = gpe_hip.grid3d(6.0, 32, 5.0, 24, 4.0, 16)
    ix, iy, iz = 3, -2, 1
    kx, ky, kz = (ix * np.pi / (G.xmax + G.dx / 2), iy * np.pi / (G.ymax + G.dy / 2), iz * np.pi / (G.zmax + G.dz / 2))
    X, Y, Z = G.mesh.x, G.mesh.y, G.mesh.z
    pw = np.exp(1j * (kx * X + ky * Y + kz * Z))
    assert rel_l2(G.lap(pw), (kx ** 2 + ky ** 2 + kz ** 2) / 2 * pw) < 1e-12
    om = 0.3
    want = ((kx ** 2 + ky ** 2 + kz ** 2) / 2 + om * (X * ky - Y * kx)) * pw
    assert rel_l2(G.lap(pw, om), want) < 1e-12


def test_grid_transform_rejects_bad_input():
    import gpe_hip
    G = gpe_hip.grid2d(5.0, 16, 5.0, 16)
    with pytest.raises(ValueError):
        G.fft(np.zeros((8, 8), dtype=complex))
    dev = G._device()
    with pytest.raises(gpe_hip.GPELibraryError):
        dev.grid_transform("lap", np.zeros(G.size, dtype=complex), float("nan"))
    # a length without a device plan is refused, not computed some other way
    G40 = gpe_hip.grid2d(5.0, 40, 5.0, 16)
    with pytest.raises(gpe_hip.GPELibraryError):
        G40.fft(np.zeros(G40.size, dtype=complex))
