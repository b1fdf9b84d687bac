"""C-order host transfers (gpe_set_state_c, gpe_get_state_c, gpe_set_potential_c,
gpe_set_coupling_c) against the column-major ones.

The reference's arrays are MATLAB (column-major) arrays; numpy callers hold C-order
arrays, which the library reorders on the device.  Both routes must give the same
device contents, so every check here is bitwise: a C-order upload read back through
the column-major download, a column-major upload read back through the C-order
download (with a pending renormalisation scale), and whole runs started from either
kind of array.  The grids include lengths that are not multiples of the 32 x 32
reorder tile.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES_2D = [(96, 48), (12, 24), (256, 64)]
SHAPES_3D = [(24, 48, 12), (64, 32, 16), (12, 12, 96)]


def _grid(gpe_hip, dims):
    if len(dims) == 2:
        return gpe_hip.grid2d(7.0, dims[0], 6.0, dims[1])
    return gpe_hip.grid3d(6.0, dims[0], 5.0, dims[1], 5.5, dims[2])


def _raw_get_colmajor(dev, comp=0):
    """gpe_get_state (column-major, pending scale applied on the host) -> C-order array."""
    from gpe_hip import _lib
    out = np.empty(dev.npts, dtype=np.complex128)
    dev.call("gpe_get_state", comp, out.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)))
    return _lib.from_colmajor_c128(out, dev.shape)


def _state(G, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.standard_normal(G.size) + 1j * rng.standard_normal(G.size)


@pytest.mark.parametrize("dims", SHAPES_2D + SHAPES_3D)
def test_c_order_round_trips_bitwise(dims):
    import gpe_hip
    from gpe_hip import _lib
    G = _grid(gpe_hip, dims)
    dev = _lib.Device(G)
    dev.set_params(1, 0.5, 0.0, 0.0, 0.0, 0.0, 10)
    psi = _state(G, sum(dims))
    assert psi.flags.c_contiguous
    dev.set_state(0, psi)                                  # C order: reordered on the device
    assert np.array_equal(_raw_get_colmajor(dev), psi)
    assert np.array_equal(dev.get_state(0), psi)           # C-order download
    f = np.asfortranarray(_state(G, sum(dims) + 1))
    dev.set_state(0, f)                                    # column-major: copied as it is
    assert np.array_equal(dev.get_state(0), f)
    assert np.array_equal(_raw_get_colmajor(dev), f)
    strided = _state(G, 7)[..., ::-1]                      # neither order: copied to C order
    dev.set_state(0, strided)
    assert np.array_equal(dev.get_state(0), strided)


@pytest.mark.parametrize("dims", [(96, 48), (24, 48, 12)])
def test_c_order_download_applies_pending_scale(dims):
    """After a damped run with Ntotal > 0 the device holds a pending renormalisation
    (solve_split.m:72); both downloads apply it with the same multiplication."""
    import gpe_hip
    G = _grid(gpe_hip, dims)
    r2 = G.mesh.x ** 2 + G.mesh.y ** 2 + (G.mesh.z ** 2 if len(dims) == 3 else 0)
    t = gpe_hip.GPEtask(G, 0.5 * r2)
    t.g, t.gamma, t.Ntotal = 0.7, 0.05, 10.0
    t.init_state = np.exp(-r2 / 3) * np.exp(1j * G.mesh.x)
    t.solve_split(0.002, 10, 1)
    dev = t._device()
    a = dev.get_state(0)
    b = _raw_get_colmajor(dev)
    assert np.array_equal(a, b)
    # the damped block lost norm; the download carries the renormalisation to Ntotal
    assert abs(G.integrate(np.abs(a) ** 2) / 10.0 - 1.0) < 1e-12


@pytest.mark.parametrize("dims", [(64, 48), (24, 48, 12)])
def test_runs_from_c_and_column_major_arrays_identical(dims):
    import gpe_hip
    from gpe_hip import _lib
    G = _grid(gpe_hip, dims)
    r2 = G.mesh.x ** 2 + G.mesh.y ** 2 + (G.mesh.z ** 2 if len(dims) == 3 else 0)
    V = 0.5 * r2 + 0.1 * G.mesh.x
    psi = np.exp(-r2 / 3) * np.exp(1j * (G.mesh.x + 0.3 * G.mesh.y))
    out = []
    for conv in (np.ascontiguousarray, np.asfortranarray):
        dev = _lib.Device(G)
        dev.set_params(1, 0.7, 0.01, 0.0, 10.0, 0.0, 10)
        dev.set_potential(0, conv(V))
        dev.set_state(0, conv(psi))
        dev.split_begin(0.002, 10)
        for j in range(1, 3):
            dev.split_outer(j)
        out.append(dev.get_state(0))
    assert np.array_equal(out[0], out[1])


def test_spinor_coupling_array_c_and_column_major_identical():
    import gpe_hip
    from gpe_hip import _lib
    G = gpe_hip.grid3d(6.0, 24, 5.0, 48, 5.5, 12)
    r2 = G.mesh.x ** 2 + G.mesh.y ** 2 + G.mesh.z ** 2
    V = 0.5 * r2
    a = np.exp(-r2 / 3) * np.exp(1j * G.mesh.x)
    b = np.exp(-r2 / 2) * np.exp(-1j * G.mesh.y)
    om = 0.5 * np.exp(2j * G.mesh.x + 0.3j * G.mesh.z)
    out = []
    for conv in (np.ascontiguousarray, np.asfortranarray):
        dev = _lib.Device(G)
        dev.set_params(2, [[0.02, 0.018], [0.018, 0.02]], 0.0, 0.0, 10.0, 0.0, 10)
        dev.set_potential(0, conv(V))
        dev.set_potential(1, conv(V))
        dev.set_state(0, conv(a))
        dev.set_state(1, conv(b))
        dev.set_coupling(conv(om))
        dev.split_begin(0.002, 10)
        dev.split_outer(1)
        out.append((dev.get_state(0), dev.get_state(1)))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
