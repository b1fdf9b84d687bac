"""gpe_detect_core_noalloc (fixed device scratch, no allocation) against gpe_detect_core
through the C ABI: the same slices give the same cores, bit for bit, on two calls in a
row (the second on a smaller grid), and an oversized request is refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _field(G, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    X, Y = G.mesh.x, G.mesh.y
    if G.ndims == 3:
        X, Y = X[:, :, 0], Y[:, :, 0]
    phi = np.exp(-(X ** 2 + Y ** 2) / 40).astype(np.complex128)
    for _ in range(12):
        xc, yc, q = rng.uniform(-5, 5), rng.uniform(-5, 5), rng.choice([-1, 1])
        z = (X - xc) + q * 1j * (Y - yc)
        phi *= z / np.sqrt(np.abs(z) ** 2 + 0.25)
    phi += 0.01 * (rng.normal(size=X.shape) + 1j * rng.normal(size=X.shape))
    if G.ndims == 3:
        phi = np.repeat(phi[:, :, None], G.nz, axis=2) * np.exp(1j * rng.uniform(0, 1, G.nz))[None, None, :]
    return phi


def _call(dev, name, x, y, refine, cap, zslice=0):
    rp, rm = np.full((cap, 2), -7.0), np.full((cap, 2), -7.0)
    n_p, n_m = C.c_int(-1), C.c_int(-1)
    fn = getattr(dev.lib, name)
    rc = fn(dev.h, 0, int(zslice), x.ctypes.data_as(C.POINTER(C.c_double)), y.ctypes.data_as(C.POINTER(C.c_double)),
            int(refine), int(cap), rp.ctypes.data_as(C.POINTER(C.c_double)), C.byref(n_p),
            rm.ctypes.data_as(C.POINTER(C.c_double)), C.byref(n_m))
    return rc, rp, n_p.value, rm, n_m.value


def _device(grid, phi):
    from gpe_hip import _lib
    dev = _lib.Device(grid)
    dev.set_state(0, phi)
    return dev


@pytest.mark.parametrize("refine", [0, 1])
def test_noalloc_matches_detect_core_two_calls(refine):
    import gpe_hip
    grids = [gpe_hip.grid3d(8, 256, 8, 192, 4, 8), gpe_hip.grid2d(8, 96, 8, 64)]   # second call: smaller slice
    for call, G in enumerate(grids):
        dev = _device(G, _field(G, 3 + call))
        x = np.ascontiguousarray(G.mesh.x[0, :, 0] if G.ndims == 3 else G.mesh.x[0, :])
        y = np.ascontiguousarray(G.mesh.y[:, 0, 0] if G.ndims == 3 else G.mesh.y[:, 0])
        zs = 5 if G.ndims == 3 else 0
        cap = 200
        rc0, rp0, np0, rm0, nm0 = _call(dev, "gpe_detect_core", x, y, refine, cap, zs)
        rc1, rp1, np1, rm1, nm1 = _call(dev, "gpe_detect_core_noalloc", x, y, refine, cap, zs)
        print(f"call {call}: grid {G.size}, {np1}+/{nm1}- cores (refine {refine})", flush=True)
        assert rc0 == 0 and rc1 == 0
        assert (np1, nm1) == (np0, nm0) and np0 + nm0 > 4
        assert np.array_equal(rp1, rp0) and np.array_equal(rm1, rm0)   # rows beyond the count untouched too
        # a capacity below the count: counts reported, only cap rows written
        rc2, rp2, np2, rm2, nm2 = _call(dev, "gpe_detect_core_noalloc", x, y, refine, 2, zs)
        assert rc2 == 0 and (np2, nm2) == (np0, nm0)
        assert np.array_equal(rp2, rp0[:2]) and np.array_equal(rm2, rm0[:2])
        dev.close()


def test_noalloc_refuses_more_than_its_scratch():
    import gpe_hip
    G = gpe_hip.grid2d(8, 64, 8, 64)
    dev = _device(G, _field(G, 9))
    x = np.ascontiguousarray(G.mesh.x[0, :])
    y = np.ascontiguousarray(G.mesh.y[:, 0])
    cap = (16 << 20) // 32            # output rows alone exceed the 16 MiB scratch
    rc, _, _, _, _ = _call(dev, "gpe_detect_core_noalloc", x, y, 0, cap)
    assert rc == -1                   # GPE_EINVAL
    msg = dev.lib.gpe_last_error(dev.h).decode()
    assert "needs" in msg and "scratch" in msg
    # the entry point still works afterwards
    rc, _, n_p, _, n_m = _call(dev, "gpe_detect_core_noalloc", x, y, 0, 64)
    assert rc == 0 and n_p + n_m > 0
    dev.close()
