"""The largest grid the library accepts: 512^3 points (2^27, 2 GiB of complex128 state).

Kernels index elements with 32-bit offsets, so this is the size at which an overflow in
an offset would show; the host transfers move more than 2^31 bytes.  The oracle takes
minutes per step here, so the checks are properties that hold at any size:
- with gamma = 0 and a fixed mu the split step is unitary (solve_split.m:55-64 apply
  unit-modulus phases and a unit-modulus kinetic factor), so the norm is conserved;
- a run with -ddt undoes a run with +ddt (time reversal of the Strang step);
- the C-order upload and download return the array bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_512_cubed_unitary_time_reversible_and_transfers_exact():
    import gpe_hip
    n = 512
    G = gpe_hip.grid3d(10.0, n, 10.0, n, 10.0, n)
    x = G.x.reshape(1, n, 1)
    y = G.y.reshape(n, 1, 1)
    z = G.z.reshape(1, 1, n)
    r2 = x ** 2 + y ** 2 + z ** 2
    V = 0.5 * r2
    psi0 = np.exp(-r2 / 8.0) * np.exp(1j * (0.7 * x - 0.4 * y + 0.2 * z))
    assert psi0.shape == G.size
    t = gpe_hip.GPEtask(G, V)
    t.g, t.gamma, t.mu_init = 50.0, 0.0, 3.0
    t.init_state = psi0
    n0 = G.integrate(np.abs(psi0) ** 2)

    dev = t._device()
    dev.set_state(0, psi0)
    assert np.array_equal(dev.get_state(0), psi0)

    fwd = t.solve_split(0.002, 10, 2)
    drift = abs(G.integrate(np.abs(fwd) ** 2) - n0) / n0
    back = t.solve_split(-0.002, 10, 4)       # resumes at current_iter = 2: two blocks back
    err = np.linalg.norm(back - psi0) / np.linalg.norm(psi0)
    print(f"512^3: 20 steps norm drift {drift:.2e}; 20 forward + 20 back rel-L2 {err:.2e}", flush=True)
    assert drift < 1e-12
    assert err < 1e-12
