"""GPU parity of the rotating frame (scalar omega: set_kinop.m:5-7,
ssft_kin_step.m:22-25, lap(phi, omega)) against the oracle."""
import numpy as np
import pytest

import oracle
from conftest import rel_l2
from test_gpu_scalar import _tasks, harm2, harm3

pytestmark = pytest.mark.gpu
TOL = 1e-11

CASES = [(2, (32, 24), (6, 5)), (2, (64, 64), (8, 8)), (2, (128, 96), (9, 8)),
         (3, (16, 24, 12), (5, 6, 4)), (3, (32, 32, 16), (6, 6, 4)), (3, (48, 64, 32), (7, 8, 6))]


@pytest.mark.parametrize("dims,n,xmax", CASES)
@pytest.mark.parametrize("gamma", [0.0, 0.02])
def test_rotating_solve_split_parity(dims, n, xmax, gamma):
    tg, to = _tasks(dims, n, xmax, harm2 if dims == 2 else harm3, gamma=gamma, omega=0.45)
    pg = tg.solve_split(0.004, 20, 3)
    po = to.solve_split(0.004, 20, 3)
    err = rel_l2(pg, po)
    print(f"rotating {dims}D {n} gamma={gamma}: rel-L2 vs oracle = {err:.3e}")
    assert err < TOL
    assert np.allclose(tg.history["mu"], to.history["mu"], rtol=TOL, atol=0)
    assert np.allclose(tg.history["n"], to.history["n"], rtol=TOL, atol=0)


@pytest.mark.parametrize("dims,n,xmax", [CASES[0], CASES[3]])
def test_rotating_energy_and_mu(dims, n, xmax):
    tg, to = _tasks(dims, n, xmax, harm2 if dims == 2 else harm3, omega=-0.3)
    eg, eo = tg.get_energy(tg.init_state), to.get_energy(to.init_state)
    assert abs(eg - eo) <= 1e-12 * abs(eo)
    # the rotation term changes the energy of the vortex state
    to.omega = 0.0
    assert abs(eo - to.get_energy(to.init_state)) > 1e-6 * abs(eo)


@pytest.mark.parametrize("dims,n,xmax", [CASES[1], CASES[4]])
def test_rotating_itp_fixed_iterations(dims, n, xmax):
    tg, to = _tasks(dims, n, xmax, harm2 if dims == 2 else harm3, omega=0.35)
    for t in (tg, to):
        t.itp_min_iter = 20
        t.itp_max_iter = 40
    with pytest.warns(UserWarning):
        pg, MUg, MU2g, EEg = tg.groundstate_itp(0.01, 1e-12, tg.init_state, nargout=4)
    with pytest.warns(UserWarning):
        po, MUo, MU2o, EEo = to.groundstate_itp(0.01, 1e-12, to.init_state, nargout=4)
    err = rel_l2(pg, po)
    print(f"rotating ITP {dims}D {n}: rel-L2 {err:.3e}")
    assert err < TOL
    assert np.allclose(MUg, MUo, rtol=TOL, atol=1e-13)
    assert np.allclose(MU2g, MU2o, rtol=1e-10, atol=1e-12) and np.allclose(EEg, EEo, rtol=1e-10, atol=1e-12)


def test_rotating_frame_vortex_eigenstate_stationary_on_gpu():
    # g = 0: (x + i y) e^{-r^2/2} has energy 2 + omega in the frame of lap(phi, omega);
    # solve_split subtracts that mu and the state stays put up to O(dt^2)
    import gpe_hip
    G = gpe_hip.grid2d(7, 48, 7, 48)
    t = gpe_hip.GPEtask(G, 0.5 * (G.mesh.x ** 2 + G.mesh.y ** 2))
    t.g, t.omega, t.Ntotal = 0.0, 0.4, 1.0
    phi0 = G.normalize((G.mesh.x + 1j * G.mesh.y) * np.exp(-(G.mesh.x ** 2 + G.mesh.y ** 2) / 2))
    t.init_state = phi0
    out = t.solve_split(0.005, 20, 10)
    assert abs(t.history["mu"][0] - 2.4) < 1e-10
    assert rel_l2(out, phi0) < 1e-4


def test_rotating_frame_rejected_for_two_components():
    import gpe_hip
    G = gpe_hip.grid2d(5, 16, 5, 16)
    V = 0.5 * (G.mesh.x ** 2 + G.mesh.y ** 2)
    t = gpe_hip.GPEtaskSpinor(G, [V, V])
    t.g = np.eye(2)
    t.Ntotal = 1.0
    t.omega = 0.3
    t.init_state = [np.exp(-(G.mesh.x ** 2 + G.mesh.y ** 2)) + 0j] * 2
    with pytest.raises(NotImplementedError):
        t.solve_split(0.01, 10, 1)
