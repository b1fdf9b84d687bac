"""Device-evaluated time-dependent potentials (BasisVtd -> gpe_vtd_set/gpe_vtd_coeffs)
against the oracle, which evaluates the same Vtd on the host (GPEtask.m:62-71)."""
import numpy as np
import pytest

import oracle
from conftest import rel_l2
from test_gpu_scalar import _tasks, harm2, harm3

pytestmark = pytest.mark.gpu


def _basis(grid):
    X, Y = grid.mesh.x, grid.mesh.y
    return [np.exp(-((X - 1.0) ** 2 + Y ** 2)), 0.5 * (X ** 2 - Y ** 2)]


def _coeffs(task, t):
    return [0.8 * np.sin(7.0 * t), 0.1 * np.cos(3.0 * t)]


@pytest.mark.parametrize("dims,n,xmax", [(2, (32, 24), (6, 5)), (2, (64, 64), (8, 8)), (3, (16, 24, 12), (5, 6, 4))])
@pytest.mark.parametrize("gamma", [0.0, 0.02])
def test_basis_vtd_matches_oracle(dims, n, xmax, gamma):
    import gpe_hip
    tg, to = _tasks(dims, n, xmax, harm2 if dims == 2 else harm3, gamma=gamma)
    tg.Vtd = gpe_hip.BasisVtd(_basis(tg.grid), _coeffs)
    to.Vtd = gpe_hip.BasisVtd(_basis(to.grid), _coeffs)   # host evaluation in the oracle
    pg, po = tg.solve_split(0.004, 20, 3), to.solve_split(0.004, 20, 3)
    err = rel_l2(pg, po)
    print(f"BasisVtd {dims}D {n} gamma={gamma}: rel-L2 {err:.2e}")
    assert err < 1e-11
    assert np.allclose(tg.history["mu"], to.history["mu"], rtol=1e-11, atol=0)


def test_basis_vtd_equals_host_upload():
    # the device-built potential and the host-uploaded one give the same run to rounding
    import gpe_hip
    tg, _ = _tasks(2, (64, 48), (8, 7), harm2)
    th, _ = _tasks(2, (64, 48), (8, 7), harm2)
    bv = gpe_hip.BasisVtd(_basis(tg.grid), _coeffs)
    tg.Vtd = bv
    th.Vtd = lambda task, t: bv(task, t)   # a plain function handle: host path
    assert rel_l2(tg.solve_split(0.004, 20, 2), th.solve_split(0.004, 20, 2)) < 1e-13


def test_basis_vtd_spinor():
    import gpe_hip
    G, O = gpe_hip.grid2d(6, 32, 5, 24), oracle.grid2d(6, 32, 5, 24)
    V = 0.5 * (G.mesh.x ** 2 + G.mesh.y ** 2)
    r2 = G.mesh.x ** 2 + G.mesh.y ** 2
    psi = [np.exp(-r2 / 3) + 0j, 0.7 * np.exp(-r2 / 2) * np.exp(0.2j * G.mesh.x)]
    tg, to = gpe_hip.GPEtaskSpinor(G, [V, V + 0.2 * G.mesh.x]), oracle.GPEtaskSpinor(O, [V, V + 0.2 * G.mesh.x])
    for t, grid in ((tg, G), (to, O)):
        t.g = np.array([[0.5, 0.45], [0.4, 0.6]])
        t.coupling = 0.3 * np.exp(0.5j)
        t.Ntotal = 4.0
        t.init_state = [p.copy() for p in psi]
        t.Vtd = gpe_hip.BasisVtd(_basis(grid), _coeffs)
    pg, po = tg.solve_split(0.004, 20, 2), to.solve_split(0.004, 20, 2)
    assert max(rel_l2(pg[0], po[0]), rel_l2(pg[1], po[1])) < 1e-11


def test_basis_vtd_rejects_bad_input():
    import gpe_hip
    G = gpe_hip.grid2d(5, 16, 5, 16)
    with pytest.raises(ValueError):
        gpe_hip.BasisVtd([G.mesh.x] * 17, lambda t, s: [0.0] * 17)
    t, _ = _tasks(2, (16, 16), (5, 5), harm2)
    t.Vtd = gpe_hip.BasisVtd([G.mesh.x], lambda task, s: [1.0, 2.0])
    with pytest.raises(ValueError, match="returned 2 values"):
        t.solve_split(0.004, 10, 1)


def test_plain_vtd_spinor_mu_at_current_time():
    # a host Vtd on the two-component path: the block-end mu reads V at task.current_time
    import gpe_hip
    G, O = gpe_hip.grid2d(6, 32, 5, 24), oracle.grid2d(6, 32, 5, 24)
    V = 0.5 * (G.mesh.x ** 2 + G.mesh.y ** 2)
    r2 = G.mesh.x ** 2 + G.mesh.y ** 2
    psi = [np.exp(-r2 / 3) + 0j, 0.7 * np.exp(-r2 / 2) + 0j]
    tg, to = gpe_hip.GPEtaskSpinor(G, [V, V]), oracle.GPEtaskSpinor(O, [V, V])
    for t, grid in ((tg, G), (to, O)):
        t.g = np.array([[0.5, 0.45], [0.4, 0.6]])
        t.coupling = 0.2
        t.Ntotal = 3.0
        t.init_state = [p.copy() for p in psi]
        t.Vtd = (lambda gx: (lambda task, time: 0.6 * np.sin(5 * time) * gx))(grid.mesh.x)
    pg, po = tg.solve_split(0.004, 20, 3), to.solve_split(0.004, 20, 3)
    assert max(rel_l2(pg[0], po[0]), rel_l2(pg[1], po[1])) < 1e-11
    assert np.allclose(tg.history["mu"], np.real(to.history["mu"]), rtol=1e-11, atol=1e-13)
