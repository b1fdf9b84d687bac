"""GPU parity of the two-component imaginary-time solvers (GPEtaskSpinor and
GPEtask2comp groundstate_itp) and of GPEtask2comp.solve_split against the oracle."""
import numpy as np
import pytest

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu
G_MAT = [[0.5, 0.45], [0.4, 0.6]]


def _pair(cls_name, dims, n, xmax, seed=6, g=G_MAT):
    import gpe_hip
    if dims == 2:
        G, O = gpe_hip.grid2d(xmax[0], n[0], xmax[1], n[1]), oracle.grid2d(xmax[0], n[0], xmax[1], n[1])
    else:
        G = gpe_hip.grid3d(xmax[0], n[0], xmax[1], n[1], xmax[2], n[2])
        O = oracle.grid3d(xmax[0], n[0], xmax[1], n[1], xmax[2], n[2])
    r2 = G.mesh.x ** 2 + G.mesh.y ** 2 + (G.mesh.z ** 2 if dims == 3 else 0)
    V1 = 0.5 * r2
    V2 = V1 + 0.3 * G.mesh.x
    tg = getattr(gpe_hip, cls_name)(G, [V1, V2], seed=seed)
    to = getattr(oracle, cls_name)(O, [V1, V2], seed=seed)
    for t in (tg, to):
        t.g = np.array(g)
    return tg, to, G


def _cmp_states(pg, po):
    return max(rel_l2(pg[0], po[0]), rel_l2(pg[1], po[1]))


@pytest.mark.parametrize("dims,n,xmax", [(2, (32, 24), (6, 5)), (3, (16, 24, 12), (5, 6, 4))])
@pytest.mark.parametrize("coupling", ["scalar", "array"])
def test_spinor_itp_fixed_iterations(dims, n, xmax, coupling):
    tg, to, G = _pair("GPEtaskSpinor", dims, n, xmax)
    cp = 0.4 * np.exp(0.6j) if coupling == "scalar" else 0.4 * np.exp(1.5j * G.mesh.x)
    for t in (tg, to):
        t.coupling = cp
        t.Ntotal = 3.0
    tg.itp_max_iter_2c = 60
    with pytest.warns(UserWarning, match="Convergence not reached"):
        pg, MUg, MU2g, EEg = tg.groundstate_itp(0.01, 1e-30, nargout=4)
    po, MUo, MU2o, EEo = to.groundstate_itp(0.01, 1e-30, nargout=4, max_iter=60)
    err = _cmp_states(pg, po)
    print(f"spinor ITP {dims}D {coupling}: state {err:.2e}, MU {np.max(np.abs(MUg - MUo)):.2e}")
    assert tg.itp_iterations == 60 and len(MUg) == 59
    assert err < 1e-11
    for a, b in ((MUg, MUo), (MU2g, MU2o), (EEg, EEo)):
        assert np.allclose(a, b, rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("nargout", [2, 3, 4])
def test_spinor_itp_converges_like_the_oracle(nargout):
    # eps = 1e-4 stops the step-size cascade at dt < 1e-4, where the stopping test
    # compares series differences far above rounding: the cascade, the stopping
    # iteration and the final state must match the oracle's
    tg, to, G = _pair("GPEtaskSpinor", 2, (32, 32), (6, 6))
    for t in (tg, to):
        t.coupling = 0.3 * np.exp(0.7j)
        t.Ntotal = 2.0
    outg = tg.groundstate_itp(0.02, 1e-4, nargout=nargout)
    outo = to.groundstate_itp(0.02, 1e-4, nargout=nargout)
    print(f"nargout {nargout}: GPU {tg.itp_iterations} it, oracle {to.itp_iterations} it")
    assert tg.itp_iterations == to.itp_iterations
    assert abs(tg.itp_dt - to.itp_dt) <= 1e-15 * to.itp_dt
    assert _cmp_states(outg[0], outo[0]) < 1e-11
    for a, b in zip(outg[1:], outo[1:]):
        assert np.allclose(a, b, rtol=1e-10, atol=1e-12)


def test_spinor_itp_coupled_oscillator_limit():
    # V1 = V2 harmonic, g = 0, uniform coupling: mu -> 1 - |Omega|
    import gpe_hip
    G = gpe_hip.grid2d(6, 32, 6, 32)
    V = 0.5 * (G.mesh.x ** 2 + G.mesh.y ** 2)
    t = gpe_hip.GPEtaskSpinor(G, [V, V], seed=0)
    t.g = np.zeros((2, 2))
    om = 0.3 * np.exp(0.7j)
    t.coupling = om
    t.Ntotal = 2.0
    phi, MU, MU2, EE = t.groundstate_itp(0.02, 1e-9, nargout=4)
    assert abs(MU[-1] - (1 - abs(om))) < 1e-6 and abs(EE[-1] - (1 - abs(om))) < 1e-6
    assert abs(G.integrate(np.abs(phi[0]) ** 2) + G.integrate(np.abs(phi[1]) ** 2) - 2.0) < 1e-12


def test_spinor_itp_tight_eps_stop_is_rounding_limited():
    # With eps = 1e-9 the cascade runs down to dt ~ 5e-6, where the energy differences
    # the stopping test compares (~eps dt) approach the energy's own rounding, so the
    # stopping iteration of two FP64 implementations can differ; both still reach the
    # same chemical potential.
    tg, to, G = _pair("GPEtaskSpinor", 2, (32, 32), (6, 6), g=[[0.0, 0.0], [0.0, 0.0]])
    for t in (tg, to):
        t.coupling = 0.3 * np.exp(0.7j)
        t.Ntotal = 2.0
    pg, MUg = tg.groundstate_itp(0.02, 1e-9, nargout=2)
    po, MUo = to.groundstate_itp(0.02, 1e-9, nargout=2)
    print(f"eps 1e-9: GPU {tg.itp_iterations} it, oracle {to.itp_iterations} it")
    assert abs(tg.itp_iterations - to.itp_iterations) <= 0.1 * to.itp_iterations
    assert abs(MUg[-1] - MUo[-1]) < 1e-9
    assert _cmp_states(pg, po) < 1e-6


@pytest.mark.parametrize("dims,n,xmax", [(2, (32, 24), (6, 5)), (3, (16, 16, 12), (5, 5, 4))])
def test_2comp_itp_fixed_iterations(dims, n, xmax):
    tg, to, G = _pair("GPEtask2comp", dims, n, xmax)
    for t in (tg, to):
        t.N1, t.N2, t.M2, t.Ntotal = 2.0, 0.7, 1.7, 2.7
    tg.itp_max_iter_2c = 50
    with pytest.warns(UserWarning, match="Convergence not reached"):
        pg, MU1g, MU2g, EEg = tg.groundstate_itp(0.01, 1e-30, nargout=4)
    po, MU1o, MU2o, EEo = to.groundstate_itp(0.01, 1e-30, nargout=4, max_iter=50)
    err = _cmp_states(pg, po)
    print(f"2comp ITP {dims}D: state {err:.2e}")
    assert err < 1e-11
    for a, b in ((MU1g, MU1o), (MU2g, MU2o), (EEg, EEo)):
        assert np.allclose(a, b, rtol=1e-11, atol=1e-12)


def test_2comp_itp_converges_like_the_oracle():
    tg, to, G = _pair("GPEtask2comp", 2, (32, 32), (6, 6))
    for t in (tg, to):
        t.N1, t.N2, t.M2 = 1.5, 0.8, 2.0
    pg, MU1g, MU2g = tg.groundstate_itp(0.02, 1e-4, nargout=3)
    po, MU1o, MU2o = to.groundstate_itp(0.02, 1e-4, nargout=3)
    print(f"2comp converge: GPU {tg.itp_iterations} it, oracle {to.itp_iterations} it")
    assert tg.itp_iterations == to.itp_iterations
    assert _cmp_states(pg, po) < 1e-11
    assert np.allclose(MU1g, MU1o, rtol=1e-10, atol=1e-12) and np.allclose(MU2g, MU2o, rtol=1e-10, atol=1e-12)
    n1 = G.integrate(np.abs(pg[0]) ** 2)
    assert abs(n1 - 1.5) < 1e-12


def test_2comp_solve_split_matches_oracle():
    tg, to, G = _pair("GPEtask2comp", 2, (32, 24), (6, 5))
    rng = np.random.Generator(np.random.PCG64(3))
    r2 = G.mesh.x ** 2 + G.mesh.y ** 2
    psi = [np.exp(-r2 / 2) * np.exp(1j * rng.uniform(0, 0.3, G.size)), 0.6 * np.exp(-r2 / 3) + 0j]
    for t in (tg, to):
        t.Ntotal = 2.0
        t.init_state = [p.copy() for p in psi]
    pg, po = tg.solve_split(0.004, 20, 2), to.solve_split(0.004, 20, 2)
    assert _cmp_states(pg, po) < 1e-11
