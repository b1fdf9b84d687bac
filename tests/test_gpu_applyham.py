"""Array-valued Hamiltonian on the device (gpe_applyham): GPEtask.applyham / applyh0
(GPEtask.m:73-95) and GPEtaskSpinor.applyham / applyham1c / applyh0 (GPEtaskSpinor.m:45-86)
against the oracle, plus the closed forms the oracle is pinned by.

The device sums the kinetic term axis by axis (ifft_a(k_a^2/2 fft_a)), the oracle takes
ifftn(kk .* fftn); the two agree to rounding.  Tolerance: relative L2 1e-12."""
import numpy as np
import pytest

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _grids(dims):
    import gpe_hip
    if dims == 2:
        return gpe_hip.grid2d(6, 64, 5, 48), oracle.grid2d(6, 64, 5, 48)
    return gpe_hip.grid3d(6, 32, 5, 24, 4, 16), oracle.grid3d(6, 32, 5, 24, 4, 16)


def _state(grid, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    M = grid.mesh
    r2 = M.x ** 2 + M.y ** 2 + (M.z ** 2 if grid.ndims == 3 else 0)
    amp = np.exp(-r2 / 6) * (1 + 0.2 * rng.normal(size=grid.size))
    return amp * np.exp(1j * (0.4 * M.x - 0.3 * M.y + rng.uniform(0, 0.5, grid.size)))


def _trap(grid):
    M = grid.mesh
    return 0.5 * (M.x ** 2 + 1.5 * M.y ** 2 + (2.0 * M.z ** 2 if grid.ndims == 3 else 0))


@pytest.mark.parametrize("dims", [2, 3])
@pytest.mark.parametrize("omega", [0.0, 0.35])
def test_applyham_matches_oracle(dims, omega):
    import gpe_hip
    G, O = _grids(dims)
    tg, to = gpe_hip.GPEtask(G, _trap(G)), oracle.GPEtask(O, _trap(O))
    for t in (tg, to):
        t.g, t.omega = 0.8, omega
    psi = _state(G, 7 + dims)
    print(f"applyham dims={dims} omega={omega}", flush=True)
    assert rel_l2(tg.applyham(psi), to.applyham(psi)) < TOL
    assert rel_l2(tg.applyh0(psi), to.applyh0(psi)) < TOL


@pytest.mark.parametrize("dims", [2, 3])
def test_applyham_plane_wave_closed_form(dims):
    import gpe_hip
    G, O = _grids(dims)
    ks = (O.kx1[5], O.ky1[-7]) + ((O.kz1[O.nz // 2],) if dims == 3 else ())
    arg = ks[0] * G.mesh.x + ks[1] * G.mesh.y + (ks[2] * G.mesh.z if dims == 3 else 0)
    A, c = 0.8 * np.exp(0.4j), 0.35
    psi = A * np.exp(1j * arg)
    kk = 0.5 * sum(k * k for k in ks)
    t = gpe_hip.GPEtask(G, c + 0 * G.mesh.x)
    t.g = 1.3
    assert rel_l2(t.applyh0(psi), (kk + c) * psi) < TOL
    assert rel_l2(t.applyham(psi), (kk + c + t.g * abs(A) ** 2) * psi) < TOL


def test_applyham_of_current_state_after_a_damped_run():
    # gamma > 0 leaves a renormalisation pending on the device (solve_split.m:72); the
    # current state the operator sees includes it
    import gpe_hip
    G, O = _grids(2)
    tg, to = gpe_hip.GPEtask(G, _trap(G)), oracle.GPEtask(O, _trap(O))
    psi = _state(G, 3)
    for t in (tg, to):
        t.g, t.gamma, t.Ntotal = 0.6, 0.05, 2.0
        t.init_state = psi.copy()
    pg, po = tg.solve_split(0.004, 20, 3), to.solve_split(0.004, 20, 3)
    assert rel_l2(pg, po) < 1e-11
    assert rel_l2(tg.applyham(), to.applyham(po)) < 1e-11
    assert rel_l2(tg.applyh0(), to.applyh0(po)) < 1e-11
    # the run goes on unchanged after the call
    assert rel_l2(tg.solve_split(0.004, 20, 1), to.solve_split(0.004, 20, 1)) < 1e-11


def test_applyham_inside_callback_with_time_dependent_potential():
    import gpe_hip
    G, O = _grids(2)
    X = G.mesh.x
    vtd = lambda task, t: 0.3 * np.sin(5 * t) * X    # noqa: E731
    tg, to = gpe_hip.GPEtask(G, _trap(G)), oracle.GPEtask(O, _trap(O))
    rec = {id(tg): [], id(to): []}
    psi = _state(G, 11)
    for t in (tg, to):
        t.g, t.Ntotal, t.Vtd = 0.7, 3.0, vtd
        t.init_state = psi.copy()
        t.user_callback = lambda task: rec[id(task)].append(
            (task.applyham(task.current_state) if task is to else task.applyham(),
             task.applyh0(task.current_state, 0.01)))
    pg, po = tg.solve_split(0.004, 20, 3), to.solve_split(0.004, 20, 3)
    assert rel_l2(pg, po) < 1e-11
    for (hg, h0g), (ho, h0o) in zip(rec[id(tg)], rec[id(to)]):
        assert rel_l2(hg, ho) < 1e-11
        assert rel_l2(h0g, h0o) < 1e-11
    # the run with the callback ends where the run without it does, bit for bit
    tq = gpe_hip.GPEtask(G, _trap(G))
    tq.g, tq.Ntotal, tq.Vtd = 0.7, 3.0, vtd
    tq.init_state = psi.copy()
    assert np.array_equal(tq.solve_split(0.004, 20, 3), pg)


@pytest.mark.parametrize("array_coupling", [False, True])
def test_spinor_applyham_matches_oracle(array_coupling):
    import gpe_hip
    G, O = _grids(3)
    V = _trap(G)
    tg = gpe_hip.GPEtaskSpinor(G, [V, V + 0.1 * G.mesh.y])
    to = oracle.GPEtaskSpinor(O, [V, V + 0.1 * O.mesh.y])
    cpl = 0.4 * np.exp(2j * G.mesh.x) if array_coupling else 0.3 * np.exp(0.9j)
    for t in (tg, to):
        t.g = np.array([[0.5, 0.45], [0.3, 0.6]])
        t.coupling = cpl
    phi = [_state(G, 21), 0.6 * _state(G, 22)]
    hg, ho = tg.applyham(phi), to.applyham(phi)
    for j in (0, 1):
        assert rel_l2(hg[j], ho[j]) < TOL
        assert rel_l2(tg.applyham1c(phi, j), to.applyham1c(phi, j)) < TOL
        # applyh0(phi, j) = lap(phi) + V_j phi (GPEtaskSpinor.m:45-54)
        want = O.lap(phi[j]) + to.getVtotal(0)[j] * phi[j]
        assert rel_l2(tg.applyh0(phi[j], j), want) < TOL


def test_spinor_applyham_plane_wave_closed_form():
    import gpe_hip
    G, O = _grids(2)
    k = (O.kx1[2], O.ky1[3])
    e = np.exp(1j * (k[0] * G.mesh.x + k[1] * G.mesh.y))
    kk = 0.5 * (k[0] ** 2 + k[1] ** 2)
    t = gpe_hip.GPEtaskSpinor(G, [0.2 + 0 * G.mesh.x, -0.1 + 0 * G.mesh.x])
    gm = np.array([[0.5, 0.45], [0.3, 0.6]])
    om = 0.3 * np.exp(0.9j)
    t.g, t.coupling = gm, om
    a, b = 0.7 * np.exp(0.2j), 0.4 * np.exp(-1.1j)
    h1, h2 = t.applyham([a * e, b * e])
    n = np.array([abs(a) ** 2, abs(b) ** 2])
    assert rel_l2(h1, ((kk + 0.2 + gm[0] @ n) * a + om * b) * e) < TOL
    assert rel_l2(h2, ((kk - 0.1 + gm[1] @ n) * b + np.conj(om) * a) * e) < TOL
