"""GPU parity of GPEtask.solve_split / groundstate_itp against the FP64 oracle.

Every test runs the HIP path through the C ABI (gpe_hip -> libgpehip.so) and the
oracle restatement on the same seeded inputs.  Tolerance: relative L2 <= 1e-11 on
the final state (the required bar), scalars to 1e-11 relative unless stated.
"""
import numpy as np
import pytest

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-11


def _pair(dims, n, xmax):
    import gpe_hip
    if dims == 2:
        return gpe_hip.grid2d(xmax[0], n[0], xmax[1], n[1]), oracle.grid2d(xmax[0], n[0], xmax[1], n[1])
    return (gpe_hip.grid3d(xmax[0], n[0], xmax[1], n[1], xmax[2], n[2]),
            oracle.grid3d(xmax[0], n[0], xmax[1], n[1], xmax[2], n[2]))


def _state(grid, seed, vortex=True):
    rng = np.random.Generator(np.random.PCG64(seed))
    r2 = grid.mesh.x ** 2 + grid.mesh.y ** 2 + (grid.mesh.z ** 2 if grid.ndims == 3 else 0)
    phi = np.exp(-r2 / 4) * np.exp(1j * rng.uniform(0, 0.1, grid.size))
    if vortex:
        phi = phi * np.exp(1j * np.arctan2(grid.mesh.x - 0.5, grid.mesh.y))
    return phi


def _tasks(dims, n, xmax, V, g=0.5, gamma=0.0, Ntotal=10.0, seed=3, **props):
    import gpe_hip
    G, O = _pair(dims, n, xmax)
    tg, to = gpe_hip.GPEtask(G, V), oracle.GPEtask(O, V)
    phi = _state(G, seed)
    for t in (tg, to):
        t.g, t.gamma, t.Ntotal = g, gamma, Ntotal
        for k, v in props.items():
            setattr(t, k, v)
        t.init_state = phi.copy()
    return tg, to


def harm2(X, Y, Z):
    return 0.5 * (X ** 2 + 2.0 * Y ** 2)


def harm3(X, Y, Z):
    return 0.5 * (X ** 2 + Y ** 2 + 4.0 * Z ** 2)


CASES_2D = [((16, 24), (6, 5)), ((32, 32), (8, 8)), ((64, 48), (8, 7)), ((12, 96), (5, 9)),
            ((128, 8), (9, 4)), ((256, 192), (10, 9)), ((384, 16), (12, 4)), ((512, 32), (14, 5))]


@pytest.mark.parametrize("n,xmax", CASES_2D)
@pytest.mark.parametrize("gamma", [0.0, 0.03])
def test_solve_split_2d_parity(n, xmax, gamma):
    tg, to = _tasks(2, n, xmax, harm2, gamma=gamma)
    pg = tg.solve_split(0.004, 20, 3)
    po = to.solve_split(0.004, 20, 3)
    err = rel_l2(pg, po)
    print(f"2D {n} gamma={gamma}: rel-L2 vs oracle = {err:.3e}")
    assert err < TOL
    assert np.allclose(tg.history["mu"], to.history["mu"], rtol=TOL, atol=0)
    assert np.allclose(tg.history["n"], to.history["n"], rtol=TOL, atol=0)
    assert tg.current_time == to.current_time and tg.current_iter == to.current_iter


CASES_3D = [((16, 16, 16), (5, 5, 5)), ((24, 16, 32), (6, 5, 7)), ((32, 48, 12), (7, 8, 4)),
            ((64, 64, 64), (8, 8, 8)), ((48, 96, 32), (7, 9, 6))]


@pytest.mark.parametrize("n,xmax", CASES_3D)
@pytest.mark.parametrize("gamma", [0.0, 0.02])
def test_solve_split_3d_parity(n, xmax, gamma):
    tg, to = _tasks(3, n, xmax, harm3, gamma=gamma)
    pg = tg.solve_split(0.002, 20, 2)
    po = to.solve_split(0.002, 20, 2)
    err = rel_l2(pg, po)
    print(f"3D {n} gamma={gamma}: rel-L2 vs oracle = {err:.3e}")
    assert err < TOL
    assert np.allclose(tg.history["mu"], to.history["mu"], rtol=TOL, atol=0)
    assert np.allclose(tg.history["n"], to.history["n"], rtol=TOL, atol=0)


def test_solve_split_options_parity():
    """n_rec fallback, decay_rate, mu_init, and resume across calls (example2 pattern)."""
    for props, ni in [({"decay_rate": 0.5}, 20), ({"mu_init": 1.7}, 20), ({}, 7), ({"n_recalc": 5}, 15)]:
        tg, to = _tasks(2, (32, 24), (7, 6), harm2, gamma=0.05, **props)
        for no in (1, 2, 3):                       # solve_split(tstep, steps_int, i)
            pg = tg.solve_split(0.004, ni, no)
            po = to.solve_split(0.004, ni, no)
        err = rel_l2(pg, po)
        print(f"options {props} ni={ni}: rel-L2 = {err:.3e}")
        assert err < TOL
        assert np.allclose(tg.history["mu"], to.history["mu"], rtol=TOL, atol=0)


def test_solve_split_without_outer_iterations_matches_oracle():
    """solve_split.m:34-38, 49, 84: with niter_outer <= current_iter the loop runs no
    iteration and phi is returned as it started: init_state on a fresh task, the
    current state on a resumed one (history untouched)."""
    tg, to = _tasks(2, (32, 24), (7, 6), harm2, gamma=0.05)
    pg, po = tg.solve_split(0.004, 10, 0), to.solve_split(0.004, 10, 0)
    assert np.array_equal(pg, po) and np.array_equal(pg, to.init_state)
    assert tg.current_iter == to.current_iter == 0
    tg.solve_split(0.004, 10, 2)
    to.solve_split(0.004, 10, 2)
    pg, po = tg.solve_split(0.004, 10, 2), to.solve_split(0.004, 10, 2)   # nothing left to run
    assert rel_l2(pg, po) < TOL
    assert tg.current_iter == to.current_iter == 2
    assert len(tg.history["mu"]) == len(to.history["mu"]) == 2


def test_solve_split_time_dependent_potential_parity():
    """Vtd evaluated per block at time2 (GPEtask.m:65-67, solve_split.m:54)."""
    def vtd(task, t):
        return 0.3 * np.sin(2.0 * t) * task.grid.mesh.x

    tg, to = _tasks(2, (32, 32), (8, 8), harm2, gamma=0.01)
    tg.Vtd = vtd
    to.Vtd = vtd
    pg = tg.solve_split(0.004, 20, 3)
    po = to.solve_split(0.004, 20, 3)
    assert rel_l2(pg, po) < TOL
    assert np.allclose(tg.history["mu"], to.history["mu"], rtol=TOL, atol=0)


def test_linear_closed_form_on_gpu():
    """g=0, V=c: two plane waves acquire e^{-i(E_j-mu)t}; pins the whole step independently of the oracle."""
    import gpe_hip
    grid = gpe_hip.grid2d(6, 16, 5, 12)
    c = 0.7
    t = gpe_hip.GPEtask(grid, c + 0 * grid.mesh.x)
    kx1 = 2 * np.pi / (grid.x[-1] - grid.x[0] + grid.dx)
    ky1 = 2 * np.pi / (grid.y[-1] - grid.y[0] + grid.dy)
    k1, k2 = (2 * kx1, 1 * ky1), (-3 * kx1, 3 * ky1)
    w1 = np.exp(1j * (k1[0] * grid.mesh.x + k1[1] * grid.mesh.y))
    w2 = np.exp(1j * (k2[0] * grid.mesh.x + k2[1] * grid.mesh.y))
    phi = 0.8 * w1 + 0.6 * w2
    t.Ntotal = grid.integrate(np.abs(phi) ** 2)
    t.init_state = phi
    out = t.solve_split(0.01, 20, 3)
    E1 = (k1[0] ** 2 + k1[1] ** 2) / 2 + c
    E2 = (k2[0] ** 2 + k2[1] ** 2) / 2 + c
    mu = 0.64 * E1 + 0.36 * E2
    T = 0.6
    expect = 0.8 * np.exp(-1j * (E1 - mu) * T) * w1 + 0.6 * np.exp(-1j * (E2 - mu) * T) * w2
    assert rel_l2(out, expect) < 1e-12
    assert abs(t.history["mu"][-1] / mu - 1) < 1e-12


@pytest.mark.parametrize("dims,n,xmax", [(2, (64, 64), (8, 8)), (3, (32, 32, 24), (6, 6, 5))])
def test_itp_fixed_iterations_parity(dims, n, xmax):
    """groundstate_itp at a fixed iteration count (itp_min_iter = itp_max_iter)."""
    import gpe_hip
    G, O = _pair(dims, n, xmax)
    V = harm2 if dims == 2 else harm3
    tg, to = gpe_hip.GPEtask(G, V, seed=5), oracle.GPEtask(O, V, seed=5)
    for t in (tg, to):
        t.g, t.Ntotal = 0.8, 50.0
        t.itp_min_iter = t.itp_max_iter = 60
    with pytest.warns(UserWarning):
        pg, MUg = tg.groundstate_itp(0.01, 1e-12, "rand", nargout=2)
    with pytest.warns(UserWarning):
        po, MUo = to.groundstate_itp(0.01, 1e-12, "rand", nargout=2)
    err = rel_l2(pg, po)
    print(f"ITP {n}: rel-L2 = {err:.3e}")
    assert err < TOL
    assert np.allclose(MUg, MUo, rtol=1e-10, atol=1e-12)


def test_itp_converged_parity():
    """groundstate_itp to convergence with MU, MU2, EE histories (nargout = 4)."""
    import gpe_hip
    G, O = _pair(2, (64, 64), (8, 8))
    tg, to = gpe_hip.GPEtask(G, harm2, seed=9), oracle.GPEtask(O, harm2, seed=9)
    for t in (tg, to):
        t.g, t.Ntotal = 0.5, 20.0
    pg, MUg, MU2g, EEg = tg.groundstate_itp(0.01, 1e-6, "tf", nargout=4)
    po, MUo, MU2o, EEo = to.groundstate_itp(0.01, 1e-6, "tf", nargout=4)
    assert tg.itp_iterations == to.itp_iterations
    assert rel_l2(pg, po) < TOL
    assert np.allclose(MUg, MUo, rtol=1e-10, atol=1e-12)
    assert np.allclose(MU2g, MU2o, rtol=1e-10, atol=1e-12)
    assert np.allclose(EEg, EEo, rtol=1e-10, atol=1e-12)


def test_energy_measure_parity():
    import gpe_hip
    G, O = _pair(3, (32, 24, 16), (6, 5, 4))
    phi = _state(G, 11)
    tg, to = gpe_hip.GPEtask(G, harm3), oracle.GPEtask(O, harm3)
    tg.g = to.g = 0.7
    eg = tg.get_energy(phi)
    eo = to.get_energy(phi)
    assert abs(eg / eo - 1) < 1e-12


def test_determinism_bitwise():
    """Two identical runs give bit-identical states (fixed-order reductions)."""
    outs = []
    for _ in range(2):
        tg, _o = _tasks(3, (32, 32, 32), (6, 6, 6), harm3, gamma=0.01)
        outs.append(tg.solve_split(0.002, 20, 2))
    assert np.array_equal(outs[0], outs[1])


def test_device_phase_accuracy():
    """Device exp/sincos (Cody-Waite + Taylor) against numpy: <= 2 ulp-scale error."""
    from gpe_hip._lib import selftest_phase
    rng = np.random.Generator(np.random.PCG64(17))
    ei = np.concatenate([rng.uniform(-10, 10, 200000), rng.uniform(-1e5, 1e5, 50000),
                         np.array([0.0, np.pi / 2, -np.pi, 1e-300, 5e-324, 3.0 * np.pi / 4])])
    er = np.concatenate([rng.uniform(-0.1, 0.1, 100000), rng.uniform(-700, 700, 150000), np.zeros(6)])
    out = selftest_phase(er, ei)
    ref = np.exp(er) * (np.cos(ei) + 1j * np.sin(ei))
    scale = np.exp(er)
    err = np.abs(out - ref) / scale
    print(f"phase max abs err / e^x: {err.max():.3e}")
    assert err.max() < 1e-15
    # NaN/Inf arguments propagate as NaN; a hugely negative real part underflows to 0
    sp = selftest_phase(np.array([np.nan, 0.0, -1000.0]), np.array([0.0, np.inf, 1.0]))
    assert np.isnan(sp[0].real) and np.isnan(sp[1].real) and sp[2] == 0


@pytest.mark.parametrize("dims,n,xmax", [(2, (64, 48), (8, 7)), (3, (24, 16, 32), (6, 5, 7))])
def test_outer_graph_replays_equal_direct_launches(dims, n, xmax, monkeypatch):
    # gpe_split_outer captures one outer iteration as a graph and replays it; every
    # replay (with eager work in the callback between them) equals the direct launches
    seen = {"graph": [], "direct": []}
    for mode in ("graph", "direct"):
        if mode == "direct":
            monkeypatch.setenv("GPE_NO_GRAPH", "1")
        tg, _ = _tasks(dims, n, xmax, harm2 if dims == 2 else harm3, gamma=0.02)
        tg.user_callback = lambda task: seen[mode].append((task.current_state.copy(), task.get_energy()))  # noqa: B023
        tg.solve_split(0.004, 20, 4)
    assert len(seen["graph"]) == len(seen["direct"]) == 4
    for (a, ea), (b, eb) in zip(seen["graph"], seen["direct"]):
        assert np.array_equal(a, b) and ea == eb


@pytest.mark.parametrize("dims", [2, 3])
def test_get_energy_halves_g_uniform_state_gpu(dims):
    """Uniform psi of norm N in V = 0: get_energy = g N^2/(2 Vol) and the chemical-potential
    integral g N^2/Vol (GPEtask.m:97-107 halves g; the kinetic term of a constant is 0)."""
    import gpe_hip
    from closed_forms import uniform_energy
    if dims == 2:
        grid = gpe_hip.grid2d(5, 16, 4, 12)
        vol = 16 * 12 * grid.dx * grid.dy
    else:
        grid = gpe_hip.grid3d(5, 16, 4, 12, 3, 8)
        vol = 16 * 12 * 8 * grid.dx * grid.dy * grid.dz
    t = gpe_hip.GPEtask(grid, 0 * grid.mesh.x)
    t.g = 0.7
    N = 3.0
    t.Ntotal = N
    psi = np.sqrt(N / vol) * np.exp(0.3j) * np.ones(grid.size)
    E, M = uniform_energy(t.g, N, vol)
    assert abs(t.get_energy(psi) / E - 1) < 1e-12
    assert abs(t.chemical_potential_integral(psi) / M - 1) < 1e-12


def test_get_energy_interaction_identity_gpu():
    """get_energy - <psi|H psi> = -(g/2) sum |psi|^4 w for a generic state at g != 0."""
    import gpe_hip
    grid = gpe_hip.grid2d(6, 32, 5, 24)
    t = gpe_hip.GPEtask(grid, lambda X, Y, Z: 0.5 * (X ** 2 + 2 * Y ** 2))
    t.g = 0.9
    t.Ntotal = 1.0
    rng = np.random.Generator(np.random.PCG64(5))
    psi = np.exp(-(grid.mesh.x ** 2 + grid.mesh.y ** 2) / 3) * np.exp(1j * rng.uniform(0, 1, grid.size))
    w = grid.dx * grid.dy
    d = t.get_energy(psi) - t.chemical_potential_integral(psi)
    assert abs(d / (-0.5 * t.g * np.sum(np.abs(psi) ** 4) * w) - 1) < 1e-11
