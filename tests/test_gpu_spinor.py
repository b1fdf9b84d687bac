"""GPU parity of GPEtaskSpinor.solve_split (fused 2x2 Rabi step) against the oracle."""
import numpy as np
import pytest

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-11


def _spinor_pair(dims, n, xmax, coupling, g, gamma=0.0, nlincpl=0.0, seed=21):
    import gpe_hip
    if dims == 2:
        G, O = gpe_hip.grid2d(xmax[0], n[0], xmax[1], n[1]), oracle.grid2d(xmax[0], n[0], xmax[1], n[1])
    else:
        G = gpe_hip.grid3d(xmax[0], n[0], xmax[1], n[1], xmax[2], n[2])
        O = oracle.grid3d(xmax[0], n[0], xmax[1], n[1], xmax[2], n[2])

    def V(X, Y, Z):
        return 0.5 * (X ** 2 + Y ** 2 + (Z ** 2 if dims == 3 else 0))

    def V2(X, Y, Z):
        return V(X, Y, Z) + 0.3 * X

    rng = np.random.Generator(np.random.PCG64(seed))
    r2 = G.mesh.x ** 2 + G.mesh.y ** 2 + (G.mesh.z ** 2 if dims == 3 else 0)
    a = np.exp(-r2 / 3) * np.exp(1j * rng.uniform(0, 0.2, G.size))
    b = 0.7 * np.exp(-r2 / 2) * np.exp(1j * rng.uniform(0, 0.2, G.size))
    tg, to = gpe_hip.GPEtaskSpinor(G, [V, V2]), oracle.GPEtaskSpinor(O, [V, V2])
    cp = coupling(G) if callable(coupling) else coupling
    for t in (tg, to):
        t.g = np.array(g)
        t.gamma = gamma
        t.nlincpl = nlincpl
        t.coupling = cp
        t.Ntotal = 5.0
        t.init_state = [a.copy(), b.copy()]
    return tg, to


@pytest.mark.parametrize("dims,n,xmax", [(2, (32, 24), (6, 5)), (2, (64, 64), (8, 8)), (3, (16, 24, 12), (5, 6, 4)),
                                         (3, (48, 32, 32), (6, 5, 5))])
@pytest.mark.parametrize("kind", ["scalar", "array", "damped", "nlin"])
def test_spinor_parity(dims, n, xmax, kind):
    g = [[0.5, 0.45], [0.4, 0.6]]
    if kind == "scalar":
        tg, to = _spinor_pair(dims, n, xmax, 0.5 * np.exp(0.3j), g)
    elif kind == "array":
        tg, to = _spinor_pair(dims, n, xmax, lambda G: 0.5 * np.exp(2j * G.mesh.x), g)
    elif kind == "damped":
        tg, to = _spinor_pair(dims, n, xmax, lambda G: 0.5 * np.exp(2j * G.mesh.x), g, gamma=0.02)
    else:
        tg, to = _spinor_pair(dims, n, xmax, 0.4, g, nlincpl=0.05)
    pg = tg.solve_split(0.002, 20, 2)
    po = to.solve_split(0.002, 20, 2)
    err = max(rel_l2(pg[0], po[0]), rel_l2(pg[1], po[1]))
    print(f"spinor {dims}D {n} {kind}: rel-L2 = {err:.3e}")
    assert err < TOL
    assert np.allclose(tg.history["mu"], np.real(to.history["mu"]), rtol=1e-11, atol=1e-13)
    assert np.all(np.abs(np.imag(to.history["mu"])) < 1e-12 * np.abs(to.history["mu"]))
    assert np.allclose(tg.history["n"], to.history["n"], rtol=1e-11, atol=0)


def test_spinor_rabi_closed_form_gpu():
    """V=0, g=0, uniform real Omega: imbalance (n1-n2)/N = cos(2 Omega t) exactly."""
    import gpe_hip
    grid = gpe_hip.grid2d(5, 16, 5, 16)
    t = gpe_hip.GPEtaskSpinor(grid, [0 * grid.mesh.x, 0 * grid.mesh.x])
    t.g = np.zeros((2, 2))
    t.coupling = 0.5
    t.Ntotal = grid.integrate(np.ones(grid.size))
    t.init_state = [np.ones(grid.size, complex), np.zeros(grid.size, complex)]
    out = t.solve_split(0.01, 30, 4)
    n1 = grid.integrate(np.abs(out[0]) ** 2)
    n2 = grid.integrate(np.abs(out[1]) ** 2)
    assert abs((n1 - n2) / t.Ntotal - np.cos(2 * 0.5 * 1.2)) < 1e-12


def test_spinor_outer_graph_replays_equal_direct_launches(monkeypatch):
    # the two-component outer iteration is replayed from a captured graph; a coupling
    # changed by the callback is baked into a new capture
    import gpe_hip
    seen = {"graph": [], "direct": []}
    for mode in ("graph", "direct"):
        if mode == "direct":
            monkeypatch.setenv("GPE_NO_GRAPH", "1")
        G = gpe_hip.grid2d(6, 32, 5, 24)
        r2 = G.mesh.x ** 2 + G.mesh.y ** 2
        t = gpe_hip.GPEtaskSpinor(G, [0.5 * r2, 0.5 * r2 + 0.2 * G.mesh.y])
        t.g = np.array([[0.5, 0.45], [0.4, 0.6]])
        t.coupling = 0.3 * np.exp(0.5j)
        t.Ntotal = 4.0
        t.init_state = [np.exp(-r2 / 3) + 0j, 0.7 * np.exp(-r2 / 2) * np.exp(0.2j * G.mesh.x)]

        def cb(task, mode=mode):
            psi = task.current_state
            seen[mode].append((psi[0].copy(), psi[1].copy(), task.get_energy()))
            if task.current_iter == 2:
                task.coupling = 0.5 * np.exp(-0.2j)
        t.user_callback = cb
        t.solve_split(0.004, 20, 4)
    assert len(seen["graph"]) == len(seen["direct"]) == 4
    for a, b in zip(seen["graph"], seen["direct"]):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]


# ------------------------------------------------------- complex Omega closed forms
from closed_forms import grid_wavenumber, rabi_two_mode, two_mode_state, uniform_spinor_energy  # noqa: E402


def test_spinor_complex_omega_is_exp_minus_iHT_gpu():
    """Scalar complex Omega = |Omega| e^{i th} (the scalar real-dt Rabi kernel): psi1 =
    cos(|Omega| T), psi2 = -i e^{-i th} sin(|Omega| T), mu = 0 (GPEtaskSpinor.m:66-70)."""
    import gpe_hip
    grid = gpe_hip.grid2d(5, 16, 5, 16)
    t = gpe_hip.GPEtaskSpinor(grid, [0 * grid.mesh.x, 0 * grid.mesh.x])
    t.g = np.zeros((2, 2))
    r, th = 0.5, 0.7
    t.coupling = r * np.exp(1j * th)
    t.Ntotal = grid.integrate(np.ones(grid.size))
    t.init_state = [np.ones(grid.size, complex), np.zeros(grid.size, complex)]
    out = t.solve_split(0.01, 30, 4)
    T = 1.2
    assert rel_l2(out[0], np.cos(r * T) * np.ones(grid.size)) < 1e-12
    assert rel_l2(out[1], -1j * np.exp(-1j * th) * np.sin(r * T) * np.ones(grid.size)) < 1e-12
    assert np.all(np.abs(t.history["mu"]) < 1e-13)


@pytest.mark.parametrize("dims", [2, 3])
@pytest.mark.parametrize("r,m,th,gamma", [(0.5, 2, 0.7, 0.0), (0.5, -3, -2.1, 0.0), (2.0, 1, 0.3, 0.02),
                                          (500.0, 2, 0.4, 0.0)])
def test_spinor_space_dependent_omega_two_mode_gpu(dims, r, m, th, gamma):
    """Omega = r e^{i(qx+th)}: the GPU step equals the 2x2 two-mode reduction
    (tests/closed_forms.py) -- the array-Omega kernels with real dt (power-series Rabi
    factors for |Omega| dt small, r = 500 reaches |Omega| dt = 1 and the sincos form)
    and with the damped complex dt."""
    import gpe_hip
    if dims == 2:
        grid = gpe_hip.grid2d(5, 16, 4, 12)
        vol = 16 * 12 * grid.dx * grid.dy
    else:
        grid = gpe_hip.grid3d(5, 16, 4, 12, 3, 8)
        vol = 16 * 12 * 8 * grid.dx * grid.dy * grid.dz
    q = grid_wavenumber(16, 5, m)
    t = gpe_hip.GPEtaskSpinor(grid, [0 * grid.mesh.x, 0 * grid.mesh.x])
    t.g = np.zeros((2, 2))
    t.gamma = gamma
    t.coupling = r * np.exp(1j * (q * grid.mesh.x + th))
    a0, b0 = 0.8, 0.6j
    t.Ntotal = (abs(a0) ** 2 + abs(b0) ** 2) * vol
    t.init_state = list(two_mode_state(grid, q, th, a0, b0))
    ddt = 0.002
    out = t.solve_split(ddt, 20, 3)
    a, b, mus, ns = rabi_two_mode(r, q, ddt, 20, 3, 10, vol, t.Ntotal, gamma, a0, b0)
    e1, e2 = two_mode_state(grid, q, th, a, b)
    err = max(rel_l2(out[0], e1), rel_l2(out[1], e2))
    print(f"two-mode {dims}D r={r} m={m} gamma={gamma}: rel-L2 {err:.3e}")
    assert err < 1e-11
    assert np.allclose(t.history["mu"], np.real(mus), rtol=1e-11, atol=1e-11 * abs(mus).max())
    assert np.allclose(t.history["n"], ns, rtol=1e-11)


@pytest.mark.parametrize("dims,n,xmax", [(2, (32, 24), (6, 5)), (3, (16, 24, 12), (5, 6, 4))])
def test_spinor_parity_large_omega(dims, n, xmax):
    """|Omega| dt ~ 1 on a generic state (trap, interactions): array Omega of magnitude
    ~500 at dt = 0.002 takes the sincos Rabi factors instead of the power series."""
    g = [[0.5, 0.45], [0.4, 0.6]]
    tg, to = _spinor_pair(dims, n, xmax, lambda G: (450 + 50 * np.cos(G.mesh.y)) * np.exp(2j * G.mesh.x), g)
    pg = tg.solve_split(0.002, 20, 2)
    po = to.solve_split(0.002, 20, 2)
    err = max(rel_l2(pg[0], po[0]), rel_l2(pg[1], po[1]))
    print(f"spinor large Omega {dims}D {n}: rel-L2 = {err:.3e}")
    assert err < TOL
    assert np.allclose(tg.history["mu"], np.real(to.history["mu"]), rtol=1e-11, atol=1e-13)


def test_spinor_get_energy_uniform_closed_form_gpu():
    """(1/2) sum_ik g_ik n_i n_k / Vol + 2 Re(Omega conj(a) b) Vol for uniform components
    in V = 0 (GPEtaskSpinor.m:100-111): the halved g and the unhalved coupling."""
    import gpe_hip
    grid = gpe_hip.grid2d(5, 16, 4, 12)
    vol = 16 * 12 * grid.dx * grid.dy
    t = gpe_hip.GPEtaskSpinor(grid, [0 * grid.mesh.x, 0 * grid.mesh.x])
    gm = np.array([[0.5, 0.45], [0.4, 0.6]])
    t.g = gm
    om = 0.3 * np.exp(0.9j)
    t.coupling = om
    a, b = 0.7 * np.exp(0.2j), 0.4 * np.exp(-1.1j)
    phi = [a * np.ones(grid.size, complex), b * np.ones(grid.size, complex)]
    t.Ntotal = (abs(a) ** 2 + abs(b) ** 2) * vol
    E = uniform_spinor_energy(gm, abs(a) ** 2 * vol, abs(b) ** 2 * vol, om, a, b, vol)
    assert abs(t.get_energy(phi) / E - 1) < 1e-12


def test_spinor_array_coupling_mutated_in_place_is_uploaded():
    """A callback that changes an array coupling in place (task.coupling *= 2) is seen by
    the next outer iteration, as MATLAB re-reads task.coupling there
    (GPEtaskSpinor/solve_split.m:61); compared against the oracle doing the same."""
    import gpe_hip
    res = {}
    for name, mod in (("gpu", gpe_hip), ("oracle", oracle)):
        G = mod.grid2d(6, 32, 5, 24)
        r2 = G.mesh.x ** 2 + G.mesh.y ** 2
        t = mod.GPEtaskSpinor(G, [0.5 * r2, 0.5 * r2 + 0.2 * G.mesh.y])
        t.g = np.array([[0.5, 0.45], [0.4, 0.6]])
        t.coupling = 0.3 * np.exp(2j * G.mesh.x)
        t.Ntotal = 4.0
        t.init_state = [np.exp(-r2 / 3) + 0j, 0.7 * np.exp(-r2 / 2) * np.exp(0.2j * G.mesh.x)]

        def cb(task):
            if task.current_iter == 1:
                task.coupling *= 2.0
            elif task.current_iter == 2:
                task.coupling[:, : task.coupling.shape[1] // 2] = 0.1
        t.user_callback = cb
        res[name] = t.solve_split(0.004, 20, 4)
    err = max(rel_l2(res["gpu"][0], res["oracle"][0]), rel_l2(res["gpu"][1], res["oracle"][1]))
    assert err < TOL


@pytest.mark.parametrize("kind", ["array", "damped", "scalar"])
def test_spinor_blocks_of_an_outer_iteration_hand_over_the_x_spectrum(kind):
    """3D, 30 inner steps of n_recalc = 10: the block ends inside an outer iteration reduce
    the stored x spectrum and the next block opens from it (no real-space round trip);
    the first and last block of the iteration use the real-space state.  Against the
    oracle, with the mu / n histories."""
    g = [[0.5, 0.45], [0.4, 0.6]]
    dims, n, xmax = 3, (48, 32, 32), (6, 5, 5)
    if kind == "scalar":
        tg, to = _spinor_pair(dims, n, xmax, 0.5 * np.exp(0.3j), g, gamma=0.01)
    else:
        tg, to = _spinor_pair(dims, n, xmax, lambda G: 0.5 * np.exp(2j * G.mesh.x), g,
                              gamma=0.02 if kind == "damped" else 0.0)
    pg = tg.solve_split(0.002, 30, 2)
    po = to.solve_split(0.002, 30, 2)
    err = max(rel_l2(pg[0], po[0]), rel_l2(pg[1], po[1]))
    print(f"spinor 3 blocks per outer {kind}: rel-L2 = {err:.3e}", flush=True)
    assert err < TOL
    assert np.allclose(tg.history["mu"], np.real(to.history["mu"]), rtol=1e-11, atol=1e-13)
    assert np.allclose(tg.history["n"], to.history["n"], rtol=1e-11, atol=0)
