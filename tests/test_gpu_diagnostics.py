"""Device vortex-core detection (gpe_detect_core / gpe_detect_core_array, the drop-in for
lib/util/detect_core.m) and reductions inside user_callback, against the oracle.

Plaquette centres (refine = 0) are compared exactly; refined positions (the closed-form
2x2 generalised eigenproblem on the device, LAPACK's QZ in the oracle) to 1e-10."""
import numpy as np
import pytest

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu

REF_TOL = 1e-10


def _vortex_field(X, Y, cores, width, xi=0.5, noise=0.0, seed=0):
    phi = np.exp(-(X ** 2 + Y ** 2) / (2 * width ** 2)).astype(np.complex128)
    for xc, yc, q in cores:
        z = (X - xc) + q * 1j * (Y - yc)
        phi *= z / np.sqrt(np.abs(z) ** 2 + xi ** 2)
    if noise:
        rng = np.random.Generator(np.random.PCG64(seed))
        phi += noise * (rng.normal(size=X.shape) + 1j * rng.normal(size=X.shape))
    return phi


def _same_cores(got, want, label):
    for k, (g, w) in enumerate(zip(got, want)):
        assert g.shape == w.shape, f"{label}: {['positive', 'negative'][k]} count {g.shape[0]} vs {w.shape[0]}"
    return got, want


@pytest.mark.parametrize("ny,nx", [(64, 64), (100, 130), (256, 256), (12, 512)])
@pytest.mark.parametrize("noise", [0.0, 0.02])
def test_detect_core_array_matches_oracle(ny, nx, noise):
    import gpe_hip
    rng = np.random.Generator(np.random.PCG64(ny * 1000 + nx))
    x = np.linspace(-8, 8, nx + 1)[:nx]
    y = np.linspace(-6, 6, ny + 1)[:ny]
    X, Y = np.meshgrid(x, y)
    cores = [(rng.uniform(-5, 5), rng.uniform(-4, 4), rng.choice([-1, 1])) for _ in range(9)]
    phi = _vortex_field(X, Y, cores, 5.0, noise=noise, seed=ny + nx)
    for refine in (0, 1):
        got = gpe_hip.detect_core(phi, X, Y, refine)
        want = oracle.detect_core(phi, X, Y, refine)
        _same_cores(got, want, f"{ny}x{nx} refine={refine}")
        print(f"detect_core {ny}x{nx} noise={noise} refine={refine}: {got[0].shape[0]}+/{got[1].shape[0]}-")
        for g, w in zip(got, want):
            if refine:
                assert np.allclose(g, w, rtol=0, atol=REF_TOL)
            else:
                assert np.array_equal(g, w)


def test_detect_core_many_cores_grow_the_output():
    # a random-phase field has hundreds of cores: more than the first output capacity
    import gpe_hip
    rng = np.random.Generator(np.random.PCG64(5))
    x = np.linspace(-8, 8, 129)[:128]
    X, Y = np.meshgrid(x, x)
    k = np.fft.fft2(rng.normal(size=(128, 128)) + 1j * rng.normal(size=(128, 128)))
    kk = np.fft.fftfreq(128)
    k *= np.exp(-(kk[:, None] ** 2 + kk[None, :] ** 2) / (2 * 0.08 ** 2))
    phi = np.fft.ifft2(k)
    for refine in (0, 1):
        got, want = gpe_hip.detect_core(phi, X, Y, refine), oracle.detect_core(phi, X, Y, refine)
        _same_cores(got, want, f"random phase refine={refine}")
        assert got[0].shape[0] > 64 and got[1].shape[0] > 64
        for g, w in zip(got, want):
            assert np.allclose(g, w, rtol=0, atol=REF_TOL if refine else 0)


def _vortex_tasks(dims=2, gamma=0.01):
    import gpe_hip
    if dims == 2:
        G, O = gpe_hip.grid2d(8, 128, 8, 128), oracle.grid2d(8, 128, 8, 128)
    else:
        G, O = gpe_hip.grid3d(7, 48, 7, 48, 4, 16), oracle.grid3d(7, 48, 7, 48, 4, 16)
    X, Y = G.mesh.x, G.mesh.y
    r2 = X ** 2 + Y ** 2 + (G.mesh.z ** 2 if dims == 3 else 0)
    V = 0.5 * r2
    phi = np.exp(-r2 / 8).astype(np.complex128)
    for xc, yc, sgn in ((2.0, 0.3, 1), (-0.4, 2.2, -1), (-1.7, -1.9, 1)):
        rr = np.sqrt((X - xc) ** 2 + (Y - yc) ** 2)
        phi *= np.tanh(rr / 0.6) * np.exp(1j * sgn * np.arctan2(Y - yc, X - xc))   # as examples/example2
    tg, to = gpe_hip.GPEtask(G, V), oracle.GPEtask(O, V)
    for t in (tg, to):
        t.g, t.gamma, t.Ntotal = 2.0, gamma, 50.0
        t.init_state = phi.copy()
    return tg, to


def test_callback_diagnostics_match_oracle_and_leave_the_run_untouched():
    # examples/example2 post_process pattern: detect_core + get_energy at every outer step
    tg, to = _vortex_tasks()
    rec_g, rec_o = [], []

    def cb_gpu(task):
        rec_g.append((task.get_energy(), task.chemical_potential_integral(), task.detect_core(refine=1),
                      task.detect_core(refine=0)))

    def cb_oracle(task):
        phi = task.current_state
        X, Y = task.grid.mesh.x, task.grid.mesh.y
        rec_o.append((task.get_energy(), None, oracle.detect_core(phi, X, Y, 1), oracle.detect_core(phi, X, Y, 0)))

    tg.user_callback, to.user_callback = cb_gpu, cb_oracle
    pg, po = tg.solve_split(0.004, 20, 4), to.solve_split(0.004, 20, 4)
    assert rel_l2(pg, po) < 1e-11
    assert len(rec_g) == len(rec_o) == 4
    for step, (g, o) in enumerate(zip(rec_g, rec_o)):
        assert abs(g[0] - o[0]) <= 1e-11 * abs(o[0]), f"energy at outer step {step + 1}"
        _same_cores(g[2], o[2], f"step {step + 1}")
        assert g[2][0].shape[0] >= 1 and g[2][1].shape[0] >= 1
        for a, b in zip(g[2], o[2]):
            assert np.allclose(a, b, rtol=0, atol=REF_TOL)
        for a, b in zip(g[3], o[3]):
            assert np.array_equal(a, b)
    # the same run without the callback ends in the same state, bit for bit
    tq, _ = _vortex_tasks()
    assert np.array_equal(tq.solve_split(0.004, 20, 4), pg)
    assert np.array_equal(tq.history["mu"], tg.history["mu"])


def test_detect_core_on_3d_slices():
    tg, to = _vortex_tasks(dims=3)
    pg, po = tg.solve_split(0.004, 10, 2), to.solve_split(0.004, 10, 2)
    assert rel_l2(pg, po) < 1e-11
    M = to.grid.mesh
    for k in (0, 5, 8, 15):
        for refine in (0, 1):
            got = tg.detect_core(refine=refine, zslice=k)
            want = oracle.detect_core(po[:, :, k], M.x[:, :, k], M.y[:, :, k], refine)
            _same_cores(got, want, f"z slice {k} refine={refine}")
            for a, b in zip(got, want):
                assert np.allclose(a, b, rtol=0, atol=REF_TOL if refine else 0)
    with pytest.raises(ValueError, match="z slice"):
        tg.detect_core()


def test_spinor_callback_energy_with_time_dependent_potential():
    # a host Vtd keeps a block-end potential on the device; get_energy inside the
    # callback reads the step potential at current_time, as the reference's does
    import gpe_hip
    G, O = gpe_hip.grid2d(6, 64, 6, 64), oracle.grid2d(6, 64, 6, 64)
    X, Y = G.mesh.x, G.mesh.y
    V = 0.5 * (X ** 2 + Y ** 2)
    r2 = X ** 2 + Y ** 2
    psi = [np.exp(-r2 / 3) * np.exp(1j * np.arctan2(Y - 0.3, X - 0.2)), 0.7 * np.exp(-r2 / 2) + 0j]
    vtd = lambda task, t: 0.3 * np.sin(5 * t) * X    # noqa: E731
    tg, to = gpe_hip.GPEtaskSpinor(G, [V, V + 0.1 * Y]), oracle.GPEtaskSpinor(O, [V, V + 0.1 * Y])
    rec = {id(tg): [], id(to): []}
    for t in (tg, to):
        t.g = np.array([[0.5, 0.45], [0.4, 0.6]])
        t.coupling = 0.3 * np.exp(0.5j)
        t.Ntotal = 4.0
        t.init_state = [p.copy() for p in psi]
        t.Vtd = vtd
        t.user_callback = lambda task: rec[id(task)].append((task.get_energy(), task.get_energy(task.current_state)))
    pg, po = tg.solve_split(0.004, 20, 3), to.solve_split(0.004, 20, 3)
    assert max(rel_l2(pg[0], po[0]), rel_l2(pg[1], po[1])) < 1e-11
    assert np.allclose(np.array(rec[id(tg)]), np.array(rec[id(to)]), rtol=1e-11, atol=0)
    # the run without a callback ends in the same state, bit for bit
    tq = gpe_hip.GPEtaskSpinor(G, [V, V + 0.1 * Y])
    tq.g, tq.coupling, tq.Ntotal, tq.Vtd = tg.g, tg.coupling, 4.0, vtd
    tq.init_state = [p.copy() for p in psi]
    pq = tq.solve_split(0.004, 20, 3)
    assert np.array_equal(pq[0], pg[0]) and np.array_equal(pq[1], pg[1])
    got = tg.detect_core(refine=1, comp=0)
    want = oracle.detect_core(po[0], O.mesh.x, O.mesh.y, 1)
    _same_cores(got, want, "spinor component 0")
    for a, b in zip(got, want):
        assert np.allclose(a, b, rtol=0, atol=REF_TOL)


def test_detect_core_rejects_bad_input():
    import gpe_hip
    x = np.linspace(-1, 1, 9)
    X, Y = np.meshgrid(x, x)
    with pytest.raises(gpe_hip.GPELibraryError, match="even sizes"):
        gpe_hip.detect_core(np.ones((9, 9), complex), X, Y)
    with pytest.raises(ValueError, match="meshgrid"):
        gpe_hip.detect_core(np.ones((9, 9), complex), Y, X)


@pytest.mark.parametrize("Ntotal", [0.0, 50.0])
def test_energies_of_other_states_and_times_inside_callback_keep_the_run(Ntotal):
    # get_energy(phi) and get_energy(time=t) inside user_callback upload what they need
    # and put the running block's state and potential back; Ntotal = 0 takes the norm of
    # the initial state (solve_split.m:17-21), which must not reset the run either
    def make():
        tg, to = _vortex_tasks()
        X = tg.grid.mesh.x
        for t in (tg, to):
            t.Ntotal = Ntotal
            t.Vtd = lambda task, tt: 0.3 * np.sin(3.0 * tt) * X    # noqa: B023
        return tg, to

    tg, to = make()
    rec = {id(tg): [], id(to): []}
    for t in (tg, to):
        t.user_callback = lambda task: rec[id(task)].append(
            (task.get_energy(task.current_state), task.get_energy(None, 0.0), task.get_energy()))
    pg, po = tg.solve_split(0.004, 20, 3), to.solve_split(0.004, 20, 3)
    assert rel_l2(pg, po) < 1e-11
    assert np.allclose(np.array(rec[id(tg)]), np.array(rec[id(to)]), rtol=1e-11, atol=0)
    tq, _ = make()
    assert np.array_equal(tq.solve_split(0.004, 20, 3), pg)
