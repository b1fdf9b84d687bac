"""Plan tuning (gpe_tune, gpe_batch_tune) and every FFT plan variant against the oracle.

The reference's GPEtask constructor plans its FFTs by measurement (GPEtask.m:28,
55-58).  Here lengths 64, 128, 256, 384 and 512 have two plans; whichever the tuner
records, solve_split must match the FP64 oracle to 1e-11, and the ensemble must
reproduce single-task runs that hold the same plans bit for bit.
"""
import numpy as np
import pytest

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def fresh_wisdom():
    from gpe_hip import _lib
    _lib.wisdom_forget()
    yield
    _lib.wisdom_forget()


def _run_pair(gpe_hip, dims, variant, steps=(0.002, 10, 2), gamma=0.01):
    from gpe_hip import _lib
    nd = len(dims)
    for a, n in enumerate(dims):
        if variant >= 0 and _lib.wisdom_get(nd, a, n) < 0 and n in (64, 128, 256, 384, 512):
            _lib.wisdom_set(nd, a, n, variant)
    if nd == 2:
        G, O = gpe_hip.grid2d(7.0, dims[0], 6.0, dims[1]), oracle.grid2d(7.0, dims[0], 6.0, dims[1])

        def V(X, Y, Z):
            return 0.5 * (X ** 2 + 1.3 * Y ** 2)
    else:
        G = gpe_hip.grid3d(6.0, dims[0], 5.0, dims[1], 5.5, dims[2])
        O = oracle.grid3d(6.0, dims[0], 5.0, dims[1], 5.5, dims[2])

        def V(X, Y, Z):
            return 0.5 * (X ** 2 + Y ** 2 + 2 * Z ** 2)
    rng = np.random.Generator(np.random.PCG64(sum(dims) + max(variant, 0)))
    r2 = G.mesh.x ** 2 + G.mesh.y ** 2 + (G.mesh.z ** 2 if nd == 3 else 0)
    phi = np.exp(-r2 / 3) * np.exp(1j * rng.uniform(0, 0.3, G.size))
    phi = phi * np.exp(1j * np.arctan2(G.mesh.x - 0.3, G.mesh.y))
    tg, to = gpe_hip.GPEtask(G, V), oracle.GPEtask(O, V)
    for t in (tg, to):
        t.g, t.gamma, t.Ntotal = 0.7, gamma, 10.0
        t.init_state = phi
    pg = tg.solve_split(*steps)
    po = to.solve_split(*steps)
    plans = tg._device().plans()
    return rel_l2(pg, po), plans, tg, to


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("dims", [(64, 128), (128, 64), (256, 384), (384, 256), (512, 64), (64, 512)])
def test_every_2d_plan_variant_matches_oracle(dims, variant):
    import gpe_hip
    err, plans, tg, to = _run_pair(gpe_hip, dims, variant)
    print(f"2D {dims} variant {variant}: plans {plans} rel-L2 {err:.3e}", flush=True)
    assert plans == [variant, variant]
    assert err < 1e-11
    assert np.allclose(tg.history["mu"], to.history["mu"], rtol=1e-11, atol=0)
    assert np.allclose(tg.history["n"], to.history["n"], rtol=1e-11, atol=0)


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("dims", [(256, 128, 64), (128, 64, 256), (64, 256, 128)])
def test_every_3d_plan_variant_matches_oracle(dims, variant):
    import gpe_hip
    err, plans, tg, to = _run_pair(gpe_hip, dims, variant, steps=(0.002, 10, 1))
    print(f"3D {dims} variant {variant}: plans {plans} rel-L2 {err:.3e}", flush=True)
    assert plans == [variant, variant, variant]
    assert err < 1e-11
    assert np.allclose(tg.history["mu"], to.history["mu"], rtol=1e-11, atol=0)


def test_tune_measures_keeps_state_and_records_wisdom():
    import gpe_hip
    from gpe_hip import _lib
    G = gpe_hip.grid2d(8.0, 256, 8.0, 128)
    dev = _lib.Device(G)
    rng = np.random.Generator(np.random.PCG64(5))
    phi = rng.standard_normal(G.size) + 1j * rng.standard_normal(G.size)
    dev.set_state(0, phi)
    t = dev.tune(remeasure=True)
    print("tune times (us) [axis, variant]:", t.tolist(), "plans", dev.plans(), flush=True)
    assert np.all(np.isfinite(t[:2])) and np.all(t[:2] > 0)
    assert np.all(np.isnan(t[2]))
    want = dev.plans()
    assert [_lib.wisdom_get(2, 0, 256), _lib.wisdom_get(2, 1, 128)] == want
    for a in range(2):      # a non-default plan is recorded only when it measured faster
        assert want[a] == 0 or t[a][1] < t[a][0]
    assert np.array_equal(dev.get_state(0), phi)           # the state survives the timing sweeps
    # without remeasure the recorded plans are kept and nothing is timed
    t2 = dev.tune()
    assert np.all(np.isnan(t2))
    assert dev.plans() == want


def test_tuned_task_matches_oracle():
    import gpe_hip
    err, plans, tg, to = _run_pair(gpe_hip, (256, 192), variant=-1)
    print(f"tuned 2D (256, 192): plans {plans} rel-L2 {err:.3e}", flush=True)
    assert err < 1e-11


@pytest.mark.parametrize("variant", [0, 1])
def test_ensemble_with_each_variant_matches_single_runs(variant):
    import gpe_hip
    from gpe_hip import _lib
    from gpe_hip.batch import GPEEnsemble
    for a in (0, 1):
        for n in (64, 128, 256, 384):
            _lib.wisdom_set(2, a, n, variant)
    rng = np.random.Generator(np.random.PCG64(7 + variant))
    shapes = [(128, 64), (64, 128), (256, 128), (384, 64), (128, 128)]
    specs = []
    for nx, ny in shapes:
        G = gpe_hip.grid2d(8.0, nx, 7.0, ny)
        V = 0.5 * (G.mesh.x ** 2 + rng.uniform(1, 2) ** 2 * G.mesh.y ** 2)
        phi = np.exp(-(G.mesh.x ** 2 + G.mesh.y ** 2) / 4) * np.exp(1j * rng.uniform(0, 0.2, G.size))
        specs.append((G, V, phi, dict(g=rng.uniform(0.2, 1.0), gamma=0.01, Ntotal=rng.uniform(5, 20))))

    def make(spec):
        G, V, phi, props = spec
        t = gpe_hip.GPEtask(G, V)
        for k, v in props.items():
            setattr(t, k, v)
        t.init_state = phi.copy()
        return t

    ens_tasks = [make(s) for s in specs]
    ens = GPEEnsemble(ens_tasks)
    assert ens.tune_times == {}                      # every pair was already in the wisdom
    outs = ens.solve_split(0.004, 20, 2)
    for i, s in enumerate(specs):
        single = make(s)
        ref = single.solve_split(0.004, 20, 2)
        assert single._device().plans() == [variant, variant]
        assert np.array_equal(outs[i], ref), (i, rel_l2(outs[i], ref))
    # and one trajectory against the oracle
    G, V, phi, props = specs[2]
    O = oracle.grid2d(8.0, 256, 7.0, 128)
    to = oracle.GPEtask(O, V)
    for k, v in props.items():
        setattr(to, k, v)
    to.init_state = phi.copy()
    assert rel_l2(outs[2], to.solve_split(0.004, 20, 2)) < 1e-11


def test_ensemble_tune_records_every_pair():
    import gpe_hip
    from gpe_hip import _lib
    from gpe_hip.batch import GPEEnsemble
    shapes = [(128, 256), (256, 128), (192, 64)]
    tasks = []
    for nx, ny in shapes:
        G = gpe_hip.grid2d(8.0, nx, 7.0, ny)
        t = gpe_hip.GPEtask(G, 0.5 * (G.mesh.x ** 2 + G.mesh.y ** 2))
        t.Ntotal, t.g = 10.0, 0.5
        t.init_state = np.exp(-(G.mesh.x ** 2 + G.mesh.y ** 2) / 4).astype(np.complex128)
        tasks.append(t)
    ens = GPEEnsemble(tasks)
    print("ensemble tune times (us):", ens.tune_times, flush=True)
    # (x, 128), (x, 256), (y, 64), (y, 128), (y, 256); 192 has one plan
    assert set(ens.tune_times) == {(0, 128), (0, 256), (1, 64), (1, 128), (1, 256)}
    for (a, n), (t0, t1) in ens.tune_times.items():
        assert t0 > 0 and t1 > 0
        assert _lib.wisdom_get(2, a, n) == (1 if t1 < 0.985 * t0 else 0)   # the tuner's 1.5 % margin
    assert _lib.wisdom_get(2, 0, 192) == -1
    ens.solve_split(0.004, 10, 1)
