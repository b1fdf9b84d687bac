"""GPU ensemble (gpe_batch): every trajectory equals its single-trajectory run."""
import numpy as np
import pytest

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _tasks(seed, shapes, gammas):
    import gpe_hip
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for (nx, ny), gam in zip(shapes, gammas):
        G = gpe_hip.grid2d(8.0, nx, 7.0, ny)
        lam = rng.uniform(1, 2)
        V = 0.5 * (G.mesh.x ** 2 + lam ** 2 * G.mesh.y ** 2)
        phi = np.exp(-(G.mesh.x ** 2 + G.mesh.y ** 2) / 4) * np.exp(1j * rng.uniform(0, 0.2, G.size))
        phi = phi * np.exp(1j * np.arctan2(G.mesh.x - rng.uniform(-1, 1), G.mesh.y))
        out.append((G, V, phi, dict(g=rng.uniform(0.2, 1.0), gamma=gam, Ntotal=rng.uniform(5, 20))))
    return out


def _make(gpe_hip, spec):
    G, V, phi, props = spec
    t = gpe_hip.GPEtask(G, V)
    for k, v in props.items():
        setattr(t, k, v)
    t.init_state = phi.copy()
    return t


SHAPES = [(32, 32), (64, 48), (48, 64), (96, 96), (128, 32), (32, 128), (64, 64), (192, 96), (256, 64)]


@pytest.mark.parametrize("use_graph", [True, False])
def test_ensemble_matches_single_runs(use_graph):
    import gpe_hip
    from gpe_hip.batch import GPEEnsemble
    specs = _tasks(1, SHAPES, [0.0, 0.02, 0.0, 0.05, 0.0, 0.01, 0.0, 0.0, 0.03])
    ens_tasks = [_make(gpe_hip, s) for s in specs]
    ens = GPEEnsemble(ens_tasks, use_graph=use_graph)
    outs = ens.solve_split(0.004, 20, 3)
    for i, s in enumerate(specs):
        single = _make(gpe_hip, s)
        ref = single.solve_split(0.004, 20, 3)
        assert np.array_equal(outs[i], ref), (i, rel_l2(outs[i], ref))
        assert np.array_equal(ens_tasks[i].history["mu"], single.history["mu"])
        assert np.array_equal(ens_tasks[i].history["n"], single.history["n"])


def test_ensemble_matches_oracle_and_resumes():
    import gpe_hip
    from gpe_hip.batch import GPEEnsemble
    specs = _tasks(2, SHAPES[:5], [0.0, 0.02, 0.0, 0.0, 0.01])
    ens_tasks = [_make(gpe_hip, s) for s in specs]
    ens = GPEEnsemble(ens_tasks)
    ens.solve_split(0.004, 20, 2)
    outs = ens.solve_split(0.004, 20, 4)      # resume from current_iter = 2
    for i, (G, V, phi, props) in enumerate(specs):
        O = oracle.grid2d(G.xmax, G.nx, G.ymax, G.ny)
        to = oracle.GPEtask(O, V)
        for k, v in props.items():
            setattr(to, k, v)
        to.init_state = phi
        to.solve_split(0.004, 20, 2)
        po = to.solve_split(0.004, 20, 4)
        err = rel_l2(outs[i], po)
        print(f"ensemble traj {i} {G.size}: rel-L2 vs oracle {err:.2e}")
        assert err < 1e-11
        assert np.allclose(ens_tasks[i].history["mu"], to.history["mu"], rtol=1e-11, atol=0)


def test_ensemble_with_long_pencils_matches_single_runs():
    # 384- and 512-point axes run 8-pencil tiles (paired on one XCD), several
    # trajectories per shape share a chain
    import gpe_hip
    from gpe_hip.batch import GPEEnsemble
    shapes = [(384, 32), (32, 384), (512, 24), (24, 512), (384, 32), (512, 24)]
    specs = _tasks(3, shapes, [0.0, 0.02, 0.0, 0.01, 0.03, 0.0])
    ens_tasks = [_make(gpe_hip, s) for s in specs]
    outs = GPEEnsemble(ens_tasks).solve_split(0.004, 20, 2)
    for i, s in enumerate(specs):
        ref = _make(gpe_hip, s).solve_split(0.004, 20, 2)
        assert np.array_equal(outs[i], ref), (i, rel_l2(outs[i], ref))
