"""Long-horizon parity: the c2 physics (V = (X^2+Y^2+16Z^2)/2, Ntotal = 1e6, g = 0.02,
gamma = 0.01, seeded Thomas-Fermi state with phase noise) for the configuration's full
1000 steps at 64^3, where the oracle finishes in seconds.  The GPU state is compared
with the oracle's every 100 steps (solve_split.m's callback, GPEtask.m:115-142).

Rounding differences between two FP64 implementations grow with the horizon.  At this
resolution the oracle run with another FFT library (numpy's instead of scipy's) stays
within 1.2e-14 of itself after 1000 steps, and the bar 1e-12 leaves room above that
floor."""
import numpy as np
import pytest

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _run(mod, n, snapshots):
    from gpe_hip import workloads
    _, V, psi0, props = workloads.c2_state(n=n)
    G = mod.grid3d(12, n, 12, n, 12, n)
    t = mod.GPEtask(G, V)
    t.Ntotal, t.g, t.gamma = props["Ntotal"], props["g"], props["gamma"]
    t.init_state = psi0
    t.user_callback = lambda task: snapshots.append(np.array(task.current_state, copy=True))
    t.solve_split(0.002, 100, 10)
    return t


def test_c2_physics_1000_steps_at_64_cubed():
    import gpe_hip
    sg, so = [], []
    tg = _run(gpe_hip, 64, sg)
    to = _run(oracle, 64, so)
    assert len(sg) == len(so) == 10
    errs = [rel_l2(a, b) for a, b in zip(sg, so)]
    print("c2 physics 64^3, rel-L2 every 100 steps: " + " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) < 1e-12
    assert np.allclose(tg.history["mu"], to.history["mu"], rtol=1e-12, atol=0)
    assert np.allclose(tg.history["n"], to.history["n"], rtol=1e-12, atol=0)
