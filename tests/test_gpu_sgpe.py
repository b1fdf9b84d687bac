"""GPU parity of the stochastic GPE (SGPEtask.solve_split) against the oracle, which
draws the same counter-based noise (oracle/philox.py)."""
import numpy as np
import pytest

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _pair(dims, n, xmax, T=0.5, ecut=0.0, g=0.4, gamma=0.05, mu=1.5, seed=21):
    import gpe_hip
    if dims == 2:
        G, O = gpe_hip.grid2d(xmax[0], n[0], xmax[1], n[1]), oracle.grid2d(xmax[0], n[0], xmax[1], n[1])
    else:
        G = gpe_hip.grid3d(xmax[0], n[0], xmax[1], n[1], xmax[2], n[2])
        O = oracle.grid3d(xmax[0], n[0], xmax[1], n[1], xmax[2], n[2])
    r2 = G.mesh.x ** 2 + G.mesh.y ** 2 + (G.mesh.z ** 2 if dims == 3 else 0)
    V = 0.5 * r2
    psi = np.exp(-r2 / 3) * np.exp(0.3j * G.mesh.x)
    tg, to = gpe_hip.SGPEtask(G, V), oracle.SGPEtask(O, V)
    for t in (tg, to):
        t.g, t.gamma, t.T, t.ecut, t.mu_init, t.noise_seed = g, gamma, T, ecut, mu, seed
        t.init_state = psi.copy()
    return tg, to


CASES = [(2, (32, 24), (6, 5)), (2, (64, 64), (8, 8)), (3, (16, 24, 12), (5, 6, 4)), (3, (32, 32, 16), (6, 6, 4))]


@pytest.mark.parametrize("dims,n,xmax", CASES)
@pytest.mark.parametrize("ecut", [0.0, 6.0])
def test_sgpe_parity(dims, n, xmax, ecut):
    tg, to = _pair(dims, n, xmax, ecut=ecut)
    pg, po = tg.solve_split(0.005, 20, 3), to.solve_split(0.005, 20, 3)
    err = rel_l2(pg, po)
    print(f"SGPE {dims}D {n} ecut={ecut}: rel-L2 {err:.2e}")
    assert err < 1e-11
    assert np.allclose(tg.history["n"], to.history["n"], rtol=1e-11, atol=0)
    assert np.all(tg.history["mu"] == 1.5)


def test_sgpe_resume_is_the_same_run():
    ta, _ = _pair(2, (32, 32), (6, 6))
    tb, _ = _pair(2, (32, 32), (6, 6))
    pa = ta.solve_split(0.005, 20, 4)
    tb.solve_split(0.005, 20, 2)
    pb = tb.solve_split(0.005, 20, 4)
    assert np.array_equal(pa, pb)


def test_sgpe_zero_temperature_is_damped_gpe_with_fixed_mu():
    import gpe_hip
    ts, _ = _pair(2, (32, 32), (6, 6), T=0.0)
    tg = gpe_hip.GPEtask(ts.grid, ts.Vtrap)
    tg.g, tg.gamma, tg.mu_init = ts.g, ts.gamma, ts.mu_init
    tg.init_state = ts.init_state.copy()
    assert rel_l2(ts.solve_split(0.005, 20, 3), tg.solve_split(0.005, 20, 3)) < 1e-13


def test_sgpe_callback_term_stops():
    ts, _ = _pair(2, (32, 24), (6, 5))
    ts.user_callback = lambda task: "TERM" if task.current_iter >= 2 else ""
    ts.solve_split(0.005, 20, 5)
    assert ts.current_iter == 2


def _with_schedule(value, fn):
    import os
    old = os.environ.get("GPE_YSPLIT")
    os.environ["GPE_YSPLIT"] = value      # read when the task's device handle is created
    try:
        return fn()
    finally:
        if old is None:
            del os.environ["GPE_YSPLIT"]
        else:
            os.environ["GPE_YSPLIT"] = old


@pytest.mark.parametrize("ecut", [0.0, 6.0])
def test_sgpe_two_sweep_step_on_a_256_point_y_axis(ecut):
    """On a 3D grid with a 256-point y axis the stochastic step runs as two sweeps (noise
    in the x sweep's row layout, the cut-off mask in the z sweep's split order); same
    noise stream, so it matches the oracle and the four-sweep step of the library."""
    n, xmax = (64, 256, 64), (6, 8, 5)
    tg, to = _with_schedule("1", lambda: _pair(3, n, xmax, ecut=ecut))
    pg = _with_schedule("1", lambda: tg.solve_split(0.005, 20, 2))
    t4, _ = _with_schedule("0", lambda: _pair(3, n, xmax, ecut=ecut))
    p4 = _with_schedule("0", lambda: t4.solve_split(0.005, 20, 2))
    po = to.solve_split(0.005, 20, 2)
    e_o, e_4 = rel_l2(pg, po), rel_l2(pg, p4)
    print(f"SGPE two-sweep 64x256x64 ecut={ecut}: vs oracle {e_o:.2e}, vs four-sweep {e_4:.2e}", flush=True)
    assert e_o < 1e-11 and e_4 < 1e-12
    assert np.allclose(tg.history["n"], to.history["n"], rtol=1e-11, atol=0)


def test_sgpe_two_sweep_step_with_more_x_tiles_than_resident():
    """64 x 256 x 256: the x sweeps run 4096 work-groups, more than the GPU holds at once,
    so the tiles of one z plane do not all run together; the block's opening and
    closing sweeps must still move the state between its two orders without a tile
    reading another's stores.  Same noise stream as the four-sweep step."""
    n, xmax = (64, 256, 256), (6, 8, 8)
    tg, _ = _with_schedule("1", lambda: _pair(3, n, xmax, ecut=6.0))
    pg = _with_schedule("1", lambda: tg.solve_split(0.005, 10, 2))
    t4, _ = _with_schedule("0", lambda: _pair(3, n, xmax, ecut=6.0))
    p4 = _with_schedule("0", lambda: t4.solve_split(0.005, 10, 2))
    e_4 = rel_l2(pg, p4)
    print(f"SGPE two-sweep 64x256x256: vs four-sweep {e_4:.2e}", flush=True)
    assert e_4 < 1e-12
    assert np.allclose(tg.history["n"], t4.history["n"], rtol=1e-11, atol=0)


def test_sgpe_two_sweep_blocks_hand_over_in_block_order():
    """30 inner steps of n_recalc = 10: the block ends inside an outer iteration keep the
    state in block order; same noise stream as the oracle and the four-sweep step."""
    n, xmax = (64, 256, 64), (6, 8, 5)
    tg, to = _with_schedule("1", lambda: _pair(3, n, xmax, ecut=6.0))
    pg = _with_schedule("1", lambda: tg.solve_split(0.005, 30, 2))
    t4, _ = _with_schedule("0", lambda: _pair(3, n, xmax, ecut=6.0))
    p4 = _with_schedule("0", lambda: t4.solve_split(0.005, 30, 2))
    po = to.solve_split(0.005, 30, 2)
    e_o, e_4 = rel_l2(pg, po), rel_l2(pg, p4)
    print(f"SGPE two-sweep 3 blocks per outer: vs oracle {e_o:.2e}, vs four-sweep {e_4:.2e}", flush=True)
    assert e_o < 1e-11 and e_4 < 1e-12
    assert np.allclose(tg.history["n"], to.history["n"], rtol=1e-11, atol=0)
