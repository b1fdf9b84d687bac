"""The two-sweep 3D step of grids with a 256-point y axis (y = a + 16 b split).

The y transform is factored 16 x 16: the factor over b runs inside the x sweep,
the factor over a inside the z sweep, each as a 16-point DFT in registers after an
LDS exchange of the tile.  Checked here: the exchange + DFT against numpy's FFT, and
GPEtask.solve_split on such grids against the FP64 oracle (rel-L2 <= 1e-11) and
against the four-sweep path of the same library (GPE_YSPLIT=0).
"""
import os

import numpy as np
import pytest

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def two_sweep_step():
    """Handles created inside these tests run the two-sweep step (the default, set
    explicitly so the tests do not depend on the environment)."""
    old = os.environ.get("GPE_YSPLIT")
    os.environ["GPE_YSPLIT"] = "1"
    yield
    if old is None:
        del os.environ["GPE_YSPLIT"]
    else:
        os.environ["GPE_YSPLIT"] = old

TOL = 1e-11


def test_row_dft16_matches_numpy():
    """The 16-point factor through the pencil -> row -> pencil exchange of one tile
    (gpe_selftest_rowdft) against numpy's FFT over the pencil index."""
    from gpe_hip import _lib
    rng = np.random.Generator(np.random.PCG64(11))
    x = rng.standard_normal(4096) + 1j * rng.standard_normal(4096)
    f, i = _lib.selftest_rowdft(x)
    X = x.reshape(256, 16)                       # [z, a]
    F = np.fft.fft(X, axis=1)
    Iv = np.fft.ifft(X, axis=1) * 16
    np.testing.assert_allclose(f.reshape(256, 16), F, rtol=0, atol=1e-13 * np.abs(F).max())
    np.testing.assert_allclose(i.reshape(256, 16), Iv, rtol=0, atol=1e-13 * np.abs(Iv).max())
    # a unit impulse at pencil 3 of point 17: W16^{3 k} on that point, zero elsewhere
    e = np.zeros(4096, dtype=complex)
    e[3 + 16 * 17] = 1.0
    f, _ = _lib.selftest_rowdft(e)
    k = np.arange(16)
    np.testing.assert_allclose(f[16 * 17: 16 * 18], np.exp(-2j * np.pi * ((3 * k) % 16) / 16), rtol=0, atol=1e-15)
    assert np.count_nonzero(f) == 16


def _pair(n, xmax):
    import gpe_hip
    return (gpe_hip.grid3d(xmax[0], n[0], xmax[1], n[1], xmax[2], n[2]),
            oracle.grid3d(xmax[0], n[0], xmax[1], n[1], xmax[2], n[2]))


def V3(X, Y, Z):
    return 0.5 * (X ** 2 + 1.5 * Y ** 2 + 3.0 * Z ** 2)


def _make(G, mod, phi, **props):
    t = mod.GPEtask(G, V3)
    t.g, t.Ntotal = 0.6, 10.0
    for k, v in props.items():
        setattr(t, k, v)
    t.init_state = phi.copy()
    return t


def _phi(G, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    r2 = G.mesh.x ** 2 + G.mesh.y ** 2 + G.mesh.z ** 2
    phi = np.exp(-r2 / 3) * np.exp(1j * rng.uniform(0, 0.2, G.size))
    return phi * np.exp(1j * np.arctan2(G.mesh.x - 0.4, G.mesh.y + 0.2))


def _four_sweep(fn):
    os.environ["GPE_YSPLIT"] = "0"
    try:
        return fn()
    finally:
        os.environ["GPE_YSPLIT"] = "1"


CASES = [((64, 256, 64), (6, 9, 5)), ((128, 256, 64), (7, 10, 5)), ((64, 256, 128), (6, 8, 7)),
         ((256, 256, 64), (9, 9, 5)), ((64, 256, 256), (6, 9, 9))]


@pytest.mark.parametrize("n,xmax", CASES)
@pytest.mark.parametrize("gamma", [0.0, 0.02])
def test_ysplit_solve_split_matches_oracle(n, xmax, gamma):
    import gpe_hip
    G, O = _pair(n, xmax)
    phi = _phi(G, sum(n))
    tg = _make(G, gpe_hip, phi, gamma=gamma)
    to = _make(O, oracle, phi, gamma=gamma)
    pg = tg.solve_split(0.002, 10, 2)
    po = to.solve_split(0.002, 10, 2)
    err = rel_l2(pg, po)
    print(f"y-split {n} gamma={gamma}: rel-L2 vs oracle {err:.3e}", flush=True)
    assert err < TOL
    assert np.allclose(tg.history["mu"], to.history["mu"], rtol=TOL, atol=0)
    assert np.allclose(tg.history["n"], to.history["n"], rtol=TOL, atol=0)
    # the step ran as two sweeps: one more outer iteration under the profiler
    dev = tg._device()
    dev.profile_enable(True)
    tg.solve_split(0.002, 10, 3)
    prof = dev.profile_read()
    dev.profile_enable(False)
    seqs = {s for (s, _ax) in prof}
    assert {20, 21, 22, 23} <= seqs and not ({4, 0, 1, 2} & seqs), seqs


def test_ysplit_equals_four_sweep_path_and_resumes():
    import gpe_hip
    n, xmax = (64, 256, 64), (6, 9, 5)
    G, O = _pair(n, xmax)
    phi = _phi(G, 5)
    props = dict(gamma=0.01, decay_rate=0.05, mu_init=0.0)
    ty = _make(G, gpe_hip, phi, **props)
    ty.solve_split(0.002, 20, 2)
    py = ty.solve_split(0.002, 20, 4)           # resume from current_iter = 2
    tf = _four_sweep(lambda: _make(G, gpe_hip, phi, **props))
    _four_sweep(lambda: tf.solve_split(0.002, 20, 2))
    pf = _four_sweep(lambda: tf.solve_split(0.002, 20, 4))
    to = _make(O, oracle, phi, **props)
    to.solve_split(0.002, 20, 2)                # a resumed run recomputes mu from the state
    po = to.solve_split(0.002, 20, 4)           # (solve_split.m:41), as the reference does
    e_yf, e_yo, e_fo = rel_l2(py, pf), rel_l2(py, po), rel_l2(pf, po)
    print(f"y-split vs 4-sweep {e_yf:.3e}, vs oracle {e_yo:.3e}; 4-sweep vs oracle {e_fo:.3e}", flush=True)
    assert e_yo < TOL and e_fo < TOL and e_yf < TOL
    assert np.allclose(ty.history["mu"], to.history["mu"], rtol=TOL, atol=0)
    assert np.allclose(ty.history["n"], to.history["n"], rtol=TOL, atol=0)


def test_ysplit_graph_replay_matches_direct_launches():
    import gpe_hip
    n, xmax = (128, 256, 64), (7, 9, 5)
    G, _ = _pair(n, xmax)
    phi = _phi(G, 8)
    ta = _make(G, gpe_hip, phi, gamma=0.02)
    pa = ta.solve_split(0.002, 10, 4)
    old = os.environ.get("GPE_NO_GRAPH")
    os.environ["GPE_NO_GRAPH"] = "1"
    try:
        tb = _make(G, gpe_hip, phi, gamma=0.02)
        pb = tb.solve_split(0.002, 10, 4)
    finally:
        if old is None:
            del os.environ["GPE_NO_GRAPH"]
        else:
            os.environ["GPE_NO_GRAPH"] = old
    assert np.array_equal(pa, pb)
    assert np.array_equal(ta.history["mu"], tb.history["mu"])


def test_ysplit_time_dependent_potential():
    """A host Vtd re-uploads V every block; the tile-order copy must follow it."""
    import gpe_hip
    n, xmax = (64, 256, 64), (6, 9, 5)
    G, O = _pair(n, xmax)
    phi = _phi(G, 9)

    def vtd(task, t):
        return 0.3 * np.sin(7.0 * t) * task.grid.mesh.x + 0.1 * np.cos(5.0 * t) * task.grid.mesh.y ** 2

    tg = _make(G, gpe_hip, phi, gamma=0.0)
    to = _make(O, oracle, phi, gamma=0.0)
    tg.Vtd = vtd
    to.Vtd = vtd
    pg = tg.solve_split(0.002, 10, 3)
    po = to.solve_split(0.002, 10, 3)
    err = rel_l2(pg, po)
    print(f"y-split with host Vtd: rel-L2 vs oracle {err:.3e}", flush=True)
    assert err < TOL


@pytest.mark.parametrize("n,xmax", [((64, 256, 64), (6, 9, 5)), ((128, 256, 64), (7, 10, 5))])
@pytest.mark.parametrize("gamma", [0.0, 0.02])
def test_ysplit_rotating_frame_matches_oracle_and_four_sweeps(n, xmax, gamma):
    """Scalar omega (set_kinop.m:5-7, ssft_kin_step.m:22-25): x stays real between the two
    sweeps, the x half steps run around the phase in the x sweep and the z sweep takes
    the (ky, x) table; against the oracle and the library's four-sweep rotating step."""
    import gpe_hip
    G, O = _pair(n, xmax)
    phi = _phi(G, 11 + sum(n))
    props = dict(gamma=gamma)
    tg = _make(G, gpe_hip, phi, **props)
    to = _make(O, oracle, phi, **props)
    t4 = _four_sweep(lambda: _make(G, gpe_hip, phi, **props))
    for t in (tg, to, t4):
        t.omega = 0.4
    pg = tg.solve_split(0.002, 10, 2)
    po = to.solve_split(0.002, 10, 2)
    p4 = _four_sweep(lambda: t4.solve_split(0.002, 10, 2))
    e_o, e_4 = rel_l2(pg, po), rel_l2(pg, p4)
    print(f"y-split rotating {n} gamma={gamma}: vs oracle {e_o:.3e}, vs four sweeps {e_4:.3e}", flush=True)
    assert e_o < TOL and e_4 < TOL
    assert np.allclose(tg.history["mu"], to.history["mu"], rtol=TOL, atol=0)
    assert np.allclose(tg.history["n"], to.history["n"], rtol=TOL, atol=0)
    dev = tg._device()
    dev.profile_enable(True)
    tg.solve_split(0.002, 10, 3)
    prof = dev.profile_read()
    dev.profile_enable(False)
    seqs = {s for (s, _ax) in prof}
    assert {20, 27, 28, 29} <= seqs and not ({12, 13, 14, 0, 1, 2} & seqs), seqs


def test_ysplit_rotating_frame_with_more_x_tiles_than_resident():
    """64 x 256 x 256 (4096 x-sweep work-groups): the rotating-frame opening and closing
    x sweeps convert between the natural and the block order through the second array;
    the run equals the library's four-sweep rotating step."""
    import gpe_hip
    n, xmax = (64, 256, 256), (6, 9, 9)
    G, _ = _pair(n, xmax)
    phi = _phi(G, 17)
    tg = _make(G, gpe_hip, phi, gamma=0.02)
    t4 = _four_sweep(lambda: _make(G, gpe_hip, phi, gamma=0.02))
    for t in (tg, t4):
        t.omega = 0.4
    pg = tg.solve_split(0.002, 10, 2)
    p4 = _four_sweep(lambda: t4.solve_split(0.002, 10, 2))
    e_4 = rel_l2(pg, p4)
    print(f"y-split rotating 64x256x256: vs four sweeps {e_4:.3e}", flush=True)
    assert e_4 < TOL
    assert np.allclose(tg.history["mu"], t4.history["mu"], rtol=TOL, atol=0)


@pytest.mark.parametrize("omega", [0.0, 0.4])
@pytest.mark.parametrize("gamma", [0.0, 0.02])
def test_ysplit_blocks_of_an_outer_iteration_hand_over_in_block_order(omega, gamma):
    """30 inner steps of n_recalc = 10: three blocks per outer iteration.  The first block
    opens from the natural order and the last closes to it; the block ends between them
    stay in block order.  Damped runs renormalise and update mu at every block end;
    against the oracle and the four-sweep step."""
    import gpe_hip
    n, xmax = (64, 256, 64), (6, 9, 5)
    G, O = _pair(n, xmax)
    phi = _phi(G, 23)
    props = dict(gamma=gamma)
    tg = _make(G, gpe_hip, phi, **props)
    to = _make(O, oracle, phi, **props)
    t4 = _four_sweep(lambda: _make(G, gpe_hip, phi, **props))
    for t in (tg, to, t4):
        t.omega = omega
    pg = tg.solve_split(0.002, 30, 2)
    po = to.solve_split(0.002, 30, 2)
    p4 = _four_sweep(lambda: t4.solve_split(0.002, 30, 2))
    e_o, e_4 = rel_l2(pg, po), rel_l2(pg, p4)
    print(f"y-split 3 blocks per outer, omega={omega} gamma={gamma}: vs oracle {e_o:.3e}, "
          f"vs four sweeps {e_4:.3e}", flush=True)
    assert e_o < TOL and e_4 < TOL
    assert np.allclose(tg.history["mu"], to.history["mu"], rtol=TOL, atol=0)
    assert np.allclose(tg.history["n"], to.history["n"], rtol=TOL, atol=0)
    dev = tg._device()
    dev.profile_enable(True)
    tg.solve_split(0.002, 30, 3)
    prof = dev.profile_read()
    dev.profile_enable(False)
    # block ends: the two inside the iteration hand over in block order (without rotation
    # in the steps' mixed form, whose kinetic sum is one z sweep, S_YKEZ; in the rotating
    # frame in real space, y kinetic sums through S_YKE); the y kinetic sums in natural
    # order are the last block end's and the initial mu of the resumed run (solve_split.m:41)
    if omega == 0.0:
        assert prof[(31, 0)][1] == 2 and prof[(32, 0)][1] == 2 and prof[(33, 2)][1] == 2, prof
    else:
        assert prof[(30, 1)][1] == 2, prof
    assert prof[(7, 1)][1] == 2, prof


def test_ysplit_blocks_hand_over_with_more_x_tiles_than_resident():
    """The block-order hand-over at 64 x 256 x 256 (4096 x-sweep work-groups): three blocks
    per outer iteration equal the four-sweep step."""
    import gpe_hip
    n, xmax = (64, 256, 256), (6, 9, 9)
    G, _ = _pair(n, xmax)
    phi = _phi(G, 29)
    tg = _make(G, gpe_hip, phi, gamma=0.02)
    t4 = _four_sweep(lambda: _make(G, gpe_hip, phi, gamma=0.02))
    pg = tg.solve_split(0.002, 30, 2)
    p4 = _four_sweep(lambda: t4.solve_split(0.002, 30, 2))
    e_4 = rel_l2(pg, p4)
    print(f"y-split 3 blocks per outer 64x256x256: vs four sweeps {e_4:.3e}", flush=True)
    assert e_4 < TOL
    assert np.allclose(tg.history["mu"], t4.history["mu"], rtol=TOL, atol=0)
