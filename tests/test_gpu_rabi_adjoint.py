"""GPU Rabi adjoint (gpe_rabi_forward / gpe_rabi_grad) against the oracle's
spinor forward and adjoint and against central finite differences on the GPU."""
import numpy as np
import pytest

import oracle
from oracle.adjoint import SpinorRabiControl
from conftest import rel_l2

pytestmark = pytest.mark.gpu

G_MAT = [[0.5, 0.45], [0.4, 0.6]]   # asymmetric, so a transposed index shows


def problem(dims, n, xmax, nsteps, ddt=0.004, seed=4, ckpt=0, tape_gib=0.0):
    import gpe_hip
    from gpe_hip.adjoint import RabiControl
    if dims == 2:
        G, O = gpe_hip.grid2d(xmax[0], n[0], xmax[1], n[1]), oracle.grid2d(xmax[0], n[0], xmax[1], n[1])
    else:
        G = gpe_hip.grid3d(xmax[0], n[0], xmax[1], n[1], xmax[2], n[2])
        O = oracle.grid3d(xmax[0], n[0], xmax[1], n[1], xmax[2], n[2])
    r2 = G.mesh.x ** 2 + G.mesh.y ** 2 + (G.mesh.z ** 2 if dims == 3 else 0)
    V1 = 0.5 * r2
    V2 = V1 + 0.2 * G.mesh.y
    rng = np.random.Generator(np.random.PCG64(seed))
    a = np.exp(-r2 / 2) * np.exp(1j * rng.uniform(0, 0.3, G.size))
    b = 0.5 * np.exp(-(r2 - G.mesh.x ** 2 + (G.mesh.x - 0.5) ** 2) / 2) + 0j
    N = O.integrate(np.abs(a) ** 2 + np.abs(b) ** 2)
    ref = SpinorRabiControl(O, V1, V2, G_MAT, ddt, nsteps, [a, b], N)
    task = gpe_hip.GPEtaskSpinor(G, [V1, V2])
    task.g = np.array(G_MAT)
    task.Ntotal = N
    task.init_state = [a.copy(), b.copy()]
    return ref, RabiControl(task, ddt, nsteps, checkpoint_blocks=ckpt, tape_gib=tape_gib), task


CASES = [(2, (32, 24), (6, 5)), (2, (64, 64), (7, 7)), (3, (16, 24, 12), (5, 6, 4))]


@pytest.mark.parametrize("dims,n,xmax", CASES)
def test_rabi_forward_matches_oracle(dims, n, xmax):
    ref, ctl, _ = problem(dims, n, xmax, 40)
    L, psi = ctl.forward(0.7, 1.3)
    po = ref.forward(0.7, 1.3)
    err = max(rel_l2(psi[0], po[0]), rel_l2(psi[1], po[1]))
    Lo = ref.loss(0.7, 1.3)
    print(f"rabi forward {dims}D {n}: rel-L2 {err:.2e}, loss {L:.15f} vs {Lo:.15f}")
    assert err < 1e-11
    assert abs(L - Lo) < 1e-12


@pytest.mark.parametrize("dims,n,xmax", CASES)
def test_rabi_grad_matches_oracle_adjoint(dims, n, xmax):
    ref, ctl, _ = problem(dims, n, xmax, 40)
    L, g = ctl.loss_and_grad(0.7, 1.3)
    Lo, go = ref.loss_and_grad(0.7, 1.3)
    rel = np.abs(g - go) / np.abs(go)
    print(f"rabi grad {dims}D {n}: {g} vs oracle {go}, rel {rel}")
    assert abs(L - Lo) < 1e-12
    assert np.all(rel < 1e-9)


@pytest.mark.parametrize("om0,k", [(0.7, 1.3), (-0.4, 2.0), (0.5, 0.0)])
def test_rabi_grad_matches_central_fd(om0, k):
    _, ctl, _ = problem(2, (48, 32), (6, 5), 60, ddt=0.005)
    _, g = ctl.loss_and_grad(om0, k)
    h = 1e-6
    fdO = (ctl.loss(om0 + h, k) - ctl.loss(om0 - h, k)) / (2 * h)
    fdk = (ctl.loss(om0, k + h) - ctl.loss(om0, k - h)) / (2 * h)
    print(f"rabi FD ({om0}, {k}): adjoint {g}, fd {[fdO, fdk]}")
    assert abs(g[0] - fdO) <= 1e-6 * abs(fdO) + 1e-9
    assert abs(g[1] - fdk) <= 1e-6 * abs(fdk) + 1e-9


@pytest.mark.parametrize("tape_gib", [-1.0, 0.0, 3 * 10 * 2 * 32 * 24 * 16 / 2 ** 30])
def test_rabi_checkpoint_segments_do_not_change_the_gradient(tape_gib):
    # tape_gib: one segment stored (all others recomputed), half the free memory
    # (everything stored), and room for exactly 3 of the 6 blocks of 10 steps
    grads = []
    for ckpt in (0, 1, 2, 3, 6):
        _, ctl, _ = problem(2, (32, 24), (6, 5), 60, ckpt=ckpt, tape_gib=tape_gib)
        grads.append(ctl.loss_and_grad(0.6, 1.1))
    for L, g in grads[1:]:
        assert L == grads[0][0]
        assert np.array_equal(g, grads[0][1])


def test_rabi_calls_leave_the_state_and_allow_solve_split():
    _, ctl, task = problem(2, (32, 24), (6, 5), 40)
    dev = ctl._prepare()
    before = [dev.get_state(0), dev.get_state(1)]
    ctl.task._dev_state_valid = True
    dev.rabi_grad(0.004, 40, 0.7, 1.3)
    after = [dev.get_state(0), dev.get_state(1)]
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
    # a plain solve_split afterwards still matches the oracle
    import gpe_hip  # noqa: F401
    O = oracle.grid2d(6, 32, 5, 24)
    to = oracle.GPEtaskSpinor(O, [np.asarray(v) for v in task.Vtrap])
    to.g = np.array(G_MAT)
    to.coupling = task.coupling = 0.5 * np.exp(0.3j)
    to.Ntotal = task.Ntotal
    to.init_state = [s.copy() for s in task.init_state]
    task._dev_state_valid = False
    pg, po = task.solve_split(0.002, 20, 1), to.solve_split(0.002, 20, 1)
    assert max(rel_l2(pg[0], po[0]), rel_l2(pg[1], po[1])) < 1e-11


def test_rabi_rejects_unsupported_runs():
    import gpe_hip
    _, ctl, task = problem(2, (32, 24), (6, 5), 40)
    dev = ctl._prepare()
    with pytest.raises(gpe_hip.GPELibraryError, match="multiple of n_recalc"):
        dev.rabi_grad(0.004, 45, 0.7, 1.3)
    task.gamma = 0.01
    task._set_params(task.Ntotal)
    with pytest.raises(gpe_hip.GPELibraryError, match="gamma = 0"):
        dev.rabi_forward(0.004, 40, 0.7, 1.3)
