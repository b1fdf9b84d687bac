"""Pins for the oracle adjoint (CPU): its forward equals GPEtask.solve_split with
the equivalent Vtd, and its gradient matches central finite differences."""
import numpy as np

import oracle
from oracle.adjoint import ControlProblem
from conftest import rel_l2


def small_problem(n=16, nsteps=40, spi=20, nbasis=2, seed=2):
    grid = oracle.grid2d(6, n, 6, n)
    rng = np.random.Generator(np.random.PCG64(seed))
    V0 = 0.5 * (grid.mesh.x ** 2 + grid.mesh.y ** 2)
    ab = rng.uniform(-2, 2, (nbasis, 2))
    G = [np.exp(-((grid.mesh.x - a) ** 2 + (grid.mesh.y - b) ** 2) / 2) for a, b in ab]
    psi0 = np.exp(-((grid.mesh.x - 1) ** 2 + grid.mesh.y ** 2) / 2) / np.sqrt(np.pi) + 0j
    psi0 = grid.normalize(psi0)
    cp = ControlProblem(grid, V0, G, g=2.0, ddt=0.01, nsteps=nsteps, steps_per_interval=spi, psi0=psi0, Ntotal=1.0)
    c = rng.normal(0, 0.5, (nbasis, cp.nint))
    c_t = rng.normal(0, 0.5, (nbasis, cp.nint))
    psi_t = cp.forward(c_t)
    return cp, c, psi_t


def test_control_forward_is_solve_split():
    cp, c, _ = small_problem()
    t = oracle.GPEtask(cp.grid, cp.V0)
    t.g, t.Ntotal = cp.g, cp.Ntotal
    t.Vtd = cp.vtd(c)
    t.init_state = cp.psi0
    ref = t.solve_split(cp.ddt, 20, 2)
    assert rel_l2(cp.forward(c), ref) < 1e-13


def test_adjoint_gradient_matches_central_fd():
    cp, c, psi_t = small_problem()
    L, grad, _ = cp.loss_and_grad(c, psi_t)
    h = 1e-6
    for k in range(c.shape[0]):
        for m in range(c.shape[1]):
            cp_ = c.copy()
            cm_ = c.copy()
            cp_[k, m] += h
            cm_[k, m] -= h
            fd = (cp.loss(cp_, psi_t) - cp.loss(cm_, psi_t)) / (2 * h)
            assert abs(fd - grad[k, m]) <= 1e-6 * max(abs(fd), 1e-3), (k, m, fd, grad[k, m])
    assert 0 < L < 1


def test_adjoint_directional_derivative():
    cp, c, psi_t = small_problem(nsteps=60, spi=30, nbasis=3, seed=5)
    _, grad, _ = cp.loss_and_grad(c, psi_t)
    v = np.random.Generator(np.random.PCG64(9)).normal(size=c.shape)
    eps = 1e-6
    fd = (cp.loss(c + eps * v, psi_t) - cp.loss(c - eps * v, psi_t)) / (2 * eps)
    assert abs(fd - np.sum(grad * v)) <= 1e-6 * abs(fd)


def spinor_problem(n=16, nsteps=30, seed=3):
    from oracle.adjoint import SpinorRabiControl
    grid = oracle.grid2d(5, n, 5, n)
    V = 0.5 * (grid.mesh.x ** 2 + grid.mesh.y ** 2)
    rng = np.random.Generator(np.random.PCG64(seed))
    a = np.exp(-(grid.mesh.x ** 2 + grid.mesh.y ** 2) / 2) * np.exp(1j * rng.uniform(0, 0.3, grid.size))
    b = 0.5 * np.exp(-((grid.mesh.x - 0.5) ** 2 + grid.mesh.y ** 2) / 2) + 0j
    N = grid.integrate(np.abs(a) ** 2 + np.abs(b) ** 2)
    g = [[0.5, 0.45], [0.45, 0.6]]
    return SpinorRabiControl(grid, V, V + 0.2 * grid.mesh.y, g, 0.01, nsteps, [a, b], N), grid, g, [a, b], N


def test_spinor_control_forward_is_solve_split():
    cp, grid, g, psi0, N = spinor_problem()
    t = oracle.GPEtaskSpinor(grid, cp.V)
    t.g = np.array(g)
    t.coupling = 0.7 * np.exp(1j * 1.3 * grid.mesh.x)
    t.Ntotal = N
    t.init_state = psi0
    ref = t.solve_split(0.01, 30, 1)
    out = cp.forward(0.7, 1.3)
    assert max(rel_l2(out[0], ref[0]), rel_l2(out[1], ref[1])) < 1e-12


def test_spinor_adjoint_matches_central_fd():
    cp, *_ = spinor_problem()
    L, grad = cp.loss_and_grad(0.7, 1.3)
    h = 1e-6
    fdO = (cp.loss(0.7 + h, 1.3) - cp.loss(0.7 - h, 1.3)) / (2 * h)
    fdk = (cp.loss(0.7, 1.3 + h) - cp.loss(0.7, 1.3 - h)) / (2 * h)
    assert abs(grad[0] - fdO) <= 1e-6 * abs(fdO) + 1e-9, (grad, fdO)
    assert abs(grad[1] - fdk) <= 1e-6 * abs(fdk) + 1e-9, (grad, fdk)
