"""GPU adjoint gradient (gpe_adjoint_grad) against the oracle adjoint and central FD."""
import numpy as np
import pytest

import oracle
from oracle.adjoint import ControlProblem
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def problem(dims, n, xmax, nsteps, spi, nbasis, seed=2, g=2.0, ddt=0.005):
    import gpe_hip
    from gpe_hip.adjoint import SplitStepControl
    if dims == 2:
        G, O = gpe_hip.grid2d(xmax, n[0], xmax, n[1]), oracle.grid2d(xmax, n[0], xmax, n[1])
    else:
        G = gpe_hip.grid3d(xmax, n[0], xmax, n[1], xmax, n[2])
        O = oracle.grid3d(xmax, n[0], xmax, n[1], xmax, n[2])
    rng = np.random.Generator(np.random.PCG64(seed))
    r2 = G.mesh.x ** 2 + G.mesh.y ** 2 + (G.mesh.z ** 2 if dims == 3 else 0)
    V0 = 0.5 * r2
    ab = rng.uniform(-xmax / 3, xmax / 3, (nbasis, 3))
    basis = []
    for a, b, cz in ab:
        d2 = (G.mesh.x - a) ** 2 + (G.mesh.y - b) ** 2 + ((G.mesh.z - cz) ** 2 if dims == 3 else 0)
        basis.append(np.exp(-d2 / 2))
    psi0 = np.exp(-((G.mesh.x - 1) ** 2 + (r2 - G.mesh.x ** 2)) / 2) + 0j
    psi0 = G.normalize(psi0)
    cp = ControlProblem(O, V0, basis, g=g, ddt=ddt, nsteps=nsteps, steps_per_interval=spi, psi0=psi0, Ntotal=1.0)
    c = rng.normal(0, 0.5, (nbasis, cp.nint))
    psi_t = cp.forward(rng.normal(0, 0.5, (nbasis, cp.nint)))
    task = gpe_hip.GPEtask(G, V0)
    task.g, task.Ntotal = g, 1.0
    task.init_state = psi0
    ctl = SplitStepControl(task, V0, basis, ddt, nsteps, spi)
    return cp, ctl, c, psi_t


@pytest.mark.parametrize("dims,n,xmax", [(2, (32, 32), 6.0), (2, (64, 48), 8.0), (3, (16, 16, 12), 5.0)])
def test_adjoint_matches_oracle(dims, n, xmax):
    cp, ctl, c, psi_t = problem(dims, n, xmax, nsteps=60, spi=30, nbasis=3)
    Lg, gg, rerr = ctl.loss_and_grad(c, psi_t)
    Lo, go, _ = cp.loss_and_grad(c, psi_t)
    print(f"adjoint {dims}D {n}: L {Lg:.6e} vs {Lo:.6e}; grad rel {rel_l2(gg, go):.2e}; recon {rerr:.2e}")
    assert abs(Lg - Lo) < 1e-12
    assert rel_l2(gg, go) < 1e-9
    assert rerr < 1e-12
    # forward alone equals the oracle forward
    assert rel_l2(ctl.forward(c), cp.forward(c)) < 1e-11


def test_adjoint_matches_central_fd_on_gpu():
    cp, ctl, c, psi_t = problem(2, (64, 64), 8.0, nsteps=100, spi=50, nbasis=4, seed=7)
    L, g, _ = ctl.loss_and_grad(c, psi_t)
    h = 1e-6
    o = cp.grid
    fds, ads = [], []
    for (k, m) in [(0, 0), (1, 1), (3, 0), (2, 1), (0, 1), (2, 0)]:
        cpl = c.copy()
        cmi = c.copy()
        cpl[k, m] += h
        cmi[k, m] -= h
        Lp = 1 - abs(o.inner(psi_t, ctl.forward(cpl))) ** 2
        Lm = 1 - abs(o.inner(psi_t, ctl.forward(cmi))) ** 2
        fd = (Lp - Lm) / (2 * h)
        print(f"  dL/dc[{k},{m}] adjoint {g[k, m]:.10e} fd {fd:.10e}")
        fds.append(fd)
        ads.append(g[k, m])
        # per component: FD rounding noise ~ eps(L) / h ~ 1e-9 absolute
        assert abs(fd - g[k, m]) <= 1e-6 * abs(fd) + 2e-9
    err = rel_l2(np.array(ads), np.array(fds))
    print(f"  gradient vs central FD (h=1e-6): rel-L2 {err:.2e}")
    assert err < 1e-6


def test_adjoint_repeatable_and_state_preserved():
    cp, ctl, c, psi_t = problem(2, (32, 32), 6.0, nsteps=40, spi=20, nbasis=2, seed=3)
    r1 = ctl.loss_and_grad(c, psi_t)
    r2 = ctl.loss_and_grad(c, psi_t)
    assert r1[0] == r2[0] and np.array_equal(r1[1], r2[1])


def test_autograd_function_wrapper():
    torch = pytest.importorskip("torch")
    cp, ctl, c, psi_t = problem(2, (32, 32), 6.0, nsteps=40, spi=20, nbasis=2, seed=4)
    f = ctl.autograd_function(psi_t)
    ct = torch.tensor(c, dtype=torch.float64, requires_grad=True)
    L = f.apply(ct)
    L.backward()
    _, go, _ = cp.loss_and_grad(c, psi_t)
    assert rel_l2(ct.grad.numpy(), go) < 1e-9
