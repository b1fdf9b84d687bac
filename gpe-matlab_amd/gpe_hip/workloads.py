"""Synthetic, seeded workloads of the benchmark configurations.

Each builder returns plain numpy inputs (grid arguments, potential, initial state,
task properties) so the same bits can be fed to the GPU path and to any checker.
"""
import numpy as np

from .grid import grid2d, grid3d


def c1_vortex_2d():
    """2D GPEtask, grid2d(17, 256, 17, 256), V = (X^2+Y^2)/2, Ntotal = 1e5, g = 0.06
    (examples/example1_vortex.m:11-35 in dimensionless units): ITP (0.02, 1e-6) from
    a seeded random guess, vortex imprint at x0 = 5, then solve_split(0.004, 50, 500)."""
    return dict(grid=("2d", (17, 256, 17, 256)), V=lambda X, Y, Z: 0.5 * (X ** 2 + Y ** 2),
                props=dict(Ntotal=1e5, g=0.06), itp=(0.02, 1e-6), split=(0.004, 50, 500), seed=0)


def imprint_vortex(grid, phi, g, x0=5.0):
    """examples/example1_vortex.m:26-30: phi * tanh(r/xi) * exp(i atan2(x-x0, y))."""
    xi = 1 / np.sqrt(2 * g * np.max(np.abs(phi)) ** 2)
    rr = np.sqrt((grid.mesh.x - x0) ** 2 + grid.mesh.y ** 2)
    angs = np.arctan2(grid.mesh.x - x0, grid.mesh.y)
    return phi * np.tanh(rr / xi) * np.exp(1j * angs)


def thomas_fermi(grid, V, N, g, eps=np.finfo(np.float64).eps):
    """groundstate_tf.m:25-37 bisection (host, one-off)."""
    mul, mur = np.min(V), np.max(V)
    NC = N * g
    mu = (mul + mur) / 2
    vvv = None
    while (mur - mul) / mu > eps:
        vvv = (mu - V) * (V < mu)
        NN = np.sum(vvv) * grid.weight
        if NN > NC:
            mur = mu
        else:
            mul = mu
        mu = (mul + mur) / 2
    return np.sqrt(vvv / g), mu


def c2_state(n=256, xmax=12.0, seed=1):
    """3D drop-in: grid3d(12,256,12,256,12,256), V = (X^2+Y^2+16Z^2)/2, Ntotal = 1e6,
    g = 0.02, gamma = 0.01; psi0 = Thomas-Fermi * exp(i theta), theta ~ U[0, 0.1]."""
    grid = grid3d(xmax, n, xmax, n, xmax, n)
    V = 0.5 * (grid.mesh.x ** 2 + grid.mesh.y ** 2 + 16 * grid.mesh.z ** 2)
    props = dict(Ntotal=1e6, g=0.02, gamma=0.01)
    tf, _ = thomas_fermi(grid, V, props["Ntotal"], props["g"])
    rng = np.random.Generator(np.random.PCG64(seed))
    theta = rng.uniform(0.0, 0.1, grid.size)
    psi0 = tf * np.exp(1j * theta)
    return grid, V, psi0, props


def c4_ensemble(ntraj=256, sizes=(128, 192, 256, 384), xmax=15.0, Ntotal=1e4, seed=4):
    """Ragged ensemble: grid sizes uniform over `sizes` (square), g ~ U[0.01, 0.1],
    V = (X^2 + lam^2 Y^2)/2 with lam ~ U[1, 2], Thomas-Fermi state with a vortex
    imprinted at a seeded position |r0| < 5.  Returns a list of
    (grid, V, psi0, props) tuples."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(ntraj):
        n = int(rng.choice(sizes))
        g = float(rng.uniform(0.01, 0.1))
        lam = float(rng.uniform(1.0, 2.0))
        r0 = 5.0 * np.sqrt(rng.uniform()) * np.exp(2j * np.pi * rng.uniform())
        grid = grid2d(xmax, n, xmax, n)
        V = 0.5 * (grid.mesh.x ** 2 + lam ** 2 * grid.mesh.y ** 2)
        tf, _ = thomas_fermi(grid, V, Ntotal, g)
        psi = imprint_vortex_at(grid, tf.astype(np.complex128), g, r0.real, r0.imag)
        out.append((grid, V, psi, dict(Ntotal=Ntotal, g=g)))
    return out


def imprint_vortex_at(grid, phi, g, x0, y0):
    """phi * tanh(r/xi) * exp(i atan2(x-x0, y-y0)) (examples/example1_vortex.m:26-30 form)."""
    xi = 1 / np.sqrt(2 * g * np.max(np.abs(phi)) ** 2)
    rr = np.sqrt((grid.mesh.x - x0) ** 2 + (grid.mesh.y - y0) ** 2)
    return phi * np.tanh(rr / xi) * np.exp(1j * np.arctan2(grid.mesh.x - x0, grid.mesh.y - y0))


def c3_control(n=512, xmax=20.0, nsteps=2000, spi=250):
    """Adjoint workload: grid2d(20, 512, 20, 512), V = (X^2+Y^2)/2 + sum_k c[k, m] G_k with
    eight Gaussian bumps G_k = exp(-((X-a_k)^2+(Y-b_k)^2)/2), (a_k, b_k) ~ U[-8, 8] and
    c ~ N(0, 0.5^2) over eight intervals of spi steps (seed 2); the target state is the
    same dynamics under an independent draw of c (seed 3).  g = 10, Ntotal = 1,
    gamma = 0, dt = 0.002, T = nsteps steps.  psi0 is a unit-norm Gaussian at x = 1."""
    grid = grid2d(xmax, n, xmax, n)
    rng = np.random.Generator(np.random.PCG64(2))
    V0 = 0.5 * (grid.mesh.x ** 2 + grid.mesh.y ** 2)
    ab = rng.uniform(-8, 8, (8, 2))
    basis = [np.exp(-((grid.mesh.x - a) ** 2 + (grid.mesh.y - b) ** 2) / 2) for a, b in ab]
    nint = (nsteps + spi - 1) // spi
    c = rng.normal(0, 0.5, (8, nint))
    ct = np.random.Generator(np.random.PCG64(3)).normal(0, 0.5, (8, nint))
    psi0 = grid.normalize(np.exp(-((grid.mesh.x - 1) ** 2 + grid.mesh.y ** 2) / 2) + 0j)
    return dict(grid=grid, grid_args=(xmax, n, xmax, n), V0=V0, basis=basis, c=c, c_target=ct, psi0=psi0,
                g=10.0, Ntotal=1.0, ddt=0.002, nsteps=nsteps, spi=spi)


def c5_spinor(n=192, xmax=12.0, seed=5):
    """Two-component workload: grid3d(12, 192, ...), V1 = V2 = (X^2+Y^2+Z^2)/2,
    g = [0.02 0.018; 0.018 0.02], Omega = 0.5 exp(2iX), Ntotal = 1e6; each component is
    the Thomas-Fermi profile / sqrt(2) times exp(i theta), theta ~ U[0, 0.1] (seeded)."""
    grid = grid3d(xmax, n, xmax, n, xmax, n)
    V = 0.5 * (grid.mesh.x ** 2 + grid.mesh.y ** 2 + grid.mesh.z ** 2)
    tf, _ = thomas_fermi(grid, V, 1e6, 0.02)
    rng = np.random.Generator(np.random.PCG64(seed))
    a = tf * np.sqrt(0.5) * np.exp(1j * rng.uniform(0, 0.1, grid.size))
    b = tf * np.sqrt(0.5) * np.exp(1j * rng.uniform(0, 0.1, grid.size))
    gm = np.array([[0.02, 0.018], [0.018, 0.02]])
    return dict(grid=grid, grid_args=(xmax, n, xmax, n, xmax, n), V=V, psi=[a, b], g=gm,
                coupling=0.5 * np.exp(2j * grid.mesh.x), omega0=0.5, k=2.0, Ntotal=1e6)
