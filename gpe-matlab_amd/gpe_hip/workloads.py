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
