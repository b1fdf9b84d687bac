"""Pins for the FP64 oracle (CPU only).

The reference ships no tests, golden vectors or data, so every oracle function is
pinned here by something the mathematics fixes: brute-force DFTs, closed-form
eigenstates, exact invariants (norm, time reversal), convergence order, and
textbook limits.  Each check is chosen so that a plausible slip (a dropped term, a
sign, an index, a transposed operand, a wrong block phase) fails it.
"""
import numpy as np
import pytest

import oracle
from oracle import Grid2D, Grid3D, GPEtask, GPEtaskSpinor
from conftest import rel_l2


def brute_dftn(a, inverse=False):
    """O(N^2) DFT along every axis: X[k] = sum_n x[n] e^{-+2 pi i nk/N} (1/N on inverse)."""
    out = a.astype(np.complex128)
    for ax, n in enumerate(a.shape):
        j = np.arange(n)
        s = 1 if inverse else -1
        W = np.exp(s * 2j * np.pi * np.outer(j, j) / n)
        out = np.moveaxis(np.tensordot(W, np.moveaxis(out, ax, 0), axes=(1, 0)), 0, ax)
        if inverse:
            out = out / n
    return out


@pytest.mark.parametrize("shape", [(4, 6), (12, 8), (6, 4, 10), (16, 12, 2)])
def test_fft_matches_brute_dft(shape, rng):
    g = Grid3D(3, shape[1], 3, shape[0], 3, shape[2]) if len(shape) == 3 else Grid2D(3, shape[1], 3, shape[0])
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    assert rel_l2(g.fft(a), brute_dftn(a)) < 1e-14
    assert rel_l2(g.ifft(a), brute_dftn(a, inverse=True)) < 1e-14


def test_k_layout_closed_form():
    # grid2d/init.m:27: k = [0:n/2, -(n/2-1):-1] * 2pi/(n dx); Nyquist is +n/2
    g = Grid2D(17, 8, 5, 6)
    dx = 34 / 7
    expect = np.array([0, 1, 2, 3, 4, -3, -2, -1]) * 2 * np.pi / (8 * dx)
    assert np.allclose(g.kx1, expect, rtol=1e-14, atol=0)
    assert g.kx1[4] > 0
    assert g.kx.shape == (6, 8) and np.allclose(g.kx[0], expect)
    assert np.allclose(g.kk, (g.kx ** 2 + g.ky ** 2) / 2)
    assert np.isclose(g.weight, (34 / 7) * (10 / 5))
    g3 = Grid3D(12, 8, 6, 4, 3, 6)
    assert g3.kk.shape == (4, 8, 6)
    # meshgrid [ny, nx, nz]: x varies along axis 1, z along axis 2
    assert np.allclose(g3.mesh.x[0, :, 0], g3.x) and np.allclose(g3.mesh.z[0, 0, :], g3.z)


def test_kinetic_step_plane_wave():
    # a plane wave on the grid is an eigenvector of ifft(exp(-kk dt) fft(.))
    g = Grid3D(4, 16, 5, 12, 3, 8)
    i, j, k = 3, -2, 4
    kx, ky, kz = g.kx1[i], g.ky1[j], g.kz1[k]
    phi = np.exp(1j * (kx * g.mesh.x + ky * g.mesh.y + kz * g.mesh.z))
    t = GPEtask(g, 0 * g.mesh.x)
    dt = 0.37j
    t.set_kinop(dt)
    out = t.ssft_kin_step(phi, dt)
    expect = np.exp(-(kx ** 2 + ky ** 2 + kz ** 2) / 2 * dt) * phi
    assert rel_l2(out, expect) < 1e-13


def _two_wave_task(gam=0.0, g=0.0, ncomp_dims=2):
    grid = Grid2D(6, 16, 5, 12)
    c = 0.7
    t = GPEtask(grid, c + 0 * grid.mesh.x)
    t.g = g
    t.gamma = gam
    k1 = (grid.kx1[2], grid.ky1[1])
    k2 = (grid.kx1[-3], grid.ky1[3])
    phi = 0.8 * np.exp(1j * (k1[0] * grid.mesh.x + k1[1] * grid.mesh.y)) + \
        0.6 * np.exp(1j * (k2[0] * grid.mesh.x + k2[1] * grid.mesh.y))
    return t, grid, phi, k1, k2, c


def test_solve_split_linear_closed_form():
    """g=0, V=c: two plane waves evolve as e^{-i(E_j - mu) t}, mu = <H> / N.

    Exercises the block structure (open half phase, n_rec full phases, closing
    half phase), the kinetic operator and the initial chemical potential."""
    t, grid, phi, k1, k2, c = _two_wave_task()
    N0 = grid.integrate(np.abs(phi) ** 2)
    t.Ntotal = N0
    t.init_state = phi
    ddt, ni, no = 0.01, 20, 3
    out = t.solve_split(ddt, ni, no)
    E1 = (k1[0] ** 2 + k1[1] ** 2) / 2 + c
    E2 = (k2[0] ** 2 + k2[1] ** 2) / 2 + c
    mu = (0.64 * E1 + 0.36 * E2) / 1.0
    T = ddt * ni * no
    expect = 0.8 * np.exp(-1j * (E1 - mu) * T) * np.exp(1j * (k1[0] * grid.mesh.x + k1[1] * grid.mesh.y)) + \
        0.6 * np.exp(-1j * (E2 - mu) * T) * np.exp(1j * (k2[0] * grid.mesh.x + k2[1] * grid.mesh.y))
    assert rel_l2(out, expect) < 1e-12
    assert np.isclose(t.history["mu"][-1], mu, rtol=1e-12)
    assert t.current_iter == no
    # time2 is the start time of the last block (solve_split.m:53,80)
    assert np.isclose(t.current_time, (no - 1) * ddt * ni + (ni // 10 - 1) * ddt * 10)


def test_solve_split_norm_and_time_reversal():
    """gamma=0: norm conserved; B(-dt) inverts B(dt) exactly (Strang symmetry)."""
    grid = Grid2D(8, 32, 8, 32)
    t = GPEtask(grid, lambda X, Y, Z: 0.5 * (X ** 2 + 2 * Y ** 2))
    t.g = 3.0
    rng = np.random.Generator(np.random.PCG64(7))
    phi0 = np.exp(-(grid.mesh.x ** 2 + grid.mesh.y ** 2) / 4) * np.exp(1j * rng.uniform(0, 0.3, grid.size))
    t.Ntotal = grid.integrate(np.abs(phi0) ** 2)
    t.init_state = phi0
    t.mu_init = 1.3
    out = t.solve_split(0.005, 10, 4)
    assert abs(t.history["n"][-1] / t.Ntotal - 1) < 1e-13
    t2 = GPEtask(grid, t.Vtrap)
    t2.g, t2.Ntotal, t2.mu_init = t.g, t.Ntotal, t.mu_init
    t2.init_state = out
    back = t2.solve_split(-0.005, 10, 4)
    assert rel_l2(back, phi0) < 1e-12


def test_solve_split_strang_second_order():
    grid = Grid2D(8, 32, 8, 32)
    V = 0.5 * (grid.mesh.x ** 2 + grid.mesh.y ** 2)
    phi0 = np.exp(-((grid.mesh.x - 1) ** 2 + grid.mesh.y ** 2) / 2) + 0j

    def run(ddt, n):
        t = GPEtask(grid, V)
        t.g = 2.0
        t.Ntotal = 1.0
        t.init_state = np.sqrt(1.0) * grid.normalize(phi0)
        t.mu_init = 1.0
        return t.solve_split(ddt, n, 1)

    ref = run(0.2 / 640, 640)
    e1 = rel_l2(run(0.2 / 20, 20), ref)
    e2 = rel_l2(run(0.2 / 40, 40), ref)
    assert 3.3 < e1 / e2 < 4.7


def test_harmonic_coherent_state_oscillates():
    """g=0, V=(x^2+y^2)/2: a displaced ground state keeps its shape, <x>(t) = x0 cos t."""
    grid = Grid2D(10, 64, 10, 64)
    t = GPEtask(grid, lambda X, Y, Z: 0.5 * (X ** 2 + Y ** 2))
    t.g = 0.0
    t.Ntotal = 1.0
    x0 = 2.0
    phi0 = np.exp(-((grid.mesh.x - x0) ** 2 + grid.mesh.y ** 2) / 2) / np.sqrt(np.pi) + 0j
    t.init_state = phi0
    t.solve_split(0.001, 50, 20)
    phi = t.current_state
    T = 1.0
    xm = grid.integrate(grid.mesh.x * np.abs(phi) ** 2)
    assert abs(xm - x0 * np.cos(T)) < 1e-5


def test_solve_split_damped_renormalises():
    """gamma>0: every block renormalises to Ntotal and updates mu (solve_split.m:68-75)."""
    grid = Grid2D(8, 32, 8, 32)
    t = GPEtask(grid, lambda X, Y, Z: 0.5 * (X ** 2 + Y ** 2))
    t.g = 1.0
    t.gamma = 0.05
    t.Ntotal = 10.0
    phi0 = np.exp(-((grid.mesh.x - 1) ** 2 + grid.mesh.y ** 2) / 2) + 0j
    t.init_state = np.sqrt(10.0) * grid.normalize(phi0)
    out = t.solve_split(0.01, 20, 3)
    assert abs(grid.integrate(np.abs(out) ** 2) / 10.0 - 1) < 1e-13
    # mu equals <phi|H phi>/N of the renormalised state at the last block
    mu = np.real(grid.inner(out, t.applyham(out))) / 10.0
    assert np.isclose(t.history["mu"][-1], mu, rtol=1e-12)
    # decay_rate: NNN = NN0 exp(-time2/tau)
    t2 = GPEtask(grid, t.Vtrap)
    t2.g, t2.gamma, t2.Ntotal, t2.decay_rate = 1.0, 0.05, 10.0, 2.0
    t2.init_state = t.init_state
    out2 = t2.solve_split(0.01, 20, 2)
    assert np.isclose(grid.integrate(np.abs(out2) ** 2), 10.0 * np.exp(-t2.current_time / 2.0), rtol=1e-13)


def test_n_rec_fallback():
    # niter_inner not a multiple of n_recalc -> n_rec = 1 (solve_split.m:25-27)
    t, grid, phi, *_ = _two_wave_task(g=0.5)
    t.Ntotal = grid.integrate(np.abs(phi) ** 2)
    t.init_state = phi
    t.solve_split(0.01, 7, 2)
    assert np.isclose(t.current_time, 0.07 + 6 * 0.01)


def test_itp_harmonic_ground_state():
    """g=0 ITP converges to the oscillator ground state: mu = 1, |phi| Gaussian."""
    grid = Grid2D(8, 64, 8, 64)
    t = GPEtask(grid, lambda X, Y, Z: 0.5 * (X ** 2 + Y ** 2))
    t.g = 0.0
    t.Ntotal = 1.0
    phi, MU, MU2, EE = t.groundstate_itp(0.01, 1e-9, "rand", nargout=4)
    gs = np.exp(-(grid.mesh.x ** 2 + grid.mesh.y ** 2) / 2) / np.sqrt(np.pi)
    assert rel_l2(np.abs(phi), gs) < 1e-3
    assert abs(MU[-1] - 1.0) < 1e-3
    assert abs(MU2[-1] - 1.0) < 1e-3
    assert abs(EE[-1] - 1.0) < 1e-3


def test_itp_thomas_fermi_limit():
    grid = Grid2D(17, 128, 17, 128)
    t = GPEtask(grid, lambda X, Y, Z: 0.5 * (X ** 2 + Y ** 2))
    t.g = 0.06
    t.Ntotal = 1e5
    phi, MU = t.groundstate_itp(0.02, 1e-6, "tf", nargout=2)
    mu_tf = np.sqrt(t.g * t.Ntotal / np.pi)
    assert abs(MU[-1] / mu_tf - 1) < 0.02  # O(1) kinetic correction on mu ~ 44
    assert np.isclose(grid.integrate(np.abs(phi) ** 2), 1e5, rtol=1e-13)


def test_groundstate_tf_closed_form():
    # 2D harmonic: integrate((mu-V)_+) = pi mu^2  ->  mu = sqrt(gN/pi)
    grid = Grid2D(17, 256, 17, 256)
    t = GPEtask(grid, lambda X, Y, Z: 0.5 * (X ** 2 + Y ** 2))
    t.g = 0.06
    t.Ntotal = 1e5
    phi, mu = t.groundstate_tf(np.finfo(float).eps)
    assert abs(mu / np.sqrt(0.06 * 1e5 / np.pi) - 1) < 2e-3
    assert np.isclose(grid.integrate(np.abs(phi) ** 2) * t.g, t.g * 1e5, rtol=1e-3)
    # 3D anisotropic: mu = 0.5 (15 g N wx wy wz / 4 pi)^(2/5)
    g3 = Grid3D(12, 64, 12, 64, 12, 64)
    t3 = GPEtask(g3, lambda X, Y, Z: 0.5 * (X ** 2 + Y ** 2 + 16 * Z ** 2))
    t3.g, t3.Ntotal = 0.02, 1e6
    _, mu3 = t3.groundstate_tf(np.finfo(float).eps)
    assert abs(mu3 / (0.5 * (15 * 0.02 * 1e6 * 4 / (4 * np.pi)) ** 0.4) - 1) < 2e-2


def test_spinor_rabi_closed_form():
    """V=0, g=0, uniform real Omega: imbalance (n1-n2)/N = cos(2 Omega t) exactly."""
    grid = Grid2D(5, 8, 5, 8)
    t = GPEtaskSpinor(grid, [0 * grid.mesh.x, 0 * grid.mesh.x])
    t.g = np.zeros((2, 2))
    t.coupling = 0.5
    t.Ntotal = grid.integrate(np.ones(grid.size))
    t.init_state = [np.ones(grid.size, complex), np.zeros(grid.size, complex)]
    out = t.solve_split(0.01, 30, 4)
    T = 1.2
    n1 = grid.integrate(np.abs(out[0]) ** 2)
    n2 = grid.integrate(np.abs(out[1]) ** 2)
    assert abs((n1 - n2) / t.Ntotal - np.cos(2 * 0.5 * T)) < 1e-13
    assert rel_l2(out[1], -1j * np.sin(0.5 * T) * np.ones(grid.size)) < 1e-13


def test_spinor_norm_and_reduction_to_scalar():
    """Complex space-dependent Omega conserves the total norm (Hermitian coupling);
    with Omega=0 and g12=0 each component follows the scalar solver's kinetic and
    phase operators (checked on one block with mu fixed to zero)."""
    grid = Grid3D(4, 16, 4, 16, 4, 12)
    V = 0.5 * (grid.mesh.x ** 2 + grid.mesh.y ** 2 + grid.mesh.z ** 2)
    t = GPEtaskSpinor(grid, [V, V])
    t.g = np.array([[0.5, 0.4], [0.4, 0.5]])
    t.coupling = 0.5 * np.exp(2j * grid.mesh.x)
    a = np.exp(-(grid.mesh.x ** 2 + grid.mesh.y ** 2 + grid.mesh.z ** 2) / 2) + 0j
    t.Ntotal = 2 * grid.integrate(np.abs(a) ** 2)
    t.init_state = [a, a * np.exp(0.3j * grid.mesh.y)]
    out = t.solve_split(0.005, 10, 2)
    n = grid.integrate(np.abs(out[0]) ** 2 + np.abs(out[1]) ** 2)
    assert abs(n / t.Ntotal - 1) < 1e-13


def _two_comp_itp_task(cls, n=32, xmax=6.0, g=((0.0, 0.0), (0.0, 0.0))):
    grid = oracle.grid2d(xmax, n, xmax, n)
    V = 0.5 * (grid.mesh.x ** 2 + grid.mesh.y ** 2)
    t = cls(grid, [V, V])
    t.g = np.array(g)
    return t, grid


def test_spinor_itp_coupled_oscillator_ground_state():
    # g = 0, uniform Omega: H = h_osc (x) 1 + [[0, Omega], [conj(Omega), 0]]; the ground
    # state has mu = 1 - |Omega| and equal populations with psi2/psi1 = -exp(-i arg Omega)
    t, grid = _two_comp_itp_task(oracle.GPEtaskSpinor)
    om = 0.3 * np.exp(0.7j)
    t.coupling = om
    t.Ntotal = 2.0
    phi, MU, MU2, EE = t.groundstate_itp(0.02, 1e-9, nargout=4)
    assert t.itp_converged
    assert abs(MU[-1] - (1 - abs(om))) < 1e-6
    assert abs(MU2[-1] - (1 - abs(om))) < 1e-6 and abs(EE[-1] - (1 - abs(om))) < 1e-6
    # the energy converges quadratically in the state error, so the state is held
    # to the square root of the energy tolerance
    n1, n2 = grid.integrate(np.abs(phi[0]) ** 2), grid.integrate(np.abs(phi[1]) ** 2)
    assert abs(n1 - 1.0) < 2e-5 and abs(n2 - 1.0) < 2e-5 and abs(n1 + n2 - 2.0) < 1e-12
    ratio = grid.inner(phi[0], phi[1]) / n1
    assert abs(ratio - (-np.exp(-1j * np.angle(om)))) < 2e-5


def test_2comp_itp_masses_and_separate_norms():
    # g = 0: each component relaxes to its own oscillator ground state, energy 1/sqrt(M)
    # in 2D; each is renormalised to its own particle number
    t, grid = _two_comp_itp_task(oracle.GPEtask2comp)
    t.N1, t.N2, t.M2 = 3.0, 0.5, 2.0
    phi, MU1, MU2 = t.groundstate_itp(0.02, 1e-9, nargout=3)
    assert t.itp_converged
    assert abs(MU1[-1] - 1.0) < 1e-6
    assert abs(MU2[-1] - 1.0 / np.sqrt(2.0)) < 1e-6
    assert abs(grid.integrate(np.abs(phi[0]) ** 2) - 3.0) < 1e-12
    assert abs(grid.integrate(np.abs(phi[1]) ** 2) - 0.5) < 1e-12


def test_two_comp_itp_history_lengths_and_cap():
    # the histories hold entries 1..i-1 of the i iterations run (groundstate_itp.m:131-142)
    t, _ = _two_comp_itp_task(oracle.GPEtaskSpinor, n=16)
    t.coupling, t.Ntotal = 0.2, 1.0
    _, MU = t.groundstate_itp(0.01, 1e-30, nargout=2, max_iter=37)
    assert not t.itp_converged and t.itp_iterations == 37 and len(MU) == 36


def test_2comp_solve_split_is_uncoupled_spinor_step():
    # the reference GPEtask2comp.solve_split is GPEtaskSpinor's with coupling 0
    grid = oracle.grid2d(5, 16, 5, 16)
    V = 0.5 * (grid.mesh.x ** 2 + grid.mesh.y ** 2)
    rng = np.random.Generator(np.random.PCG64(8))
    psi = [np.exp(-(grid.mesh.x ** 2 + grid.mesh.y ** 2) / 2) * np.exp(1j * rng.uniform(0, 1, grid.size))
           for _ in range(2)]
    a, b = oracle.GPEtask2comp(grid, [V, V]), oracle.GPEtaskSpinor(grid, [V, V])
    for t in (a, b):
        t.g = np.array([[0.3, 0.2], [0.2, 0.4]])
        t.Ntotal = 2.0
        t.init_state = [p.copy() for p in psi]
    pa, pb = a.solve_split(0.01, 10, 2), b.solve_split(0.01, 10, 2)
    assert np.array_equal(pa[0], pb[0]) and np.array_equal(pa[1], pb[1])


def _lz_spectral(grid, phi):
    # Lz = -i (x d/dy - y d/dx) (lib/@grid2d/lz.m) with spectral derivatives, numpy.fft
    kx = grid.kx1[None, :]
    ky = grid.ky1[:, None]
    dx = np.fft.ifft(1j * kx * np.fft.fft(phi, axis=1), axis=1)
    dy = np.fft.ifft(1j * ky * np.fft.fft(phi, axis=0), axis=0)
    return -1j * (grid.mesh.x * dy - grid.mesh.y * dx)


def test_rotating_lap_is_lap_plus_omega_lz():
    # grid2d/lap.m with omega: the cross terms -kx y omega and +ky x omega are omega*Lz
    grid = oracle.grid2d(6, 32, 5, 24)
    rng = np.random.Generator(np.random.PCG64(4))
    phi = np.exp(-(grid.mesh.x ** 2 + grid.mesh.y ** 2) / 3) * (1 + 0.3 * rng.normal(size=grid.size))
    om = 0.37
    assert rel_l2(grid.lap(phi, om) - grid.lap(phi), om * _lz_spectral(grid, phi)) < 1e-12


@pytest.mark.parametrize("m", [1, -1])
def test_rotating_frame_vortex_eigenstate_is_stationary(m):
    # g = 0, V = (x^2+y^2)/2: (x + i m y) e^{-r^2/2} has H = 2 and Lz = m, so in the frame
    # of lap(phi, omega) its energy is 2 + m omega; solve_split subtracts mu = <H_omega>,
    # so the state is stationary up to the splitting error O(dt^2)
    grid = oracle.grid2d(7, 48, 7, 48)
    V = 0.5 * (grid.mesh.x ** 2 + grid.mesh.y ** 2)
    phi0 = (grid.mesh.x + 1j * m * grid.mesh.y) * np.exp(-(grid.mesh.x ** 2 + grid.mesh.y ** 2) / 2)
    errs = []
    for dt in (0.01, 0.005):
        t = GPEtask(grid, V)
        t.g, t.omega, t.Ntotal = 0.0, 0.4, 1.0
        t.init_state = grid.normalize(phi0)
        out = t.solve_split(dt, 20, int(round(1.0 / (20 * dt))))
        assert abs(t.history["mu"][0] - (2 + m * 0.4)) < 1e-10
        errs.append(rel_l2(out, t.init_state))
    assert errs[0] < 1e-4 and 3.0 < errs[0] / errs[1] < 5.0   # second order in dt


def test_rotating_frame_itp_keeps_the_gaussian_ground_state():
    # for omega < 1 the ground state of H + omega Lz in the isotropic trap is still the
    # Gaussian (Lz = 0), mu = 1
    grid = oracle.grid2d(6, 32, 6, 32)
    t = GPEtask(grid, 0.5 * (grid.mesh.x ** 2 + grid.mesh.y ** 2))
    t.g, t.omega, t.Ntotal = 0.0, 0.5, 1.0
    t.itp_adjust_stepsize = True
    phi, MU = t.groundstate_itp(0.01, 1e-7, "rand", nargout=2)
    assert t.itp_iterations < t.itp_max_iter
    assert abs(MU[-1] - 1.0) < 1e-4
    gauss = grid.normalize(np.exp(-(grid.mesh.x ** 2 + grid.mesh.y ** 2) / 2) + 0j)
    assert abs(abs(grid.inner(gauss, phi)) - 1.0) < 1e-6


def test_philox_known_answers():
    # Random123 known-answer vectors for philox4x32-10
    from oracle.philox import philox4x32_10
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, want in kat:
        out = philox4x32_10(*[np.uint32(c) for c in ctr], *[np.uint32(k) for k in key])
        assert tuple(int(o) for o in out) == want


def test_complex_normals_statistics():
    from oracle.philox import complex_normals
    n = 1 << 20
    re, im = complex_normals(n, 12345, 7)
    for x in (re, im):
        assert abs(x.mean()) < 5 / np.sqrt(n)
        assert abs(x.var() - 1) < 5 * np.sqrt(2 / n)
    assert abs(np.mean(re * im)) < 5 / np.sqrt(n)
    re2, _ = complex_normals(n, 12346, 7)
    assert abs(np.mean(re * re2)) < 5 / np.sqrt(n)       # neighbouring steps are independent


def _sgpe(grid, V, T, g=0.3, gamma=0.05, ecut=0.0, mu=2.0, seed=11):
    t = oracle.SGPEtask(grid, V)
    t.g, t.gamma, t.T, t.ecut, t.mu_init, t.noise_seed = g, gamma, T, ecut, mu, seed
    r2 = grid.mesh.x ** 2 + grid.mesh.y ** 2
    t.init_state = np.exp(-r2 / 2) * (1 + 0.1j * grid.mesh.x)
    return t


def test_sgpe_at_zero_temperature_is_the_damped_gpe_with_fixed_mu():
    grid = oracle.grid2d(6, 32, 6, 32)
    V = 0.5 * (grid.mesh.x ** 2 + grid.mesh.y ** 2)
    s = _sgpe(grid, V, T=0.0)
    ref = GPEtask(grid, V)
    ref.g, ref.gamma, ref.mu_init = 0.3, 0.05, 2.0        # mu_init > 0: no renormalisation
    ref.init_state = s.init_state.copy()
    assert rel_l2(s.solve_split(0.01, 20, 3), ref.solve_split(0.01, 20, 3)) < 1e-13


def test_sgpe_noise_response_is_linear_in_sqrt_temperature():
    # g = 0: the step is linear, so psi(T) - psi(0) scales with sqrt(T gamma ddt / weight)
    grid = oracle.grid2d(6, 32, 6, 32)
    V = 0.5 * (grid.mesh.x ** 2 + grid.mesh.y ** 2)
    out = {T: _sgpe(grid, V, T, g=0.0).solve_split(0.01, 10, 2) for T in (0.0, 0.5, 2.0)}
    d1, d4 = out[0.5] - out[0.0], out[2.0] - out[0.0]
    assert rel_l2(d4, 2 * d1) < 1e-12
    assert np.linalg.norm(d1) > 1e-3 * np.linalg.norm(out[0.0])


def test_sgpe_energy_cutoff_empties_the_masked_modes():
    # V = 0, g = 0: every inner step ends with the masked kinetic factor followed by
    # uniform phases, so no power is left at kk >= ecut (up to the FFT round trip)
    grid = oracle.grid2d(6, 32, 6, 32)
    s = _sgpe(grid, 0.0 * grid.mesh.x, T=1.0, g=0.0, ecut=4.0)
    phi = s.solve_split(0.01, 10, 1)
    spec = np.abs(grid.fft(phi)) ** 2
    assert spec[grid.kk >= 4.0].max() < 1e-25 * spec.max() and spec[grid.kk < 4.0].max() > 0.0


# --------------------------------------------------------------- complex Omega and g/2
# The Rabi step's complex phase convention (GPEtaskSpinor/solve_split.m:61-65, 90-91
# against the Hamiltonian of GPEtaskSpinor.m:66-70) and the halved interaction of
# get_energy (GPEtask.m:97-107, GPEtaskSpinor.m:100-111), pinned by closed forms.
from closed_forms import (grid_wavenumber, rabi_two_mode, two_mode_state, uniform_energy,  # noqa: E402
                          uniform_spinor_energy)


def test_spinor_complex_omega_is_exp_minus_iHT():
    """V = g = 0, uniform psi1 = 1, psi2 = 0, Omega = |Omega| e^{i th}: the step is
    exp(-i H T) for H = [[0, Omega], [conj(Omega), 0]] exactly (products of rotations),
    so psi1 = cos(|Omega| T), psi2 = -i e^{-i th} sin(|Omega| T), and mu stays 0
    (Re(Omega conj(psi1) psi2) = 0).  Swapping e^{+i th} and e^{-i th} gives psi2 a
    phase e^{+i th} and fails."""
    grid = Grid2D(5, 8, 5, 8)
    t = GPEtaskSpinor(grid, [0 * grid.mesh.x, 0 * grid.mesh.x])
    t.g = np.zeros((2, 2))
    r, th = 0.5, 0.7
    t.coupling = r * np.exp(1j * th)
    t.Ntotal = grid.integrate(np.ones(grid.size))
    t.init_state = [np.ones(grid.size, complex), np.zeros(grid.size, complex)]
    out = t.solve_split(0.01, 30, 4)
    T = 1.2
    assert rel_l2(out[0], np.cos(r * T) * np.ones(grid.size)) < 1e-13
    assert rel_l2(out[1], -1j * np.exp(-1j * th) * np.sin(r * T) * np.ones(grid.size)) < 1e-13
    assert np.all(np.abs(t.history["mu"]) < 1e-14)


@pytest.mark.parametrize("r,m,th,gamma", [(0.5, 2, 0.7, 0.0), (0.5, -3, -2.1, 0.0), (2.0, 1, 0.3, 0.02),
                                          (500.0, 2, 0.4, 0.0)])
def test_spinor_space_dependent_omega_two_mode_reduction(r, m, th, gamma):
    """Omega = r e^{i(qx + th)} with q a grid wave number keeps psi in the two-mode
    subspace (psi1 = a, psi2 = b e^{-i(qx+th)}); the oracle must equal the 2x2 product
    of the reduction (tests/closed_forms.py), including mu, n and the damped
    renormalisation.  r = 500 puts |Omega| dt = 1 (the large-angle regime)."""
    grid = Grid2D(5, 16, 4, 12)
    q = grid_wavenumber(16, 5, m)
    t = GPEtaskSpinor(grid, [0 * grid.mesh.x, 0 * grid.mesh.x])
    t.g = np.zeros((2, 2))
    t.gamma = gamma
    t.coupling = r * np.exp(1j * (q * grid.mesh.x + th))
    vol = 16 * 12 * grid.dx * grid.dy
    a0, b0 = 0.8, 0.6j
    t.Ntotal = (abs(a0) ** 2 + abs(b0) ** 2) * vol
    t.init_state = list(two_mode_state(grid, q, th, a0, b0))
    ddt = 0.002
    out = t.solve_split(ddt, 20, 3)
    a, b, mus, ns = rabi_two_mode(r, q, ddt, 20, 3, 10, vol, t.Ntotal, gamma, a0, b0)
    e1, e2 = two_mode_state(grid, q, th, a, b)
    assert rel_l2(out[0], e1) < 1e-12 and rel_l2(out[1], e2) < 1e-12
    assert np.allclose(t.history["mu"], mus, rtol=1e-11, atol=1e-11 * abs(mus).max())
    assert np.allclose(t.history["n"], ns, rtol=1e-12)


def test_two_mode_reduction_matches_exp_minus_iHT():
    """The reduction itself: q = 0, gamma = 0, mu = 0 gives a = cos(rT), b = -i sin(rT)."""
    a, b, mus, _ = rabi_two_mode(0.5, 0.0, 0.01, 30, 4, 10, 1.0, 1.0)
    assert abs(a - np.cos(0.6)) < 1e-14 and abs(b + 1j * np.sin(0.6)) < 1e-14
    assert np.all(np.abs(mus) < 1e-15)


@pytest.mark.parametrize("dims", [2, 3])
def test_get_energy_halves_g_uniform_state(dims):
    """Uniform psi of norm N in V = 0 on the periodic box: get_energy = g N^2/(2 Vol),
    <psi|H psi> = g N^2/Vol (the kinetic term of a constant is 0)."""
    if dims == 2:
        grid = Grid2D(5, 16, 4, 12)
        vol = 16 * 12 * grid.dx * grid.dy
    else:
        grid = Grid3D(5, 16, 4, 12, 3, 8)
        vol = 16 * 12 * 8 * grid.dx * grid.dy * grid.dz
    t = GPEtask(grid, 0 * grid.mesh.x)
    t.g = 0.7
    N = 3.0
    psi = np.sqrt(N / vol) * np.exp(0.3j) * np.ones(grid.size)
    E, M = uniform_energy(t.g, N, vol)
    assert abs(t.get_energy(psi) / E - 1) < 1e-13
    assert abs(np.real(grid.inner(psi, t.applyham(psi))) / M - 1) < 1e-13


def test_get_energy_interaction_identity():
    """Any state: get_energy - <psi|H psi> = -(g/2) sum |psi|^4 w at g != 0."""
    grid = Grid2D(6, 32, 5, 24)
    t = GPEtask(grid, lambda X, Y, Z: 0.5 * (X ** 2 + 2 * Y ** 2))
    t.g = 0.9
    rng = np.random.Generator(np.random.PCG64(5))
    psi = np.exp(-(grid.mesh.x ** 2 + grid.mesh.y ** 2) / 3) * np.exp(1j * rng.uniform(0, 1, grid.size))
    w = grid.dx * grid.dy
    E = t.get_energy(psi)
    M = np.real(grid.inner(psi, t.applyham(psi)))
    assert abs((E - M) / (-0.5 * t.g * np.sum(np.abs(psi) ** 4) * w) - 1) < 1e-12


def test_spinor_get_energy_uniform_closed_form():
    """Uniform components a, b in V = 0 with uniform Omega: (1/2) sum_ik g_ik n_i n_k / Vol
    plus the coupling energy 2 Re(Omega conj(a) b) Vol (not halved)."""
    grid = Grid2D(5, 16, 4, 12)
    vol = 16 * 12 * grid.dx * grid.dy
    t = GPEtaskSpinor(grid, [0 * grid.mesh.x, 0 * grid.mesh.x])
    gm = np.array([[0.5, 0.45], [0.4, 0.6]])
    t.g = gm
    om = 0.3 * np.exp(0.9j)
    t.coupling = om
    a, b = 0.7 * np.exp(0.2j), 0.4 * np.exp(-1.1j)
    phi = [a * np.ones(grid.size), b * np.ones(grid.size)]
    n1, n2 = abs(a) ** 2 * vol, abs(b) ** 2 * vol
    E = uniform_spinor_energy(gm, n1, n2, om, a, b, vol)
    assert abs(t.get_energy(phi) / E - 1) < 1e-13


def _plane_wave(grid, idx):
    ks = (grid.kx1[idx[0]], grid.ky1[idx[1]]) + ((grid.kz1[idx[2]],) if grid.ndims == 3 else ())
    arg = ks[0] * grid.mesh.x + ks[1] * grid.mesh.y + (ks[2] * grid.mesh.z if grid.ndims == 3 else 0)
    return np.exp(1j * arg), 0.5 * sum(k * k for k in ks)


@pytest.mark.parametrize("dims", [2, 3])
def test_applyham_plane_wave_eigenvalue(dims):
    """A plane wave A e^{ik.x} in a uniform potential c: applyh0 = (k^2/2 + c) psi and
    applyham = (k^2/2 + c + g |A|^2) psi (GPEtask.m:73-95; the Nyquist index included)."""
    grid = Grid2D(5, 16, 4, 12) if dims == 2 else Grid3D(5, 16, 4, 12, 3, 8)
    idx = (3, -2, 4) if dims == 3 else (8, -5)
    e, kk = _plane_wave(grid, idx)
    A = 0.8 * np.exp(0.4j)
    c = 0.35
    t = GPEtask(grid, c + 0 * grid.mesh.x)
    t.g = 1.3
    psi = A * e
    assert rel_l2(t.applyh0(psi), (kk + c) * psi) < 1e-13
    assert rel_l2(t.applyham(psi), (kk + c + t.g * abs(A) ** 2) * psi) < 1e-13


def test_spinor_applyham_plane_wave_closed_form():
    """psi_j = c_j e^{ik.x}, uniform V_j, uniform complex Omega: component 1 couples
    through Omega, component 2 through conj(Omega) (GPEtaskSpinor.m:55-76), and each takes
    sum_k g_jk |c_k|^2 with the matrix read as g(j, k)."""
    grid = Grid2D(5, 16, 4, 12)
    e, kk = _plane_wave(grid, (2, 3))
    t = GPEtaskSpinor(grid, [0.2 + 0 * grid.mesh.x, -0.1 + 0 * grid.mesh.x])
    gm = np.array([[0.5, 0.45], [0.3, 0.6]])
    t.g = gm
    om = 0.3 * np.exp(0.9j)
    t.coupling = om
    a, b = 0.7 * np.exp(0.2j), 0.4 * np.exp(-1.1j)
    h1, h2 = t.applyham([a * e, b * e])
    n = np.array([abs(a) ** 2, abs(b) ** 2])
    assert rel_l2(h1, ((kk + 0.2 + gm[0] @ n) * a + om * b) * e) < 1e-13
    assert rel_l2(h2, ((kk - 0.1 + gm[1] @ n) * b + np.conj(om) * a) * e) < 1e-13


@pytest.mark.parametrize("m", [1, -1])
def test_rotating_applyh0_vortex_eigenvalue(m):
    """(x + i m y) e^{-r^2/2} in V = r^2/2: applyh0 with a scalar omega gives
    (2 + m omega) psi up to the grid's spectral error (GPEtask.m:85-89, grid2d/lap.m:8-9)."""
    grid = oracle.grid2d(8, 64, 8, 64)
    V = 0.5 * (grid.mesh.x ** 2 + grid.mesh.y ** 2)
    psi = (grid.mesh.x + 1j * m * grid.mesh.y) * np.exp(-(grid.mesh.x ** 2 + grid.mesh.y ** 2) / 2)
    t = GPEtask(grid, V)
    t.omega = 0.4
    assert rel_l2(t.applyh0(psi), (2 + m * 0.4) * psi) < 1e-10
