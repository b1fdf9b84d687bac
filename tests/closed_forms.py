"""Closed forms shared by the oracle pins and their GPU twins (test infrastructure).

Two-mode reduction of GPEtaskSpinor.solve_split (lib/@GPEtaskSpinor/solve_split.m:13-121)
-----------------------------------------------------------------------------------
With V = 0, g = 0 and Omega(x) = r exp(i (q x + th)), where q is a wave number of the
periodic grid (q N dx = 2 pi m), the states

    psi1 = a,    psi2 = b exp(-i (q x + th))        (a, b uniform complex amplitudes)

form an invariant two-dimensional subspace of every operator of the step:
  * H psi1 = lap(psi1) + Omega psi2 = r b,
    H psi2 = lap(psi2) + conj(Omega) psi1 = (q^2/2 b + r a) exp(-i (q x + th))
    (GPEtaskSpinor.m:55-75: component 1 couples through Omega, component 2 through
    its conjugate);
  * the kinetic factor exp(-kk dt) multiplies a by 1 and b by exp(-q^2/2 dt);
  * the Rabi matrix cosh(dt r) I - sinh(dt r) [[0, e^{i phi}], [e^{-i phi}, 0]]
    (solve_split.m:61-65, 90-91) acts on (a, b) as [[C, -S], [-S, C]] with
    C = cosh(dt r), S = sinh(dt r): the phase e^{i phi} of Omega cancels against the
    phase of psi2 exactly when component 1 takes Omega and component 2 takes
    conj(Omega).  With the two swapped, psi2 would leave the subspace;
  * the potential phase is the uniform factor exp(mu dt).

So the split-step's output is a product of 2x2 matrices, computed here from this
derivation alone; the block structure (half kinetic step, n_rec x [phase . Rabi,
kinetic], closing half step, renormalisation when gamma > 0, mu from <psi|H psi>)
follows the cited lines.  The exact (continuous-time) answer for q = 0, gamma = 0,
mu = 0 is the textbook exp(-i H T): a = cos(rT), b = -i sin(rT).
"""
import numpy as np


def rabi_two_mode(r, q, ddt, niter_inner, niter_outer, n_rec, vol, ntotal, gamma=0.0, a0=1.0, b0=0.0):
    """Amplitudes (a, b) after solve_split(ddt, niter_inner, niter_outer) and the mu
    history (one value per outer iteration), for the subspace described above."""
    dt = ddt * 1j / (1 + 1j * gamma)
    kk = 0.5 * q * q
    C, S = np.cosh(dt * r), np.sinh(dt * r)
    v = np.array([a0, b0], dtype=np.complex128)

    def mu_of(v, nnn):
        # <psi|H psi> / N with the two-mode H = [[0, r], [r, q^2/2]] per unit volume
        hv = np.array([r * v[1], r * v[0] + kk * v[1]])
        return (np.vdot(v, hv) * vol) / nnn

    NNN = ntotal
    mu = np.real(mu_of(v, ntotal))
    mus, ns = [], []
    ncur = ntotal
    for _j in range(niter_outer):
        for _jj in range(niter_inner // n_rec):
            v = v * np.array([1.0, np.exp(-kk * dt * 0.5)])
            for _ in range(n_rec):
                v = np.exp(mu * dt) * np.array([C * v[0] - S * v[1], C * v[1] - S * v[0]])
                v = v * np.array([1.0, np.exp(-kk * dt)])
            v = v * np.array([1.0, np.exp(kk * dt * 0.5)])
            ncur = float(np.real(np.vdot(v, v)) * vol)
            if gamma > 0:
                v = v * np.sqrt(NNN / ncur)
            mu = mu_of(v, NNN)
        mus.append(mu)
        ns.append(ncur)
    return v[0], v[1], np.array(mus), np.array(ns)


def two_mode_state(grid, q, th, a, b):
    """(psi1, psi2) on the grid for amplitudes (a, b)."""
    ph = np.exp(-1j * (q * grid.mesh.x + th))
    return a * np.ones(grid.size, np.complex128), b * ph


def grid_wavenumber(n, xmax, m):
    """The m-th wave number of an n-point axis from linspace(-xmax, xmax, n)."""
    dx = 2 * xmax / (n - 1)
    return 2 * np.pi * m / (n * dx)


def uniform_energy(g, ntot, vol):
    """get_energy of a uniform state of norm ntot in V = 0 on a periodic box of volume
    vol: only the interaction term survives, (g/2) integral |psi|^4 = g ntot^2 / (2 vol)
    (GPEtask.m:97-107 halves g).  The chemical-potential integral <psi|H psi> has the
    full g: g ntot^2 / vol."""
    return 0.5 * g * ntot * ntot / vol, g * ntot * ntot / vol


def uniform_spinor_energy(gmat, n1, n2, omega, a, b, vol):
    """get_energy of uniform components (amplitudes a, b; norms n_i = |amp|^2 vol) in
    V = 0 with a uniform coupling omega: (1/2) sum_ik g_ik n_i n_k / vol from the halved
    g (GPEtaskSpinor.m:100-111) plus the coupling energy 2 Re(omega conj(a) b) vol,
    which the halving does not touch."""
    gm = np.asarray(gmat)
    n = np.array([n1, n2])
    return 0.5 * float(n @ gm @ n) / vol + 2 * np.real(omega * np.conj(a) * b) * vol
