"""FP64 discrete adjoint of GPEtask.solve_split for a time-dependent control potential.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  The reference has no adjoint;
this module defines the gradient the GPU entry point must reproduce, written from
the derivation below and independently of the GPU algorithm (it stores every
intermediate state instead of recomputing them).

Control problem (the differentiation configuration):
    V(x, t) = V0(x) + sum_k c[k, m] G_k(x),   m = interval of the block holding t
    forward = solve_split.m:49-79 with gamma = 0 (every block applies
              P(+h/2) [K P(+h)]^n P(-h/2) with the density of each map's input;
              V is evaluated at the block start, solve_split.m:54)
    loss    = 1 - |<psi_target | psi(T)>|^2,  <a|b> = sum conj(a) b dV (grid1d/inner.m:12)

Adjoint (lambda = 2 dL/d conj(psi), so dL = Re sum conj(lambda) d psi):
    phase map  w = e^{-i theta} u,  theta = tau (V - mu + g |u|^2),  |w| = |u|
        lambda_u = e^{+i theta} (lambda_w - 2 tau g Im(conj(w) lambda_w) w)
        dL/dV(x) += tau Im(conj(lambda_w) w)
    kinetic    w = F^-1 e^{-kk dt} F u  (unitary):  lambda_u = F^-1 e^{+kk dt} F lambda_w
    terminal   lambda_T = -2 <psi_target|psi_T> psi_target dV
    dL/dc[k, m] = sum over the blocks of interval m, sum_x G_k(x) dL/dV(x)
The chemical potential only adds a global phase, which the loss ignores, so it is
held fixed at its initial value (solve_split.m:41-46 with gamma = 0).
"""
import numpy as np


class ControlProblem:
    def __init__(self, grid, V0, basis, g, ddt, nsteps, steps_per_interval, psi0, Ntotal, n_rec=10, mu_init=0.0):
        if nsteps % n_rec or steps_per_interval % n_rec:
            raise ValueError("nsteps and steps_per_interval must be multiples of n_rec")
        self.grid = grid
        self.V0 = V0
        self.G = np.asarray(basis)
        self.g = g
        self.ddt = ddt
        self.nsteps = nsteps
        self.spi = steps_per_interval
        self.n_rec = n_rec
        self.psi0 = psi0
        self.Ntotal = Ntotal
        self.mu_init = mu_init
        self.kinop = np.exp(-grid.kk * (1j * ddt))          # set_kinop.m:9 with dt = i ddt
        self.nint = (nsteps + steps_per_interval - 1) // steps_per_interval

    def potential(self, c, m):
        V = self.V0
        for k in range(self.G.shape[0]):
            V = V + c[k, m] * self.G[k]
        return V

    def vtd(self, c):
        """Vtd(task, t) of GPEtask.m:65-67 for this control (interval from the step index)."""
        def f(task, t):
            m = int(round(t / self.ddt)) // self.spi
            return self.potential(c, m) - self.V0
        return f

    def mu0(self, c):
        gr = self.grid
        V = self.potential(c, 0)
        phi = self.psi0
        h = V * phi + gr.lap(phi) + self.g * np.abs(phi) ** 2 * phi
        muc = np.real(gr.inner(phi, h)) / self.Ntotal     # solve_split.m:41
        return self.mu_init if self.mu_init > 0 else muc

    def _phase(self, u, V, mu, tau):
        theta = tau * (V - mu + self.g * np.abs(u) ** 2)
        return np.exp(-1j * theta) * u

    def forward(self, c, keep=False):
        gr = self.grid
        mu = self.mu0(c)
        h = self.ddt
        phi = self.psi0.copy()
        tape = []
        for b in range(self.nsteps // self.n_rec):
            V = self.potential(c, (b * self.n_rec) // self.spi)
            phi = self._phase(phi, V, mu, h / 2)                      # solve_split.m:55-56
            if keep:
                tape.append(("P", b, h / 2, phi))
            for _ in range(self.n_rec):
                phi = gr.ifft(self.kinop * gr.fft(phi))               # 59
                phi = self._phase(phi, V, mu, h)                       # 60-61
                if keep:
                    tape.append(("P", b, h, phi))
            phi = self._phase(phi, V, mu, -h / 2)                     # 63-64
            if keep:
                tape.append(("P", b, -h / 2, phi))
        return (phi, tape, mu) if keep else phi

    def loss(self, c, psi_t):
        phi = self.forward(c)
        o = self.grid.inner(psi_t, phi)
        return 1.0 - abs(o) ** 2

    def loss_and_grad(self, c, psi_t):
        gr = self.grid
        phiT, tape, mu = self.forward(c, keep=True)
        o = gr.inner(psi_t, phiT)
        L = 1.0 - abs(o) ** 2
        lam = -2.0 * o * psi_t * gr.weight
        grad = np.zeros_like(np.asarray(c, dtype=float))
        gV = np.zeros(gr.size)
        cur_m = None
        # walk the tape backwards; every entry is the output w of a phase map
        idx = len(tape) - 1
        while idx >= 0:
            _, b, tau, w = tape[idx]
            m = (b * self.n_rec) // self.spi
            if cur_m is not None and m != cur_m:
                for k in range(self.G.shape[0]):
                    grad[k, cur_m] += np.sum(self.G[k] * gV)
                gV[:] = 0.0
            cur_m = m
            V = self.potential(c, m)
            theta = tau * (V - mu + self.g * np.abs(w) ** 2)
            gV += tau * np.imag(np.conj(lam) * w)
            xi = np.imag(np.conj(w) * lam)
            lam = np.exp(1j * theta) * (lam - 2.0 * tau * self.g * xi * w)
            # the entry before a full-step phase is a kinetic step (block-open phases have none)
            if tau == self.ddt:
                lam = gr.ifft(np.conj(self.kinop) * gr.fft(lam))
            idx -= 1
        for k in range(self.G.shape[0]):
            grad[k, cur_m] += np.sum(self.G[k] * gV)
        return L, grad, lam


class SpinorRabiControl:
    """Gradient of the final population imbalance of GPEtaskSpinor.solve_split
    (lib/@GPEtaskSpinor/solve_split.m:1-121, gamma = 0) with respect to (Omega0, k)
    of the coupling Omega(x) = Omega0 exp(i k X).

    Loss L = (N1 - N2) / Ntotal at the end, N_j = integrate(|psi_j|^2).
    The step is w = D(theta(u)) R u per point with R = [[c, -i s e^{ikX}],
    [-i s e^{-ikX}, c]], c = cos(h Omega0), s = sin(h Omega0) (solve_split.m:61-65,90-91
    with dt = i h), theta_j = h (V_j - mu + sum_k g_jk |u_k|^2) (78-83), then the
    kinetic steps (92-94, 73-75, 97-99).  Adjoint (lambda = 2 dL/d conj(psi)):
        mu_j = e^{i theta_j} lambda_j,   q_j = Im(conj(lambda_j) w_j),
        a_k = h sum_j g_jk q_j,          lambda_u = R^H mu + 2 a u,
        dL/dp += Re sum_x mu^H (dR/dp) u   for p = Omega0, k.
    The chemical potential only multiplies both components by a common phase,
    which the loss ignores, so it is held at the values of the forward run
    (real part, as the device computes it).
    """

    def __init__(self, grid, V1, V2, g, ddt, nsteps, psi0, Ntotal, n_rec=10):
        self.grid = grid
        self.V = [V1, V2]
        self.g = np.asarray(g, dtype=float)
        self.h = ddt
        self.nsteps = nsteps
        self.n_rec = n_rec
        self.psi0 = [psi0[0].copy(), psi0[1].copy()]
        self.Ntotal = Ntotal
        kk = grid.kk
        self.ekk = np.exp(-kk * 1j * ddt)
        self.ekkp = np.exp(-kk * 1j * ddt * 0.5)
        self.ekkm = np.exp(kk * 1j * ddt * 0.5)
        self.X = grid.mesh.x

    def _kin(self, p, op):
        return self.grid.ifft(op * self.grid.fft(p))

    def _mu(self, phi, Om):
        gr = self.grid
        tot = 0.0
        for j in range(2):
            res = self.V[j] + self.g[j, 0] * np.abs(phi[0]) ** 2 + self.g[j, 1] * np.abs(phi[1]) ** 2
            res = res * phi[j] + gr.lap(phi[j])
            res = res + (Om * phi[1] if j == 0 else np.conj(Om) * phi[0])
            tot = tot + gr.inner(phi[j], res)
        return np.real(tot)

    def forward(self, Om0, k, keep=False):
        h = self.h
        Om = Om0 * np.exp(1j * k * self.X)
        c, s = np.cos(h * Om0), np.sin(h * Om0)
        sp = 1j * s * np.exp(1j * k * self.X)
        sm = 1j * s * np.exp(-1j * k * self.X)
        phi = [self.psi0[0].copy(), self.psi0[1].copy()]
        mu = self._mu(phi, Om) / self.Ntotal
        tape = []
        for b in range(self.nsteps // self.n_rec):
            phi = [self._kin(p, self.ekkp) for p in phi]
            for i in range(self.n_rec):
                u = phi
                n1, n2 = np.abs(u[0]) ** 2, np.abs(u[1]) ** 2
                th1 = h * (self.V[0] - mu + self.g[0, 0] * n1 + self.g[0, 1] * n2)
                th2 = h * (self.V[1] - mu + self.g[1, 0] * n1 + self.g[1, 1] * n2)
                w1 = np.exp(-1j * th1) * (c * u[0] - sp * u[1])
                w2 = np.exp(-1j * th2) * (c * u[1] - sm * u[0])
                if keep:
                    tape.append((u, (w1, w2), (th1, th2)))
                phi = [self._kin(w1, self.ekk), self._kin(w2, self.ekk)]
            phi = [self._kin(p, self.ekkm) for p in phi]
            mu = self._mu(phi, Om) / self.Ntotal
        return (phi, tape) if keep else phi

    def loss(self, Om0, k):
        phi = self.forward(Om0, k)
        gr = self.grid
        return (gr.integrate(np.abs(phi[0]) ** 2) - gr.integrate(np.abs(phi[1]) ** 2)) / self.Ntotal

    def loss_and_grad(self, Om0, k):
        gr = self.grid
        h = self.h
        phi, tape = self.forward(Om0, k, keep=True)
        L = (gr.integrate(np.abs(phi[0]) ** 2) - gr.integrate(np.abs(phi[1]) ** 2)) / self.Ntotal
        f = 2.0 * gr.weight / self.Ntotal
        lam = [f * phi[0], -f * phi[1]]
        c, s = np.cos(h * Om0), np.sin(h * Om0)
        e = np.exp(1j * k * self.X)
        sp, sm = 1j * s * e, 1j * s * np.conj(e)
        dR0 = ((-h * s, -1j * h * c * e), (-1j * h * c * np.conj(e), -h * s))
        dRk = ((0.0, s * self.X * e), (-s * self.X * np.conj(e), 0.0))
        gO = gk = 0.0
        idx = len(tape) - 1
        for b in reversed(range(self.nsteps // self.n_rec)):
            lam = [self._kin(l, np.conj(self.ekkm)) for l in lam]
            for i in reversed(range(self.n_rec)):
                u, w, th = tape[idx]
                idx -= 1
                lam = [self._kin(l, np.conj(self.ekk)) for l in lam]
                m1, m2 = np.exp(1j * th[0]) * lam[0], np.exp(1j * th[1]) * lam[1]
                q1 = np.imag(np.conj(lam[0]) * w[0])
                q2 = np.imag(np.conj(lam[1]) * w[1])
                a1 = h * (self.g[0, 0] * q1 + self.g[1, 0] * q2)
                a2 = h * (self.g[0, 1] * q1 + self.g[1, 1] * q2)
                for dR, which in ((dR0, 0), (dRk, 1)):
                    t1 = dR[0][0] * u[0] + dR[0][1] * u[1]
                    t2 = dR[1][0] * u[0] + dR[1][1] * u[1]
                    val = np.sum(np.real(np.conj(m1) * t1 + np.conj(m2) * t2))
                    if which == 0:
                        gO += val
                    else:
                        gk += val
                l1 = np.conj(c) * m1 + np.conj(-sm) * m2 + 2 * a1 * u[0]
                l2 = np.conj(-sp) * m1 + np.conj(c) * m2 + 2 * a2 * u[1]
                lam = [l1, l2]
            lam = [self._kin(l, np.conj(self.ekkp)) for l in lam]
        return L, np.array([gO, gk])
