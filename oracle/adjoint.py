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
