"""Oracle solvers: restatement of lib/@GPEtask, lib/@GPEtaskSpinor and lib/@GPEtask2comp.

Test infrastructure only (see ``oracle/__init__.py``).  The control flow, the order
of the pointwise operations and the block structure follow the MATLAB files line by
line; the cited line numbers are those of the reference files.

Deviations, all forced by the environment and stated here:
  * ``rng('shuffle')`` (GPEtask.m:49) is replaced by a seeded numpy PCG64 generator
    (``task.seed``), so random initial guesses are reproducible.
  * ``dispstat``/``toc``/``imagesc`` console and plot output is not reproduced.
  * The rotating frame is restated for a scalar ``omega`` on GPEtask (set_kinop.m:5-7,
    ssft_kin_step.m:22-25, lap with omega); an array ``omega`` (the Crank-Nicolson
    Lz branch) and omega != 0 on the two-component classes raise.
"""
import numpy as np


def _matlab_eps():
    return np.finfo(np.float64).eps


class GPEtask:
    """GPEtask.m:1-167 (handle class; Python objects are references as well)."""

    def __init__(self, grid, trappot, seed=0):
        # GPEtask.m:41-60
        self.grid = grid
        self.g = 0.0
        self.omega = 0.0
        self.n_crank = 3
        self.n_recalc = 10
        self.gamma = 0.0
        self.decay_rate = 0.0
        self.Ntotal = 0.0
        self.mu_init = 0.0
        self.Vtd = None
        self.V0 = 0
        self.current_state = None
        self.current_time = 0.0
        self.current_iter = 0
        self.current_mu = 0.0
        self.current_n = None
        self.user_callback = None
        self.itp_max_iter = 5000
        self.itp_min_iter = 50
        self.itp_adjust_stepsize = False
        self.show_image = 0
        self.init_state = np.zeros(grid.size, dtype=np.complex128)   # GPEtask.m:43
        if callable(trappot):                                         # GPEtask.m:44-48
            self.Vtrap = trappot(grid.mesh.x, grid.mesh.y, grid.mesh.z)
        else:
            self.Vtrap = trappot
        self.rng = np.random.Generator(np.random.PCG64(seed))          # replaces GPEtask.m:49-52
        self.history = {"mu": np.zeros(0), "n": np.zeros(0)}           # GPEtask.m:54
        self.kinop = None
        self.nlinop = None

    # ------------------------------------------------------------------ helpers
    def _check_omega(self):
        if not np.isscalar(self.omega) or self.omega != 0:
            raise NotImplementedError("rotating frame (omega != 0) is outside the restated path")

    def _check_scalar_omega(self):
        if not np.isscalar(self.omega):
            raise NotImplementedError("an array omega (Crank-Nicolson Lz step) is outside the restated path")

    def getVtotal(self, time):
        # GPEtask.m:62-71
        if time == 0 and np.size(self.V0) > 1:
            return self.V0
        if callable(self.Vtd):
            return self.Vtrap + self.Vtd(self, time)
        return self.Vtrap

    def applyham(self, phi, time=None):
        # GPEtask.m:73-78
        if time is None:
            time = self.current_time
        return self.applyh0(phi, time) + self.g * np.abs(phi) ** 2 * phi

    def applyh0(self, phi=None, time=None):
        # GPEtask.m:80-95 (scalar omega: V.*phi + lap(phi, omega))
        self._check_scalar_omega()
        if phi is None:
            phi = self.current_state
        if time is None:
            time = self.current_time
        res = self.getVtotal(time) * phi
        if self.omega != 0:
            res = res + self.grid.lap(phi, self.omega)
        else:
            res = res + self.grid.lap(phi)
        return res

    def get_energy(self, phi=None, time=None):
        # GPEtask.m:97-107 (g halved for the interaction energy)
        if phi is None:
            phi = self.current_state
        if time is None:
            time = self.current_time
        tmp = self.g
        self.g = 0.5 * self.g
        try:
            res = np.real(self.grid.inner(phi, self.applyham(phi, time)))
        finally:
            self.g = tmp
        return res

    def set_kinop(self, dt):
        # set_kinop.m:5-10
        self._check_scalar_omega()
        g = self.grid
        if self.omega != 0:
            self.kinop = np.exp(-(g.kx ** 2 / 2 - g.kx * g.mesh.y * self.omega) * dt / 2)      # 6
            self.kinop2 = np.exp(-(g.ky ** 2 / 2 + g.kz ** 2 / 2 + g.ky * g.mesh.x * self.omega) * dt)   # 7
        else:
            self.kinop = np.exp(-g.kk * dt)                                                  # 9

    def ssft_kin_step(self, phi, dt):
        # ssft_kin_step.m:22-27
        self._check_scalar_omega()
        g = self.grid
        if self.omega != 0:
            phi = g.ifftx(self.kinop * g.fftx(phi))                                          # 23
            phi = g.iffty(g.ifftz(self.kinop2 * g.fftz(g.ffty(phi))))                        # 24
            return g.ifftx(self.kinop * g.fftx(phi))                                         # 25
        return g.ifft(self.kinop * g.fft(phi))                                               # 27

    @staticmethod
    def _hist_set(arr, step, val):
        # MATLAB `history.mu(step) = val` grows the row with zeros (GPEtask.m:123,125)
        if step > arr.size:
            arr = np.concatenate([arr, np.zeros(step - arr.size, dtype=np.result_type(arr, val))])
        elif np.iscomplexobj(val) and not np.iscomplexobj(arr):
            arr = arr.astype(np.complex128)
        arr[step - 1] = val
        return arr

    def ext_callback(self, phi, step, time, mu, n):
        # GPEtask.m:115-142 (no snapshot dir, plot or console line)
        self.current_state = phi
        self.current_time = time
        self.current_iter = step
        self.current_mu = mu
        self.history["mu"] = self._hist_set(self.history["mu"], step, mu)
        self.current_n = n
        self.history["n"] = self._hist_set(self.history["n"], step, n)
        res_text = ""
        if callable(self.user_callback):
            r = self.user_callback(self)
            res_text = r if isinstance(r, str) else ""
        return res_text

    def process_init_state(self, phi0):
        # GPEtask.m:144-164
        if isinstance(phi0, str):
            if self.Ntotal > 0:
                if phi0 == "tf":
                    res, _ = self.groundstate_tf(_matlab_eps())
                else:
                    r = self.rng.random(self.grid.size) + 1j * self.rng.random(self.grid.size)
                    res = np.sqrt(self.Ntotal) * self.grid.normalize(r)
            else:
                V = self.getVtotal(0)
                res = np.real(np.sqrt((self.mu_init - V + 0j) / self.g))
            return res
        if self.Ntotal > 0:
            return np.sqrt(self.Ntotal) * self.grid.normalize(phi0)
        return phi0

    # ------------------------------------------------------------------ solvers
    def groundstate_tf(self, eps, N=None, V=None, mul=None, mur=None):
        # groundstate_tf.m:1-37 (bisection on mu so that integrate((mu-V)_+) = N g)
        if mul is None or mur is None:
            if V is None:
                if N is None:
                    N = self.Ntotal
                V = self.getVtotal(0)
            mul = np.min(V)
            mur = np.max(V)
        g1 = np.atleast_1d(np.asarray(self.g, dtype=float)).ravel()[0]
        NC = N * g1
        mu = (mul + mur) / 2
        vvv = None
        while (mur - mul) / mu > eps:
            vvv = (mu - V) * (V < mu)
            NN = self.grid.integrate(vvv)
            if NN > NC:
                mur = mu
            else:
                mul = mu
            mu = (mul + mur) / 2
        phi = np.sqrt(vvv / g1)
        return phi, mu

    def solve_split(self, ddt, niter_inner, niter_outer):
        """solve_split.m:1-84 -- Strang split-step with per-block open/close phases."""
        grid = self.grid
        g = self.g
        if self.Ntotal > 0:                                      # 17-21
            NN0 = self.Ntotal
        else:
            NN0 = grid.integrate(np.abs(self.init_state) ** 2)
        NNN = NN0
        gam = self.gamma
        n_rec = self.n_recalc                                    # 24-27
        if niter_inner % n_rec != 0:
            n_rec = 1
        tau = self.decay_rate
        start = self.current_iter

        dt = ddt * 1j / (1 + 1j * gam)                           # 31
        self.set_kinop(dt)                                       # 32

        phi = self.current_state if start > 0 else self.init_state   # 34-38

        self.nlinop = np.real(phi * np.conj(phi))                # 40
        muc = np.real(grid.inner(phi, self.applyham(phi))) / NN0     # 41
        mu = self.mu_init if self.mu_init > 0 else muc           # 42-46
        dt_outer = ddt * niter_inner
        time2 = 0.0
        ncur = NN0
        for j in range(start + 1, niter_outer + 1):              # 49
            time = (j - 1) * dt_outer
            for jj in range(1, niter_inner // n_rec + 1):        # 51
                time2 = time + (jj - 1) * ddt * n_rec            # 53
                VV = self.getVtotal(time2) - mu                  # 54
                self.nlinop = np.exp((-VV - g * self.nlinop) * dt * 0.5)   # 55
                phi = self.nlinop * phi                          # 56
                for _ in range(n_rec):                           # 58-62
                    phi = self.ssft_kin_step(phi, dt)
                    self.nlinop = np.exp((-VV - g * phi * np.conj(phi)) * dt)
                    phi = self.nlinop * phi
                self.nlinop = np.exp((VV + g * phi * np.conj(phi)) * dt * 0.5)   # 63
                phi = self.nlinop * phi                          # 64
                self.nlinop = np.real(phi * np.conj(phi))        # 65

                ncur = grid.integrate(self.nlinop)               # 67
                if gam > 0 and self.mu_init == 0:                # 68-78
                    if tau > 0:
                        NNN = NN0 * np.exp(-time2 / tau)
                    phi = phi * np.sqrt(NNN / ncur)
                    self.nlinop = self.nlinop * (NNN / ncur)
                    muc = np.real(grid.inner(phi, self.applyham(phi, time2))) / NNN
                    mu = muc
                else:
                    muc = np.real(grid.inner(phi, self.applyham(phi, time2))) / ncur
            self.ext_callback(phi, j, time2, muc, ncur)          # 80
        return phi

    def groundstate_itp(self, dt, eps, phi0="rand", nargout=1):
        """groundstate_itp.m:1-112 -- imaginary-time split step with renormalisation.

        ``nargout`` plays MATLAB's nargout: 1 -> phi, 2 -> (phi, MU), 3 -> (+MU2),
        4 -> (+EE).  MU2/EE are only evaluated when requested, as in the reference.
        """
        grid = self.grid
        V = self.getVtotal(0)                                    # 20
        nnn = self.Ntotal if self.Ntotal > 0 else 1              # 21-25
        phi = self.process_init_state(phi0)                      # 27
        self.current_state = phi                                 # 28
        self.set_kinop(dt)                                       # 30
        MU = np.zeros(self.itp_max_iter)                         # 31-33
        MU2 = np.zeros(self.itp_max_iter)
        EE = np.zeros(self.itp_max_iter)
        i = 0
        dtmin = np.sqrt(eps)
        self.nlinop = np.exp(-dt * 0.5 * (np.abs(phi) ** 2 * self.g + V))   # 36
        while True:
            i += 1
            phi = self.nlinop * phi                              # 40
            phi = self.ssft_kin_step(phi, dt)                    # 41
            phi = self.nlinop * phi                              # 42

            self.nlinop = np.abs(phi) ** 2                       # 44
            if self.Ntotal > 0:                                  # 45-51
                mu = np.sqrt(self.Ntotal / grid.integrate(self.nlinop))
                MU[i - 1] = np.log(mu) / dt
            else:
                mu = np.exp(self.mu_init * dt)
                MU[i - 1] = grid.integrate(self.nlinop) * mu ** 2
            phi = phi * mu                                       # 52

            if i > self.itp_min_iter and i % 10 == 0:            # 54
                if nargout >= 3:
                    MU2[i - 1] = np.real(grid.inner(phi, self.applyham(phi)))
                if nargout >= 4:
                    EE[i - 1] = self.get_energy(phi)
                delta = np.max(np.abs(np.abs(phi) - np.abs(self.current_state))) / (dt * 10)   # 61
                self.current_state = phi
                if delta < eps:
                    if nargout >= 3:
                        dtmin = dt * np.sqrt(eps / abs(MU[i - 1] - MU2[i - 1] / nnn))
                    if (not self.itp_adjust_stepsize) or dt < dtmin:
                        break
                    dt = dt / 2
                    self.set_kinop(dt)

            self.nlinop = np.exp(-dt * 0.5 * (self.nlinop * self.g * mu ** 2 + V))   # 76

            if i >= self.itp_max_iter:                           # 78-81
                import warnings
                warnings.warn("Convergence not reached")
                break

        MU = MU[:i]
        out = [phi]
        if nargout >= 2:                                         # 84-91
            if self.Ntotal > 0:
                self.current_mu = MU[-1]
                self.current_n = self.Ntotal
            out.append(MU)
        if nargout >= 3:                                         # 92-100
            if self.Ntotal > 0:
                MU2 = MU2[:i] / self.Ntotal
            else:
                MU2 = MU2[:i] / MU
                self.current_mu = MU2[-1]
                self.current_n = MU[-1]
            out.append(MU2)
        if nargout >= 4:                                         # 101-108
            EE = EE[:i] / self.Ntotal if self.Ntotal > 0 else EE[:i] / MU
            out.append(EE)
        self.init_state = phi                                    # 111
        self.itp_iterations = i
        return out[0] if nargout == 1 else tuple(out)


class GPEtaskSpinor(GPEtask):
    """GPEtaskSpinor.m:1-146 -- two coherently coupled components (omega = 0)."""

    def __init__(self, grid, trappot, seed=0):
        # GPEtaskSpinor.m:12-22
        super().__init__(grid, trappot, seed)
        self.coupling = 0.0
        self.nlincpl = 0.0
        self.ncomp = 0
        if isinstance(trappot, (list, tuple)):
            self.ncomp = len(trappot)
            self.Vtrap = [t(grid.mesh.x, grid.mesh.y, grid.mesh.z) if callable(t) else t
                          for t in trappot]

    def getVtotal(self, time):
        # GPEtaskSpinor.m:24-44
        if self.Vtd is None:
            return self.Vtrap
        if time == 0 and np.size(self.V0) > 1:
            return self.V0
        if callable(self.Vtd):
            return [self.Vtrap[i] + self.Vtd(self, time) for i in range(self.ncomp)]
        if isinstance(self.Vtd, (list, tuple)):
            return [self.Vtrap[i] + self.Vtd[i](self, time) for i in range(self.ncomp)]
        return self.Vtrap

    def applyham1c(self, phi, j, time=None):
        # GPEtaskSpinor.m:55-75 (j is 0-based here)
        self._check_omega()
        if time is None:
            time = self.current_time
        res = self.getVtotal(time)[j]
        g = np.asarray(self.g)
        for k in range(self.ncomp):
            res = res + g[j, k] * np.abs(phi[k]) ** 2
        res = res * phi[j] + self.grid.lap(phi[j])
        if j == 0:
            res = res + self.coupling * phi[1]
        else:
            res = res + np.conj(self.coupling) * phi[0]
        return res

    def applyham(self, phi, time=None):
        # GPEtaskSpinor.m:77-85
        if time is None:
            time = self.current_time
        res = list(phi)
        for j in range(self.ncomp):
            res[j] = self.applyham1c(phi, j, time)
        return res

    def inner(self, left, right):
        # GPEtaskSpinor.m:87-92
        res = 0
        for j in range(self.ncomp):
            res = res + self.grid.inner(left[j], right[j])
        return res

    def get_energy(self, phi=None, time=None):
        # GPEtaskSpinor.m:100-111
        if time is None:
            time = self.current_time
        if phi is None:
            phi = self.current_state
        tmp = self.g
        self.g = 0.5 * np.asarray(self.g)
        try:
            res = np.real(self.inner(phi, self.applyham(phi, time)))
        finally:
            self.g = tmp
        return res

    def solve_split(self, ddt, niter_inner, niter_outer):
        """@GPEtaskSpinor/solve_split.m:1-121 -- K/2 [(P R) K]^n K^-1/2 per block."""
        self._check_omega()
        ncomp = np.asarray(self.g).shape[0]                      # 15
        grid = self.grid
        g = np.asarray(self.g)
        gam = self.gamma
        n_rec = self.n_recalc
        tau = self.decay_rate
        start = self.current_iter
        phi = list(self.current_state) if start > 0 else list(self.init_state)   # 24-28
        if self.Ntotal > 0:                                      # 30-38
            NN0 = self.Ntotal
        else:
            raise ValueError("GPEtaskSpinor.solve_split needs Ntotal > 0 "
                             "(solve_split.m:37 fails on a cell init_state otherwise)")
        NNN = NN0
        ncur = NN0
        dt = ddt * 1j / (1 + 1j * gam)                           # 42
        ekk = np.exp(-grid.kk * dt)                              # 43
        ekkp = np.exp(-grid.kk * dt * 0.5)                       # 44
        ekkm = np.exp(grid.kk * dt * 0.5)                        # 45
        tmp2 = [None] * ncomp

        hphi = self.applyham(phi)                                # 52
        mu = 0
        for j in range(ncomp):                                   # 53-56
            mu = mu + np.real(grid.inner(phi[j], hphi[j]))
        mu = mu / NN0                                            # 57
        dt_outer = ddt * niter_inner
        time2 = 0.0
        for j in range(start + 1, niter_outer + 1):              # 60
            coupl = self.coupling                                # 61-65
            cang = np.angle(coupl)
            cosom = np.cosh(dt * np.abs(coupl))
            sinomm = np.sinh(dt * np.abs(coupl)) * np.exp(-1j * cang)
            sinomp = np.sinh(dt * np.abs(coupl)) * np.exp(1j * cang)
            time = (j - 1) * dt_outer
            # 67: 1:niter_inner/n_rec truncates toward the floor in MATLAB's colon
            for jj in range(1, int(np.floor(niter_inner / n_rec)) + 1):
                time2 = time + (jj - 1) * ddt * n_rec            # 69
                VV = self.getVtotal(time2)                       # 70
                for i in range(ncomp):                           # 73-75
                    phi[i] = grid.ifft(ekkp * grid.fft(phi[i]))
                for _ in range(n_rec):                           # 76-96
                    for i in range(ncomp):
                        tmp2[i] = VV[i] - mu
                        for k in range(ncomp):
                            tmp2[i] = tmp2[i] + g[i, k] * np.abs(phi[k]) ** 2
                    if self.nlincpl != 0:                        # 84-89
                        tmp3 = coupl + self.nlincpl * (np.abs(phi[0]) ** 2 + np.abs(phi[1]) ** 2)
                        cosom = np.cosh(dt * tmp3)
                        sinomm = np.sinh(dt * tmp3)
                        sinomp = sinomm
                    t0 = np.exp(-tmp2[0] * dt) * (cosom * phi[0] - sinomp * phi[1])   # 90
                    t1 = np.exp(-tmp2[1] * dt) * (cosom * phi[1] - sinomm * phi[0])   # 91
                    tmp2[0], tmp2[1] = t0, t1
                    for i in range(ncomp):                       # 92-94
                        phi[i] = grid.ifft(ekk * grid.fft(tmp2[i]))
                for i in range(ncomp):                           # 97-99
                    phi[i] = grid.ifft(ekkm * grid.fft(phi[i]))
                if gam > 0:                                      # 100-115
                    if tau > 0:
                        NNN = NN0 * np.exp(-time2 / tau)
                    ncur = np.real(grid.integrate(phi[0] * np.conj(phi[0]) + phi[1] * np.conj(phi[1])))
                    for i in range(ncomp):
                        phi[i] = phi[i] * np.sqrt(NNN / ncur)
                    hphi = self.applyham(phi)
                    mu = self.inner(phi, hphi) / NNN
                else:
                    ncur = np.real(grid.integrate(phi[0] * np.conj(phi[0]) + phi[1] * np.conj(phi[1])))
                    hphi = self.applyham(phi)
                    mu = self.inner(phi, hphi) / NNN
            self.ext_callback(list(phi), j, time2, mu, ncur)     # 117
        return phi

    def _itp_init(self, phi, ncomp):
        # GPEtaskSpinor/groundstate_itp.m:26-28 (rand, seeded here)
        if phi is None:
            g = self.grid
            phi = []
            for _ in range(ncomp):
                r = self.rng.random(g.size) + 1j * self.rng.random(g.size)
                phi.append(g.normalize(r) / np.sqrt(ncomp))
        return [p.astype(np.complex128, copy=True) for p in phi]

    def groundstate_itp(self, dt, eps, phi=None, nargout=1, max_iter=500000):
        """@GPEtaskSpinor/groundstate_itp.m:1-145 (omega = 0): K/2 . (P R) . K/2 with the
        density taken before the first kinetic half step, total-norm renormalisation
        to Ntotal, and a step-size cascade dt -> dt/1.5 on the energy-like series EE."""
        self._check_omega()
        grid = self.grid
        ncomp = np.asarray(self.g).shape[0]                      # 20-21
        self.ncomp = ncomp
        VV = self.getVtotal(0)                                   # 22
        coupl = self.coupling                                    # 24
        g = np.asarray(self.g)
        phi = self._itp_init(phi, ncomp)                         # 26-28
        ekk = np.exp(-grid.kk * dt * 0.5)                        # 29
        MU, MU2, EE = [], [], []
        i = 1                                                    # 39
        iswitch = 20                                             # 40
        tmp2 = [None] * ncomp
        cang = np.angle(coupl)                                   # 43-46
        cosom = np.cosh(dt * np.abs(coupl))
        sinomm = np.sinh(dt * np.abs(coupl)) * np.exp(-1j * cang)
        sinomp = np.sinh(dt * np.abs(coupl)) * np.exp(1j * cang)
        converged = True
        while True:
            for j in range(ncomp):                               # 49-54
                tmp2[j] = VV[j]
                for k in range(ncomp):
                    tmp2[j] = tmp2[j] + g[j, k] * np.abs(phi[k]) ** 2
            phi[0] = grid.ifft(ekk * grid.fft(phi[0]))           # 67-68
            phi[1] = grid.ifft(ekk * grid.fft(phi[1]))
            t0 = np.exp(-tmp2[0] * dt) * (cosom * phi[0] - sinomp * phi[1])   # 70
            t1 = np.exp(-tmp2[1] * dt) * (cosom * phi[1] - sinomm * phi[0])   # 71
            phi[0] = grid.ifft(ekk * grid.fft(t0))               # 78-79
            phi[1] = grid.ifft(ekk * grid.fft(t1))
            ntot = np.real(grid.integrate(phi[0] * np.conj(phi[0]) + phi[1] * np.conj(phi[1])))   # 82
            mu = np.sqrt(self.Ntotal / ntot)                     # 83
            MU.append(np.log(mu) / dt)                           # 84
            for j in range(ncomp):                               # 86-88
                phi[j] = phi[j] * mu
            self.current_state = phi
            if nargout >= 3:                                     # 91-95
                MU2.append(np.real(self.inner(phi, self.applyham(phi))) / self.Ntotal)
            else:
                MU2.append(MU[-1])
            if nargout >= 4:                                     # 96-100
                EE.append(self.get_energy(phi) / self.Ntotal)
            else:
                EE.append(MU2[-1])
            if (i - iswitch) > 10 and i % 10 == 0:               # 102-123
                delta = (abs(EE[i - 1] - EE[i - 10]) / 9 + abs(EE[i - 1] - EE[i - 2])) / dt
                if delta < eps:
                    if dt < eps or dt < 5e-6:
                        break
                    dt = dt / 1.5
                    ekk = np.exp(-grid.kk * dt * 0.5)
                    cosom = np.cosh(dt * np.abs(coupl))
                    sinomm = np.sinh(dt * np.abs(coupl)) * np.exp(-1j * cang)
                    sinomp = np.sinh(dt * np.abs(coupl)) * np.exp(1j * cang)
                    iswitch = i
            if i >= max_iter:                                    # 124-127 (cap 500000)
                converged = False
                break
            i = i + 1
        self.itp_iterations = i
        self.itp_dt = dt
        self.itp_converged = converged
        out = [phi]
        if nargout >= 2:                                         # 131-142: entries 1..i-1
            out.append(np.array(MU[:i - 1]))
        if nargout >= 3:
            out.append(np.array(MU2[:i - 1]))
        if nargout >= 4:
            out.append(np.array(EE[:i - 1]))
        self.init_state = phi                                    # 143-144
        self.current_state = phi
        return out[0] if nargout == 1 else tuple(out)


class GPEtask2comp(GPEtaskSpinor):
    """lib/@GPEtask2comp (GPEtask2comp.m:1-154): incoherent two-component mixture,
    per-component particle numbers N1, N2 and masses M1, M2 (omega = 0).

    Its solve_split.m is the same file as GPEtaskSpinor's, which reads
    task.coupling and task.nlincpl -- properties GPEtask2comp does not define, so
    the reference raises there.  Here both are fixed at 0 (the evident intent: no
    inter-component coupling), so solve_split runs the uncoupled step."""

    def __init__(self, grid, trappot, seed=0):
        super().__init__(grid, trappot, seed)                    # GPEtask2comp.m:16-26
        self.ncomp = 2
        self.N1 = None
        self.N2 = None
        self.mu1 = None
        self.mu2 = None
        self.M1 = 1.0
        self.M2 = 1.0

    def applyham1c(self, phi, j, time=None):
        # GPEtask2comp.m:51-64: no coupling term
        self._check_omega()
        if time is None:
            time = self.current_time
        res = self.getVtotal(time)[j]
        g = np.asarray(self.g)
        for k in range(self.ncomp):
            res = res + g[j, k] * np.abs(phi[k]) ** 2
        return res * phi[j] + self.grid.lap(phi[j])

    def get_mu(self, phi=None, time=None):
        # GPEtask2comp.m:76-86
        if phi is None:
            phi = self.current_state
        if time is None:
            time = self.current_time
        return [np.real(self.grid.inner(phi[0], self.applyham1c(phi, 0, time)) / self.N1),
                np.real(self.grid.inner(phi[1], self.applyham1c(phi, 1, time)) / self.N2)]

    def groundstate_itp(self, dt, eps, phi=None, nargout=1, max_iter=500000):
        """@GPEtask2comp/groundstate_itp.m:1-157 (omega = 0): K_j/2 . P . K_j/2 with
        kinetic factors exp(-kk dt/2/M_j), densities taken before the first kinetic
        half step, each component renormalised to N1 / N2; MU2 holds component 2's
        log(mu2)/dt and is the convergence series unless nargout >= 4."""
        self._check_omega()
        grid = self.grid
        VV = self.getVtotal(0)                                   # 22
        g = np.asarray(self.g)
        phi = self._itp_init(phi, 2)                             # 26-28
        ekk1 = np.exp(-grid.kk * dt * 0.5 / self.M1)             # 29-30
        ekk2 = np.exp(-grid.kk * dt * 0.5 / self.M2)
        MU1, MU2, EE = [], [], []
        i = 1
        iswitch = 20
        converged = True
        while True:
            t0 = VV[0] + g[0, 0] * np.abs(phi[0]) ** 2 + g[0, 1] * np.abs(phi[1]) ** 2   # 58
            t1 = VV[1] + g[1, 0] * np.abs(phi[0]) ** 2 + g[1, 1] * np.abs(phi[1]) ** 2   # 59
            phi[0] = grid.ifft(ekk1 * grid.fft(phi[0]))          # 67-68
            phi[1] = grid.ifft(ekk2 * grid.fft(phi[1]))
            phi[0] = np.exp(-t0 * dt) * phi[0]                   # 70-71
            phi[1] = np.exp(-t1 * dt) * phi[1]
            phi[0] = grid.ifft(ekk1 * grid.fft(phi[0]))          # 78-79
            phi[1] = grid.ifft(ekk2 * grid.fft(phi[1]))
            n1c = np.real(grid.integrate(phi[0] * np.conj(phi[0])))   # 82-85
            n2c = np.real(grid.integrate(phi[1] * np.conj(phi[1])))
            mu1 = np.sqrt(self.N1 / n1c)
            mu2 = np.sqrt(self.N2 / n2c)
            MU1.append(np.log(mu1) / dt)                         # 86-87
            MU2.append(np.log(mu2) / dt)
            phi[0] = phi[0] * mu1                                # 89-90
            phi[1] = phi[1] * mu2
            self.current_state = phi
            if nargout >= 4:                                     # 98-102
                EE.append(self.get_energy(phi) / self.Ntotal)
            else:
                EE.append(MU2[-1])
            if (i - iswitch) > 10 and i % 10 == 0:               # 113-135
                delta = (abs(EE[i - 1] - EE[i - 10]) / 9 + abs(EE[i - 1] - EE[i - 2])) / dt
                if delta < eps:
                    if dt < eps or dt < 5e-6:
                        break
                    dt = dt / 1.5
                    ekk1 = np.exp(-grid.kk * dt * 0.5 / self.M1)
                    ekk2 = np.exp(-grid.kk * dt * 0.5 / self.M2)
                    iswitch = i
            if i >= max_iter:                                    # 136-139
                converged = False
                break
            i = i + 1
        self.itp_iterations = i
        self.itp_dt = dt
        self.itp_converged = converged
        out = [phi]
        if nargout >= 2:                                         # 143-154
            out.append(np.array(MU1[:i - 1]))
        if nargout >= 3:
            out.append(np.array(MU2[:i - 1]))
        if nargout >= 4:
            out.append(np.array(EE[:i - 1]))
        self.init_state = phi                                    # 155-156
        self.current_state = phi
        return out[0] if nargout == 1 else tuple(out)


class SGPEtask(GPEtask):
    """lib/@SGPEtask (SGPEtask.m:1-47): the stochastic GPE (omega = 0).

    The complex Gaussian noise of solve_split.m:59 comes from the counter-based
    generator of ``oracle/philox.py`` (seed ``noise_seed``, one counter per global
    step and grid point) instead of MATLAB's ``randn`` stream; the global step of
    outer iteration j, block jj, inner step i is (j-1)*niter_inner + (jj-1)*n_rec + i-1.
    """

    def __init__(self, grid, trappot, seed=0):
        super().__init__(grid, trappot, seed)
        self.T = 0.0                 # SGPEtask.m:5
        self.snapshots = []          # SGPEtask.m:6
        self.ss_maxlength = 100      # SGPEtask.m:7
        self.ecut = 0.0              # SGPEtask.m:8
        self.noise_seed = 0

    def ext_callback(self, phi, step, time, mu, n):
        # SGPEtask.m:20-44: as GPEtask's, returning the user callback's text
        return super().ext_callback(phi, step, time, mu, n)

    def solve_split(self, ddt, niter_inner, niter_outer):
        """SGPEtask/solve_split.m:1-81."""
        from .philox import complex_normals
        self._check_omega()                                      # 61-68 (Crank-Nicolson Lz) not restated
        grid = self.grid
        g = self.g
        gam = self.gamma
        n_rec = self.n_recalc
        start = self.current_iter
        dt = ddt * 1j / (1 + 1j * gam)                           # 29
        phi = self.current_state if start > 0 else self.init_state   # 31-35
        phi = np.array(phi, dtype=np.complex128)
        tmp2 = np.real(phi * np.conj(phi))                       # 37
        if self.ecut > 0 and self.T > 0:                         # 39-44
            mask = grid.kk < self.ecut
            ekk = np.exp(-grid.kk * dt) * mask
        else:
            ekk = np.exp(-grid.kk * dt)
        dt_outer = ddt * niter_inner                             # 45
        npts = phi.size
        variance = np.sqrt(self.T * gam * ddt / grid.weight)     # 47
        time2 = 0.0
        ncur = None
        for j in range(start + 1, niter_outer + 1):              # 49
            mu = self.mu_init                                    # 50
            time = (j - 1) * dt_outer
            for jj in range(1, int(np.floor(niter_inner / n_rec)) + 1):   # 52
                time2 = time + (jj - 1) * ddt * n_rec
                VV = self.getVtotal(time2)                       # 54
                phi = np.exp((mu - VV - g * tmp2) * dt / 2) * phi   # 55
                for i in range(1, n_rec + 1):                    # 57-70
                    step = (j - 1) * niter_inner + (jj - 1) * n_rec + (i - 1)
                    re, im = complex_normals(npts, step, self.noise_seed)
                    # 59: randn(sz) + 1i*randn(sz), column-major point order
                    noise = (re + 1j * im).reshape(phi.shape, order="F")
                    phi = phi + variance * noise
                    phi = grid.ifft(ekk * grid.fft(phi))         # 60
                    phi = np.exp((mu - VV - g * phi * np.conj(phi)) * dt) * phi   # 69
                phi = np.exp((VV - mu + g * phi * np.conj(phi)) * dt / 2) * phi   # 72
                tmp2 = np.real(phi * np.conj(phi))               # 73
                ncur = grid.integrate(tmp2)                      # 75
            err = self.ext_callback(phi, j, time2, mu, ncur)     # 77-80
            if err == "TERM":
                break
        return phi
