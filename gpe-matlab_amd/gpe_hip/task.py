"""GPEtask / GPEtaskSpinor with the reference's API, running on the HIP library.

``task.solve_split(ddt, niter_inner, niter_outer)`` and
``task.groundstate_itp(dt, eps, phi0)`` keep the reference signatures, property
names, side effects (current_state/current_iter/current_time/history, user
callback) and resume semantics; their bodies call the C ABI (include/gpe_hip.h),
which runs every step on the GPU.  There is no CPU fallback: without the built
library every solver call raises ``GPELibraryError``.

Reference: lib/@GPEtask/GPEtask.m, solve_split.m, groundstate_itp.m,
groundstate_tf.m; lib/@GPEtaskSpinor/GPEtaskSpinor.m, solve_split.m.
"""
import warnings

import numpy as np

from . import _lib
from .vtd import BasisVtd

_UNSET = object()


class GPEtask:
    """Drop-in for lib/@GPEtask (GPEtask.m:1-167), omega == 0."""

    _ncomp = 1

    def __init__(self, grid, trappot, seed=None):
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
        self.Vcurrent = 0
        self.init_state = np.zeros(grid.size, dtype=np.complex128)
        self.current_time = 0.0
        self.current_iter = 0
        self.current_mu = 0.0
        self.current_n = None
        self.history = {"mu": np.zeros(0), "n": np.zeros(0)}
        self.user_callback = None
        self.itp_max_iter = 5000
        self.itp_min_iter = 50
        self.itp_adjust_stepsize = False
        self.show_image = 0
        # GPEtask.m:28, 55-58: measure the FFT plans of this grid when the task first
        # reaches the device (plans already in the process's wisdom are reused)
        self.fftw_set_wisdom = True
        if callable(trappot):
            self.Vtrap = trappot(grid.mesh.x, grid.mesh.y, grid.mesh.z)
        else:
            self.Vtrap = trappot
        # GPEtask.m:49 seeds from the clock (rng('shuffle')); a seed makes it reproducible
        self.rng = np.random.Generator(np.random.PCG64(seed))
        self._dev = None
        self._host_state = None      # host copy of current_state
        self._dev_state_valid = False  # device holds current_state (maybe newer than host copy)
        self._host_stale = False
        self._in_callback = False    # inside user_callback: the solver's state, V and params are resident

    # ------------------------------------------------------------------ state
    def _device(self):
        if self._dev is None:
            self._dev = _lib.Device(self.grid)
            if self.fftw_set_wisdom:
                self._dev.tune()
        return self._dev

    @property
    def current_state(self):
        if self._host_stale:
            self._host_state = self._fetch()
            self._host_stale = False
        return self._host_state

    @current_state.setter
    def current_state(self, value):
        self._host_state = value
        self._host_stale = False
        self._dev_state_valid = False

    def _fetch(self):
        return self._device().get_state(0)

    def _upload(self, phi):
        self._device().set_state(0, phi)

    def _split_result(self, start, niter_outer):
        # solve_split.m:34-38, 84: phi starts as current_state (resumed) or init_state
        # and is returned as it stands when the outer loop runs no iteration
        if start == 0 and niter_outer <= start:
            return self.init_state
        return self.current_state

    def getVtotal(self, time):
        # GPEtask.m:62-71
        if time == 0 and np.size(self.V0) > 1:
            return self.V0
        if callable(self.Vtd):
            return self.Vtrap + self.Vtd(self, time)
        return self.Vtrap

    def _check_omega(self):
        # a scalar omega selects the rotating-frame kinetic step (set_kinop.m:5-7); an
        # array omega (the Crank-Nicolson Lz branch, ssft_kin_step.m:13-21) is not accelerated
        if not np.isscalar(self.omega):
            raise NotImplementedError("an array omega (Crank-Nicolson Lz step) is not part of the accelerated path")

    def _set_params(self, Ntotal_eff):
        dev = self._device()
        dev.set_params(self._ncomp, self.g, self.gamma, self.decay_rate, Ntotal_eff, self.mu_init, self.n_recalc)
        dev.set_omega(float(self.omega))

    def _hist_set(self, key, step, val):
        arr = self.history[key]
        if step > arr.size:
            arr = np.concatenate([arr, np.zeros(step - arr.size, dtype=np.result_type(arr, val))])
        elif np.iscomplexobj(val) and not np.iscomplexobj(arr):
            arr = arr.astype(np.complex128)
        arr[step - 1] = val
        self.history[key] = arr

    def ext_callback(self, step, time, mu, n):
        # GPEtask.m:115-142: the state stays on the device and is fetched on first access
        self._host_stale = True
        self._dev_state_valid = True
        self.current_time = time
        self.current_iter = step
        self.current_mu = mu
        self._hist_set("mu", step, mu)
        self.current_n = n
        self._hist_set("n", step, n)
        res_text = ""
        if callable(self.user_callback):
            self._in_callback = True
            try:
                r = self.user_callback(self)
            finally:
                self._in_callback = False
            res_text = r if isinstance(r, str) else ""
        return res_text

    def _resident(self, phi, time):
        """True inside user_callback for the current state at current_time: the device then
        holds that state, the potential of current_time (the last block's, uploaded or built
        at its start time2 = current_time) and the run's parameters, so reductions need no
        upload and leave the run untouched."""
        return phi is None and self._in_callback and (time is None or time == self.current_time)

    # ------------------------------------------------------------------ energies
    def get_energy(self, phi=None, time=None):
        """GPEtask.m:97-107 -- real <phi|H_{g/2} phi>, reduced on the device."""
        m = self._measure(phi, time)
        return (m[1] + m[3]) + 0.5 * self.g * m[2]

    def chemical_potential_integral(self, phi=None, time=None):
        """real(grid.inner(phi, applyham(phi))) (GPEtask.m:73-78 + grid1d/inner.m:12)."""
        m = self._measure(phi, time)
        return (m[1] + m[3]) + self.g * m[2]

    def applyham(self, phi=None, time=None):
        """GPEtask.m:73-78 -- H phi = applyh0(phi) + g |phi|^2 phi as an array, on the device
        (phi = None: current_state)."""
        return self._apply(None if phi is None else [phi], time, False)[0]

    def applyh0(self, phi=None, time=None):
        """GPEtask.m:80-95 -- getVtotal(time) phi + lap(phi[, omega]) as an array, on the device."""
        return self._apply(None if phi is None else [phi], time, True)[0]

    def _potential_at(self, time):
        self._device().set_potential(0, self.getVtotal(time))

    def _apply(self, phis, time, h0):
        """The Hamiltonian of the arrays `phis` (None: the current state) with the potential
        of `time` (default current_time).  Inside user_callback the running solve_split's
        state, potential and parameters are left as they were, as for the reductions."""
        self._check_omega()
        dev = self._device()
        if self._resident(phis, time):
            return dev.applyham(None, self._ncomp, h0)
        if not self._in_callback:
            self._set_params(self.Ntotal if self.Ntotal > 0 else 1.0)
        if time is None:
            time = self.current_time
        self._potential_at(time)
        if phis is None and not self._dev_state_valid:
            if self.current_state is None:
                raise ValueError("no current state")
            self._upload(self.current_state)
            self._dev_state_valid = True
        res = dev.applyham(phis, self._ncomp, h0)
        if self._in_callback and time != self.current_time:
            self._potential_at(self.current_time)   # the running block's potential
        return res

    # ------------------------------------------------------------------ diagnostics
    def detect_core(self, refine=0, zslice=None, comp=0):
        """lib/util/detect_core.m:1-66 on the resident state, without copying it to the host:
        detect_core(phi(:,:,zslice+1), mesh.x(:,:,zslice+1), mesh.y(:,:,zslice+1), refine)
        with phi = current_state (component ``comp`` of a two-component task).  zslice is
        0-based and required on a 3D grid.  Returns (resp, resm), (n, 2) arrays of [x, y]."""
        dev = self._device()
        mesh = self.grid.mesh
        if self.grid.ndims == 3:
            if zslice is None:
                raise ValueError("detect_core on a 3D grid needs a z slice")
            x, y = mesh.x[0, :, 0], mesh.y[:, 0, 0]
        else:
            if zslice not in (None, 0):
                raise ValueError("a 2D grid has only z slice 0")
            zslice = 0
            x, y = mesh.x[0, :], mesh.y[:, 0]
        if not self._dev_state_valid:
            if self.current_state is None:
                raise ValueError("no current state")
            self._upload(self.current_state)
            self._dev_state_valid = True
        return dev.detect_core(x, y, refine, comp, zslice)

    def _measure(self, phi, time):
        self._check_omega()
        dev = self._device()
        if self._resident(phi, time):
            return dev.measure()
        if not self._in_callback:                # inside a callback the run's parameters stay
            self._set_params(self.Ntotal if self.Ntotal > 0 else 1.0)
        if time is None:
            time = self.current_time
        dev.set_potential(0, self.getVtotal(time))
        if phi is None:
            if not self._dev_state_valid:
                self._upload(self.current_state)
                self._dev_state_valid = True
            m = dev.measure()
        else:
            saved = self._fetch() if self._dev_state_valid else None
            self._upload(phi)
            m = dev.measure()
            if saved is not None:
                self._upload(saved)
        if self._in_callback and time != self.current_time:
            dev.set_potential(0, self.getVtotal(self.current_time))   # the running block's potential
        return m

    # ------------------------------------------------------------------ solvers
    def groundstate_tf(self, eps, N=None, V=None, mul=None, mur=None):
        """groundstate_tf.m:1-37 -- one-off host bisection for the Thomas-Fermi profile."""
        if mul is None or mur is None:
            if V is None:
                if N is None:
                    N = self.Ntotal
                V = self.getVtotal(0)
            mul = np.min(V)
            mur = np.max(V)
        g1 = float(np.ravel(self.g)[0])
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
        return np.sqrt(vvv / g1), mu

    def process_init_state(self, phi0):
        # GPEtask.m:144-164
        if isinstance(phi0, str):
            if self.Ntotal > 0:
                if phi0 == "tf":
                    res, _ = self.groundstate_tf(np.finfo(np.float64).eps)
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

    def _static_potential(self):
        return not callable(self.Vtd) and np.size(self.V0) <= 1

    def solve_split(self, ddt, niter_inner, niter_outer):
        """solve_split.m:1-84 on the GPU; returns the final state."""
        self._check_omega()
        dev = self._device()
        start = self.current_iter
        if self.Ntotal > 0:                                     # solve_split.m:17-21
            NN0 = self.Ntotal
        else:
            self._set_params(1.0)
            dev.set_potential(0, self.getVtotal(self.current_time))
            self._upload(self.init_state)
            self._dev_state_valid = False
            NN0 = dev.measure()[0]
        self._set_params(NN0)
        if start > 0:                                           # 34-38
            if not self._dev_state_valid:
                self._upload(self.current_state)
        else:
            self._upload(self.init_state)
        self._dev_state_valid = True
        Vb = self.getVtotal(self.current_time)                 # applyham(phi) at line 41
        dev.set_potential(0, Vb)
        dev.split_begin(ddt, niter_inner)
        static = self._static_potential()
        n_rec = self.n_recalc if niter_inner % self.n_recalc == 0 else 1
        if static and Vb is not self.Vtrap:
            dev.set_potential(0, self.Vtrap)
        dt_outer = ddt * niter_inner
        device_vtd = isinstance(self.Vtd, BasisVtd)
        if device_vtd:
            dev.vtd_set(self.Vtd.basis, [self.Vtrap])
        for j in range(start + 1, niter_outer + 1):             # 49
            if static:
                time2, muc, ncur = dev.split_outer(j)
            else:
                time = (j - 1) * dt_outer
                for jj in range(1, niter_inner // n_rec + 1):
                    time2 = time + (jj - 1) * ddt * n_rec
                    if device_vtd and not (time2 == 0 and np.size(self.V0) > 1):
                        dev.vtd_coeffs(self.Vtd.amplitudes(self, time2), 0)   # 54 on the device
                    else:
                        dev.set_potential(0, self.getVtotal(time2))   # 54
                    muc, ncur = dev.split_block(time2)
            self.ext_callback(j, time2, muc, ncur)              # 80
        return self._split_result(start, niter_outer)

    def groundstate_itp(self, dt, eps, phi0="rand", nargout=1):
        """groundstate_itp.m:1-112 on the GPU.  ``nargout`` plays MATLAB's nargout."""
        self._check_omega()
        dev = self._device()
        V = self.getVtotal(0)                                   # 20
        phi = self.process_init_state(phi0)                     # 27
        self.current_state = phi                                # 28
        self._set_params(self.Ntotal)
        dev.set_potential(0, V)
        self._upload(phi)
        MU, MU2, EE, i, dt_end, conv = dev.itp(dt, eps, self.itp_max_iter, self.itp_min_iter,
                                               self.itp_adjust_stepsize, nargout)
        if not conv:
            warnings.warn("Convergence not reached")           # 78-81
        self._dev_state_valid = True
        self._host_stale = True
        phi = self.current_state
        self.itp_iterations = i
        out = [phi]
        if nargout >= 2:                                        # 84-91
            if self.Ntotal > 0:
                self.current_mu = MU[-1]
                self.current_n = self.Ntotal
            out.append(MU)
        if nargout >= 3:                                        # 92-100
            if self.Ntotal > 0:
                MU2 = MU2 / self.Ntotal
            else:
                MU2 = MU2 / MU
                self.current_mu = MU2[-1]
                self.current_n = MU[-1]
            out.append(MU2)
        if nargout >= 4:                                        # 101-108
            EE = EE / self.Ntotal if self.Ntotal > 0 else EE / MU
            out.append(EE)
        self.init_state = phi                                   # 111
        return out[0] if nargout == 1 else tuple(out)


class SGPEtask(GPEtask):
    """Drop-in for lib/@SGPEtask (SGPEtask.m:1-47): the stochastic GPE at temperature T
    with an optional energy cut-off ecut (omega = 0).

    Deviation: MATLAB's randn stream is replaced by a counter-based Philox4x32-10
    stream keyed by ``noise_seed`` (one complex normal variate per grid point and
    global step), so runs are reproducible and resumable; the oracle draws the same
    variates."""

    def __init__(self, grid, trappot, seed=None):
        super().__init__(grid, trappot, seed)
        self.T = 0.0                 # SGPEtask.m:5
        self.snapshots = []          # SGPEtask.m:6
        self.ss_maxlength = 100      # SGPEtask.m:7
        self.ecut = 0.0              # SGPEtask.m:8
        self.noise_seed = 0

    def _check_omega(self):
        if not np.isscalar(self.omega) or self.omega != 0:
            raise NotImplementedError("the stochastic GPE runs without rotation (omega = 0)")

    def solve_split(self, ddt, niter_inner, niter_outer):
        """SGPEtask/solve_split.m:1-81 on the GPU; returns the final state."""
        self._check_omega()
        dev = self._device()
        start = self.current_iter
        self._set_params(self.Ntotal if self.Ntotal > 0 else 1.0)
        if start > 0:                                           # 31-35
            if not self._dev_state_valid:
                self._upload(self.current_state)
        else:
            self._upload(self.init_state)
        self._dev_state_valid = True
        static = self._static_potential()
        dev.set_potential(0, self.Vtrap if static else self.getVtotal(self.current_time))
        dev.sgpe_begin(ddt, niter_inner, self.T, self.ecut, self.noise_seed)
        n_rec = self.n_recalc
        dt_outer = ddt * niter_inner
        mu = self.mu_init                                       # 50
        for j in range(start + 1, niter_outer + 1):             # 49
            if static:
                time2, ncur = dev.sgpe_outer(j)
            else:
                time = (j - 1) * dt_outer
                for jj in range(1, niter_inner // n_rec + 1):   # 52
                    time2 = time + (jj - 1) * ddt * n_rec       # 53
                    dev.set_potential(0, self.getVtotal(time2))   # 54
                    ncur = dev.sgpe_block((j - 1) * niter_inner + (jj - 1) * n_rec)
            err = self.ext_callback(j, time2, mu, ncur)         # 77-80
            if err == "TERM":
                break
        return self._split_result(start, niter_outer)


class GPEtaskSpinor(GPEtask):
    """Drop-in for lib/@GPEtaskSpinor (GPEtaskSpinor.m:1-146): two coherently coupled
    components with a scalar or space-dependent complex Rabi coupling.

    Deviation: the chemical potential is reduced as a real number (its imaginary
    part is rounding noise: the coupled Hamiltonian is Hermitian)."""

    _ncomp = 2

    def __init__(self, grid, trappot, seed=None):
        # GPEtaskSpinor.m:12-22
        super().__init__(grid, trappot, seed)
        self.coupling = 0.0
        self.nlincpl = 0.0
        self.ncomp = 0
        if isinstance(trappot, (list, tuple)):
            self.ncomp = len(trappot)
            self.Vtrap = [t(grid.mesh.x, grid.mesh.y, grid.mesh.z) if callable(t) else t for t in trappot]

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

    def _check_omega(self):
        if not np.isscalar(self.omega) or self.omega != 0:
            raise NotImplementedError("the two-component classes run without rotation (omega = 0)")

    def _set_params(self, Ntotal_eff):
        dev = self._device()
        dev.set_params(2, np.asarray(self.g, dtype=float).reshape(2, 2), self.gamma, self.decay_rate,
                       Ntotal_eff, self.mu_init, self.n_recalc, self.nlincpl)
        dev.set_omega(0.0)

    def _fetch(self):
        d = self._device()
        return [d.get_state(0), d.get_state(1)]

    def _upload(self, phi):
        d = self._device()
        d.set_state(0, phi[0])
        d.set_state(1, phi[1])

    def _upload_potential(self, V):
        d = self._device()
        d.set_potential(0, V[0])
        d.set_potential(1, V[1])

    def _upload_coupling(self):
        self._device().set_coupling(self.coupling)

    def get_energy(self, phi=None, time=None):
        """GPEtaskSpinor.m:100-111 -- real <phi|H_{g/2} phi>, reduced on the device."""
        m = self._measure_sp(phi, time)
        return (m[1] + m[3]) + m[4] + 0.5 * m[2]

    def applyham(self, phi=None, time=None):
        """GPEtaskSpinor.m:77-86 -- [applyham1c(phi, 1), applyham1c(phi, 2)] as arrays, on the
        device (phi = None: current_state)."""
        return self._apply(None if phi is None else list(phi), time, False)

    def applyham1c(self, phi, j, time=None):
        """GPEtaskSpinor.m:55-76 -- component j (0-based here; 1-based in MATLAB) of H phi:
        (V_j + sum_k g_jk |phi_k|^2) phi_j + lap(phi_j) + Omega phi_2 (j = 0) or
        conj(Omega) phi_1 (j = 1)."""
        return self.applyham(phi, time)[j]

    def applyh0(self, phi, j, time=None):
        """GPEtaskSpinor.m:45-54 -- lap(phi) + V_j phi for one array phi and component j
        (0-based)."""
        return self._apply([phi, phi], time, True)[j]

    def _potential_at(self, time):
        self._upload_potential(self.getVtotal(time))
        self._upload_coupling()

    def _measure_sp(self, phi, time):
        self._check_omega()
        dev = self._device()
        if self._in_callback:
            dev.set_potential_mu(0, None)   # the step potential at current_time; the next block-end
            if self._resident(phi, time):   # potential is set again before it is used
                return dev.measure()
        else:
            self._set_params(self.Ntotal if self.Ntotal > 0 else 1.0)
        if time is None:
            time = self.current_time
        self._upload_potential(self.getVtotal(time))
        self._upload_coupling()
        saved = self._fetch() if (phi is not None and self._dev_state_valid) else None
        if phi is not None:
            self._upload(phi)
        elif not self._dev_state_valid:
            self._upload(self.current_state)
            self._dev_state_valid = True
        m = dev.measure()
        if saved is not None:
            self._upload(saved)
        if self._in_callback and time != self.current_time:
            self._upload_potential(self.getVtotal(self.current_time))   # the running block's potential
        return m

    def solve_split(self, ddt, niter_inner, niter_outer):
        """@GPEtaskSpinor/solve_split.m:1-121 on the GPU; returns [psi1, psi2]."""
        self._check_omega()
        dev = self._device()
        if np.asarray(self.g).shape != (2, 2):
            raise ValueError("GPEtaskSpinor needs a 2x2 g matrix")
        if not self.Ntotal > 0:
            raise ValueError("GPEtaskSpinor.solve_split needs Ntotal > 0 (solve_split.m:30-38)")
        start = self.current_iter
        self._set_params(self.Ntotal)
        if start > 0:
            if not self._dev_state_valid:
                self._upload(self.current_state)
        else:
            self._upload(self.init_state)
        self._dev_state_valid = True
        self._upload_potential(self.getVtotal(self.current_time))
        self._upload_coupling()
        dev.split_begin(ddt, niter_inner)
        static = self.Vtd is None
        n_rec = self.n_recalc
        dt_outer = ddt * niter_inner
        cpl = self.coupling
        # what was uploaded: a callback may rebind task.coupling or change an array
        # coupling in place (task.coupling *= 2), and MATLAB's value semantics make both
        # visible at the next outer iteration (61-65)
        cpl_snap = np.array(cpl, copy=True) if self.user_callback is not None else None
        device_vtd = isinstance(self.Vtd, BasisVtd)
        if device_vtd:
            dev.vtd_set(self.Vtd.basis, list(self.Vtrap))
        try:
            for j in range(start + 1, niter_outer + 1):
                if self.coupling is not cpl or (cpl_snap is not None and not np.array_equal(cpl, cpl_snap)):
                    cpl = self.coupling
                    self._upload_coupling()
                    if self.user_callback is not None:
                        cpl_snap = np.array(cpl, copy=True)
                if static:
                    time2, mu, ncur = dev.split_outer(j)
                else:
                    # the block-end mu applies applyham at task.current_time (109-114)
                    tc = self.current_time
                    if device_vtd and not (tc == 0 and np.size(self.V0) > 1):
                        dev.vtd_coeffs(self.Vtd.amplitudes(self, tc), 1)
                    else:
                        Vc = self.getVtotal(tc)
                        dev.set_potential_mu(0, Vc[0])
                        dev.set_potential_mu(1, Vc[1])
                    time = (j - 1) * dt_outer
                    for jj in range(1, niter_inner // n_rec + 1):
                        time2 = time + (jj - 1) * ddt * n_rec
                        if device_vtd and not (time2 == 0 and np.size(self.V0) > 1):
                            dev.vtd_coeffs(self.Vtd.amplitudes(self, time2))
                        else:
                            self._upload_potential(self.getVtotal(time2))
                        mu, ncur = dev.split_block(time2)
                self.ext_callback(j, time2, mu, ncur)
        finally:
            if not static:
                dev.set_potential_mu(0, None)
        return self._split_result(start, niter_outer)

    # two-component imaginary time ------------------------------------------------
    itp_max_iter_2c = 500000   # the two-component solvers stop at 500000 iterations

    def _itp_init(self, phi, ncomp):
        # GPEtaskSpinor/groundstate_itp.m:26-28 (rand, from the task's seeded generator)
        if phi is None:
            g = self.grid
            phi = []
            for _ in range(ncomp):
                r = self.rng.random(g.size) + 1j * self.rng.random(g.size)
                phi.append(g.normalize(r) / np.sqrt(ncomp))
        return [np.asarray(p, dtype=np.complex128) for p in phi]

    def _itp2_finish(self, res, nargout, names):
        MU, MU2, EE, i, dt_end, conv = res
        if not conv:
            warnings.warn("Convergence not reached")
        self._dev_state_valid = True
        self._host_stale = True
        phi = self.current_state
        self.itp_iterations = i
        self.itp_dt = dt_end
        out = [phi]
        series = {"MU": MU, "MU2": MU2, "EE": EE}
        for n in names[:max(0, nargout - 1)]:
            out.append(series[n])
        self.init_state = phi
        return out[0] if nargout == 1 else tuple(out)

    def groundstate_itp(self, dt, eps, phi0=None, nargout=1):
        """@GPEtaskSpinor/groundstate_itp.m:1-145 on the GPU (omega = 0).  Returns phi
        and, by ``nargout``, MU, MU2 (applyham-based) and EE (energy) over Ntotal."""
        self._check_omega()
        dev = self._device()
        ncomp = np.asarray(self.g).shape[0]                      # 20-21
        self.ncomp = ncomp
        VV = self.getVtotal(0)                                   # 22
        phi = self._itp_init(phi0, ncomp)                        # 26-28
        self._set_params(self.Ntotal)
        self._upload_potential(VV)
        self._upload_coupling()
        self._upload(phi)
        res = dev.itp2(dt, eps, self.itp_max_iter_2c, nargout)
        return self._itp2_finish(res, nargout, ["MU", "MU2", "EE"])


class GPEtask2comp(GPEtaskSpinor):
    """Drop-in for lib/@GPEtask2comp (GPEtask2comp.m:1-154): incoherent two-component
    mixture with particle numbers N1, N2 and masses M1, M2.

    Deviation: its solve_split.m is the same file as GPEtaskSpinor's and reads
    task.coupling / task.nlincpl, properties this class does not define, so the
    reference raises there; here both stay 0 and solve_split runs the uncoupled
    two-component step."""

    def __init__(self, grid, trappot, seed=None):
        super().__init__(grid, trappot, seed)                    # GPEtask2comp.m:16-26
        self.ncomp = 2
        self.N1 = None
        self.N2 = None
        self.mu1 = None
        self.mu2 = None
        self.M1 = 1.0
        self.M2 = 1.0

    def groundstate_itp(self, dt, eps, phi0=None, nargout=1):
        """@GPEtask2comp/groundstate_itp.m:1-157 on the GPU (omega = 0).  Returns phi
        and, by ``nargout``, MU1, MU2 (log(mu_j)/dt of each component) and EE
        (energy over Ntotal)."""
        self._check_omega()
        if self.N1 is None or self.N2 is None:
            raise ValueError("GPEtask2comp.groundstate_itp needs N1 and N2")
        dev = self._device()
        VV = self.getVtotal(0)                                   # 22
        phi = self._itp_init(phi0, 2)                            # 26-28
        self.coupling = 0.0
        self._set_params(self.Ntotal)
        self._upload_potential(VV)
        self._upload_coupling()
        self._upload(phi)
        res = dev.itp2(dt, eps, self.itp_max_iter_2c, nargout, mixture=True, N1=self.N1, N2=self.N2,
                       M1=self.M1, M2=self.M2)
        return self._itp2_finish(res, nargout, ["MU", "MU2", "EE"])
