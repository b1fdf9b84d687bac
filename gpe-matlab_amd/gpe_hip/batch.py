"""Ensembles of independent 2D GPEtask.solve_split runs advanced together.

The reference runs one task at a time; ``GPEEnsemble`` takes a list of 2D
``GPEtask`` objects (grid sizes, potentials, g, gamma may differ; n_recalc is
common) and runs ``solve_split`` on all of them through one gpe_batch: every
sweep is one launch per distinct pencil length for the whole ensemble, and with
``use_graph`` a whole outer iteration is replayed from a captured hipGraph.
Each task ends with the same state, history and current_* values that
``task.solve_split`` would give it.
"""
import numpy as np

from . import _lib


class GPEEnsemble:
    def __init__(self, tasks, use_graph=True, tune=True):
        if not tasks:
            raise ValueError("empty ensemble")
        self.tasks = list(tasks)
        self.use_graph = use_graph
        self._b = _lib.Batch([t.grid for t in self.tasks])
        # FFT plans of the ensemble's lengths, measured on its own sweeps where the
        # process's plan wisdom holds none (the counterpart of GPEtask.m:55-58)
        self.tune_times = self._b.tune() if tune else {}

    def solve_split(self, ddt, niter_inner, niter_outer):
        """solve_split.m:1-84 for every task; returns the list of final states."""
        starts = {t.current_iter for t in self.tasks}
        if len(starts) != 1:
            raise ValueError("all tasks must be at the same current_iter")
        start = starts.pop()
        b = self._b
        for i, t in enumerate(self.tasks):
            t._check_omega()
            if not t.Ntotal > 0:
                raise ValueError("ensemble tasks need Ntotal > 0")
            if callable(t.Vtd) or np.size(t.V0) > 1:
                raise ValueError("ensemble tasks need a static potential")
            p = _lib.make_params(1, t.g, t.gamma, t.decay_rate, t.Ntotal, t.mu_init, t.n_recalc)
            phi = t.current_state if start > 0 else t.init_state
            b.set(i, p, t.Vtrap, phi)
        b.split_begin(ddt, niter_inner, self.use_graph)
        for j in range(start + 1, niter_outer + 1):
            time2, muc, ncur = b.split_outer(j)
            for i, t in enumerate(self.tasks):
                t._ens = (self, i)
                t.current_time = time2
                t.current_iter = j
                t.current_mu = muc[i]
                t._hist_set("mu", j, muc[i])
                t.current_n = ncur[i]
                t._hist_set("n", j, ncur[i])
        return [self.state(i) for i in range(len(self.tasks))]

    def state(self, i):
        psi = self._b.get_state(i)
        self.tasks[i].current_state = psi
        return psi
