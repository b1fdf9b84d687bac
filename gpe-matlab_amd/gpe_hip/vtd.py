"""Device-evaluated time-dependent potentials.

GPEtask.Vtd is a function handle ``Vtd(task, time)`` added to Vtrap once per block
(GPEtask.m:62-71, solve_split.m:53-54).  ``BasisVtd`` is such a function for the
common case of fixed fields with time-dependent amplitudes,

    Vtd(task, t) = c_0(t) G_0 + c_1(t) G_1 + ...,

and lets solve_split rebuild the potential on the GPU from a resident basis
(gpe_vtd_set / gpe_vtd_coeffs) instead of evaluating it on the host and copying the
grid to the device every block.  Called directly it evaluates on the host with the
same summation order, so it also works wherever a plain Vtd does.
"""
import numpy as np


class BasisVtd:
    MAXBASIS = 16

    def __init__(self, basis, coeffs):
        """basis: up to 16 real arrays of the grid's shape; coeffs(task, time) -> sequence
        of as many real amplitudes."""
        self.basis = [np.asarray(b, dtype=np.float64) for b in basis]
        if not 1 <= len(self.basis) <= self.MAXBASIS:
            raise ValueError(f"BasisVtd takes 1..{self.MAXBASIS} fields")
        self.coeffs = coeffs

    def amplitudes(self, task, time):
        c = np.asarray(self.coeffs(task, time), dtype=np.float64).ravel()
        if c.size != len(self.basis):
            raise ValueError(f"coeffs returned {c.size} values for {len(self.basis)} fields")
        return c

    def __call__(self, task, time):
        c = self.amplitudes(task, time)
        acc = c[0] * self.basis[0]
        for k in range(1, len(self.basis)):
            acc = acc + c[k] * self.basis[k]
        return acc
