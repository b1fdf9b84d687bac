"""Adjoint gradients of split-step dynamics with respect to control amplitudes.

The reference has no adjoint; this is the gradient entry point of the library
(gpe_adjoint_grad in include/gpe_hip.h).  The forward model is
GPEtask.solve_split with gamma = 0 and the time-dependent potential

    V(x, t) = V0(x) + sum_k c[k, m] G_k(x),   m = interval holding the block of t,

evaluated at block starts as solve_split.m:54 does.  The loss is
1 - |<psi_target|psi(T)>|^2.  ``SplitStepControl`` plays the role of a
torch.autograd.Function: ``forward`` runs the dynamics, ``loss_and_grad`` returns
the loss and dL/dc from one forward and one backward pass on the GPU, and
``autograd_function()`` wraps both as a real torch.autograd.Function.
"""
import numpy as np


class SplitStepControl:
    def __init__(self, task, V0, basis, ddt, nsteps, steps_per_interval, checkpoint_blocks=0):
        self.task = task
        self.V0 = np.asarray(V0, dtype=np.float64)
        self.basis = [np.asarray(b, dtype=np.float64) for b in basis]
        self.ddt = float(ddt)
        self.nsteps = int(nsteps)
        self.spi = int(steps_per_interval)
        self.ckpt = int(checkpoint_blocks)
        self.nintervals = (self.nsteps + self.spi - 1) // self.spi
        if task.gamma != 0:
            raise ValueError("the adjoint path needs gamma = 0")
        self._ready = False

    def _prepare(self):
        t = self.task
        dev = t._device()
        t._set_params(t.Ntotal)
        dev.control_set(self.V0, self.basis)
        dev.set_state(0, t.init_state)
        t._dev_state_valid = False
        self._ready = True
        return dev

    def _coeffs(self, c):
        c = np.asarray(c, dtype=np.float64)
        if c.shape != (len(self.basis), self.nintervals):
            raise ValueError(f"coefficients must have shape {(len(self.basis), self.nintervals)}")
        return c

    def forward(self, coeffs):
        """Final state psi(T) of the controlled dynamics from task.init_state."""
        dev = self._prepare()
        return dev.control_forward(self.ddt, self.nsteps, self.spi, self._coeffs(coeffs), self.ckpt)

    def loss_and_grad(self, coeffs, psi_target):
        """(loss, dL/dc, reconstruction error of the backward state sweep)."""
        dev = self._prepare()
        return dev.adjoint_grad(self.ddt, self.nsteps, self.spi, self._coeffs(coeffs), psi_target, self.ckpt)

    def autograd_function(self, psi_target):
        """A torch.autograd.Function computing the loss of float64 coefficients."""
        import torch

        ctl = self

        class _Loss(torch.autograd.Function):
            @staticmethod
            def forward(ctx, c):
                L, g, _ = ctl.loss_and_grad(c.detach().cpu().numpy(), psi_target)
                ctx.save_for_backward(torch.as_tensor(g, dtype=c.dtype, device=c.device))
                return torch.as_tensor(L, dtype=c.dtype, device=c.device)

            @staticmethod
            def backward(ctx, gL):
                (g,) = ctx.saved_tensors
                return gL * g

        return _Loss


class RabiControl:
    """Gradient of the final population imbalance L = (N1 - N2) / Ntotal of a
    GPEtaskSpinor.solve_split run (gamma = 0) with respect to (Omega0, k) of the
    coupling Omega(x) = Omega0 exp(i k x) (gpe_rabi_grad in include/gpe_hip.h).

    The run starts from ``task.init_state`` with ``task.g``, ``task.Vtrap`` (static
    potentials) and ``task.n_recalc``; ``nsteps`` must be a multiple of n_recalc.
    The two-component step is not explicitly invertible, so the backward pass
    reads the state before every step from a tape: segments of about sqrt(nsteps)
    steps are stored by the forward pass while ``tape_gib`` (0: half the free device
    memory, < 0: one segment) holds them, and recomputed from checkpoints otherwise.
    """

    def __init__(self, task, ddt, nsteps, checkpoint_blocks=0, tape_gib=0.0):
        if task.gamma != 0:
            raise ValueError("the Rabi adjoint needs gamma = 0")
        if not task.Ntotal > 0:
            raise ValueError("the Rabi adjoint needs Ntotal > 0")
        if task.Vtd is not None:
            raise ValueError("the Rabi adjoint supports static potentials only")
        self.task = task
        self.ddt = float(ddt)
        self.nsteps = int(nsteps)
        self.ckpt = int(checkpoint_blocks)
        self.tape_gib = float(tape_gib)

    def _prepare(self):
        t = self.task
        dev = t._device()
        t._set_params(t.Ntotal)
        t._upload_potential(t.getVtotal(0))
        psi0 = t.init_state
        t._upload(psi0)
        t._dev_state_valid = False
        return dev

    def forward(self, omega0, k):
        """(L, [psi1(T), psi2(T)])"""
        return self._prepare().rabi_forward(self.ddt, self.nsteps, omega0, k, self.ckpt)

    def loss(self, omega0, k):
        return self._prepare().rabi_forward(self.ddt, self.nsteps, omega0, k, self.ckpt, want_state=False)[0]

    def loss_and_grad(self, omega0, k):
        """(L, array([dL/dOmega0, dL/dk]))"""
        return self._prepare().rabi_grad(self.ddt, self.nsteps, omega0, k, self.ckpt, self.tape_gib)

    def autograd_function(self):
        """A torch.autograd.Function mapping a float64 tensor (Omega0, k) to L."""
        import torch

        ctl = self

        class _Loss(torch.autograd.Function):
            @staticmethod
            def forward(ctx, p):
                v = p.detach().cpu().numpy()
                L, g = ctl.loss_and_grad(float(v[0]), float(v[1]))
                ctx.save_for_backward(torch.as_tensor(g, dtype=p.dtype, device=p.device))
                return torch.as_tensor(L, dtype=p.dtype, device=p.device)

            @staticmethod
            def backward(ctx, gL):
                (g,) = ctx.saved_tensors
                return gL * g

        return _Loss
