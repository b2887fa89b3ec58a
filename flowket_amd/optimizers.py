"""Parameter updates of the VMC loop.

``ComplexValuesStochasticReconfiguration`` follows
optimizers/stochastic_reconfiguration/optimizer.py:14-183 (same constructor
arguments, same CG stopping rule linear_equations.py:87-131, same metric
names).  ConvNetAutoregressive2D has real parameters and a complex log psi,
which the reference optimizer does not accept (complex_values_optimizer.py:
12-25); the real-parameter form used here is
    S = Re[(O-Obar)^H (O-Obar)]/B + shift*I,   F = Re[(O-Obar)^H (E_loc-E)]/B,
    theta <- theta - lr * S^{-1} F,
with S applied implicitly (fk_o_times_v / fk_oh_times_u, FP32 products,
FP64 reductions) and never formed.  The energy gradient of the Adam path,
d<E>/d theta = 2 Re[sum_b O_b conj(E_loc,b - E)]/B (loss 2 Re(logpsi y),
optimization/loss.py:4-5), uses the same kernels.
"""
import math

import torch

from . import _lib


def _complex_eloc_pairs(local_energies, energy):
    e = (local_energies - energy).to(torch.complex128)
    return torch.view_as_real(e).to(torch.float32).contiguous()


class _OFactor:
    """O (per-sample gradients) of one batch with its column means."""

    def __init__(self, machine, configurations, stream=None):
        self.o_re, self.o_im, self.logpsi = machine.per_sample_gradients(configurations, stream=stream)
        self.B, self.P = self.o_re.shape
        dev = self.o_re.device
        self.mean_re = torch.empty(self.P, dtype=torch.float32, device=dev)
        self.mean_im = torch.empty(self.P, dtype=torch.float32, device=dev)
        _lib.call('fk_o_mean', _lib.ptr(self.o_re), _lib.ptr(self.o_im), self.B, self.P, _lib.ptr(self.mean_re),
                  _lib.ptr(self.mean_im), _lib.stream_ptr(stream))
        self.scratch = torch.empty(int(_lib.load().fk_sr_scratch_doubles()), dtype=torch.float64, device=dev)
        self.u = torch.empty((self.B, 2), dtype=torch.float32, device=dev)
        self.stream = stream

    def oh(self, u, scale, shift=0.0, v=None, out=None):
        """scale * Re[(O-Obar)^H u] + shift * v."""
        if out is None:
            out = torch.empty(self.P, dtype=torch.float32, device=self.o_re.device)
        _lib.call('fk_oh_times_u', _lib.ptr(self.o_re), _lib.ptr(self.o_im), _lib.ptr(self.mean_re),
                  _lib.ptr(self.mean_im), _lib.ptr(u), self.B, self.P, float(scale), float(shift), _lib.ptr(v),
                  _lib.ptr(out), _lib.ptr(self.scratch), _lib.stream_ptr(self.stream))
        return out

    def s_matvec(self, v, diag_shift, out=None):
        """S v = Re[(O-Obar)^H ((O-Obar) v)]/B + shift v (optimizer.py:78-81)."""
        _lib.call('fk_o_times_v', _lib.ptr(self.o_re), _lib.ptr(self.o_im), _lib.ptr(self.mean_re),
                  _lib.ptr(self.mean_im), _lib.ptr(v), self.B, self.P, _lib.ptr(self.u), _lib.ptr(self.scratch),
                  _lib.stream_ptr(self.stream))
        return self.oh(self.u, 1.0 / self.B, diag_shift, v, out)

    def dot(self, a, b):
        out = torch.empty(1, dtype=torch.float64, device=a.device)
        _lib.call('fk_dot_f64', _lib.ptr(a), _lib.ptr(b), a.numel(), _lib.ptr(self.scratch), _lib.ptr(out),
                  _lib.stream_ptr(self.stream))
        return float(out.item())


def energy_gradient(machine, configurations, local_energies, energy=None):
    """d<E>/d theta (real parameters) = 2 Re[sum_b O_b conj(E_loc,b - E)]/B."""
    of = _OFactor(machine, configurations)
    if energy is None:
        energy = local_energies.mean()
    u = _complex_eloc_pairs(local_energies, energy)
    return of.oh(u, 2.0 / of.B)


def conjugate_gradient(apply, dot, rhs, tol=1e-4, max_iter=20):
    """linear_equations.py:34-137 (no preconditioner): x0 = 0, stop when
    i >= max_iter or ||r|| <= tol * ||rhs|| (norms as float32).  Vector
    updates run on the device (fk_axpy / fk_xpby); dot products in FP64."""
    x = torch.zeros_like(rhs)
    r = rhs.clone()
    p = rhs.clone()
    gamma = dot(r, r)
    tol_abs = float(torch.tensor(tol, dtype=torch.float32)) * float(torch.tensor(math.sqrt(gamma),
                                                                                   dtype=torch.float32))
    z = torch.empty_like(rhs)
    i = 0
    s = _lib.stream_ptr()
    n = rhs.numel()
    while (max_iter is None or i < max_iter) and float(torch.tensor(math.sqrt(gamma), dtype=torch.float32)) > tol_abs:
        apply(p, z)
        alpha = gamma / dot(p, z)
        _lib.call('fk_axpy', alpha, _lib.ptr(p), _lib.ptr(x), n, s)
        _lib.call('fk_axpy', -alpha, _lib.ptr(z), _lib.ptr(r), n, s)
        g_new = dot(r, r)
        _lib.call('fk_xpby', _lib.ptr(r), g_new / gamma, _lib.ptr(p), n, s)
        gamma = g_new
        i += 1
    return x, i, math.sqrt(gamma)


class ComplexValuesStochasticReconfiguration:
    """optimizers/stochastic_reconfiguration/optimizer.py:14-31 (real-parameter
    SR, see module docstring).  ``step(configurations, local_energies)``
    performs one update of ``machine.params`` and returns the metrics."""

    def __init__(self, machine, predictions_jacobian=None, lr=0.01, diag_shift=0.05, iterative_solver=True,
                 compute_jvp_instead_of_full_jacobian=False, conjugate_gradient_tol=1e-3,
                 iterative_solver_max_iterations=200, use_cholesky=True, **kwargs):
        if not iterative_solver:
            raise NotImplementedError('the direct (Cholesky) solver forms S explicitly; only the iterative '
                                      'solver is implemented')
        self.machine = machine
        self.lr = lr
        self.diag_shift = diag_shift
        self.conjugate_gradient_tol = conjugate_gradient_tol
        self.iterative_solver_max_iterations = iterative_solver_max_iterations
        self.metrics = {}

    def solve(self, configurations, local_energies, energy=None):
        """S^{-1} F for the batch; returns (x, iterations, residual_norm)."""
        of = _OFactor(self.machine, configurations)
        if energy is None:
            energy = local_energies.mean()
        u = _complex_eloc_pairs(local_energies, energy)
        force = of.oh(u, 1.0 / of.B)
        x, it, res = conjugate_gradient(lambda v, out: of.s_matvec(v, self.diag_shift, out), of.dot, force,
                                        tol=self.conjugate_gradient_tol, max_iter=self.iterative_solver_max_iterations)
        return x, it, res

    def step(self, configurations, local_energies, energy=None):
        x, it, res = self.solve(configurations, local_energies, energy)
        self.machine.params.add_(x, alpha=-self.lr)
        self.machine.mark_params_updated()
        self.metrics = {'stochastic_reconfiguration/conjugate_gradient_iterations': it,
                        'stochastic_reconfiguration/conjugate_gradient_residual_norm': res}
        return self.metrics


class Adam:
    """Keras Adam (beta 0.9 / 0.999) on the machine's flat parameters with the
    VMC energy gradient (configs 1 and 3 of the reference examples,
    examples/basic_autoregressive_2d.py:24)."""

    def __init__(self, machine, lr=1e-3, beta_1=0.9, beta_2=0.999, epsilon=1e-7):
        self.machine = machine
        self.opt = torch.optim.Adam([machine.params], lr=lr, betas=(beta_1, beta_2), eps=epsilon)

    def step(self, configurations, local_energies, energy=None):
        g = energy_gradient(self.machine, configurations, local_energies, energy)
        self.machine.params.grad = g
        self.opt.step()
        self.machine.params.grad = None
        self.machine.mark_params_updated()
        return g
