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
    """O (per-sample gradients) of one batch with its column means (zero means
    when ``centered`` is False: products with O itself)."""

    def __init__(self, machine, configurations, stream=None, centered=True):
        self.o_re, self.o_im, self.logpsi = machine.per_sample_gradients(configurations, stream=stream)
        self.B, self.P = self.o_re.shape
        dev = self.o_re.device
        if centered:
            self.mean_re = torch.empty(self.P, dtype=torch.float32, device=dev)
            self.mean_im = torch.empty(self.P, dtype=torch.float32, device=dev)
            _lib.call('fk_o_mean', _lib.ptr(self.o_re), _lib.ptr(self.o_im), self.B, self.P, _lib.ptr(self.mean_re),
                      _lib.ptr(self.mean_im), _lib.stream_ptr(stream))
        else:
            self.mean_re = torch.zeros(self.P, dtype=torch.float32, device=dev)
            self.mean_im = torch.zeros(self.P, dtype=torch.float32, device=dev)
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


def loss_gradient(machine, configurations, coefficients):
    """Gradient of the VMC loss sum_b 2 Re(log psi_b y_b) (optimization/
    loss.py:4-5) over one (mini-)batch of configurations with coefficients y
    (the second output of VariationalMonteCarlo.next_batch, conj(E_loc - E)/B):
    2 Re sum_b O_b y_b, uncentered, so that the sums over the mini-batches of
    a batch add up to the batch's energy gradient."""
    of = _OFactor(machine, configurations, centered=False)
    y = torch.as_tensor(coefficients).to(of.o_re.device).to(torch.complex128).reshape(-1)
    u = torch.view_as_real(torch.conj(y).resolve_conj()).to(torch.float32).contiguous()
    return of.oh(u, 2.0)


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
    """Keras Adam (tf.keras.optimizers.Adam, the optimizer of the reference's
    examples, e.g. examples/basic_autoregressive_2d.py:24) on the machine's
    flat parameters with the VMC energy gradient.  The update is Keras's, not
    torch's: lr_t = lr*sqrt(1-beta_2^t)/(1-beta_1^t) and
    p -= lr_t*m/(sqrt(v)+epsilon), epsilon added to the uncorrected sqrt(v)
    (fk_adam_step, FP32 on the device)."""

    def __init__(self, machine, lr=1e-3, beta_1=0.9, beta_2=0.999, epsilon=1e-7):
        self.machine = machine
        self.lr, self.beta_1, self.beta_2, self.epsilon = lr, beta_1, beta_2, epsilon
        self.iterations = 0
        self._m = torch.zeros_like(machine.params)
        self._v = torch.zeros_like(machine.params)

    @property
    def params(self):
        return self.machine.params

    def step(self, configurations, local_energies, energy=None):
        g = energy_gradient(self.machine, configurations, local_energies, energy)
        self.apply_gradient(g)
        return g

    def apply_gradient(self, g):
        """One Adam update of the machine's parameters with gradient g."""
        p = self.machine.params
        if p.device.type != 'cuda':
            raise _lib.FlowketHipError('Adam updates run on the GPU (fk_adam_step); the parameters are on '
                                       f'{p.device}')
        _lib.require_gpu()
        g = g.to(device=p.device, dtype=torch.float32).contiguous()
        if g.numel() != p.numel():
            raise ValueError(f'gradient has {g.numel()} values, the parameters {p.numel()}')
        self.iterations += 1
        _lib.call('fk_adam_step', _lib.ptr(p), _lib.ptr(self._m), _lib.ptr(self._v), _lib.ptr(g), p.numel(),
                  float(self.lr), float(self.beta_1), float(self.beta_2), float(self.epsilon), self.iterations,
                  _lib.stream_ptr())
        self.machine.mark_params_updated()

    def get_weights(self):
        """Optimizer state as a list of tensors (Keras ``optimizer.weights``
        order: iterations, first moments, second moments)."""
        it = torch.tensor([float(self.iterations)], dtype=torch.float64)
        return [it, self._m.detach().clone(), self._v.detach().clone()]

    def set_weights(self, weights):
        it, m, v = weights
        p = self.machine.params
        self.iterations = int(torch.as_tensor(it).reshape(-1)[0].item())
        self._m.copy_(torch.as_tensor(m).to(p.device, p.dtype).reshape(p.shape))
        self._v.copy_(torch.as_tensor(v).to(p.device, p.dtype).reshape(p.shape))


class SGD:
    """Plain gradient descent on a flat parameter tensor (Keras SGD without
    momentum): params <- params - lr * g."""

    def __init__(self, params, lr=0.01, on_update=None):
        self.params = params
        self.lr = lr
        self.iterations = 0
        self.on_update = on_update

    def apply_gradient(self, g):
        with torch.no_grad():
            self.params -= self.lr * g.to(self.params.dtype)
        self.iterations += 1
        if self.on_update is not None:
            self.on_update()

    def get_weights(self):
        return [torch.tensor([float(self.iterations)], dtype=torch.float64)]

    def set_weights(self, weights):
        self.iterations = int(torch.as_tensor(weights[0]).reshape(-1)[0].item())


class AccumulateGradientOptimizer:
    """optimizers/accumulate_gradient_optimizer.py:13-89 around an optimizer
    of this package (one with ``params`` and ``apply_gradient``): gradients are
    summed -- or averaged, dividing each by update_params_frequency -- over
    update_params_frequency calls, and the wrapped optimizer steps on the
    accumulated gradient every update_params_frequency-th call, after which
    the accumulator is cleared.  With ema_decay > 0 an exponential moving
    average of the parameters is kept at every update and set_weights_ema
    writes its bias-corrected value (e / (1 - decay^updates)) into them."""

    def __init__(self, orig_optimizer, update_params_frequency, accumulate_sum_or_mean=True, ema_decay=0):
        if update_params_frequency < 1:
            raise ValueError('update_params_frequency must be >= 1')
        self.orig = orig_optimizer
        self.update_params_frequency = int(update_params_frequency)
        self.accumulate_sum_or_mean = accumulate_sum_or_mean
        self.ema_decay = ema_decay
        self.accumulated_iterations = 0
        self.total_iterations = 0
        self._acc = torch.zeros_like(orig_optimizer.params)
        self._ema = torch.zeros_like(orig_optimizer.params) if ema_decay > 0 else None

    @property
    def params(self):
        return self.orig.params

    def set_update_params_frequency(self, update_params_frequency):
        self.update_params_frequency = int(update_params_frequency)

    def apply_gradient(self, g):
        self.accumulated_iterations += 1
        g = g.to(self._acc.dtype)
        if not self.accumulate_sum_or_mean:
            g = g / self.update_params_frequency
        self._acc += g
        if self.accumulated_iterations % self.update_params_frequency == 0:
            self.orig.apply_gradient(self._acc.clone())
            self._acc.zero_()
            if self._ema is not None:
                self.total_iterations += 1
                with torch.no_grad():
                    self._ema.mul_(self.ema_decay).add_((1.0 - self.ema_decay) * self.orig.params)

    def step(self, machine, configurations, coefficients):
        """One mini-batch of VariationalMonteCarlo (configurations and its
        loss coefficients) into the accumulator."""
        g = loss_gradient(machine, configurations, coefficients)
        self.apply_gradient(g)
        return g

    def get_weights(self):
        """The wrapped optimizer's state, then the call counters, the
        accumulator and (with ema_decay > 0) the parameter average."""
        counters = torch.tensor([float(self.accumulated_iterations), float(self.total_iterations)],
                                dtype=torch.float64)
        out = list(self.orig.get_weights()) + [counters, self._acc.detach().clone()]
        if self._ema is not None:
            out.append(self._ema.detach().clone())
        return out

    def set_weights(self, weights):
        n_orig = len(self.orig.get_weights())
        self.orig.set_weights(weights[:n_orig])
        rest = weights[n_orig:]
        counters = torch.as_tensor(rest[0]).reshape(-1)
        self.accumulated_iterations = int(counters[0].item())
        self.total_iterations = int(counters[1].item())
        self._acc.copy_(torch.as_tensor(rest[1]).to(self._acc.device, self._acc.dtype).reshape(self._acc.shape))
        if self._ema is not None:
            if len(rest) < 3:
                raise ValueError('the saved state has no parameter average (ema_decay was 0)')
            self._ema.copy_(torch.as_tensor(rest[2]).to(self._ema.device, self._ema.dtype).reshape(self._ema.shape))

    def set_weights_ema(self):
        if self._ema is None or self.total_iterations == 0:
            return
        with torch.no_grad():
            self.orig.params.copy_(self._ema / (1.0 - self.ema_decay ** self.total_iterations))
        mark = getattr(getattr(self.orig, 'machine', None), 'mark_params_updated', None)
        if mark is not None:
            mark()


def convert_to_accumulate_gradient_optimizer(orig_optimizer, update_params_frequency, accumulate_sum_or_mean=True,
                                             ema_decay=0):
    """optimizers/accumulate_gradient_optimizer.py:13-89 (no Horovod: one GPU)."""
    return AccumulateGradientOptimizer(orig_optimizer, update_params_frequency, accumulate_sum_or_mean, ema_decay)
