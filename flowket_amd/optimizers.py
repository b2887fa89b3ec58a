"""Parameter updates of the VMC loop.

``ComplexValuesStochasticReconfiguration`` follows
optimizers/stochastic_reconfiguration/optimizer.py:14-183 (same constructor
arguments, the same three solvers -- CG on a materialised O, CG on implicit
products, direct Cholesky / LU -- the CG stopping rule of
linear_equations.py:87-131 and the same metric names).  ConvNetAutoregressive2D
has real parameters and a complex log psi, which the reference optimizer does
not accept (complex_values_optimizer.py:12-25); the real-parameter form used
here is
    S = Re[(O-Obar)^H (O-Obar)]/B + shift*I,   F = Re[(O-Obar)^H (E_loc-E)]/B,
    theta <- theta - lr * S^{-1} F,
over the machine's trainable values (with weight normalisation: v and g).
For a holomorphic log psi of complex parameters stored as (real, imag) pairs
(the RBM) this real system is the reference's complex one written in real
coordinates, so the update is the same.  S is never formed by the iterative
solvers: with a materialised O it is applied as fk_o_times_v /
fk_oh_times_u, with compute_jvp_instead_of_full_jacobian as the machine's
forward-mode product O t (centred) followed by its batched backward product
Re[O^H u] (optimizer.py:116-124, optimizers/utils.py:5-34), so O is never
stored.  The energy gradient of the Adam path,
d<E>/d theta = 2 Re[sum_b O_b conj(E_loc,b - E)]/B (loss 2 Re(logpsi y),
optimization/loss.py:4-5), is one batched backward product.
"""
import math

import torch

from . import _lib


def _complex_pairs(z):
    z = torch.as_tensor(z).reshape(-1).to(torch.complex128)
    return torch.view_as_real(z).to(torch.float32).contiguous()


def _complex_eloc_pairs(local_energies, energy):
    return _complex_pairs(torch.as_tensor(local_energies).to(torch.complex128) - energy)


def _implicit(machine):
    f = getattr(machine, 'supports_implicit_products', None)
    return f is not None and f()


class _OFactor:
    """O (per-sample gradients over the trainable values) of one batch with
    its column means (zero means when ``centered`` is False: products with O
    itself).  ``jacobian`` (configurations -> (o_re, o_im)) replaces the
    machine's own per_sample_gradients when given."""

    def __init__(self, machine, configurations, stream=None, centered=True, jacobian=None):
        if jacobian is not None:
            self.o_re, self.o_im = [torch.as_tensor(t, dtype=torch.float32).contiguous()
                                    for t in jacobian(configurations)[:2]]
        else:
            self.o_re, self.o_im, _ = machine.per_sample_gradients(configurations, stream=stream)
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
        self.matvec_ws = None
        self.block_rows = 0

    def oh(self, u, scale, shift=0.0, v=None, out=None):
        """scale * Re[(O-Obar)^H u] + shift * v."""
        if out is None:
            out = torch.empty(self.P, dtype=torch.float32, device=self.o_re.device)
        _lib.call('fk_oh_times_u', _lib.ptr(self.o_re), _lib.ptr(self.o_im), _lib.ptr(self.mean_re),
                  _lib.ptr(self.mean_im), _lib.ptr(u), self.B, self.P, float(scale), float(shift), _lib.ptr(v),
                  _lib.ptr(out), _lib.ptr(self.scratch), _lib.stream_ptr(self.stream))
        return out

    def force(self, u):
        return self.oh(u, 1.0 / self.B)

    def s_matvec(self, v, diag_shift, out=None):
        """S v = Re[(O-Obar)^H ((O-Obar) v)]/B + shift v (optimizer.py:78-81)
        through fk_sr_matvec: two passes over O (block_rows = 0, the faster
        form on MI355X) or one pass in row blocks of block_rows."""
        if out is None:
            out = torch.empty(self.P, dtype=torch.float32, device=self.o_re.device)
        if self.block_rows and self.matvec_ws is None:
            nbytes = int(_lib.load().fk_sr_matvec_workspace_bytes(self.B, self.P))
            self.matvec_ws = torch.empty(nbytes // 8 + 2, dtype=torch.float64, device=self.o_re.device)
        ws = self.matvec_ws if self.block_rows else None
        _lib.call('fk_sr_matvec', _lib.ptr(self.o_re), _lib.ptr(self.o_im), _lib.ptr(self.mean_re),
                  _lib.ptr(self.mean_im), _lib.ptr(v), self.B, self.P, 1.0 / self.B, float(diag_shift),
                  _lib.ptr(out), _lib.ptr(self.u), int(self.block_rows), _lib.ptr(ws),
                  0 if ws is None else ws.numel() * 8, _lib.ptr(self.scratch), _lib.stream_ptr(self.stream))
        return out

    def s_matvec_two_pass(self, v, diag_shift, out=None):
        """The same product as the two separate calls fk_o_times_v, fk_oh_times_u."""
        _lib.call('fk_o_times_v', _lib.ptr(self.o_re), _lib.ptr(self.o_im), _lib.ptr(self.mean_re),
                  _lib.ptr(self.mean_im), _lib.ptr(v), self.B, self.P, _lib.ptr(self.u), _lib.ptr(self.scratch),
                  _lib.stream_ptr(self.stream))
        return self.oh(self.u, 1.0 / self.B, diag_shift, v, out)

    def dense_s(self):
        """S without the shift, formed (P, P) in FP64 for the direct solver:
        (O-Obar) in FP64 and one library GEMM per part."""
        a = self.o_re.double() - self.mean_re.double()
        b = self.o_im.double() - self.mean_im.double()
        return (a.T @ a + b.T @ b) / self.B


class _ImplicitS:
    """S products through the machine's implicit O products
    (get_stochastic_reconfiguration_matrix_vector_product_via_jvp,
    optimizer.py:116-124): w = O v (forward mode), w -= mean(w), then
    Re[O^H w]/B (one batched backward pass) + shift v."""

    def __init__(self, machine, configurations):
        self.machine = machine
        self.x = machine._spins(configurations)
        self.B = self.x.shape[0]
        self.P = machine.n_trainable
        self.scratch = torch.empty(int(_lib.load().fk_sr_scratch_doubles()), dtype=torch.float64,
                                   device=machine.device)

    def force(self, u):
        # the cotangent (E_loc - E) sums to zero, so Re[O^H u] = Re[(O-Obar)^H u]
        return self.machine.vjp(self.x, u, 1.0 / self.B)

    def s_matvec(self, v, diag_shift, out=None):
        w = self.machine.jvp(self.x, v)
        _lib.call('fk_center_pairs', _lib.ptr(w), self.B, _lib.ptr(self.scratch), _lib.stream_ptr())
        g = self.machine.vjp(self.x, w, 1.0 / self.B)
        if out is None:
            out = g
        else:
            out.copy_(g)
        if diag_shift:
            _lib.call('fk_axpy', float(diag_shift), _lib.ptr(v), _lib.ptr(out), out.numel(), _lib.stream_ptr())
        return out


def energy_gradient(machine, configurations, local_energies, energy=None):
    """d<E>/d theta over the trainable values =
    2 Re[sum_b O_b conj(E_loc,b - E)]/B: the machine's batched backward
    product when it has one, else through a materialised O."""
    if energy is None:
        energy = torch.as_tensor(local_energies).mean()
    u = _complex_eloc_pairs(local_energies, energy)
    B = u.shape[0]
    if _implicit(machine):
        return machine.vjp(configurations, u.to(machine.device), 2.0 / B)
    of = _OFactor(machine, configurations)
    return of.oh(u.to(of.o_re.device), 2.0 / of.B)


def loss_gradient(machine, configurations, coefficients):
    """Gradient of the VMC loss sum_b 2 Re(log psi_b y_b) (optimization/
    loss.py:4-5) over one (mini-)batch of configurations with coefficients y
    (the second output of VariationalMonteCarlo.next_batch, conj(E_loc - E)/B):
    2 Re sum_b O_b y_b, uncentered, so that the sums over the mini-batches of
    a batch add up to the batch's energy gradient."""
    y = torch.as_tensor(coefficients).reshape(-1).to(torch.complex128)
    u = _complex_pairs(torch.conj(y).resolve_conj())
    if _implicit(machine):
        return machine.vjp(configurations, u.to(machine.device), 2.0)
    of = _OFactor(machine, configurations, centered=False)
    return of.oh(u.to(of.o_re.device), 2.0)


def conjugate_gradient(apply, rhs, tol=1e-4, max_iter=20, check_every=10):
    """linear_equations.py:34-137 (no preconditioner): x0 = 0, stop when
    i >= max_iter or ||r|| <= tol * ||rhs|| (norms as float32).  The scalars
    live on the device (fk_cg_init / fk_cg_step): ``apply(p, z)`` writes
    z = A p, and the host reads the done flag only every ``check_every``
    iterations (None: never -- every iteration is launched, the ones after
    convergence change nothing, and the solve can be captured in a graph).
    Returns (x, iterations, residual norm)."""
    n = rhs.numel()
    dev = rhs.device
    x = torch.empty_like(rhs)
    r = torch.empty_like(rhs)
    p = torch.empty_like(rhs)
    z = torch.empty_like(rhs)
    state = torch.zeros(int(_lib.load().fk_cg_state_doubles()), dtype=torch.float64, device=dev)
    scratch = torch.empty(int(_lib.load().fk_sr_scratch_doubles()), dtype=torch.float64, device=dev)
    s = _lib.stream_ptr()
    max_iter = 2 ** 31 - 1 if max_iter is None else int(max_iter)
    _lib.call('fk_cg_init', _lib.ptr(rhs), _lib.ptr(x), _lib.ptr(r), _lib.ptr(p), n, float(tol), max_iter,
              _lib.ptr(state), _lib.ptr(scratch), s)
    for i in range(max_iter):
        if check_every and i % check_every == 0 and state[4].item() != 0.0:
            break
        apply(p, z)
        _lib.call('fk_cg_step', _lib.ptr(x), _lib.ptr(r), _lib.ptr(p), _lib.ptr(z), n, _lib.ptr(state),
                  _lib.ptr(scratch), s)
    st = state.cpu()
    return x, int(st[5].item()), math.sqrt(st[0].item())


class ComplexValuesStochasticReconfiguration:
    """optimizers/stochastic_reconfiguration/optimizer.py:14-31 (real-parameter
    SR, see module docstring).  ``step(configurations, local_energies)``
    performs one update of the machine's trainable values and returns the
    metrics.

    - ``predictions_jacobian``: optional callable configurations ->
      (o_re, o_im) (B, n_trainable) used instead of the machine's own
      per-sample gradients (the reference takes the Jacobian the same way).
    - ``compute_jvp_instead_of_full_jacobian`` with the iterative solver:
      implicit S products, O never formed (optimizer.py:37-40,116-124).
    - ``iterative_solver=False``: S formed in FP64 and solved by Cholesky
      (``use_cholesky``) or LU (optimizer.py:55-67); ``add_s_matrix_stats``
      reports its rank and smallest eigenvalue (:173-183).
    """

    def __init__(self, machine, predictions_jacobian=None, lr=0.01, diag_shift=0.05, iterative_solver=True,
                 compute_jvp_instead_of_full_jacobian=False, use_energy_loss=False, add_s_matrix_stats=False,
                 conjugate_gradient_tol=1e-3, iterative_solver_max_iterations=200, use_cholesky=True, name=None,
                 cg_check_every=10):
        if use_energy_loss:
            raise NotImplementedError('use_energy_loss=True (the branch the reference marks "todo fix") is not '
                                      'implemented; the force is the Jacobian product of optimizer.py:100-108')
        if predictions_jacobian is not None and not callable(predictions_jacobian):
            raise TypeError('predictions_jacobian must be a callable configurations -> (o_re, o_im)')
        if compute_jvp_instead_of_full_jacobian and iterative_solver and not _implicit(machine):
            raise NotImplementedError('compute_jvp_instead_of_full_jacobian needs the machine\'s implicit '
                                      'products (ConvNetAutoregressive2D with a multiple of 32 channels)')
        self.machine = machine
        self.predictions_jacobian = predictions_jacobian
        self.lr = lr
        self.diag_shift = diag_shift
        self.iterative_solver = iterative_solver
        self.compute_jvp_instead_of_full_jacobian = compute_jvp_instead_of_full_jacobian
        self.add_s_matrix_stats = add_s_matrix_stats
        self.conjugate_gradient_tol = conjugate_gradient_tol
        self.iterative_solver_max_iterations = iterative_solver_max_iterations
        self.use_cholesky = use_cholesky
        self.name = name or 'ComplexValuesStochasticReconfiguration'
        self.cg_check_every = cg_check_every
        self.iterations = 0
        self.metrics = {}

    def _products(self, configurations):
        if self.iterative_solver and self.compute_jvp_instead_of_full_jacobian:
            return _ImplicitS(self.machine, configurations)
        return _OFactor(self.machine, configurations, jacobian=self.predictions_jacobian)

    def solve(self, configurations, local_energies, energy=None):
        """S^{-1} F for the batch; returns (x, metrics)."""
        prod = self._products(configurations)
        if energy is None:
            energy = torch.as_tensor(local_energies).mean()
        u = _complex_eloc_pairs(local_energies, energy).to(self.machine.device)
        force = prod.force(u)
        if self.iterative_solver:
            x, it, res = conjugate_gradient(lambda v, out: prod.s_matvec(v, self.diag_shift, out), force,
                                            tol=self.conjugate_gradient_tol,
                                            max_iter=self.iterative_solver_max_iterations,
                                            check_every=self.cg_check_every)
            return x, {'stochastic_reconfiguration/conjugate_gradient_iterations': it,
                       'stochastic_reconfiguration/conjugate_gradient_residual_norm': res}
        P = prod.P
        need = 8 * P * P * 3
        if need > 0.5 * torch.cuda.get_device_properties(self.machine.device).total_memory:
            raise MemoryError(f'the direct solver forms S ({P} x {P} in FP64, {need / 2 ** 30:.0f} GiB with '
                              'the factorisation); use the iterative solver')
        s = prod.dense_s()
        metrics = {}
        if self.add_s_matrix_stats:
            ev = torch.linalg.eigvalsh(s).abs()
            tol = torch.finfo(torch.float32).eps * P * s.abs().max()
            kept = ev[ev > tol]
            metrics['stochastic_reconfiguration/s_matrix_rank'] = int(kept.numel())
            metrics['stochastic_reconfiguration/s_matrix_min_eigval'] = float(kept.min()) if kept.numel() else 0.0
        s += self.diag_shift * torch.eye(P, dtype=s.dtype, device=s.device)
        f = force.double()[:, None]
        if self.use_cholesky:
            x = torch.cholesky_solve(f, torch.linalg.cholesky(s))
        else:
            x = torch.linalg.solve(s, f)
        return x[:, 0].to(torch.float32), metrics

    def step(self, configurations, local_energies, energy=None):
        """One update w <- w - lr * S^{-1} F (optimizer.py:33-42)."""
        x, metrics = self.solve(configurations, local_energies, energy)
        self.machine.trainable.add_(x, alpha=-self.lr)
        self.machine.mark_params_updated()
        self.iterations += 1
        self.metrics = metrics
        return metrics


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
        self._m = torch.zeros_like(machine.trainable)
        self._v = torch.zeros_like(machine.trainable)

    @property
    def params(self):
        return self.machine.trainable

    def step(self, configurations, local_energies, energy=None):
        g = energy_gradient(self.machine, configurations, local_energies, energy)
        self.apply_gradient(g)
        return g

    def apply_gradient(self, g):
        """One Adam update of the machine's trainable values with gradient g."""
        p = self.machine.trainable
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
        p = self.machine.trainable
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
