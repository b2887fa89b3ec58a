"""Exact enumeration path: FlowKet's ExactVariational, ExactObservable and
ExactSampler (optimization/exact_variational.py:10-161,
samplers/exact_sampler.py:7-20) over all 2^N configurations on the GPU.

log psi of every configuration comes from the machine's own kernel in
batches; the local energies then need no further network evaluation, because
every connected configuration of an enumerated state is itself enumerated
(fk_exact_local_energy locates it by index).  Probabilities, the energy, its
variance and the gradient coefficients are FP64 with deterministic
reductions.
"""
import numpy as np
import torch

from . import _lib


def _conj_pair(z):
    return complex(z[0], -z[1])


class ExactVariational:
    """optimization/exact_variational.py:103-161.  ``model`` is a machine of
    this package (ConvNetAutoregressive2D or RBM), ``operator`` one of its
    operators; ``batch_size`` must divide 2^N (as the reference requires)."""

    def __init__(self, model, operator, batch_size):
        _lib.require_gpu()
        self.model = model
        self.operator = operator
        self.device = model.device
        self.input_size = tuple(operator.hilbert_state_shape)
        self.number_of_spins = int(np.prod(self.input_size))
        if self.number_of_spins > 30:
            raise ValueError('exact enumeration needs 2^N states in memory; N <= 30')
        self.num_of_states = 2 ** self.number_of_spins
        self._set_batch_size(batch_size)
        n = self.number_of_spins
        self.states = torch.empty((self.num_of_states, n), dtype=torch.int8, device=self.device)
        _lib.call('fk_exact_states', n, 0, self.num_of_states, _lib.ptr(self.states), _lib.stream_ptr())
        self._logpsi = torch.empty((self.num_of_states, 2), dtype=torch.float32, device=self.device)
        self.probs = torch.empty(self.num_of_states, dtype=torch.float64, device=self.device)
        self._scalars = torch.empty(2, dtype=torch.float64, device=self.device)
        self._ws = torch.empty(int(_lib.load().fk_exact_workspace_doubles(self.num_of_states)), dtype=torch.float64,
                               device=self.device)
        self.wave_function_norm_squared = None
        self.energy_observable = ExactObservable(self, operator, calculate_variance_of_the_local_operator=True)

    def _set_batch_size(self, batch_size):
        """exact_variational.py:119-125."""
        if batch_size > self.num_of_states:
            batch_size = self.num_of_states
        if self.num_of_states % batch_size != 0:
            raise Exception('In exact the batch size must divide the total number of states in the system')
        self.batch_size = batch_size
        self.num_of_batch_until_full_cycle = self.num_of_states // self.batch_size

    @property
    def wave_function(self):
        """log psi of every configuration, (2^N,) complex128 (the reference
        keeps the model's complex64 output widened to complex128)."""
        return torch.complex(self._logpsi[:, 0].double(), self._logpsi[:, 1].double())

    @property
    def log_probs(self):
        return 2.0 * self._logpsi[:, 0].double() - self._scalars[0]

    def _update_wave_function_arrays(self):
        """exact_variational.py:127-133."""
        for i in range(0, self.num_of_states, self.batch_size):
            out, _ = self.model.forward(self.states[i:i + self.batch_size])
            self._logpsi[i:i + self.batch_size].copy_(out)
        _lib.call('fk_exact_normalize', _lib.ptr(self._logpsi), self.num_of_states, _lib.ptr(self.probs),
                  _lib.ptr(self._scalars), _lib.ptr(self._ws), _lib.stream_ptr())
        self.wave_function_norm_squared = float(np.exp(self._scalars[0].item()))

    def _update_local_energy(self):
        """exact_variational.py:135-138."""
        self.energy_observable.update_local_energy()
        self.energy_grad_coefficients = self.energy_observable.coefficients

    def machine_updated(self):
        self._update_wave_function_arrays()
        self._update_local_energy()

    def states_batch(self, i):
        return self.states[i:i + self.batch_size].reshape((-1,) + self.input_size)

    def to_generator(self):
        """exact_variational.py:153-158: (configurations, coefficients) of
        every batch after each machine update."""
        while True:
            self.machine_updated()
            for i in range(0, self.num_of_states, self.batch_size):
                yield self.states_batch(i), self.energy_grad_coefficients[i:i + self.batch_size]

    def __iter__(self):
        return self.to_generator()


class ExactObservable:
    """optimization/exact_variational.py:10-100 for one operator: the find_conn
    of every configuration is taken once; update_local_energy evaluates the
    local energies of the current psi by index, the energy
    (``current_energy`` = sum_k p_k conj(E_loc(k)), the reference's form),
    the variance of Re E_loc and the gradient coefficients."""

    def __init__(self, exact_variational, operator, calculate_variance_of_the_local_operator=False):
        self.exact_variational = exact_variational
        self.operator = operator
        self.calculate_variance_of_the_local_operator = calculate_variance_of_the_local_operator
        ev = exact_variational
        r = operator.find_conn_device(ev.states, want_conn=False)
        self._mel, self._use = r['mel'], r['use']
        self._tables = operator.tables(ev.device)
        self.local_energies = torch.empty((ev.num_of_states, 2), dtype=torch.float64, device=ev.device)
        self.coefficients_pairs = torch.empty((ev.num_of_states, 2), dtype=torch.float64, device=ev.device)
        self._moments = torch.empty(3, dtype=torch.float64, device=ev.device)

    def update_local_energy(self):
        ev = self.exact_variational
        _lib.call('fk_exact_local_energy', _lib.ptr(ev._logpsi), ev.num_of_states, ev.number_of_spins,
                  self.operator.kind, _lib.ptr(self._tables['conn_a']), _lib.ptr(self._tables['conn_b']),
                  _lib.ptr(self._mel), _lib.ptr(self._use), self.operator.num_connections,
                  _lib.ptr(self.local_energies), _lib.stream_ptr())
        _lib.call('fk_exact_energy', _lib.ptr(ev.probs), _lib.ptr(self.local_energies), ev.num_of_states,
                  _lib.ptr(self.coefficients_pairs), _lib.ptr(self._moments), _lib.ptr(ev._ws), _lib.stream_ptr())
        m = self._moments.cpu().numpy()
        self.current_energy = _conj_pair(m)
        if self.calculate_variance_of_the_local_operator:
            self.current_local_energy_variance = float(m[2])

    @property
    def naive_energies(self):
        """E_loc of every configuration, (2^N,) complex128."""
        return torch.complex(self.local_energies[:, 0], self.local_energies[:, 1])

    @property
    def coefficients(self):
        """p_k conj(E_loc(k) - E), (2^N,) complex128."""
        return torch.complex(self.coefficients_pairs[:, 0], self.coefficients_pairs[:, 1])


class ExactSampler:
    """samplers/exact_sampler.py:7-20: configurations drawn from |psi|^2 of
    the last ``machine_updated``.  Uniforms come from a seeded torch
    generator, or are injected with ``uniforms`` (batch,) per call."""

    def __init__(self, exact_variational, batch_size, seed=0, uniforms=None):
        self.exact_variational = exact_variational
        self.batch_size = batch_size
        self.input_size = exact_variational.input_size
        self.uniforms = uniforms
        self._gen = torch.Generator(device=exact_variational.device)
        self._gen.manual_seed(seed)

    def __iter__(self):
        return self

    def __next__(self):
        ev = self.exact_variational
        if self.uniforms is not None:
            u = torch.as_tensor(self.uniforms, dtype=torch.float64).reshape(-1).to(ev.device).contiguous()
            if u.numel() != self.batch_size:
                raise ValueError(f'ExactSampler: {u.numel()} injected uniforms for a batch of {self.batch_size}')
        else:
            u = torch.rand(self.batch_size, generator=self._gen, device=ev.device, dtype=torch.float64)
        out = torch.empty((self.batch_size, ev.number_of_spins), dtype=torch.int8, device=ev.device)
        _lib.call('fk_exact_sample', _lib.ptr(ev.probs), ev.num_of_states, ev.number_of_spins, _lib.ptr(u),
                  self.batch_size, _lib.ptr(out), _lib.ptr(ev._ws), _lib.stream_ptr())
        return out.reshape((self.batch_size,) + self.input_size)
