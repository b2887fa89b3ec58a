"""VMC driver (optimization/variational_monte_carlo.py,
optimization/mini_batch_generator.py, optimization/loss.py)."""
import abc
import math
import time

import torch

from .observables import BaseObservable, Observable


class MiniBatchGenerator:
    """optimization/mini_batch_generator.py:5-45."""

    def __init__(self, batch_size, mini_batch_size):
        self.set_batch_size(batch_size, mini_batch_size)

    @abc.abstractmethod
    def next_batch(self):
        pass

    def set_batch_size(self, batch_size, mini_batch_size=None):
        if mini_batch_size is None:
            mini_batch_size = batch_size
        if mini_batch_size > batch_size:
            mini_batch_size = batch_size
        self.batch_size = batch_size
        self.mini_batch_size = mini_batch_size
        self._idx = self.batch_size
        self.update_params_frequency = math.ceil(self.batch_size / float(self.mini_batch_size))
        return self.update_params_frequency

    def next_mini_batch_size(self):
        if self._idx + self.mini_batch_size > self.batch_size:
            self._x, self._y = self.next_batch()
            self._idx = 0
        self._idx += self.mini_batch_size
        return self._x[self._idx - self.mini_batch_size:self._idx], self._y[self._idx - self.mini_batch_size:self._idx]

    def next(self):
        return self.next_mini_batch_size()

    def __iter__(self):
        return self

    def __next__(self):
        return self.next()

    def to_generator(self):
        while True:
            yield next(self)


class VariationalMonteCarlo(MiniBatchGenerator):
    """optimization/variational_monte_carlo.py:12-50."""

    def __init__(self, model, operator, sampler, mini_batch_size=None, wave_function_evaluation_batch_size=None,
                 **kwargs):
        super().__init__(sampler.batch_size, mini_batch_size)
        self.model = model
        self.operator = operator
        self.sampler = sampler
        self.wave_function = model
        self.current_batch = None
        self.energy_observable = operator if isinstance(operator, BaseObservable) else Observable(operator)

    def set_sampler(self, sampler, mini_batch_size=None):
        self.sampler = sampler
        return self.set_batch_size(sampler.batch_size, mini_batch_size)

    def _update_batch_local_energy(self):
        self.current_energy, self.current_local_energy_variance, self.current_local_energy = \
            self.energy_observable.estimate(self.wave_function, self.current_batch)

    def loss_coefficients(self):
        """conj(E_loc - E) (:38-40)."""
        return torch.conj(self.current_local_energy - self.current_energy)

    def next_batch(self):
        self.start_time = time.time()
        self.current_batch = next(self.sampler)
        self.sampling_end_time = time.time()
        self._update_batch_local_energy()
        self.local_energy_end_time = time.time()
        return self.current_batch, self.loss_coefficients() / self.batch_size


def loss_for_energy_minimization(y_true, y_pred):
    """optimization/loss.py:4-5: 2 Re(logpsi * y)."""
    return 2.0 * torch.real(y_pred * y_true)
