"""Samplers with FlowKet's iterator interface (deepar/samplers/base_sampler.py).

``next(sampler)`` returns a (batch_size, *input_size) int8 device tensor of
+-1 spins (the reference returns float64 numpy; ``.cpu().numpy()`` gives the
same values).  Randomness: a seeded torch generator on the device, or
injected uniforms for draw-by-draw parity with the reference rule.
"""
import abc
import copy

import numpy as np
import torch

from . import _lib


class Sampler(abc.ABC):
    """deepar/samplers/base_sampler.py:4-25."""

    def __init__(self, input_size, batch_size, mini_batch_size=None):
        self.input_size = tuple(input_size)
        self._set_batch_size(batch_size, mini_batch_size)

    def _set_batch_size(self, batch_size, mini_batch_size=None):
        if mini_batch_size is None:
            mini_batch_size = batch_size
        if batch_size < mini_batch_size:
            mini_batch_size = batch_size
        self.batch_size = batch_size
        self.mini_batch_size = mini_batch_size

    def copy_with_new_batch_size(self, batch_size, mini_batch_size=None):
        new = copy.copy(self)
        new._set_batch_size(batch_size, mini_batch_size)
        return new

    def __iter__(self):
        return self

    @abc.abstractmethod
    def __next__(self):
        pass


def _machine_of(model):
    return getattr(model, 'machine', model)


class AutoregressiveSampler(Sampler):
    """Exact autoregressive sampling by H*W full forwards
    (deepar/samplers/autoregressive.py:10-48, flowket alias with
    zero_base=False, samplers/__init__.py:8-14).  Unsampled sites are 0 in
    the forwards; site (h, w) is +1 iff exp(log p(+1)) > u (strict)."""

    def __init__(self, conditional_log_probs_machine, batch_size, mini_batch_size=None, seed=None,
                 uniforms=None, **kwargs):
        self.machine = _machine_of(conditional_log_probs_machine)
        super().__init__(self.machine.input_size, batch_size, mini_batch_size)
        self.generator = None
        self.seed = seed
        self.uniforms = uniforms

    def _draw_uniforms(self, device):
        if self.uniforms is not None:
            u = torch.as_tensor(self.uniforms, dtype=torch.float64).to(device)
            return u.reshape(self.batch_size, -1).contiguous()
        if self.generator is None:
            self.generator = torch.Generator(device=device)
            if self.seed is not None:
                self.generator.manual_seed(int(self.seed))
        return torch.rand((self.batch_size, int(np.prod(self.input_size))), dtype=torch.float64,
                          device=device, generator=self.generator)

    def __next__(self):
        _lib.require_gpu()
        m = self.machine
        n = m.H * m.W
        u = self._draw_uniforms(m.device)
        batch = torch.zeros((self.batch_size, n), dtype=torch.int8, device=m.device)
        s = _lib.stream_ptr()
        for site in range(n):
            _, cond = m.forward(batch, want_cond=True)
            _lib.call('fk_ar_draw_site', _lib.ptr(cond), _lib.ptr(u), self.batch_size, n, site, _lib.ptr(batch), s)
        return batch.reshape((self.batch_size,) + self.input_size)
