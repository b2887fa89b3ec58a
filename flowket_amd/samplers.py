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


class FastAutoregressiveSampler(AutoregressiveSampler):
    """deepar/samplers/fast_autoregressive.py:13-76: same distribution and,
    given the same uniforms, the same draws as AutoregressiveSampler, with each
    (layer, site) activation computed once (fk_fast_sample: one wavefront per
    16 samples walking the lattice, activation cache in device memory).
    ``mini_batch_size`` bounds the cache: batches larger than it are drawn in
    chunks (:30-35)."""

    def __init__(self, conditional_log_probs_machine, batch_size, mini_batch_size=None, seed=None,
                 uniforms=None, precision=None, **kwargs):
        super().__init__(conditional_log_probs_machine, batch_size, mini_batch_size, seed=seed, uniforms=uniforms)
        self.precision = precision
        self._ws = None
        self.last_cond_logp = None
        self.keep_cond_logp = False

    def _workspace(self, rows, prec_code):
        m = self.machine
        nbytes = _lib.load().fk_fast_sampler_workspace_bytes(m.H, m.W, m.depth, m.num_of_channels, prec_code, rows)
        if nbytes < 0:
            raise _lib.FlowketHipError(
                f'fast sampler does not implement channels={m.num_of_channels} (supported: 32, 64)')
        if self._ws is None or self._ws.numel() < nbytes:
            self._ws = torch.empty(nbytes, dtype=torch.uint8, device=m.device)
        return self._ws

    def __next__(self):
        _lib.require_gpu()
        from .machines import _ACT, _PREC
        m = self.machine
        precision = self.precision or m.precision
        prec = _PREC[precision]
        n = m.H * m.W
        u = self._draw_uniforms(m.device)
        wbuf, abuf = m._pack(precision)
        out = torch.empty((self.batch_size, n), dtype=torch.int8, device=m.device)
        cond = torch.empty((self.batch_size, n, 2), dtype=torch.float32, device=m.device) \
            if self.keep_cond_logp else None
        chunk = self.mini_batch_size
        s = _lib.stream_ptr()
        for start in range(0, self.batch_size, chunk):
            rows = min(chunk, self.batch_size - start)
            ws = self._workspace(rows, prec)
            _lib.call('fk_fast_sample', _lib.ptr(wbuf), _lib.ptr(abuf), prec, m.H, m.W, m.depth,
                      m.num_of_channels, _ACT[m.activation], rows, _lib.ptr(u[start:start + rows]),
                      _lib.ptr(ws), _lib.ptr(out[start:start + rows]),
                      _lib.ptr(cond[start:start + rows]) if cond is not None else None, s)
        self.last_cond_logp = cond
        return out.reshape((self.batch_size,) + self.input_size)
