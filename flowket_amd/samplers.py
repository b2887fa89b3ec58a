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
            n = int(np.prod(self.input_size))
            if u.numel() != self.batch_size * n:
                raise ValueError(f'{u.numel()} injected uniforms for a batch of {self.batch_size} x {n} sites')
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


def sum_correlations(correlations):
    """Geyer pair truncation (samplers/metropolis_hastings.py:8-12)."""
    for i in range(1, correlations.shape[0] // 2):
        if correlations[2 * i - 1] + correlations[2 * i] < 0:
            return correlations[:2 * i].sum()
    return correlations.sum()


class MetropolisHastingsSampler(Sampler):
    """samplers/metropolis_hastings.py:15-89.  ``num_of_chains`` chains
    (uniform +-1 start), per ``next``: (samples_per_chain // discard_ratio) *
    unused_sampels burn-in sweeps, then (unused_sampels + 1) sweeps per
    recorded sample; one sweep = one proposal per chain.  RBM machines run the
    whole chain in one fused kernel (fk_mh_rbm, incremental theta updates);
    other machines use fk_mh_propose / the machine's log psi / fk_mh_accept
    per sweep.  ``randomness`` may inject (sites, moves, uniforms) arrays of
    shape (sweeps, chains) for draw-by-draw parity."""

    kind = 0

    def __init__(self, machine, batch_size, num_of_chains=1, unused_sampels=0, discard_ratio=10, seed=None,
                 mini_batch_size=None, **kwargs):
        self.machine = _machine_of(machine)
        super().__init__(self.machine.input_size, batch_size, mini_batch_size)
        if self.batch_size % num_of_chains != 0:
            raise Exception('Num of samplers must divide the batch size')
        self.num_of_chains = num_of_chains
        self.unused_sampels = unused_sampels
        self.discard_ratio = discard_ratio
        self.acceptance_ratio = 1.0
        self.seed = int(seed) if seed is not None else int(np.random.randint(0, 2 ** 62))
        self._calls = 0
        self.randomness = None
        dev = self.machine.device
        g = torch.Generator().manual_seed(self.seed)
        n = int(np.prod(self.input_size))
        self.sample = (torch.randint(0, 2, (num_of_chains, n), generator=g) * 2 - 1).to(torch.int8).to(dev)
        shape = self.input_size if len(self.input_size) == 2 else (1, n)
        self.H, self.W = int(shape[0]), int(shape[1])

    def _counts(self):
        spc = self.batch_size // self.num_of_chains
        burn = (spc // self.discard_ratio) * self.unused_sampels if self.discard_ratio > 0 else 0
        return spc, burn

    def _nonfinite_flag(self):
        """Device flag the acceptance kernels set when a candidate's log psi
        is not finite (cleared here; read by _raise_if_nonfinite)."""
        if getattr(self, '_nonfinite', None) is None:
            self._nonfinite = torch.zeros(1, dtype=torch.int32, device=self.machine.device)
        else:
            self._nonfinite.zero_()
        return self._nonfinite

    def _raise_if_nonfinite(self):
        if int(self._nonfinite.item()):
            raise Exception('candidates_machine_values has not finite element')

    def _checked_randomness(self, steps, dev):
        """The injected (sites, moves, uniforms) arrays as device tensors,
        after checking that they cover every sweep and chain and index inside
        the lattice (the kernels read them without bounds checks)."""
        sites, moves, us = self.randomness
        C, n = self.num_of_chains, self.sample.shape[1]
        sites = torch.as_tensor(np.asarray(sites)).to(torch.int32)
        us = torch.as_tensor(np.asarray(us), dtype=torch.float64)
        if tuple(sites.shape) != (steps, C) or tuple(us.shape) != (steps, C):
            raise ValueError(f'injected sites and uniforms must have shape {(steps, C)} (sweeps, chains), got '
                             f'{tuple(sites.shape)} and {tuple(us.shape)}')
        if sites.numel() and (int(sites.min()) < 0 or int(sites.max()) >= n):
            raise ValueError(f'injected sites must lie in [0, {n})')
        if self.kind == 1:
            if moves is None:
                raise ValueError('exchange moves need injected moves')
            moves = torch.as_tensor(np.asarray(moves)).to(torch.int32)
            if tuple(moves.shape) != (steps, C):
                raise ValueError(f'injected moves must have shape {(steps, C)}')
            if moves.numel() and (int(moves.min()) < 0 or int(moves.max()) > 8):
                raise ValueError('injected moves must lie in [0, 9)')
            moves = moves.to(dev).contiguous()
        else:
            moves = None
        return sites.to(dev).contiguous(), moves, us.to(dev).contiguous()

    def __next__(self):
        _lib.require_gpu()
        spc, burn = self._counts()
        n = self.sample.shape[1]
        dev = self.machine.device
        out = torch.empty((self.num_of_chains, spc, n), dtype=torch.int8, device=dev)
        steps = burn + spc * (self.unused_sampels + 1)
        flag = self._nonfinite_flag()
        self._injected = None if self.randomness is None else self._checked_randomness(steps, dev)
        from .machines import RBM
        if isinstance(self.machine, RBM):
            accepts = torch.zeros(self.num_of_chains, dtype=torch.int64, device=dev)
            r_site = r_move = r_u = None
            if self._injected is not None:
                r_site, r_move, r_u = self._injected
            w = [_lib.ptr(t) for t in self.machine.parts()]
            _lib.call('fk_mh_rbm', *w, n, self.machine.n_hidden, self.H, self.W, self.kind, self.num_of_chains,
                      burn, spc, self.unused_sampels, _lib.ptr(self.sample), _lib.ptr(out), _lib.ptr(accepts), None,
                      _lib.ptr(r_site), _lib.ptr(r_move), _lib.ptr(r_u),
                      (self.seed * 1000003 + self._calls) & 0xFFFFFFFFFFFFFFFF, _lib.ptr(flag), _lib.stream_ptr())
            acc_total = int(accepts.sum().item())
        else:
            acc_total = self._generic(out, burn, spc, steps)
        self._raise_if_nonfinite()
        self._calls += 1
        self.acceptance_ratio = acc_total / (self.batch_size * (self.unused_sampels + 1))
        return out.reshape((self.batch_size,) + self.input_size)

    def _plain_sweeps(self, out, n_prop, acc_total, salt, logpsi=None):
        """n_prop proposals per chain from the current states (no burn-in),
        recording the final states into out[:, 0]; per-chain accept counts
        are added to acc_total (int64, device) without a host round trip.
        For the RBM the final states' log psi is written to ``logpsi``
        ((chains, 2) float32) when given; returns whether it was."""
        from .machines import RBM
        C, n = self.num_of_chains, self.sample.shape[1]
        dev = self.machine.device
        accepts = torch.empty(C, dtype=torch.int64, device=dev)
        if isinstance(self.machine, RBM):
            w = [_lib.ptr(t) for t in self.machine.parts()]
            _lib.call('fk_mh_rbm', *w, n, self.machine.n_hidden, self.H, self.W, self.kind, C, 0, 1, n_prop - 1,
                      _lib.ptr(self.sample), _lib.ptr(out), _lib.ptr(accepts), _lib.ptr(logpsi), None, None, None,
                      (self.seed * 1000003 + salt) & 0xFFFFFFFFFFFFFFFF, _lib.ptr(self._nonfinite),
                      _lib.stream_ptr())
            acc_total += accepts
            return logpsi is not None
        m = self.machine
        gen = torch.Generator(device=dev)
        gen.manual_seed((self.seed * 7919 + salt) & 0x7FFFFFFFFFFFFFFF)
        cur, _ = m.forward(self.sample)
        cur = cur.contiguous()
        cand = torch.empty_like(self.sample)
        accepted = torch.empty(C, dtype=torch.int32, device=dev)
        s = _lib.stream_ptr()
        for _ in range(n_prop):
            site = torch.randint(0, n, (C,), generator=gen, device=dev, dtype=torch.int32)
            move = torch.randint(0, 9, (C,), generator=gen, device=dev, dtype=torch.int32) if self.kind == 1 else None
            u = torch.rand((C,), generator=gen, device=dev, dtype=torch.float64)
            _lib.call('fk_mh_propose', _lib.ptr(self.sample), C, self.H, self.W, self.kind, _lib.ptr(site),
                      _lib.ptr(move), _lib.ptr(cand), s)
            lc, _ = m.forward(cand)
            _lib.call('fk_mh_accept', _lib.ptr(self.sample), _lib.ptr(cur), _lib.ptr(cand), _lib.ptr(lc),
                      _lib.ptr(u), C, n, _lib.ptr(accepted), _lib.ptr(self._nonfinite), s)
            acc_total += accepted
        out[:, 0] = self.sample
        return False

    def _draw(self, step, gen, dev):
        C, n = self.num_of_chains, self.sample.shape[1]
        if self._injected is not None:
            sites, moves, us = self._injected
            return sites[step], None if moves is None else moves[step], us[step]
        site = torch.randint(0, n, (C,), generator=gen, device=dev, dtype=torch.int32)
        move = torch.randint(0, 9, (C,), generator=gen, device=dev, dtype=torch.int32) if self.kind == 1 else None
        u = torch.rand((C,), generator=gen, device=dev, dtype=torch.float64)
        return site, move, u

    def _generic(self, out, burn, spc, steps):
        m = self.machine
        dev = m.device
        C, n = self.num_of_chains, self.sample.shape[1]
        gen = torch.Generator(device=dev)
        gen.manual_seed(self.seed * 7919 + self._calls)
        cur, _ = m.forward(self.sample)   # machine_updated (:39-40)
        cur = cur.contiguous()
        cand = torch.empty_like(self.sample)
        accepted = torch.empty(C, dtype=torch.int32, device=dev)
        acc_total = torch.zeros((), dtype=torch.int64, device=dev)
        s = _lib.stream_ptr()
        for step in range(steps):
            site, move, u = self._draw(step, gen, dev)
            _lib.call('fk_mh_propose', _lib.ptr(self.sample), C, self.H, self.W, self.kind, _lib.ptr(site),
                      _lib.ptr(move), _lib.ptr(cand), s)
            lc, _ = m.forward(cand)
            _lib.call('fk_mh_accept', _lib.ptr(self.sample), _lib.ptr(cur), _lib.ptr(cand), _lib.ptr(lc),
                      _lib.ptr(u), C, n, _lib.ptr(accepted), _lib.ptr(self._nonfinite), s)
            if step >= burn:
                acc_total += accepted.sum()
                rs = step - burn
                if (rs + 1) % (self.unused_sampels + 1) == 0:
                    out[:, rs // (self.unused_sampels + 1)] = self.sample
        return int(acc_total.item())

    def calc_r_hat_value(self, estimated_values):
        """Gelman-Rubin R-hat, variogram correlations and ESS
        (samplers/metropolis_hastings.py:67-89) of this sampler's batch."""
        return calc_r_hat_value(estimated_values, self.num_of_chains, self.batch_size)


def calc_r_hat_value(estimated_values, num_of_chains, batch_size=None):
    """Gelman-Rubin R-hat, variogram correlations and ESS of values laid out
    (chains, samples per chain) (samplers/metropolis_hastings.py:67-89):
    returns (r_hat, var, correlations_sum, effective_sample_size)."""
    v = np.asarray(estimated_values.cpu() if isinstance(estimated_values, torch.Tensor) else estimated_values)
    if batch_size is None:
        batch_size = v.size
    spc = batch_size // num_of_chains
    per = v.reshape(num_of_chains, spc)
    cm = per.mean(axis=1)
    mean = cm.mean()
    between = (spc / (num_of_chains - 1.0)) * ((cm - mean) ** 2).sum()
    if spc == 1:
        return 1.0, between, 0.0, batch_size
    within = ((per - cm[:, None]) ** 2).sum(axis=1).mean() / (spc - 1)
    var = ((spc - 1) * within + between) / spc
    r_hat = np.sqrt(var / within)
    vt = np.array([((per[:, t + 1:] - per[:, :spc - t - 1]) ** 2).mean(axis=1).mean() for t in range(spc - 1)])
    corr = 1 - vt / (2 * var)
    cs = sum_correlations(corr)
    return r_hat, var, cs, batch_size / (1 + 2 * cs)


class MetropolisHastingsLocal(MetropolisHastingsSampler):
    """Flip one uniformly chosen site per chain (:119-127)."""
    kind = 0


class MetropolisHastingsExchange(MetropolisHastingsSampler):
    """Swap a uniformly chosen site with a neighbour at offset in
    {-1, 0, 1}^2, periodic (:130-145)."""
    kind = 1
