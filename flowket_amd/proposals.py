"""Learned autoregressive proposal for the Markov-chain path.

A masked MLP (MADE) q(x) = prod_i q(x_i | x_<i) over the N spins, two hidden
layers (width 256 by default), BF16 weights with FP32 arithmetic, trained by
maximum likelihood on the chains' own samples (fk_made_fit_step).  It is used
as an *independent* Metropolis-Hastings proposal with the exact ratio
    min(1, |psi(x')|^2 q(x) / (|psi(x)|^2 q(x'))),
so the chain stays unbiased; the plain local / exchange chain runs between
proposals and takes over completely while the acceptance of the proposal is
below ``min_acceptance``.  This has no counterpart in the reference (its
MetropolisHastingsGlobal, samplers/metropolis_hastings.py:148-157, treats
every proposal as symmetric).
"""
import numpy as np
import torch

from . import _lib
from .samplers import MetropolisHastingsSampler, _machine_of


class MADEProposal:
    def __init__(self, n_sites, hidden=256, seed=0, device=None, lr=1e-3):
        self.N = int(n_sites)
        self.Hd = int(hidden)
        if device is None:
            device = torch.device('cuda', torch.cuda.current_device())
        self.device = torch.device(device)
        N, Hd = self.N, self.Hd
        self.sizes = [Hd * N, Hd, Hd * Hd, Hd, N * Hd, N]
        g = torch.Generator().manual_seed(int(seed))
        deg_h = 1 + torch.arange(Hd) % (N - 1)
        deg_in = torch.arange(N) + 1
        m1 = (deg_in[None, :] <= deg_h[:, None]).float()            # (Hd, N)
        m2t = (deg_h[:, None] <= deg_h[None, :]).float()            # [k][t]: deg k <= deg t
        m3 = (deg_h[None, :] <= torch.arange(N)[:, None]).float()   # (N, Hd)

        def glorot(shape, fan_in, fan_out):
            lim = float(np.sqrt(6.0 / (fan_in + fan_out)))
            return (torch.rand(shape, generator=g) * 2 - 1) * lim

        parts = [glorot((Hd, N), N, Hd) * m1, torch.zeros(Hd), glorot((Hd, Hd), Hd, Hd) * m2t, torch.zeros(Hd),
                 glorot((N, Hd), Hd, N) * m3, torch.zeros(N)]
        self.master = torch.cat([p.reshape(-1) for p in parts]).float().to(self.device).contiguous()
        self.adam_m = torch.zeros_like(self.master)
        self.adam_v = torch.zeros_like(self.master)
        self.w1 = torch.empty(Hd * N, dtype=torch.bfloat16, device=self.device)
        self.w2t = torch.empty(Hd * Hd, dtype=torch.bfloat16, device=self.device)
        self.w3 = torch.empty(N * Hd, dtype=torch.bfloat16, device=self.device)
        self._refresh_bf16()
        self.lr = lr
        self.steps = 0

    def _views(self):
        out, off = [], 0
        for n in self.sizes:
            out.append(self.master[off:off + n])
            off += n
        return out

    def _refresh_bf16(self):
        w1, _, w2t, _, w3, _ = self._views()
        self.w1.copy_(w1)
        self.w2t.copy_(w2t)
        self.w3.copy_(w3)

    def _wargs(self):
        _, b1, _, b2, _, b3 = self._views()
        return [_lib.ptr(self.w1), _lib.ptr(b1), _lib.ptr(self.w2t), _lib.ptr(b2), _lib.ptr(self.w3), _lib.ptr(b3)]

    def log_prob(self, x):
        x = x.reshape(x.shape[0], self.N).to(torch.int8).contiguous()
        out = torch.empty(x.shape[0], dtype=torch.float32, device=self.device)
        _lib.call('fk_made_log_prob', *self._wargs(), self.N, self.Hd, _lib.ptr(x), x.shape[0], _lib.ptr(out),
                  _lib.stream_ptr())
        return out

    def sample(self, n, uniforms=None, generator=None):
        if uniforms is None:
            uniforms = torch.rand((n, self.N), dtype=torch.float64, device=self.device, generator=generator)
        uniforms = uniforms.to(self.device).contiguous()
        x = torch.empty((n, self.N), dtype=torch.int8, device=self.device)
        lq = torch.empty(n, dtype=torch.float32, device=self.device)
        _lib.call('fk_made_sample', *self._wargs(), self.N, self.Hd, _lib.ptr(uniforms), n, _lib.ptr(x), _lib.ptr(lq),
                  _lib.stream_ptr())
        return x, lq

    def fit(self, x, steps=50, lr=None):
        """Full-batch Adam on -mean log q(x); returns the final mean NLL."""
        x = x.reshape(x.shape[0], self.N).to(torch.int8).contiguous()
        B = x.shape[0]
        ws = torch.empty(int(_lib.load().fk_made_fit_workspace_floats(self.N, self.Hd, B)), dtype=torch.float32,
                         device=self.device)
        _, b1, _, b2, _, b3 = self._views()
        for _ in range(steps):
            self.steps += 1
            _lib.call('fk_made_fit_step', _lib.ptr(self.master), _lib.ptr(self.adam_m), _lib.ptr(self.adam_v),
                      _lib.ptr(self.w1), _lib.ptr(b1), _lib.ptr(self.w2t), _lib.ptr(b2), _lib.ptr(self.w3),
                      _lib.ptr(b3), self.N, self.Hd, _lib.ptr(x), B, _lib.ptr(ws), float(lr or self.lr), 0.9, 0.999,
                      1e-7, self.steps, _lib.stream_ptr())
        return float(-self.log_prob(x).mean().item())


class MetropolisHastingsLearnedProposal(MetropolisHastingsSampler):
    """Chains of ``base_kind`` (0 local flips, 1 exchange) with one
    independent learned-proposal step per recorded sample.  The proposal is
    refit every ``refit_every`` calls on the samples the chains produced; while
    its acceptance is below ``min_acceptance`` (or before the first fit) only
    the plain chain runs."""

    def __init__(self, machine, batch_size, num_of_chains=1, unused_sampels=0, discard_ratio=10, seed=None,
                 base_kind=1, hidden=256, refit_every=20, fit_steps=100, min_acceptance=0.05, proposal_lr=1e-3,
                 fit_batch=2048, **kwargs):
        super().__init__(machine, batch_size, num_of_chains=num_of_chains, unused_sampels=unused_sampels,
                         discard_ratio=discard_ratio, seed=seed, **kwargs)
        self.kind = base_kind
        n = self.sample.shape[1]
        self.proposal = MADEProposal(n, hidden=hidden, seed=self.seed % (2 ** 31), device=self.machine.device,
                                     lr=proposal_lr)
        self.refit_every = refit_every
        self.fit_steps = fit_steps
        self.fit_batch = fit_batch
        self.min_acceptance = min_acceptance
        self.proposal_fitted = False
        self.proposal_acceptance = 0.0
        self.last_nll = None
        self.used_proposal = False
        self._outer_calls = 0
        self._gen = torch.Generator(device=self.machine.device)
        self._gen.manual_seed(self.seed * 31 + 7)

    def _plain(self):
        return MetropolisHastingsSampler.__next__(self)

    def __next__(self):
        self.used_proposal = self.proposal_fitted and self.proposal_acceptance >= self.min_acceptance
        if self.used_proposal:
            batch = self._with_proposal()
        else:
            batch = self._plain()
        if self._outer_calls % self.refit_every == 0:
            flat = batch.reshape(self.batch_size, -1)
            if self.batch_size > self.fit_batch:
                pick = torch.randperm(self.batch_size, device=flat.device, generator=self._gen)[:self.fit_batch]
                flat = flat[pick]
            self.last_nll = self.proposal.fit(flat, steps=self.fit_steps)
            self.proposal_fitted = True
            # acceptance of one proposal step from the current chain states,
            # measured without moving the chains
            self.proposal_acceptance = self._probe()
        self._outer_calls += 1
        return batch

    def _independent_step(self, logq_cur, logpsi_cur):
        C, n = self.num_of_chains, self.sample.shape[1]
        m = self.machine
        cand, lq = self.proposal.sample(C, generator=self._gen)
        lc, _ = m.forward(cand)
        u = torch.rand(C, dtype=torch.float64, device=m.device, generator=self._gen)
        accepted = torch.empty(C, dtype=torch.int32, device=m.device)
        _lib.call('fk_mh_accept_independent', _lib.ptr(self.sample), _lib.ptr(logpsi_cur), _lib.ptr(logq_cur),
                  _lib.ptr(cand), _lib.ptr(lc.contiguous()), _lib.ptr(lq), _lib.ptr(u), C, n, _lib.ptr(accepted),
                  _lib.stream_ptr())
        return accepted

    def _probe(self):
        m = self.machine
        lp, _ = m.forward(self.sample)
        lp = lp.contiguous()
        lq = self.proposal.log_prob(self.sample)
        saved = self.sample.clone()
        acc = self._independent_step(lq, lp)
        self.sample.copy_(saved)
        return float(acc.float().mean().item())

    def _with_proposal(self):
        spc, burn = self._counts()
        C, n = self.num_of_chains, self.sample.shape[1]
        m = self.machine
        out = torch.empty((C, spc, n), dtype=torch.int8, device=m.device)
        saved_counts = (self.batch_size, self.unused_sampels, self.discard_ratio)
        acc_prop = 0
        acc_local = 0.0
        for i in range(spc):
            lp, _ = m.forward(self.sample)
            lp = lp.contiguous()
            lq = self.proposal.log_prob(self.sample)
            acc_prop += int(self._independent_step(lq, lp).sum().item())
            # plain chain sweeps between records: one recorded sample per chain
            self.batch_size, self.discard_ratio = C, 0
            rec = MetropolisHastingsSampler.__next__(self)
            acc_local += self.acceptance_ratio
            self.batch_size, self.unused_sampels, self.discard_ratio = saved_counts
            out[:, i] = rec.reshape(C, n)
        self.proposal_acceptance = acc_prop / (C * spc)
        self.acceptance_ratio = acc_local / spc
        return out.reshape((self.batch_size,) + self.input_size)
