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
        self._repack()

    def _repack(self):
        """Degree-sorted masked copy that log q and sampling read."""
        if getattr(self, 'packed', None) is None:
            nb = int(_lib.load().fk_made_packed_bytes(self.N, self.Hd))
            self.packed = torch.empty(nb, dtype=torch.uint8, device=self.device)
        _lib.call('fk_made_pack', *self._wargs(), self.N, self.Hd, _lib.ptr(self.packed), _lib.stream_ptr())

    def _wargs(self):
        _, b1, _, b2, _, b3 = self._views()
        return [_lib.ptr(self.w1), _lib.ptr(b1), _lib.ptr(self.w2t), _lib.ptr(b2), _lib.ptr(self.w3), _lib.ptr(b3)]

    def log_prob(self, x):
        x = x.reshape(x.shape[0], self.N).to(torch.int8).contiguous()
        out = torch.empty(x.shape[0], dtype=torch.float32, device=self.device)
        _lib.call('fk_made_log_prob', _lib.ptr(self.packed), self.N, self.Hd, _lib.ptr(x), x.shape[0],
                  _lib.ptr(out), _lib.stream_ptr())
        return out

    def sample(self, n, uniforms=None, generator=None):
        if uniforms is None:
            uniforms = torch.rand((n, self.N), dtype=torch.float64, device=self.device, generator=generator)
        uniforms = uniforms.to(self.device).contiguous()
        x = torch.empty((n, self.N), dtype=torch.int8, device=self.device)
        lq = torch.empty(n, dtype=torch.float32, device=self.device)
        _lib.call('fk_made_sample', _lib.ptr(self.packed), self.N, self.Hd, _lib.ptr(uniforms), n, _lib.ptr(x),
                  _lib.ptr(lq), _lib.stream_ptr())
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
        self._repack()
        return float(-self.log_prob(x).mean().item())

    def fit_validated(self, x, max_steps=200, chunk=10, holdout=8, lr=None):
        """Adam on -mean log q over (holdout-1)/holdout of x, stopped when the
        NLL of the held-out rows stops improving (the best weights are kept).
        Returns the held-out NLL."""
        x = x.reshape(x.shape[0], self.N).to(torch.int8).contiguous()
        nv = max(1, x.shape[0] // holdout)
        xv, xt = x[:nv].contiguous(), x[nv:].contiguous()
        best = float(-self.log_prob(xv).mean().item())
        keep = (self.master.clone(), self.adam_m.clone(), self.adam_v.clone(), self.steps)
        done = 0
        while done < max_steps:
            self.fit(xt, steps=chunk, lr=lr)
            done += chunk
            nll = float(-self.log_prob(xv).mean().item())
            if nll >= best:
                break
            best = nll
            keep = (self.master.clone(), self.adam_m.clone(), self.adam_v.clone(), self.steps)
        self.master.copy_(keep[0])
        self.adam_m.copy_(keep[1])
        self.adam_v.copy_(keep[2])
        self.steps = keep[3]
        self._refresh_bf16()
        return best


class MetropolisHastingsLearnedProposal(MetropolisHastingsSampler):
    """Chains of ``base_kind`` (0 local flips, 1 exchange) with one
    independent learned-proposal step per recorded sample.  The proposal is
    refit every ``refit_every`` calls on the samples the chains produced; while
    its acceptance is below ``min_acceptance`` (or before the first fit) only
    the plain chain runs."""

    def __init__(self, machine, batch_size, num_of_chains=1, unused_sampels=0, discard_ratio=10, seed=None,
                 base_kind=1, hidden=256, refit_every=20, fit_steps=200, min_acceptance=0.05, proposal_lr=1e-3,
                 fit_batch=8192, **kwargs):
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
            pick = torch.randperm(self.batch_size, device=flat.device, generator=self._gen)[:self.fit_batch]
            self.last_nll = self.proposal.fit_validated(flat[pick], max_steps=self.fit_steps)
            self.proposal_fitted = True
            # acceptance of one proposal step from the current chain states,
            # measured without moving the chains
            self.proposal_acceptance = self._probe()
        self._outer_calls += 1
        return batch

    def _independent_step(self, logq_cur, logpsi_cur, cand=None, lq=None, lc=None, u=None, accepted=None):
        C, n = self.num_of_chains, self.sample.shape[1]
        m = self.machine
        if cand is None:
            cand, lq = self.proposal.sample(C, generator=self._gen)
            lc, _ = m.forward(cand)
            u = torch.rand(C, dtype=torch.float64, device=m.device, generator=self._gen)
        if accepted is None:
            accepted = torch.empty(C, dtype=torch.int32, device=m.device)
        _lib.call('fk_mh_accept_independent', _lib.ptr(self.sample), _lib.ptr(logpsi_cur), _lib.ptr(logq_cur),
                  _lib.ptr(cand), _lib.ptr(lc), _lib.ptr(lq), _lib.ptr(u), C, n, _lib.ptr(accepted),
                  _lib.ptr(self._nonfinite), _lib.stream_ptr())
        return accepted

    def _probe(self):
        m = self.machine
        self._nonfinite_flag()
        lp, _ = m.forward(self.sample)
        lp = lp.contiguous()
        lq = self.proposal.log_prob(self.sample)
        saved = self.sample.clone()
        acc = self._independent_step(lq, lp)
        self.sample.copy_(saved)
        return float(acc.float().mean().item())

    def _with_proposal(self):
        """Per recorded sample: one independent proposal from the MADE, then
        (unused_sampels + 1) plain proposals.  All candidates of the call are
        drawn (and their log psi evaluated) in one launch each; the loop has
        no host synchronisation."""
        spc, _ = self._counts()
        C, n = self.num_of_chains, self.sample.shape[1]
        m = self.machine
        dev = m.device
        self._nonfinite_flag()
        out = torch.empty((C, spc, n), dtype=torch.int8, device=dev)
        cand, lqc = self.proposal.sample(C * spc, generator=self._gen)
        lcand, _ = m.forward(cand)
        cand = cand.reshape(spc, C, n)
        lqc = lqc.reshape(spc, C)
        lcand = lcand.reshape(spc, C, 2)
        us = torch.rand((spc, C), dtype=torch.float64, device=dev, generator=self._gen)
        acc_prop = torch.zeros(C, dtype=torch.int64, device=dev)
        acc_loc = torch.zeros(C, dtype=torch.int64, device=dev)
        accepted = torch.empty(C, dtype=torch.int32, device=dev)
        sweep_out = torch.empty((C, 1, n), dtype=torch.int8, device=dev)
        lp = None
        for i in range(spc):
            if lp is None:
                lp, _ = m.forward(self.sample)
                lp = lp.contiguous()
            lq = self.proposal.log_prob(self.sample)
            self._independent_step(lq, lp, cand[i], lqc[i], lcand[i].contiguous(), us[i], accepted)
            acc_prop += accepted
            lp_next = torch.empty((C, 2), dtype=torch.float32, device=dev)
            if not self._plain_sweeps(sweep_out, self.unused_sampels + 1, acc_loc, self._calls * 1000003 + i,
                                      logpsi=lp_next):
                lp_next = None
            lp = lp_next
            out[:, i] = sweep_out[:, 0]
        self._calls += 1
        self._raise_if_nonfinite()
        self.proposal_acceptance = float(acc_prop.sum().item()) / (C * spc)
        self.acceptance_ratio = float(acc_loc.sum().item()) / (C * spc * (self.unused_sampels + 1))
        return out.reshape((self.batch_size,) + self.input_size)
