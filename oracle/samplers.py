"""Sampler rules restated from the reference, with injected randomness so the
GPU path can be compared draw for draw."""
import numpy as np


def ar_sample(cond_log_probs_fn, uniforms):
    """``AutoregressiveSampler.__next__`` (deepar/samplers/autoregressive.py:29-48)
    with zero_base=False (samplers/__init__.py:8-14): raster order, unsampled
    sites are 0 during the forwards (:30), s = +1 iff exp(logp[...,0]) > u
    (strict, :42-44).  ``uniforms`` has shape (B, H, W)."""
    B, H, W = uniforms.shape
    batch = np.zeros((B, H, W))
    for h in range(H):
        for w in range(W):
            lp = cond_log_probs_fn(batch)
            up = np.exp(lp[:, h, w, 0]) > uniforms[:, h, w]
            batch[:, h, w] = 2.0 * up - 1.0
    return batch


def mh_local_candidates(sample, sites):
    """``MetropolisHastingsLocal._next_candidates`` (samplers/metropolis_hastings
    .py:119-127): flip one site per chain; ``sites`` are flat site indices."""
    cand = sample.reshape(sample.shape[0], -1).copy()
    cand[np.arange(cand.shape[0]), sites] *= -1
    return cand.reshape(sample.shape)


def mh_exchange_candidates(sample, pos, move):
    """``MetropolisHastingsExchange._next_candidates`` (:130-145): per chain a
    site ``pos`` (per-dim indices, shape (chains, ndim)) and an offset ``move``
    in {-1,0,1}^ndim; swap the two sites (mod size)."""
    shape = sample.shape[1:]
    cand = sample.copy()
    n = sample.shape[0]
    for c in range(n):
        a = tuple(pos[c])
        b = tuple((pos[c][d] + move[c][d]) % shape[d] for d in range(len(shape)))
        cand[(c,) + a], cand[(c,) + b] = sample[(c,) + b], sample[(c,) + a]
    return cand


def mh_accept(logpsi_cand, logpsi_cur, u):
    """``MetropolisHastingsSymmetricProposal._sweep`` acceptance (:99-113):
    accept iff exp(min(2*Re(logpsi' - logpsi), 0)) > U."""
    log_ratio = np.minimum(2.0 * (np.real(logpsi_cand) - np.real(logpsi_cur)), 0.0)
    return np.exp(log_ratio) > u


def sum_correlations(correlations):
    """Geyer pair truncation (metropolis_hastings.py:8-12)."""
    for i in range(1, correlations.shape[0] // 2):
        if correlations[2 * i - 1] + correlations[2 * i] < 0:
            return correlations[:2 * i].sum()
    return correlations.sum()


def r_hat_and_ess(values, num_of_chains):
    """``calc_r_hat_value`` (metropolis_hastings.py:67-89); ``values`` laid out
    (chains, samples_per_chain) flattened."""
    batch = values.shape[0]
    spc = batch // num_of_chains
    v = values.reshape(num_of_chains, spc)
    cm = v.mean(axis=1)
    mean = cm.mean()
    between = (spc / (num_of_chains - 1.0)) * ((cm - mean) ** 2).sum()
    if spc == 1:
        return 1.0, between, 0.0, batch
    within = ((v - cm[:, None]) ** 2).sum(axis=1).mean() / (spc - 1)
    var = ((spc - 1) * within + between) / spc
    r_hat = np.sqrt(var / within)
    vt = np.zeros(spc - 1)
    for t in range(spc - 1):
        d = v[:, t + 1:] - v[:, :spc - t - 1]
        vt[t] = (d * d).mean(axis=1).mean()
    corr = 1 - vt / (2 * var)
    cs = sum_correlations(corr)
    return r_hat, var, cs, batch / (1 + 2 * cs)
