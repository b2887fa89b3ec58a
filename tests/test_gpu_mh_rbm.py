"""RBM log psi / local energies and Metropolis-Hastings chains on the GPU vs
the oracle (draw-for-draw with injected randomness) and vs exact |psi|^2."""
import numpy as np
import pytest
import torch

from oracle import exact as oex
from oracle import observables as oob
from oracle import operators as oop
from oracle import rbm as orbm
from oracle import samplers as osa
from tests.conftest import gpu_banner

pytestmark = pytest.mark.gpu


def _rbm(shape, alpha, stddev, seed=0):
    import flowket_amd as fk
    return fk.RBM(shape, alpha=alpha, stddev=stddev, seed=seed)


def _oracle_lp(m, spins):
    w_re, w_im, b_re, b_im, a_re, a_im = [t.double().cpu().numpy() for t in m.parts()]
    N, M = m.n_visible, m.n_hidden
    return orbm.log_psi(np.asarray(spins).reshape(len(spins), -1), w_re.reshape(N, M), w_im.reshape(N, M), b_re,
                        b_im, a_re, a_im)


@pytest.mark.parametrize('shape,alpha,stddev', [((8, 8), 4, 0.01), ((4, 4), 2, 0.3), ((3, 5), 1, 1.0)])
def test_rbm_log_psi(rng, shape, alpha, stddev):
    gpu_banner(f'rbm log psi {shape} alpha={alpha} sd={stddev}')
    m = _rbm(shape, alpha, stddev)
    s = rng.choice([-1, 1], size=(50,) + shape).astype(np.int8)
    got = m.log_psi(s).cpu().numpy().astype(np.complex128)
    ref = _oracle_lp(m, s)
    scale = np.abs(ref.real).max() + 1
    assert np.abs(got.real - ref.real).max() / scale < 1e-5
    # the phase is defined modulo 2 pi per hidden unit
    d = np.angle(np.exp(1j * (got.imag - ref.imag)))
    assert np.abs(d).max() < 1e-4


def test_rbm_local_energy(rng):
    import flowket_amd as fk
    gpu_banner('rbm local energy 4x4 heisenberg')
    m = _rbm((4, 4), 2, 0.2)
    op = fk.Heisenberg(hilbert_state_shape=(4, 4), pbc=True)
    s = rng.choice([-1, 1], size=(40, 4, 4)).astype(np.int8)
    got = fk.Observable(op).local_values(m, s).cpu().numpy()
    conn, mel, use = oop.heisenberg_find_conn(s.astype(np.float64), pbc=True)
    ref = oob.local_values_unbalanced(lambda x: _oracle_lp(m, x).astype(np.complex64), conn, mel, use)
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-4


# Tolerance of an acceptance decision: log psi is FP32 on the GPU (1e-5
# relative), so a decision may flip only where u lies within the propagated
# error of the acceptance probability exp(min(2 dRe, 0)).
LOGPSI_RTOL = 1e-5


def _oracle_chain(lp_fn, state, sites, moves, us, kind, shape, n_burn, spc, unused):
    """The reference chain on the oracle's log psi; also returns, per chain,
    the first step at which |u - p_accept| fell within the FP32 tolerance
    (n_steps if never)."""
    cur = state.astype(np.float64).copy()
    lcur = lp_fn(cur)
    C = cur.shape[0]
    out = np.zeros((C, spc) + tuple(shape))
    n_steps = n_burn + spc * (unused + 1)
    first_tie = np.full(C, n_steps)
    acc = 0
    H, W = shape
    for step in range(n_burn + spc * (unused + 1)):
        if kind == 0:
            cand = osa.mh_local_candidates(cur, sites[step])
        else:
            pos = np.stack([sites[step] // W, sites[step] % W], axis=1)
            mv = np.stack([moves[step] // 3 - 1, moves[step] % 3 - 1], axis=1)
            cand = osa.mh_exchange_candidates(cur, pos, mv)
        lc = lp_fn(cand)
        a = osa.mh_accept(lc, lcur, us[step])
        p = np.exp(np.minimum(2.0 * (lc.real - lcur.real), 0.0))
        tol = 2.0 * p * LOGPSI_RTOL * (np.abs(lc.real) + np.abs(lcur.real) + 1.0)
        tie = np.abs(us[step] - p) <= tol
        first_tie = np.where(tie & (first_tie == n_steps), step, first_tie)
        cur[a] = cand[a]
        lcur[a] = lc[a]
        if step >= n_burn:
            acc += a.sum()
            rs = step - n_burn
            if (rs + 1) % (unused + 1) == 0:
                out[:, rs // (unused + 1)] = cur
    return out, acc, first_tie


def _check_divergence(got, ref, first_tie, n_burn, unused):
    """Chains must be identical to the oracle's up to the first near-tie
    acceptance decision: every recorded sample taken before that step
    matches exactly.  Returns the fraction of chains with no near tie."""
    C, spc = got.shape[:2]
    for c in range(C):
        for r in range(spc):
            rec_step = n_burn + (r + 1) * (unused + 1) - 1
            if rec_step < first_tie[c]:
                assert np.array_equal(got[c, r], ref[c, r]), (c, r, first_tie[c])
    n_steps = n_burn + spc * (unused + 1)
    return float((first_tie == n_steps).mean())


@pytest.mark.parametrize('kind', [0, 1])
def test_mh_rbm_chains_match_oracle_rule(rng, kind):
    import flowket_amd as fk
    gpu_banner(f'mh rbm kind={kind} vs oracle')
    shape = (4, 4)
    m = _rbm(shape, 2, 0.3, seed=5)
    C, spc, unused, discard = 64, 8, 3, 4
    cls = fk.MetropolisHastingsLocal if kind == 0 else fk.MetropolisHastingsExchange
    sampler = cls(m, C * spc, num_of_chains=C, unused_sampels=unused, discard_ratio=discard, seed=3)
    spc_, burn = sampler._counts()
    steps = burn + spc * (unused + 1)
    sites = rng.integers(0, 16, size=(steps, C)).astype(np.int32)
    moves = rng.integers(0, 9, size=(steps, C)).astype(np.int32)
    us = rng.random((steps, C))
    sampler.randomness = (sites, moves if kind == 1 else None, us)
    start = sampler.sample.cpu().numpy().reshape((C,) + shape)
    got = next(sampler).cpu().numpy().reshape(C, spc, *shape)
    ref, acc, first_tie = _oracle_chain(lambda x: _oracle_lp(m, x), start, sites, moves, us, kind, shape, burn, spc,
                                        unused)
    clean = _check_divergence(got, ref, first_tie, burn, unused)
    assert clean > 0.9, clean
    assert abs(sampler.acceptance_ratio - acc / (C * spc * (unused + 1))) < 0.02


def test_mh_generic_machine_matches_oracle_rule(rng):
    """MH-Local around a ConvNetAutoregressive2D (fused forward per sweep)."""
    import flowket_amd as fk
    from oracle import convnet as ocn
    from tests.test_gpu_convnet import _machine, _oracle_params
    gpu_banner('mh generic convnet vs oracle')
    shape = (4, 3)
    m = _machine(shape, 2, 16, 'relu', 'fp32', bias_scale=0.3)
    C, spc, unused = 32, 4, 2
    sampler = fk.MetropolisHastingsLocal(m, C * spc, num_of_chains=C, unused_sampels=unused, discard_ratio=2, seed=1)
    _, burn = sampler._counts()
    steps = burn + spc * (unused + 1)
    sites = rng.integers(0, 12, size=(steps, C)).astype(np.int32)
    us = rng.random((steps, C))
    sampler.randomness = (sites, None, us)
    start = sampler.sample.cpu().numpy().reshape((C,) + shape)
    got = next(sampler).cpu().numpy().reshape(C, spc, *shape)
    params = _oracle_params(m)
    ref, _, first_tie = _oracle_chain(lambda x: ocn.log_psi(params, x, 2, activation='relu'), start, sites, None, us,
                                      0, shape, burn, spc, unused)
    clean = _check_divergence(got, ref, first_tie, burn, unused)
    assert clean > 0.9, clean


def test_mh_rbm_distribution_exact(rng):
    """Long MH-Local chains on 2x3 spins reproduce |psi|^2
    (tests/test_samplers.py:49-82 z-statistic)."""
    import flowket_amd as fk
    gpu_banner('mh rbm distribution 2x3')
    m = _rbm((2, 3), 2, 0.6, seed=9)
    n = 2 ** 14
    sampler = fk.MetropolisHastingsLocal(m, n, num_of_chains=n // 64, unused_sampels=12, discard_ratio=2, seed=11)
    x = next(sampler).cpu().numpy().reshape(n, -1)
    idx = oex.binary_array_to_decimal_array(x)
    states = oex.all_states((2, 3))
    p = np.exp(2 * _oracle_lp(m, states).real)
    p /= p.sum()
    exact = rng.choice(64, size=n, p=p)
    a = np.bincount(idx, minlength=64)
    b = np.bincount(exact, minlength=64)
    z = ((np.square(a - b) - a - b) / (a + b + 1e-20)).sum()
    assert z <= np.sqrt(n)


def test_mh_nonfinite_candidates_raise(rng):
    """A candidate with a non-finite log psi is never accepted and the
    sampler raises, as MetropolisHastingsSymmetricProposal._sweep does
    (samplers/metropolis_hastings.py:103-105): the acceptance kernel on
    injected values, the generic chain and the fused RBM chain."""
    import flowket_amd as fk
    from flowket_amd import _lib
    gpu_banner('mh non-finite candidates')
    C, N = 256, 12
    st = torch.as_tensor(rng.choice([-1, 1], size=(C, N)).astype(np.int8)).cuda()
    cand = -st
    lp_cur = torch.zeros((C, 2), dtype=torch.float32, device='cuda')
    lp_cand = torch.full((C, 2), 5.0, dtype=torch.float32, device='cuda')
    lp_cand[3, 0] = float('nan')
    lp_cand[9, 1] = float('inf')
    u = torch.full((C,), 0.5, dtype=torch.float64, device='cuda')
    acc = torch.empty(C, dtype=torch.int32, device='cuda')
    flag = torch.zeros(1, dtype=torch.int32, device='cuda')
    _lib.call('fk_mh_accept', _lib.ptr(st), _lib.ptr(lp_cur), _lib.ptr(cand), _lib.ptr(lp_cand), _lib.ptr(u), C, N,
              _lib.ptr(acc), _lib.ptr(flag), _lib.stream_ptr())
    a = acc.cpu().numpy()
    assert int(flag.item()) == 1
    assert a[3] == 0 and a[9] == 0 and a.sum() == C - 2
    # generic path: a network whose parameters are NaN
    from tests.test_gpu_convnet import _machine
    m = _machine((3, 4), 2, 16, 'relu', 'fp32')
    p = m.params.clone()
    p[:] = float('nan')
    m.set_params(p)
    with pytest.raises(Exception, match='not finite'):
        next(fk.MetropolisHastingsLocal(m, 64, num_of_chains=16, unused_sampels=1, seed=1))
    # fused RBM chain
    r = _rbm((3, 4), 2, 0.1)
    q = r.params.clone()
    q[:] = float('inf')
    r.set_params(q)
    with pytest.raises(Exception, match='not finite'):
        next(fk.MetropolisHastingsLocal(r, 64, num_of_chains=16, unused_sampels=1, seed=1))


def test_mh_injected_randomness_checked():
    """Injected arrays that do not cover every sweep and chain, or index
    outside the lattice, are refused on the host."""
    import flowket_amd as fk
    gpu_banner('mh injected randomness checks')
    m = _rbm((4, 4), 1, 0.1)
    sampler = fk.MetropolisHastingsLocal(m, 64, num_of_chains=16, unused_sampels=1, discard_ratio=2, seed=3)
    spc, burn = sampler._counts()
    steps = burn + spc * 2
    sampler.randomness = (np.zeros((steps - 1, 16), np.int32), None, np.zeros((steps - 1, 16)))
    with pytest.raises(ValueError):
        next(sampler)
    sampler.randomness = (np.full((steps, 16), 16, np.int32), None, np.zeros((steps, 16)))
    with pytest.raises(ValueError):
        next(sampler)
