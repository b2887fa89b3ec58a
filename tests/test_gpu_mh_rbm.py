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


def _oracle_chain(lp_fn, state, sites, moves, us, kind, shape, n_burn, spc, unused):
    cur = state.astype(np.float64).copy()
    lcur = lp_fn(cur)
    C = cur.shape[0]
    out = np.zeros((C, spc) + tuple(shape))
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
        cur[a] = cand[a]
        lcur[a] = lc[a]
        if step >= n_burn:
            acc += a.sum()
            rs = step - n_burn
            if (rs + 1) % (unused + 1) == 0:
                out[:, rs // (unused + 1)] = cur
    return out, acc


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
    ref, acc = _oracle_chain(lambda x: _oracle_lp(m, x), start, sites, moves, us, kind, shape, burn, spc, unused)
    same = np.all(got.reshape(C, -1) == ref.reshape(C, -1), axis=1)
    assert same.mean() > 0.95, same.mean()
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
    ref, _ = _oracle_chain(lambda x: ocn.log_psi(params, x, 2, activation='relu'), start, sites, None, us, 0, shape,
                           burn, spc, unused)
    same = np.all(got.reshape(C, -1) == ref.reshape(C, -1), axis=1)
    assert same.mean() > 0.9, same.mean()


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
