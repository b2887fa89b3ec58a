"""Learned MADE proposal on the GPU: log q, sampling and the maximum-
likelihood fit step against the float64 oracle (same BF16-rounded weights),
the independent-proposal acceptance rule, and exactness of the Markov chain
that uses it."""
import numpy as np
import pytest
import torch

from oracle import exact as oex
from oracle import made as omd
from oracle import rbm as orbm
from tests.conftest import gpu_banner

pytestmark = pytest.mark.gpu


def _proposal(N, Hd, seed=0, scale=None):
    import flowket_amd as fk
    q = fk.MADEProposal(N, hidden=Hd, seed=seed)
    if scale is not None:
        # larger random weights so every layer matters
        g = torch.Generator().manual_seed(seed + 1)
        live = torch.as_tensor(np.concatenate([a.ravel() for a in _live(N, Hd)]), dtype=torch.float32)
        q.master.copy_((torch.randn(q.master.numel(), generator=g) * scale * live).to(q.master.device))
        q._refresh_bf16()
    return q


def _live(N, Hd):
    m1, m2, m3 = omd.masks(N, Hd)
    return [m1, np.ones(Hd), m2.T, np.ones(Hd), m3, np.ones(N)]


def _oracle_params(q):
    """The weights the kernels use: BF16 copies of the matrices, FP32 biases."""
    v = q.master.detach().cpu().double().numpy().copy()
    n1, n2 = q.Hd * q.N, q.Hd * q.Hd
    o2 = n1 + q.Hd
    o3 = o2 + n2 + q.Hd
    for lo, n in ((0, n1), (o2, n2), (o3, q.N * q.Hd)):
        v[lo:lo + n] = torch.as_tensor(v[lo:lo + n]).float().bfloat16().double().numpy()
    return omd.unpack(v, q.N, q.Hd)


@pytest.mark.parametrize('N,Hd', [(6, 16), (16, 64), (100, 256)])
def test_made_log_prob_vs_oracle(rng, N, Hd):
    gpu_banner(f'made log q N={N} Hd={Hd}')
    q = _proposal(N, Hd, seed=2, scale=0.3)
    x = rng.choice([-1, 1], size=(77, N)).astype(np.int8)
    got = q.log_prob(torch.as_tensor(x).cuda()).double().cpu().numpy()
    ref = omd.log_q(_oracle_params(q), x, hidden_bf16=True)
    # FP32 accumulation order may move a hidden value across a bf16 rounding
    # boundary: a few 1e-3 of a logit at most
    assert np.abs(got - ref).max() < 2e-3 * (np.abs(ref).max() + 1)


def test_made_normalised_on_gpu():
    gpu_banner('made sum_x q(x) == 1, N=10')
    N = 10
    q = _proposal(N, 32, seed=4, scale=0.5)
    states = torch.as_tensor(oex.all_states((N,)).reshape(-1, N).astype(np.int8)).cuda()
    total = torch.exp(q.log_prob(states).double()).sum().item()
    assert total == pytest.approx(1.0, abs=2e-5)


@pytest.mark.parametrize('N,Hd', [(6, 16), (40, 128)])
def test_made_sample_vs_oracle(rng, N, Hd):
    gpu_banner(f'made sample N={N} Hd={Hd}')
    q = _proposal(N, Hd, seed=6, scale=0.4)
    B = 300
    u = rng.random((B, N))
    x, lq = q.sample(B, uniforms=torch.as_tensor(u).cuda())
    x = x.cpu().numpy()
    params = _oracle_params(q)
    ref_x, ref_lq = omd.sample(params, u, hidden_bf16=True)
    same = np.all(x == ref_x, axis=1)
    # a draw may differ only where u lies within the logit tolerance of the
    # conditional (|d logit| <= 2e-3, so |d p| <= 5e-4) at the first site
    # whose draw differs; the conditionals along the oracle's own path
    # (site i depends on sites < i only, so the final samples give them all)
    z = omd.logits(params, ref_x, True)
    pz = 1.0 / (1.0 + np.exp(-z))
    for b in np.nonzero(~same)[0]:
        i = int(np.nonzero(x[b] != ref_x[b])[0][0])
        assert abs(pz[b, i] - u[b, i]) <= 1e-3, (b, i, pz[b, i], u[b, i])
    assert same.mean() > 0.95, same.mean()
    assert np.abs(lq.double().cpu().numpy()[same] - ref_lq[same]).max() < 2e-3 * (np.abs(ref_lq).max() + 1)
    # the returned log q is the log q of the returned sample, bit for bit
    # (sampling and log q share the arithmetic)
    assert torch.equal(lq, q.log_prob(torch.as_tensor(x).cuda()))


def test_made_sample_distribution():
    gpu_banner('made sampling z-test N=6')
    N = 6
    q = _proposal(N, 16, seed=8, scale=0.6)
    n = 2 ** 15
    x, _ = q.sample(n, generator=torch.Generator(device='cuda').manual_seed(3))
    states = oex.all_states((N,)).reshape(-1, N)
    p = np.exp(q.log_prob(torch.as_tensor(states.astype(np.int8)).cuda()).double().cpu().numpy())
    rng = np.random.default_rng(5)
    a = np.bincount(oex.binary_array_to_decimal_array(x.cpu().numpy()), minlength=2 ** N)
    b = np.bincount(rng.choice(2 ** N, size=n, p=p / p.sum()), minlength=2 ** N)
    z = ((np.square(a - b) - a - b) / (a + b + 1e-20)).sum()
    assert z <= np.sqrt(n)


@pytest.mark.parametrize('N,Hd,B', [(6, 16, 40), (30, 64, 257)])
def test_made_fit_step_gradient_vs_oracle(rng, N, Hd, B):
    """After one Adam step the first moment is (1 - beta1) * gradient: compare
    it with the oracle's backpropagated gradient of -mean log q."""
    gpu_banner(f'made fit gradient N={N} Hd={Hd} B={B}')
    q = _proposal(N, Hd, seed=10, scale=0.3)
    params = _oracle_params(q)
    x = rng.choice([-1, 1], size=(B, N)).astype(np.int8)
    q.fit(torch.as_tensor(x).cuda(), steps=1)
    got = q.adam_m.double().cpu().numpy() / 0.1
    ref = omd.pack(*omd.grad_mean_nll(params, x))
    assert np.abs(got - ref).max() < 1e-4 * np.abs(ref).max()
    # masked weights stay exactly zero
    live = np.concatenate([a.ravel() for a in _live(N, Hd)])
    assert np.all(q.master.cpu().numpy()[live == 0] == 0)


def test_made_fit_lowers_nll():
    """Maximum likelihood on samples of a fixed RBM |psi|^2 (6 spins) brings the
    proposal close to it."""
    gpu_banner('made fit to rbm distribution')
    N = 6
    rng = np.random.default_rng(12)
    w_re, w_im = rng.normal(scale=0.5, size=(N, 12)), np.zeros((N, 12))
    states = oex.all_states((N,)).reshape(-1, N)
    lp = orbm.log_psi(states, w_re, w_im, np.zeros(12), np.zeros(12), rng.normal(scale=0.3, size=N), np.zeros(N))
    p = np.exp(2 * lp.real)
    p /= p.sum()
    x = states[rng.choice(2 ** N, size=4096, p=p)].astype(np.int8)
    q = _proposal(N, 64, seed=0)
    xd = torch.as_tensor(x).cuda()
    before = float(-q.log_prob(xd).mean())
    after = q.fit(xd, steps=300, lr=1e-2)
    entropy = float(-(p * np.log(p)).sum())
    assert after < before - 0.2
    assert after < entropy + 0.1


def test_mh_accept_independent_rule(rng):
    gpu_banner('independent-proposal acceptance rule')
    import flowket_amd as fk
    from flowket_amd import _lib
    C, N = 1000, 12
    st = torch.as_tensor(rng.choice([-1, 1], size=(C, N)).astype(np.int8)).cuda()
    cand = torch.as_tensor(rng.choice([-1, 1], size=(C, N)).astype(np.int8)).cuda()
    lp_cur = torch.as_tensor(rng.normal(size=(C, 2)), dtype=torch.float32).cuda()
    lp_cand = torch.as_tensor(rng.normal(size=(C, 2)), dtype=torch.float32).cuda()
    lq_cur = torch.as_tensor(rng.normal(size=C) - 8, dtype=torch.float32).cuda()
    lq_cand = torch.as_tensor(rng.normal(size=C) - 8, dtype=torch.float32).cuda()
    u = torch.as_tensor(rng.random(C)).cuda()
    ref_st, ref_lp, ref_lq = st.clone(), lp_cur.clone(), lq_cur.clone()
    acc = torch.empty(C, dtype=torch.int32, device='cuda')
    flag = torch.zeros(1, dtype=torch.int32, device='cuda')
    _lib.call('fk_mh_accept_independent', _lib.ptr(st), _lib.ptr(lp_cur), _lib.ptr(lq_cur), _lib.ptr(cand),
              _lib.ptr(lp_cand), _lib.ptr(lq_cand), _lib.ptr(u), C, N, _lib.ptr(acc), _lib.ptr(flag),
              _lib.stream_ptr())
    assert int(flag.item()) == 0
    r = (2 * (lp_cand[:, 0] - ref_lp[:, 0]) + ref_lq - lq_cand).double().cpu().numpy()
    expect = np.exp(np.minimum(r, 0)) > u.cpu().numpy()
    a = acc.cpu().numpy().astype(bool)
    assert np.array_equal(a, expect)
    assert 0.1 < a.mean() < 0.9
    ex = torch.as_tensor(expect).cuda()
    assert torch.equal(st, torch.where(ex[:, None], cand, ref_st))
    assert torch.equal(lp_cur, torch.where(ex[:, None], lp_cand, ref_lp))
    assert torch.equal(lq_cur, torch.where(ex, lq_cand, ref_lq))
    del fk


def test_mh_learned_proposal_distribution_exact():
    """Chains that mix exchange sweeps with learned independent proposals
    still sample |psi|^2 exactly (z-statistic of tests/test_samplers.py:63-82)."""
    gpu_banner('mh learned proposal distribution 2x3')
    import flowket_amd as fk
    m = fk.RBM((2, 3), alpha=2, stddev=0.6, seed=9)
    n = 2 ** 14
    s = fk.MetropolisHastingsLearnedProposal(m, n, num_of_chains=n // 64, unused_sampels=1, discard_ratio=2,
                                             seed=11, base_kind=0, hidden=32, refit_every=3, fit_steps=100,
                                             proposal_lr=1e-2)
    next(s)
    x = next(s).cpu().numpy().reshape(n, -1)
    assert s.used_proposal and s.proposal_acceptance > 0.3
    w_re, w_im, b_re, b_im, a_re, a_im = [t.double().cpu().numpy() for t in m.parts()]
    states = oex.all_states((2, 3))
    lp = orbm.log_psi(states.reshape(64, -1), w_re.reshape(6, 12), w_im.reshape(6, 12), b_re, b_im, a_re, a_im)
    p = np.exp(2 * lp.real)
    p /= p.sum()
    rng = np.random.default_rng(2)
    a = np.bincount(oex.binary_array_to_decimal_array(x), minlength=64)
    b = np.bincount(rng.choice(64, size=n, p=p), minlength=64)
    z = ((np.square(a - b) - a - b) / (a + b + 1e-20)).sum()
    assert z <= np.sqrt(n)
