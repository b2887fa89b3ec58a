"""Pins for the CPU oracle: published/closed-form energies, the reference's own
fixtures and invariants, and independent restatements.  CPU only."""
import numpy as np
import pytest

from oracle import convnet as ocn
from oracle import exact as oex
from oracle import observables as oob
from oracle import operators as oop
from oracle import rbm as orbm
from oracle import samplers as osa
from oracle import sr as osr


# ---------------------------------------------------------------- operators
def test_heisenberg_4x4_pbc_ground_energy():
    """ED of the oracle's matrix elements; -44.9139328337 (Pauli convention,
    = 4x the spin-1/2 value -11.2284832084 of the 4x4 periodic lattice)."""
    e = oex.ground_energy(lambda s: oop.heisenberg_find_conn(s, pbc=True), (4, 4))
    assert e == pytest.approx(-44.9139328337, abs=1e-8)


def _golden():
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), 'golden', 'reference_fixtures.json')) as f:
        return json.load(f)


@pytest.mark.parametrize('k', range(3))
def test_ground_energies_published(k):
    """Ground energies the reference's examples print as the target
    (examples/basic_autoregressive_2d.py:39,
    examples/basic_autoregressive_monte_carlo_gradient.py:41,
    examples/j1j2_2d_monte_carlo_4.py:43), by ED of the oracle's matrix
    elements."""
    g = _golden()['ground_energies'][k]
    shape = tuple(g['shape'])
    if g['model'] == 'ising':
        fc = lambda s: oop.ising_find_conn(s, h=g['h'], j=g['j'], pbc=g['pbc'])  # noqa: E731
    else:
        fc = lambda s: oop.j1j2_find_conn(s, j2=g['j2'], pbc=g['pbc'])  # noqa: E731
    assert oex.ground_energy(fc, shape) == pytest.approx(g['energy'], abs=1e-9)


def test_heisenberg_chain_bethe_limit():
    """1-D periodic chain, 10 sites: E0 = -4*4.515446354 (spin-1/2 value
    -4.515446354 from exact diagonalisation literature)."""
    e = oex.ground_energy(lambda s: oop.heisenberg_find_conn(s, pbc=True), (10,))
    assert e == pytest.approx(-4 * 4.515446354, abs=1e-7)


def test_heisenberg_invariants(rng):
    s = rng.choice([-1, 1], size=(50, 5, 6)).astype(np.float64)
    conn, mel, use = oop.heisenberg_find_conn(s, pbc=True)
    assert conn.shape[0] == 2 * 30 + 1
    assert np.array_equal(conn[0], s)
    for c in range(1, conn.shape[0]):
        diff = (conn[c] != s).reshape(50, -1).sum(axis=1)
        assert np.all(diff[use[c]] == 2) and np.all(diff[~use[c]] == 0)
    # degree-4 periodic lattice: even number of antiparallel bonds -> steps of 4
    assert np.all((mel[0] - 60) % 4 == 0)
    assert np.all(mel[1:][use[1:]] == -2.0)


def test_ising_off_diagonal_and_flip(rng):
    s = rng.choice([-1, 1], size=(7, 3, 4)).astype(np.float64)
    conn, mel, use = oop.ising_find_conn(s, h=2.5, pbc=True)
    assert np.all(use) and np.all(mel[1:] == -2.5)
    for q in range(12):
        assert np.array_equal(conn[1 + q].reshape(7, -1)[:, q], -s.reshape(7, -1)[:, q])


def test_j1j2_edge_count():
    assert len(oop.j1j2_edges(10, 10, True)) == 400
    e = oop.j1j2_edges(10, 10, True)
    assert sum(1 for x in e if x[2] == 1) == 200 and sum(1 for x in e if x[2] == 2) == 200
    # every site has 4 NN and 4 NNN neighbours on the torus
    deg = np.zeros(100, int)
    for a, b, _ in e:
        deg[a] += 1
        deg[b] += 1
    assert np.all(deg == 8)


# ---------------------------------------------------------------- E_loc
def reference_fixture(rng):
    """tests/test_variational.py:18-46 fixture (3 samples x 7 spins, complex
    mel, use mask): (conn (5, 3, 7), mel (5, 3), use (5, 3), linear complex
    wave-function weights w (7,))."""
    g = _golden()['unbalanced_vs_balanced_local_energy']
    sample = np.array(g['sample'])
    nc = g['num_connections']
    conn = rng.choice([-1, 1], size=(nc,) + sample.shape).astype(np.float64)
    conn[0] = sample
    mel = (np.array(g['hamiltonian_values_re']) + 1j * np.array(g['hamiltonian_values_im'])).T
    use = np.array(g['all_use_conn']).T
    w = rng.normal(size=sample.shape[1]) + 1j * rng.normal(size=sample.shape[1])
    return conn, mel, use, w


def test_unbalanced_equals_balanced_reference_fixture():
    conn, mel, use, w = reference_fixture(np.random.default_rng(0))

    def wf(x):
        return (x.reshape(x.shape[0], -1) @ w).astype(np.complex64)

    a = oob.local_values_unbalanced(wf, conn, mel, use)
    b = oob.local_values_balanced(wf, conn, mel)
    assert np.allclose(np.mean(a), np.mean(b))


def test_exact_local_energy_identity(rng):
    """<E_loc>_{|psi|^2} = <psi|H|psi>/<psi|psi> by full enumeration."""
    params = ocn.glorot_params(rng, 3, 16, bias_scale=0.3)
    states = oex.all_states((3, 4))
    lp = ocn.log_psi(params, states, 3, activation='elu')

    def fc(s):
        return oop.heisenberg_find_conn(s, pbc=True)

    conn, mel, use = fc(states)
    el = oob.local_values(lambda x: ocn.log_psi(params, x, 3, activation='elu'), conn, mel, use)
    p = np.exp(2 * lp.real)
    p /= p.sum()
    e_mc = (p * el).sum()
    e_ex = oex.exact_energy(fc, (3, 4), lp)
    assert abs(e_mc - e_ex) < 1e-9


def _random_log_psi(rng, n):
    """Complex log amplitudes of moderate spread over 2^n states."""
    return 0.7 * rng.standard_normal(2 ** n) + 1j * rng.uniform(-np.pi, np.pi, 2 ** n)


@pytest.mark.parametrize('op', ['heisenberg', 'ising', 'j1j2'])
def test_exact_variational_energy_is_rayleigh_quotient(rng, op):
    """ExactObservable's current_energy (optimization/exact_variational.py:
    80-95) equals <psi|H|psi>/<psi|psi> from the independently assembled
    sparse H, for a complex psi; the coefficients sum to zero and equal
    p conj(E_loc - E)."""
    fc = {'heisenberg': lambda s: oop.heisenberg_find_conn(s, pbc=True),
          'ising': lambda s: oop.ising_find_conn(s, h=0.7, j=1.3, pbc=True),
          'j1j2': lambda s: oop.j1j2_find_conn(s, j2=0.5, pbc=False)}[op]
    shape = (3, 3)
    lp = _random_log_psi(rng, 9)
    r = oex.exact_variational(lp, fc, shape)
    rq = oex.exact_energy(fc, shape, lp)
    assert abs(np.conj(r['current_energy']) - rq) < 1e-10 * max(1.0, abs(rq))
    assert abs(r['probs'].sum() - 1.0) < 1e-12
    assert abs(r['coefficients'].sum()) < 1e-12
    e = np.conj(r['current_energy'])
    assert np.allclose(r['coefficients'], r['probs'] * np.conj(r['eloc'] - e), atol=1e-14, rtol=1e-10)
    # variance = <(Re E_loc - Re E)^2> by its definition
    assert r['variance'] == pytest.approx(np.sum(r['probs'] * (r['eloc'].real - e.real) ** 2), rel=1e-10)


def test_exact_variational_at_the_ground_state():
    """For psi = the exact ground state every local energy equals E0 and the
    variance and gradient coefficients vanish (a closed form)."""
    fc = lambda s: oop.ising_find_conn(s, h=1.0, j=1.0, pbc=True)  # noqa: E731
    shape = (3, 4)
    H = oex.sparse_hamiltonian(fc, shape)
    import scipy.sparse.linalg
    w, v = scipy.sparse.linalg.eigsh(H, k=1, which='SA', tol=1e-14)
    psi = v[:, 0] * np.sign(v[np.argmax(np.abs(v[:, 0])), 0])
    assert psi.min() > 0  # transverse-field Ising: a positive ground state on every configuration
    r = oex.exact_variational(np.log(psi), fc, shape)
    assert np.abs(r['eloc'] - w[0]).max() < 1e-8
    assert r['variance'] < 1e-14
    assert np.abs(r['coefficients']).max() < 1e-10


def test_exact_sample_rule():
    """samplers/exact_sampler.py:14-20 through numpy's choice: index = first k
    with cdf_k > u (brute force on a tiny distribution)."""
    p = np.array([0.1, 0.0, 0.25, 0.4, 0.25])
    u = np.array([0.0, 0.05, 0.1, 0.1000001, 0.34, 0.35, 0.74, 0.75, 0.999])
    idx, cdf = oex.exact_sample(p, u)
    brute = [int(np.nonzero(cdf > x)[0][0]) for x in u]
    assert list(idx) == brute
    assert list(idx) == [0, 0, 2, 2, 2, 3, 3, 4, 4]


# ---------------------------------------------------------------- network
def test_convnet_normalised(rng):
    """tests/test_autoregressive.py:18-30."""
    params = ocn.glorot_params(rng, 4, 16, bias_scale=0.2)
    states = oex.all_states((4, 4))
    lp = ocn.log_psi(params, states, 4)
    assert np.exp(2 * lp.real).sum() == pytest.approx(1.0, abs=1e-10)


def test_convnet_causal(rng):
    params = ocn.glorot_params(rng, 3, 16, bias_scale=0.2)
    s = rng.choice([-1, 1], size=(1, 4, 5)).astype(np.float64)
    base = ocn.conditional_log_probs(params, s, 3).reshape(20, 2)
    for k in range(20):
        t = s.reshape(1, -1).copy()
        t[0, k:] *= -1
        c = ocn.conditional_log_probs(params, t.reshape(1, 4, 5), 3).reshape(20, 2)
        assert np.abs(c[:k + 1] - base[:k + 1]).max() < 1e-12
        if k < 19:
            assert np.abs(c[k + 1:] - base[k + 1:]).max() > 0


def test_convnet_matches_torch_conv2d(rng):
    """Every convolution restated with torch.nn.functional.conv2d (NCHW,
    explicit padding) -- an independent implementation of Keras 'valid'
    convolution after ZeroPadding2D."""
    torch = pytest.importorskip('torch')
    import torch.nn.functional as F
    depth, C = 3, 16
    params = ocn.glorot_params(rng, depth, C, bias_scale=0.3)
    s = rng.choice([-1, 1], size=(5, 4, 5)).astype(np.float64)

    def conv(x, kb, pad):
        k, b = kb
        (pt, pb), (pl, pr) = pad
        xt = F.pad(x, (pl, pr, pt, pb))
        w = torch.as_tensor(k).permute(3, 2, 0, 1)
        return F.conv2d(xt, w, torch.as_tensor(b))

    act = F.relu

    def block(v, h, layers, mask=False):
        v_out = conv(v, layers[0], ((2, 0), (1, 1)))
        x = act(conv(h, layers[1], ((0, 0), (2, 0))))
        if mask:
            x = F.pad(x, (1, 0, 0, 0))[..., :-1]
        x = conv(x, layers[2], ((0, 0), (0, 0)))
        y = F.pad(act(v_out), (0, 0, 1, 0))[..., :-1, :]
        y = conv(y, layers[3], ((0, 0), (0, 0)))
        cat = torch.cat([act(x), act(y)], dim=1)
        return v_out, conv(cat, layers[4], ((2, 0), (2, 0)))

    x = torch.as_tensor(s)[:, None]
    it = iter(range(0, 5 * ocn.num_causal_blocks(depth), 5))
    i0 = next(it)
    v, h = block(x, x, params[i0:i0 + 5])
    v, h = act(v), act(h)
    for _ in range(depth - 2):
        vi, hi = v, h
        i0 = next(it)
        v, h = block(v, h, params[i0:i0 + 5])
        v, h = act(v), act(h)
        i0 = next(it)
        v, h = block(v, h, params[i0:i0 + 5])
        v, h = act(vi + v), act(hi + h)
    i0 = next(it)
    _, h = block(v, h, params[i0:i0 + 5], mask=True)
    logits = conv(act(h), params[-1], ((0, 0), (0, 0))).permute(0, 2, 3, 1).numpy()
    ref = ocn.unnormalized_logits(params, s, depth)
    assert np.abs(logits - ref).max() < 1e-12


def test_convnet_closed_form_first_site():
    """Site (0,0) sees no input: its logits are the network's response to an
    all-zero receptive field.  With every kernel zero the conditional is
    softmax of the head bias: log p(+1) = 2*(b0 - 0.5*logsumexp(2b0, 2b1))."""
    params = ocn.glorot_params(np.random.default_rng(3), 2, 16)
    params = [(np.zeros_like(k), np.zeros_like(b)) for k, b in params]
    params[-1] = (params[-1][0], np.array([0.3, -0.2, 0.7, 1.1]))
    s = np.ones((1, 3, 3))
    lp = ocn.log_psi(params, s, 2)
    lse = 0.5 * np.log(np.exp(0.6) + np.exp(-0.4))
    assert lp[0].real == pytest.approx(9 * (0.3 - lse), abs=1e-12)
    assert lp[0].imag == pytest.approx(9 * 0.7, abs=1e-12)


def test_flops_formula():
    # 2*H*W*[(10C^2+12C) + (2 depth-3) 22 C^2 + 4C]
    assert ocn.flops_per_forward(10, 10, 20, 32) == 2 * 100 * ((10 * 1024 + 12 * 32) + 37 * 22 * 1024 + 128)
    assert ocn.num_params(20, 32) == 849156
    assert ocn.num_params(24, 64) == 4108804
    assert ocn.num_params(5, 16) == 42756


# ---------------------------------------------------------------- samplers
def test_ar_sampler_distribution(rng):
    """tests/test_samplers.py:63-82 z-statistic against exact |psi|^2."""
    params = ocn.glorot_params(rng, 2, 16, bias_scale=0.5)
    n = 2 ** 12
    u = rng.random((n, 2, 3))
    samp = osa.ar_sample(lambda x: ocn.conditional_log_probs(params, x, 2), u)
    idx = oex.binary_array_to_decimal_array(samp.reshape(n, -1))
    states = oex.all_states((2, 3))
    p = np.exp(2 * ocn.log_psi(params, states, 2).real)
    exact = rng.choice(len(p), size=n, p=p / p.sum())
    x = np.bincount(idx, minlength=64)
    y = np.bincount(exact, minlength=64)
    z = ((np.square(x - y) - x - y) / (x + y + 1e-20)).sum()
    assert z <= np.sqrt(n)


def test_mh_detailed_balance(rng):
    """MH local chains on 6 spins: long-run histogram matches |psi|^2."""
    logw = rng.normal(size=64) * 0.8
    states = oex.all_states((6,))

    def lp(s):
        return 0.5 * logw[oex.binary_array_to_decimal_array(s)]

    chains = 512
    cur = rng.choice([-1, 1], size=(chains, 6)).astype(np.float64)
    cl = lp(cur)
    hist = np.zeros(64)
    for sweep in range(600):
        cand = osa.mh_local_candidates(cur, rng.integers(0, 6, size=chains))
        cnd = lp(cand)
        acc = osa.mh_accept(cnd, cl, rng.random(chains))
        cur[acc] = cand[acc]
        cl[acc] = cnd[acc]
        if sweep >= 100:
            hist += np.bincount(oex.binary_array_to_decimal_array(cur), minlength=64)
    p = np.exp(logw) / np.exp(logw).sum()
    emp = hist / hist.sum()
    assert np.abs(emp - p).max() < 0.01
    _ = states


def test_ess_iid_and_correlated(rng):
    v = rng.normal(size=64 * 200)
    r_hat, var, cs, ess = osa.r_hat_and_ess(v, 64)
    assert abs(r_hat - 1) < 0.05 and ess > 0.5 * v.size
    # AR(1) with rho = 0.9: tau = (1+rho)/(1-rho) = 19 -> ESS ~ N/19
    x = np.zeros((64, 400))
    for t in range(1, 400):
        x[:, t] = 0.9 * x[:, t - 1] + rng.normal(size=64)
    _, _, _, ess2 = osa.r_hat_and_ess(x.reshape(-1), 64)
    assert 0.3 * x.size / 19 < ess2 < 3 * x.size / 19


# ---------------------------------------------------------------- RBM
def test_lncosh_identities(rng):
    """tests/test_tensorflow_complex_numbers_ops.py:9-33: lncosh = log cosh
    and d/dz lncosh = tanh (1e-8)."""
    z = rng.normal(size=50) * 3 + 1j * rng.normal(size=50)
    assert np.allclose(np.exp(orbm.lncosh(z)), np.cosh(z), rtol=1e-10)
    h = 1e-6
    d = (orbm.lncosh(z + h) - orbm.lncosh(z - h)) / (2 * h)
    assert np.allclose(d, np.tanh(z), atol=1e-8)


def test_rbm_brute_force(rng):
    N, M = 5, 7
    w_re, w_im = rng.normal(size=(N, M)) * 0.3, rng.normal(size=(N, M)) * 0.3
    b_re, b_im = rng.normal(size=M) * 0.3, rng.normal(size=M) * 0.3
    a_re, a_im = rng.normal(size=N) * 0.3, rng.normal(size=N) * 0.3
    s = rng.choice([-1, 1], size=(4, N)).astype(np.float64)
    got = orbm.log_psi(s, w_re, w_im, b_re, b_im, a_re, a_im)
    W, b, a = w_re - 1j * w_im, b_re - 1j * b_im, a_re - 1j * a_im
    for k in range(4):
        ref = np.log(np.exp(s[k] @ a) * np.prod(np.cosh(s[k] @ W + b)))
        assert abs(np.exp(got[k]) - np.exp(ref)) < 1e-10 * abs(np.exp(ref))


# ---------------------------------------------------------------- SR
@pytest.mark.parametrize('B,P', [(1, 1), (128, 16), (128, 40)])
def test_sr_cg_vs_pinv(rng, B, P):
    """tests/test_stochastic_reconfiguration.py:32-72: CG solution within 1e-5
    of the pinv solution of the explicit S (shift 0.01)."""
    O = rng.normal(size=(B, P)) + 1j * rng.normal(size=(B, P))
    v = rng.normal(size=P) + 1j * rng.normal(size=P)
    S = osr.s_matrix(O, 0.01)
    ref = osr.pinv_solve(S, v)
    x, it, r = osr.conjugate_gradient(lambda p: osr.s_matvec(O, p, 0.01), v, tol=1e-6, max_iter=None)
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-5
    assert np.allclose(osr.s_matvec(O, v, 0.01), S @ v)


def test_real_parameter_sr_closed_form(rng):
    B, P = 64, 10
    O = rng.normal(size=(B, P)) + 1j * rng.normal(size=(B, P))
    el = rng.normal(size=B) + 1j * rng.normal(size=B)
    x, S, F = osr.real_parameter_sr(O, el, 1e-3)
    stacked = np.concatenate([osr.centered(O).real, osr.centered(O).imag], axis=0)
    S2 = stacked.T @ stacked / B + 1e-3 * np.eye(P)
    assert np.allclose(S, S2)
    assert np.allclose(S @ x, F)


# ---------------------------------------------------------------- learned proposal (MADE)
def _made_params(rng, N, Hd, scale=0.7):
    from oracle import made as omd
    v = rng.normal(scale=scale, size=Hd * N + Hd + Hd * Hd + Hd + N * Hd + N)
    return omd.unpack(v, N, Hd)


@pytest.mark.parametrize('N,Hd', [(5, 7), (6, 16), (2, 3)])
def test_made_normalised(rng, N, Hd):
    """sum_x q(x) == 1 over all 2^N states: every mask is autoregressive."""
    from oracle import made as omd
    p = _made_params(rng, N, Hd)
    states = oex.all_states((N,)).reshape(-1, N)
    assert np.exp(omd.log_q(p, states)).sum() == pytest.approx(1.0, abs=1e-12)


def test_made_mask_error_breaks_normalisation(rng):
    """The normalisation pin is sensitive: letting logit o see hidden units of
    degree o+1 (one off) makes sum_x q(x) != 1."""
    from oracle import made as omd
    N, Hd = 5, 8
    p = list(_made_params(rng, N, Hd))
    dh = omd.degrees(N, Hd)
    v = rng.normal(size=(N, Hd))
    p[4] = v * (dh[None, :] <= np.arange(N)[:, None] + 1)
    states = oex.all_states((N,)).reshape(-1, N)
    assert abs(np.exp(omd.log_q(p, states)).sum() - 1.0) > 1e-3


def test_made_gradient_vs_finite_differences(rng):
    from oracle import made as omd
    N, Hd = 4, 6
    p = [a.copy() for a in _made_params(rng, N, Hd)]
    x = rng.choice([-1.0, 1.0], size=(7, N))
    g = omd.grad_mean_nll(p, x)
    m = omd.masks(N, Hd)
    live = [m[0], np.ones(Hd), m[1], np.ones(Hd), m[2], np.ones(N)]
    h = 1e-6
    for a, ga, lv in zip(p, g, live):
        for idx in np.ndindex(a.shape):
            if lv[idx] == 0:
                assert ga[idx] == 0
                continue
            old = a[idx]
            a[idx] = old + h
            fp = -omd.log_q(p, x).mean()
            a[idx] = old - h
            fm = -omd.log_q(p, x).mean()
            a[idx] = old
            assert ga[idx] == pytest.approx((fp - fm) / (2 * h), abs=1e-6)


def test_made_sampler_distribution(rng):
    """Samples drawn site by site follow q (z-statistic of
    tests/test_samplers.py:63-82), and the returned log q matches log_q."""
    from oracle import made as omd
    N, Hd = 4, 8
    p = _made_params(rng, N, Hd, scale=1.0)
    n = 2 ** 14
    x, lq = omd.sample(p, rng.random((n, N)))
    assert np.allclose(lq, omd.log_q(p, x))
    states = oex.all_states((N,)).reshape(-1, N)
    q = np.exp(omd.log_q(p, states))
    a = np.bincount(oex.binary_array_to_decimal_array(x), minlength=2 ** N)
    b = np.bincount(rng.choice(2 ** N, size=n, p=q / q.sum()), minlength=2 ** N)
    z = ((np.square(a - b) - a - b) / (a + b + 1e-20)).sum()
    assert z <= np.sqrt(n)


def test_made_bf16_round_matches_torch(rng):
    """The oracle's bfloat16 rounding (ties to even) equals torch's."""
    import torch
    from oracle import made as omd
    x = np.concatenate([rng.normal(scale=10.0, size=5000), [0.0, -0.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9]])
    ref = torch.as_tensor(x, dtype=torch.float32).bfloat16().double().numpy()
    assert np.array_equal(omd.bf16_round(x), ref)
