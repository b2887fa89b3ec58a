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


@pytest.mark.parametrize('act', ['relu', 'elu'])
def test_convnet_matches_torch_conv2d(rng, act):
    """The network restated with torch.nn.functional.conv2d (oracle/
    convnet_torch.py: NCHW, explicit padding) -- an independent implementation
    of Keras 'valid' convolution after ZeroPadding2D -- equals the numpy
    oracle's logits and log psi in float64."""
    from oracle import convnet_torch as oct_
    depth, C = 3, 16
    params = ocn.glorot_params(rng, depth, C, bias_scale=0.3)
    s = rng.choice([-1, 1], size=(5, 4, 5)).astype(np.float64)
    tp = oct_.torch_params(params)
    logits = oct_.unnormalized_logits(tp, s, depth, act).permute(0, 2, 3, 1).numpy()
    ref = ocn.unnormalized_logits(params, s, depth, activation=act)
    assert np.abs(logits - ref).max() < 1e-12
    lp = oct_.log_psi(tp, s, depth, act, chunk=2)
    assert np.abs(lp - ocn.log_psi(params, s, depth, activation=act)).max() < 1e-12


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


# ------------------------------------------------ connection-slot known answers
# (slot, a, b) of every bond, transcribed by hand from
# HeisenbergFindConn.calc_conn_and_mel (operators/heisenberg.py:63-68,86-121):
# the mel/use views are reshaped (H, W, dim, B), so site (i, j) owns slots
# 1 + (i*W + j)*dim + d, where d counts only the lattice axes longer than one
# (down first, then right) and dim is the rank of the operator's shape.  A
# slot not listed is never written: use False, mel 0, conn = the sample.
HEIS_SLOT_KAT = {
    ((2, 3), True): [(1, 0, 3), (2, 0, 1), (3, 1, 4), (4, 1, 2), (5, 2, 5), (6, 2, 0), (7, 3, 0), (8, 3, 4),
                     (9, 4, 1), (10, 4, 5), (11, 5, 2), (12, 5, 3)],
    ((2, 3), False): [(1, 0, 3), (2, 0, 1), (3, 1, 4), (4, 1, 2), (5, 2, 5), (8, 3, 4), (10, 4, 5)],
    ((5,), True): [(1, 0, 1), (2, 1, 2), (3, 2, 3), (4, 3, 4), (5, 4, 0)],
    ((5,), False): [(1, 0, 1), (2, 1, 2), (3, 2, 3), (4, 3, 4)],
    ((1, 4), True): [(1, 0, 1), (3, 1, 2), (5, 2, 3), (7, 3, 0)],
    ((1, 4), False): [(1, 0, 1), (3, 1, 2), (5, 2, 3)],
    ((4, 1), True): [(1, 0, 1), (3, 1, 2), (5, 2, 3), (7, 3, 0)],
}
HEIS_NC = {(2, 3): 13, (5,): 6, (1, 4): 9, (4, 1): 9}


@pytest.mark.parametrize('shape,pbc', list(HEIS_SLOT_KAT))
def test_heisenberg_slot_known_answers(shape, pbc):
    """Every slot of the oracle's Heisenberg find_conn against the hand
    transcription: which bond it holds (the swapped pair), its use flag and
    element, on all 2^N configurations (N <= 6)."""
    n = int(np.prod(shape))
    kat = {c: (a, b) for c, a, b in HEIS_SLOT_KAT[(shape, pbc)]}
    states = np.array([[1 if (k >> q) & 1 else -1 for q in range(n)] for k in range(2 ** n)], np.float64)
    conn, mel, use = oop.heisenberg_find_conn(states.reshape((-1,) + shape), pbc=pbc)
    assert conn.shape[0] == HEIS_NC[shape]
    conn = conn.reshape(conn.shape[0], 2 ** n, n)
    for c in range(1, conn.shape[0]):
        if c not in kat:
            assert not use[c].any() and np.all(mel[c] == 0) and np.array_equal(conn[c], states)
            continue
        a, b = kat[c]
        anti = states[:, a] != states[:, b]
        swapped = states.copy()
        swapped[:, a], swapped[:, b] = states[:, b], states[:, a]
        assert np.array_equal(use[c], anti)
        assert np.array_equal(mel[c], np.where(anti, -2.0, 0.0))
        assert np.array_equal(conn[c], swapped)
    diag = sum(states[:, a] * states[:, b] for a, b in kat.values())
    assert np.array_equal(mel[0], diag)


@pytest.mark.parametrize('shape,pbc', list(HEIS_SLOT_KAT))
def test_heisenberg_operator_tables_known_answers(shape, pbc):
    """The GPU operator's connection table (built on the host) holds the same
    bonds in the same slots."""
    import flowket_amd as fk
    op = fk.Heisenberg(hilbert_state_shape=shape, pbc=pbc)
    kat = {c: (a, b) for c, a, b in HEIS_SLOT_KAT[(shape, pbc)]}
    assert op.num_connections == HEIS_NC[shape]
    for c in range(1, op.num_connections):
        got = (int(op._conn_a[c]), int(op._conn_b[c]))
        assert got == kat.get(c, (-1, -1)), c


def _numpy_pairwise_f32(vals):
    """numpy's float32 pairwise summation written out (the algorithm the GPU
    kernel implements): < 8 values one by one; <= 128 values in 8 interleaved
    float32 partial sums, combined ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)), then
    the remainder; longer runs split at n/2 rounded down to a multiple of 8."""
    f = np.float32
    n = len(vals)
    if n < 8:
        res = f(0.0)
        for v in vals:
            res = f(res + v)
        return res
    if n <= 128:
        r = [f(v) for v in vals[:8]]
        i = 8
        while i < n - (n % 8):
            for q in range(8):
                r[q] = f(r[q] + vals[i + q])
            i += 8
        res = f(f(f(r[0] + r[1]) + f(r[2] + r[3])) + f(f(r[4] + r[5]) + f(r[6] + r[7])))
        for v in vals[i:]:
            res = f(res + v)
        return res
    n2 = n // 2
    n2 -= n2 % 8
    return f(_numpy_pairwise_f32(vals[:n2]) + _numpy_pairwise_f32(vals[n2:]))


ISING_F32_CASES = [((10, 10), True, 0.3), ((20, 20), False, 0.3), ((300,), True, 0.7), ((12, 12), False, 1.1),
                   ((1, 9), True, 0.1), ((9, 1), False, 0.1), ((5,), True, 0.3), ((17, 13), True, -0.37)]


@pytest.mark.parametrize('shape,pbc,j', ISING_F32_CASES)
def test_ising_diagonal_float32_rule(rng, shape, pbc, j):
    """The oracle's Ising diagonal (numpy, as ising.py:25-41 computes it) equals
    per-site float32 values folded term by term (each term in float64, rounded
    after every subtraction) and summed by the pairwise rule above, bit for
    bit, for non-dyadic j; and it differs from a plain float64 sum for some
    sample, so the rule is observable."""
    s = rng.choice([-1, 1], size=(40,) + shape).astype(np.float64)
    _, mel, _ = oop.ising_find_conn(s, h=1.0, j=j, pbc=pbc)
    s2 = s if len(shape) == 2 else s[:, :, None]
    H, W = s2.shape[1:]
    want = np.empty(40)
    plain = np.empty(40)
    for b in range(40):
        vals, tot = [], 0.0
        for i in range(H):
            for q in range(W):
                v = np.float32(0.0)
                terms = []
                if len(shape) == 1 or H > 1:
                    terms.append(s2[b, (i + 1) % H, q] if (i + 1 < H or pbc) else 0.0)
                if len(shape) == 2 and W > 1:
                    terms.append(s2[b, i, (q + 1) % W] if (q + 1 < W or pbc) else 0.0)
                for t in terms:
                    v = np.float32(float(v) - j * s2[b, i, q] * t)
                    tot -= j * s2[b, i, q] * t
                vals.append(v)
        want[b] = float(_numpy_pairwise_f32(vals))
        plain[b] = tot
    assert np.array_equal(mel[0], want)
    if int(np.prod(shape)) >= 100:
        assert not np.array_equal(mel[0], plain)


@pytest.mark.parametrize('chains,spc', [(16, 64), (8, 5), (32, 1), (4, 2)])
def test_calc_r_hat_value_matches_oracle(rng, chains, spc):
    """The package's R-hat / ESS (samplers.calc_r_hat_value) equals the
    oracle's restatement of metropolis_hastings.py:67-89 on fixed correlated
    chains, including one sample per chain (R-hat 1, ESS = batch)."""
    from flowket_amd.samplers import calc_r_hat_value
    x = np.zeros((chains, spc))
    x[:, 0] = rng.normal(size=chains)
    for t in range(1, spc):
        x[:, t] = 0.8 * x[:, t - 1] + rng.normal(size=chains)
    x += rng.normal(size=(chains, 1)) * 0.3
    got = calc_r_hat_value(x.reshape(-1), chains)
    ref = osa.r_hat_and_ess(x.reshape(-1), chains)
    assert np.allclose(np.array(got, dtype=float), np.array(ref, dtype=float), rtol=1e-12, atol=0)
    if spc == 1:
        assert got[0] == 1.0 and got[2] == 0.0 and got[3] == chains


def test_keras_adam_first_step_closed_form(rng):
    """Keras Adam's first step is p - lr*g/(|g| + eps/sqrt(1-beta_2)); a
    torch-style update (eps after the bias correction) differs from it."""
    from oracle import optim as oopt
    p0 = rng.normal(size=50).astype(np.float32)
    g = (rng.normal(size=50) * 10.0 ** rng.integers(-9, 1, size=50)).astype(np.float32)
    lr, b2, eps = 1e-2, 0.999, 1e-7
    got = oopt.keras_adam(p0, [g], lr=lr, beta_2=b2, epsilon=eps)
    want = p0.astype(np.float64) - lr * g / (np.abs(g) + eps / np.sqrt(1 - b2))
    assert np.allclose(got, want, rtol=1e-6, atol=1e-9)
    torch_style = p0.astype(np.float64) - lr * g / (np.abs(g) + eps)
    assert not np.allclose(got, torch_style, rtol=1e-6, atol=1e-9)
