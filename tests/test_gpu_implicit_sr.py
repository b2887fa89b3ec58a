"""Products with O that never form it (batched VJP, forward-mode JVP), the
weight-normalisation chain rule, the RBM's per-sample O, the three SR solvers
(CG on a materialised O, CG on implicit products, direct Cholesky / LU) and
the device-side CG, on the GPU against materialised O, float64 finite
differences of the oracle and dense float64 algebra."""
import numpy as np
import pytest
import torch

from oracle import convnet as ocn
from oracle import rbm as orbm
from oracle import sr as osr
from tests.conftest import gpu_banner

pytestmark = pytest.mark.gpu


def _net(shape, depth, C, act, wn=False, seed=5, scale=0.3):
    import flowket_amd as fk
    m = fk.ConvNetAutoregressive2D(shape, depth=depth, num_of_channels=C, activation=act, weights_normalization=wn,
                                   precision='fp32', seed=seed)
    g = torch.Generator().manual_seed(seed + 1)
    p = m.params.cpu().clone()
    off = 0
    for kshape, bshape in m.shapes:   # non-zero biases so every bias path is exercised
        nk, nb = int(np.prod(kshape)), int(np.prod(bshape))
        p[off + nk:off + nk + nb] = scale * torch.randn(nb, generator=g)
        off += nk + nb
    m.set_params(p)
    return m


def _dense(o_re, o_im):
    return o_re.double().cpu().numpy() + 1j * o_im.double().cpu().numpy()


NETS = [((5, 6), 3, 32, 'relu', False), ((4, 4), 2, 64, 'elu', False), ((3, 7), 4, 32, 'elu', True),
        ((4, 5), 3, 32, 'relu', True)]


@pytest.mark.parametrize('shape,depth,C,act,wn', NETS)
def test_vjp_and_jvp_equal_products_with_materialised_o(rng, shape, depth, C, act, wn):
    """machine.vjp(x, u, scale) = scale Re[O^H u] and machine.jvp(x, t) = O t
    with O from the per-sample gradient kernel (over the trainable values,
    v and g with weight normalisation)."""
    gpu_banner(f'vjp/jvp vs O shape={shape} depth={depth} C={C} {act} wn={wn}')
    m = _net(shape, depth, C, act, wn)
    B = 37
    x = rng.choice([-1, 1], size=(B,) + shape).astype(np.int8)
    o_re, o_im, _ = m.per_sample_gradients(x)
    O = _dense(o_re, o_im)
    assert O.shape == (B, m.n_trainable)
    u = rng.normal(size=(B, 2))
    got = m.vjp(x, torch.as_tensor(u, dtype=torch.float32), 0.5).double().cpu().numpy()
    ref = 0.5 * (O.real.T @ u[:, 0] + O.imag.T @ u[:, 1])
    assert np.abs(got - ref).max() < 2e-5 * np.abs(ref).max()
    t = rng.normal(size=m.n_trainable)
    w = m.jvp(x, torch.as_tensor(t, dtype=torch.float32)).double().cpu().numpy()
    ref_w = O @ t
    assert np.abs(w[:, 0] - ref_w.real).max() < 2e-5 * np.abs(ref_w).max()
    assert np.abs(w[:, 1] - ref_w.imag).max() < 2e-5 * np.abs(ref_w).max()


def _wn_log_psi(m, spins, vec, depth, act):
    """Oracle log psi of a weight-normalised network from its trainable
    values [params | g] (deepar/layers/wrappers.py:123-135)."""
    P = m.n_params
    params, off, goff = [], 0, 0
    for idx, (ks, bs) in enumerate(m.shapes):
        nk, nb = int(np.prod(ks)), int(np.prod(bs))
        k = vec[off:off + nk].reshape(ks)
        if idx < len(m.shapes) - 1:
            g = vec[P + goff:P + goff + ks[3]]
            goff += ks[3]
            k = ocn.weight_norm(k, g, exponential_norm=True)
        params.append((k, vec[off + nk:off + nk + nb]))
        off += nk + nb
    return ocn.log_psi(params, spins, depth, activation=act)


@pytest.mark.parametrize('C', [32, 4])
def test_weight_normalised_gradients_vs_finite_differences(rng, C):
    """O over (v, g) of a weight-normalised network (the reference's default)
    against central finite differences of the float64 oracle, on a random
    subset of columns that includes g entries (MFMA kernel for 32 channels,
    scalar kernel for 4)."""
    gpu_banner(f'weight-normalised O vs finite differences C={C}')
    depth, act = 3, 'elu'
    m = _net((4, 4), depth, C, act, wn=True, seed=9)
    g = m.wn_g.cpu().clone() + 0.2 * torch.randn(m.wn_g.numel(), generator=torch.Generator().manual_seed(3))
    m.set_wn_g(g)
    B = 5
    x = rng.choice([-1, 1], size=(B, 4, 4)).astype(np.int8)
    o_re, o_im, _ = m.per_sample_gradients(x)
    O = _dense(o_re, o_im)
    vec = m.trainable.double().cpu().numpy()
    cols = np.concatenate([rng.choice(m.n_params, size=30, replace=False),
                           m.n_params + rng.choice(m.n_trainable - m.n_params, size=20, replace=False)])
    h = 1e-6
    for k in cols:
        e = np.zeros_like(vec)
        e[k] = h
        fd = (_wn_log_psi(m, x, vec + e, depth, act) - _wn_log_psi(m, x, vec - e, depth, act)) / (2 * h)
        assert np.abs(O[:, k] - fd).max() <= 1e-4 * (np.abs(O).max() + 1e-3), k


def test_rbm_per_sample_gradients_vs_finite_differences(rng):
    """fk_rbm_grad against central finite differences of the oracle RBM over
    every stored real and imaginary part (rbm.py:35-48, base_layer.py:30-35)."""
    import flowket_amd as fk
    gpu_banner('RBM O vs finite differences')
    m = fk.RBM((3, 4), alpha=2, stddev=0.3, seed=4)
    B = 6
    x = rng.choice([-1, 1], size=(B, 3, 4)).astype(np.int8)
    o_re, o_im, lp = m.per_sample_gradients(x)
    O = _dense(o_re, o_im)
    vec = m.params.double().cpu().numpy()
    N, M = m.n_visible, m.n_hidden
    sizes = [N * M, N * M, M, M, N, N]

    def lp_of(v):
        parts, off = [], 0
        for n in sizes:
            parts.append(v[off:off + n])
            off += n
        return orbm.log_psi(x.reshape(B, -1), parts[0].reshape(N, M), parts[1].reshape(N, M), *parts[2:])

    h = 1e-6
    for k in range(vec.size):
        e = np.zeros_like(vec)
        e[k] = h
        fd = (lp_of(vec + e) - lp_of(vec - e)) / (2 * h)
        assert np.abs(O[:, k] - fd).max() < 1e-4 * (np.abs(O).max()), k
    assert np.abs(lp.cpu().numpy()[:, 0] - lp_of(vec).real).max() < 1e-4


def test_implicit_s_products_and_solve_equal_materialised(rng):
    """S v through the implicit products equals S v through the materialised
    O (whose products are pinned to dense float64 algebra in
    test_gpu_grad_sr.py); the force agrees; the implicit SR solve satisfies
    S x = F (S applied through the materialised O) and equals the
    materialised solve.  (A dense solve of this network's ~79k parameters
    would be the reference test's pinv, at a size the CPU cannot afford.)"""
    import flowket_amd as fk
    from flowket_amd.optimizers import _ImplicitS, _OFactor, _complex_eloc_pairs
    gpu_banner('implicit vs materialised SR 4x5 C32')
    m = _net((4, 5), 3, 32, 'relu', wn=True, seed=12)
    B = 64
    x = rng.choice([-1, 1], size=(B, 4, 5)).astype(np.int8)
    shift = 1e-2
    of, im = _OFactor(m, x), _ImplicitS(m, x)
    v = torch.as_tensor(rng.normal(size=m.n_trainable), dtype=torch.float32).cuda()
    a = of.s_matvec(v, shift).double().cpu().numpy()
    b = im.s_matvec(v, shift).double().cpu().numpy()
    assert np.abs(a - b).max() < 2e-5 * np.abs(a).max()
    el = torch.as_tensor(rng.normal(size=B) + 1j * rng.normal(size=B)).cuda()
    u = _complex_eloc_pairs(el, el.mean()).cuda()
    F = of.force(u)
    assert torch.allclose(im.force(u), F, rtol=1e-4, atol=2e-5 * F.abs().max().item())
    sols = {}
    for jvp in (False, True):
        sr = fk.ComplexValuesStochasticReconfiguration(m, lr=0.1, diag_shift=shift, conjugate_gradient_tol=1e-5,
                                                       iterative_solver_max_iterations=300,
                                                       compute_jvp_instead_of_full_jacobian=jvp)
        sol, met = sr.solve(x, el)
        res = met['stochastic_reconfiguration/conjugate_gradient_residual_norm']
        assert met['stochastic_reconfiguration/conjugate_gradient_iterations'] > 0
        r = (of.s_matvec(sol, shift) - F).double().norm().item()
        assert r <= 1e-4 * F.double().norm().item() and res <= 2e-5 * F.double().norm().item(), (r, res)
        sols[jvp] = sol.double().cpu().numpy()
    assert np.linalg.norm(sols[True] - sols[False]) / np.linalg.norm(sols[False]) < 2e-3


@pytest.mark.parametrize('use_cholesky', [True, False])
def test_direct_solver(rng, use_cholesky):
    """iterative_solver=False (optimizer.py:55-67): S formed and solved by
    Cholesky or LU equals the dense float64 solution; s-matrix stats match
    numpy's rank and smallest kept eigenvalue."""
    import flowket_amd as fk
    gpu_banner(f'direct SR solver cholesky={use_cholesky}')
    m = _net((3, 3), 2, 4, 'elu', seed=13)
    B = 50
    x = rng.choice([-1, 1], size=(B, 3, 3)).astype(np.int8)
    el = torch.as_tensor(rng.normal(size=B) + 1j * rng.normal(size=B)).cuda()
    shift = 1e-3
    sr = fk.ComplexValuesStochasticReconfiguration(m, diag_shift=shift, iterative_solver=False,
                                                   use_cholesky=use_cholesky, add_s_matrix_stats=True)
    sol, met = sr.solve(x, el)
    o_re, o_im, _ = m.per_sample_gradients(x)
    O = _dense(o_re, o_im)
    x_ref, S, _ = osr.real_parameter_sr(O, el.cpu().numpy(), shift)
    assert np.linalg.norm(sol.double().cpu().numpy() - x_ref) / np.linalg.norm(x_ref) < 1e-4
    S0 = S - shift * np.eye(S.shape[0])
    ev = np.abs(np.linalg.eigvalsh(S0))
    tol = np.finfo(np.float32).eps * S0.shape[0] * np.abs(S0).max()
    assert met['stochastic_reconfiguration/s_matrix_rank'] == int((ev > tol).sum())
    assert met['stochastic_reconfiguration/s_matrix_min_eigval'] == pytest.approx(ev[ev > tol].min(), rel=1e-6)


@pytest.mark.parametrize('n,tol,max_iter', [(60, 1e-6, 500), (60, 1e-3, 500), (200, 1e-8, 7), (50, 0.0, 0)])
def test_device_cg_follows_the_reference_rule(rng, n, tol, max_iter):
    """The device-side CG (fk_cg_init / fk_cg_step) against the oracle's
    restatement of linear_equations.py:87-131 on a dense SPD system: same
    iteration count, solution and residual; with check_every=None (all
    max_iter steps launched) the result is unchanged."""
    from flowket_amd.optimizers import conjugate_gradient
    gpu_banner(f'device CG n={n} tol={tol} max_iter={max_iter}')
    A = rng.normal(size=(n, n))
    A = A @ A.T / n + 0.1 * np.eye(n)
    b = rng.normal(size=n)
    Ad = torch.as_tensor(A, dtype=torch.float32).cuda()
    bd = torch.as_tensor(b, dtype=torch.float32).cuda()

    def apply(p, z):
        torch.matmul(Ad, p, out=z)

    x_ref, it_ref, _ = osr.conjugate_gradient(lambda p: A @ p, b, tol=tol, max_iter=max_iter)
    for check in (10, 1, None):
        x, it, res = conjugate_gradient(apply, bd, tol=tol, max_iter=max_iter, check_every=check)
        if max_iter == 0:
            assert it == 0 and torch.count_nonzero(x).item() == 0
            continue
        assert abs(it - it_ref) <= 1, (it, it_ref)
        assert np.linalg.norm(x.double().cpu().numpy() - x_ref) <= 1e-3 * np.linalg.norm(x_ref)
        r = b - A @ x.double().cpu().numpy()
        assert res == pytest.approx(np.linalg.norm(r), rel=1e-2, abs=1e-5)


def test_sr_step_updates_and_metric_names(rng):
    """step() applies w <- w - lr S^{-1} F to the trainable values
    (optimizer.py:33-42) and reports the reference's metric names
    (:126-152); with weight normalisation g moves too."""
    import flowket_amd as fk
    gpu_banner('SR step update')
    m = _net((4, 4), 3, 32, 'elu', wn=True, seed=21)
    B = 128
    x = rng.choice([-1, 1], size=(B, 4, 4)).astype(np.int8)
    el = torch.as_tensor(rng.normal(size=B) + 1j * rng.normal(size=B)).cuda()
    sr = fk.ComplexValuesStochasticReconfiguration(m, lr=0.05, diag_shift=1e-2, conjugate_gradient_tol=1e-5,
                                                   compute_jvp_instead_of_full_jacobian=True)
    before = m.trainable.clone()
    sol, _ = sr.solve(x, el)
    met = sr.step(x, el)
    assert set(met) == {'stochastic_reconfiguration/conjugate_gradient_iterations',
                        'stochastic_reconfiguration/conjugate_gradient_residual_norm'}
    assert sr.iterations == 1
    assert torch.allclose(m.trainable, before - 0.05 * sol, rtol=1e-5, atol=1e-7)
    assert not torch.equal(m.wn_g, before[m.n_params:])
    with pytest.raises(NotImplementedError):
        fk.ComplexValuesStochasticReconfiguration(m, use_energy_loss=True)


@pytest.mark.parametrize('jvp', [False, True])
def test_sr_training_lowers_the_exact_energy(jvp):
    """SR on the 4x4 periodic Heisenberg model (32 channels, weight
    normalisation on, the reference's defaults) from fast samples: the
    exact energy of the network (all 65,536 states) drops by more than 5
    over 25 steps, with both the materialised and the implicit products."""
    import flowket_amd as fk
    gpu_banner(f'SR training 4x4 jvp={jvp}')
    m = fk.ConvNetAutoregressive2D((4, 4), depth=3, num_of_channels=32, activation='relu',
                                   weights_normalization=True, precision='fp32', seed=1)
    op = fk.Heisenberg(hilbert_state_shape=(4, 4), pbc=True)
    ev = fk.ExactVariational(m, op, 2 ** 16)
    ev.machine_updated()
    e0 = ev.energy_observable.current_energy.real
    vmc = fk.VariationalMonteCarlo(m, op, fk.FastAutoregressiveSampler(m, 512, seed=3))
    sr = fk.ComplexValuesStochasticReconfiguration(m, lr=0.05, diag_shift=1e-2, conjugate_gradient_tol=1e-4,
                                                   iterative_solver_max_iterations=50,
                                                   compute_jvp_instead_of_full_jacobian=jvp)
    for _ in range(25):
        x, _ = vmc.next_batch()
        sr.step(x, vmc.current_local_energy, vmc.current_energy)
    ev.machine_updated()
    e1 = ev.energy_observable.current_energy.real
    assert e1 < e0 - 5.0, (e0, e1)
    assert e1 > -44.9139328337 - 1e-6   # variational bound
