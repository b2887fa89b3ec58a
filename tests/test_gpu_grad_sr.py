"""Per-sample gradients O_k, the SR matvec/solve and the energy gradient on
the GPU, against float64 finite differences of the oracle network and dense
float64 linear algebra."""
import numpy as np
import pytest
import torch

from oracle import convnet as ocn
from oracle import sr as osr
from tests.conftest import gpu_banner
from tests.test_gpu_convnet import _machine, _oracle_params

pytestmark = pytest.mark.gpu


def _fd_jacobian(machine, spins, depth, act, h=1e-6):
    """Central finite differences of the float64 oracle log psi w.r.t. every
    flat parameter: (B, P) complex."""
    flat = machine.effective_params().double().cpu().numpy()
    shapes = machine.shapes

    def lp(vec):
        params, off = [], 0
        for ks, bs in shapes:
            nk, nb = int(np.prod(ks)), int(np.prod(bs))
            params.append((vec[off:off + nk].reshape(ks), vec[off + nk:off + nk + nb]))
            off += nk + nb
        return ocn.log_psi(params, spins, depth, activation=act)

    J = np.zeros((spins.shape[0], flat.size), dtype=np.complex128)
    for k in range(flat.size):
        e = np.zeros_like(flat)
        e[k] = h
        J[:, k] = (lp(flat + e) - lp(flat - e)) / (2 * h)
    return J


@pytest.mark.parametrize('shape,depth,C,act', [((3, 3), 3, 4, 'elu'), ((2, 4), 2, 4, 'relu'),
                                              ((3, 4), 4, 2, 'elu')])
def test_per_sample_gradients_vs_finite_differences(rng, shape, depth, C, act):
    gpu_banner(f'per-sample gradients shape={shape} depth={depth} C={C} {act}')
    import flowket_amd as fk
    m = fk.ConvNetAutoregressive2D(shape, depth=depth, num_of_channels=C, activation=act,
                                   weights_normalization=False, precision='fp32', seed=3)
    g = torch.Generator().manual_seed(4)
    m.set_params(m.params.cpu() + 0.2 * torch.randn(m.n_params, generator=g))
    s = rng.choice([-1, 1], size=(5,) + shape).astype(np.int8)
    o_re, o_im, lp = m.per_sample_gradients(s)
    J = _fd_jacobian(m, s.astype(np.float64), depth, act)
    got = o_re.double().cpu().numpy() + 1j * o_im.double().cpu().numpy()
    scale = np.abs(J).max()
    assert np.abs(got - J).max() / scale < 1e-4
    ref_lp = ocn.log_psi(_oracle_params(m), s, depth, activation=act)
    assert np.abs(lp[:, 0].cpu().numpy() - ref_lp.real).max() < 1e-4 * (np.abs(ref_lp.real).max() + 1)


def test_gradients_match_fused_forward_logpsi(rng):
    """The gradient kernel's forward and the fused forward agree (FP32)."""
    gpu_banner('gradient kernel forward vs fused forward 10x10 depth 4 C 32')
    m = _machine((10, 10), 4, 32, 'relu', 'fp32', bias_scale=0.2)
    s = rng.choice([-1, 1], size=(16, 10, 10)).astype(np.int8)
    _, _, lp = m.per_sample_gradients(s)
    ref = m.log_psi(s)
    assert torch.allclose(lp[:, 0], ref.real, rtol=1e-5, atol=1e-4)
    assert torch.allclose(lp[:, 1], ref.imag, rtol=1e-5, atol=1e-4)


def test_sr_matvec_solve_and_energy_gradient(rng):
    """S v, the CG solution and 2 Re[O^H (E_loc - E)]/B against dense float64
    algebra on the same O (tests/test_stochastic_reconfiguration.py:32-109)."""
    gpu_banner('SR matvec / CG / energy gradient')
    import flowket_amd as fk
    from flowket_amd.optimizers import _OFactor, _complex_eloc_pairs
    m = _machine((4, 4), 3, 4, 'elu', 'fp32', bias_scale=0.3)
    B = 96
    s = rng.choice([-1, 1], size=(B, 4, 4)).astype(np.int8)
    of = _OFactor(m, s)
    O = of.o_re.double().cpu().numpy() + 1j * of.o_im.double().cpu().numpy()
    P = O.shape[1]
    shift = 0.01
    v = rng.normal(size=P)
    Sv = of.s_matvec(torch.as_tensor(v, dtype=torch.float32).cuda(), shift).double().cpu().numpy()
    _, S, _ = osr.real_parameter_sr(O, np.zeros(B, complex), shift)
    assert np.linalg.norm(Sv - S @ v) / np.linalg.norm(S @ v) < 1e-5
    el = torch.as_tensor(rng.normal(size=B) + 1j * rng.normal(size=B)).cuda()
    x_ref, _, F = osr.real_parameter_sr(O, el.cpu().numpy(), shift)
    sr = fk.ComplexValuesStochasticReconfiguration(m, lr=0.1, diag_shift=shift, conjugate_gradient_tol=1e-6,
                                                   iterative_solver_max_iterations=500)
    x, met = sr.solve(s, el)
    it = met['stochastic_reconfiguration/conjugate_gradient_iterations']
    res = met['stochastic_reconfiguration/conjugate_gradient_residual_norm']
    x = x.double().cpu().numpy()
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) < 1e-3
    assert it > 0 and res <= 1e-6 * np.linalg.norm(F) * 1.01
    g = fk.energy_gradient(m, s, el).double().cpu().numpy()
    e = el.cpu().numpy()
    g_ref = 2 * np.real(O.T @ np.conj(e - e.mean())) / B
    assert np.abs(g - g_ref).max() / np.abs(g_ref).max() < 1e-5
    u = _complex_eloc_pairs(el, el.mean())
    assert u.shape == (B, 2)


def test_vmc_adam_lowers_energy_4x4():
    """Config-1 shape: 4x4 periodic Heisenberg, depth 5, 16 channels, ELU,
    exact AR samples, Adam lr 1e-3: the energy goes down over 40 steps."""
    gpu_banner('VMC Adam 4x4 depth 5 C16 40 steps')
    import flowket_amd as fk
    m = fk.ConvNetAutoregressive2D((4, 4), depth=5, num_of_channels=16, activation='elu',
                                   weights_normalization=False, precision='fp32', seed=0)
    op = fk.Heisenberg(hilbert_state_shape=(4, 4), pbc=True)
    sampler = fk.AutoregressiveSampler(m, 1024, seed=1)
    vmc = fk.VariationalMonteCarlo(m, op, sampler)
    opt = fk.Adam(m, lr=1e-3)
    energies = []
    for _ in range(40):
        configs, _ = vmc.next_batch()
        energies.append(vmc.current_energy.real)
        opt.step(configs, vmc.current_local_energy, vmc.current_energy)
    assert np.mean(energies[-5:]) < np.mean(energies[:5]) - 1.0
    assert np.mean(energies[-5:]) > -44.9139328337 - 0.5  # variational bound (statistical slack)


def _torch_log_psi(flat, shapes, spins, depth, act_name):
    """float64 torch.nn.functional restatement of the network (the same
    restatement the oracle tests pin against oracle.convnet) on a flat
    parameter tensor, for autograd Jacobians."""
    import torch.nn.functional as F
    params, off = [], 0
    for ks, bs in shapes:
        nk, nbias = int(np.prod(ks)), int(np.prod(bs))
        params.append((flat[off:off + nk].reshape(ks), flat[off + nk:off + nk + nbias]))
        off += nk + nbias
    act = F.relu if act_name == 'relu' else F.elu

    def conv(x, kb, pad):
        k, b = kb
        (pt, pb), (pl, pr) = pad
        return F.conv2d(F.pad(x, (pl, pr, pt, pb)), k.permute(3, 2, 0, 1), b)

    def block(v, h, layers, mask=False):
        v_out = conv(v, layers[0], ((2, 0), (1, 1)))
        x = act(conv(h, layers[1], ((0, 0), (2, 0))))
        if mask:
            x = F.pad(x, (1, 0, 0, 0))[..., :-1]
        x = conv(x, layers[2], ((0, 0), (0, 0)))
        y = F.pad(act(v_out), (0, 0, 1, 0))[..., :-1, :]
        y = conv(y, layers[3], ((0, 0), (0, 0)))
        return v_out, conv(torch.cat([act(x), act(y)], dim=1), layers[4], ((2, 0), (2, 0)))

    x = torch.as_tensor(spins, dtype=torch.float64)[:, None]
    nbk = ocn.num_causal_blocks(depth)
    it = iter(range(0, 5 * nbk, 5))
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
    z = conv(act(h), params[-1], ((0, 0), (0, 0))).permute(0, 2, 3, 1)      # (B, H, W, 4)
    re, im = z[..., 0:2], z[..., 2:4]
    re = re - 0.5 * torch.logsumexp(2 * re, dim=-1, keepdim=True)
    idx = torch.as_tensor((1 - np.asarray(spins, dtype=np.int64)) // 2)[..., None]
    return torch.gather(re, -1, idx)[..., 0].sum(dim=(1, 2)), torch.gather(im, -1, idx)[..., 0].sum(dim=(1, 2))


@pytest.mark.parametrize('shape,depth,C,act', [((5, 6), 3, 32, 'elu'), ((4, 4), 2, 64, 'relu'),
                                              ((3, 7), 4, 32, 'relu')])
def test_mfma_per_sample_gradients_vs_autograd(rng, shape, depth, C, act):
    """The MFMA gradient kernel (widths that are multiples of 32) against
    float64 autograd per-sample Jacobians of the torch restatement, itself
    checked against the oracle's log psi."""
    gpu_banner(f'MFMA per-sample gradients shape={shape} depth={depth} C={C} {act}')
    import flowket_amd as fk
    m = fk.ConvNetAutoregressive2D(shape, depth=depth, num_of_channels=C, activation=act,
                                   weights_normalization=False, precision='fp32', seed=6)
    g = torch.Generator().manual_seed(7)
    m.set_params(m.params.cpu() * (1.0 + 0.3 * torch.randn(m.n_params, generator=g)))
    B = 3
    s = rng.choice([-1, 1], size=(B,) + shape).astype(np.int8)
    o_re, o_im, lp = m.per_sample_gradients(s)
    flat = m.params.double().cpu().clone().requires_grad_(True)
    re, im = _torch_log_psi(flat, m.shapes, s, depth, act)
    ref = ocn.log_psi(_oracle_params(m), s, depth, activation=act)
    assert np.abs(re.detach().numpy() - ref.real).max() < 1e-9
    J_re = torch.stack([torch.autograd.grad(re[b], flat, retain_graph=True)[0] for b in range(B)]).numpy()
    J_im = torch.stack([torch.autograd.grad(im[b], flat, retain_graph=True)[0] for b in range(B)]).numpy()
    got_re, got_im = o_re.double().cpu().numpy(), o_im.double().cpu().numpy()
    scale = max(np.abs(J_re).max(), np.abs(J_im).max())
    assert np.abs(got_re - J_re).max() / scale < 1e-4
    assert np.abs(got_im - J_im).max() / scale < 1e-4
    assert np.abs(lp[:, 0].cpu().numpy() - ref.real).max() < 1e-4 * (np.abs(ref.real).max() + 1)


@pytest.mark.parametrize('B,P', [(37, 1000), (37, 1001), (5, 4)])
def test_sr_products_c_abi(rng, B, P):
    """fk_o_times_v / fk_oh_times_u on random O against dense float64 algebra,
    on both the float4 (P % 4 == 0) and the scalar column paths."""
    from flowket_amd import _lib
    gpu_banner(f'SR products B={B} P={P}')
    O = rng.normal(size=(B, P)) + 1j * rng.normal(size=(B, P))
    v = rng.normal(size=P)
    ore = torch.as_tensor(O.real, dtype=torch.float32).contiguous().cuda()
    oim = torch.as_tensor(O.imag, dtype=torch.float32).contiguous().cuda()
    mre, mim = torch.empty(P, device='cuda'), torch.empty(P, device='cuda')
    _lib.call('fk_o_mean', _lib.ptr(ore), _lib.ptr(oim), B, P, _lib.ptr(mre), _lib.ptr(mim), _lib.stream_ptr())
    scratch = torch.empty(int(_lib.load().fk_sr_scratch_doubles()), dtype=torch.float64, device='cuda')
    vd = torch.as_tensor(v, dtype=torch.float32).cuda()
    u = torch.empty((B, 2), dtype=torch.float32, device='cuda')
    _lib.call('fk_o_times_v', _lib.ptr(ore), _lib.ptr(oim), _lib.ptr(mre), _lib.ptr(mim), _lib.ptr(vd), B, P,
              _lib.ptr(u), _lib.ptr(scratch), _lib.stream_ptr())
    Oc = O.astype(np.complex64).astype(np.complex128)
    Oc = Oc - Oc.mean(axis=0)
    ref_u = Oc @ v.astype(np.float32).astype(np.float64)
    got_u = u[:, 0].double().cpu().numpy() + 1j * u[:, 1].double().cpu().numpy()
    assert np.abs(got_u - ref_u).max() < 1e-4 * np.abs(ref_u).max()
    out = torch.empty(P, device='cuda')
    _lib.call('fk_oh_times_u', _lib.ptr(ore), _lib.ptr(oim), _lib.ptr(mre), _lib.ptr(mim), _lib.ptr(u), B, P,
              0.5, 0.25, _lib.ptr(vd), _lib.ptr(out), _lib.ptr(scratch), _lib.stream_ptr())
    uu = got_u
    ref = 0.5 * np.real(np.conj(Oc).T @ uu) + 0.25 * v.astype(np.float32)
    assert np.abs(out.double().cpu().numpy() - ref).max() < 1e-4 * np.abs(ref).max()


@pytest.mark.parametrize('B,P,rows', [(37, 1000, 0), (37, 1000, 1), (37, 1000, 5), (64, 2048, 32), (70, 1001, 8),
                                      (3, 4, 16)])
def test_sr_matvec_one_pass_c_abi(rng, B, P, rows):
    """fk_sr_matvec (row blocks: partial dots of one block and Re[O^H u] of the
    previous one per launch) against dense float64 algebra and the two-pass
    products; ragged last blocks, one-row blocks, and P % 4 != 0 (two-pass
    path)."""
    from flowket_amd import _lib
    gpu_banner(f'SR one-pass matvec B={B} P={P} rows={rows}')
    O = rng.normal(size=(B, P)) + 1j * rng.normal(size=(B, P))
    v = rng.normal(size=P)
    ore = torch.as_tensor(O.real, dtype=torch.float32).contiguous().cuda()
    oim = torch.as_tensor(O.imag, dtype=torch.float32).contiguous().cuda()
    mre, mim = torch.empty(P, device='cuda'), torch.empty(P, device='cuda')
    _lib.call('fk_o_mean', _lib.ptr(ore), _lib.ptr(oim), B, P, _lib.ptr(mre), _lib.ptr(mim), _lib.stream_ptr())
    scratch = torch.empty(int(_lib.load().fk_sr_scratch_doubles()), dtype=torch.float64, device='cuda')
    nbytes = int(_lib.load().fk_sr_matvec_workspace_bytes(B, P))
    ws = torch.empty(nbytes // 8, dtype=torch.float64, device='cuda')
    vd = torch.as_tensor(v, dtype=torch.float32).cuda()
    u = torch.empty((B, 2), dtype=torch.float32, device='cuda')
    outs = []
    for _ in range(2):  # the second call reuses the workspace's stale contents
        out = torch.empty(P, device='cuda')
        _lib.call('fk_sr_matvec', _lib.ptr(ore), _lib.ptr(oim), _lib.ptr(mre), _lib.ptr(mim), _lib.ptr(vd), B, P,
                  0.5, 0.25, _lib.ptr(out), _lib.ptr(u), rows, _lib.ptr(ws), nbytes, _lib.ptr(scratch),
                  _lib.stream_ptr())
        outs.append(out.double().cpu().numpy())
    assert np.array_equal(outs[0], outs[1])
    Oc = O.astype(np.complex64).astype(np.complex128)
    Oc = Oc - Oc.mean(axis=0)
    v32 = v.astype(np.float32).astype(np.float64)
    ref_u = Oc @ v32
    got_u = u[:, 0].double().cpu().numpy() + 1j * u[:, 1].double().cpu().numpy()
    assert np.abs(got_u - ref_u).max() < 1e-5 * np.abs(ref_u).max()
    ref = 0.5 * np.real(np.conj(Oc).T @ ref_u) + 0.25 * v32
    assert np.abs(outs[0] - ref).max() < 1e-5 * np.abs(ref).max()
    # the two-pass products on the same inputs
    u2 = torch.empty((B, 2), dtype=torch.float32, device='cuda')
    _lib.call('fk_o_times_v', _lib.ptr(ore), _lib.ptr(oim), _lib.ptr(mre), _lib.ptr(mim), _lib.ptr(vd), B, P,
              _lib.ptr(u2), _lib.ptr(scratch), _lib.stream_ptr())
    out2 = torch.empty(P, device='cuda')
    _lib.call('fk_oh_times_u', _lib.ptr(ore), _lib.ptr(oim), _lib.ptr(mre), _lib.ptr(mim), _lib.ptr(u2), B, P,
              0.5, 0.25, _lib.ptr(vd), _lib.ptr(out2), _lib.ptr(scratch), _lib.stream_ptr())
    assert np.abs(outs[0] - out2.double().cpu().numpy()).max() < 1e-5 * np.abs(ref).max()
    # arguments the entry point refuses
    with pytest.raises(RuntimeError):
        _lib.call('fk_sr_matvec', _lib.ptr(ore), _lib.ptr(oim), _lib.ptr(mre), _lib.ptr(mim), _lib.ptr(vd), B, P,
                  0.5, 0.25, _lib.ptr(out), _lib.ptr(u), 33, _lib.ptr(ws), nbytes, _lib.ptr(scratch),
                  _lib.stream_ptr())


def test_mini_batch_loss_gradients_add_up_to_the_energy_gradient():
    """VariationalMonteCarlo with mini_batch_size = B/4: the loss gradients of
    the four mini-batches (optimization/loss.py:4-5) sum to the energy
    gradient of the whole batch, which is what the accumulating optimizer
    applies (optimizers/accumulate_gradient_optimizer.py:13-89)."""
    gpu_banner('mini-batch gradient accumulation 4x4')
    import flowket_amd as fk
    m = fk.ConvNetAutoregressive2D((4, 4), depth=3, num_of_channels=32, activation='elu',
                                   weights_normalization=False, precision='fp32', seed=2)
    op = fk.Heisenberg(hilbert_state_shape=(4, 4), pbc=True)
    B = 256
    vmc = fk.VariationalMonteCarlo(m, op, fk.FastAutoregressiveSampler(m, B, seed=4), mini_batch_size=B // 4)
    assert vmc.update_params_frequency == 4
    parts, xs = [], []
    for _ in range(4):
        x, y = next(vmc)
        xs.append(x)
        parts.append(fk.loss_gradient(m, x, y))
    acc = torch.stack(parts).sum(0)
    full = fk.energy_gradient(m, torch.cat(xs), vmc.current_local_energy, vmc.current_energy)
    assert torch.allclose(acc, full, rtol=1e-4, atol=1e-5 * full.abs().max().item())
    # the wrapper steps once per batch on that sum
    sgd = fk.SGD(m.params, lr=1e-3, on_update=m.mark_params_updated)
    opt = fk.convert_to_accumulate_gradient_optimizer(sgd, vmc.update_params_frequency)
    before = m.params.clone()
    for x, g in zip(xs, parts):
        opt.apply_gradient(g)
    assert sgd.iterations == 1
    # (params are FP32: the difference carries their rounding, a few 1e-8)
    assert torch.allclose(before - m.params, 1e-3 * acc, rtol=1e-4, atol=2e-7)
