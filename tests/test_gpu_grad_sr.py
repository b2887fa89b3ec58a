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
    x, it, res = sr.solve(s, el)
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
