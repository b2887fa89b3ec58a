"""Symmetrised wave functions on the GPU (fk_transform_spins,
fk_ensemble_combine and the relabelled connected-configuration forward)
against the oracle restatement, and the trained 12x12 Ising network of the
reference's experiments against its published evaluation."""
import os

import numpy as np
import pytest
import torch

from oracle import convnet as ocn
from oracle import ensemble as oen
from tests.conftest import gpu_banner
from tests.test_gpu_convnet import _machine, _oracle_params

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'ising_12x12_h3_weights.npz')


def test_transform_spins(rng):
    import flowket_amd as fk
    from flowket_amd.ensemble import _rot90_src
    gpu_banner('transform spins 7x7')
    x = (rng.integers(0, 2, size=(300, 7, 7)) * 2 - 1).astype(np.int8)
    m = fk.make_up_down_invariant(fk.make_2d_obc_invariants(fk.RBM((7, 7), alpha=1)))
    inner = m.model
    xs = torch.from_numpy(x.reshape(300, 49)).cuda()
    imgs = oen.d4_images(x)
    for k in range(8):
        assert np.array_equal(inner._transform(xs, k, None).cpu().numpy(), imgs[k].reshape(300, 49))
        assert np.array_equal(inner.members[k][0], _rot90_src(7, 7, k % 4, k >= 4))
    assert np.array_equal(m._transform(xs, 1, None).cpu().numpy(), -x.reshape(300, 49))


def test_symmetrised_log_psi_vs_oracle(rng):
    """up/down(obc(ConvNetAutoregressive2D)) on 5x5 (FP32): log psi of every
    configuration against the oracle's nested ensemble in FP64."""
    import flowket_amd as fk
    gpu_banner('symmetrised log psi 5x5')
    m = _machine((5, 5), 4, 16, 'elu', 'fp32', seed=2, bias_scale=0.3)
    sym = fk.make_up_down_invariant(fk.make_2d_obc_invariants(m))
    x = (rng.integers(0, 2, size=(512, 5, 5)) * 2 - 1).astype(np.int8)
    got = sym.log_psi(torch.from_numpy(x)).cpu().numpy()
    params = _oracle_params(m)
    ref = oen.evaluation_log_psi(lambda y: ocn.log_psi(params, y, 4, activation='elu'), x.astype(np.float64))
    assert np.abs(got.real - ref.real).max() < 2e-5 * max(1.0, np.abs(ref.real).max())
    assert np.abs(np.angle(np.exp(1j * (got.imag - ref.imag)))).max() < 2e-5


@pytest.mark.parametrize('name', ['heisenberg', 'ising', 'j1j2'])
def test_symmetrised_local_energy_relabelled_vs_materialised(name, rng):
    """E_loc of the symmetrised network through the relabelled connection
    tables (no connected configurations stored) equals E_loc through the
    materialised connected configurations (balanced layout)."""
    import flowket_amd as fk
    gpu_banner(f'symmetrised E_loc 4x4 {name}')
    shape = (4, 4)
    op = {'heisenberg': lambda: fk.Heisenberg(hilbert_state_shape=shape, pbc=True),
          'ising': lambda: fk.Ising(h=3.0, hilbert_state_shape=shape, pbc=False),
          'j1j2': lambda: fk.J1J2(hilbert_state_shape=shape, j2=0.5, pbc=False)}[name]()
    m = _machine(shape, 3, 16, 'relu', 'fp32', seed=5, bias_scale=0.2)
    sym = fk.make_up_down_invariant(fk.make_2d_obc_invariants(m))
    x = next(fk.AutoregressiveSampler(m, 256, seed=7))
    obs = fk.Observable(op)
    a = obs.local_values(sym, x).cpu().numpy()
    b = obs.local_values_balanced(sym, obs._spins(x)).cpu().numpy()
    assert np.abs(a - b).max() < 1e-5 * max(1.0, np.abs(b).max())


# experiments/README.md:38-47 (2^15 samples): field -> (energy, local-energy
# variance, |Mz|); experiments/ising_runner.py:26-31 gives the DMRG energies
PUBLISHED = {'2': (-346.9817926, 0.00125, 0.783929189, -346.982),
             '2_5': (-395.6618438, 0.00421, 0.57235633, -395.654),
             '3': (-457.0420317, 0.000821, 0.1622390747, -457.039),
             '3_5': (-524.5172088, 0.000647, 0.1106881036, -524.510),
             '4': (-593.5389339, 0.000777, 0.09679073758, -593.536)}


def _trained_ising(precision, h='3'):
    import flowket_amd as fk
    w = np.load(FIXTURE.replace('_h3_', '_h%s_' % h))
    m = fk.ConvNetAutoregressive2D((12, 12), depth=10, num_of_channels=32, precision=precision,
                                   params=torch.from_numpy(w['params']))
    m.set_wn_g(torch.from_numpy(w['g']))
    return m


@pytest.mark.parametrize('h', sorted(PUBLISHED))
def test_trained_ising_12x12_reproduces_the_published_evaluation(h):
    """experiments/README.md:38-47: each symmetrised trained network evaluated
    on samples of its base network (experiments/run_evaluation.py) gives the
    published energy, local-energy variance and |Mz| (2^15 samples there).
    2048 samples here (FP32): the energy agrees within 5 combined standard
    errors, the variance within a factor 2.5 and |Mz| within 0.017 (5
    standard errors of 2048 samples for a per-sample spread below 0.15)."""
    import flowket_amd as fk
    from flowket_amd.evaluation import evaluate
    gpu_banner(f'trained Ising 12x12 h={h} evaluation')
    e_pub, var_pub, mz_pub, e_dmrg = PUBLISHED[h]
    m = _trained_ising('fp32', h)
    op = fk.Ising(hilbert_state_shape=(12, 12), pbc=False, h=float(h.replace('_', '.')))
    sym = fk.make_up_down_invariant(fk.make_2d_obc_invariants(m))
    B, steps = 1024, 2
    vmc = fk.VariationalMonteCarlo(sym, op, fk.FastAutoregressiveSampler(m, B, seed=11))
    r = evaluate(vmc, steps, true_ground_state_energy=e_dmrg)
    e, var = r['energy/energy'], r['energy/local_energy_variance']
    se = np.sqrt(var / (B * steps) + var_pub / 2 ** 15)
    assert abs(e - e_pub) < 5 * se + 1e-4, (e, var)
    assert var_pub / 2.5 < var < var_pub * 2.5, var
    assert abs(r['energy/relative_error']) < 5e-5
    assert abs(r['observables/abs_sigma_z'] - mz_pub) < 0.017, r


def test_ensemble_combine_through_the_c_abi(rng):
    """fk_ensemble_combine on K = 3 blocks with a device row count: the live
    rows equal the oracle's probabilistic ensemble (FP64 inside, FP32 out),
    rows past *m_dev are not written, and one member gives itself back."""
    from flowket_amd import _lib
    gpu_banner('ensemble combine C-ABI')
    K, m, live = 3, 1000, 777
    x = np.stack([rng.normal(size=(K, m)) * 5, rng.uniform(-4, 4, size=(K, m))], axis=-1).astype(np.float32)
    lp = torch.from_numpy(x).cuda().contiguous()
    out = torch.full((m, 2), 123.0, dtype=torch.float32, device='cuda')
    mdev = torch.tensor([live], dtype=torch.int64, device='cuda')
    _lib.call('fk_ensemble_combine', _lib.ptr(lp), K, m, m, _lib.ptr(mdev), _lib.ptr(out), _lib.stream_ptr())
    got = out.cpu().numpy()
    z = x[..., 0].astype(np.float64) + 1j * x[..., 1].astype(np.float64)
    ref = oen.probabilistic_ensemble(z.T)
    assert np.abs(got[:live, 0] - ref.real[:live]).max() < 1e-5 * max(1.0, np.abs(ref.real).max())
    assert np.abs(np.angle(np.exp(1j * (got[:live, 1] - ref.imag[:live])))).max() < 1e-5
    assert (got[live:] == 123.0).all()
    one = torch.empty((m, 2), dtype=torch.float32, device='cuda')
    _lib.call('fk_ensemble_combine', _lib.ptr(lp), 1, m, m, None, _lib.ptr(one), _lib.stream_ptr())
    o = one.cpu().numpy()
    assert np.abs(o[:, 0] - x[0, :, 0]).max() < 1e-5 * np.abs(x[0, :, 0]).max()
    assert np.abs(np.angle(np.exp(1j * (o[:, 1] - x[0, :, 1])))).max() < 1e-5


def test_pbc_invariant_log_psi_vs_oracle(rng):
    """make_pbc_invariants (16 translations around the 8 dihedral members) of
    a 4x4 FP32 network against the oracle, and its connected-configuration
    E_loc through relabelled tables against the materialised layout."""
    import flowket_amd as fk
    gpu_banner('pbc-invariant log psi 4x4')
    m = _machine((4, 4), 3, 16, 'elu', 'fp32', seed=6, bias_scale=0.3)
    sym = fk.make_pbc_invariants(m)
    assert len(sym.members) == 16 and len(sym.model.members) == 8
    x = (rng.integers(0, 2, size=(64, 4, 4)) * 2 - 1).astype(np.int8)
    got = sym.log_psi(torch.from_numpy(x)).cpu().numpy()
    params = _oracle_params(m)
    ref = oen.pbc_invariant_log_psi(lambda y: ocn.log_psi(params, y, 3, activation='elu'), x.astype(np.float64))
    assert np.abs(got.real - ref.real).max() < 2e-5 * max(1.0, np.abs(ref.real).max())
    assert np.abs(np.angle(np.exp(1j * (got.imag - ref.imag)))).max() < 2e-5
    op = fk.Heisenberg(hilbert_state_shape=(4, 4), pbc=True)
    obs = fk.Observable(op)
    xs = torch.from_numpy(x[:32])
    a = obs.local_values(sym, xs).cpu().numpy()
    b = obs.local_values_balanced(sym, obs._spins(xs)).cpu().numpy()
    assert np.abs(a - b).max() < 1e-5 * max(1.0, np.abs(b).max())
