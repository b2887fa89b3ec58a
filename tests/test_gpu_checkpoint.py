"""Checkpoint and resume on the GPU (callbacks/checkpoint.py:9-73): training
through the exact enumeration path (no sampling noise) interrupted by a
save and a load into fresh objects ends with the parameters of an
uninterrupted run, bit for bit; and the magnetisation observables."""
import numpy as np
import pytest
import torch

from tests.conftest import gpu_banner

pytestmark = pytest.mark.gpu


def _train(m, opt, ev, n_updates):
    gen = iter(ev)
    for _ in range(n_updates * ev.num_of_batch_until_full_cycle):
        x, c = next(gen)
        opt.step(m, x, c)


def test_resume_equals_uninterrupted(tmp_path):
    import flowket_amd as fk
    gpu_banner('checkpoint resume 4x4 exact path')
    op = fk.Heisenberg(hilbert_state_shape=(4, 4), pbc=True)

    def make(seed):
        m = fk.ConvNetAutoregressive2D((4, 4), depth=3, num_of_channels=16, activation='elu',
                                       weights_normalization=False, precision='fp32', seed=seed)
        ev = fk.ExactVariational(m, op, batch_size=2 ** 14)
        opt = fk.convert_to_accumulate_gradient_optimizer(fk.Adam(m, lr=1e-2), ev.num_of_batch_until_full_cycle,
                                                          ema_decay=0.5)
        return m, ev, opt

    m_ref, ev_ref, opt_ref = make(3)
    _train(m_ref, opt_ref, ev_ref, 4)
    m_a, ev_a, opt_a = make(3)
    _train(m_a, opt_a, ev_a, 2)
    path = str(tmp_path / 'ck')
    cb = fk.CheckpointByTime(path, m_a, opt_a, save_frequency_in_minutes=0)
    cb.on_batch_end(7)
    m_b, ev_b, opt_b = make(11)
    m_b.load_weights(path)
    assert fk.load_optimizer_weights(opt_b, path) == (0, 7)
    _train(m_b, opt_b, ev_b, 2)
    assert torch.equal(m_b.params, m_ref.params)
    ev_b.machine_updated()
    ev_ref.machine_updated()
    assert ev_b.energy_observable.current_energy == ev_ref.energy_observable.current_energy


def test_sigma_z_observables(rng):
    import flowket_amd as fk
    gpu_banner('sigma_z observables')
    for shape in [(12, 12), (7, 9), (2, 3)]:
        x = (rng.integers(0, 2, size=(1000,) + shape) * 2 - 1).astype(np.int8)
        n = int(np.prod(shape))
        s = fk.SigmaZ().local_values(None, torch.from_numpy(x)).cpu().numpy()
        a = fk.AbsSigmaZ().local_values(None, torch.from_numpy(x)).cpu().numpy()
        ref = x.reshape(1000, -1).sum(axis=1) / n
        assert np.array_equal(s.real, ref) and np.array_equal(a.real, np.abs(ref))
        assert not s.imag.any()
        mean, var, _ = fk.AbsSigmaZ().estimate(None, torch.from_numpy(x))
        assert mean.real == pytest.approx(np.abs(ref).mean(), abs=1e-14)
