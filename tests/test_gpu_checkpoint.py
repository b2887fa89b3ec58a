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


def test_adam_matches_keras_rule(rng):
    """fk_adam_step over 20 steps equals the Keras Adam restatement
    (epsilon on the uncorrected sqrt(v), bias-corrected learning rate), with
    gradients over nine orders of magnitude so epsilon matters."""
    import flowket_amd as fk
    from oracle import optim as oopt
    gpu_banner('adam vs keras rule')
    m = fk.ConvNetAutoregressive2D((4, 4), depth=2, num_of_channels=8, weights_normalization=False, seed=1)
    p0 = m.params.cpu().numpy().copy()
    grads = [(rng.normal(size=m.n_params) * 10.0 ** rng.integers(-9, 1, size=m.n_params)).astype(np.float32)
             for _ in range(20)]
    opt = fk.Adam(m, lr=1e-2, epsilon=1e-7)
    for g in grads:
        opt.apply_gradient(torch.from_numpy(g))
    ref = oopt.keras_adam(p0, grads, lr=1e-2, epsilon=1e-7)
    got = m.params.cpu().numpy()
    assert np.allclose(got, ref, rtol=2e-5, atol=1e-7)
    assert opt.iterations == 20


def test_adam_resume_equals_uninterrupted(tmp_path, rng):
    """Plain Adam: 5 updates, save, load into fresh objects, 5 more: equal
    to 10 uninterrupted updates bit for bit."""
    import flowket_amd as fk
    gpu_banner('adam resume')

    def make(seed):
        m = fk.ConvNetAutoregressive2D((4, 4), depth=3, num_of_channels=8, weights_normalization=False, seed=seed)
        return m, fk.Adam(m, lr=1e-2)

    grads = [torch.from_numpy(rng.normal(size=make(0)[0].n_params).astype(np.float32)) for _ in range(10)]
    m_ref, o_ref = make(7)
    for g in grads:
        o_ref.apply_gradient(g)
    m_a, o_a = make(7)
    for g in grads[:5]:
        o_a.apply_gradient(g)
    path = str(tmp_path / 'ck')
    fk.save_weights(m_a, path)
    fk.save_optimizer_weights(o_a, path, epoch=0, batch=5)
    m_b, o_b = make(99)
    fk.load_weights(m_b, path)
    fk.load_optimizer_weights(o_b, path)
    for g in grads[5:]:
        o_b.apply_gradient(g)
    assert torch.equal(m_b.params, m_ref.params)
