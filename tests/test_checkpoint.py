"""Checkpoint and resume (callbacks/checkpoint.py:9-73) on the CPU: weight and
optimizer-state round trips, resumed updates identical to uninterrupted ones,
and the time-based save schedule."""
import os

import numpy as np
import pytest
import torch


def _machine(seed=0, wn=True):
    import flowket_amd as fk
    return fk.ConvNetAutoregressive2D((4, 4), depth=3, num_of_channels=8, weights_normalization=wn, seed=seed,
                                      device='cpu')


def _grads(n, P, seed=1):
    g = torch.Generator().manual_seed(seed)
    return [torch.randn(P, generator=g) for _ in range(n)]


def test_weights_round_trip(tmp_path):
    m = _machine(seed=3)
    path = str(tmp_path / 'w.safetensors')
    m.save_weights(path)
    m2 = _machine(seed=4)
    assert not torch.equal(m2.params, m.params)
    m2.load_weights(path)
    assert torch.equal(m2.params, m.params) and torch.equal(m2.wn_g, m.wn_g)
    assert torch.equal(m2.effective_params(), m.effective_params())
    with pytest.raises(ValueError):
        _machine(wn=False).load_weights(path)
    import flowket_amd as fk
    with pytest.raises(ValueError):
        fk.ConvNetAutoregressive2D((4, 4), depth=2, num_of_channels=8, device='cpu').load_weights(path)


def test_adam_state_round_trip(tmp_path):
    """Adam's state (iterations, moments) survives save + load into a fresh
    optimizer (the updates themselves run on the GPU: tests/test_gpu_checkpoint.py)."""
    import flowket_amd as fk
    m = _machine()
    opt = fk.Adam(m, lr=1e-2)
    g = torch.Generator().manual_seed(2)
    opt.set_weights([torch.tensor([5.0], dtype=torch.float64), torch.randn(m.n_trainable, generator=g),
                     torch.rand(m.n_trainable, generator=g)])
    path = str(tmp_path / 'ck')
    fk.save_optimizer_weights(opt, path, epoch=1, batch=2)
    opt2 = fk.Adam(_machine(seed=5))
    assert fk.load_optimizer_weights(opt2, path) == (1, 2)
    assert opt2.iterations == 5
    for a, b in zip(opt.get_weights(), opt2.get_weights()):
        assert torch.equal(a, b)
    with pytest.raises(fk.FlowketHipError):
        opt2.apply_gradient(torch.zeros(m.n_trainable))


@pytest.mark.parametrize('wrap', ['accumulate_sgd'])
def test_resumed_updates_equal_uninterrupted(tmp_path, wrap):
    """k updates, save, fresh machine and optimizer, load, the remaining
    updates: parameters equal those of an uninterrupted run bit for bit
    (host-side SGD; the Adam variants run on the GPU)."""
    import flowket_amd as fk

    def make(seed):
        m = _machine(seed=seed)
        if wrap == 'adam':
            return m, fk.Adam(m, lr=1e-2)
        if wrap == 'accumulate_adam_ema':
            return m, fk.convert_to_accumulate_gradient_optimizer(fk.Adam(m, lr=1e-2), update_params_frequency=3,
                                                                  accumulate_sum_or_mean=False, ema_decay=0.9)
        return m, fk.convert_to_accumulate_gradient_optimizer(
            fk.SGD(m.params, lr=0.05, on_update=m.mark_params_updated), update_params_frequency=2)

    m_ref, opt_ref = make(7)
    grads = _grads(10, m_ref.n_params)
    for g in grads:
        opt_ref.apply_gradient(g)
    m_a, opt_a = make(7)
    for g in grads[:5]:
        opt_a.apply_gradient(g)
    path = str(tmp_path / 'ck')
    fk.save_weights(m_a, path)
    fk.save_optimizer_weights(opt_a, path, epoch=2, batch=5)
    m_b, opt_b = make(99)
    fk.load_weights(m_b, path)
    assert fk.load_optimizer_weights(opt_b, path) == (2, 5)
    for g in grads[5:]:
        opt_b.apply_gradient(g)
    assert torch.equal(m_b.params, m_ref.params)
    if wrap == 'accumulate_adam_ema':
        opt_b.set_weights_ema()
        opt_ref.set_weights_ema()
        assert torch.equal(m_b.params, m_ref.params)


def test_missing_optimizer_state_returns_none(tmp_path):
    import flowket_amd as fk
    m = _machine()
    assert fk.load_optimizer_weights(fk.Adam(m), str(tmp_path / 'none')) is None


def test_checkpoint_by_time_schedule(tmp_path):
    """Saves when save_frequency_in_minutes have passed since the last save
    (checked at each batch end) and once at train end with epoch + 1
    (checkpoint.py:58-73); the path is formatted from the logs."""
    import flowket_amd as fk
    now = [0.0]
    m = _machine()
    opt = fk.Adam(m)
    cb = fk.CheckpointByTime(str(tmp_path / 'ck_{tag}'), m, opt, save_frequency_in_minutes=1, clock=lambda: now[0])
    cb.on_epoch_begin(0)
    for batch, t in enumerate([10.0, 59.0, 61.0, 100.0, 125.0]):
        now[0] = t
        cb.on_batch_end(batch, {'tag': 'b%d' % batch})
    cb.on_train_end({'tag': 'end'})
    assert [(os.path.basename(p), e, b) for p, e, b in cb.saved] == [('ck_b2', 0, 2), ('ck_b4', 0, 4),
                                                                            ('ck_end', 1, 0)]
    assert os.path.exists(str(tmp_path / 'ck_end_optimizer.safetensors'))


def test_safetensors_files_hold_no_code(tmp_path):
    """The optimizer file is tensors plus string metadata."""
    import flowket_amd as fk
    m = _machine()
    opt = fk.Adam(m)
    opt.set_weights([torch.tensor([1.0], dtype=torch.float64), torch.ones(m.n_trainable),
                     torch.ones(m.n_trainable)])
    fk.save_optimizer_weights(opt, str(tmp_path / 'x'), 0, 0)
    from safetensors import safe_open
    with safe_open(str(tmp_path / 'x_optimizer.safetensors'), framework='pt') as f:
        assert sorted(f.keys()) == ['w0000', 'w0001', 'w0002']
        assert f.metadata()['count'] == '3'
        assert np.array_equal(f.get_tensor('w0000').numpy(), [1.0])


def test_rbm_weights_round_trip(tmp_path):
    import flowket_amd as fk
    a = fk.RBM((3, 4), alpha=2, stddev=0.3, seed=1, device='cpu')
    b = fk.RBM((3, 4), alpha=2, stddev=0.3, seed=2, device='cpu')
    path = str(tmp_path / 'rbm.safetensors')
    a.save_weights(path)
    b.load_weights(path)
    assert torch.equal(a.params, b.params)
    with pytest.raises(ValueError):
        fk.RBM((3, 4), alpha=1, device='cpu').load_weights(path)
