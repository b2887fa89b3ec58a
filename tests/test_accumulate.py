"""Gradient accumulation (optimizers/accumulate_gradient_optimizer.py:13-89)
with the known answers of the reference's tests/test_gradients_aggregation.py:
a one-weight linear model w = 1 whose loss mean(w x) has gradient x, SGD with
lr 1.  CPU only (host-side optimizer logic)."""
import pytest
import torch

import flowket_amd as fk


def _linear_model(update_params_frequency, accumulate_sum_or_mean, ema_decay=0):
    """tests/test_gradients_aggregation.py:13-34."""
    w = torch.ones(1, dtype=torch.float64)
    sgd = fk.SGD(w, lr=1.0)
    opt = fk.convert_to_accumulate_gradient_optimizer(sgd, update_params_frequency=update_params_frequency,
                                                      accumulate_sum_or_mean=accumulate_sum_or_mean,
                                                      ema_decay=ema_decay)

    def fit(x):
        opt.apply_gradient(torch.tensor([x], dtype=torch.float64))

    return opt, sgd, w, fit


def test_update_just_when_need():
    """tests/test_gradients_aggregation.py:37-51."""
    opt, sgd, w, fit = _linear_model(2, False)
    assert w.item() == 1.0
    fit(2.0)
    assert w.item() == 1.0 and sgd.iterations == 0
    fit(3.0)
    assert w.item() == -1.5 and sgd.iterations == 1


def test_changing_the_update_frequency():
    """tests/test_gradients_aggregation.py:54-73."""
    opt, sgd, w, fit = _linear_model(1, False)
    opt.set_update_params_frequency(2)
    fit(2.0)
    assert w.item() == 1.0 and sgd.iterations == 0
    fit(3.0)
    assert w.item() == -1.5 and sgd.iterations == 1
    opt.set_update_params_frequency(1)
    fit(1.0)
    assert w.item() == -2.5


def test_reset_after_update():
    """tests/test_gradients_aggregation.py:76-81."""
    opt, sgd, w, fit = _linear_model(1, False)
    fit(2.0)
    fit(3.0)
    assert w.item() == -4.0


def test_update_ema_just_when_need():
    """tests/test_gradients_aggregation.py:84-99."""
    opt, sgd, w, fit = _linear_model(2, False, .9)
    fit(2.0)
    assert w.item() == 1.0 and sgd.iterations == 0
    fit(3.0)
    assert w.item() == -1.5 and sgd.iterations == 1
    opt.set_weights_ema()
    assert w.item() == pytest.approx(-1.5)


def test_ema():
    """tests/test_gradients_aggregation.py:102-111."""
    opt, sgd, w, fit = _linear_model(1, False, .9)
    fit(2.0)
    assert w.item() == -1.0
    fit(3.0)
    assert w.item() == -4.0
    opt.set_weights_ema()
    assert w.item() == pytest.approx(-2.5789473684210527)


def test_sum_mode_adds_unscaled():
    """accumulate_sum_or_mean=True adds the gradients as they are."""
    opt, sgd, w, fit = _linear_model(2, True)
    fit(2.0)
    fit(3.0)
    assert w.item() == 1.0 - 5.0


def test_frequency_must_be_positive():
    with pytest.raises(ValueError):
        fk.convert_to_accumulate_gradient_optimizer(fk.SGD(torch.ones(1)), 0)
