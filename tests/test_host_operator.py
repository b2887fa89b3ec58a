"""Operators given only by a numpy find_conn (operators/operator.py:14-21):
the device layout Operator.find_conn_device builds from it (CPU tensors)."""
import numpy as np
import pytest
import torch

import flowket_amd as fk


class _Field(fk.Operator):
    """sum_i (a sigma^x_i + b sigma^y_i) + d: connection 0 is the sample
    (diagonal d), connection 1 + i the sample with site i flipped (matrix
    element a - i b s_i); the last slot is unused."""

    def __init__(self, shape, a, b, d=0.5, break_first=False):
        super().__init__(shape)
        self.a, self.b, self.d, self.break_first = a, b, d, break_first

    def find_conn(self, sample):
        B = sample.shape[0]
        x = sample.reshape(B, -1)
        n = x.shape[1]
        conn = np.repeat(x[None], n + 2, axis=0).astype(np.float64)
        for i in range(n):
            conn[1 + i, :, i] *= -1
        if self.break_first:
            conn[0, :, 0] *= -1
        mel = np.zeros((n + 2, B), dtype=np.complex128)
        mel[0] = self.d
        mel[1:n + 1] = self.a - 1j * self.b * x.T
        mel[n + 1] = 7.0          # unused slot: must not reach the kernels
        use = np.ones((n + 2, B), dtype=bool)
        use[n + 1] = False
        return conn.reshape((n + 2, B) + sample.shape[1:]), mel, use


def test_find_conn_device_layout(rng):
    shape = (2, 3)
    x = rng.choice([-1, 1], size=(5,) + shape).astype(np.int8)
    spins = torch.as_tensor(x.reshape(5, -1))
    for b_coef in (0.0, 0.5):
        op = _Field(shape, 0.25, b_coef)
        r = op.find_conn_device(spins)
        conn, mel, use = op.find_conn(x)
        assert r['conn'].dtype == torch.int8 and r['conn'].shape == (8, 5, 6)
        assert np.array_equal(r['conn'].numpy(), conn.reshape(8, 5, 6).astype(np.int8))
        assert np.array_equal(r['use'].numpy(), use.astype(np.uint8))
        assert np.array_equal(r['mel'].numpy(), np.where(use, mel.real, 0.0))
        if b_coef == 0.0:
            assert 'mel_im' not in r          # real operators carry no imaginary plane
        else:
            assert np.array_equal(r['mel_im'].numpy(), np.where(use, mel.imag, 0.0))
        assert op.n_sites == 6


def test_find_conn_device_requires_the_sample_first(rng):
    x = rng.choice([-1, 1], size=(3, 2, 2)).astype(np.int8)
    with pytest.raises(ValueError, match='connection 0'):
        _Field((2, 2), 1.0, 0.0, break_first=True).find_conn_device(torch.as_tensor(x.reshape(3, -1)))
