"""Fast autoregressive sampler vs the plain sampler and the oracle rule."""
import numpy as np
import pytest
import torch

from oracle import convnet as ocn
from oracle import samplers as osa
from tests.conftest import gpu_banner
from tests.test_gpu_convnet import _machine, _oracle_params

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('shape,depth,C,act,prec,B', [((6, 5), 3, 32, 'elu', 'bf16', 70),
                                                     ((4, 4), 2, 32, 'relu', 'fp32', 33),
                                                     ((10, 10), 4, 32, 'relu', 'bf16', 48),
                                                     ((5, 7), 3, 64, 'elu', 'fp32', 20),
                                                     ((5, 7), 3, 64, 'elu', 'bf16', 20),
                                                     ((6, 6), 3, 48, 'relu', 'bf16', 40),
                                                     ((4, 5), 2, 24, 'elu', 'fp32', 30),
                                                     ((10, 10), 3, 32, 'relu', 'fp16', 40),
                                                     ((5, 6), 2, 64, 'elu', 'fp16', 17),
                                                     # rows of 11-16 pixels (two staging batches, > 64 KB
                                                     # of LDS), 17 pixels (per-pixel vertical pass), depth 20
                                                     ((4, 13), 3, 32, 'elu', 'fp16', 20),
                                                     ((3, 16), 4, 32, 'relu', 'bf16', 17),
                                                     ((3, 17), 3, 32, 'relu', 'bf16', 16),
                                                     ((10, 10), 20, 32, 'relu', 'fp16', 32)])
def test_fast_equals_plain_given_uniforms(rng, shape, depth, C, act, prec, B):
    """Same uniforms -> the same draws as AutoregressiveSampler (identical
    arithmetic), and the conditionals it drew from equal the full forward's."""
    import flowket_amd as fk
    gpu_banner(f'fast sampler shape={shape} depth={depth} C={C} {act} {prec} B={B}')
    m = _machine(shape, depth, C, act, prec, bias_scale=0.4)
    u = rng.random((B,) + shape)
    plain = next(fk.AutoregressiveSampler(m, B, uniforms=u))
    fs = fk.FastAutoregressiveSampler(m, B, uniforms=u)
    fs.keep_cond_logp = True
    fast = next(fs)
    assert torch.equal(plain, fast)
    full = m.conditional_log_probs(fast)
    got = fs.last_cond_logp.reshape(full.shape)
    assert torch.equal(got, full)


def test_fast_matches_oracle_rule_fp32(rng):
    import flowket_amd as fk
    gpu_banner('fast sampler fp32 vs oracle rule 5x4')
    m = _machine((5, 4), 3, 32, 'elu', 'fp32', bias_scale=0.5)
    B = 128
    u = rng.random((B, 5, 4))
    got = next(fk.FastAutoregressiveSampler(m, B, uniforms=u)).cpu().numpy().astype(np.float64)
    params = _oracle_params(m)
    ref = osa.ar_sample(lambda x: ocn.conditional_log_probs(params, x, 3, activation='elu'), u)
    lp = ocn.conditional_log_probs(params, ref, 3, activation='elu')
    margin = np.abs(np.exp(lp[..., 0]) - u).reshape(B, -1).min(axis=1)
    ok = margin > 1e-6
    assert ok.mean() > 0.9
    assert np.array_equal(got[ok], ref[ok])


def test_fast_sampler_chunks(rng):
    """mini_batch_size < batch_size draws in chunks with the same result."""
    import flowket_amd as fk
    gpu_banner('fast sampler chunked')
    m = _machine((4, 6), 2, 32, 'relu', 'bf16')
    u = rng.random((50, 4, 6))
    a = next(fk.FastAutoregressiveSampler(m, 50, uniforms=u))
    b = next(fk.FastAutoregressiveSampler(m, 50, mini_batch_size=16, uniforms=u))
    assert torch.equal(a, b)
