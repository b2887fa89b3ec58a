"""Keras HDF5 weight import (experiments/run_evaluation.py:13-14,
experiments/weights/*.h5) on the CPU."""
import os

import numpy as np
import pytest
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, 'golden', 'ising_12x12_h3_weights.npz')
H5 = '/root/reference/experiments/weights/ising_3.h5'


def test_fixture_has_the_trained_network():
    """398,340 values: 90 weight-normalised layers (kernel, bias, g) and the
    1x1 head (kernel, bias) of depth 10, 32 channels."""
    from flowket_amd.machines import layer_shapes
    w = np.load(FIXTURE)
    shapes = layer_shapes(10, 32)
    assert w['params'].size == sum(int(np.prod(k)) + int(np.prod(b)) for k, b in shapes)
    assert w['g'].size == sum(k[3] for k, _ in shapes[:-1])
    assert w['params'].size + w['g'].size == 398340
    assert np.isfinite(w['params']).all() and np.isfinite(w['g']).all()


def test_creation_order():
    from flowket_amd.keras_weights import _creation_index
    names = ['conv2d_90', 'weight_normalization_86', 'weight_normalization', 'weight_normalization_85']
    assert sorted(names, key=_creation_index) == ['weight_normalization', 'weight_normalization_85',
                                                   'weight_normalization_86', 'conv2d_90']


@pytest.mark.skipif(not os.path.exists(H5), reason='the reference weight files are not present')
def test_load_weights_matches_fixture():
    """load_weights of the HDF5 file: 272 datasets in Keras' save order, and
    the machine's parameters equal the committed fixture."""
    import flowket_amd as fk
    w = fk.load_keras_weights(H5)
    assert len(w) == 272 and sum(a.size for _, a in w) == 398340
    # Keras saves the last block's hconv before its vconv (topological order)
    names = [n.split('/')[0] for n, _ in w]
    assert names.index('weight_normalization_86') < names.index('weight_normalization_85')
    for h in ['2', '2_5', '3', '3_5', '4']:
        m = fk.ConvNetAutoregressive2D((12, 12), depth=10, num_of_channels=32, device='cpu')
        m.load_weights(H5.replace('ising_3', 'ising_' + h))
        fx = np.load(FIXTURE.replace('_h3_', '_h%s_' % h))
        assert np.array_equal(m.params.numpy(), fx['params'])
        assert np.array_equal(m.wn_g.numpy(), fx['g'])


@pytest.mark.skipif(not os.path.exists(H5), reason='the reference weight files are not present')
def test_load_weights_rejects_a_different_network():
    import flowket_amd as fk
    with pytest.raises(ValueError):
        fk.ConvNetAutoregressive2D((12, 12), depth=8, num_of_channels=32, device='cpu').load_weights(H5)
    with pytest.raises(ValueError):
        fk.ConvNetAutoregressive2D((12, 12), depth=10, num_of_channels=16, device='cpu').load_weights(H5)
    with pytest.raises(ValueError):
        fk.ConvNetAutoregressive2D((12, 12), depth=10, num_of_channels=32, weights_normalization=False,
                                   device='cpu').load_weights(H5)


def test_parser_rejects_other_files(tmp_path):
    from flowket_amd.keras_weights import load_keras_weights
    p = tmp_path / 'x.h5'
    p.write_bytes(b'not hdf5' * 16)
    with pytest.raises(ValueError):
        load_keras_weights(str(p))


def test_effective_kernel_of_the_loaded_weights():
    """Weight normalisation of the loaded network: every output channel of a
    wrapped kernel has L2 norm exp(g) (deepar/layers/wrappers.py:123-135)."""
    import flowket_amd as fk
    fx = np.load(FIXTURE)
    m = fk.ConvNetAutoregressive2D((12, 12), depth=10, num_of_channels=32, device='cpu',
                                   params=torch.from_numpy(fx['params']))
    m.set_wn_g(torch.from_numpy(fx['g']))
    eff = m.effective_params().double().numpy()
    kshape = m.shapes[0][0]
    k0 = eff[:int(np.prod(kshape))].reshape(kshape)
    assert np.allclose(np.sqrt((k0 ** 2).sum(axis=(0, 1, 2))), np.exp(fx['g'][:kshape[3]]), rtol=1e-5)
