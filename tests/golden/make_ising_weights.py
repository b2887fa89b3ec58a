"""Writes ising_12x12_h3_weights.npz: the trained 12x12 open-boundary
transverse-field Ising (h = 3) ConvNetAutoregressive2D (depth 10, 32
channels, weight normalisation) of the reference's experiments, read from its
HDF5 weight file with this package's own parser (the file is data; nothing of
the reference is run).  ``params`` is the flat kernel/bias vector in network
order, ``g`` the weight-normalisation g vector."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from flowket_amd.keras_weights import convnet_params_from_keras, load_keras_weights  # noqa: E402
from flowket_amd.machines import layer_shapes  # noqa: E402

SRC = '/root/reference/experiments/weights/ising_3.h5'


def main():
    params, g = convnet_params_from_keras(load_keras_weights(SRC), layer_shapes(10, 32), True)
    np.savez_compressed(os.path.join(HERE, 'ising_12x12_h3_weights.npz'), params=params, g=g)
    print(params.size, g.size)


if __name__ == '__main__':
    main()
