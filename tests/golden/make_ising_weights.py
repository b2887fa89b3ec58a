"""Writes ising_12x12_h{2,2_5,3,3_5,4}_weights.npz: the trained 12x12
open-boundary transverse-field Ising ConvNetAutoregressive2D networks (depth
10, 32 channels, weight normalisation) of the reference's experiments, one
per field h, read from its HDF5 weight files with this package's own parser
(the files are data; nothing of the reference is run).  ``params`` is the
flat kernel/bias vector in network order, ``g`` the weight-normalisation g
vector."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from flowket_amd.keras_weights import convnet_params_from_keras, load_keras_weights  # noqa: E402
from flowket_amd.machines import layer_shapes  # noqa: E402

SRC = '/root/reference/experiments/weights/ising_%s.h5'
FIELDS = ['2', '2_5', '3', '3_5', '4']


def main():
    for h in FIELDS:
        params, g = convnet_params_from_keras(load_keras_weights(SRC % h), layer_shapes(10, 32), True)
        np.savez_compressed(os.path.join(HERE, 'ising_12x12_h%s_weights.npz' % h), params=params, g=g)
        print(h, params.size, g.size)


if __name__ == '__main__':
    main()
