"""Complex RBM restated (machines/rbm.py:35-48, layers/complex/dense.py:11-50,
layers/complex/base_layer.py:8-35, layers/complex/tensorflow_ops.py:69-85)."""
import numpy as np


def lncosh(z):
    """tensorflow_ops.py:79-85: |Re z| - log 2 + log(e^{z-|Re z|} + e^{-z-|Re z|})
    with the complex log taken as log|w| + i*atan2(Im w, Re w)."""
    z = np.asarray(z, dtype=np.complex128)
    a = np.abs(z.real)
    w = np.exp(z - a) + np.exp(-z - a)
    return a - np.log(2.0) + (np.log(np.abs(w)) + 1j * np.arctan2(w.imag, w.real))


def log_psi(spins, w_re, w_im, b_re, b_im, a_re, a_im):
    """RBM.build_predictions (rbm.py:36-48).  Each complex weight is stored as
    (real, imag) and used as real - 1j*imag (base_layer.py:30-35).
    Shapes: W (N, M), b (M,), a (N,).  Returns complex128 (B,)."""
    x = np.asarray(spins, dtype=np.float64).reshape(len(spins), -1)
    W = w_re - 1j * w_im
    b = b_re - 1j * b_im
    a = a_re - 1j * a_im
    theta = x @ W + b
    return x @ a + lncosh(theta).sum(axis=1)
