"""Keras Adam restated in numpy float32 -- TEST INFRASTRUCTURE ONLY.

The reference trains with ``tf.keras.optimizers.Adam`` (examples/
basic_autoregressive_2d.py:24; TensorFlow 1.10-1.14 is a third-party
dependency absent here).  Its published update (Keras ``Adam.get_updates``):
    t += 1
    lr_t = lr * sqrt(1 - beta_2^t) / (1 - beta_1^t)
    m = beta_1 * m + (1 - beta_1) * g
    v = beta_2 * v + (1 - beta_2) * g^2
    p = p - lr_t * m / (sqrt(v) + epsilon)
with epsilon added to the uncorrected sqrt(v) (torch's Adam adds it after the
bias correction instead).  Pinned by the closed form of the first step in
tests/test_oracle.py.
"""
import numpy as np


def keras_adam(params, grads, lr=1e-3, beta_1=0.9, beta_2=0.999, epsilon=1e-7):
    """Parameters after one Keras Adam update per gradient in ``grads``."""
    f = np.float32
    p = np.asarray(params, dtype=f).copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for t, g in enumerate(grads, start=1):
        g = np.asarray(g, dtype=f)
        lr_t = f(lr * np.sqrt(1.0 - beta_2 ** t) / (1.0 - beta_1 ** t))
        m = f(beta_1) * m + f(1.0 - beta_1) * g
        v = f(beta_2) * v + f(1.0 - beta_2) * g * g
        p = p - lr_t * m / (np.sqrt(v) + f(epsilon))
    return p
