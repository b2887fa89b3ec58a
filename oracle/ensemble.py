"""Symmetrised wave functions restated in float64 numpy (machines/ensemble.py:14-69,
layers/dihedral_4_invariants.py:8-30).  Test infrastructure only.

The symmetries act on (B, H, W) images as tf.image.rot90 (counter-clockwise
quarter turns, numpy.rot90 on axes (1, 2)) and tf.image.flip_left_right
(reversal of the last axis); the up/down symmetry negates the spins.
"""
import numpy as np


def probabilistic_ensemble(x):
    """ensemble.py:14-20 over the last axis of complex log psi values x."""
    k = x.shape[-1]
    two_re = 2.0 * x.real
    mx = two_re.max(axis=-1, keepdims=True)
    re = 0.5 * (mx[..., 0] + np.log(np.exp(two_re - mx).sum(axis=-1))) - 0.5 * np.log(k)
    im = np.angle(np.exp(1j * x.imag).mean(axis=-1))
    return re + 1j * im


def d4_images(x):
    """The eight inputs of make_2d_obc_invariants (ensemble.py:44-49): rot90
    by 0..3, then each of them flipped left-right."""
    rots = [np.rot90(x, k, axes=(1, 2)) for k in range(4)]
    return rots + [r[:, :, ::-1] for r in rots]


def obc_invariant_log_psi(log_psi_fn, x):
    """make_2d_obc_invariants(model) (ensemble.py:44-52)."""
    return probabilistic_ensemble(np.stack([log_psi_fn(np.ascontiguousarray(t)) for t in d4_images(x)], axis=-1))


def up_down_invariant_log_psi(log_psi_fn, x):
    """make_up_down_invariant(model) (ensemble.py:65-69)."""
    return probabilistic_ensemble(np.stack([log_psi_fn(x), log_psi_fn(-x)], axis=-1))


def evaluation_log_psi(log_psi_fn, x):
    """The model experiments/run_evaluation.py:16-19 evaluates:
    up/down invariance around the obc (dihedral) invariance."""
    return up_down_invariant_log_psi(lambda y: obc_invariant_log_psi(log_psi_fn, y), x)


def pbc_invariant_log_psi(log_psi_fn, x, apply_also_obc_invariants=True):
    """make_pbc_invariants(model) (ensemble.py:55-62): every translation of
    the lattice (tensorflow.roll of the image by (i, j)) around the
    obc-invariant model."""
    inner = (lambda y: obc_invariant_log_psi(log_psi_fn, y)) if apply_also_obc_invariants else log_psi_fn
    H, W = x.shape[1:]
    members = [inner(np.ascontiguousarray(np.roll(x, (i, j), axis=(1, 2)))) for i in range(H) for j in range(W)]
    return probabilistic_ensemble(np.stack(members, axis=-1))
