"""Symmetrised wave functions: machines/ensemble.py:14-69.

``make_2d_obc_invariants(model)`` averages log psi over the eight lattice
symmetries of a square open lattice (Rot90 by 0..3 quarter turns, each with
and without FlipLeftRight, layers/dihedral_4_invariants.py:8-30),
``make_up_down_invariant(model)`` over x and -x, and ``make_pbc_invariants``
over every translation (optionally of the obc-invariant model).  Each is a
probabilistic ensemble (ensemble.py:14-20): |psi|^2 is the mean of the
members' |psi|^2 and the phase is the argument of the mean of their phase
factors.  Ensembles nest, as in experiments/run_evaluation.py:16-19.

A symmetry is a gather y[q] = sign * x[src[q]] over the flattened lattice.
For connected configurations the operator's connection table is relabelled
instead of materialising them: flipping (or exchanging) site a of x is
flipping site inv[a] of y, where inv is the inverse of src, so each member
runs the machine's fused connected-configuration forward unchanged.
"""
import contextlib
import itertools

import numpy as np
import torch

from . import _lib


def _rot90_src(H, W, k, flip):
    """Flat source index of every output site of tf.image.rot90(x, k)
    (counter-clockwise quarter turns) followed, when ``flip``, by
    tf.image.flip_left_right."""
    idx = np.arange(H * W, dtype=np.int32).reshape(H, W)
    out = np.rot90(idx, k)
    if flip:
        out = out[:, ::-1]
    return np.ascontiguousarray(out).reshape(-1)


def _roll_src(shape, shifts):
    """tensorflow.roll of the lattice by ``shifts`` (layers Roll)."""
    idx = np.arange(int(np.prod(shape)), dtype=np.int32).reshape(shape)
    return np.ascontiguousarray(np.roll(idx, shifts, axis=tuple(range(len(shape))))).reshape(-1)


class _RelabelledOperator:
    """The operator's connection tables with every site a replaced by inv[a]
    (the diagonal tables are not used by the forward and are kept)."""

    def __init__(self, operator, inv):
        self.operator = operator
        self.kind = operator.kind
        self.hilbert_state_shape = operator.hilbert_state_shape
        self.num_connections = operator.num_connections
        self._inv = np.asarray(inv, dtype=np.int32)
        self._tables = {}

    def tables(self, device):
        key = str(device)
        if key not in self._tables:
            tb = dict(self.operator.tables(device))
            inv = torch.as_tensor(self._inv, dtype=torch.int32, device=device)
            n = inv.numel()
            for name in ('conn_a', 'conn_b'):
                c = tb[name]
                ok = (c >= 0) & (c < n)
                tb[name] = torch.where(ok, inv[c.clamp(0, n - 1).long()], c).to(torch.int32).contiguous()
            self._tables[key] = tb
        return self._tables[key]


class SymmetrizedMachine:
    """Probabilistic ensemble of ``model`` over the symmetries ``members``
    (list of (src, sign)); ``model`` is a machine of this package or another
    SymmetrizedMachine."""

    def __init__(self, model, members):
        self.model = model
        self.members = []
        n = None
        for src, sign in members:
            src = np.asarray(src, dtype=np.int32)
            n = src.size if n is None else n
            if src.size != n or not np.array_equal(np.sort(src), np.arange(n)):
                raise ValueError('every symmetry must be a permutation of the lattice sites')
            inv = np.empty_like(src)
            inv[src] = np.arange(n, dtype=np.int32)
            self.members.append((src, int(sign), inv))
        self.input_size = model.input_size
        self.device = model.device
        self.n_sites = n
        self._src_dev = {}
        self._ops = {}

    def _src(self, k):
        key = (k, str(self.device))
        if key not in self._src_dev:
            self._src_dev[key] = torch.as_tensor(self.members[k][0], device=self.device)
        return self._src_dev[key]

    def _relabelled(self, operator, k):
        key = (id(operator), k)
        if key not in self._ops:
            self._ops[key] = (operator, _RelabelledOperator(operator, self.members[k][2]))
        return self._ops[key][1]

    def _transform(self, spins, k, stream):
        out = torch.empty_like(spins)
        _lib.call('fk_transform_spins', _lib.ptr(spins), spins.shape[0], self.n_sites, _lib.ptr(self._src(k)),
                  self.members[k][1], _lib.ptr(out), _lib.stream_ptr(stream))
        return out

    def _combine(self, parts, m, m_dev, stream):
        K = len(parts)
        buf = torch.empty((K, m, 2), dtype=torch.float32, device=self.device)
        with (torch.cuda.stream(stream) if stream is not None else contextlib.nullcontext()):
            for k, p in enumerate(parts):
                buf[k].copy_(p[:m])
        out = torch.zeros((m, 2), dtype=torch.float32, device=self.device)
        _lib.call('fk_ensemble_combine', _lib.ptr(buf), K, m, m, _lib.ptr(m_dev), _lib.ptr(out),
                  _lib.stream_ptr(stream))
        return out

    def _spins(self, x):
        if not isinstance(x, torch.Tensor):
            x = torch.as_tensor(np.asarray(x))
        x = x.to(self.device)
        if x.dtype != torch.int8:
            x = x.to(torch.int8)
        return x.reshape(x.shape[0], self.n_sites).contiguous()

    def log_psi_connected(self, spins, pairs, m, operator, m_dev=None, precision=None, stream=None):
        """log psi (m, 2) float32 of the connected configurations in ``pairs``
        (first m rows, or *m_dev rows when given); with ``pairs`` None, of the
        configurations ``spins`` themselves."""
        _lib.require_gpu()
        if pairs is None:
            m, m_dev = spins.shape[0], None
        elif m is None:
            m = pairs.shape[0]
        parts = [self.model.log_psi_connected(self._transform(spins, k, stream), pairs, m,
                                              None if pairs is None else self._relabelled(operator, k),
                                              m_dev=m_dev, precision=precision, stream=stream)
                 for k in range(len(self.members))]
        return self._combine(parts, m, m_dev, stream)

    def forward(self, spins, precision=None, stream=None):
        """(logpsi (B, 2) float32, None) of a batch (B, n_sites) int8."""
        return self.log_psi_connected(spins, None, None, None, precision=precision, stream=stream), None

    def log_psi(self, x):
        """(B,) complex64 log psi."""
        out, _ = self.forward(self._spins(x))
        return torch.complex(out[:, 0], out[:, 1])

    def predict(self, x, batch_size=None):
        return self.log_psi(x)[:, None]

    __call__ = predict


def make_2d_obc_invariants(model):
    """ensemble.py:44-52 (square lattices: quarter turns keep the shape)."""
    H, W = model.input_size
    if H != W:
        raise ValueError('the dihedral symmetries of a lattice need H == W')
    return SymmetrizedMachine(model, [(_rot90_src(H, W, k, flip), 1) for flip in (False, True) for k in range(4)])


def make_up_down_invariant(model):
    """ensemble.py:65-69."""
    n = int(np.prod(model.input_size))
    ident = np.arange(n, dtype=np.int32)
    return SymmetrizedMachine(model, [(ident, 1), (ident, -1)])


def make_pbc_invariants(model, apply_also_obc_invariants=True):
    """ensemble.py:55-62: every translation of the lattice."""
    if apply_also_obc_invariants:
        model = make_2d_obc_invariants(model)
    shape = tuple(model.input_size)
    shifts = itertools.product(*[range(s) for s in shape])
    return SymmetrizedMachine(model, [(_roll_src(shape, list(s)), 1) for s in shifts])
