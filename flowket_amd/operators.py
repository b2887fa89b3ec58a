"""Hamiltonians with FlowKet's ``find_conn`` interface, evaluated on the GPU.

Mirrors ``flowket.operators`` (operators/operator.py, heisenberg.py, ising.py,
j1j2.py).  Each operator is compiled once into a connection table
(connection c -> sites a, b and its off-diagonal element) plus a diagonal bond
table; ``fk_find_conn`` evaluates every (sample, connection) pair in one
launch, bit-exact with the reference layout.
"""
import abc

import numpy as np
import torch

from . import _lib


def cube_shape(number_of_spins_in_each_dimention=20, cube_dimention=1, column_or_row=True):
    """operators/operator.py:33-37."""
    if cube_dimention == 1:
        return [number_of_spins_in_each_dimention, 1] if column_or_row else [1, number_of_spins_in_each_dimention]
    return [number_of_spins_in_each_dimention, ] * cube_dimention


class Operator(abc.ABC):
    """operators/operator.py:6-30."""

    def __init__(self, hilbert_state_shape):
        self.hilbert_state_shape = tuple(int(s) for s in hilbert_state_shape)
        self.max_number_of_local_connections = None

    @abc.abstractmethod
    def find_conn(self, sample):
        pass

    def use_state(self, state):
        return True

    def random_states(self, num_of_states):
        return np.random.choice([-1, 1], size=(num_of_states,) + self.hilbert_state_shape)

    @property
    def n_sites(self):
        return int(np.prod(self.hilbert_state_shape))

    def find_conn_device(self, spins, want_conn=True, stream=None):
        """For an operator defined only by a numpy ``find_conn`` (a subclass
        of this class, as in operators/operator.py:14-21, or a wrapper of a
        netket operator, operators/netket_operator.py:44-62): the subclass's
        find_conn on the host, its (conn, mel, use) uploaded as the device
        layout of fk_find_conn -- conn (Nc, B, N) int8, mel (Nc, B) float64
        real parts with unused slots zeroed, use (Nc, B) uint8 -- and, when
        any element is complex, mel_im (Nc, B) float64 imaginary parts.
        As the reference's estimators require
        (observables/monte_carlo/operator.py:30-42), connection 0 of every
        sample must be the sample itself and used: its log psi is the
        denominator of every ratio.  A find_conn that breaks this raises."""
        B = spins.shape[0]
        x = spins.detach().to('cpu').numpy().reshape((B,) + self.hilbert_state_shape)
        conn, mel, use = self.find_conn(x)
        mel = np.asarray(mel)
        use = np.asarray(use, dtype=bool)
        nc = mel.shape[0]
        if mel.shape != (nc, B) or use.shape != (nc, B):
            raise ValueError(f'find_conn returned mel {mel.shape} / use {use.shape}, expected ({nc}, {B})')
        c0 = np.asarray(conn)[0].reshape(B, -1)
        if not (use[0].all() and np.array_equal(c0, x.reshape(B, -1))):
            raise ValueError('find_conn: connection 0 of every sample must be the sample itself, marked used '
                             '(observables/monte_carlo/operator.py:30-42)')
        m = np.where(use, mel, 0)
        dev = spins.device
        out = {'mel': torch.as_tensor(np.ascontiguousarray(np.real(m), dtype=np.float64)).to(dev),
               'use': torch.as_tensor(use.astype(np.uint8)).to(dev)}
        if np.iscomplexobj(m) and np.any(np.imag(m) != 0):
            out['mel_im'] = torch.as_tensor(np.ascontiguousarray(np.imag(m), dtype=np.float64)).to(dev)
        if want_conn:
            c = np.asarray(conn).reshape(nc, B, -1)
            out['conn'] = torch.as_tensor(np.ascontiguousarray(c).astype(np.int8)).to(dev)
        return out


class OperatorOnGrid(Operator, abc.ABC):
    """operators/operator.py:40-46."""

    def __init__(self, hilbert_state_shape=None, pbc=True):
        if hilbert_state_shape is None:
            hilbert_state_shape = cube_shape()
        super().__init__(hilbert_state_shape)
        self.pbc = pbc


def _as_device_spins(sample, n_sites, device=None):
    if isinstance(sample, torch.Tensor):
        t = sample
    else:
        t = torch.as_tensor(np.asarray(sample))
    if device is None:
        device = torch.device('cuda', torch.cuda.current_device())
    t = t.to(device=device).reshape(t.shape[0], n_sites)
    if t.dtype != torch.int8:
        t = t.to(torch.int8)
    return t.contiguous()


class _BondTableOperator(OperatorOnGrid):
    """Shared GPU machinery: a connection table (conn_a, conn_b, conn_mel)
    and a diagonal table (diag_a, diag_b, diag_j) evaluated by fk_find_conn."""

    kind = _lib.FK_CONN_SWAP
    diag_rule = _lib.FK_DIAG_SUM_F64

    def _set_tables(self, conn_a, conn_b, conn_mel, diag_a, diag_b, diag_j):
        self._conn_a = np.asarray(conn_a, dtype=np.int32)
        self._conn_b = np.asarray(conn_b, dtype=np.int32)
        self._conn_mel = np.asarray(conn_mel, dtype=np.float64)
        self._diag_a = np.asarray(diag_a, dtype=np.int32)
        self._diag_b = np.asarray(diag_b, dtype=np.int32)
        self._diag_j = np.asarray(diag_j, dtype=np.float64)
        self._device_tables = {}

    @property
    def n_sites(self):
        return int(np.prod(self.hilbert_state_shape))

    @property
    def num_connections(self):
        return int(self._conn_a.shape[0])

    def tables(self, device):
        """Connection/diagonal tables resident on `device` (uploaded once)."""
        key = str(device)
        if key not in self._device_tables:
            def up(a):
                return torch.as_tensor(a).to(device)
            # the diagonal table always gets at least one (zero-weight) entry
            da, db, dj = self._diag_a, self._diag_b, self._diag_j
            if da.size == 0:
                da, db, dj = np.zeros(1, np.int32), np.zeros(1, np.int32), np.zeros(1)
            self._device_tables[key] = dict(
                conn_a=up(self._conn_a), conn_b=up(self._conn_b), conn_mel=up(self._conn_mel),
                diag_a=up(da), diag_b=up(db), diag_j=up(dj), n_diag=int(self._diag_a.size))
        return self._device_tables[key]

    def row_buckets(self, W, device, first=2, last=None):
        """Per connection slot, the bucket of the row-reuse forward on a
        lattice W sites wide: 0 for slot 0 (the sample itself), else 1 + the
        first lattice row the pass computes.  A connection whose lowest
        changed site lies on row r is computed from row min(r, last) when r >=
        first (a restricted pass needs two rows above it), else in full (row
        0).  last defaults to H - 2.  Returns (device int32 table, number of
        buckets, slots per bucket)."""
        n = self.n_sites
        H = n // W
        first = max(2, int(first))
        last = H - 2 if last is None else min(int(last), H - 1)
        key = ('row_buckets', int(W), first, last, str(device))
        if key not in self._device_tables:
            a, b = self._conn_a, self._conn_b
            bk = np.zeros(a.size, np.int32)
            for c in range(1, a.size):
                if a[c] < 0:
                    bk[c] = 1
                    continue
                q0 = min(a[c], b[c]) if (self.kind != _lib.FK_CONN_FLIP and b[c] >= 0) else a[c]
                r = q0 // W
                bk[c] = 1 + (0 if r < first else min(r, last))
            sizes = np.bincount(bk, minlength=H + 1)
            self._device_tables[key] = (torch.as_tensor(bk).to(device), H + 1, sizes)
        return self._device_tables[key]

    def find_conn_device(self, spins, want_conn=True, stream=None):
        """spins (B, n_sites) int8 on the GPU -> dict of device tensors
        conn (Nc, B, n_sites) int8 or None, mel (Nc, B) float64, use (Nc, B)
        uint8, n_used (B,) int32."""
        _lib.require_gpu()
        B = spins.shape[0]
        n = self.n_sites
        nc = self.num_connections
        dev = spins.device
        tb = self.tables(dev)
        mel = torch.empty((nc, B), dtype=torch.float64, device=dev)
        use = torch.empty((nc, B), dtype=torch.uint8, device=dev)
        n_used = torch.empty((B,), dtype=torch.int32, device=dev)
        conn = torch.empty((nc, B, n), dtype=torch.int8, device=dev) if want_conn else None
        _lib.call('fk_find_conn', _lib.ptr(spins), B, n, self.kind, _lib.ptr(tb['conn_a']), _lib.ptr(tb['conn_b']),
                  _lib.ptr(tb['conn_mel']), nc, _lib.ptr(tb['diag_a']), _lib.ptr(tb['diag_b']),
                  _lib.ptr(tb['diag_j']), tb['n_diag'], self.diag_rule, _lib.ptr(conn), _lib.ptr(mel), _lib.ptr(use),
                  _lib.ptr(n_used), _lib.stream_ptr(stream))
        return dict(conn=conn, mel=mel, use=use, n_used=n_used)

    def find_conn(self, sample, as_numpy=None):
        """FlowKet's triple (conn (Nc, B, *shape), mel (Nc, B), use (Nc, B)).
        numpy input gives numpy output (conn float64 like the reference);
        torch input gives device tensors (conn int8, use bool)."""
        numpy_in = not isinstance(sample, torch.Tensor)
        if as_numpy is None:
            as_numpy = numpy_in
        spins = _as_device_spins(sample, self.n_sites)
        r = self.find_conn_device(spins, want_conn=True)
        B = spins.shape[0]
        conn = r['conn'].reshape((self.num_connections, B) + self.hilbert_state_shape)
        use = r['use'].to(torch.bool)
        if as_numpy:
            return conn.cpu().numpy().astype(np.float64), r['mel'].cpu().numpy(), use.cpu().numpy()
        return conn, r['mel'], use

    def compact(self, spins, stream=None):
        """Ragged form for the local-energy path: mel, use, offsets (B+1,)
        int64 and pairs (Nc*B, 2) int32 of which offsets[B] are valid."""
        r = self.find_conn_device(spins, want_conn=False, stream=stream)
        B = spins.shape[0]
        nc = self.num_connections
        offsets = torch.empty((B + 1,), dtype=torch.int64, device=spins.device)
        pairs = torch.empty((nc * B, 2), dtype=torch.int32, device=spins.device)
        _lib.call('fk_compact_connections', _lib.ptr(r['use']), B, nc, _lib.ptr(offsets), _lib.ptr(pairs), None,
                  _lib.stream_ptr(stream))
        r.update(offsets=offsets, pairs=pairs)
        return r


class Heisenberg(_BondTableOperator):
    """operators/heisenberg.py:6-44 (Marshall-rotated off-diagonal -2 when
    unitary_rotation, else +2; J = 1)."""

    kind = _lib.FK_CONN_SWAP

    def __init__(self, total_sz=0.0, unitary_rotation=True, **kwargs):
        super().__init__(**kwargs)
        self.off_diag = -2.0 if unitary_rotation else 2.0
        self.total_sz = total_sz
        self.total_size = int(np.prod(self.hilbert_state_shape))
        self.dim = len(self.hilbert_state_shape)
        assert self.dim <= 2
        self.max_number_of_local_connections = self.total_size * self.dim + 1
        H, W = (self.hilbert_state_shape if self.dim == 2 else (self.hilbert_state_shape[0], 1))
        nc = self.max_number_of_local_connections
        ca = -np.ones(nc, np.int32)
        cb = -np.ones(nc, np.int32)
        da, db = [], []
        for i in range(H):
            for j in range(W):
                d = 0
                for axis, size in ((0, H), (1, W)):
                    if size <= 1:
                        continue
                    ni, nj = (i + 1, j) if axis == 0 else (i, j + 1)
                    wrap = ni >= H or nj >= W
                    c = 1 + (i * W + j) * self.dim + d
                    if not (wrap and not self.pbc):
                        a, b = i * W + j, (ni % H) * W + (nj % W)
                        ca[c], cb[c] = a, b
                        da.append(a)
                        db.append(b)
                    d += 1
        self._set_tables(ca, cb, np.full(nc, self.off_diag), da, db, np.ones(len(da)))

    def random_states(self, num_of_states):
        """operators/heisenberg.py:21-32."""
        if self.total_sz is not None:
            size = self.total_size // 2
            states = np.zeros((num_of_states, self.total_size))
            for i in range(num_of_states):
                states[i] = np.concatenate([np.ones(size + int(self.total_sz)),
                                            np.full(size - int(self.total_sz), -1)])
                np.random.shuffle(states[i])
            return states.reshape((num_of_states,) + self.hilbert_state_shape)
        return super().random_states(num_of_states)

    def use_state(self, state):
        if self.total_sz is None:
            return True
        return state.sum() == self.total_sz


class Ising(_BondTableOperator):
    """operators/ising.py:8-46: H = -j sum s s' - h sum sigma_x.  The diagonal
    follows the reference's float32 arithmetic (ising.py:25-41: per-site
    float32 values, float32 pairwise sum), bit for bit for any j, as the
    reference computes it for float64 +-1 samples (the samplers' output)."""

    kind = _lib.FK_CONN_FLIP
    diag_rule = _lib.FK_DIAG_ISING_F32

    def __init__(self, h=1.0, j=1.0, **kwargs):
        super().__init__(**kwargs)
        self.h = h
        self.j = j
        n = int(np.prod(self.hilbert_state_shape))
        self.max_number_of_local_connections = n + 1
        shape = self.hilbert_state_shape
        assert len(shape) <= 2
        ca = np.concatenate([[-1], np.arange(n)]).astype(np.int32)
        cb = -np.ones(n + 1, np.int32)
        da, db = [], []
        if len(shape) == 1:
            L = shape[0]
            for i in range(L):
                if i + 1 < L or self.pbc:
                    da.append(i)
                    db.append((i + 1) % L)
        else:
            H, W = shape
            for i in range(H):
                for jj in range(W):
                    if H > 1 and (i + 1 < H or self.pbc):
                        da.append(i * W + jj)
                        db.append(((i + 1) % H) * W + jj)
                    if W > 1 and (jj + 1 < W or self.pbc):
                        da.append(i * W + jj)
                        db.append(i * W + (jj + 1) % W)
        self._set_tables(ca, cb, np.full(n + 1, -float(h)), da, db, np.full(len(da), -float(j)))


def j1j2_edges(L1, L2, pbc):
    """Bond list of operators/j1j2.py:7-35 as (site_a, site_b, color), color 1
    = J1 (nearest neighbour), 2 = J2 (next-nearest neighbour), in the
    reference's append order."""
    e = []
    for hh in range(L1):
        for w in range(L2 - 1):
            e.append((w + L2 * hh, w + 1 + L2 * hh, 1))
            if hh < L1 - 1:
                e += [(w + L2 * hh, w + L2 * (hh + 1), 1), (w + L2 * hh, w + 1 + L2 * (hh + 1), 2)]
            elif pbc:
                e += [(w + L2 * hh, w, 1), (w + L2 * hh, w + 1, 2)]
            if hh > 0:
                e.append((w + L2 * hh, w + 1 + L2 * (hh - 1), 2))
            elif pbc:
                e.append((w + L2 * hh, w + 1 + L2 * (L1 - 1), 2))
        w = L2 - 1
        if pbc:
            e += [(L2 - 1 + L2 * hh, L2 * hh, 1), (w + L2 * hh, L2 * ((hh + 1) % L1), 2),
                  (w + L2 * hh, L2 * ((L1 + hh - 1) % L1), 2)]
        if hh < L1 - 1:
            e.append((w + L2 * hh, w + L2 * (hh + 1), 1))
        elif pbc:
            e.append((w + L2 * hh, w, 1))
    return e


class J1J2(_BondTableOperator):
    """J1-J2 model of operators/j1j2.py:58-69 without netket: bond operator
    J_b*(sz sz + sx sx + sy sy) (:37-55) gives diagonal J_b s_a s_b and +2 J_b
    for the swapped antiparallel pair; padding/use rule of
    operators/netket_operator.py:46-66 (use = mel != 0, unused slots are
    all-zero configurations).  Slot 1 + k belongs to bond k of j1j2_edges."""

    kind = _lib.FK_CONN_SWAP_ZERO_PAD

    def __init__(self, hilbert_state_shape, j2=0.5, pbc=False):
        assert len(hilbert_state_shape) == 2
        super().__init__(hilbert_state_shape=hilbert_state_shape, pbc=pbc)
        self.j2 = j2
        L1, L2 = self.hilbert_state_shape
        n = L1 * L2
        self.max_number_of_local_connections = n * 2 * 2 + 1
        nc = self.max_number_of_local_connections
        edges = j1j2_edges(L1, L2, pbc)
        J = {1: 1.0, 2: float(j2)}
        ca = -np.ones(nc, np.int32)
        cb = -np.ones(nc, np.int32)
        cm = np.zeros(nc)
        for k, (a, b, col) in enumerate(edges):
            ca[1 + k], cb[1 + k], cm[1 + k] = a, b, 2.0 * J[col]
        self._set_tables(ca, cb, cm, [e[0] for e in edges], [e[1] for e in edges], [J[e[2]] for e in edges])
        self.edges = edges


j1j2_two_dim_operator = J1J2
