"""HIP-graph replay of the local-energy step.

One evaluation of the ragged local-energy chain -- fk_find_conn,
fk_compact_connections (count / scan / index map), the machine's forward over
the ragged (sample, connection) list with the list length read from device
memory, fk_local_energy and fk_energy_moments -- is a fixed sequence of
launches whose sizes depend only on the batch, so it is captured once into a
HIP graph and replayed: no host round trip and no per-launch CPU cost.  New
configurations are copied into the graph's static input buffer; after a
parameter update the packed weight buffers the graph reads are refreshed in
place.  (operators/operator.py:13-54 + observables/monte_carlo/operator.py
:13-54 of the reference are the eager counterpart.)
"""
import torch

from . import _lib


class LocalEnergyGraph:
    """E_loc of ``batch`` configurations of ``operator`` under ``machine``
    (a ConvNetAutoregressive2D or RBM), as a replayable graph.

    ``__call__(spins)`` returns (eloc (B,) complex128, moments (4,) float64 =
    mean re, mean im, var re, -) as device tensors owned by the graph (valid
    until the next call)."""

    def __init__(self, machine, operator, batch, precision=None):
        _lib.require_gpu()
        self.machine = machine
        self.operator = operator
        self.batch = int(batch)
        self.precision = precision
        dev = machine.device
        self.spins = torch.zeros((self.batch, operator.n_sites), dtype=torch.int8, device=dev)
        self.spins[:, ::2] = 1
        self.spins[:, 1::2] = -1
        operator.tables(dev)
        # eager warm-up on a side stream: uploads tables, packs weights and
        # performs every one-time launch attribute change outside the capture
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(side):
            self._body()
        torch.cuda.current_stream(dev).wait_stream(side)
        torch.cuda.synchronize(dev)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.eloc, self.moments = self._body()

    def _refresh_weights(self):
        """ConvNet weights are read from packed buffers that the machine
        repacks in place after a parameter update (RBM parameters are read
        directly)."""
        m = self.machine
        if hasattr(m, '_pack'):
            m._pack(self.precision or m.precision)

    def _body(self):
        op, m, B = self.operator, self.machine, self.batch
        r = op.compact(self.spins)
        offsets, pairs = r['offsets'], r['pairs']
        kw = {} if self.precision is None else {'precision': self.precision}
        if getattr(m, 'uses_row_reuse', lambda o: False)(op):
            logpsi = m.log_psi_connected_rows(self.spins, r, op, **kw)
        else:
            logpsi = m.log_psi_connected(self.spins, pairs, pairs.shape[0], op, m_dev=offsets[B:], **kw)
        eloc = torch.empty((B, 2), dtype=torch.float64, device=self.spins.device)
        s = _lib.stream_ptr()
        _lib.call('fk_local_energy', _lib.ptr(offsets), _lib.ptr(pairs), _lib.ptr(logpsi), _lib.ptr(r['mel']), None,
                  B, _lib.ptr(eloc), s)
        mom = torch.empty(4, dtype=torch.float64, device=self.spins.device)
        _lib.call('fk_energy_moments', _lib.ptr(eloc), B, _lib.ptr(mom), s)
        self.offsets = offsets
        return eloc, mom

    def __call__(self, spins):
        x = spins if isinstance(spins, torch.Tensor) else torch.as_tensor(spins)
        x = x.reshape(self.batch, -1)
        if x.shape[1] != self.operator.n_sites:
            raise ValueError('configurations do not match the operator')
        self.spins.copy_(x.to(self.spins.device, torch.int8))
        self._refresh_weights()
        self.graph.replay()
        return torch.complex(self.eloc[:, 0], self.eloc[:, 1]), self.moments

    @property
    def num_connections(self):
        """Used connections of the last replay (device scalar)."""
        return self.offsets[self.batch:]
