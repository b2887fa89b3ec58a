"""Local-energy estimators (observables/monte_carlo/operator.py,
observables/monte_carlo/observable.py) on the GPU.

For a machine with a fused connected-configuration path the whole local
energy runs without a host round trip: fk_find_conn (mel/use),
fk_compact_connections (ragged index map), the network over the ragged
(sample, connection) list with the count read from device memory (for
ConvNetAutoregressive2D with row reuse: the samples' pass, then each bucket of
connections evaluated from the first row it changes), then fk_local_energy.  Any other callable wave function goes through the dense
(balanced) layout.
"""
import abc

import numpy as np
import torch

from . import _lib


class BaseObservable(abc.ABC):
    """observables/monte_carlo/observable.py:5-14."""

    @abc.abstractmethod
    def local_values(self, wave_function, configurations):
        pass

    def estimate(self, wave_function, configurations):
        lv = self.local_values(wave_function, configurations)
        mean, var = energy_moments(lv)
        return mean, var, lv


def energy_moments(local_values, stream=None):
    """(mean complex, population variance of the real part) reduced in FP64
    on the device (numpy.mean / numpy.var of observable.py:12-13)."""
    lv = torch.view_as_real(local_values.to(torch.complex128)).contiguous()
    out = torch.empty(4, dtype=torch.float64, device=lv.device)
    _lib.call('fk_energy_moments', _lib.ptr(lv), lv.shape[0], _lib.ptr(out), _lib.stream_ptr(stream))
    o = out.cpu().numpy()
    return complex(o[0], o[1]), float(o[2])


class Observable(BaseObservable):
    """observables/monte_carlo/operator.py:13-54."""

    def __init__(self, operator):
        self.operator = operator
        self.last_num_connections = None

    def _spins(self, configurations):
        op = self.operator
        x = configurations
        if not isinstance(x, torch.Tensor):
            x = torch.as_tensor(np.asarray(x))
        dev = torch.device('cuda', torch.cuda.current_device())
        return x.to(dev).to(torch.int8).reshape(x.shape[0], op.n_sites).contiguous()

    def local_values(self, wave_function, configurations, stream=None):
        """E_loc(s) = sum over used connections of mel * exp(logpsi(s') -
        logpsi(s)), returned as a (B,) complex128 device tensor."""
        _lib.require_gpu()
        spins = self._spins(configurations)
        B = spins.shape[0]
        op = self.operator
        if hasattr(wave_function, 'log_psi_connected') and hasattr(op, 'compact'):
            r = op.compact(spins, stream=stream)
            offsets, pairs = r['offsets'], r['pairs']
            upper = pairs.shape[0]
            if getattr(wave_function, 'uses_row_reuse', lambda o: False)(op):
                logpsi = wave_function.log_psi_connected_rows(spins, r, op, stream=stream)
            else:
                logpsi = wave_function.log_psi_connected(spins, pairs, upper, op, m_dev=offsets[B:], stream=stream)
            eloc = torch.empty((B, 2), dtype=torch.float64, device=spins.device)
            _lib.call('fk_local_energy', _lib.ptr(offsets), _lib.ptr(pairs), _lib.ptr(logpsi), _lib.ptr(r['mel']),
                      _lib.ptr(r.get('mel_im')), B, _lib.ptr(eloc), _lib.stream_ptr(stream))
            self.last_num_connections = offsets[B:]
            return torch.complex(eloc[:, 0], eloc[:, 1])
        return self.local_values_balanced(wave_function, spins, stream=stream)

    def local_values_balanced(self, wave_function, spins, stream=None):
        """operator.py:35-42: every slot evaluated (mel = 0 for unused).
        Operators without a connection table (a subclass of Operator with a
        numpy find_conn) come here too; complex matrix elements keep their
        imaginary parts."""
        op = self.operator
        r = op.find_conn_device(spins, want_conn=True, stream=stream)
        nc, B = r['mel'].shape
        flat = r['conn'].reshape(nc * B, *op.hilbert_state_shape)
        lp = wave_function(flat)
        lp = lp.reshape(-1)
        lp = torch.view_as_real(lp.to(torch.complex64)).contiguous()
        eloc = torch.empty((B, 2), dtype=torch.float64, device=spins.device)
        _lib.call('fk_local_energy_dense', _lib.ptr(lp), _lib.ptr(r['mel']), _lib.ptr(r.get('mel_im')), nc, B,
                  _lib.ptr(eloc), _lib.stream_ptr(stream))
        return torch.complex(eloc[:, 0], eloc[:, 1])


class _SigmaZBase(BaseObservable):
    """observables/monte_carlo/sigma_z.py:7-20 (LambdaObservable of the
    configurations only; the wave function is not evaluated)."""

    absolute = 0

    def local_values(self, wave_function, configurations, stream=None):
        _lib.require_gpu()
        x = configurations
        if not isinstance(x, torch.Tensor):
            x = torch.as_tensor(np.asarray(x))
        dev = torch.device('cuda', torch.cuda.current_device())
        B = x.shape[0]
        spins = x.to(dev).to(torch.int8).reshape(B, -1).contiguous()
        out = torch.empty((B, 2), dtype=torch.float64, device=dev)
        _lib.call('fk_sigma_z', _lib.ptr(spins), B, spins.shape[1], self.absolute, _lib.ptr(out),
                  _lib.stream_ptr(stream))
        return torch.complex(out[:, 0], out[:, 1])


class SigmaZ(_SigmaZBase):
    """Mean spin per site of each configuration."""
    absolute = 0


class AbsSigmaZ(_SigmaZBase):
    """|mean spin per site| of each configuration."""
    absolute = 1
