"""Checkpoint and resume: callbacks/checkpoint.py:9-73.

The reference saves Keras HDF5 weights plus ``<path>_optimizer.pkl``, a
pickle of (optimizer weights, epoch, batch).  Here both files are
safetensors (tensors and string metadata only, nothing executable is
stored or loaded):

- ``save_weights(machine, path)``: the flat parameters, and the
  weight-normalisation g vector when the machine has one;
- ``save_optimizer_weights(optimizer, path, epoch, batch)`` writes
  ``<path>_optimizer.safetensors``; ``load_optimizer_weights`` restores the
  state and returns (epoch, batch), or None when the file is absent
  (checkpoint.py:18-26).

``machine.load_weights(path)`` of ConvNetAutoregressive2D also reads the
reference's own HDF5 weight files.  As in the reference, sampler and random
generator states are not saved.
"""
import os
import time

import torch

_HDF5_MAGIC = b'\x89HDF\r\n\x1a\n'


def _st():
    import safetensors.torch
    return safetensors.torch


def save_weights(machine, filepath):
    tensors = {'params': machine.params.detach().float().cpu().contiguous()}
    g = getattr(machine, 'wn_g', None)
    if getattr(machine, 'weights_normalization', False) and g is not None:
        tensors['wn_g'] = torch.as_tensor(g).detach().float().cpu().contiguous()
    _st().save_file(tensors, filepath, metadata={'format': 'flowket_amd weights', 'n_params': str(machine.n_params)})


def is_hdf5(filepath):
    with open(filepath, 'rb') as f:
        return f.read(8) == _HDF5_MAGIC


def load_weights(machine, filepath):
    """Weights written by ``save_weights`` into ``machine`` (parameter count
    and weight normalisation must match)."""
    t = _st().load_file(filepath)
    if 'params' not in t:
        raise ValueError(f'{filepath} holds no parameters')
    if t['params'].numel() != machine.n_params:
        raise ValueError(f'{filepath} holds {t["params"].numel()} parameters, the machine has {machine.n_params}')
    wn = getattr(machine, 'weights_normalization', False)
    if wn != ('wn_g' in t):
        raise ValueError(f'{filepath}: weight normalisation of the file and of the machine differ')
    machine.set_params(t['params'])
    if wn:
        machine.set_wn_g(t['wn_g'])
    return machine


def _optimizer_path(filepath):
    return filepath + '_optimizer.safetensors'


def save_optimizer_weights(optimizer, filepath, epoch, batch):
    """checkpoint.py:9-15."""
    weights = optimizer.get_weights()
    tensors = {'w%04d' % i: torch.as_tensor(w).detach().cpu().contiguous() for i, w in enumerate(weights)}
    _st().save_file(tensors, _optimizer_path(filepath),
                    metadata={'epoch': str(int(epoch)), 'batch': str(int(batch)), 'count': str(len(weights))})


def load_optimizer_weights(optimizer, filepath):
    """checkpoint.py:18-26: (epoch, batch), or None without a saved state."""
    path = _optimizer_path(filepath)
    if not os.path.exists(path):
        return None
    from safetensors import safe_open
    with safe_open(path, framework='pt') as f:
        meta = f.metadata()
        weights = [f.get_tensor('w%04d' % i) for i in range(int(meta['count']))]
    optimizer.set_weights(weights)
    return int(meta['epoch']), int(meta['batch'])


class CheckpointByTime:
    """callbacks/checkpoint.py:29-73: saves every ``save_frequency_in_minutes``
    (checked at the end of each batch) and at the end of training.
    ``filepath`` may hold format fields filled from the logs."""

    def __init__(self, filepath, machine, optimizer=None, save_frequency_in_minutes=30, clock=time.time):
        self.filepath = filepath
        self.machine = machine
        self.optimizer = optimizer
        self.save_frequency_in_minutes = save_frequency_in_minutes
        self.clock = clock
        self.last_save_time = clock()
        self.current_epoch = 0
        self.saved = []

    def _save(self, logs, batch):
        filepath = self.filepath.format(**(logs or {}))
        save_weights(self.machine, filepath)
        if self.optimizer is not None:
            save_optimizer_weights(self.optimizer, filepath, self.current_epoch, batch)
        self.last_save_time = self.clock()
        self.saved.append((filepath, self.current_epoch, batch))

    def on_batch_end(self, batch, logs=None):
        if self.clock() - self.last_save_time < self.save_frequency_in_minutes * 60:
            return
        self._save(logs, batch)

    def on_epoch_begin(self, epoch, logs=None):
        self.current_epoch = epoch

    def on_train_end(self, logs=None):
        self.current_epoch += 1
        self._save(logs, 0)
