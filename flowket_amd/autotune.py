"""Schedule autotuner for the fused ConvNetAutoregressive2D forward.

Every (channels, precision) has a few compiled launch schedules (loop order x
waves per workgroup, fk_convar_num_schedules).  ``tune_forward`` times each on
the caller's own workload -- events recorded on the stream the kernels are
launched on -- selects the fastest for that network (fk_convar_set_schedule,
keyed by lattice, depth, width, activation and precision) and records it in a
JSON cache keyed by device and the same network description, so later
processes start from the tuned choice (``apply_cached``).  The cache path is $FLOWKET_AMD_TUNE_CACHE or
~/.cache/flowket_amd/forward_schedules.json.  All schedules compute the same
function; they differ only in FP32 summation order.  The row-reuse
policies give the same log psi up to FP32 summation order too (rows above a
connection's first changed row are the sample's own, taken from its cache).
"""
import json
import os
import threading

import torch

from . import _lib

_lock = threading.Lock()


def cache_path():
    return os.environ.get('FLOWKET_AMD_TUNE_CACHE',
                          os.path.join(os.path.expanduser('~'), '.cache', 'flowket_amd', 'forward_schedules.json'))


def _key(machine, prec):
    dev = torch.cuda.get_device_name(machine.device)
    return '|'.join(str(v) for v in (dev, machine.H, machine.W, machine.num_of_channels, prec, machine.activation,
                                     machine.depth))


def _read_cache():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _write_cache(entries):
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + '.tmp%d' % os.getpid()
        with open(tmp, 'w') as f:
            json.dump(entries, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only loses the cache, not the choice


def schedules(channels, prec):
    lib = _lib.load()
    n = int(lib.fk_convar_num_schedules(channels, prec))
    out = []
    for i in range(n):
        t, w = _lib.ctypes.c_int32(), _lib.ctypes.c_int32()
        _lib.call('fk_convar_schedule_info', channels, prec, i, _lib.ctypes.byref(t), _lib.ctypes.byref(w))
        out.append({'index': i, 'tile_outer': int(t.value), 'waves': int(w.value)})
    return out


def _act(machine):
    from .machines import _ACT
    return _ACT[machine.activation]


def set_schedule(machine, prec, index):
    """Select schedule ``index`` for this network (lattice, depth, width,
    activation, precision); other networks keep theirs."""
    _lib.call('fk_convar_set_schedule', machine.H, machine.W, machine.depth, machine.num_of_channels, _act(machine),
              prec, int(index))


def get_schedule(machine, prec):
    return int(_lib.load().fk_convar_get_schedule(machine.H, machine.W, machine.depth, machine.num_of_channels,
                                                  _act(machine), prec))


def apply_cached(machine, prec):
    """Select the cached schedule for this machine/precision, if any."""
    with _lock:
        entry = _read_cache().get(_key(machine, prec))
    if entry is not None and 0 <= int(entry['schedule']) < len(schedules(machine.num_of_channels, prec)):
        set_schedule(machine, prec, int(entry['schedule']))
        if 'row_reuse' in entry and hasattr(machine, 'set_row_reuse_policy'):
            machine.set_row_reuse_policy(int(entry['row_reuse']))
        return int(entry['schedule'])
    return None


def _time(run, stream, reps, warmup):
    for _ in range(warmup):
        run()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(reps):
        run()
    end.record(stream)
    end.synchronize()
    return start.elapsed_time(end) / reps


def tune_forward(machine, run, precision=None, reps=3, warmup=1, persist=True, row_reuse=True):
    """Times ``run()`` (a closure that launches the machine's forward on the
    current stream, e.g. one ragged local-energy evaluation) under every
    schedule and keeps the fastest; then, with ``row_reuse`` and a machine
    that has row-reuse policies (off, or the first row that takes restricted
    passes), times those under the chosen schedule and keeps the fastest.
    Returns {'schedule': i, 'ms': [...], 'schedules': [...]} plus
    'row_reuse' and 'row_reuse_ms' when policies were compared."""
    from .machines import _PREC
    prec = _PREC[precision or machine.precision]
    table = schedules(machine.num_of_channels, prec)
    stream = torch.cuda.current_stream(machine.device)
    times = []
    for s in table:
        set_schedule(machine, prec, s['index'])
        times.append(_time(run, stream, reps, warmup))
    best = min(range(len(table)), key=lambda i: times[i])
    set_schedule(machine, prec, table[best]['index'])
    out = {'schedule': table[best]['index'], 'ms': times, 'schedules': table}
    policies = machine.row_reuse_policies() if row_reuse and hasattr(machine, 'row_reuse_policies') else []
    if len(policies) > 1:
        rtimes = []
        for pol in policies:
            machine.set_row_reuse_policy(pol)
            rtimes.append(_time(run, stream, reps, warmup))
        rbest = policies[min(range(len(policies)), key=lambda i: rtimes[i])]
        machine.set_row_reuse_policy(rbest)
        out['row_reuse'] = rbest
        out['row_reuse_ms'] = dict(zip(policies, rtimes))
    if persist:
        with _lock:
            entries = _read_cache()
            entry = {'schedule': table[best]['index'], 'ms': times}
            if 'row_reuse' in out:
                entry['row_reuse'] = out['row_reuse']
            entries[_key(machine, prec)] = entry
            _write_cache(entries)
    return out
