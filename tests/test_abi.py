"""The C-ABI library builds, loads, and exports every declared symbol; host-only
entry points (packing maps) behave.  No GPU compute here."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'flowket_hip.h')


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(fk_[a-z0-9_]+)\s*\(', text)))


@pytest.fixture(scope='module')
def lib():
    from flowket_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from flowket_amd import build
        build.build()
    return _lib.load()


def test_header_symbols_exported(lib):
    syms = declared_symbols()
    assert len(syms) >= 10
    for s in syms:
        assert hasattr(lib, s), s


def test_binding_covers_header():
    from flowket_amd import _lib
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_version(lib):
    from flowket_amd import _lib
    assert lib.fk_version() >= 1
    assert isinstance(lib.fk_last_error(), bytes)
    _ = _lib


def _maps(depth, C, prec):
    from flowket_amd import _lib
    nw, na = ctypes.c_int64(), ctypes.c_int64()
    rc = _lib.load().fk_convar_packed_sizes(depth, C, prec, ctypes.byref(nw), ctypes.byref(na))
    assert rc == 0
    wmap = np.empty(nw.value, np.int32)
    amap = np.empty(na.value, np.int32)
    rc = _lib.load().fk_convar_pack_maps(depth, C, prec, wmap.ctypes.data_as(ctypes.c_void_p),
                                         amap.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    return wmap, amap


@pytest.mark.parametrize('depth,C,prec', [(5, 16, 0), (4, 32, 0), (20, 32, 1), (3, 64, 1), (2, 32, 1), (1, 16, 0),
                                          (3, 48, 1), (2, 24, 0), (2, 8, 0), (3, 40, 1), (2, 6, 1),
                                          (3, 32, 2), (2, 48, 2)])
def test_pack_maps_cover_every_parameter_once(lib, depth, C, prec):
    """Every network parameter appears in the packed buffers exactly once
    (weights of MFMA convolutions once in the fragment buffer, biases and the
    first block's 1-channel kernels once in the aux buffer)."""
    from flowket_amd.machines import layer_shapes
    wmap, amap = _maps(depth, C, prec)
    n = sum(int(np.prod(k)) + int(np.prod(b)) for k, b in layer_shapes(depth, C))
    used = np.concatenate([wmap[wmap >= 0], amap[amap >= 0]])
    assert used.min() >= 0 and used.max() < n
    counts = np.bincount(used, minlength=n)
    assert np.all(counts == 1)


def test_unsupported_shapes_are_errors(lib):
    from flowket_amd import _lib
    nw, na = ctypes.c_int64(), ctypes.c_int64()
    for C, prec in ((25, 1), (66, 1), (72, 0), (0, 0)):
        assert _lib.load().fk_convar_packed_sizes(3, C, prec, ctypes.byref(nw), ctypes.byref(na)) == 3
        assert b'not implemented' in _lib.load().fk_last_error()
    # padded widths share the wider kernel's packing
    assert _lib.load().fk_convar_packed_sizes(3, 48, 1, ctypes.byref(nw), ctypes.byref(na)) == 0
    n48 = nw.value
    assert _lib.load().fk_convar_packed_sizes(3, 64, 1, ctypes.byref(nw), ctypes.byref(na)) == 0
    assert nw.value == n48


def test_schedule_table(lib):
    """Host-only schedule selection: table sizes, range checks, set/get/reset."""
    from flowket_amd import _lib
    L = _lib.load()
    assert L.fk_convar_num_schedules(32, 1) == 4
    assert L.fk_convar_num_schedules(16, 0) == 4
    assert L.fk_convar_num_schedules(64, 1) == 2
    assert L.fk_convar_num_schedules(24, 1) == 4
    assert L.fk_convar_num_schedules(48, 1) == 2
    assert L.fk_convar_num_schedules(25, 1) == 0
    assert L.fk_convar_num_schedules(32, 2) == 2
    assert L.fk_convar_num_schedules(32, 3) == 0
    t, w = ctypes.c_int32(), ctypes.c_int32()
    seen = set()
    for i in range(4):
        assert L.fk_convar_schedule_info(32, 1, i, ctypes.byref(t), ctypes.byref(w)) == 0
        seen.add((t.value, w.value))
    assert len(seen) == 4 and (0, 8) in seen
    assert L.fk_convar_schedule_info(32, 1, 4, ctypes.byref(t), ctypes.byref(w)) == 1
    assert L.fk_convar_set_schedule(10, 10, 20, 32, 0, 1, 4) == 1
    assert b'out of range' in L.fk_last_error()
    assert L.fk_convar_get_schedule(10, 10, 20, 32, 0, 1) == 0
    assert L.fk_convar_set_schedule(10, 10, 20, 32, 0, 1, 2) == 0
    assert L.fk_convar_get_schedule(10, 10, 20, 32, 0, 1) == 2
    assert L.fk_convar_get_schedule(10, 9, 20, 32, 0, 1) == 0
    # another network on the same lattice and width keeps its own schedule
    assert L.fk_convar_get_schedule(10, 10, 19, 32, 0, 1) == 0
    assert L.fk_convar_get_schedule(10, 10, 20, 32, 1, 1) == 0
    assert L.fk_convar_set_schedule(10, 10, 20, 32, 0, 1, -1) == 0
    assert L.fk_convar_get_schedule(10, 10, 20, 32, 0, 1) == 0


def test_forward_workspace_query_lds_resident(lib):
    """Host-only: lattices whose activations fit in LDS need no workspace;
    unsupported widths report -1."""
    from flowket_amd import _lib
    L = _lib.load()
    assert L.fk_convar_forward_workspace_bytes(10, 10, 20, 32, 0, 1) == 0
    assert L.fk_convar_forward_workspace_bytes(4, 4, 5, 16, 1, 0) == 0
    assert L.fk_convar_forward_workspace_bytes(10, 10, 20, 66, 0, 1) == -1
