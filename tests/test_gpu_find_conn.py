"""find_conn and the ragged index map on the GPU vs the oracle: bit-exact."""
import numpy as np
import pytest
import torch

from oracle import operators as oop
from tests.conftest import gpu_banner

pytestmark = pytest.mark.gpu


def _spins(rng, B, shape):
    return rng.choice([-1, 1], size=(B,) + tuple(shape)).astype(np.int8)


HEIS_CASES = [((4, 4), True), ((4, 4), False), ((10, 10), True), ((3, 5), False), ((2, 3), True), ((7,), True),
              ((7,), False), ((1, 6), True), ((2, 2), True), ((20, 20), True)]


@pytest.mark.parametrize('shape,pbc', HEIS_CASES)
@pytest.mark.parametrize('rotation', [True, False])
def test_heisenberg_bit_exact(rng, shape, pbc, rotation):
    import flowket_amd as fk
    gpu_banner(f'heisenberg shape={shape} pbc={pbc} rot={rotation}')
    B = 37
    s = _spins(rng, B, shape)
    op = fk.Heisenberg(hilbert_state_shape=shape, pbc=pbc, unitary_rotation=rotation)
    conn, mel, use = op.find_conn(s)
    rc, rm, ru = oop.heisenberg_find_conn(s.astype(np.float64), pbc=pbc, unitary_rotation=rotation)
    assert conn.shape == rc.shape and op.max_number_of_local_connections == rc.shape[0]
    assert np.array_equal(conn, rc)
    assert np.array_equal(mel, rm)
    assert np.array_equal(use, ru)


ISING_CASES = [((4, 4), True, 3.0, 1.0), ((4, 4), False, 2.5, 1.0), ((12, 12), False, 3.5, 1.0), ((16,), False, 3.0, 1.0),
               ((16,), True, 1.0, 2.0), ((3, 1), True, 1.0, 1.0), ((5, 6), True, 0.75, -0.5),
               # non-dyadic couplings: the reference's float32 diagonal rule (ising.py:25-41)
               ((10, 10), True, 3.0, 0.3), ((20, 20), False, 1.0, 0.3), ((300,), True, 1.0, 0.7),
               ((12, 12), False, 2.0, 1.1), ((9, 1), False, 1.0, 0.1), ((17, 13), True, 0.5, -0.37)]


@pytest.mark.parametrize('shape,pbc,h,j', ISING_CASES)
def test_ising_bit_exact(rng, shape, pbc, h, j):
    import flowket_amd as fk
    gpu_banner(f'ising shape={shape} pbc={pbc} h={h} j={j}')
    s = _spins(rng, 29, shape)
    op = fk.Ising(h=h, j=j, hilbert_state_shape=shape, pbc=pbc)
    conn, mel, use = op.find_conn(s)
    rc, rm, ru = oop.ising_find_conn(s.astype(np.float64), h=h, j=j, pbc=pbc)
    assert np.array_equal(conn, rc.astype(np.float64))
    assert np.array_equal(mel, rm)
    assert np.array_equal(use, ru)


@pytest.mark.parametrize('shape,pbc,j2', [((4, 4), False, 0.5), ((4, 4), True, 0.5), ((10, 10), True, 0.5),
                                          ((3, 5), True, 0.25), ((4, 6), False, 0.0)])
def test_j1j2_bit_exact(rng, shape, pbc, j2):
    import flowket_amd as fk
    gpu_banner(f'j1j2 shape={shape} pbc={pbc} j2={j2}')
    s = _spins(rng, 23, shape)
    op = fk.J1J2(hilbert_state_shape=shape, j2=j2, pbc=pbc)
    conn, mel, use = op.find_conn(s)
    rc, rm, ru = oop.j1j2_find_conn(s.astype(np.float64), j2=j2, pbc=pbc)
    assert np.array_equal(conn, rc)
    assert np.array_equal(mel, rm)
    assert np.array_equal(use, ru)


@pytest.mark.parametrize('opname', ['heisenberg', 'ising', 'j1j2'])
def test_compact_index_map(rng, opname):
    import flowket_amd as fk
    gpu_banner(f'compact {opname}')
    shape = (6, 5)
    B = 3001
    s = _spins(rng, B, shape)
    if opname == 'heisenberg':
        op = fk.Heisenberg(hilbert_state_shape=shape, pbc=True)
        _, _, use = oop.heisenberg_find_conn(s.astype(np.float64), pbc=True)
    elif opname == 'ising':
        op = fk.Ising(h=1.0, hilbert_state_shape=shape, pbc=True)
        _, _, use = oop.ising_find_conn(s.astype(np.float64), h=1.0, pbc=True)
    else:
        op = fk.J1J2(hilbert_state_shape=shape, j2=0.5, pbc=True)
        _, _, use = oop.j1j2_find_conn(s.astype(np.float64), j2=0.5, pbc=True)
    d = torch.as_tensor(s).cuda().reshape(B, -1)
    r = op.compact(d)
    counts, offsets, b_idx, c_idx = oop.compact_connections(use)
    off = r['offsets'].cpu().numpy()
    total = int(off[-1])
    assert np.array_equal(off[:-1], offsets) and total == counts.sum()
    assert np.array_equal(r['n_used'].cpu().numpy(), counts)
    pairs = r['pairs'][:total].cpu().numpy()
    assert np.array_equal(pairs[:, 0], b_idx) and np.array_equal(pairs[:, 1], c_idx)


def test_empty_and_single(rng):
    import flowket_amd as fk
    gpu_banner('find_conn empty/single batch')
    op = fk.Heisenberg(hilbert_state_shape=(4, 4), pbc=True)
    conn, mel, use = op.find_conn(np.zeros((0, 4, 4), np.int8))
    assert conn.shape == (33, 0, 4, 4) and mel.shape == (33, 0)
    s = _spins(rng, 1, (4, 4))
    conn, mel, use = op.find_conn(s)
    rc, rm, ru = oop.heisenberg_find_conn(s.astype(np.float64), pbc=True)
    assert np.array_equal(conn, rc) and np.array_equal(mel, rm) and np.array_equal(use, ru)
