"""fk_convar_forward_noalloc (the allocation-free forward) against
fk_convar_forward_ws with a caller workspace, through the C ABI: bit-identical
outputs on two calls in a row, the second with a smaller input, for a lattice
whose horizontal stack lives in the static scratch (20x20, 64 channels) and
one that needs no scratch (10x10, 32 channels); plus the size check."""
import numpy as np
import pytest
import torch

from tests.conftest import gpu_banner

pytestmark = pytest.mark.gpu


def _net(shape, depth, C, prec, seed):
    import flowket_amd as fk
    return fk.ConvNetAutoregressive2D(shape, depth=depth, num_of_channels=C, activation='relu',
                                      weights_normalization=False, precision=prec, seed=seed)


def _call(name, m, spins, pairs, m_count, op, with_cond, ws=None):
    from flowket_amd import _lib
    from flowket_amd.machines import _ACT, _PREC
    prec = _PREC[m.precision]
    wbuf, abuf = m._pack(m.precision)
    out = torch.full((m_count, 2), float('nan'), dtype=torch.float32, device='cuda')
    cond = torch.full((m_count, m.H, m.W, 2), float('nan'), dtype=torch.float32, device='cuda') \
        if with_cond else None
    kind, ca, cb = 0, None, None
    if pairs is not None:
        tb = op.tables(torch.device('cuda'))
        kind, ca, cb = op.kind, tb['conn_a'], tb['conn_b']
    args = [_lib.ptr(wbuf), _lib.ptr(abuf), prec, m.H, m.W, m.depth, m.num_of_channels, _ACT[m.activation],
            _lib.ptr(spins), _lib.ptr(pairs), None, int(m_count), kind, _lib.ptr(ca), _lib.ptr(cb), _lib.ptr(out),
            _lib.ptr(cond)]
    if name == 'fk_convar_forward_ws':
        args += [_lib.ptr(ws), 0 if ws is None else ws.numel()]
    _lib.call(name, *args, _lib.stream_ptr())
    torch.cuda.synchronize()
    return out, cond


@pytest.mark.parametrize('shape,depth,C,prec', [((20, 20), 3, 64, 'bf16'), ((20, 20), 2, 64, 'fp16'),
                                                ((10, 10), 4, 32, 'bf16')])
def test_noalloc_equals_workspace_forward(rng, shape, depth, C, prec):
    import flowket_amd as fk
    from flowket_amd import _lib
    from flowket_amd.machines import _ACT, _PREC
    gpu_banner(f'noalloc forward {shape} depth={depth} C={C} {prec}')
    m = _net(shape, depth, C, prec, seed=2)
    H, W = shape
    n = int(_lib.load().fk_convar_forward_workspace_bytes(H, W, depth, C, _ACT['relu'], _PREC[prec]))
    assert 0 <= n <= int(_lib.load().fk_convar_forward_noalloc_scratch_bytes())
    ws = torch.empty(max(n, 1), dtype=torch.uint8, device='cuda') if n > 0 else None
    op = fk.Heisenberg(hilbert_state_shape=shape, pbc=True)
    for B in (9, 4):   # two calls in a row, the second smaller
        x = torch.as_tensor(rng.choice([-1, 1], size=(B, H * W)).astype(np.int8)).cuda()
        a, ca = _call('fk_convar_forward_ws', m, x, None, B, op, True, ws)
        b, cb = _call('fk_convar_forward_noalloc', m, x, None, B, op, True)
        assert torch.equal(a, b) and torch.equal(ca, cb)
        assert torch.isfinite(a).all()
        r = op.compact(x)
        total = int(r['offsets'][-1])
        pairs = r['pairs'][:total].contiguous()
        a, _ = _call('fk_convar_forward_ws', m, x, pairs, total, op, False, ws)
        b, _ = _call('fk_convar_forward_noalloc', m, x, pairs, total, op, False)
        assert torch.equal(a, b)


def test_noalloc_size_limit():
    """The static buffer is 32 MiB; a network whose scratch exceeds it is
    refused before anything is launched (checked on the host query)."""
    from flowket_amd import _lib
    gpu_banner('noalloc size limit')
    L = _lib.load()
    cap = int(L.fk_convar_forward_noalloc_scratch_bytes())
    assert cap == 32 << 20
    need = int(L.fk_convar_forward_workspace_bytes(20, 20, 24, 64, 0, _lib.FK_PREC_BF16))
    assert 0 < need <= cap
    rc = L.fk_convar_forward_noalloc(None, None, _lib.FK_PREC_BF16, 0, 20, 24, 64, 0, None, None, None, 1, 0, None,
                                     None, None, None, None)
    assert rc == 1 and b'bad shape' in L.fk_last_error()
