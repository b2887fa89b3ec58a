"""Fused ConvNetAutoregressive2D forward on the GPU vs the float64 oracle."""
import numpy as np
import pytest
import torch

from oracle import convnet as ocn
from oracle import exact as oex
from tests.conftest import gpu_banner

pytestmark = pytest.mark.gpu

# FP32 kernels: log psi within 1e-5 relative; BF16: 2e-2 relative (tolerances
# of the FP32 and BF16 paths).
FP32_RTOL = 1e-5
BF16_RTOL = 2e-2
FP16_RTOL = 3e-3
RTOL = {'fp32': FP32_RTOL, 'bf16': BF16_RTOL, 'fp16': FP16_RTOL}


def _oracle_params(machine):
    flat = machine.effective_params().detach().double().cpu().numpy()
    out, off = [], 0
    for kshape, bshape in machine.shapes:
        nk, nb = int(np.prod(kshape)), int(np.prod(bshape))
        out.append((flat[off:off + nk].reshape(kshape), flat[off + nk:off + nk + nb]))
        off += nk + nb
    return out


def _machine(shape, depth, C, act, prec, seed=0, bias_scale=0.3, wn=False):
    import flowket_amd as fk
    m = fk.ConvNetAutoregressive2D(shape, depth=depth, num_of_channels=C, activation=act,
                                   weights_normalization=wn, precision=prec, seed=seed)
    if bias_scale:
        # non-zero biases so every bias path is exercised
        g = torch.Generator().manual_seed(seed + 1)
        p = m.params.cpu().clone()
        off = 0
        for kshape, bshape in m.shapes:
            nk, nb = int(np.prod(kshape)), int(np.prod(bshape))
            p[off + nk:off + nk + nb] = bias_scale * torch.randn(nb, generator=g)
            off += nk + nb
        m.set_params(p)
    return m


def _check_logpsi(got, ref, rtol):
    got = got.detach().cpu().numpy().astype(np.complex128)
    scale = np.abs(ref.real).max() + 1.0
    err_re = np.abs(got.real - ref.real).max() / scale
    # phases: compare directly (absolute, same scale)
    err_im = np.abs(got.imag - ref.imag).max() / scale
    assert err_re < rtol and err_im < rtol, (err_re, err_im)


CASES = [
    ((4, 4), 5, 16, 'elu', 'fp32'),
    ((4, 4), 2, 16, 'relu', 'fp32'),
    ((4, 3), 2, 16, 'relu', 'fp32'),
    ((6, 5), 3, 32, 'relu', 'fp32'),
    ((10, 10), 4, 32, 'elu', 'fp32'),
    ((10, 10), 4, 32, 'elu', 'bf16'),
    ((7, 9), 3, 32, 'relu', 'bf16'),
    ((10, 10), 3, 64, 'elu', 'fp32'),
    ((10, 10), 3, 64, 'elu', 'bf16'),
    ((4, 4), 1, 32, 'relu', 'bf16'),
    # widths without their own kernel: zero-padded to the next compiled one
    ((10, 10), 3, 48, 'elu', 'bf16'),
    ((6, 6), 2, 24, 'relu', 'fp32'),
    ((5, 4), 2, 8, 'elu', 'fp32'),
    ((4, 4), 2, 6, 'relu', 'bf16'),
    # lattices whose five activation buffers exceed the LDS: the horizontal
    # stack lives in global scratch
    ((16, 16), 3, 64, 'relu', 'bf16'),
    ((20, 20), 2, 64, 'elu', 'fp16'),
    ((16, 16), 2, 32, 'relu', 'fp32'),
    # fp16 MFMA
    ((10, 10), 4, 32, 'elu', 'fp16'),
    ((7, 9), 3, 64, 'relu', 'fp16'),
    ((6, 5), 2, 20, 'elu', 'fp16'),
]


@pytest.mark.parametrize('shape,depth,C,act,prec', CASES)
def test_forward_matches_oracle(rng, shape, depth, C, act, prec):
    gpu_banner(f'convnet forward shape={shape} depth={depth} C={C} act={act} prec={prec}')
    m = _machine(shape, depth, C, act, prec)
    B = 67
    s = rng.choice([-1, 1], size=(B,) + shape).astype(np.int8)
    got = m.log_psi(s)
    ref = ocn.log_psi(_oracle_params(m), s, depth, activation=act)
    _check_logpsi(got, ref, RTOL[prec])
    cond = m.conditional_log_probs(s).cpu().numpy()
    rcond = ocn.conditional_log_probs(_oracle_params(m), s, depth, activation=act)
    tol = {'fp32': 1e-4, 'bf16': 5e-2, 'fp16': 8e-3}[prec]
    assert np.abs(cond - rcond).max() < tol


def test_forward_weight_normalisation(rng):
    gpu_banner('convnet forward with weight normalisation')
    m = _machine((6, 6), 3, 32, 'relu', 'fp32', wn=True)
    s = rng.choice([-1, 1], size=(40, 6, 6)).astype(np.int8)
    got = m.log_psi(s)
    ref = ocn.log_psi(_oracle_params(m), s, 3, activation='relu')
    _check_logpsi(got, ref, FP32_RTOL)


def test_normalised_distribution_4x4():
    """tests/test_autoregressive.py:18-30: sum_s |psi(s)|^2 == 1 (1e-5)."""
    gpu_banner('convnet normalisation 4x4 depth 4 C 16')
    m = _machine((4, 4), 4, 16, 'relu', 'fp32')
    states = oex.all_states((4, 4)).astype(np.int8)
    lp = m.log_psi(states).cpu().numpy().astype(np.complex128)
    from scipy.special import logsumexp
    assert np.exp(logsumexp(2 * lp.real)) == pytest.approx(1.0, abs=1e-5)


def test_causality(rng):
    """Conditional at site k is unchanged by any perturbation at sites >= k."""
    gpu_banner('convnet causality 5x4')
    m = _machine((5, 4), 3, 16, 'elu', 'fp32')
    s = rng.choice([-1, 1], size=(1, 5, 4)).astype(np.int8)
    base = m.conditional_log_probs(s).cpu().numpy()[0]
    for k in range(20):
        t = s.copy().reshape(1, -1)
        t[0, k:] = rng.choice([-1, 1], size=20 - k)
        c = m.conditional_log_probs(t.reshape(1, 5, 4)).cpu().numpy()[0].reshape(20, 2)
        assert np.abs(c[:k + 1] - base.reshape(20, 2)[:k + 1]).max() < 1e-5


def test_connected_equals_dense(rng):
    """Ragged (sample, connection) evaluation == evaluation of the materialised
    connected configurations."""
    import flowket_amd as fk
    gpu_banner('connected vs dense 8x8 heisenberg bf16+fp32')
    for prec in ('fp32', 'bf16'):
        m = _machine((8, 8), 3, 32, 'elu', prec)
        op = fk.Heisenberg(hilbert_state_shape=(8, 8), pbc=True)
        B = 50
        s = torch.as_tensor(rng.choice([-1, 1], size=(B, 64)).astype(np.int8)).cuda()
        r = op.compact(s)
        total = int(r['offsets'][-1])
        lp = m.log_psi_connected(s, r['pairs'], total, op)
        conn = op.find_conn_device(s)['conn']
        pairs = r['pairs'][:total].long()
        dense = conn[pairs[:, 1], pairs[:, 0]]
        lp2, _ = m.forward(dense.contiguous())
        assert torch.equal(lp, lp2)


@pytest.mark.parametrize('shape,depth,C,act,prec', [((10, 10), 4, 32, 'elu', 'bf16'), ((7, 9), 3, 32, 'relu', 'bf16'),
                                                    ((6, 5), 3, 16, 'relu', 'fp32'), ((10, 10), 3, 64, 'elu', 'bf16'),
                                                    ((4, 4), 2, 32, 'relu', 'fp32'),
                                                    ((8, 7), 3, 32, 'elu', 'fp16')])
def test_every_schedule_matches_oracle(rng, shape, depth, C, act, prec):
    """Each compiled launch schedule (loop order x waves per workgroup)
    computes the same function, on plain rows and on ragged connected
    configurations with the count read from device memory."""
    import flowket_amd as fk
    from flowket_amd import _lib, autotune
    gpu_banner(f'every schedule shape={shape} depth={depth} C={C} {act} {prec}')
    m = _machine(shape, depth, C, act, prec)
    from flowket_amd.machines import _PREC
    pc = _PREC[prec]
    s = rng.choice([-1, 1], size=(45,) + shape).astype(np.int8)
    ref = ocn.log_psi(_oracle_params(m), s, depth, activation=act)
    op = fk.Heisenberg(hilbert_state_shape=shape, pbc=True)
    sd = torch.as_tensor(s.reshape(45, -1)).cuda()
    r = op.compact(sd)
    B = sd.shape[0]
    outs = []
    try:
        for sc in autotune.schedules(C, pc):
            autotune.set_schedule(m, pc, sc['index'])
            got = m.log_psi(s)
            _check_logpsi(got, ref, RTOL[prec])
            lp = m.log_psi_connected(sd, r['pairs'], r['pairs'].shape[0], op, m_dev=r['offsets'][B:])
            outs.append(lp[:int(r['offsets'][-1])].clone())
    finally:
        autotune.set_schedule(m, pc, -1)
    for o in outs[1:]:
        assert torch.allclose(o, outs[0], rtol=1e-5, atol=1e-4)


def test_autotune_selects_and_caches(rng, tmp_path, monkeypatch):
    import flowket_amd as fk
    from flowket_amd import _lib, autotune
    gpu_banner('autotune 8x8 depth 3 C32 bf16')
    monkeypatch.setenv('FLOWKET_AMD_TUNE_CACHE', str(tmp_path / 'tune.json'))
    m = _machine((8, 8), 3, 32, 'relu', 'bf16')
    op = fk.Heisenberg(hilbert_state_shape=(8, 8), pbc=True)
    sd = torch.as_tensor(rng.choice([-1, 1], size=(256, 64)).astype(np.int8)).cuda()
    obs = fk.Observable(op)
    ref = obs.local_values(m, sd)
    try:
        res = autotune.tune_forward(m, lambda: obs.local_values(m, sd), reps=2)
        assert len(res['ms']) == 4 and all(t > 0 for t in res['ms'])
        assert autotune.get_schedule(m, _lib.FK_PREC_BF16) == res['schedule']
        assert (tmp_path / 'tune.json').exists()
        autotune.set_schedule(m, _lib.FK_PREC_BF16, -1)
        assert autotune.apply_cached(m, _lib.FK_PREC_BF16) == res['schedule']
        got = obs.local_values(m, sd)
        assert torch.allclose(got, ref, rtol=1e-4, atol=1e-3)
    finally:
        autotune.set_schedule(m, _lib.FK_PREC_BF16, -1)


def test_fp16_closer_to_fp32_than_bf16(rng):
    """On the same network and configurations the fp16 path's log psi is
    closer to the FP32 result than the bf16 path's (finer rounding)."""
    gpu_banner('fp16 vs bf16 accuracy 10x10 depth 6 C32')
    m = _machine((10, 10), 6, 32, 'relu', 'fp32', bias_scale=0.2)
    s = rng.choice([-1, 1], size=(64, 10, 10)).astype(np.int8)
    ref = m.log_psi(s).real.double()
    e16 = (m.log_psi(s, precision='fp16').real.double() - ref).abs().max().item()
    eb16 = (m.log_psi(s, precision='bf16').real.double() - ref).abs().max().item()
    assert e16 < eb16 / 3, (e16, eb16)


def test_global_scratch_forward_connected_and_workspace(rng):
    """20x20 / 64 channels (the largest workload's lattice): the ragged
    connected evaluation equals the evaluation of the materialised
    configurations, the C ABI without a workspace refuses the lattice with
    FK_ERR_ARG, and the workspace query is consistent."""
    import flowket_amd as fk
    from flowket_amd import _lib
    gpu_banner('global-scratch forward 20x20 C64 bf16 connected')
    m = _machine((20, 20), 2, 64, 'relu', 'bf16')
    n = int(_lib.load().fk_convar_forward_workspace_bytes(20, 20, 2, 64, _lib.FK_ACT_RELU, _lib.FK_PREC_BF16))
    assert n > 0
    assert int(_lib.load().fk_convar_forward_workspace_bytes(10, 10, 2, 32, _lib.FK_ACT_RELU,
                                                             _lib.FK_PREC_BF16)) == 0
    op = fk.Heisenberg(hilbert_state_shape=(20, 20), pbc=True)
    B = 6
    s = torch.as_tensor(rng.choice([-1, 1], size=(B, 400)).astype(np.int8)).cuda()
    r = op.compact(s)
    total = int(r['offsets'][-1])
    lp = m.log_psi_connected(s, r['pairs'], total, op)
    conn = op.find_conn_device(s)['conn']
    pairs = r['pairs'][:total].long()
    dense = conn[pairs[:, 1], pairs[:, 0]]
    lp2, _ = m.forward(dense.contiguous())
    assert torch.equal(lp, lp2)
    wbuf, abuf = m._pack('bf16')
    out = torch.empty((B, 2), dtype=torch.float32, device='cuda')
    rc = _lib.load().fk_convar_forward(_lib.ptr(wbuf), _lib.ptr(abuf), _lib.FK_PREC_BF16, 20, 20, 2, 64,
                                       _lib.FK_ACT_RELU, _lib.ptr(s), None, None, B, 0, None, None,
                                       _lib.ptr(out), None, _lib.stream_ptr())
    assert rc == 1 and b'workspace' in _lib.load().fk_last_error()


def test_global_scratch_forward_both_wave_counts(rng):
    """The 20x20 / 64-channel forward under its 8-wave and 4-wave schedules:
    a tile's arithmetic does not depend on the wave that runs it, so the two
    agree bit for bit (and the 8-wave one is pinned against the oracle by the
    HG cases above)."""
    import flowket_amd as fk
    from flowket_amd import _lib, autotune
    gpu_banner('global-scratch forward 20x20 C64 bf16, 8 vs 4 waves')
    m = _machine((20, 20), 3, 64, 'relu', 'bf16', seed=4)
    s = torch.as_tensor(rng.choice([-1, 1], size=(40, 400)).astype(np.int8)).cuda()
    try:
        autotune.set_schedule(m, _lib.FK_PREC_BF16, 0)
        a, _ = m.forward(s)
        autotune.set_schedule(m, _lib.FK_PREC_BF16, 1)
        assert autotune.schedules(64, _lib.FK_PREC_BF16)[1]['waves'] == 4
        b, _ = m.forward(s)
    finally:
        autotune.set_schedule(m, _lib.FK_PREC_BF16, -1)
    assert torch.equal(a, b)


def test_schedule_is_per_network(rng):
    """Selecting a schedule for one network leaves another network on the same
    lattice and width (different depth) bit-identical."""
    from flowket_amd import _lib, autotune
    gpu_banner('schedule per network')
    a = _machine((10, 10), 4, 32, 'relu', 'bf16', seed=1)
    b = _machine((10, 10), 3, 32, 'relu', 'bf16', seed=2)
    x = rng.choice([-1, 1], size=(64, 10, 10)).astype(np.int8)
    before = b.log_psi(x).clone()
    try:
        for sc in autotune.schedules(32, _lib.FK_PREC_BF16)[1:]:
            autotune.set_schedule(a, _lib.FK_PREC_BF16, sc['index'])
            a.log_psi(x)
            assert autotune.get_schedule(b, _lib.FK_PREC_BF16) == 0
            assert torch.equal(b.log_psi(x), before)
    finally:
        autotune.set_schedule(a, _lib.FK_PREC_BF16, -1)
