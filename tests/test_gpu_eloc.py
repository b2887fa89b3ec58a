"""Local energies on the GPU vs the oracle E_loc with oracle log psi."""
import numpy as np
import pytest
import torch

from oracle import convnet as ocn
from oracle import observables as oob
from oracle import operators as oop
from tests.conftest import gpu_banner
from tests.test_gpu_convnet import _machine, _oracle_params

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('opname,shape,prec', [('heisenberg', (6, 6), 'fp32'), ('ising', (5, 5), 'fp32'),
                                               ('j1j2', (6, 6), 'fp32'), ('heisenberg', (10, 10), 'bf16')])
def test_local_energy(rng, opname, shape, prec):
    import flowket_amd as fk
    gpu_banner(f'eloc {opname} {shape} {prec}')
    depth, C, act = 3, 32, 'elu'
    m = _machine(shape, depth, C, act, prec, bias_scale=0.1)
    B = 64
    s = rng.choice([-1, 1], size=(B,) + shape).astype(np.int8)
    if opname == 'heisenberg':
        op = fk.Heisenberg(hilbert_state_shape=shape, pbc=True)
        conn, mel, use = oop.heisenberg_find_conn(s.astype(np.float64), pbc=True)
    elif opname == 'ising':
        op = fk.Ising(h=3.0, hilbert_state_shape=shape, pbc=False)
        conn, mel, use = oop.ising_find_conn(s.astype(np.float64), h=3.0, pbc=False)
    else:
        op = fk.J1J2(hilbert_state_shape=shape, j2=0.5, pbc=True)
        conn, mel, use = oop.j1j2_find_conn(s.astype(np.float64), j2=0.5, pbc=True)
    params = _oracle_params(m)

    def wf(x):
        return ocn.log_psi(params, x, depth, activation=act).astype(np.complex64)

    ref = oob.local_values_unbalanced(wf, conn, mel, use)
    obs = fk.Observable(op)
    got = obs.local_values(m, s).cpu().numpy()
    scale = np.abs(ref).max()
    tol = 1e-4 if prec == 'fp32' else 5e-2
    assert np.abs(got - ref).max() / scale < tol
    mean, var, lv = obs.estimate(m, s)
    rmean, rvar, _ = oob.estimate(got)
    assert abs(mean - rmean) <= 1e-12 * max(1.0, abs(rmean))
    assert abs(var - rvar) <= 1e-10 * max(1.0, rvar)


def test_balanced_equals_unbalanced(rng):
    """tests/test_variational.py:18-46 on the GPU: the dense and ragged paths
    agree."""
    import flowket_amd as fk
    gpu_banner('eloc balanced vs unbalanced')
    m = _machine((5, 6), 3, 16, 'relu', 'fp32')
    op = fk.Heisenberg(hilbert_state_shape=(5, 6), pbc=True)
    s = rng.choice([-1, 1], size=(40, 5, 6)).astype(np.int8)
    obs = fk.Observable(op)
    a = obs.local_values(m, s)
    b = obs.local_values_balanced(lambda x: m.log_psi(x), obs._spins(s))
    assert torch.allclose(a, b, rtol=1e-12, atol=1e-12)


def test_autoregressive_sampler_matches_oracle_rule(rng):
    """AutoregressiveSampler with injected uniforms == the reference rule on
    oracle conditionals, except where u lies within 1e-6 of exp(log p)."""
    import flowket_amd as fk
    from oracle import samplers as osa
    gpu_banner('AR sampler 4x4 C16 vs oracle')
    m = _machine((4, 4), 3, 16, 'elu', 'fp32', bias_scale=0.5)
    B = 256
    u = rng.random((B, 4, 4))
    sampler = fk.AutoregressiveSampler(m, B, uniforms=u)
    got = next(sampler).cpu().numpy().astype(np.float64)
    params = _oracle_params(m)
    ref = osa.ar_sample(lambda x: ocn.conditional_log_probs(params, x, 3, activation='elu'), u)
    # the draw margins along the reference trajectory
    lp = ocn.conditional_log_probs(params, ref, 3, activation='elu')
    # conditional of site k only depends on earlier sites, so the final batch gives all margins
    margin = np.abs(np.exp(lp[..., 0]) - u).reshape(B, -1).min(axis=1)
    ok = margin > 1e-6
    assert ok.mean() > 0.95
    assert np.array_equal(got[ok], ref[ok])


def test_reduction_kernels_on_reference_fixture():
    """The E_loc reductions (ragged fk_local_energy over the compacted use
    mask, and fk_local_energy_dense over every slot) on the reference's
    unbalanced/balanced fixture (tests/test_variational.py:18-46) with the
    log psi of a linear complex wave function."""
    from flowket_amd import _lib
    from tests.test_oracle import reference_fixture
    gpu_banner('E_loc reduction kernels on the reference fixture')
    conn, mel, use, w = reference_fixture(np.random.default_rng(3))
    nc, B, n = conn.shape
    lp = (conn.reshape(nc * B, n) @ w).astype(np.complex64)          # (nc*B,)
    ref = oob.local_values_unbalanced(lambda x: (x.reshape(x.shape[0], -1) @ w).astype(np.complex64),
                                      conn, mel, use)
    # dense: every slot
    lpd = torch.view_as_real(torch.as_tensor(lp)).contiguous().cuda()
    # the operators of the path have real matrix elements: the kernels carry
    # float64 mel, so the fixture's real parts are used
    mel_re = torch.as_tensor(np.ascontiguousarray(mel.real)).cuda()
    out = torch.empty((B, 2), dtype=torch.float64, device='cuda')
    _lib.call('fk_local_energy_dense', _lib.ptr(lpd), _lib.ptr(mel_re), nc, B, _lib.ptr(out), _lib.stream_ptr())
    got_dense = out[:, 0].cpu().numpy() + 1j * out[:, 1].cpu().numpy()
    ref_re = oob.local_values_balanced(lambda x: (x.reshape(x.shape[0], -1) @ w).astype(np.complex64), conn,
                                       mel.real)
    assert np.allclose(got_dense, ref_re, rtol=1e-6)
    # ragged: compact the use mask, evaluate only used slots
    use_d = torch.as_tensor(use.astype(np.uint8)).contiguous().cuda()
    offsets = torch.empty(B + 1, dtype=torch.int64, device='cuda')
    pairs = torch.empty((nc * B, 2), dtype=torch.int32, device='cuda')
    _lib.call('fk_compact_connections', _lib.ptr(use_d), B, nc, _lib.ptr(offsets), _lib.ptr(pairs), None,
              _lib.stream_ptr())
    M = int(offsets[-1])
    assert M == int(use.sum())
    pr = pairs[:M].cpu().numpy()
    lpr = torch.view_as_real(torch.as_tensor(lp.reshape(nc, B)[pr[:, 1], pr[:, 0]])).contiguous().cuda()
    _lib.call('fk_local_energy', _lib.ptr(offsets), _lib.ptr(pairs), _lib.ptr(lpr), _lib.ptr(mel_re), B,
              _lib.ptr(out), _lib.stream_ptr())
    got = out[:, 0].cpu().numpy() + 1j * out[:, 1].cpu().numpy()
    ref_u = oob.local_values_unbalanced(lambda x: (x.reshape(x.shape[0], -1) @ w).astype(np.complex64), conn,
                                        mel.real, use)
    assert np.allclose(got, ref_u, rtol=1e-6)
    # the fixture's point: balanced and unbalanced agree on the mean
    assert np.allclose(np.mean(ref), np.mean(oob.local_values_balanced(
        lambda x: (x.reshape(x.shape[0], -1) @ w).astype(np.complex64), conn, mel)))
