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
    unbalanced/balanced fixture (tests/test_variational.py:18-46), whose
    matrix elements are complex, with the log psi of a linear complex wave
    function: real and imaginary parts of mel both enter (mel / mel_im)."""
    from flowket_amd import _lib
    from tests.test_oracle import reference_fixture
    gpu_banner('E_loc reduction kernels on the reference fixture (complex mel)')
    conn, mel, use, w = reference_fixture(np.random.default_rng(3))
    assert np.iscomplexobj(mel) and np.any(mel.imag != 0)
    nc, B, n = conn.shape
    lp = (conn.reshape(nc * B, n) @ w).astype(np.complex64)          # (nc*B,)

    def psi(x):
        return (x.reshape(x.shape[0], -1) @ w).astype(np.complex64)

    ref = oob.local_values_unbalanced(psi, conn, mel, use)
    mel_re = torch.as_tensor(np.ascontiguousarray(mel.real)).cuda()
    mel_im = torch.as_tensor(np.ascontiguousarray(mel.imag)).cuda()
    # dense: every slot
    lpd = torch.view_as_real(torch.as_tensor(lp)).contiguous().cuda()
    out = torch.empty((B, 2), dtype=torch.float64, device='cuda')
    _lib.call('fk_local_energy_dense', _lib.ptr(lpd), _lib.ptr(mel_re), _lib.ptr(mel_im), nc, B, _lib.ptr(out),
              _lib.stream_ptr())
    got_dense = out[:, 0].cpu().numpy() + 1j * out[:, 1].cpu().numpy()
    assert np.allclose(got_dense, oob.local_values_balanced(psi, conn, mel), rtol=1e-6)
    # without the imaginary parts the result differs (the test would catch a
    # dropped mel_im)
    _lib.call('fk_local_energy_dense', _lib.ptr(lpd), _lib.ptr(mel_re), None, nc, B, _lib.ptr(out),
              _lib.stream_ptr())
    got_re = out[:, 0].cpu().numpy() + 1j * out[:, 1].cpu().numpy()
    assert np.allclose(got_re, oob.local_values_balanced(psi, conn, mel.real), rtol=1e-6)
    assert not np.allclose(got_re, got_dense, rtol=1e-3)
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
    _lib.call('fk_local_energy', _lib.ptr(offsets), _lib.ptr(pairs), _lib.ptr(lpr), _lib.ptr(mel_re),
              _lib.ptr(mel_im), B, _lib.ptr(out), _lib.stream_ptr())
    got = out[:, 0].cpu().numpy() + 1j * out[:, 1].cpu().numpy()
    assert np.allclose(got, ref, rtol=1e-6)
    # the fixture's point: balanced and unbalanced agree on the mean
    assert np.allclose(np.mean(ref), np.mean(oob.local_values_balanced(
        lambda x: (x.reshape(x.shape[0], -1) @ w).astype(np.complex64), conn, mel)))


class _SigmaYField:
    """H = sum_i sigma^y_i on a small lattice, an operator given only by a
    numpy find_conn (as a subclass of Operator or a netket wrapper would be):
    connection 0 is the sample itself (diagonal element 0, as the
    reference's estimators require), connection 1 + i the sample with site i
    flipped, with matrix element <s|sigma^y_i|s'> = -i s_i (sigma^y =
    [[0, -i], [i, 0]] on (up, down))."""

    @staticmethod
    def make(shape):
        import flowket_amd as fk

        class SigmaYField(fk.Operator):
            def find_conn(self, sample):
                B = sample.shape[0]
                x = sample.reshape(B, -1)
                n = x.shape[1]
                conn = np.repeat(x[None], n + 1, axis=0).astype(np.float64)
                for i in range(n):
                    conn[1 + i, :, i] *= -1
                mel = np.zeros((n + 1, B), dtype=np.complex128)
                mel[1:] = -1j * x.T
                use = np.ones((n + 1, B), dtype=bool)
                return conn.reshape((n + 1, B) + sample.shape[1:]), mel, use

        return SigmaYField(shape)


def test_complex_matrix_elements_rayleigh_quotient():
    """Observable.local_values with an operator whose matrix elements are
    purely imaginary: sum_s |psi(s)|^2 E_loc(s) over all 2^9 states equals
    psi^H H psi / psi^H psi with H assembled independently from Kronecker
    products of the Pauli matrix (the real parts alone give 0)."""
    import flowket_amd as fk
    gpu_banner('complex mel: sigma^y field, Rayleigh quotient')
    shape = (3, 3)
    n = 9
    m = fk.ConvNetAutoregressive2D(shape, depth=2, num_of_channels=16, activation='elu',
                                   weights_normalization=False, precision='fp32', seed=5)
    op = _SigmaYField.make(shape)
    states = np.array([[1 if (k >> i) & 1 else -1 for i in range(n)] for k in range(2 ** n)], dtype=np.int8)
    el = fk.Observable(op).local_values(m, states.reshape((-1,) + shape)).cpu().numpy()
    lp, _ = m.forward(torch.as_tensor(states).cuda())
    lp = lp.double().cpu().numpy()
    psi = np.exp(lp[:, 0] + 1j * lp[:, 1])
    p = np.abs(psi) ** 2
    # H from Kronecker products; state k has site i = +1 (up, basis index 0)
    # when bit i of k is set
    sy = np.array([[0, -1j], [1j, 0]])
    eye = np.eye(2)
    Hm = np.zeros((2 ** n, 2 ** n), dtype=np.complex128)
    for i in range(n):
        ops = [sy if q == i else eye for q in range(n)]
        term = ops[0]
        for o in ops[1:]:
            term = np.kron(term, o)
        Hm += term
    # Kronecker index: site 0 is the most significant factor, basis 0 = up
    kron_index = np.array([sum((0 if states[k, q] == 1 else 1) << (n - 1 - q) for q in range(n))
                           for k in range(2 ** n)])
    v = np.zeros(2 ** n, dtype=np.complex128)
    v[kron_index] = psi
    rq = np.vdot(v, Hm @ v) / np.vdot(v, v)
    e = np.sum(p * el) / np.sum(p)
    assert abs(rq) > 1e-3
    assert abs(e - rq) < 1e-5 * max(1.0, abs(rq)), (e, rq)
