"""Exact enumeration path on the GPU (ExactVariational / ExactObservable /
ExactSampler) against the FP64 oracle restatement and closed forms."""
import numpy as np
import pytest
import torch

from oracle import convnet as ocn
from oracle import exact as oex
from oracle import operators as oop
from tests.conftest import gpu_banner
from tests.test_gpu_convnet import _machine, _oracle_params

pytestmark = pytest.mark.gpu


def _ops(fk, name, shape):
    if name == 'heisenberg':
        return (fk.Heisenberg(hilbert_state_shape=shape, pbc=True),
                lambda s: oop.heisenberg_find_conn(s, pbc=True))
    if name == 'ising':
        return (fk.Ising(h=0.7, j=1.3, hilbert_state_shape=shape, pbc=True),
                lambda s: oop.ising_find_conn(s, h=0.7, j=1.3, pbc=True))
    return (fk.J1J2(hilbert_state_shape=shape, j2=0.5, pbc=False),
            lambda s: oop.j1j2_find_conn(s, j2=0.5, pbc=False))


def test_exact_states_enumeration():
    from flowket_amd import _lib
    gpu_banner('exact states 3x4')
    n = 12
    st = torch.empty((2 ** n, n), dtype=torch.int8, device='cuda')
    _lib.call('fk_exact_states', n, 0, 2 ** n, _lib.ptr(st), _lib.stream_ptr())
    ref = oex.all_states((3, 4)).reshape(2 ** n, n)
    assert np.array_equal(st.cpu().numpy(), ref.astype(np.int8))
    # a sub-range
    part = torch.empty((100, n), dtype=torch.int8, device='cuda')
    _lib.call('fk_exact_states', n, 1000, 100, _lib.ptr(part), _lib.stream_ptr())
    assert np.array_equal(part.cpu().numpy(), ref[1000:1100].astype(np.int8))


@pytest.mark.parametrize('name', ['heisenberg', 'ising', 'j1j2'])
def test_exact_variational_vs_oracle(name):
    """Probabilities, local energies, energy, variance and gradient
    coefficients of a 4x4 ConvNetAutoregressive2D (FP32) over all 65,536
    configurations, against the oracle fed the oracle network's log psi."""
    import flowket_amd as fk
    gpu_banner(f'exact variational 4x4 {name}')
    shape = (4, 4)
    depth, C, act = 3, 16, 'elu'
    m = _machine(shape, depth, C, act, 'fp32', seed=3, bias_scale=0.2)
    op, fc = _ops(fk, name, shape)
    ev = fk.ExactVariational(m, op, batch_size=4096)
    ev.machine_updated()
    params = _oracle_params(m)
    lp = ocn.log_psi(params, oex.all_states(shape), depth, activation=act)
    r = oex.exact_variational(lp, fc, shape)
    probs = ev.probs.cpu().numpy()
    assert np.abs(probs - r['probs']).max() < 1e-5 * r['probs'].max()
    assert abs(probs.sum() - 1.0) < 1e-12
    el = ev.energy_observable.naive_energies.cpu().numpy()
    scale = np.abs(r['eloc']).max()
    assert np.abs(el - r['eloc']).max() < 1e-4 * scale
    e = ev.energy_observable.current_energy
    assert abs(e - r['current_energy']) < 1e-5 * max(1.0, abs(r['current_energy']))
    v = ev.energy_observable.current_local_energy_variance
    assert v == pytest.approx(r['variance'], rel=1e-4)
    coef = ev.energy_grad_coefficients.cpu().numpy()
    assert np.abs(coef - r['coefficients']).max() < 1e-5 * np.abs(r['coefficients']).max()
    # the exact counterpart of the estimator's coefficients sums to zero
    assert abs(coef.sum()) < 1e-12
    # the generator yields every configuration once per update
    gen = iter(ev)
    xs, cs = zip(*[next(gen) for _ in range(ev.num_of_batch_until_full_cycle)])
    assert sum(x.shape[0] for x in xs) == 2 ** 16
    assert np.allclose(torch.cat(cs).cpu().numpy(), coef)


def test_exact_local_energy_at_the_ground_state():
    """log psi = the exact transverse-field Ising ground state (3x4, every
    amplitude positive): every local energy equals E0 and the variance and the
    coefficients vanish, to FP32 rounding of log psi."""
    import flowket_amd as fk
    from flowket_amd import _lib
    gpu_banner('exact local energy at the ground state')
    shape = (3, 4)
    fc = lambda s: oop.ising_find_conn(s, h=1.0, j=1.0, pbc=True)  # noqa: E731
    H = oex.sparse_hamiltonian(fc, shape)
    import scipy.sparse.linalg
    w, v = scipy.sparse.linalg.eigsh(H, k=1, which='SA', tol=1e-14)
    psi = v[:, 0] * np.sign(v[np.argmax(np.abs(v[:, 0])), 0])
    N = 2 ** 12
    lp = torch.zeros((N, 2), dtype=torch.float32)
    lp[:, 0] = torch.from_numpy(np.log(psi)).float()
    lp = lp.cuda()
    op = fk.Ising(h=1.0, j=1.0, hilbert_state_shape=shape, pbc=True)
    st = torch.empty((N, 12), dtype=torch.int8, device='cuda')
    _lib.call('fk_exact_states', 12, 0, N, _lib.ptr(st), _lib.stream_ptr())
    rc = op.find_conn_device(st, want_conn=False)
    tb = op.tables(st.device)
    eloc = torch.empty((N, 2), dtype=torch.float64, device='cuda')
    _lib.call('fk_exact_local_energy', _lib.ptr(lp), N, 12, op.kind, _lib.ptr(tb['conn_a']), _lib.ptr(tb['conn_b']),
              _lib.ptr(rc['mel']), _lib.ptr(rc['use']), op.num_connections, _lib.ptr(eloc), _lib.stream_ptr())
    ws = torch.empty(int(_lib.load().fk_exact_workspace_doubles(N)), dtype=torch.float64, device='cuda')
    probs = torch.empty(N, dtype=torch.float64, device='cuda')
    sc = torch.empty(2, dtype=torch.float64, device='cuda')
    _lib.call('fk_exact_normalize', _lib.ptr(lp), N, _lib.ptr(probs), _lib.ptr(sc), _lib.ptr(ws), _lib.stream_ptr())
    mom = torch.empty(3, dtype=torch.float64, device='cuda')
    coef = torch.empty((N, 2), dtype=torch.float64, device='cuda')
    _lib.call('fk_exact_energy', _lib.ptr(probs), _lib.ptr(eloc), N, _lib.ptr(coef), _lib.ptr(mom), _lib.ptr(ws),
              _lib.stream_ptr())
    el = eloc.cpu().numpy()
    assert np.abs(el[:, 0] - w[0]).max() < 1e-5 * abs(w[0])
    assert np.abs(el[:, 1]).max() == 0.0
    mom = mom.cpu().numpy()
    assert mom[0] == pytest.approx(w[0], rel=1e-6)
    assert mom[2] < 1e-8 * w[0] ** 2
    assert np.abs(probs.cpu().numpy() - psi ** 2).max() < 1e-6


def test_exact_sampler_vs_oracle_rule(rng):
    """ExactSampler on injected uniforms picks the oracle's searchsorted
    index, except for uniforms within 1e-9 of a CDF boundary."""
    import flowket_amd as fk
    gpu_banner('exact sampler 4x4')
    shape = (4, 4)
    m = fk.RBM(shape, alpha=2, stddev=0.3, seed=4)
    op, _ = _ops(fk, 'heisenberg', shape)
    ev = fk.ExactVariational(m, op, batch_size=65536)
    ev.machine_updated()
    B = 20000
    u = rng.random(B)
    s = fk.ExactSampler(ev, B, uniforms=u)
    x = next(s).reshape(B, 16).cpu().numpy()
    got = oex.binary_array_to_decimal_array(x)
    ref, cdf = oex.exact_sample(ev.probs.cpu().numpy(), u)
    near = np.min(np.abs(cdf[None, :] - u[:, None]), axis=1) < 1e-9
    assert near.mean() < 0.01
    assert np.array_equal(got[~near], ref[~near])
    # and the drawn states follow |psi|^2 (seeded generator)
    s2 = fk.ExactSampler(ev, B, seed=5)
    y = oex.binary_array_to_decimal_array(next(s2).reshape(B, 16).cpu().numpy())
    p = ev.probs.cpu().numpy()
    top = np.argsort(p)[-8:]
    counts = np.array([(y == k).sum() for k in top])
    z = (counts - B * p[top]) / np.sqrt(B * p[top] * (1 - p[top]))
    assert np.abs(z).max() < 5


def test_sampled_energy_agrees_with_exact():
    """8192 exact autoregressive samples estimate the exact energy of the same
    psi (4x4, depth 5, 16 channels, ELU) within 5 standard errors, with the
    exact variance; the exact energy lies above the ED ground energy."""
    import flowket_amd as fk
    gpu_banner('sampled vs exact energy 4x4')
    m = fk.ConvNetAutoregressive2D((4, 4), depth=5, num_of_channels=16, activation='elu',
                                   weights_normalization=False, precision='fp32', seed=0)
    op = fk.Heisenberg(hilbert_state_shape=(4, 4), pbc=True)
    ev = fk.ExactVariational(m, op, batch_size=16384)
    ev.machine_updated()
    E = ev.energy_observable.current_energy.real
    var = ev.energy_observable.current_local_energy_variance
    B = 8192
    x = next(fk.AutoregressiveSampler(m, B, seed=3))
    el = fk.Observable(op).local_values(m, x).real.cpu().numpy()
    assert abs(el.mean() - E) < 5 * np.sqrt(var / B)
    assert el.var() == pytest.approx(var, rel=0.1)
    assert E > -44.9139328337


def test_exact_evaluate_logs():
    """exact_evaluate: the energy logs equal the exact observable's, and
    sigma_z / |sigma_z| equal sum_k p_k m_k over all configurations."""
    import flowket_amd as fk
    from flowket_amd.evaluation import exact_evaluate
    gpu_banner('exact evaluate 4x4')
    m = fk.RBM((4, 4), alpha=2, stddev=0.3, seed=9)
    op = fk.Ising(h=1.0, hilbert_state_shape=(4, 4), pbc=True)
    ev = fk.ExactVariational(m, op, batch_size=65536)
    logs = exact_evaluate(ev, true_ground_state_energy=-34.0)
    e = ev.energy_observable.current_energy.real
    assert logs['energy/energy'] == e
    assert logs['energy/relative_error'] == pytest.approx((-34.0 - e) / -34.0, rel=1e-12)
    p = ev.probs.cpu().numpy()
    mz = oex.all_states((4, 4)).reshape(65536, 16).sum(axis=1) / 16.0
    assert logs['observables/sigma_z'] == pytest.approx(float(np.dot(p, mz)), abs=1e-12)
    assert logs['observables/abs_sigma_z'] == pytest.approx(float(np.dot(p, np.abs(mz))), abs=1e-12)


def test_exact_sampler_checks_injected_uniforms():
    """A uniform array of the wrong length is refused before any launch."""
    import flowket_amd as fk
    gpu_banner('exact sampler uniforms length check')
    m = fk.ConvNetAutoregressive2D((2, 3), depth=2, num_of_channels=16, weights_normalization=False, seed=0)
    op = fk.Heisenberg(hilbert_state_shape=(2, 3), pbc=False)
    ev = fk.ExactVariational(m, op, 64)
    with pytest.raises(ValueError):
        next(fk.ExactSampler(ev, 16, uniforms=np.full(15, 0.5)))
    with pytest.raises(ValueError):
        next(fk.AutoregressiveSampler(m, 4, uniforms=np.full((4, 5), 0.5)))
