"""The networks of the benchmark configurations at their real depth and width,
against the FP64 oracle (oracle/convnet_torch.py in float64, equal to the numpy
oracle to 1e-12):

  10x10 periodic Heisenberg, depth 20, 32 channels (log psi, conditionals,
        local energies; fp32 / fp16 / bf16);
  10x10 periodic J1-J2 (J2 = 0.5), depth 20, 48 channels, through the
        graph-captured ragged local-energy step (fp16 / bf16);
  20x20 periodic Heisenberg, depth 24, 64 channels (log psi of samples and
        of ragged connected configurations; fp16 / bf16).

Weights are the seeded Glorot initialisation with zero biases (Keras Conv2D
defaults), configurations come from the network's own fast sampler.

Bars (each test states which one it checks):
  log psi: 1e-5 relative (FP32) and 2e-2 (BF16); FP16 meets 3e-3.
  E_loc: 1e-4 relative per sample (FP32, the bar).  16-bit activations do not
  meet 1e-4 per sample: FP16's measured envelope is stated instead, and the
  energy per site -- the mean over the samples -- is checked against the
  1e-4 bar.  BF16 meets neither 1e-4 bar; its envelope is stated.
"""
import numpy as np
import pytest
import torch

from oracle import convnet_torch as oct_
from oracle import observables as oob
from oracle import operators as oop
from tests.conftest import gpu_banner
from tests.test_gpu_convnet import _oracle_params

pytestmark = pytest.mark.gpu

LOGPSI_RTOL = {'fp32': 1e-5, 'fp16': 3e-3, 'bf16': 2e-2}


def _net(shape, depth, C, prec, seed=0):
    import flowket_amd as fk
    return fk.ConvNetAutoregressive2D(shape, depth=depth, num_of_channels=C, activation='relu',
                                      weights_normalization=False, precision=prec, seed=seed)


def _fp64(m):
    return oct_.torch_params(_oracle_params(m), torch.float64)


def _samples(m, B, seed):
    import flowket_amd as fk
    x = next(fk.FastAutoregressiveSampler(m, B, seed=seed))
    return x.reshape(B, -1).to(torch.int8)


def _rel(got, ref):
    scale = np.abs(ref.real).max() + 1.0
    return max(np.abs(got.real - ref.real).max(), np.abs(got.imag - ref.imag).max()) / scale


@pytest.mark.parametrize('prec', ['fp32', 'fp16', 'bf16'])
def test_heisenberg_10x10_depth20_logpsi(prec):
    gpu_banner(f'full depth 10x10 d20 C32 log psi {prec}')
    m = _net((10, 10), 20, 32, prec, seed=3)
    x = _samples(m, 32, seed=5)
    xs = x.cpu().numpy().reshape(32, 10, 10)
    ref = oct_.log_psi(_fp64(m), xs, 20, 'relu')
    got = m.log_psi(x).cpu().numpy().astype(np.complex128)
    assert _rel(got, ref) < LOGPSI_RTOL[prec], _rel(got, ref)
    # conditionals (2 Re of the normalised log conditionals, absolute)
    tp = _fp64(m)
    z = oct_.unnormalized_logits(tp, xs, 20, 'relu')
    re = z[:, 0:2] - 0.5 * torch.logsumexp(2 * z[:, 0:2], dim=1, keepdim=True)
    rcond = (2 * re).permute(0, 2, 3, 1).numpy()
    cond = m.conditional_log_probs(x).cpu().numpy()
    assert np.abs(cond - rcond).max() < {'fp32': 1e-4, 'fp16': 1e-2, 'bf16': 6e-2}[prec]


def _oracle_eloc(m, xs, op_fn, depth):
    conn, mel, use = op_fn(xs.astype(np.float64))
    tp = _fp64(m)

    def wf(c):
        return oct_.log_psi(tp, c, depth, 'relu').astype(np.complex64)

    return oob.local_values_unbalanced(wf, conn, mel, use)


@pytest.mark.parametrize('prec', ['fp32', 'fp16', 'bf16'])
def test_heisenberg_10x10_depth20_local_energy(prec):
    """Ragged local energies of 12 sampled configurations (about 1,200
    network evaluations) at depth 20 / 32 channels."""
    import flowket_amd as fk
    gpu_banner(f'full depth 10x10 d20 C32 E_loc {prec}')
    m = _net((10, 10), 20, 32, prec, seed=3)
    B = 12
    x = _samples(m, B, seed=7)
    xs = x.cpu().numpy().reshape(B, 10, 10)
    ref = _oracle_eloc(m, xs, lambda s: oop.heisenberg_find_conn(s, pbc=True), 20)
    op = fk.Heisenberg(hilbert_state_shape=(10, 10), pbc=True)
    got = fk.Observable(op).local_values(m, x).cpu().numpy()
    per_sample = np.abs(got - ref) / np.abs(ref)
    print(f'[full-depth] {prec} E_loc max rel err {per_sample.max():.3e} rms {np.sqrt((per_sample ** 2).mean()):.3e}',
          flush=True)
    # FP32: the E_loc bar; FP16 / BF16: the measured envelope (neither meets
    # 1e-4 per sample: 16-bit activations; FP16's energy per site is checked
    # against the bar on a full batch below)
    assert per_sample.max() < {'fp32': 1e-4, 'fp16': 1e-2, 'bf16': 5e-2}[prec], per_sample.max()


def test_heisenberg_10x10_depth20_energy_per_site_fp16():
    """The energy per site of 2048 sampled configurations: FP16 against the
    FP32 path (itself within 1e-4 per sample of the FP64 oracle above) within
    the 1e-4 bar; BF16's difference is recorded (it does not meet it)."""
    import flowket_amd as fk
    gpu_banner('full depth 10x10 d20 C32 energy per site fp16 vs fp32, 2048 samples')
    m = _net((10, 10), 20, 32, 'fp32', seed=3)
    x = _samples(m, 2048, seed=23)
    op = fk.Heisenberg(hilbert_state_shape=(10, 10), pbc=True)
    obs = fk.Observable(op)
    e = {}
    for prec in ('fp32', 'fp16', 'bf16'):
        m.precision = prec
        e[prec] = obs.local_values(m, x).cpu().numpy().real.mean() / 100
    print(f'[full-depth] energy/site fp32 {e["fp32"]:.7f} fp16 err {abs(e["fp16"] - e["fp32"]):.2e} '
          f'bf16 err {abs(e["bf16"] - e["fp32"]):.2e}', flush=True)
    assert abs(e['fp16'] - e['fp32']) < 1e-4


@pytest.mark.parametrize('prec', ['fp16', 'bf16'])
def test_j1j2_10x10_depth20_c48_graph(prec):
    """The graph-captured ragged step (find_conn, compaction, forward with a
    device-side count, E_loc, moments) on the 48-channel J1-J2 network, on 4
    sampled configurations (about 800 network evaluations), replayed twice."""
    import flowket_amd as fk
    gpu_banner(f'full depth J1J2 10x10 d20 C48 graph {prec}')
    m = _net((10, 10), 20, 48, prec, seed=11)
    B = 4
    op = fk.J1J2(hilbert_state_shape=(10, 10), j2=0.5, pbc=True)
    g = fk.LocalEnergyGraph(m, op, B)
    x = _samples(m, B, seed=13)
    xs = x.cpu().numpy().reshape(B, 10, 10)
    ref = _oracle_eloc(m, xs, lambda s: oop.j1j2_find_conn(s, j2=0.5, pbc=True), 20)
    for _ in range(2):
        eloc, mom = g(x)
        got = eloc.cpu().numpy()
        per_sample = np.abs(got - ref) / np.abs(ref)
        assert per_sample.max() < {'fp16': 1e-2, 'bf16': 5e-2}[prec], per_sample.max()
        assert mom[0].item() == pytest.approx(got.real.mean(), rel=1e-12)


@pytest.mark.parametrize('prec', ['fp16', 'bf16'])
def test_heisenberg_20x20_depth24_c64(prec, rng):
    """log psi of 6 sampled 20x20 configurations and of 24 of their connected
    configurations through the ragged (pairs) forward, depth 24 / 64
    channels (horizontal stack in global scratch)."""
    import flowket_amd as fk
    gpu_banner(f'full depth 20x20 d24 C64 {prec}')
    m = _net((20, 20), 24, 64, prec, seed=17)
    B = 6
    x = _samples(m, B, seed=19)
    xs = x.cpu().numpy().reshape(B, 20, 20)
    tp = _fp64(m)
    ref = oct_.log_psi(tp, xs, 24, 'relu')
    got = m.log_psi(x).cpu().numpy().astype(np.complex128)
    assert _rel(got, ref) < LOGPSI_RTOL[prec], _rel(got, ref)
    op = fk.Heisenberg(hilbert_state_shape=(20, 20), pbc=True)
    r = op.compact(x)
    total = int(r['offsets'][-1])
    pick = np.sort(rng.choice(total, size=24, replace=False))
    pairs = r['pairs'][torch.as_tensor(pick).cuda()].contiguous()
    lp = m.log_psi_connected(x, pairs, 24, op).cpu().numpy()
    conn, _, _ = oop.heisenberg_find_conn(xs.astype(np.float64), pbc=True)
    pr = pairs.cpu().numpy()
    cfg = conn[pr[:, 1], pr[:, 0]]
    rc = oct_.log_psi(tp, cfg, 24, 'relu')
    assert _rel(lp[:, 0] + 1j * lp[:, 1], rc) < LOGPSI_RTOL[prec]
