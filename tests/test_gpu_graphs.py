"""The HIP-graph replay of the ragged local-energy step against the eager
launches (same kernels, same order: identical results), over several replays
with eager work and synchronisations between them, after parameter updates,
and against the oracle."""
import numpy as np
import pytest
import torch

from oracle import observables as oob
from oracle import operators as oop
from tests.conftest import gpu_banner

pytestmark = pytest.mark.gpu


def _check_replays(machine, op, B, rng, update):
    import flowket_amd as fk
    g = fk.LocalEnergyGraph(machine, op, B)
    obs = fk.Observable(op)
    n = op.n_sites
    for rep in range(4):
        s = torch.as_tensor(rng.choice([-1, 1], size=(B, n)).astype(np.int8)).cuda()
        el, mom = g(s)
        el, mom = el.clone(), mom.clone()
        # eager work between replays
        ref = obs.local_values(machine, s)
        mean, var = fk.energy_moments(ref)
        torch.cuda.synchronize()
        assert torch.equal(el, ref), rep
        assert mom[0].item() == pytest.approx(mean.real, rel=1e-12, abs=1e-12)
        assert mom[2].item() == pytest.approx(var, rel=1e-10, abs=1e-12)
        if update and rep == 1:
            gen = torch.Generator().manual_seed(rep)
            machine.set_params(machine.params.cpu() * (1 + 0.05 * torch.randn(machine.n_params, generator=gen)))
    return g


def test_graph_convnet_bf16_heisenberg(rng):
    from tests.test_gpu_convnet import _machine
    import flowket_amd as fk
    gpu_banner('local-energy graph: convnet bf16 8x8 heisenberg')
    m = _machine((8, 8), 3, 32, 'relu', 'bf16')
    op = fk.Heisenberg(hilbert_state_shape=(8, 8), pbc=True)
    _check_replays(m, op, 96, rng, update=True)


def test_graph_convnet_fp32_j1j2_vs_oracle(rng):
    from tests.test_gpu_convnet import _machine, _oracle_params
    from oracle import convnet as ocn
    import flowket_amd as fk
    gpu_banner('local-energy graph: convnet fp32 4x4 J1J2 vs oracle')
    m = _machine((4, 4), 2, 16, 'elu', 'fp32')
    op = fk.J1J2(hilbert_state_shape=(4, 4), pbc=True, j2=0.5)
    g = _check_replays(m, op, 40, rng, update=True)
    s = rng.choice([-1, 1], size=(40, 4, 4)).astype(np.int8)
    el, _ = g(torch.as_tensor(s.reshape(40, 16)).cuda())
    conn, mel, use = oop.j1j2_find_conn(s.astype(np.float64), pbc=True, j2=0.5)
    ref = oob.local_values_unbalanced(lambda x: ocn.log_psi(_oracle_params(m), x, 2, activation='elu'), conn, mel,
                                      use)
    assert np.abs(el.cpu().numpy() - ref).max() / np.abs(ref).max() < 1e-4


def test_graph_rbm_ising(rng):
    import flowket_amd as fk
    gpu_banner('local-energy graph: rbm 6x6 ising')
    m = fk.RBM((6, 6), alpha=2, stddev=0.1, seed=3)
    op = fk.Ising(hilbert_state_shape=(6, 6), pbc=True, h=2.0)
    _check_replays(m, op, 64, rng, update=True)
