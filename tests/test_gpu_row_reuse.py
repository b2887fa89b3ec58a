"""Row reuse in the connected-configuration forward (fk_bucket_connections +
fk_convar_forward_rows): log psi of every used connection equals the plain
ragged forward's (same arithmetic per pixel; only the order of the final
FP32 site sum differs), and local energies match with and without it, for the
three operators, every precision and both kernel families (lattices in LDS,
horizontal stack in global scratch)."""
import numpy as np
import pytest
import torch

from tests.conftest import gpu_banner

pytestmark = pytest.mark.gpu

CASES = [((10, 10), 4, 32, 'fp16', 'heisenberg'), ((10, 10), 3, 32, 'bf16', 'j1j2'),
         ((10, 10), 3, 32, 'fp32', 'ising'), ((6, 5), 3, 16, 'fp32', 'heisenberg'),
         ((7, 9), 3, 48, 'bf16', 'heisenberg'), ((20, 20), 2, 64, 'fp16', 'heisenberg'),
         ((12, 12), 3, 32, 'fp16', 'ising'), ((3, 4), 2, 32, 'fp32', 'j1j2')]


def _op(name, shape):
    import flowket_amd as fk
    if name == 'heisenberg':
        return fk.Heisenberg(hilbert_state_shape=shape, pbc=True)
    if name == 'ising':
        return fk.Ising(h=2.0, j=0.3, hilbert_state_shape=shape, pbc=False)
    return fk.J1J2(hilbert_state_shape=shape, j2=0.5, pbc=True)


@pytest.mark.parametrize('shape,depth,C,prec,opname', CASES)
def test_row_reuse_equals_plain_forward(rng, shape, depth, C, prec, opname):
    import flowket_amd as fk
    gpu_banner(f'row reuse {shape} depth={depth} C={C} {prec} {opname}')
    m = fk.ConvNetAutoregressive2D(shape, depth=depth, num_of_channels=C, activation='elu',
                                   weights_normalization=False, precision=prec, seed=4)
    op = _op(opname, shape)
    B = 37
    x = torch.as_tensor(rng.choice([-1, 1], size=(B, shape[0] * shape[1])).astype(np.int8)).cuda()
    r = op.compact(x)
    total = int(r['offsets'][-1])
    plain = m.log_psi_connected(x, r['pairs'], total, op)[:total].double().cpu().numpy()
    rows = m.log_psi_connected_rows(x, r, op)[:total].double().cpu().numpy()
    scale = np.abs(plain[:, 0]).max() + 1.0
    # the activations are bit-identical; the site sums differ in FP32 order
    assert np.abs(rows - plain).max() / scale < 2e-6, np.abs(rows - plain).max() / scale
    # every bucket was evaluated (no configuration left unwritten)
    assert np.isfinite(rows).all()
    obs = fk.Observable(op)
    e_rows = obs.local_values(m, x).cpu().numpy()
    m.row_reuse = False
    e_plain = obs.local_values(m, x).cpu().numpy()
    assert np.abs(e_rows - e_plain).max() <= 1e-5 * np.abs(e_plain).max()


def test_row_reuse_graph_replays(rng):
    """The graph-captured local-energy step with row reuse: replays equal the
    eager evaluation, for two different batches in a row."""
    import flowket_amd as fk
    gpu_banner('row reuse graph replays 10x10')
    m = fk.ConvNetAutoregressive2D((10, 10), depth=3, num_of_channels=32, activation='relu',
                                   weights_normalization=False, precision='fp16', seed=9)
    op = fk.Heisenberg(hilbert_state_shape=(10, 10), pbc=True)
    g = fk.LocalEnergyGraph(m, op, 64)
    obs = fk.Observable(op)
    for seed in (1, 2):
        x = torch.as_tensor(np.random.default_rng(seed).choice([-1, 1], size=(64, 100)).astype(np.int8)).cuda()
        eager = obs.local_values(m, x).clone()
        for _ in range(3):
            e, _ = g(x)
            assert torch.equal(e, eager)
