"""Benchmark of the VMC hot path on one MI355X.

Workload (the headline configuration): 2-D antiferromagnetic Heisenberg on a
10x10 periodic lattice (Marshall-rotated, J = 1), ConvNetAutoregressive2D with
depth 20, 32 channels, 3x3 kernels, ReLU (the reference default), no weight
normalisation, random Glorot weights from a seed, batch 8192 configurations.

One step = one pass of the hot path over the batch:
    find_conn -> ragged index map -> log psi of every used connected
    configuration (fused FP16 MFMA forward, FP32 accumulation) -> E_loc
    reduction -> FP64 energy mean/variance.
The headline ``value`` is local-energy evaluations per second of that step
(16-bit MFMA network, ragged batched evaluation; FP16 because it meets the
1e-4 energy-per-site bar at the BF16 rate, BF16 is reported beside it).
Reported beside it: the same local energies through the plain FP32
per-sample-padded path (every one of the 201 connection slots of every sample
evaluated in FP32), the samples/s of the samplers, accuracy against the FP64
reference, and the CPU baseline (an FP32 torch-CPU forward of the same
network) on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

H = W = 10
DEPTH = 20
CHANNELS = 32
BATCH = 8192
ACT = 'relu'
PEAK_BF16 = 2.5e15      # MI355X dense BF16 MFMA FLOP/s
PEAK_FP32 = 157.3e12    # MI355X FP32 FLOP/s
PEAK_HBM = 8.0e12       # MI355X HBM3E bytes/s
# HBM bytes per connected configuration moved by the forward launches of the
# step, measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE over
# `bench.py --profile-forward 1024 --schedule 0` (103,432 configurations),
# each scaled by the reading of a 256 MiB device copy (`bench.py
# --profile-copy`: FETCH x 2.0, WRITE x 1.0):
#   whole-lattice forward (`--plain-forward`):
#     (13,941.4 KiB x 2 + 5,900.9 KiB) / 103,432 = 334 B
#   row reuse from row 5 (`--row-first 5`; six launches, the sample pass
#   writing 486 kB of cache per sample and the restricted passes reading
#   their halo rows): (682,823 KiB x 2 + 511,385 KiB) / 103,432 = 18,584 B
FWD_HBM_BYTES_PER_CONFIG = {0: 334.5, 5: 18584.0}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--batch', type=int, default=BATCH)
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU oracle baseline')
    ap.add_argument('--no-fp32', action='store_true', help='skip the FP32 plain-path comparison')
    ap.add_argument('--no-sr', action='store_true', help='skip the SR step (per-sample gradients + CG) leg')
    ap.add_argument('--no-mh', action='store_true', help='skip the Metropolis-Hastings RBM leg')
    ap.add_argument('--no-large', action='store_true', help='skip the 20x20 depth-24 64-channel leg')
    ap.add_argument('--profile-forward', type=int, default=0, metavar='N',
                    help='only run two fused-forward launches over the connected configurations of N random '
                         'samples (a short program for rocprofv3 counter passes) and exit')
    ap.add_argument('--row-first', type=int, default=None, metavar='R',
                    help='with --profile-forward: the first lattice row whose connections take restricted '
                         'row-reuse passes (default: the machine default)')
    ap.add_argument('--plain-forward', action='store_true',
                    help='evaluate every connected configuration with the whole-lattice forward (no row reuse)')
    ap.add_argument('--profile-grad', type=int, default=0, metavar='N',
                    help='only run the per-sample gradient kernel twice, then one batched VJP and one JVP, over '
                         'N fast samples of the headline network (a short program for rocprofv3) and exit')
    ap.add_argument('--profile-sampler', type=int, default=0, metavar='N',
                    help='only draw N fast samples of the headline network twice (FP16; a short program for '
                         'rocprofv3) and exit')
    ap.add_argument('--sr-large', action='store_true',
                    help='only run one SR step of the largest configuration (20x20, depth 24, 64 channels, '
                         '16384 fast samples, implicit S products, 50 CG iterations; minutes) and exit')
    ap.add_argument('--profile-copy', action='store_true',
                    help='only run two 256 MiB device-to-device copies (the calibration kernel for the '
                         'FETCH_SIZE / WRITE_SIZE counters) and exit')
    ap.add_argument('--precision', default='fp16', choices=['fp16', 'bf16', 'fp32'])
    ap.add_argument('--schedule', type=int, default=None,
                    help='force a forward launch schedule (index into the compiled table) instead of autotuning')
    return ap.parse_args()


def profile_forward(n, precision, schedule, plain=False, row_first=None):
    import torch
    import flowket_amd as fk
    from flowket_amd import _lib, autotune
    m = fk.ConvNetAutoregressive2D((H, W), depth=DEPTH, num_of_channels=CHANNELS, activation=ACT,
                                   weights_normalization=False, precision=precision, seed=1234)
    if schedule is not None:
        from flowket_amd.machines import _PREC
        autotune.set_schedule(m, _PREC[precision], schedule)
    if row_first is not None:
        m.set_row_reuse_policy(row_first)
    op = fk.Heisenberg(hilbert_state_shape=(H, W), pbc=True)
    gen = torch.Generator().manual_seed(0)
    spins = (torch.randint(0, 2, (n, H * W), generator=gen) * 2 - 1).to(torch.int8).cuda()
    r = op.compact(spins)
    total = int(r['offsets'][-1])
    rows = not plain and m.uses_row_reuse(op)
    for _ in range(2):
        if rows:
            m.log_psi_connected_rows(spins, r, op)
        else:
            m.log_psi_connected(spins, r['pairs'], total, op)
    torch.cuda.synchronize()
    print(json.dumps({'profile_forward': {'samples': n, 'configurations': total, 'precision': precision,
                                          'schedule': schedule, 'row_reuse': rows,
                                          'row_first': m.row_reuse_policy()}}))
    return 0


def profile_grad(n):
    import torch
    import flowket_amd as fk
    m = fk.ConvNetAutoregressive2D((H, W), depth=DEPTH, num_of_channels=CHANNELS, activation=ACT,
                                   weights_normalization=False, precision='fp16', seed=1234)
    spins = next(fk.FastAutoregressiveSampler(m, n, seed=3)).reshape(n, H * W).contiguous()
    for _ in range(2):
        o_re, o_im, _ = m.per_sample_gradients(spins)
        del o_re, o_im
    cot = torch.randn(n, 2, device=spins.device)
    m.vjp(spins, cot)
    m.jvp(spins, torch.randn(m.n_trainable, device=spins.device))
    torch.cuda.synchronize()
    print(json.dumps({'profile_grad': {'samples': n, 'params': m.n_params}}))
    return 0


def profile_sampler(n):
    import torch
    import flowket_amd as fk
    m = fk.ConvNetAutoregressive2D((H, W), depth=DEPTH, num_of_channels=CHANNELS, activation=ACT,
                                   weights_normalization=False, precision='fp16', seed=1234)
    fs = fk.FastAutoregressiveSampler(m, n, seed=7)
    for _ in range(2):
        next(fs)
    torch.cuda.synchronize()
    print(json.dumps({'profile_sampler': {'samples': n}}))
    return 0


def profile_copy():
    import torch
    src = torch.ones(256 << 20, dtype=torch.uint8, device='cuda')
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({'profile_copy': {'bytes': src.numel()}}))
    return 0


def flops_per_forward(C=CHANNELS, depth=DEPTH, h=H, w=W):
    nb = 2 * depth - 2
    return 2 * h * w * ((10 * C * C + 12 * C) + (nb - 1) * 22 * C * C + 4 * C)


def bench_j1j2(batch=4096, plain_samples=1024, reps=5):
    """The J1-J2 configuration: 10x10 periodic, J2/J1 = 0.5, depth 20, 48
    channels (run on the 64-channel kernel; the padding channels stay zero),
    bf16, 4096 fast samples.  The ragged set of used connections (1 to 401
    per sample) goes through the hipGraph-captured local-energy chain, with
    the forward schedule autotuned on this workload first; the comparison is
    the plain per-sample-padded FP32 path (all 401 slots of every sample,
    timed on 1024 of the samples).  FLOPs count the 48 real channels."""
    import torch
    import flowket_amd as fk
    from flowket_amd import _lib, autotune
    C, depth = 48, 20
    m16 = fk.ConvNetAutoregressive2D((H, W), depth=depth, num_of_channels=C, activation=ACT,
                                     weights_normalization=False, precision='bf16', seed=21)
    m32 = fk.ConvNetAutoregressive2D((H, W), depth=depth, num_of_channels=C, activation=ACT,
                                     weights_normalization=False, precision='fp32', params=m16.params)
    op = fk.J1J2(hilbert_state_shape=(H, W), j2=0.5, pbc=True)
    obs = fk.Observable(op)
    fs = fk.FastAutoregressiveSampler(m16, batch, seed=22)
    next(fs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    spins = next(fs).reshape(batch, H * W).contiguous()
    torch.cuda.synchronize()
    sample_s = time.perf_counter() - t0
    tune = autotune.tune_forward(m16, lambda: obs.local_values(m16, spins), reps=2, persist=False)
    g = fk.LocalEnergyGraph(m16, op, batch)
    el16, _ = g(spins)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        el16, mom = g(spins)
    torch.cuda.synchronize()
    step_s = (time.perf_counter() - t0) / reps
    n_conn = int(g.num_connections.item())
    e16 = el16.real.mean().item() / (H * W)
    # plain FP32 path on a subset
    sub = spins[:plain_samples].contiguous()
    nc = op.num_connections
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = op.find_conn_device(sub, want_conn=True)
    lp, _ = m32.forward(r['conn'].reshape(nc * plain_samples, H * W))
    el32 = torch.empty((plain_samples, 2), dtype=torch.float64, device=spins.device)
    _lib.call('fk_local_energy_dense', _lib.ptr(lp), _lib.ptr(r['mel']), None, nc, plain_samples, _lib.ptr(el32),
              _lib.stream_ptr())
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t0
    plain_rate = plain_samples / plain_s
    sub16 = el16[:plain_samples].real.cpu().numpy()
    sub32 = el32[:, 0].cpu().numpy()
    return {'workload': '10x10 periodic J1-J2 (J2/J1 0.5), depth 20, 48 channels (64-channel kernel), bf16, '
                        '4096 fast samples, hipGraph-captured ragged local energies',
            'fast_samples_per_s': round(batch / sample_s, 1),
            'connected_configurations': n_conn, 'mean_connections_per_sample': round(n_conn / batch, 1),
            'E_loc_evals_per_s': round(batch / step_s, 1), 'ms_per_step': round(1e3 * step_s, 2),
            'forward_tflops_48ch': round(n_conn * flops_per_forward(C, depth) / step_s / 1e12, 1),
            'forward_schedule': tune['schedule'], 'schedule_ms': [round(t, 2) for t in tune['ms']],
            'fp32_plain_E_loc_evals_per_s': round(plain_rate, 1),
            'speedup_vs_fp32_plain': round((batch / step_s) / plain_rate, 2),
            'energy_per_site_bf16': e16,
            'eloc_rms_rel_err_bf16_vs_fp32_subset': float(np.sqrt(np.mean((sub16 - sub32) ** 2))
                                                         / np.sqrt(np.mean(sub32 ** 2)))}


def bench_sr(machine, op, spins, shift=1e-3, tol=1e-6, max_iter=50):
    """One SR step of the headline configuration: local energies (bf16 ragged
    path), per-sample O_k of every sample (FP32 gradient kernel, O stored
    B x P complex as two FP32 planes), force and CG on the implicit
    S = Re[(O-Obar)^H (O-Obar)]/B + shift (FP32 products, FP64 reductions)."""
    import torch
    import flowket_amd as fk
    from flowket_amd.optimizers import _OFactor, _complex_eloc_pairs, conjugate_gradient
    B = spins.shape[0]
    eloc = fk.Observable(op).local_values(machine, spins)
    E = eloc.mean()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
    ev[0].record()
    of = _OFactor(machine, spins)
    ev[1].record()
    u = _complex_eloc_pairs(eloc, E)
    force = of.oh(u, 1.0 / of.B)
    x, it, res = conjugate_gradient(lambda v, out: of.s_matvec(v, shift, out), force, tol=tol, max_iter=max_iter)
    ev[2].record()
    # one S.v product on its own (one pass over O)
    z = torch.empty_like(force)
    ev[3].record()
    of.s_matvec(force, shift, z)
    ev[4].record()
    torch.cuda.synchronize()
    P = of.P
    grad_ms = ev[0].elapsed_time(ev[1])
    cg_ms = ev[1].elapsed_time(ev[2])
    mv_ms = ev[3].elapsed_time(ev[4])
    o_bytes = 2 * 4 * B * P

    def timed(fn):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        r = fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1), r

    two_ms, _ = timed(lambda: of.s_matvec_two_pass(force, shift))
    rows_ms = {}
    for r in (16, 32):
        of.block_rows = r
        rows_ms[r], z1 = timed(lambda: of.s_matvec(force, shift))
        rows_ms[r] = round(rows_ms[r], 3)
    of.block_rows = 0
    fnorm = float(torch.linalg.vector_norm(force.double()).item())
    out = {'batch': B, 'params': P, 'grad_ms': round(grad_ms, 2),
           'grad_tflops': round(B * 2 * flops_per_forward() / (grad_ms * 1e-3) / 1e12, 2),
           'grad_o_write_gbs': round(o_bytes / (grad_ms * 1e-3) / 1e9, 1),
           'cg_ms': round(cg_ms, 2), 'cg_iterations': it, 'cg_residual': res, 'cg_force_norm': fnorm,
           'cg_relative_residual': res / fnorm if fnorm > 0 else 0.0, 'cg_tol': tol,
           'cg_converged': bool(res <= tol * fnorm),
           'matvec_ms': round(mv_ms, 3), 'matvec_gbs': round(2 * o_bytes / (mv_ms * 1e-3) / 1e9, 1),
           'matvec_hbm_frac': round(2 * o_bytes / (mv_ms * 1e-3) / PEAK_HBM, 4),
           'matvec_note': 'the default S.v reads O twice (matvec_gbs = 2 x bytes of O / time); '
                          'matvec_one_pass_ms: one pass over O from HBM in row blocks of the given rows, '
                          're-read from the last-level cache',
           'matvec_one_pass_ms': rows_ms,
           'matvec_two_calls_ms': round(two_ms, 3),
           'one_pass_vs_two_pass_rel_diff': float(
               (torch.linalg.vector_norm((z1 - z).double()) / torch.linalg.vector_norm(z.double())).item()),
           'step_ms': round(grad_ms + cg_ms, 2), 'o_gib': round(o_bytes / 2 ** 30, 2)}
    # the same S.v without O: forward-mode JVP, then one batched VJP
    if machine.supports_implicit_products():
        from flowket_amd.optimizers import _ImplicitS
        im = _ImplicitS(machine, spins)
        zi = im.s_matvec(force, shift)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        zi = im.s_matvec(force, shift)
        e1.record()
        torch.cuda.synchronize()
        out['implicit_matvec_ms'] = round(e0.elapsed_time(e1), 2)
        out['implicit_vs_materialised_rel_diff'] = float(
            (torch.linalg.vector_norm((zi - z).double()) / torch.linalg.vector_norm(z.double())).item())
    del of
    torch.cuda.empty_cache()
    return out


def bench_sr_training(steps=10, batch=BATCH, lr=0.05, shift=1e-3):
    """A short SR optimisation of the headline network and Hamiltonian
    (fast samples, FP16 local energies, FP32 O_k, CG on the materialised
    S = Re[(O-Obar)^H (O-Obar)]/B + shift with tol 1e-6 and at most 50
    iterations): the sampled energy per step and the time of a step."""
    import torch
    import flowket_amd as fk
    m = fk.ConvNetAutoregressive2D((H, W), depth=DEPTH, num_of_channels=CHANNELS, activation=ACT,
                                   weights_normalization=False, precision='fp16', seed=1234)
    op = fk.Heisenberg(hilbert_state_shape=(H, W), pbc=True)
    vmc = fk.VariationalMonteCarlo(m, op, fk.FastAutoregressiveSampler(m, batch, seed=41))
    sr = fk.ComplexValuesStochasticReconfiguration(m, lr=lr, diag_shift=shift, conjugate_gradient_tol=1e-6,
                                                   iterative_solver_max_iterations=50)
    energies, cg_iters = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        x, _ = vmc.next_batch()
        met = sr.step(x, vmc.current_local_energy, vmc.current_energy)
        energies.append(float(torch.as_tensor(vmc.current_energy).real) / (H * W))
        cg_iters.append(int(met['stochastic_reconfiguration/conjugate_gradient_iterations']))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    del sr, vmc
    torch.cuda.empty_cache()
    return {'workload': f'10x10 periodic Heisenberg, depth {DEPTH}, {CHANNELS} channels, batch {batch} fast samples, '
                        f'SR lr {lr} shift {shift}, CG tol 1e-6 max 50 (materialised O)',
            'steps': steps, 'ms_per_step': round(1e3 * dt / steps, 1),
            'energy_per_site': [round(e, 5) for e in energies], 'cg_iterations': cg_iters,
            'energy_decreased': bool(np.mean(energies[-3:]) < np.mean(energies[:3]))}


def bench_mh(seconds_burn=None):
    """Metropolis-Hastings on the 8x8 periodic Heisenberg RBM (alpha 4,
    stddev 0.01, 1024 chains, exchange moves): ESS/s of the plain chain and
    of the chain with the learned MADE proposal (64 -> 256 -> 256 -> 64,
    refit every 20 batches).  One recorded sample per 64 proposals per chain
    (one sweep of the lattice); 500 sweeps of burn-in before timing.  The
    timed window is one whole refit period (20 batches), so the cost of the
    refit that falls in it is amortized into ESS/s."""
    import torch
    import flowket_amd as fk
    m = fk.RBM((8, 8), alpha=4, stddev=0.01, seed=5)
    op = fk.Heisenberg(hilbert_state_shape=(8, 8), pbc=True)
    obs = fk.Observable(op)
    chains, spc = 1024, 16
    B = chains * spc
    res = {}
    for name, cls, kw in (('exchange', fk.MetropolisHastingsExchange, {}),
                          ('exchange+learned', fk.MetropolisHastingsLearnedProposal,
                           {'base_kind': 1, 'hidden': 256, 'refit_every': 20, 'fit_steps': 100})):
        s = cls(m, B, num_of_chains=chains, unused_sampels=63, discard_ratio=spc, seed=3, **kw)
        # burn-in: 500 sweeps of 64 proposals per chain
        for _ in range(500 // (spc + 1) + 1):
            next(s)
        torch.cuda.synchronize()
        ess, es, secs = [], [], 0.0
        acc = []
        window = 20
        for _ in range(window):
            t0 = time.perf_counter()
            x = next(s)
            torch.cuda.synchronize()
            secs += time.perf_counter() - t0
            el = obs.local_values(m, x).real
            r_hat, var, cs, e = s.calc_r_hat_value(el)
            ess.append(float(e))
            es.append(float(el.mean().item()))
            acc.append(float(getattr(s, 'proposal_acceptance', s.acceptance_ratio)))
        res[name] = {'ess_per_s': round(sum(ess) / secs, 1), 'ess_per_batch': round(float(np.mean(ess)), 1),
                     'tau_int': round(B / float(np.mean(ess)), 2), 'batch_ms': round(1e3 * secs / window, 2),
                     'energy_per_site': round(float(np.mean(es)) / 64, 5),
                     'energy_spread': round(float(np.std(es)) / 64, 5), 'acceptance': round(float(np.mean(acc)), 3)}
        if hasattr(s, 'used_proposal'):
            res[name]['used_proposal'] = bool(s.used_proposal)
            res[name]['proposal_nll'] = s.last_nll
    return res


# FP64 exact-diagonalisation ground energy of the 4x4 periodic Heisenberg
# model in the Pauli convention (pinned by the oracle tests).
E0_HEISENBERG_4X4_PBC = -44.9139328337


def bench_small(steps=100):
    """The small plain-method configuration: 4x4 periodic Heisenberg,
    ConvNetAutoregressive2D depth 5, 16 channels, ELU, FP32, Glorot weights
    from the seed; 1024 exact autoregressive samples per step, 100 Adam steps
    at lr 1e-3.  The exact energy of psi over all 65,536 configurations (the
    exact enumeration path) before and after, against the ED ground energy."""
    import torch
    import flowket_amd as fk
    m = fk.ConvNetAutoregressive2D((4, 4), depth=5, num_of_channels=16, activation='elu',
                                   weights_normalization=False, precision='fp32', seed=11)
    op = fk.Heisenberg(hilbert_state_shape=(4, 4), pbc=True)
    ev = fk.ExactVariational(m, op, batch_size=65536)
    ev.machine_updated()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        ev.machine_updated()
    torch.cuda.synchronize()
    exact_ms = 1e3 * (time.perf_counter() - t0) / reps
    e_start = ev.energy_observable.current_energy.real
    vmc = fk.VariationalMonteCarlo(m, op, fk.AutoregressiveSampler(m, 1024, seed=12))
    opt = fk.Adam(m, lr=1e-3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sampled = []
    for _ in range(steps):
        configs, _ = vmc.next_batch()
        sampled.append(vmc.current_energy.real)
        opt.step(configs, vmc.current_local_energy, vmc.current_energy)
    torch.cuda.synchronize()
    step_ms = 1e3 * (time.perf_counter() - t0) / steps
    ev.machine_updated()
    e_end = ev.energy_observable.current_energy.real
    return {'workload': '4x4 periodic Heisenberg, depth 5, 16 channels, elu, fp32, 1024 samples, Adam lr 1e-3',
            'steps': steps, 'ms_per_step': round(step_ms, 3),
            'exact_energy_start': e_start, 'exact_energy_end': e_end,
            'sampled_energy_last10': float(np.mean(sampled[-10:])),
            'ed_ground_energy': E0_HEISENBERG_4X4_PBC,
            'exact_energy_end_per_site_above_ed': (e_end - E0_HEISENBERG_4X4_PBC) / 16,
            'exact_variance_end': ev.energy_observable.current_local_energy_variance,
            'exact_pass_ms': round(exact_ms, 3),
            'exact_states_per_s': round(65536 / (exact_ms * 1e-3), 1)}


def bench_trained_ising(n_samples=2 ** 15, mini_batch=2 ** 12, precision='fp16'):
    """The reference's published evaluation (experiments/README.md:38-47,
    experiments/run_evaluation.py:12-26): its trained 12x12 open-boundary
    transverse-field Ising network (h = 3, depth 10, 32 channels, weight
    normalisation; weights imported from its HDF5 file into
    tests/golden/ising_12x12_h3_weights.npz), samples from the base network's
    fast sampler, local energies of the up/down x dihedral symmetrised network
    (16 members), 2^15 samples as published."""
    import torch
    import flowket_amd as fk
    from flowket_amd.evaluation import evaluate
    w = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'tests', 'golden',
                             'ising_12x12_h3_weights.npz'))
    m = fk.ConvNetAutoregressive2D((12, 12), depth=10, num_of_channels=32, precision=precision,
                                   params=torch.from_numpy(w['params']))
    m.set_wn_g(torch.from_numpy(w['g']))
    op = fk.Ising(hilbert_state_shape=(12, 12), pbc=False, h=3.0)
    sym = fk.make_up_down_invariant(fk.make_2d_obc_invariants(m))
    from flowket_amd import autotune
    tune_spins = next(fk.FastAutoregressiveSampler(m, mini_batch, seed=20))
    obs = fk.Observable(op)
    tune = autotune.tune_forward(m, lambda: obs.local_values(sym, tune_spins), reps=1, persist=False)
    vmc = fk.VariationalMonteCarlo(sym, op, fk.FastAutoregressiveSampler(m, mini_batch, seed=21))
    evaluate(vmc, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    steps = n_samples // mini_batch
    r = evaluate(vmc, steps, true_ground_state_energy=-457.039)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    e, var = r['energy/energy'], r['energy/local_energy_variance']
    published_e, published_var = -457.0420317, 0.000821
    se = float(np.sqrt(var / n_samples + published_var / 2 ** 15))
    return {'workload': '12x12 obc Ising h=3, trained depth 10, 32 channels, %s, 2^15 samples, '
                        '16-member symmetrised E_loc' % precision,
            'seconds': round(dt, 3), 'samples_per_s': round(n_samples / dt, 1),
            'network_evals_per_s': round(n_samples * 145 * 16 / dt, 1),
            'energy': e, 'local_energy_variance': var, 'relative_error_vs_dmrg': r['energy/relative_error'],
            'abs_sigma_z': r['observables/abs_sigma_z'], 'published_abs_sigma_z': 0.1622390747,
            'published_energy': published_e, 'published_variance': published_var,
            'energy_diff_in_standard_errors': round((e - published_e) / se, 2),
            'forward_schedule': tune['schedule'], 'schedule_ms': [round(t, 2) for t in tune['ms']]}


def bench_large(n_samples=16384, n_eloc=256, n_sv=512):
    """The largest configuration's network and lattice: 20x20 periodic
    Heisenberg, depth 24, 64 channels, bf16.  Fast sampling of n_samples
    configurations and the ragged local energies of the first n_eloc (a
    bounded slice of the 16384-sample step), with the forward's MFMA roofline
    fraction (its horizontal stack lives in global scratch: the five 20x20x64
    lattices exceed the LDS)."""
    import torch
    import flowket_amd as fk
    Hl = Wl = 20
    depth, C = 24, 64
    m = fk.ConvNetAutoregressive2D((Hl, Wl), depth=depth, num_of_channels=C, activation=ACT,
                                   weights_normalization=False, precision='bf16', seed=99)
    op = fk.Heisenberg(hilbert_state_shape=(Hl, Wl), pbc=True)
    fs = fk.FastAutoregressiveSampler(m, n_samples, seed=5)
    next(fs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x = next(fs)
    torch.cuda.synchronize()
    sps = n_samples / (time.perf_counter() - t0)
    x = x.reshape(n_samples, Hl * Wl)[:max(n_eloc, n_sv)].contiguous()
    del fs
    torch.cuda.empty_cache()
    spins = x[:n_eloc].contiguous()
    obs = fk.Observable(op)
    from flowket_amd import autotune
    tune = autotune.tune_forward(m, lambda: obs.local_values(m, spins), reps=1, persist=False)
    obs.local_values(m, spins)
    r = op.compact(spins)
    total = int(r['offsets'][-1])
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    el = obs.local_values(m, spins)
    e1.record()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    fwd = 2 * Hl * Wl * ((10 * C * C + 12 * C) + (2 * depth - 3) * 22 * C * C + 4 * C)
    out = {'workload': '20x20 periodic Heisenberg, depth 24, 64 channels, bf16 (horizontal stack in global scratch)',
           'params': m.n_params, 'fast_samples_per_s': round(sps, 1), 'eloc_samples': n_eloc,
           'connected_configurations': total, 'E_loc_evals_per_s': round(n_eloc / dt, 2),
           'connected_logpsi_per_s': round(total / dt, 1),
           'forward_tflops': round(total * fwd / dt / 1e12, 1),
           'forward_frac_of_bf16_peak': round(total * fwd / dt / PEAK_BF16, 4),
           'forward_schedule': tune['schedule'], 'schedule_ms': [round(t, 2) for t in tune['ms']],
           'row_reuse_first_row': tune.get('row_reuse'),
           'energy_per_site': float(el.real.mean().item()) / (Hl * Wl)}
    # SR at this size never forms O (16384 x 4.1M complex would be 538 GB):
    # one implicit S.v (JVP, then a batched VJP) over n_sv samples, and the
    # CG step it implies at 16384 samples and 50 iterations
    from flowket_amd.optimizers import _ImplicitS
    xs = x[:n_sv].contiguous()
    im = _ImplicitS(m, xs)
    v = torch.randn(m.n_trainable, device=xs.device)
    im.s_matvec(v, 1e-3)
    torch.cuda.synchronize()
    e0.record()
    im.s_matvec(v, 1e-3)
    e1.record()
    torch.cuda.synchronize()
    sv_ms = e0.elapsed_time(e1)
    out['implicit_sr'] = {'samples': n_sv, 'matvec_ms': round(sv_ms, 1),
                          'matvec_us_per_sample': round(1e3 * sv_ms / n_sv, 1),
                          'projected_cg_step_s_16384x50': round(50 * sv_ms * 16384 / n_sv / 1e3, 1)}
    del im
    torch.cuda.empty_cache()
    return out


def bench_sr_large(n_samples=16384, chunk=4096, shift=1e-3, tol=1e-6, max_iter=50, lr=0.01):
    """One SR step of the largest configuration: 20x20 periodic Heisenberg,
    depth 24, 64 channels (4.1M parameters), bf16 network, 16384 fast
    samples, local energies (row reuse, in chunks of 4096 samples), then CG
    on S = Re[(O-Obar)^H (O-Obar)]/B + shift without forming O (each S.v one
    forward-mode JVP and one batched VJP over all samples, FP32 MFMA),
    tol 1e-6, at most 50 iterations; the update is applied.  Prints a
    progress line per CG iteration."""
    import torch
    import flowket_amd as fk
    from flowket_amd.optimizers import _ImplicitS, _complex_eloc_pairs, conjugate_gradient
    Hl = Wl = 20
    m = fk.ConvNetAutoregressive2D((Hl, Wl), depth=24, num_of_channels=64, activation=ACT,
                                   weights_normalization=False, precision='bf16', seed=99)
    op = fk.Heisenberg(hilbert_state_shape=(Hl, Wl), pbc=True)
    obs = fk.Observable(op)
    out = {'workload': '20x20 periodic Heisenberg, depth 24, 64 channels, bf16 network, %d fast samples, SR shift %g, '
                       'CG tol %g max %d, implicit S products (FP32 MFMA)' % (n_samples, shift, tol, max_iter),
           'params': m.n_params}

    def tick(msg):
        print(json.dumps({'progress': msg, 't': round(time.perf_counter() - t_start, 1)}), flush=True)

    t_start = time.perf_counter()
    fs = fk.FastAutoregressiveSampler(m, n_samples, seed=5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x = next(fs).reshape(n_samples, Hl * Wl).contiguous()
    torch.cuda.synchronize()
    out['sample_s'] = round(time.perf_counter() - t0, 2)
    tick('sampled')
    t0 = time.perf_counter()
    els = []
    for c0 in range(0, n_samples, chunk):
        els.append(obs.local_values(m, x[c0:c0 + chunk].contiguous()))
        torch.cuda.synchronize()
        tick('local energies %d' % (c0 + chunk))
    el = torch.cat(els)
    out['local_energy_s'] = round(time.perf_counter() - t0, 2)
    E = el.mean()
    out['energy_per_site'] = float(E.real) / (Hl * Wl)
    m._rows_ws = {}
    torch.cuda.empty_cache()
    t0 = time.perf_counter()
    im = _ImplicitS(m, x)
    u = _complex_eloc_pairs(el, E)
    force = im.force(u.to(x.device))
    torch.cuda.synchronize()
    out['force_s'] = round(time.perf_counter() - t0, 2)
    tick('force')
    calls = [0]

    def apply(v, z):
        im.s_matvec(v, shift, z)
        calls[0] += 1
        torch.cuda.synchronize()
        tick('S.v %d' % calls[0])

    t0 = time.perf_counter()
    sol, it, res = conjugate_gradient(apply, force, tol=tol, max_iter=max_iter, check_every=1)
    torch.cuda.synchronize()
    cg_s = time.perf_counter() - t0
    fnorm = float(torch.linalg.vector_norm(force.double()).item())
    m.trainable.add_(sol, alpha=-lr)
    m.mark_params_updated()
    out.update({'cg_s': round(cg_s, 2), 'cg_iterations': it, 'cg_relative_residual': res / fnorm if fnorm else 0.0,
                'matvec_s': round(cg_s / max(calls[0], 1), 2),
                'sr_step_s': round(out['force_s'] + cg_s, 2),
                'total_step_s': round(out['sample_s'] + out['local_energy_s'] + out['force_s'] + cg_s, 2)})
    return out


def _oracle_params(machine):
    flat = machine.effective_params().double().cpu().numpy()
    params, off = [], 0
    for ks, bs in machine.shapes:
        nk, nb = int(np.prod(ks)), int(np.prod(bs))
        params.append((flat[off:off + nk].reshape(ks), flat[off + nk:off + nk + nb]))
        off += nk + nb
    return params


def cpu_eloc(machine, spins_np, dtype):
    """Local energies on the host: the oracle's numpy find_conn and unbalanced
    E_loc rule with the network's log psi from the torch restatement
    (oracle/convnet_torch.py) in ``dtype`` (float32: the arithmetic of the
    reference's TensorFlow CPU path; float64: the FP64 reference)."""
    import torch
    from oracle import convnet_torch as oct_
    from oracle import observables as oob
    from oracle import operators as oop
    tp = oct_.torch_params(_oracle_params(machine), dtype)
    conn, mel, use = oop.heisenberg_find_conn(spins_np.astype(np.float64), pbc=True)
    eloc = oob.local_values(lambda x: oct_.log_psi(tp, x, DEPTH, ACT).astype(np.complex64), conn, mel, use)
    return eloc, int(use.sum())


def cpu_baseline(machine, spins_np, n_samples=64, runs=5):
    """The CPU baseline on a bounded sample: local energies of the first
    n_samples configurations with an FP32 torch-CPU forward of the same
    network (find_conn and the E_loc rule from the oracle), the median of
    ``runs`` runs on all of the host's cores; plus the FP64 oracle's local
    energies of the same samples for the accuracy figures."""
    import torch
    cores = len(os.sched_getaffinity(0))
    threads = torch.get_num_threads()
    torch.set_num_threads(cores)
    s = spins_np[:n_samples].reshape(n_samples, H, W)
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        _, nfwd = cpu_eloc(machine, s, torch.float32)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    el64, _ = cpu_eloc(machine, s, torch.float64)
    torch.set_num_threads(threads)
    return {'value': n_samples / dt, 'unit': 'E_loc evals/s', 'cores': cores, 'kind': 'port',
            'sample': f'{n_samples} configurations of the benchmark batch: numpy find_conn + unbalanced E_loc rule '
                      f'with an FP32 torch-CPU forward of the same network ({nfwd} network evaluations), median '
                      f'of {runs} runs ({dt:.2f} s each)'}, el64


def main():
    args = parse()
    if args.profile_copy:
        return profile_copy()
    if args.profile_forward:
        return profile_forward(args.profile_forward, args.precision, args.schedule, args.plain_forward,
                               args.row_first)
    if args.profile_grad:
        return profile_grad(args.profile_grad)
    if args.profile_sampler:
        return profile_sampler(args.profile_sampler)
    if args.sr_large:
        print(json.dumps({'sr_large': bench_sr_large()}))
        return 0
    if args.gpus != 1:
        print(json.dumps({'error': 'the VMC path here runs on one GPU (samples shard across GPUs by design, '
                                   'but no multi-GPU launch is implemented); --gpus must be 1'}))
        return 2
    import torch
    import flowket_amd as fk
    from flowket_amd import _lib
    _lib.require_gpu()
    dev = torch.device('cuda', 0)
    torch.cuda.set_device(dev)
    B = args.batch

    # the headline network computes in FP16 MFMA (FP32 accumulation): the
    # 16-bit path that meets the 1e-4 energy-per-site bar; the same weights
    # in BF16 and FP32 run beside it
    m16 = fk.ConvNetAutoregressive2D((H, W), depth=DEPTH, num_of_channels=CHANNELS, activation=ACT,
                                     weights_normalization=False, precision='fp16', seed=1234)
    mb = fk.ConvNetAutoregressive2D((H, W), depth=DEPTH, num_of_channels=CHANNELS, activation=ACT,
                                    weights_normalization=False, precision='bf16', params=m16.params)
    m32 = fk.ConvNetAutoregressive2D((H, W), depth=DEPTH, num_of_channels=CHANNELS, activation=ACT,
                                     weights_normalization=False, precision='fp32', params=m16.params)
    op = fk.Heisenberg(hilbert_state_shape=(H, W), pbc=True)
    obs = fk.Observable(op)

    # ---- samplers: fast (each (layer, site) once) and plain (H*W forwards)
    def time_sampler(sampler, reps):
        next(sampler)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            x = next(sampler)
        torch.cuda.synchronize()
        return x, B * reps / (time.perf_counter() - t0)

    fast16 = fk.FastAutoregressiveSampler(m16, B, seed=7)
    spins, sps_fast16 = time_sampler(fast16, 3)
    spins = spins.reshape(B, H * W).contiguous()
    _, sps_fastb = time_sampler(fk.FastAutoregressiveSampler(mb, B, seed=8), 2)
    _, sps_fast32 = time_sampler(fk.FastAutoregressiveSampler(m32, B, seed=8), 1)
    sps = {'fast_autoregressive_fp16': round(sps_fast16, 1), 'fast_autoregressive_bf16': round(sps_fastb, 1),
           'fast_autoregressive_fp32': round(sps_fast32, 1)}
    if not args.no_fp32:
        t0 = time.perf_counter()
        next(fk.AutoregressiveSampler(m16, B, seed=9))
        torch.cuda.synchronize()
        sps['plain_autoregressive_fp16'] = round(B / (time.perf_counter() - t0), 1)
        t0 = time.perf_counter()
        next(fk.AutoregressiveSampler(m32, B, seed=9))
        torch.cuda.synchronize()
        sps['plain_autoregressive_fp32'] = round(B / (time.perf_counter() - t0), 1)
        sps['fast_fp16_vs_plain_fp32'] = round(sps_fast16 / sps['plain_autoregressive_fp32'], 1)

    # ---- the step
    fwd_times = []

    def step(record):
        r = op.compact(spins)
        offsets, pairs = r['offsets'], r['pairs']
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        if m16.uses_row_reuse(op):
            logpsi = m16.log_psi_connected_rows(spins, r, op)
        else:
            logpsi = m16.log_psi_connected(spins, pairs, pairs.shape[0], op, m_dev=offsets[B:])
        e1.record()
        eloc = torch.empty((B, 2), dtype=torch.float64, device=dev)
        _lib.call('fk_local_energy', _lib.ptr(offsets), _lib.ptr(pairs), _lib.ptr(logpsi), _lib.ptr(r['mel']), None,
                  B, _lib.ptr(eloc), _lib.stream_ptr())
        mom = torch.empty(4, dtype=torch.float64, device=dev)
        _lib.call('fk_energy_moments', _lib.ptr(eloc), B, _lib.ptr(mom), _lib.stream_ptr())
        if record:
            fwd_times.append((e0, e1))
        return offsets, mom

    # ---- launch schedule of the fused forward: forced, or autotuned on this step
    from flowket_amd import autotune
    if args.plain_forward:
        m16.set_row_reuse_policy(0)
    if args.schedule is not None:
        autotune.set_schedule(m16, _lib.FK_PREC_FP16, args.schedule)
        tune = {'schedule': args.schedule, 'forced': True}
    else:
        tune = autotune.tune_forward(m16, lambda: step(False), reps=2, persist=False, row_reuse=not args.plain_forward)
        tune = {'schedule': tune['schedule'], 'ms_per_schedule': [round(t, 2) for t in tune['ms']],
                'schedules': tune['schedules'], 'row_reuse_first_row': tune.get('row_reuse'),
                'ms_per_row_reuse_policy': {str(k): round(v, 2) for k, v in tune.get('row_reuse_ms', {}).items()}}

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        offsets, mom = step(True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = B * args.steps / elapsed
    n_conn = int(offsets[-1].item())
    fwd_ms = float(np.mean([a.elapsed_time(b) for a, b in fwd_times]))
    # reference-algorithm work: a full forward per used connection; executed
    # work with row reuse: the samples' full forwards plus, per bucket of
    # connections first changing row R0 >= 2, the rows R0 .. H-1 only
    flop_launch = n_conn * flops_per_forward()
    executed = flop_launch
    if m16.uses_row_reuse(op):
        boff = m16._rows_workspace(_lib.FK_PREC_FP16, B, op.num_connections, H + 1)['boff'].cpu().numpy()
        counts = np.diff(boff)
        rows = np.array([H] + [H if k - 1 < 2 else H - (k - 1) for k in range(1, H + 1)], dtype=np.float64)
        executed = float((counts * rows / H).sum()) * flops_per_forward()
    achieved = flop_launch / (fwd_ms * 1e-3)
    policy = m16.row_reuse_policy() if m16.uses_row_reuse(op) else 0
    mom = mom.cpu().numpy()

    out = {
        'metric': 'samples/s & E_loc evals/s, 10x10 Heisenberg: bf16 batched vs fp32 plain',
        'value': round(value, 2),
        'unit': 'E_loc evals/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(ms_per_step, 3),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'fp16',
        'data': 'synthetic: exact (fast) autoregressive samples of a seeded random-weight network',
        'config': {'workload': '10x10 periodic Heisenberg, ConvNetAutoregressive2D depth 20 C 32 relu, '
                               'batch 8192, fp16 MFMA (fp32 accumulate) ragged local energies', 'batch': B,
                   'connected_evaluations_per_step': n_conn, 'params': m16.n_params},
        'energy_per_site': float(mom[0]) / (H * W),
        'roofline': {'bound': 'mfma', 'achieved': round(achieved / 1e12, 2), 'peak': PEAK_BF16 / 1e12,
                     'unit': 'TFLOP/s', 'frac': round(achieved / PEAK_BF16, 4),
                     'traffic': (round(FWD_HBM_BYTES_PER_CONFIG[policy] * n_conn)
                                 if policy in FWD_HBM_BYTES_PER_CONFIG else None),
                     'row_reuse_first_row': policy,
                     'kernel': 'convar_fwd_kernel<fp16,32>', 'kernel_ms': round(fwd_ms, 3),
                     'flop_per_launch': flop_launch,
                     'note': 'achieved = reference-algorithm FLOPs (a full forward per used connection) per '
                             'second of the forward launches; with row reuse the kernels execute fewer',
                     'executed_flop_per_launch': round(executed),
                     'executed_tflops': round(executed / (fwd_ms * 1e-3) / 1e12, 2),
                     'executed_frac': round(executed / (fwd_ms * 1e-3) / PEAK_BF16, 4)},
        'forward_schedule': tune,
        'samples_per_s': sps,
        'connected_logpsi_per_s': round(n_conn * args.steps / elapsed, 1),
    }

    if not args.no_fp32:
        # plain FP32 path: every slot of every sample (per-sample padding), FP32 MFMA
        nc = op.num_connections
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = op.find_conn_device(spins, want_conn=True)
        flat = r['conn'].reshape(nc * B, H * W)
        lp, _ = m32.forward(flat)
        eloc32 = torch.empty((B, 2), dtype=torch.float64, device=dev)
        _lib.call('fk_local_energy_dense', _lib.ptr(lp), _lib.ptr(r['mel']), None, nc, B, _lib.ptr(eloc32),
                  _lib.stream_ptr())
        torch.cuda.synchronize()
        dt32 = time.perf_counter() - t0
        e32 = torch.complex(eloc32[:, 0], eloc32[:, 1]).mean().item()
        out['fp32_plain'] = {'E_loc_evals_per_s': round(B / dt32, 2), 'ms_per_step': round(dt32 * 1e3, 1),
                             'evaluations': nc * B, 'energy_per_site': e32.real / (H * W)}
        out['speedup_vs_fp32_plain'] = round(value / (B / dt32), 2)

    if not args.no_sr:
        out['vmc_sr_step'] = bench_sr(m16, op, spins)
        out['vmc_sr_training'] = bench_sr_training()
    if not args.no_mh:
        out['mh_rbm_8x8'] = bench_mh()
    if not args.no_large:
        out['large_20x20_d24_c64'] = bench_large()
    out['small_4x4_exact'] = bench_small()
    if not args.no_large:
        out['j1j2_10x10_c48'] = bench_j1j2()
        out['trained_ising_12x12_eval'] = bench_trained_ising()
    # ---- the same ragged step with BF16 MFMA (same speed class, coarser
    # rounding: it does not meet the 1e-4 energy-per-site bar)
    elb = obs.local_values(mb, spins)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        obs.local_values(mb, spins)
    torch.cuda.synchronize()
    dtb = (time.perf_counter() - t0) / args.steps
    out['bf16'] = {'E_loc_evals_per_s': round(B / dtb, 2), 'ms_per_step': round(dtb * 1e3, 3),
                   'speedup_vs_fp32_plain': round(B / dtb / out['fp32_plain']['E_loc_evals_per_s'], 2)
                   if 'fp32_plain' in out else None}

    # ---- accuracy: the same samples through the FP32 ragged path (all B),
    # and the FP64 reference on the CPU-baseline sample
    el16 = obs.local_values(m16, spins)
    el32 = obs.local_values(m32, spins)
    e32 = el32.real.mean().item() / (H * W)
    acc = {'bar': 'energy per site within 1e-4 of FP64; E_loc within 1e-4 relative per sample',
           'energy_per_site_fp32': e32,
           'eloc_std_per_site_fp32': el32.real.std().item() / (H * W),
           'eloc_stat_err_per_site': el32.real.std().item() / (H * W) / np.sqrt(B)}
    for name, el in (('fp16', el16), ('bf16', elb)):
        e = el.real.mean().item() / (H * W)
        acc[f'energy_per_site_{name}'] = e
        acc[f'energy_per_site_err_{name}_vs_fp32'] = abs(e - e32)
        acc[f'eloc_rms_rel_err_{name}_vs_fp32'] = float(((el - el32).abs().pow(2).mean().sqrt()
                                                         / el32.abs().pow(2).mean().sqrt()).item())
    if not args.no_cpu:
        cb, el64 = cpu_baseline(m32, spins.cpu().numpy())
        out['cpu_baseline'] = cb
        n64 = len(el64)
        acc['fp64_reference_samples'] = n64
        for name, el in (('fp32', el32), ('fp16', el16), ('bf16', elb)):
            g = el[:n64].cpu().numpy()
            rel = np.abs(g - el64) / np.abs(el64)
            acc[f'eloc_max_rel_err_{name}_vs_fp64'] = float(rel.max())
            acc[f'eloc_rms_rel_err_{name}_vs_fp64'] = float(np.sqrt(np.mean(rel ** 2)))
            acc[f'energy_per_site_err_{name}_vs_fp64'] = float(abs(g.real.mean() - el64.real.mean()) / (H * W))
        # the 1e-4 energy-per-site bar: FP32 is within 1e-4 per sample of the
        # FP64 reference; each 16-bit path against FP32 on all B samples
        fp32_ok = acc['eloc_max_rel_err_fp32_vs_fp64'] <= 1e-4
        acc['within_1e-4_energy_per_site'] = {
            name: bool(fp32_ok and acc[f'energy_per_site_err_{name}_vs_fp32'] <= 1e-4) for name in ('fp16', 'bf16')}
        acc['within_1e-4_energy_per_site']['fp32'] = bool(fp32_ok)
    out['accuracy'] = acc
    print(json.dumps(out))
    return 0


if __name__ == '__main__':
    sys.exit(main())
