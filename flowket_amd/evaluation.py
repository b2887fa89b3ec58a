"""Monte Carlo evaluation of a trained wave function: evaluation/evaluate.py:15-28
with the logs of the default statistics callbacks
(callbacks/monte_carlo/__init__.py:10-29): the energy logs of
local_energy_stats.py:15-20 and the sigma_z / abs_sigma_z observables of
observable.py:16-18.

``evaluate(generator, steps)`` draws ``steps`` batches from a
VariationalMonteCarlo generator and returns the mean over batches of the
batch energy, the batch variance of Re E_loc, the batch means of the
observables and, given the true ground-state energy, the relative error
(E_true - E) / E_true -- the numbers experiments/run_evaluation.py:12-26
prints.
"""
import numpy as np

from .observables import AbsSigmaZ, SigmaZ


def _mcmc_logs(generator):
    """callbacks/monte_carlo/mcmc_stats.py:11-17 for a Metropolis-Hastings
    sampler: acceptance ratio, R-hat, correlation sum and effective sample
    size of Re E_loc."""
    sampler = generator.sampler
    r_hat, _, correlations_sum, ess = sampler.calc_r_hat_value(generator.current_local_energy.real)
    return {'mcmc/acceptance_ratio': float(sampler.acceptance_ratio), 'mcmc/energy_r_hat': float(r_hat),
            'mcmc/energy_effective_sample_size': float(ess),
            'mcmc/energy_correlations_sum': float(correlations_sum)}


def evaluate(generator, steps, true_ground_state_energy=None, observables=None, mcmc_stats=False):
    """``mcmc_stats`` adds the MCMCStats logs (Metropolis-Hastings samplers
    only)."""
    if observables is None:
        observables = {'sigma_z': SigmaZ(), 'abs_sigma_z': AbsSigmaZ()}
    if mcmc_stats and not hasattr(generator.sampler, 'calc_r_hat_value'):
        raise ValueError('mcmc_stats needs a Metropolis-Hastings sampler')
    logs_arr = []
    for _ in range(int(steps)):
        next(generator)
        energy = float(np.real(generator.current_energy))
        logs = {'energy/energy': energy,
                'energy/local_energy_variance': float(np.real(generator.current_local_energy_variance))}
        if true_ground_state_energy is not None:
            logs['energy/relative_error'] = (true_ground_state_energy - energy) / true_ground_state_energy
        for name, obs in observables.items():
            value, _, _ = obs.estimate(generator.wave_function, generator.current_batch)
            logs['observables/%s' % name] = float(np.real(value))
        if mcmc_stats:
            logs.update(_mcmc_logs(generator))
        logs_arr.append(logs)
    if not logs_arr:
        return {}
    return {key: float(np.mean([logs[key] for logs in logs_arr])) for key in logs_arr[0]}


def exact_evaluate(exact_variational, true_ground_state_energy=None):
    """evaluation/evaluate.py:31-37 with the exact statistics callbacks
    (callbacks/exact/local_energy.py:12-17, callbacks/exact/sigma_z.py:20-22):
    the exact energy and variance of the current machine, and the
    probability-weighted sigma_z and |sigma_z| over all 2^N configurations
    (FP64 device reductions)."""
    import torch
    from . import _lib
    ev = exact_variational
    ev.machine_updated()
    obs = ev.energy_observable
    energy = float(np.real(obs.current_energy))
    logs = {'energy/energy': energy, 'energy/local_energy_variance': float(obs.current_local_energy_variance)}
    if true_ground_state_energy is not None:
        logs['energy/relative_error'] = (true_ground_state_energy - energy) / true_ground_state_energy
    vals = torch.empty((ev.num_of_states, 2), dtype=torch.float64, device=ev.device)
    coef = torch.empty_like(vals)
    mom = torch.empty(3, dtype=torch.float64, device=ev.device)
    for name, absolute in (('abs_sigma_z', 1), ('sigma_z', 0)):
        _lib.call('fk_sigma_z', _lib.ptr(ev.states), ev.num_of_states, ev.number_of_spins, absolute, _lib.ptr(vals),
                  _lib.stream_ptr())
        _lib.call('fk_exact_energy', _lib.ptr(ev.probs), _lib.ptr(vals), ev.num_of_states, _lib.ptr(coef),
                  _lib.ptr(mom), _lib.ptr(ev._ws), _lib.stream_ptr())
        logs['observables/%s' % name] = float(mom[0].item())
    return logs
