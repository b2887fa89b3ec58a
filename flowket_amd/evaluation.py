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


def evaluate(generator, steps, true_ground_state_energy=None, observables=None):
    if observables is None:
        observables = {'sigma_z': SigmaZ(), 'abs_sigma_z': AbsSigmaZ()}
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
        logs_arr.append(logs)
    if not logs_arr:
        return {}
    return {key: float(np.mean([logs[key] for logs in logs_arr])) for key in logs_arr[0]}
