"""evaluate() with the MCMCStats logs (callbacks/monte_carlo/mcmc_stats.py:11-17)
of a Metropolis-Hastings sampler."""
import numpy as np
import pytest

from tests.conftest import gpu_banner

pytestmark = pytest.mark.gpu


def test_evaluate_mcmc_logs():
    import flowket_amd as fk
    from flowket_amd.evaluation import evaluate
    gpu_banner('evaluate with MCMC logs, 4x4 RBM')
    m = fk.RBM((4, 4), alpha=2, stddev=0.2, seed=3)
    op = fk.Ising(h=2.0, hilbert_state_shape=(4, 4), pbc=True)
    sampler = fk.MetropolisHastingsLocal(m, 1024, num_of_chains=16, unused_sampels=16, seed=4)
    vmc = fk.VariationalMonteCarlo(m, op, sampler)
    r = evaluate(vmc, 1, mcmc_stats=True)
    assert 0.0 < r['mcmc/acceptance_ratio'] <= 1.0
    from oracle import samplers as osa
    r_hat, _, cs, ess = osa.r_hat_and_ess(vmc.current_local_energy.real.double().cpu().numpy(), 16)
    assert r['mcmc/energy_r_hat'] == pytest.approx(float(r_hat))
    assert r['mcmc/energy_effective_sample_size'] == pytest.approx(float(ess))
    assert r['mcmc/energy_correlations_sum'] == pytest.approx(float(cs))
    assert 0.0 < r['mcmc/energy_effective_sample_size'] <= 1024 * 1.0001
    assert np.isfinite(r['mcmc/energy_r_hat'])
    with pytest.raises(ValueError):
        evaluate(fk.VariationalMonteCarlo(m, op, fk.ExactSampler(fk.ExactVariational(m, op, 65536), 64)), 1,
                 mcmc_stats=True)
