"""flowket_amd -- MI355X-native VMC hot path with FlowKet's public surface.

The compute runs in libflowket_hip.so (hand-written HIP kernels for gfx950,
C ABI in include/flowket_hip.h); this package is the Python host side that
mirrors FlowKet's operator / machine / sampler / VMC classes.
"""
from ._lib import FlowketHipError  # noqa: F401
from .operators import Operator, OperatorOnGrid, cube_shape, Heisenberg, Ising, J1J2, j1j2_edges  # noqa: F401
from .machines import ConvNetAutoregressive2D, RBM  # noqa: F401
from .observables import BaseObservable, Observable, SigmaZ, AbsSigmaZ, energy_moments  # noqa: F401
from .samplers import (Sampler, AutoregressiveSampler, FastAutoregressiveSampler,  # noqa: F401
                       MetropolisHastingsSampler, MetropolisHastingsLocal, MetropolisHastingsExchange,
                       calc_r_hat_value)
from .optimization import VariationalMonteCarlo, MiniBatchGenerator, loss_for_energy_minimization  # noqa: F401
from .optimizers import (ComplexValuesStochasticReconfiguration, Adam, SGD, energy_gradient, loss_gradient,  # noqa: F401
                         conjugate_gradient, AccumulateGradientOptimizer, convert_to_accumulate_gradient_optimizer)

from . import autotune  # noqa: F401
from .graphs import LocalEnergyGraph  # noqa: F401
from .proposals import MADEProposal, MetropolisHastingsLearnedProposal  # noqa: F401
from .exact import ExactVariational, ExactObservable, ExactSampler  # noqa: F401
from .ensemble import (SymmetrizedMachine, make_2d_obc_invariants, make_up_down_invariant,  # noqa: F401
                       make_pbc_invariants)
from .keras_weights import load_keras_weights  # noqa: F401
from .checkpoint import (CheckpointByTime, save_optimizer_weights, load_optimizer_weights,  # noqa: F401
                         save_weights, load_weights)

__version__ = '0.1.0'
