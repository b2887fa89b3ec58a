"""CPU oracle for the FlowKet VMC hot path -- TEST INFRASTRUCTURE ONLY.

This package is a plain numpy (float64) restatement of the reference algorithms
that the HIP path in ``flowket_amd`` replaces.  Every function cites the
reference file:line it follows (paths relative to the FlowKet source tree,
``src/flowket/...``).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import anything from here, and only as the checker (or, for
``bench.py``, as the timed CPU baseline).  The product path in ``flowket_amd``
never imports this package and has no CPU fallback: it fails loudly when the
HIP library is missing.

Pinning (how each function is checked; the tests are tests/test_oracle.py
unless named):
  operators.*_find_conn      ED ground energies of the oracle's matrix
                             elements vs published values (4x4 Ising h=3 obc,
                             16-site chain, 4x4 J1J2 obc) and closed forms
                             (4x4 Heisenberg pbc, 10-site chain); structural
                             invariants; hand-transcribed (slot, a, b) known
                             answers of every Heisenberg connection slot
                             (heisenberg.py:63-68,86-121) on all 2^N states;
                             the Ising float32 diagonal against the rule
                             written out (per-site float32 rounding, numpy's
                             pairwise float32 sum).
  observables.local_values_* the reference's own unbalanced-vs-balanced fixture
                             (tests/test_variational.py:22-32) and the exact
                             identity <E_loc>_{|psi|^2} = <psi|H|psi>/<psi|psi>.
  convnet.*                  sum_s |psi(s)|^2 == 1 (tests/test_autoregressive.py
                             :18-30), strict causality of every conditional, a
                             torch.nn.functional.conv2d restatement of every
                             convolution, and a closed form of the first site.
  samplers.*                 exact-distribution z-test (tests/test_samplers.py
                             :63-82), detailed balance on 2^N enumerations,
                             ESS of i.i.d. and AR(1) chains, and the package's
                             calc_r_hat_value on fixed chains.
  rbm.*                      lncosh identities (tests/test_tensorflow_complex_
                             numbers_ops.py:9-33) and brute-force sums.
  made.*                     sum_x q(x) == 1 on all 2^N states, backprop vs
                             central finite differences, sampler z-test
                             (a learned proposal with no reference
                             counterpart: pinned by mathematics only).
  sr.*                       dense pinv solve (tests/test_stochastic_
                             reconfiguration.py:19-72), explicit S.v.
  exact.*                    the Rayleigh quotient of an independently
                             assembled sparse H, and E_loc == E0 at the
                             ground state.
  optim.keras_adam           the closed form of the first step.
  ensemble.*                 the probabilistic ensemble's closed forms
                             (tests/test_ensemble.py).
"""
