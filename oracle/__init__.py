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

Pinning (how each function is checked, see ``tests/test_oracle_*.py``):
  operators.find_conn_*      brute-force ED spectra vs published/closed-form
                             ground energies, structural invariants, and the
                             reference loop restated independently for tiny
                             lattices.
  observables.local_values_* the reference's own unbalanced-vs-balanced fixture
                             (tests/test_variational.py:22-32) and the exact
                             identity <E_loc>_{|psi|^2} = <psi|H|psi>/<psi|psi>.
  convnet.forward            sum_s |psi(s)|^2 == 1 (tests/test_autoregressive.py
                             :18-30), strict causality of every conditional, a
                             torch.nn.functional.conv2d restatement of every
                             convolution, and 4-parameter closed forms.
  samplers.*                 exact-distribution z-test (tests/test_samplers.py
                             :63-82) and detailed balance on 2^N enumerations.
  rbm.*                      lncosh identities (tests/test_tensorflow_complex_
                             numbers_ops.py:9-33) and brute-force sums.
  made.*                     sum_x q(x) == 1 on all 2^N states, backprop vs
                             central finite differences, sampler z-test
                             (a learned proposal with no reference
                             counterpart: pinned by mathematics only).
  sr.*                       dense pinv solve (tests/test_stochastic_
                             reconfiguration.py:19-72), explicit S.v.
"""
