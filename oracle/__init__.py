"""CPU oracle for the SpecAnn alpha-sweep hot path — TEST INFRASTRUCTURE ONLY.

A numpy/scipy restatement of the reference's algorithm (Vgueca/SpecAnn,
models/spectral_annealing.py, utils/predictors.py, utils/eigsh.py, models/spec_adj.py,
models/spec_signed_lap.py), used as the checker of the HIP path: by ``tests/``, by
``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of ``bench.py``. The product path
(``specann_amd``) never imports it.

How each function is pinned (checked against something other than itself):
  normalize, alpha_values, alpha_scaling, shift_rho, gershgorin  closed forms (tests/test_oracle.py)
  n_matvec / s_matvec        dense diag(s) J diag(s) products
  leading_eigpair            numpy.linalg.eigh on dense matrices; ARPACK cross-check
  diagonal_predictor         closed form + first-order identity in delta alpha
  secant_predictor           closed form (sign alignment, fallbacks)
  tangent                    central finite differences of leading_eigpair in alpha, dense pinv
  round_spins, spin_bits     exact; bit layout checked against numpy.packbits
  energy_int64               brute force -1/2 s^T J s in float64 and exhaustive ground states (n <= 20)
  greedy_colouring           networkx.greedy_color with the same node order
  local_search               a scalar one-node-at-a-time loop (tests/test_oracle.py), 1-flip optimality,
                             monotone energy, exhaustive optimum bound on tiny instances
  normal_vector, mix3        SplitMix64 published constants; normality checked statistically
  cpu_sweep                  composed of the above; sweep-level invariants
No function is left "parity unpinned" at this level; the CuPy eigsh internals themselves are not
restated (third-party, absent here): the eigensolver is pinned by eigenpair invariants instead.
"""
