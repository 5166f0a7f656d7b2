"""CPU oracle for the POLFED hot path -- TEST INFRASTRUCTURE ONLY.

This package is a plain-numpy (plus one C file) restatement of the reference
Polfed.jl algorithms on the hot path.  It exists only to *check* the HIP
library: it may be imported by ``tests/``, by ``__graft_entry__.smoke()`` and
by the ``cpu_baseline`` leg of ``bench.py`` -- never by the product package
``polfed.jl_amd`` (imported as ``polfed_amd``), which must fail loudly when its
HIP library is missing instead of falling back to anything here.

Every function cites the reference file:line it restates.  The reference is
Julia (not installed here), so the oracle is pinned by (a) the reference's own
known-answer tests (``tests/golden/reference_kat.json``), (b) closed forms and
invariants (dense eigendecompositions, Chebyshev identities, Pascal identities
for ranks), and (c) brute force on tiny inputs.  Parity status per module:

* ``models``     -- pinned: reference KAT matrices, brute-force basis, rank
                    closed form (bit-exact).
* ``packing``    -- pinned: exact flat/starts arrays of test/optimization.jl.
* ``clenshaw``   -- pinned: closed forms, numpy chebval, dense spectra.
* ``dos``        -- pinned: trace on identity, D-moments, rho(0)=4/pi, kernel
                    values of test/test_spectral_helpers.jl, exact traces.
* ``spectral``   -- pinned: bounds +-0.5 for flat rho, K fixed, bisection KAT.
                    QuadGK adaptive integration is restated with
                    scipy.integrate.quad at the same tolerances: agreement to
                    tolerance only ("parity pinned to 1e-7 by the reference
                    test", exact K may differ by one when a bisection decision
                    sits on the integration error).
* ``lanczos``    -- pinned: exact spectra of dense problems, the small-POLFED
                    closed-form answer {-0.2, 0.55, 1.8}, orthogonality and
                    block-tridiagonal invariants.
* ``cpu_ref.c``  -- pinned against ``packing``/``clenshaw`` on the same inputs.
"""
