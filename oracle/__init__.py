"""CPU oracle for the qradient statevector ``grad_run`` path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package (``qradient_amd``)
imports, calls or links anything in this directory.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use
it, and only as the checker or as the timed CPU baseline -- never as the
thing that is measured on the GPU or shipped.

Contents
--------
``statevector``          numpy complex128 restatement of the reference algorithm,
                          matrix-free (reshape butterflies), with a
                          ``dtype`` switch used to model complex64 rounding.
``reference_algorithm``  the reference's own formulation (scipy CSR gate
                          operators, 0/1 Z masks, stored state history),
                          used as the CPU baseline and as a second,
                          independently constructed cross-check.

Every function cites the reference file:line it restates.  How each part is
pinned (dense brute force, parameter shift, finite differences, closed forms,
special angles) is listed in ``tests/test_oracle_pins.py``.  The reference
ships no golden vectors or known-answer tests for this path (its tests are
``pass`` stubs, reference tests/mc_clean.py:1-11), so the pins are
mathematical.  No oracle function is left "parity unpinned".
"""
