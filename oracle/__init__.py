"""CPU oracle for the ladder-DGA hot path (test infrastructure only).

This package is a NumPy complex128 restatement of the reference algorithm
(Julpe/DGAmore) for the nonlocal ladder-DGA hot path. It exists only to
check the HIP implementation and to time a CPU baseline:

* only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
  ``bench.py`` may import it;
* the product package ``dgamore_amd`` never imports it and has no CPU
  fallback.

Every function cites the reference file:line it restates. Parity status per
function is listed in ``dga_oracle.PARITY``.
"""
