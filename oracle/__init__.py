"""CPU oracle for the soliton_solver relaxation core -- TEST INFRASTRUCTURE ONLY.

This package is a plain NumPy FP64 restatement of the reference algorithm
(Paulnleask/soliton_solver v1.1.1), written from reading its source as text.
It exists to check the HIP product path; it is never the product:

* only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
  ``bench.py`` may import or run anything under ``oracle/``;
* ``soliton_solver_amd`` (the product) never imports it and fails loudly when
  its HIP library is missing.

Layout of every array here is the reference's host layout: a field ``F`` has
shape ``(nf, xlen, ylen)`` so that ``F.ravel()`` is the reference's flat
``idx_field(a, x, y) = y + x*ylen + a*xlen*ylen`` order
(core/utils.py:16-42).  First derivatives are ``D1[a, i, x, y]`` (i = 0 for d/dx,
1 for d/dy); second derivatives are ``D2[a, k, x, y]`` with k = 0 (xx), 1 (yy),
2 (the mixed entry).

The checks that pin each function to the reference (closed forms, identities,
invariants) are in ``tests/test_oracle_pinning.py`` and
``tests/test_oracle_pinning_charges.py``.
"""
