"""Per-theory FP64 restatements of the reference's kernels (oracle).

TEST INFRASTRUCTURE ONLY: imported by the tests, ``__graft_entry__.smoke``
and the CPU-baseline leg of ``bench.py`` as the checker, never by the product
path.
"""
