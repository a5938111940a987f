"""FP64 CPU restatement of plainmp's hot path.

Test infrastructure only: imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of ``bench.py``, always as the checker.  The
product package ``plainmp_amd`` never imports it.
"""
