"""ORACLE — test infrastructure only.

Checkers for the matched-filter hot path:

* ``oracle.cpu``: ctypes bindings to ``libfmf_oracle.so`` (the C restatement of
  the reference ``arch='cpu'`` / ``'precise'`` / no-sum / ``variable_precise``
  algorithms, ``fmf_oracle.c``) and, where it was built, to the compiled
  reference under ``oracle/_ref``.
* ``oracle.fp64``: an FP64 NumPy restatement of the same mathematics (exact
  normalized cross-correlation), used as the precision yardstick.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this package, and only as the checker or the
timed CPU baseline.  The product package never imports it.
"""
