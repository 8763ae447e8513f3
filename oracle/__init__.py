"""ORACLE — test infrastructure only.

CPU restatement of the MOLSCAT/BOUND log-derivative propagation and
matching path (FP64), used to check the HIP engine.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
may import it, and only as the checker; the product package
``molscat_amd`` never imports it and fails loudly when its own HIP
library is missing.
"""
