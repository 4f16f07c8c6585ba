"""Oracle: CPU restatement of the tinyvlm reference backend.

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline
leg of ``bench.py`` as the checker; the HIP product path never imports, links or calls it.
See ``oracle/vlm_oracle.h`` for the functions, the reference file:line each one follows and the
list of checks that pin it.
"""
