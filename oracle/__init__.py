"""CPU oracle for the Mi3 inference round (test infrastructure only).

Nothing in the product path imports this package.  It is used by `tests/`,
by `__graft_entry__.smoke()` and by the `cpu_baseline` leg of `bench.py`, and
only as the checker or as the timed CPU port, never as the thing measured on
the GPU.

Contents
--------
potts_oracle.py   numpy restatements (fp64 where the reference accumulates in
                  fp32) of every kernel on the inference-round path, each
                  citing the reference file:line it follows.
philox.py         numpy Philox4x32-10 (Salmon et al., SC'11) and the stream
                  layout used by the sampler, positions and independent
                  reseeding.
oracle.c          the same algorithms in plain C (OpenMP over walkers) for
                  sizes numpy cannot finish quickly; also the timed CPU port
                  of the sampler (`cpu_baseline.kind = "port"`).
cboracle.py       ctypes loader for the compiled oracle.c.
Makefile          builds oracle/_build/liboracle.so with gcc.

Pinning (what fixes each function independently of its own formula) is
listed per function in the docstrings and collected in tests/test_oracle.py.
"""
