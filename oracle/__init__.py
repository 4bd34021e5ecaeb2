"""CPU checkers for the tree-gravity path. Test infrastructure only.

``liboracle.so`` is the C restatement (``gz_oracle.c``); ``_ref/libgzref.so`` is the reference's
own ``shmem/tree.cc`` compiled by ``oracle/Makefile``. Only the tests, the smoke check and the
CPU-baseline leg of the benchmark import this package.
"""
