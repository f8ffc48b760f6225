"""TEST INFRASTRUCTURE ONLY: the oracle.

A CPU restatement of the reference's CPU multigrid (``MultiGridScheme<T>``,
VLP/Model/CPU/MultiGridScheme.cs) in double, float and double-double, plus an
independent DST-I direct solver.  Imported only by ``tests/``, by the smoke
check in ``__graft_entry__.py`` and by the CPU-baseline leg of ``bench.py``,
always as the checker.  The product library (``vlp2d_amd``) never imports it.
"""
from .oracle import Oracle, load, mixed_solve, plain_solve, forward_cycles, adjoint_cycles  # noqa: F401
from .dst import dst_solve, dst_adjoint  # noqa: F401
