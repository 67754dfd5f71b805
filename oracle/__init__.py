"""CPU oracle for the oa-fdtd time-stepping path.

TEST INFRASTRUCTURE: imported only by ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py``, to compare against. The product package never imports it.
"""

from .numpy_engine import OracleEngine, oracle_factory

__all__ = ["OracleEngine", "oracle_factory"]
