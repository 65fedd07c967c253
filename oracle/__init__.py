"""CPU oracle for the recursive-block SpTRSV path.

Test infrastructure only: tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg use it as the checker.  The product library
never imports, links or calls anything here.
"""
from .pyoracle import *  # noqa: F401,F403
