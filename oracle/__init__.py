"""FP64 CPU oracle for the KSpaceJet reconstruction hot path.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` may import this package, and only as the
checker.  The product package ``kspacejet_amd`` never imports it.
"""
from .oracle import *  # noqa: F401,F403
