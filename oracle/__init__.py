"""CPU oracle for the HCORE tile low-rank hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the hcore_amd product.  See
hcore_oracle.c for the reference lines each function restates.
"""
from .oracle import *  # noqa: F401,F403
