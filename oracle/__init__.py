"""Scalar FP64 host oracle (TEST INFRASTRUCTURE ONLY).

Loaded by tests/, __graft_entry__.smoke() and the CPU-baseline leg of
bench.py as the checker of the HIP engine; the engine never imports it.
The C++ restatement lives in oracle.cpp; this module binds it with ctypes.
"""
from .binding import Oracle, load_oracle, ORACLE_LIB  # noqa: F401
