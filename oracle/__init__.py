"""CPU oracle for the pfem2dealii particle hot path (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package; the product library never does.
"""
from .pyoracle import Oracle, FemOracle, build_oracle, MODE_CANONICAL, MODE_REFERENCE  # noqa: F401
