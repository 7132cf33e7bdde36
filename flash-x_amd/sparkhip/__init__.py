"""Host-side Python plumbing around libspark_hip.so (the MI355X Spark RK stage).

capi      ctypes binding of include/spark_hip.h
problems  synthetic block sets of the benchmark configurations
"""
from . import capi, problems  # noqa: F401
