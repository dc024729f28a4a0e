"""FP64 CPU oracle of the MMSB tensor-method pipeline -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this package,
and only as the checker (or the timed CPU baseline); the product path never calls it.
core.py restates the pipeline in numpy / scipy; stgd_cpu.c (bound by cstgd.py, built by this
folder's Makefile) restates the STGD loop in C / OpenMP; time_pipeline.py times both per phase.
"""
from .core import *  # noqa: F401,F403
