"""FP64 CPU oracle of the MMSB tensor-method pipeline -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this package,
and only as the checker (or the timed CPU baseline); the product path never calls it.
"""
from .core import *  # noqa: F401,F403
