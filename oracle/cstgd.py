"""ctypes binding of stgd_cpu.c, the FP64 C / OpenMP restatement of the STGD loop -- TEST
INFRASTRUCTURE ONLY (a second restatement checked against core.stgd, and the compiled CPU
baseline of the benchmark). Built by this folder's Makefile into _build/libstgd_cpu.so."""
import ctypes
import os

import numpy as np

from .core import coeffs_for

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "libstgd_cpu.so")
PHASES = ("contractions", "t_hat", "step", "renorm")
_lib = None


def available() -> bool:
    return os.path.exists(LIB_PATH)


def _load():
    global _lib
    if _lib is None:
        if not available():
            raise ImportError(f"{LIB_PATH} is not built (make -C oracle)")
        lib = ctypes.CDLL(LIB_PATH)
        d, i32, i64 = ctypes.POINTER(ctypes.c_double), ctypes.c_int32, ctypes.c_int64
        f = ctypes.c_double
        lib.stgd_cpu_f64.argtypes = [d, d, d, i64, i32, d, d, d, d, f, f, f, f, f, i64, i64, i32, i32, i32, d, d]
        lib.stgd_cpu_f64.restype = ctypes.c_int
        _lib = lib
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def stgd(Y_A, Y_B, Y_C, mu_A, mu_B, mu_C, phi0, alpha0, eta, batch, epochs=1, steps=None, theta=1.0,
         ns_iters=2, ns_degree=3, scaling="paper", threads=0):
    """Same arguments and minibatch order as core.stgd (renorm_mode "ns", a renormalisation every
    step); returns (Phi, log of ||dPhi||_F^2 per step, wall seconds per phase)."""
    lib = _load()
    ya, yb, yc = (np.ascontiguousarray(y, np.float64) for y in (Y_A, Y_B, Y_C))
    n, k = ya.shape
    if yb.shape != (n, k) or yc.shape != (n, k) or phi0.shape != (k, k):
        raise ValueError("Y_A, Y_B, Y_C must be n x k and Phi k x k")
    mus = [np.ascontiguousarray(v, np.float64) for v in (mu_A, mu_B, mu_C)]
    if any(v.shape != (k,) for v in mus):
        raise ValueError("the whitened means must have k entries")
    if batch <= 0:
        raise ValueError("the minibatch must hold at least one node")
    phi = np.array(phi0, np.float64, order="C", copy=True)
    total = steps if steps is not None else epochs * ((n + batch - 1) // batch)
    log = np.zeros(max(total, 1), np.float64)
    ph = np.zeros(4, np.float64)
    c3, c1, c0 = coeffs_for(alpha0, scaling)
    rc = lib.stgd_cpu_f64(_ptr(ya), _ptr(yb), _ptr(yc), n, k, _ptr(mus[0]), _ptr(mus[1]), _ptr(mus[2]), _ptr(phi),
                          c3, c1, c0, eta, theta, batch, total, ns_iters, ns_degree, threads, _ptr(log), _ptr(ph))
    if rc != 0:
        raise ValueError(f"stgd_cpu_f64 returned {rc}")
    return phi, log[:total], dict(zip(PHASES, ph.tolist()))
