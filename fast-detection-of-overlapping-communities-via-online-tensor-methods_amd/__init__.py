"""ctypes binding of libmmsb_amd.so, the MI355X-native MMSB tensor-method library.

The library's C-ABI (include/mmsb_stpm.h) is the product boundary. This module only declares
argument types, converts status codes into exceptions and offers small helpers for passing
torch tensors (device memory and streams) through the ABI. It never computes anything itself:
if the shared library is missing it raises ImportError instead of falling back to CPU code.

torch is imported before the library is opened so that the process holds a single HIP runtime:
torch ships its own libamdhip64.so.7 and the library's DT_NEEDED entry then resolves to it.
"""
from __future__ import annotations

import atexit
import ctypes
import os

try:  # one HIP runtime per process: torch's copy must be resident first
    import torch  # noqa: F401
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libmmsb_amd.so")

MMSB_OK = 0
MMSB_ERR_ARG = -1
MMSB_ERR_IO = -2
MMSB_ERR_HIP = -3
MMSB_ERR_NOMEM = -4
MMSB_ERR_NUMERIC = -5
MMSB_ERR_STATE = -6


class MMSBError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mmsb status {code}: {msg}")
        self.code = code


_c_int32_p = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p


class CsrC(ctypes.Structure):
    _fields_ = [("n_rows", ctypes.c_int32), ("n_cols", ctypes.c_int32), ("nnz", ctypes.c_int64),
                ("rowptr", ctypes.POINTER(ctypes.c_int32)), ("colidx", ctypes.POINTER(ctypes.c_int32)),
                ("vals", ctypes.POINTER(ctypes.c_float))]


class GraphC(ctypes.Structure):
    _fields_ = [("xa", CsrC), ("xb", CsrC), ("xc", CsrC), ("ba", CsrC), ("ca", CsrC),
                ("n_nodes", ctypes.c_int64), ("n_edges", ctypes.c_int64), ("k", ctypes.c_int32),
                ("theta", ctypes.POINTER(ctypes.c_double)), ("part", ctypes.POINTER(ctypes.c_int8)),
                ("local", ctypes.POINTER(ctypes.c_int32)), ("pi", ctypes.POINTER(ctypes.c_double))]


class GenParamsC(ctypes.Structure):
    _fields_ = [("n_x", ctypes.c_int32), ("n_a", ctypes.c_int32), ("n_b", ctypes.c_int32), ("n_c", ctypes.c_int32),
                ("k", ctypes.c_int32), ("alpha0", ctypes.c_double), ("p_in", ctypes.c_double),
                ("p_out", ctypes.c_double), ("seed", ctypes.c_uint64), ("degree_corrected", ctypes.c_int32),
                ("pareto_shape", ctypes.c_double), ("mode", ctypes.c_int32), ("threads", ctypes.c_int32),
                ("verbose", ctypes.c_int32)]


GEN_MODES = {"auto": 0, "exact": 1, "balls": 2}


class StpmConfigC(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("na", ctypes.c_int32), ("nb", ctypes.c_int32), ("nc", ctypes.c_int32),
                ("ncom", ctypes.c_int32), ("alpha_0", ctypes.c_double), ("learn_rate", ctypes.c_double),
                ("thresh", ctypes.c_double), ("pinv_tol", ctypes.c_double), ("sgd_iter_min", ctypes.c_int64),
                ("sgd_iter_max", ctypes.c_int64), ("binary", ctypes.c_int32), ("synthetic", ctypes.c_int32),
                ("folder_name", ctypes.c_char * 512), ("file_prefix", ctypes.c_char * 64),
                ("alpha_0_str", ctypes.c_char * 32), ("learn_rate_str", ctypes.c_char * 32),
                ("minibatch", ctypes.c_int32), ("epochs", ctypes.c_int32), ("theta", ctypes.c_double),
                ("renorm_every", ctypes.c_int32), ("ns_iters", ctypes.c_int32), ("ns_degree", ctypes.c_int32),
                ("precision", ctypes.c_int32),
                ("polish_epochs", ctypes.c_int32), ("oversample", ctypes.c_int32), ("power_iters", ctypes.c_int32),
                ("m2_scaling", ctypes.c_int32), ("member_tau", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("ba_first_is_row", ctypes.c_int32), ("feof_repeat", ctypes.c_int32), ("device", ctypes.c_int32),
                ("verbose", ctypes.c_int32), ("use_graphs", ctypes.c_int32),
                ("y_mode", ctypes.c_int32), ("renorm_mode", ctypes.c_int32),
                ("csr_host", ctypes.c_int32)]


PREC = {"fp32": 0, "bf16": 1}
M2_SCALING = {"paper": 0, "gpu": 1}
Y_MODE = {"auto": 0, "stored": 1, "recompute": 2}
RENORM = {"ns": 0, "qr": 1}
BLOCKS = {"XA": 0, "XB": 1, "XC": 2, "BA": 3, "CA": 4}
QUANTITIES = ["mu_a", "mu_b", "mu_c", "W", "U", "V", "sigma", "LB", "LC", "lam_m2", "YA", "YB", "YC",
              "muw_a", "muw_b", "muw_c", "phi", "frob", "lambda", "lambda_p", "alpha_hat",
              "pi_x", "pi_a", "pi_b", "pi_c", "pin_x", "pin_a", "pin_b", "pin_c",
              "count_x", "count_a", "count_b", "count_c", "times", "sweeps"]
Q = {name: i for i, name in enumerate(QUANTITIES)}

# name -> (restype, argtypes); every symbol declared in include/mmsb_stpm.h appears here.
SIGNATURES: dict = {
    "mmsb_last_error": (ctypes.c_char_p, []),
    "mmsb_version": (ctypes.c_char_p, []),
    "mmsb_spmm_f32": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, _vp, _vp, _vp,
                                     ctypes.c_float, _vp, ctypes.c_int32, ctypes.c_int32, _vp, ctypes.c_int32]),
    "mmsb_spmm_f32_ordered": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, _vp, _vp, _vp,
                                             ctypes.c_float, _vp, ctypes.c_int32, ctypes.c_int32, _vp, ctypes.c_int32,
                                             _vp]),
    "mmsb_spmm_f32_host_csr": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, _vp, _vp, _vp,
                                              ctypes.c_float, _vp, ctypes.c_int32, ctypes.c_int32, _vp, ctypes.c_int32,
                                              ctypes.c_int64]),
    "mmsb_csr_row_order": (ctypes.c_int, [_vp, ctypes.c_int32, _vp, _vp]),
    "mmsb_gen_params_default": (None, [ctypes.POINTER(GenParamsC)]),
    "mmsb_generate": (ctypes.c_int, [ctypes.POINTER(GenParamsC), ctypes.c_int32, ctypes.POINTER(ctypes.POINTER(GraphC))]),
    "mmsb_graph_free": (None, [ctypes.POINTER(GraphC)]),
    "mmsb_read_coo": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(CsrC)]),
    "mmsb_csr_free": (None, [ctypes.POINTER(CsrC)]),
    "mmsb_write_coo": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(CsrC)]),
    "mmsb_write_values": (ctypes.c_int, [ctypes.c_char_p, _vp, ctypes.c_int64]),
    "mmsb_write_pi_triplets": (ctypes.c_int, [ctypes.c_char_p, _vp, ctypes.c_int64, ctypes.c_int32]),
    "stpm_config_default": (None, [ctypes.POINTER(StpmConfigC)]),
    "stpm_config_size": (ctypes.c_size_t, []),
    "mmsb_create": (ctypes.c_int, [ctypes.POINTER(StpmConfigC), ctypes.POINTER(_vp)]),
    "mmsb_destroy": (None, [_vp]),
    "mmsb_set_block": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.POINTER(CsrC)]),
    "mmsb_read_inputs": (ctypes.c_int, [_vp]),
    "mmsb_whiten": (ctypes.c_int, [_vp]),
    "mmsb_set_phi": (ctypes.c_int, [_vp, _vp]),
    "mmsb_set_yvectors": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int32, _vp, _vp, _vp, _vp, _vp, _vp]),
    "mmsb_stgd": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int32]),
    "mmsb_stgd_timed": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int32, _vp]),
    "mmsb_set_learn_rate": (ctypes.c_int, [_vp, ctypes.c_double]),
    "mmsb_batch_create": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, _vp]),
    "mmsb_batch_run": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "mmsb_batch_autotune": (ctypes.c_int, [_vp, _vp, _vp]),
    "mmsb_batch_destroy": (None, [_vp]),
    "mmsb_eigen": (ctypes.c_int, [_vp]),
    "mmsb_post": (ctypes.c_int, [_vp]),
    "mmsb_write_outputs": (ctypes.c_int, [_vp]),
    "mmsb_get": (ctypes.c_int, [_vp, ctypes.c_int32, _vp, ctypes.c_int64]),
    "mmsb_get_size": (ctypes.c_int64, [_vp, ctypes.c_int32]),
    "mmsb_get_rows": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, _vp, ctypes.c_int64]),
    "mmsb_get_f32": (ctypes.c_int, [_vp, ctypes.c_int32, _vp, ctypes.c_int64]),
    "mmsb_sync": (ctypes.c_int, [_vp]),
    "stpm_set_config": (ctypes.c_int, [ctypes.POINTER(StpmConfigC)]),
    "mmsb_tensor_coeffs": (ctypes.c_int, [ctypes.c_double, ctypes.c_int32, _vp]),
    "mmsb_stgd_grad_workspace": (ctypes.c_size_t, [ctypes.c_int32, ctypes.c_int32]),
    "mmsb_stgd_grad_f64": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, _vp, _vp, _vp, ctypes.c_int32, _vp, _vp, _vp,
                                          _vp, _vp, ctypes.c_double, _vp, _vp, _vp]),
    "mmsb_stgd_grad_f32": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, _vp, _vp, _vp, ctypes.c_int32, _vp, _vp, _vp,
                                          _vp, _vp, ctypes.c_double, _vp, _vp, _vp]),
    "mmsb_power_method": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, _vp]),
    "mmsb_request_stop": (ctypes.c_int, [_vp]),
    "stpm_force_exit": (None, []),
    "mmsb_grad": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int32, _vp, ctypes.c_int32, ctypes.c_double, ctypes.c_int32,
                                 _vp, _vp]),
    "mmsb_stgd_grad_f32_noalloc": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, _vp, _vp, _vp, ctypes.c_int32, _vp,
                                                   _vp, _vp, _vp, _vp, ctypes.c_double, _vp, _vp]),
    "mmsb_write_triplets": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_char_p]),
    "mmsb_release": (None, []),
    "mmsb_set_strip_limit": (ctypes.c_int64, [ctypes.c_int64]),
    "pre_proc": (ctypes.c_int, []),
    "GPU_Task": (_vp, [_vp]),
    "post_proc": (ctypes.c_int, []),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.stpm_config_size() != ctypes.sizeof(StpmConfigC):
        raise ImportError(f"stpm_config layout mismatch: the library has {lib.stpm_config_size()} bytes, the binding "
                          f"{ctypes.sizeof(StpmConfigC)}; rebuild the library")
    return lib


lib = _load()
# Free the library's process-wide pinned staging, copy stream and sort temporaries while the HIP
# runtime is still alive (Python's exit handlers run before the runtime's own teardown).
atexit.register(lib.mmsb_release)


def last_error() -> str:
    return lib.mmsb_last_error().decode()


def check(status: int) -> int:
    if status != MMSB_OK:
        raise MMSBError(status, last_error())
    return status


def ptr(t) -> int | None:
    """Device (or host) address of a torch tensor or numpy array; None stays NULL."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    return t.ctypes.data


def stream_handle(stream=None) -> int | None:
    """hipStream_t of a torch stream (default: torch's current stream on the current device)."""
    if torch is None or not torch.cuda.is_available():
        return None
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def spmm_f32(rowptr, colidx, vals, B, Y, r: int, n: int, alpha: float = 1.0, stream=None, order=None):
    """Y[:m, :r] = alpha * G @ B[:, :r] through mmsb_spmm_f32, or through mmsb_spmm_f32_ordered with the
    rows visited in `order` (all device tensors)."""
    m = rowptr.numel() - 1
    args = (stream_handle(stream), m, n, colidx.numel(), ptr(rowptr), ptr(colidx), ptr(vals), float(alpha), ptr(B),
            B.stride(0), r, ptr(Y), Y.stride(0))
    check(lib.mmsb_spmm_f32(*args) if order is None else lib.mmsb_spmm_f32_ordered(*args, ptr(order)))
    return Y


def spmm_f32_host_csr(rowptr, colidx, vals, B, Y, r: int, n: int, alpha: float = 1.0, chunk_nnz: int = 1 << 21,
                      stream=None):
    """Y[:m, :r] = alpha * G @ B[:, :r] through mmsb_spmm_f32_host_csr: G as host numpy arrays (int32
    rowptr / colidx, float32 vals or None), B and Y device tensors; returns when Y is complete."""
    import numpy as np
    rp = np.ascontiguousarray(rowptr, np.int32)
    ci = np.ascontiguousarray(colidx, np.int32)
    v = None if vals is None else np.ascontiguousarray(vals, np.float32)
    check(lib.mmsb_spmm_f32_host_csr(stream_handle(stream), rp.size - 1, n, ci.size, rp.ctypes.data, ci.ctypes.data,
                                     None if v is None else v.ctypes.data, float(alpha), ptr(B), B.stride(0), r,
                                     ptr(Y), Y.stride(0), int(chunk_nnz)))
    return Y


def csr_row_order(rowptr, stream=None):
    """The rows of a device CSR matrix by descending length (mmsb_csr_row_order), as an int32 tensor."""
    import torch
    m = rowptr.numel() - 1
    order = torch.empty(m, dtype=torch.int32, device=rowptr.device)
    check(lib.mmsb_csr_row_order(stream_handle(stream), m, ptr(rowptr), ptr(order)))
    return order


def _csr_to_numpy(c: CsrC):
    import numpy as np
    rowptr = np.ctypeslib.as_array(c.rowptr, shape=(c.n_rows + 1,)).copy()
    colidx = np.ctypeslib.as_array(c.colidx, shape=(max(c.nnz, 1),))[: c.nnz].copy()
    vals = None if not c.vals else np.ctypeslib.as_array(c.vals, shape=(max(c.nnz, 1),))[: c.nnz].copy()
    return rowptr, colidx, vals, c.n_rows, c.n_cols


def generate(n_x, n_a, n_b, n_c, k, alpha0, p_in, p_out, seed=0, degree_corrected=False, pareto_shape=2.5,
             mode="auto", threads=0, keep_truth=False, verbose=False) -> dict:
    """Seeded MMSB graph from the library generator; returns numpy copies of the five CSR blocks
    (rowptr, colidx, vals, n_rows, n_cols) and, optionally, the ground truth."""
    import numpy as np
    p = GenParamsC(n_x, n_a, n_b, n_c, k, alpha0, p_in, p_out, seed, int(degree_corrected), pareto_shape,
                   GEN_MODES[mode], threads, int(verbose))
    g = ctypes.POINTER(GraphC)()
    check(lib.mmsb_generate(ctypes.byref(p), int(keep_truth), ctypes.byref(g)))
    try:
        G = g.contents
        out = {name: _csr_to_numpy(getattr(G, name.lower())) for name in ("XA", "XB", "XC", "BA", "CA")}
        out["n_edges"] = G.n_edges
        if keep_truth:
            n = G.n_nodes
            out["theta"] = np.ctypeslib.as_array(G.theta, shape=(n,)).copy()
            out["part"] = np.ctypeslib.as_array(G.part, shape=(n,)).copy()
            out["local"] = np.ctypeslib.as_array(G.local, shape=(n,)).copy()
            out["pi"] = np.ctypeslib.as_array(G.pi, shape=(n * k,)).copy().reshape(n, k)
        return out
    finally:
        lib.mmsb_graph_free(g)


def config(**kw) -> StpmConfigC:
    """stpm_config with the library defaults, overridden by keyword (strings for the char fields)."""
    c = StpmConfigC()
    lib.stpm_config_default(ctypes.byref(c))
    for key, val in kw.items():
        if key == "precision" and isinstance(val, str):
            val = PREC[val]
        if key == "m2_scaling" and isinstance(val, str):
            val = M2_SCALING[val]
        if key == "y_mode" and isinstance(val, str):
            val = Y_MODE[val]
        if key == "renorm_mode" and isinstance(val, str):
            val = RENORM[val]
        if isinstance(val, str):
            val = val.encode()
        if not hasattr(c, key):
            raise AttributeError(key)
        setattr(c, key, val)
    return c


def _csr_c(rowptr, colidx, vals, n_rows, n_cols):
    import numpy as np
    rp = np.ascontiguousarray(rowptr, np.int32)
    ci = np.ascontiguousarray(colidx, np.int32)
    v = None if vals is None else np.ascontiguousarray(vals, np.float32)
    c = CsrC(n_rows, n_cols, ci.size, rp.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
             ci.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
             None if v is None else v.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return c, (rp, ci, v)


def read_coo(path, n_rows, n_cols, first_is_row=True, weighted=False, duplicates="one", feof_repeat=False):
    dup = {"one": 0, "last": 1, "sum": 2}[duplicates]
    m = CsrC()
    check(lib.mmsb_read_coo(str(path).encode(), n_rows, n_cols, int(first_is_row), int(weighted), dup,
                            int(feof_repeat), ctypes.byref(m)))
    try:
        return _csr_to_numpy(m)
    finally:
        lib.mmsb_csr_free(ctypes.byref(m))


class Pipeline:
    """Owns an mmsb_ctx: the whitening / STGD / post-processing state on one GPU."""

    def __init__(self, cfg: StpmConfigC):
        self.cfg = cfg
        self.k = cfg.ncom
        self.ctx = _vp()
        check(lib.mmsb_create(ctypes.byref(cfg), ctypes.byref(self.ctx)))

    def close(self):
        if self.ctx:
            lib.mmsb_destroy(self.ctx)
            self.ctx = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_block(self, name, rowptr, colidx, vals, n_rows, n_cols):
        c, keep = _csr_c(rowptr, colidx, vals, n_rows, n_cols)
        check(lib.mmsb_set_block(self.ctx, BLOCKS[name], ctypes.byref(c)))

    def set_graph(self, g: dict):
        for name in BLOCKS:
            if name in g and g[name] is not None:
                self.set_block(name, *g[name])

    def read_inputs(self):
        check(lib.mmsb_read_inputs(self.ctx))

    def whiten(self):
        check(lib.mmsb_whiten(self.ctx))

    def set_phi(self, phi):
        import numpy as np
        a = np.ascontiguousarray(phi, np.float64)
        check(lib.mmsb_set_phi(self.ctx, a.ctypes.data))

    def set_yvectors(self, ya, yb, yc, mua, mub, muc):
        import numpy as np
        arrs = [np.ascontiguousarray(x, np.float64) for x in (ya, yb, yc, mua, mub, muc)]
        n, k = arrs[0].shape
        check(lib.mmsb_set_yvectors(self.ctx, n, k, *[a.ctypes.data for a in arrs]))

    def stgd(self, steps=-1, precision="fp32"):
        check(lib.mmsb_stgd(self.ctx, int(steps), PREC[precision] if isinstance(precision, str) else precision))

    def set_learn_rate(self, eta):
        check(lib.mmsb_set_learn_rate(self.ctx, float(eta)))

    def request_stop(self):
        """Ends a running (or the next) stgd() at its next epoch boundary (thread-safe)."""
        check(lib.mmsb_request_stop(self.ctx))

    def stgd_timed(self, steps, precision="fp32"):
        """Runs `steps` direct STGD steps with HIP events around each phase; returns the summed
        device milliseconds of (P1, P2, renormalisation)."""
        import numpy as np
        out = np.zeros(3, np.float64)
        check(lib.mmsb_stgd_timed(self.ctx, int(steps), PREC[precision] if isinstance(precision, str) else precision,
                                  out.ctypes.data))
        return tuple(float(v) for v in out)

    def eigen(self):
        check(lib.mmsb_eigen(self.ctx))

    def post(self):
        check(lib.mmsb_post(self.ctx))

    def write_outputs(self):
        check(lib.mmsb_write_outputs(self.ctx))

    def write_triplets(self, pi_path=None, white_path=None):
        check(lib.mmsb_write_triplets(self.ctx, None if pi_path is None else str(pi_path).encode(),
                                      None if white_path is None else str(white_path).encode()))

    def get(self, name):
        import numpy as np
        q = Q[name]
        n = lib.mmsb_get_size(self.ctx, q)
        if n <= 0:
            raise MMSBError(MMSB_ERR_STATE, f"{name} not available")
        out = np.empty(n, np.float64)
        check(lib.mmsb_get(self.ctx, q, out.ctypes.data, n))
        if name in ("W", "U", "V", "LB", "LC", "YA", "YB", "YC", "phi") or name.startswith("pi"):
            out = out.reshape(-1, self.k)
        return out

    def grad(self, phi, row0, nrows, scaling="gpu", theta=1.0, recompute=False):
        """Loss and adjoint over X rows [row0, row0 + nrows) (through mmsb_grad)."""
        import numpy as np
        a = np.ascontiguousarray(phi, np.float64)
        g = np.empty_like(a)
        loss = np.zeros(1)
        check(lib.mmsb_grad(self.ctx, int(row0), int(nrows), a.ctypes.data, M2_SCALING[scaling], float(theta),
                            int(recompute), g.ctypes.data, loss.ctypes.data))
        return float(loss[0]), g

    def power_method(self, iters, orthogonalize=True):
        import numpy as np
        lam = np.zeros(self.k)
        check(lib.mmsb_power_method(self.ctx, int(iters), int(orthogonalize), lam.ctypes.data))
        return lam

    def get_f32(self, name):
        import numpy as np
        n = lib.mmsb_get_size(self.ctx, Q[name])
        if n <= 0:
            raise MMSBError(MMSB_ERR_STATE, f"{name} not available")
        out = np.empty(n, np.float32)
        check(lib.mmsb_get_f32(self.ctx, Q[name], out.ctypes.data, n))
        return out.reshape(-1, self.k)

    def get_rows(self, name, row0, nrows):
        import numpy as np
        q = Q[name]
        tall = name in ("W", "U", "V", "LB", "LC", "YA", "YB", "YC") or name.startswith("pi")
        cols = self.k if tall else 1
        out = np.empty(nrows * cols, np.float64)
        check(lib.mmsb_get_rows(self.ctx, q, int(row0), int(nrows), out.ctypes.data, out.size))
        return out.reshape(nrows, cols) if tall else out


def tensor_coeffs(alpha0: float, scaling: str = "gpu"):
    import numpy as np
    out = np.zeros(3)
    check(lib.mmsb_tensor_coeffs(float(alpha0), M2_SCALING[scaling], out.ctypes.data))
    return out


class Batch:
    """Independent pipelines advanced together, one persistent launch per epoch (mmsb_batch_*)."""

    def __init__(self, pipelines, precision="fp32"):
        self.members = list(pipelines)
        arr = (ctypes.c_void_p * len(self.members))(*[p.ctx for p in self.members])
        out = ctypes.c_void_p()
        check(lib.mmsb_batch_create(arr, len(self.members), PREC[precision] if isinstance(precision, str) else precision,
                                    ctypes.byref(out)))
        self.h = out

    def run(self, epochs=1):
        check(lib.mmsb_batch_run(self.h, int(epochs)))

    def autotune(self):
        g, ms = ctypes.c_int32(), ctypes.c_double()
        check(lib.mmsb_batch_autotune(self.h, ctypes.byref(g), ctypes.byref(ms)))
        return g.value, ms.value

    def close(self):
        if self.h:
            lib.mmsb_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def stgd_grad(ya, yb, yc, mu_a, mu_b, mu_c, phi, coeffs, theta=1.0, stream=None):
    """Loss and adjoint on device tensors (FP64 or FP32 y, FP64 mu/phi) through mmsb_stgd_grad_*.
    Returns (loss tensor [1], grad tensor [k, k]) on the same device."""
    import numpy as np
    nb, k = ya.shape
    dev = ya.device
    work = torch.empty(int(lib.mmsb_stgd_grad_workspace(k, nb)), dtype=torch.uint8, device=dev)
    grad = torch.empty((k, k), dtype=torch.float64, device=dev)
    loss = torch.empty(1, dtype=torch.float64, device=dev)
    co = np.ascontiguousarray(coeffs, np.float64)
    fn = lib.mmsb_stgd_grad_f64 if ya.dtype == torch.float64 else lib.mmsb_stgd_grad_f32
    check(fn(stream_handle(stream), k, nb, ptr(ya), ptr(yb), ptr(yc), ya.stride(0), ptr(mu_a), ptr(mu_b), ptr(mu_c),
             ptr(phi), co.ctypes.data, float(theta), ptr(grad), ptr(loss), ptr(work)))
    torch.cuda.current_stream().synchronize() if stream is None else stream.synchronize()
    return loss, grad
