"""CPU checks of the C-ABI boundary: the in-tree library loads and exports every function
declared in include/*.h, and the Python binding declares a signature for each of them."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        text = re.sub(r"typedef[^;]*;", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_]\w*[\s\*]+\**\s*([A-Za-z_]\w*)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_parses():
    names = declared_functions()
    assert "mmsb_spmm_f32" in names and "mmsb_last_error" in names


def test_library_exports_every_declared_symbol(mmsb):
    lib = ctypes.CDLL(mmsb.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_binding_covers_header(mmsb):
    assert declared_functions() <= set(mmsb.SIGNATURES), declared_functions() - set(mmsb.SIGNATURES)


def test_version_and_error_strings(mmsb):
    assert mmsb.lib.mmsb_version().decode().startswith("mmsb-amd")
    assert isinstance(mmsb.last_error(), str)


def test_argument_errors_are_status_codes_not_exits(mmsb):
    # bad leading dimension: rejected on the host before any launch, no GPU needed
    st = mmsb.lib.mmsb_spmm_f32(None, 4, 4, 0, None, None, None, 1.0, None, 3, 3, None, 3)
    assert st == mmsb.MMSB_ERR_ARG
    assert "multiples of 4" in mmsb.last_error()


def test_host_checks_of_stop_and_y_mode(mmsb):
    """Host-side validation of the stop request and the neighbour-vector mode (no GPU call)."""
    import ctypes
    assert mmsb.lib.mmsb_request_stop(None) == mmsb.MMSB_ERR_ARG
    mmsb.lib.stpm_force_exit()                      # no program-level context: nothing to stop
    assert mmsb.config(y_mode="recompute").y_mode == 2 and mmsb.config().y_mode == 0
    out = ctypes.c_void_p()
    cfg = mmsb.config(nx=100, na=100, nb=100, nc=100, ncom=4, y_mode=7)
    assert mmsb.lib.mmsb_create(ctypes.byref(cfg), ctypes.byref(out)) == mmsb.MMSB_ERR_ARG
    assert "y_mode" in mmsb.last_error()


def test_cli_argument_errors(mmsb):
    """./algo rejects a missing or unknown option and an unknown KEY=VALUE with exit status 1 and the
    usage text, before touching a GPU (the reference's main prints its choices, stpm.c:1068-1118)."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(mmsb.__file__), "algo")
    for args in ([], ["GPU_C"], ["GPU_A", "NOPE=1"], ["GPU_A", "NX"]):
        r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=60)
        assert r.returncode == 1, args
        assert "usage: ./algo" in r.stdout, args


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 ends with a message and status 2, before any GPU use."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 2
    assert "one GPU" in (r.stdout + r.stderr)


def test_release_and_strip_limit_without_a_gpu(mmsb):
    """mmsb_release with nothing allocated touches no device and may be called repeatedly; the strip
    limit setter returns the previous value and restores the 8 GiB default for a negative argument."""
    mmsb.lib.mmsb_release()
    mmsb.lib.mmsb_release()
    prev = mmsb.lib.mmsb_set_strip_limit(12345)
    assert mmsb.lib.mmsb_set_strip_limit(-1) == 12345
    assert mmsb.lib.mmsb_set_strip_limit(prev) == 8 << 30


def test_renorm_mode_field_round_trips(mmsb):
    assert mmsb.config().renorm_mode == 0 and mmsb.config(renorm_mode="qr").renorm_mode == 1


def test_csr_host_field_defaults_to_resident_blocks(mmsb):
    """stpm_config.csr_host is the last field of the struct (the binding's layout must match the
    header's size) and defaults to 0 (blocks resident in HBM)."""
    import ctypes
    cfg = mmsb.config()
    assert cfg.csr_host == 0 and mmsb.config(csr_host=1).csr_host == 1
    assert type(cfg)._fields_[-1][0] == "csr_host"
    assert ctypes.sizeof(cfg) == mmsb.lib.stpm_config_size()
