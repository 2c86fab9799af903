"""The C-ABI library loads (no GPU needed) and exports every function that
include/umhip.h declares; host-side tables match the oracle."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "umhip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(um_[a-z0-9_]+)\s*\(", src)))


def _lib():
    from umagnus_amd import build, umhip
    if not os.path.exists(umhip.LIB_PATH):
        build.build()
    return umhip.load()


def test_header_and_binding_agree():
    from umagnus_amd import umhip
    assert _declared() == sorted(umhip.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = _lib()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_calls_without_gpu_fail_loudly():
    """No CPU fallback: with no device the library reports an error."""
    import torch
    if torch.cuda.is_available():
        return
    lib = _lib()
    lib.um_last_error.restype = ctypes.c_char_p
    rc = lib.um_init(0)
    assert rc != 0 and lib.um_last_error()


def test_exchange_lut_matches_oracle(oracle):
    from umagnus_amd.params import exchange_lut
    vals = {0: 1.3e-11, 1: 2e-11, 2: -1e-11, 7: 5e-12}
    got = exchange_lut(vals, dtype=np.float64)
    want = oracle.exchange_lut(vals)
    assert np.array_equal(got, want)
