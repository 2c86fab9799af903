"""um_rk45dp_error_noalloc against um_rk45dp_error through the C ABI.

The allocation-free variant keeps the error kernel's max slots in the
library's per-GPU device array and writes (err, rlerr, maxdm) to a caller
device buffer.  Two calls in a row, the second on a smaller input, must give
exactly what um_rk45dp_error returns for the same stages: the second call
proves the shared slots are cleared between calls.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from cases import GAMMA

pytestmark = pytest.mark.gpu

E_DP = [35 / 384 - 5179 / 57600, 500 / 1113 - 7571 / 16695, 125 / 192 - 393 / 640,
        -2187 / 6784 + 92097 / 339200, 11 / 84 - 187 / 2100, -1 / 40]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu():
    import torch
    from umagnus_amd import umhip
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    umhip.init(0)
    return umhip


def _stages(shape, dtype, seed, nk):
    import torch
    rng = np.random.default_rng(seed)
    ks = [rng.normal(0, 1.0, (3,) + shape).astype(dtype) for _ in range(nk)]
    ks[-1][:, 0, 0, :3] = 0                      # |delM| = 0 cells
    return [[torch.as_tensor(np.ascontiguousarray(k[c])).cuda() for c in range(3)] for k in ks]


def _noalloc(gpu, ks, e, h, out):
    import torch
    flat = (C.c_void_p * (3 * len(ks)))(*[t.data_ptr() for v in ks for t in v])
    rc = gpu.load().um_rk45dp_error_noalloc(gpu._prec(ks[0][0]), len(ks), flat, (C.c_double * len(e))(*e),
                                            C.c_double(h), ks[0][0].numel(), C.c_void_p(out.data_ptr()))
    gpu._check(rc)
    torch.cuda.synchronize()
    return tuple(out.cpu().numpy().tolist())


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_matches_rk45dp_error(gpu, prec):
    import torch
    dtype = np.float32 if prec == 0 else np.float64
    h = float(np.float32(1e-14 * GAMMA))
    out = torch.full((3,), float("nan"), dtype=torch.float64, device="cuda")
    # the second call is smaller: stale maxima from the first would show
    for shape, seed, scale in (((4, 24, 40), 1, 3.0), ((2, 10, 12), 2, 1.0)):
        _say(f"noalloc prec={prec} shape={shape}")
        ks = _stages(shape, dtype, seed, len(E_DP))
        for v in ks:
            for t in v:
                t.mul_(scale)
        want = gpu.RK45DPError(ks, E_DP, h)
        got = _noalloc(gpu, ks, E_DP, h, out)
        assert all(np.isfinite(got)), got
        np.testing.assert_array_equal(np.array(got), np.array(want))


def test_noalloc_rejects_bad_arguments(gpu):
    import torch
    ks = _stages((1, 2, 4), np.float32, 3, 6)
    flat = (C.c_void_p * 18)(*[t.data_ptr() for v in ks for t in v])
    e = (C.c_double * 6)(*E_DP)
    lib = gpu.load()
    assert lib.um_rk45dp_error_noalloc(0, 6, flat, e, C.c_double(1.0), 8, None) != 0
    out = torch.zeros(3, dtype=torch.float64, device="cuda")
    assert lib.um_rk45dp_error_noalloc(0, 7, flat, e, C.c_double(1.0), 8, C.c_void_p(out.data_ptr())) != 0
    # the library still works afterwards
    got = _noalloc(gpu, ks, E_DP, 1.0, out)
    assert np.array_equal(np.array(got), np.array(gpu.RK45DPError(ks, E_DP, 1.0)))
