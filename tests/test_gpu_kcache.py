"""The demag tensor cache on the GPU (mag.DemagKernel / NewellDemagKernel with a
cache directory): a cold plan writes the reference's files, holding the
float32 rounding of the GPU tensor; a warm plan reads them and gives the same
field bit for bit; an unreadable file is recomputed and rewritten; the
simulation handle uses the same cache."""
import glob
import os
import sys

import numpy as np
import pytest

from cases import random_unit, setup

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu():
    import torch
    from umagnus_amd import umhip
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    umhip.init(0)
    yield umhip
    umhip.set_kernel_cache(None)


def _t(a, dtype):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype)).cuda()


def _field(gpu, N, cs, m, Ms, prec, **kw):
    dt = np.float32 if prec == 0 else np.float64
    conv = gpu.DemagConvolution(N, cs, precision=prec, **kw)
    B = [_t(np.zeros_like(m[0]), dt) for _ in range(3)]
    conv.Exec(B, [_t(m[k], dt) for k in range(3)], None, _t(Ms, dt))
    return np.stack([b.cpu().numpy() for b in B])


IJ = {"00": 0, "01": 5, "02": 4, "11": 1, "12": 3, "22": 2}


@pytest.mark.parametrize("N,pbc,newell", [((12, 10, 3), (0, 0, 0), False), ((16, 8, 1), (0, 2, 0), False),
                                          ((10, 8, 4), (0, 0, 0), True)])
def test_cache_cold_warm(gpu, tmp_path, N, pbc, newell):
    _say("kcache", N, pbc, newell)
    cs = (2e-9, 3e-9, 2.5e-9)
    m = random_unit((N[2], N[1], N[0]), 3)
    Ms = np.full((N[2], N[1], N[0]), 8e5)
    kw = dict(pbc=pbc, newell=newell)
    gpu.set_kernel_cache(str(tmp_path))
    try:
        cold = _field(gpu, N, cs, m, Ms, 1, **kw)
        base = gpu.kernel_cache_basename(str(tmp_path), N, cs, pbc=pbc, newell=newell)
        files = sorted(glob.glob(glob.escape(base) + "*.ovf"))
        want = ["00", "01", "11", "22"] if N[2] == 1 else ["00", "01", "02", "11", "12", "22"]
        assert [f[len(base):-4] for f in files] == want
        K = (gpu.demag_kernel_newell(N, cs, pbc=pbc) if newell else gpu.demag_kernel_bf(N, cs, pbc=pbc))
        P = K[0].shape[::-1]
        for ij in want:
            got = gpu.load_kernel(base + ij + ".ovf", P)
            assert np.array_equal(got, K[IJ[ij]].astype(np.float32).astype(np.float64)), ij
        warm = _field(gpu, N, cs, m, Ms, 1, **kw)
        assert np.array_equal(cold, warm)
        # a damaged entry is recomputed and rewritten
        with open(files[0], "r+b") as f:
            f.write(b"# OOMMF OVF 9.9")
        again = _field(gpu, N, cs, m, Ms, 1, **kw)
        assert np.array_equal(cold, again)
        assert gpu.load_kernel(files[0], P).shape == K[0].shape
    finally:
        gpu.set_kernel_cache(None)
    # without the cache: the FP64 tensor, within float32 rounding of the cached one
    plain = _field(gpu, N, cs, m, Ms, 1, **kw)
    assert np.abs(plain - cold).max() <= 1e-6 * np.abs(plain).max()


def test_sim_uses_cache(gpu, tmp_path):
    N, cs = (16, 12, 2), (2e-9, 2e-9, 1.5e-9)
    m = random_unit((N[2], N[1], N[0]), 4)
    gpu.set_kernel_cache(str(tmp_path))
    try:
        conv = _field(gpu, N, cs, m, np.full((N[2], N[1], N[0]), 8e5), 0)
        s = gpu.Sim(N, cs, dtype=np.float32)
        setup(s, gpu, Ms=8e5, Aex=1.3e-11, alpha=0.1, dtype=np.float32)
        s.set_m(m)
        Bs = s.demag_field()
        assert np.abs(Bs - conv).max() <= 1e-6 * np.abs(conv).max()
        assert len(os.listdir(tmp_path)) == 6
    finally:
        gpu.set_kernel_cache(None)
