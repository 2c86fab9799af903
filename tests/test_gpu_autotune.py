"""The launch-schedule autotuner on the GPU: every radix schedule of the
generic Stockham engine, forced through a tuning file, gives the demag field of
the oracle; tuned plans (including the ensemble grid sizes with a factor 3,
odd periodic Nx and 2D) stay within the convolution tolerances; results are
kept per shape, written to the tuning file and reused by mode 2.

Tolerances as for the convolution parity tests: FP64 1e-12, FP32 2e-6
normwise against the FP64 oracle."""
import sys

import numpy as np
import pytest

from cases import random_unit

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
    umhip.set_autotune(0)


def _t(a, dtype):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype)).cuda()


def _check(gpu, oracle, N, pbc, prec, cs=(2e-9, 3e-9, 2.5e-9)):
    rng = np.random.default_rng(sum(N))
    m = random_unit((N[2], N[1], N[0]), sum(N) + 1)
    Ms = rng.uniform(5e5, 1e6, (N[2], N[1], N[0]))
    Bo = oracle.Demag(N, cs, pbc=pbc, dtype=np.float64).exec(m, Ms)
    dt = np.float32 if prec == 0 else np.float64
    conv = gpu.DemagConvolution(N, cs, pbc=pbc, precision=prec)
    B = [_t(np.zeros_like(m[0]), dt) for _ in range(3)]
    conv.Exec(B, [_t(m[k], dt) for k in range(3)], None, _t(Ms, dt))
    Bg = np.stack([b.cpu().numpy() for b in B]).astype(np.float64)
    err = np.abs(Bg - Bo).max() / np.abs(Bo).max()
    assert err < (1e-12 if prec == 1 else 2e-6), (N, pbc, prec, conv.schedule(), err)
    return conv.schedule()


def _key(oracle, N, pbc, prec):
    P = oracle.pad_size(N, pbc)
    return "%s N=%d,%d,%d P=%d,%d,%d" % (("f32", "f64")[prec], *N, *P)


SHAPES = [((96, 48, 2), (0, 0, 0)), ((48, 96, 3), (0, 0, 0)), ((192, 64, 1), (0, 0, 0)),
          ((40, 24, 6), (0, 0, 0)), ((7, 6, 3), (1, 0, 0)), ((24, 20, 4), (0, 0, 2))]


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("N,pbc", SHAPES)
def test_every_radix_schedule(gpu, oracle, tmp_path, N, pbc, prec):
    _say("schedules", N, pbc, prec)
    gpu.set_autotune(0)
    base = _check(gpu, oracle, N, pbc, prec)
    for sc in range(4):
        f = tmp_path / ("s%d.txt" % sc)
        v = base[:5] + [sc, sc, sc]
        v[2] = 0     # Stockham X inverse, so its schedule is used too
        f.write_text("%s: %s\n" % (_key(oracle, N, pbc, prec), " ".join(map(str, v))))
        gpu.set_autotune(2, str(f))
        got = _check(gpu, oracle, N, pbc, prec)
        assert got == v, (got, v)
    gpu.set_autotune(0)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("N,pbc", SHAPES + [((256, 256, 1), (0, 0, 0)), ((64, 64, 8), (0, 0, 0)),
                                            ((128, 96, 4), (0, 0, 0))])
def test_tuned_plans(gpu, oracle, tmp_path, N, pbc, prec):
    _say("tuned", N, pbc, prec)
    f = tmp_path / "tune.txt"
    gpu.set_autotune(1, str(f))
    try:
        sched = _check(gpu, oracle, N, pbc, prec)
        lines = f.read_text().splitlines()
        assert len(lines) == 1 and lines[0].startswith(_key(oracle, N, pbc, prec) + ":")
        assert [int(x) for x in lines[0].split(":")[1].split()] == sched
        # known shape: reused without a second entry
        assert _check(gpu, oracle, N, pbc, prec) == sched
        assert len(f.read_text().splitlines()) == 1
    finally:
        gpu.set_autotune(0)
