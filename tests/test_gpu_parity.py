"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Tolerances (written per test): FP64 GPU vs FP64 oracle to ~1e-12 relative;
FP32 GPU vs FP64 oracle to 1e-5 relative for fields (normwise: max|diff| /
max|field|); integer work (Philox) bit-exact.
"""
import math
import sys

import numpy as np
import pytest

from cases import GAMMA, MU0, strip_m, random_unit, relerr, setup

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
    _say("device:", umhip.device_info())
    return umhip


def _t(a, dtype):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype)).cuda()


def _h(t):
    return t.detach().cpu().numpy()


# ---- demag tensor (FP64 on the GPU) --------------------------------------------------------
# The offsets next to the source integrate thousands of points (31104 at
# accuracy 6, 1e6 at 16) whose terms cancel to the tensor element; the GPU sums
# them as wavefront trees, the oracle sequentially as the reference does, whose
# own rounding reaches ~n eps of the terms: 3e-13 of the largest element at
# accuracy 6.  The reference stores the tensor in float32 (6e-8).
TENSOR_TOL = 2e-12
@pytest.mark.parametrize("N,cs", [((3, 4, 1), (1e-9, 2e-9, 0.5e-9)), ((3, 4, 2), (1e-9, 2e-9, 3e-9)),
                                  ((6, 5, 3), (2e-9, 3e-9, 1.5e-9)), ((16, 8, 7), (2e-9, 2e-9, 2e-9))])
def test_demag_tensor_vs_oracle(gpu, oracle, N, cs):
    _say("tensor", N)
    Kg = gpu.demag_kernel_bf(N, cs)
    Ko = oracle.demag_kernel_bf(N, cs)
    for c in range(6):
        scale = max(np.abs(Ko[c]).max(), 1e-300)
        assert np.abs(Kg[c] - Ko[c]).max() <= TENSOR_TOL * scale, c


@pytest.mark.parametrize("N,cs,pbc,acc", [((5, 4, 1), (2.5e-9, 2.5e-9, 5e-9), (0, 0, 0), 16),
                                          ((4, 3, 2), (2e-9, 3e-9, 2.5e-9), (2, 1, 0), 10),
                                          ((9, 7, 3), (1e-9, 1.2e-9, 0.8e-9), (0, 0, 0), 7),
                                          ((6, 6, 2), (1e-9, 1e-9, 1e-9), (0, 0, 0), 1)])
def test_demag_tensor_accuracy_vs_oracle(gpu, oracle, N, cs, pbc, acc):
    """DemagAccuracy away from 6: the offsets next to the source, whose
    integration grid grows as accuracy^5, are split over whole wavefronts (and
    at accuracy 1 none is); with PBC two near images can wrap onto one element.
    The points per offset grow as accuracy^5 and the rounding of a sequential
    sum as their square root, so the bound scales as (accuracy / 6)^2.5."""
    _say("tensor accuracy", N, pbc, acc)
    Kg = gpu.demag_kernel_bf(N, cs, pbc=pbc, accuracy=acc)
    Ko = oracle.demag_kernel_bf(N, cs, pbc=pbc, accuracy=acc)
    tol = TENSOR_TOL * max(1.0, (acc / 6) ** 2.5)
    for c in range(6):
        scale = max(np.abs(Ko[c]).max(), 1e-300)
        assert np.abs(Kg[c] - Ko[c]).max() <= tol * scale, c


# ---- demag convolution ---------------------------------------------------------------
DEMAG_GOLDEN = [
    ((3, 4, 1), (1e-9, 2e-9, 0.5e-9), (-0.15768295526, -0.05676037073, -0.78447151184), 1e-6),
    ((3, 4, 2), (1e-9, 2e-9, 3e-9), (-0.52710835138, -0.20121671756, -0.27112555503), 1e-6),
    ((2, 2, 2), (1e-9, 1e-9, 1e-9), (-1 / 3, -1 / 3, -1 / 3), 5e-4),
    ((5, 5, 5), (1e-9, 1e-9, 1e-9), (-1 / 3, -1 / 3, -1 / 3), 5e-4),
]


@pytest.mark.parametrize("N,cs,expect,tol", DEMAG_GOLDEN)
@pytest.mark.parametrize("prec", [0, 1])
def test_demag_golden(gpu, N, cs, expect, tol, prec):
    """test/demagSmall.mx3, demagSmall3D.mx3, demagcube.mx3, demagodd.mx3."""
    _say("demag golden", N, prec)
    dt = np.float32 if prec == 0 else np.float64
    conv = gpu.DemagConvolution(N, cs, precision=prec)
    for c in range(3):
        m = np.zeros((3, N[2], N[1], N[0]))
        m[c] = 1
        mt = [_t(m[k], dt) for k in range(3)]
        B = [_t(np.zeros_like(m[0]), dt) for _ in range(3)]
        conv.Exec(B, mt, None, 1 / MU0)
        avg = [float(_h(B[k]).mean()) for k in range(3)]
        for k in range(3):
            want = expect[c] if k == c else 0.0
            assert abs(avg[k] - want) < tol, (c, k, avg[k], want)


CONV_SIZES = [(3, 4, 1), (16, 8, 1), (20, 12, 1), (5, 4, 2), (6, 5, 3), (8, 6, 4), (7, 3, 5), (16, 12, 6),
              (24, 10, 7), (48, 20, 3), (96, 12, 2), (64, 64, 1), (32, 24, 8),
              # register Z pass (Pz = 2 Nz, Nz in 8/16/32): partial kx groups, unpaired ky rows
              (70, 5, 8), (40, 9, 16), (12, 8, 32),
              # four-step Y pass (Py = 64, 256, 1024), with even Nz (two planes per work-group) and odd Nz
              (12, 32, 4), (10, 128, 3), (6, 512, 2), (9, 512, 1),
              # four-step X passes (Nx = 128 .. 1024)
              (128, 6, 3), (256, 5, 2), (512, 3, 2), (1024, 2, 1),
              # small grids: X rows and 2D conv columns split over more work-groups
              (128, 32, 1), (256, 256, 1), (512, 64, 1), (1024, 8, 1), (32, 256, 1), (256, 64, 8),
              (256, 256, 8), (128, 128, 16), (64, 64, 64), (1536, 128, 1), (3072, 32, 1), (3072, 64, 1),
              # ensemble sizes: X lengths 32/48/96/192 (full and small-grid rows), Y 96/192/384, Pz 3/5/7
              (32, 48, 2), (48, 96, 3), (96, 192, 4), (192, 32, 2), (48, 48, 1), (96, 96, 1), (192, 192, 1),
              (32, 192, 1), (192, 256, 4), (256, 96, 3), (96, 32, 4), (48, 192, 2)]


@pytest.mark.parametrize("N", CONV_SIZES)
@pytest.mark.parametrize("prec", [1, 0])
def test_demag_conv_vs_oracle(gpu, oracle, N, prec):
    """B_demag of a random m with per-cell Ms and geometry vs the FP64 oracle."""
    _say("demag conv", N, prec)
    cs = (2e-9, 3e-9, 1.5e-9)
    rng = np.random.default_rng(sum(N) + 7 * prec)
    m = random_unit((N[2], N[1], N[0]), sum(N))
    Ms = rng.uniform(5e5, 1e6, (N[2], N[1], N[0]))
    vol = rng.uniform(0.2, 1.0, (N[2], N[1], N[0]))
    ref = oracle.Demag(N, cs, dtype=np.float64).exec(m, Ms, vol)
    dt = np.float32 if prec == 0 else np.float64
    conv = gpu.DemagConvolution(N, cs, precision=prec)
    B = [_t(np.zeros_like(m[0]), dt) for _ in range(3)]
    conv.Exec(B, [_t(m[k], dt) for k in range(3)], _t(vol, dt), _t(Ms, dt))
    got = np.stack([_h(b) for b in B])
    tol = 1e-12 if prec == 1 else 2e-6
    assert relerr(got, ref) <= tol, relerr(got, ref)


@pytest.mark.parametrize("N,prec,tol", [((12, 10, 3), 1, 1e-12),
                                         # Nz = 8 (Pz = 16): the register Z pass, whose plan stores the
                                         # tensor spectrum ky-major and reads it back in [kz][ky][kx] order
                                         ((12, 10, 8), 1, 1e-12), ((40, 10, 8), 0, 1e-6)])
def test_demag_kspec_vs_oracle(gpu, oracle, N, prec, tol):
    cs = (2e-9, 3e-9, 1.5e-9)
    conv = gpu.DemagConvolution(N, cs, precision=prec)
    d = oracle.Demag(N, cs, dtype=np.float64)
    for c in range(6):
        got = conv.kspec(c).astype(np.float64)
        want = np.zeros_like(got)
        assert oracle.fn("demag_kspec", np.float64)(__import__("ctypes").c_void_p(d.h), c, oracle.ptr(want)) == 0
        assert relerr(got, want) < tol, (c, relerr(got, want))


# ---- unfused field terms vs oracle ----------------------------------------------------------
@pytest.mark.parametrize("prec", [1, 0])
def test_field_terms_vs_oracle(gpu, oracle, prec):
    _say("field terms", prec)
    import ctypes as C
    dt = np.float32 if prec == 0 else np.float64
    tol = 1e-13 if prec == 1 else 2e-6
    N = (13, 9, 4)
    shape = (N[2], N[1], N[0])
    n = N[0] * N[1] * N[2]
    cs = (2e-9, 2.5e-9, 3e-9)
    rng = np.random.default_rng(11)
    m = random_unit(shape, 3).astype(dt)
    m[:, 1, 2, 3] = 0  # a hole: missing neighbour handling
    regions = rng.integers(0, 3, shape).astype(np.uint8)
    Ms = rng.uniform(5e5, 1e6, shape).astype(dt)
    from umagnus_amd.params import exchange_lut
    aLUT = exchange_lut({0: 1.3e-11, 1: 2e-11, 2: 1e-11}, dtype=dt)
    dLUT = exchange_lut({0: 3e-3, 1: -1e-3, 2: 2e-3}, dtype=dt)
    R = oracle._rt(dt)
    p = oracle.ptr
    for pbc in (0, 1 | 2, 4):
        for openbc in (0, 1):
            Bo = np.zeros((3,) + shape, dt)
            oracle.fn("adddmi", dt)(p(Bo[0]), p(Bo[1]), p(Bo[2]), p(m[0]), p(m[1]), p(m[2]), p(Ms), R(1.0),
                                    p(aLUT), p(dLUT), p(regions), R(cs[0]), R(cs[1]), R(cs[2]), N[0], N[1], N[2],
                                    pbc, openbc)
            Bg = [_t(np.zeros(shape), dt) for _ in range(3)]
            gpu.AddDMI(Bg, [_t(m[k], dt) for k in range(3)], _t(aLUT, dt), _t(dLUT, dt), _t(Ms, dt),
                       _t(regions, np.uint8), cs, pbc=pbc, OpenBC=bool(openbc))
            assert relerr(np.stack([_h(b) for b in Bg]), Bo) <= tol, ("dmi", pbc, openbc)
        Bo = np.zeros((3,) + shape, dt)
        w = [dt(2 / c**2) for c in cs]
        oracle.fn("addexchange", dt)(p(Bo[0]), p(Bo[1]), p(Bo[2]), p(m[0]), p(m[1]), p(m[2]), p(Ms), R(1.0),
                                     p(aLUT), p(regions), R(w[0]), R(w[1]), R(w[2]), N[0], N[1], N[2], pbc)
        Bg = [_t(np.zeros(shape), dt) for _ in range(3)]
        gpu.AddExchange(Bg, [_t(m[k], dt) for k in range(3)], _t(aLUT, dt), _t(Ms, dt), _t(regions, np.uint8), cs,
                        pbc=pbc)
        assert relerr(np.stack([_h(b) for b in Bg]), Bo) <= tol, ("exchange", pbc)
    # anisotropies: per-cell K1, uniform K2 / axes
    K1 = rng.uniform(1e5, 8e5, shape).astype(dt)
    Bo = np.zeros((3,) + shape, dt)
    oracle.fn("adduniaxial2", dt)(p(Bo[0]), p(Bo[1]), p(Bo[2]), p(m[0]), p(m[1]), p(m[2]), p(Ms), R(1.0), p(K1),
                                  R(1.0), None, R(-2e5), None, R(0.3), None, R(-0.2), None, R(0.9), n)
    Bg = [_t(np.zeros(shape), dt) for _ in range(3)]
    gpu.AddUniaxialAnisotropy2(Bg, [_t(m[k], dt) for k in range(3)], _t(Ms, dt), _t(K1, dt), -2e5, (0.3, -0.2, 0.9))
    assert relerr(np.stack([_h(b) for b in Bg]), Bo) <= tol, "uniaxial"
    Bo = np.zeros((3,) + shape, dt)
    oracle.fn("addcubic2", dt)(p(Bo[0]), p(Bo[1]), p(Bo[2]), p(m[0]), p(m[1]), p(m[2]), p(Ms), R(1.0), None,
                               R(4e5), None, R(-1e5), None, R(2e4), None, R(1.0), None, R(1.0), None, R(0.0), None,
                               R(-1.0), None, R(1.0), None, R(0.0), n)
    Bg = [_t(np.zeros(shape), dt) for _ in range(3)]
    gpu.AddCubicAnisotropy2(Bg, [_t(m[k], dt) for k in range(3)], _t(Ms, dt), 4e5, -1e5, 2e4, (1.0, 1.0, 0.0),
                            (-1.0, 1.0, 0.0))
    assert relerr(np.stack([_h(b) for b in Bg]), Bo) <= tol, "cubic"
    # LL torque, no-precession torque
    H = rng.standard_normal((3,) + shape).astype(dt)
    alpha = rng.uniform(0, 1, shape).astype(dt)
    To = np.zeros((3,) + shape, dt)
    oracle.fn("lltorque2", dt)(p(To[0]), p(To[1]), p(To[2]), p(m[0]), p(m[1]), p(m[2]), p(H[0]), p(H[1]),
                               p(H[2]), p(alpha), R(1.0), n)
    Tg = [_t(np.zeros(shape), dt) for _ in range(3)]
    gpu.LLTorque(Tg, [_t(m[k], dt) for k in range(3)], [_t(H[k], dt) for k in range(3)], _t(alpha, dt))
    assert relerr(np.stack([_h(b) for b in Tg]), To) <= tol, "lltorque"
    To = np.zeros((3,) + shape, dt)
    oracle.fn("llnoprecess", dt)(p(To[0]), p(To[1]), p(To[2]), p(m[0]), p(m[1]), p(m[2]), p(H[0]), p(H[1]),
                                 p(H[2]), n)
    gpu.LLNoPrecess(Tg, [_t(m[k], dt) for k in range(3)], [_t(H[k], dt) for k in range(3)])
    assert relerr(np.stack([_h(b) for b in Tg]), To) <= tol, "llnoprecess"
    # region LUT add and zero mask
    lut = [rng.standard_normal(256).astype(dt) for _ in range(3)]
    Do = H.copy()
    oracle.fn("regionaddv", dt)(p(Do[0]), p(Do[1]), p(Do[2]), p(lut[0]), p(lut[1]), p(lut[2]), p(regions), n)
    Dg = [_t(H[k], dt) for k in range(3)]
    gpu.RegionAddV(Dg, [_t(l, dt) for l in lut], _t(regions, np.uint8))
    assert np.array_equal(np.stack([_h(b) for b in Dg]), Do), "regionaddv"
    mask = np.zeros(256, dt)
    mask[1] = 1
    zo = H[0].copy()
    oracle.fn("zeromask", dt)(p(zo), p(mask), p(regions), n)
    zg = _t(H[0], dt)
    gpu.ZeroMask(zg, _t(mask, dt), _t(regions, np.uint8))
    assert np.array_equal(_h(zg), zo), "zeromask"


@pytest.mark.parametrize("prec", [1, 0])
def test_vector_ops_and_reductions(gpu, oracle, prec):
    """Madd, Normalize, VecNorm, Div and opencl/reduce_test.go:46-177 known answers."""
    import torch
    dt = np.float32 if prec == 0 else np.float64
    rng = np.random.default_rng(5)
    n = 1000
    a, b, c = [rng.standard_normal(n).astype(dt) for _ in range(3)]
    d = _t(np.zeros(n), dt)
    gpu.Madd(d, [_t(a, dt), _t(b, dt), _t(c, dt)], [1.0, 0.25, -3.0])
    want = (dt(1.0) * a) + (dt(0.25) * b + dt(-3.0) * c)
    assert np.abs(_h(d) - want).max() <= 4 * np.finfo(dt).eps * np.abs(want).max()
    v = [_t(x, dt) for x in (a, b, c)]
    gpu.Normalize(v)
    nrm = np.sqrt(sum(_h(x) ** 2 for x in v))
    assert np.abs(nrm - 1).max() < 8 * np.finfo(dt).eps
    # known answers
    in1 = _t(np.arange(1000), dt)
    assert gpu.Sum(in1) == 499500
    assert gpu.Dot(_t([1, 2, 3, 4, 5], dt), _t([5, 4, 3, -1, 2], dt)) == 5 + 8 + 9 - 4 + 10
    assert gpu.MaxAbs(in1) == 999
    assert gpu.MaxAbs(_t(-np.arange(100000) / 100, dt)) == float(dt(999.99))
    N = 32
    A = [_t(np.full(N, x), dt) for x in (1, 2, 3)]
    B = [_t(np.full(N, x), dt) for x in (4, 5, 6)]
    assert gpu.MaxVecDiff(A, B) == math.sqrt(27)
    B[0][5], B[1][5], B[2][5] = 6, 7, 9
    assert gpu.MaxVecDiff(A, B) == math.sqrt(25 + 25 + 36)
    assert abs(gpu.MaxVecDiff(A, B, sqrt3_quirk=True) - math.sqrt(3 * 86)) < 1e-12
    assert gpu.MaxVecNorm(A) == math.sqrt(14)
    A[0][5], A[1][5], A[2][5] = 6, 7, 8
    assert gpu.MaxVecNorm(A) == math.sqrt(36 + 49 + 64)
    nan = _t(np.ones(N), dt)
    nan[7] = float("nan")
    assert math.isnan(gpu.MaxVecNorm([nan, _t(np.ones(N), dt), _t(np.ones(N), dt)]))
    torch.cuda.synchronize()


@pytest.mark.parametrize("prec", [1, 0])
def test_thermal_field_vs_oracle(gpu, oracle, prec):
    """Philox normals in the kernel equal the oracle's; B = eta sqrt(k alpha T / Ms)."""
    dt = np.float32 if prec == 0 else np.float64
    n = 4096
    seed, step, stream = 0x1234ABCD5678, 17, 3
    k = 2 * 1.380650424e-23 / (GAMMA * 8e-27 * 1e-14)
    Bg = [_t(np.zeros(n), dt) for _ in range(3)]
    gpu.ThermalField(Bg, seed, step, stream, k, 8e5, 300.0, 0.02)
    eta = oracle.thermal_normals(n, seed, step, stream, dtype=np.float64)
    want = eta * math.sqrt(k * 0.02 * 300.0 / 8e5)
    got = np.stack([_h(b) for b in Bg]).astype(np.float64)
    tol = 1e-12 if prec == 1 else 2e-6
    assert np.abs(got - want).max() <= tol * np.abs(want).max()


# ---- simulation level: effective field and torque ---------------------------------------------
def _both(oracle, gpu, N, cs, prec, **params):
    dt = np.float32 if prec == 0 else np.float64
    so = oracle.Sim(N, cs, dtype=np.float64)
    sg = gpu.Sim(N, cs, dtype=dt)
    setup(so, oracle, dtype=np.float64, **params)
    setup(sg, gpu, dtype=dt, **params)
    return so, sg


@pytest.mark.parametrize("prec", [1, 0])
def test_sim_beff_torque_film_physics(gpu, oracle, prec):
    """The 512x512x32 workload's physics (DMI, Ku1 || z, alpha 0.1, random m) on a small 3D grid."""
    N, cs = (32, 24, 6), (2e-9, 2e-9, 2e-9)
    so, sg = _both(oracle, gpu, N, cs, prec, Ms=1.1e6, Aex=1.6e-11, Dind=3e-3, Ku1=8e5, alpha=0.1)
    m = random_unit((N[2], N[1], N[0]), 21)
    so.set_m(m)
    sg.set_m(m)
    tol = 1e-12 if prec == 1 else 1e-5
    b_o, b_g = so.beff(), sg.beff()
    assert relerr(b_g, b_o) <= tol, relerr(b_g, b_o)
    t_o, t_g = so.torque(), sg.torque()
    assert relerr(t_g, t_o) <= tol, relerr(t_g, t_o)


@pytest.mark.parametrize("prec", [1, 0])
def test_sim_beff_regions_cubic_thermal(gpu, oracle, prec):
    """Regions (harmonic-mean exchange), cubic anisotropy, region B_ext, frozen spins, T > 0."""
    N, cs = (20, 16, 3), (2e-9, 2.5e-9, 3e-9)
    rng = np.random.default_rng(8)
    regions = rng.integers(0, 4, (N[2], N[1], N[0])).astype(np.uint8)
    Ms = rng.uniform(6e5, 1e6, (N[2], N[1], N[0]))
    lut = [np.zeros(256), np.zeros(256), np.zeros(256)]
    lut[0][1], lut[2][2] = 0.05, -0.02
    frozen = np.zeros(256)
    frozen[3] = 1
    so, sg = _both(oracle, gpu, N, cs, prec, Ms=Ms, Aex={0: 1e-11, 1: 2e-11, 2: 1.5e-11, 3: 1.2e-11},
                   Kc=(4e5, -1e5, 0.0), c1=(1, 1, 0), c2=(-1, 1, 0), alpha=0.05, regions=regions,
                   bext_lut=lut, frozen=frozen, temp=300.0, seed=99)
    for s in (so, sg):
        s.set_solver_params(dt=1e-14)
    m = random_unit((N[2], N[1], N[0]), 5)
    so.set_m(m)
    sg.set_m(m)
    tol = 1e-11 if prec == 1 else 1e-5
    assert relerr(sg.beff(), so.beff()) <= tol
    assert relerr(sg.torque(), so.torque()) <= tol


@pytest.mark.parametrize("prec", [1, 0])
@pytest.mark.parametrize("N,openbc", [((24, 12, 4), 1), ((16, 8, 1), 0), ((12, 20, 3), 0), ((8, 6, 2), 1),
                                      ((64, 12, 3), 1), ((128, 10, 1), 0)])
def test_sim_beff_dmi_regions_vacuum(gpu, oracle, prec, N, openbc):
    """Exchange + interfacial DMI with a region map (per-pair A and D), vacuum cells
    (m = 0: the free-boundary branches of dmi.cl) and open / closed edges.  Nx = 64
    and 128 in FP32 take the torque through the fused X-inverse + field kernel
    (with a partly filled last row block)."""
    _say("dmi regions", N, openbc, prec)
    cs = (2e-9, 2.5e-9, 1.5e-9)
    rng = np.random.default_rng(N[0] * 7 + openbc)
    shape = (N[2], N[1], N[0])
    regions = rng.integers(0, 3, shape).astype(np.uint8)
    so, sg = _both(oracle, gpu, N, cs, prec, Ms=9e5, Aex={0: 1e-11, 1: 2e-11, 2: 1.5e-11},
                   Dind={0: 2e-3, 1: -1e-3, 2: 3e-3}, Ku1=5e5, alpha=0.1, regions=regions, openbc=openbc)
    m = random_unit(shape, N[1])
    m[:, rng.random(shape) < 0.15] = 0.0
    so.set_m(m)
    sg.set_m(m)
    tol = 1e-12 if prec == 1 else 1e-5
    assert relerr(sg.beff(), so.beff()) <= tol
    assert relerr(sg.torque(), so.torque()) <= tol


FUSED_CASES = [((64, 12, 3), "dmi_regions"), ((128, 10, 1), "exchange"), ((256, 9, 2), "dmi"),
               ((512, 6, 4), "dmi"), ((512, 7, 3), "none_regions"), ((1024, 5, 2), "exchange_regions"),
               ((256, 8, 5), "dmi_pbc")]


@pytest.mark.parametrize("N,terms", FUSED_CASES)
def test_fused_xinv_field_vs_oracle(gpu, oracle, N, terms):
    """The fused X-inverse + field + torque kernel (k_xi4_field) at every row
    length it covers (Nx = 64 ... 1024), with and without a region map, with
    no exchange, exchange only and exchange + interfacial DMI, partly filled
    row blocks (Ny not a multiple of the rows per work-group), 2D and periodic
    grids: the launch trace shows the fused kernel and no separate field
    kernel; the torque is within 1e-6 (normwise) of the two-kernel path
    (k_xi4 + k_field_torque4, UMHIP_FUSED_FIELD=0: the same arithmetic, with
    the compiler free to contract products into FMAs differently) and within
    1e-5 per cell of the FP64 oracle."""
    import os
    from cases import percell_relerr, percell_torque_relerr
    _say("fused", N, terms)
    cs = (2e-9, 2.5e-9, 1.5e-9)
    shape = (N[2], N[1], N[0])
    rng = np.random.default_rng(N[0] + N[1])
    params = dict(Ms=9e5, Aex=1.3e-11, Ku1=5e5, alpha=0.1)
    if "dmi" in terms:
        params["Dind"] = 2e-3
    if "none" in terms:
        params["Aex"] = 0.0
    if "regions" in terms:
        params["regions"] = rng.integers(0, 3, shape).astype(np.uint8)
        params["Aex"] = {0: 1e-11, 1: 2e-11, 2: 1.5e-11} if params["Aex"] else 0.0
        if "Dind" in params:
            params["Dind"] = {0: 2e-3, 1: -1e-3, 2: 3e-3}
    so = oracle.Sim(N, cs, dtype=np.float64, pbc=(0, 2, 1) if "pbc" in terms else (0, 0, 0))
    setup(so, oracle, dtype=np.float64, **params)
    m = random_unit(shape, 4)
    so.set_m(m)
    got = {}
    for fused in ("0", "1"):
        os.environ["UMHIP_FUSED_FIELD"] = fused
        try:
            sg = gpu.Sim(N, cs, dtype=np.float32, pbc=(0, 2, 1) if "pbc" in terms else (0, 0, 0))
            setup(sg, gpu, dtype=np.float32, **params)
            sg.set_m(m)
            sg.torque()
            with gpu.KernelTrace() as tr:
                got[fused] = sg.torque()
        finally:
            os.environ.pop("UMHIP_FUSED_FIELD", None)
        fk = [n for n in tr.names if n.startswith("void um::k_xi4_field<float, %d," % N[0])]
        sk = [n for n in tr.names if "k_field_torque" in n]
        assert (len(fk), len(sk)) == ((1, 0) if fused == "1" else (0, 1)), tr.names
    d01 = relerr(got["1"], got["0"])
    _say("fused vs two kernels", d01)
    assert d01 <= 1e-6
    b_o, t_o = so.beff(), so.torque()
    eb = percell_relerr(sg.beff(), b_o)
    et = percell_torque_relerr(got["1"], t_o, b_o)
    _say("per cell: B", eb, "tau", et)
    assert et <= 1e-5 and relerr(got["1"], t_o) <= 1e-5


# ---- steppers ---------------------------------------------------------------------------------
def _macrospin(mod, solver, fixdt, t_end, dtype, alpha=0.1, Bz=0.1):
    s = mod.Sim((1, 1, 1), (1e-9, 1e-9, 1e-9), dtype=dtype)
    setup(s, mod, Ms=8e5, Aex=0.0, alpha=alpha, Bext=(0.0, 0.0, Bz), demag=False, dtype=dtype)
    th0, ph0 = 1.2, 0.3
    s.set_m(np.array([math.sin(th0) * math.cos(ph0), math.sin(th0) * math.sin(ph0), math.cos(th0)]).reshape(3, 1, 1, 1))
    s.set_solver(solver)
    s.set_solver_params(dt=1e-14, fixdt=fixdt)
    s.run(t_end)
    t_end = s.state()["t"]
    g = GAMMA / (1 + alpha**2)
    th = 2 * math.atan(math.tan(th0 / 2) * math.exp(-g * alpha * Bz * t_end))
    ph = ph0 + g * Bz * t_end
    want = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
    return s.get_m().reshape(3), want


@pytest.mark.parametrize("solver,fixdt,tol", [(5, 0.0, 2e-5), (5, 1e-13, 1e-7), (4, 1e-13, 1e-6), (2, 1e-14, 1e-4)])
def test_macrospin_analytic(gpu, solver, fixdt, tol):
    m, want = _macrospin(gpu, solver, fixdt, 0.2e-9, np.float64)
    assert np.abs(m - want).max() < tol
    m, want = _macrospin(gpu, solver, fixdt, 0.2e-9, np.float32)
    assert np.abs(m - want).max() < max(tol, 2e-5)


def test_gammaLL_heun_golden(gpu):
    """test/gammaLL.mx3:6-28."""
    for dt in (np.float64, np.float32):
        N, c = 16, 5e-9
        s = gpu.Sim((N, N, 1), (c, c, c), dtype=dt)
        setup(s, gpu, Ms=1 / MU0, Aex=10e-12, alpha=0.001, Bext=(0.0, 0.0, 1.0), dtype=dt)
        s.set_solver(gpu.SOLVER_HEUN)
        s.set_solver_params(dt=1e-15, maxerr=1e-4)
        s.set_m(np.ones((3, 1, N, N)))
        s.run(0.1e-9)
        s.set_solver_params(dt=s.state()["dt"], maxerr=1e-4, gammaLL=2 * GAMMA)
        s.run(0.1e-9)
        got = s.average_m()
        assert np.abs(got - np.array([-0.286423, -0.746625, 0.5970549])).max() < 1.2e-4, (dt, got)


def test_rk4_golden(gpu):
    """test/rk4.mx3:6-28 (RK4 relax, fixed-step reversal, adaptive reversal)."""
    for dt in (np.float64, np.float32):
        s = gpu.Sim((128, 32, 1), (500e-9 / 128, 125e-9 / 32, 3e-9), dtype=dt)
        setup(s, gpu, Ms=800e3, Aex=13e-12, alpha=3.0, dtype=dt)
        m = np.zeros((3, 1, 32, 128))
        m[0], m[1] = 1, 0.1
        s.set_m(m)
        s.set_solver(gpu.SOLVER_RK4)
        s.set_solver_params(dt=1e-15)
        s.run(1e-9)
        s.set_scalar(gpu.P_ALPHA, 0.02)
        s.set_vector(gpu.P_BEXT, (-24.6e-3, 4.3e-3, 0.0))
        s.set_solver_params(dt=s.state()["dt"], fixdt=1e-14)
        s.run(0.1e-9)
        got = s.average_m()
        assert np.abs(got - np.array([0.59293, 0.63278, -0.08152])).max() < 1e-3, (dt, got)
        s.set_solver_params(dt=1e-14, fixdt=0.0)
        s.run(0.4e-9)
        got = s.average_m()
        assert np.abs(got - np.array([-0.8740, -0.2713, 0.01795])).max() < 1e-3, (dt, got)


@pytest.mark.parametrize("prec", [1, 0])
def test_strip_trajectory_vs_oracle(gpu, oracle, prec):
    """The 128x32x1 strip (RK45DP adaptive, MaxErr 1e-5): <m>(t) within 1e-4 of the FP64 oracle."""
    N, cs = (128, 32, 1), (3.90625e-9, 3.90625e-9, 3e-9)
    so, sg = _both(oracle, gpu, N, cs, prec, Ms=8e5, Aex=1.3e-11, alpha=0.02, Bext=(-24.6e-3, 4.3e-3, 0.0))
    m = strip_m(1)
    so.set_m(m)
    sg.set_m(m)
    for k in range(4):
        so.run(0.025e-9)
        sg.run(0.025e-9)
        d = np.abs(sg.average_m() - so.average_m()).max()
        _say("t", so.state()["t"], "d<m>", d, "steps", so.state()["nsteps"], sg.state()["nsteps"])
        assert d < (1e-9 if prec == 1 else 1e-4)


# ---- the timed kernels at the timed size --------------------------------------------------------
TIMED_512x512x32 = ["void um::k_xf4<float, 512>(", "void um::k_coly<float, 32, -1, 16, true, 1>(",
                    "void um::k_convz<float, 32, 8, 4>(", "void um::k_coly<float, 32, 1, 16, true, 1>(",
                    "void um::k_xi4_field<float, 512, false, 2, true, false>("]


def test_timed_kernels_512x512x32_vs_oracle(gpu, oracle):
    """The 512x512x32 FP32 workload (DMI, Ku1 || z, alpha 0.1, random m) through
    exactly the kernels the benchmark times -- the launch trace names the five
    evaluation kernels, the last the X inverse fused with the field and torque
    (uniform parameters, no region map) -- against the FP64 oracle at full size, per cell:
    |dB_i| / |B_i| <= 1e-5 for the effective field and |dtau_i| / |B_i| <= 1e-5
    for the torque (the field bounds the torque of a unit m), and normwise."""
    from cases import percell_relerr, percell_torque_relerr
    N, cs = (512, 512, 32), (2e-9, 2e-9, 2e-9)
    m = random_unit((N[2], N[1], N[0]), 2)
    params = dict(Ms=1.1e6, Aex=1.6e-11, Dind=3e-3, Ku1=8e5, alpha=0.1)
    so = oracle.Sim(N, cs, dtype=np.float64)
    setup(so, oracle, dtype=np.float64, **params)
    so.set_m(m)
    b_o = so.beff()
    t_o = so.torque()
    del so
    _say("oracle done")
    sg = gpu.Sim(N, cs, dtype=np.float32)
    setup(sg, gpu, dtype=np.float32, **params)
    sg.set_m(m)
    sg.torque()                       # plan and tensor setup outside the trace
    with gpu.KernelTrace() as tr:
        t_g = sg.torque()
    names = tr.names
    _say("traced", names)
    for k in TIMED_512x512x32:
        assert sum(n.startswith(k) for n in names) == 1, (k, names)
    b_g = sg.beff()
    eb, et = percell_relerr(b_g, b_o), percell_torque_relerr(t_g, t_o, b_o)
    _say("per cell: B", eb, "tau", et, "normwise:", relerr(b_g, b_o), relerr(t_g, t_o))
    assert eb <= 1e-5 and et <= 1e-5
    assert relerr(b_g, b_o) <= 1e-5 and relerr(t_g, t_o) <= 1e-5


@pytest.mark.parametrize("N,expect", [((512, 64, 32), ["void um::k_xf4<float, 512>(",
                                                       "void um::k_xi4_field<float, 512, false, 2, true, false>(",
                                                       "void um::k_convz<float, 32, 8, 4>("]),
                                      ((32, 512, 32), ["void um::k_coly<float, 32, -1, 16, true, 1>(",
                                                       "void um::k_coly<float, 32, 1, 16, true, 1>(",
                                                       "void um::k_convz<float, 32, 8, 4>(",
                                                       "void um::k_field_torque4<float, false, 2, false, true>("])])
def test_timed_kernels_slabs_vs_oracle(gpu, oracle, N, expect):
    """Smaller grids that select the timed X passes (Nx = 512, >= 2048 rows:
    X forward and the fused X inverse + field kernel) or the timed Y passes
    (Ny = 512) with the timed Z pass and the four-cell field kernel of uniform
    parameters, config-2 physics, FP32 vs the FP64 oracle per cell."""
    from cases import percell_relerr, percell_torque_relerr
    cs = (2e-9, 2e-9, 2e-9)
    m = random_unit((N[2], N[1], N[0]), 3)
    params = dict(Ms=1.1e6, Aex=1.6e-11, Dind=3e-3, Ku1=8e5, alpha=0.1)
    so, sg = _both(oracle, gpu, N, cs, 0, **params)
    so.set_m(m)
    sg.set_m(m)
    sg.torque()
    with gpu.KernelTrace() as tr:
        t_g = sg.torque()
    for k in expect:
        assert any(n.startswith(k) for n in tr.names), (k, tr.names)
    b_o, t_o, b_g = so.beff(), so.torque(), sg.beff()
    eb, et = percell_relerr(b_g, b_o), percell_torque_relerr(t_g, t_o, b_o)
    _say(N, "per cell: B", eb, "tau", et)
    assert eb <= 1e-5 and et <= 1e-5


# ---- full-size properties (512 x 512 x 32) -----------------------------------------------------
def test_512x512x32_full_size_properties(gpu):
    """The 512x512x32 workload at full size: FP32 field within 1e-5 of the FP64 GPU path, demag
    linearity and symmetry (<D u, v> = <u, D v>), unit |m| after RK45DP steps."""
    import torch
    N, cs = (512, 512, 32), (2e-9, 2e-9, 2e-9)
    m = random_unit((N[2], N[1], N[0]), 2).astype(np.float32)
    params = dict(Ms=1.1e6, Aex=1.6e-11, Dind=3e-3, Ku1=8e5, alpha=0.1)
    s64 = gpu.Sim(N, cs, dtype=np.float64)
    setup(s64, gpu, dtype=np.float64, **params)
    s64.set_m(m)
    b64 = s64.beff()
    del s64
    torch.cuda.empty_cache()
    s32 = gpu.Sim(N, cs, dtype=np.float32)
    setup(s32, gpu, dtype=np.float32, **params)
    s32.set_m(m)
    b32 = s32.beff()
    e = relerr(b32, b64)
    _say("512x512x32 FP32 vs FP64 Beff relerr", e)
    assert e <= 1e-5
    # demag symmetry on the FP32 path: <D u, v> = <u, D v>
    conv = gpu.DemagConvolution(N, cs, precision=0)
    u = [torch.randn(N[2], N[1], N[0], device="cuda") for _ in range(3)]
    v = [torch.randn(N[2], N[1], N[0], device="cuda") for _ in range(3)]
    Du = [torch.zeros_like(u[0]) for _ in range(3)]
    Dv = [torch.zeros_like(u[0]) for _ in range(3)]
    conv.Exec(Du, u, None, 1 / MU0)
    conv.Exec(Dv, v, None, 1 / MU0)
    a = sum(float((Du[c].double() * v[c].double()).sum()) for c in range(3))
    b = sum(float((u[c].double() * Dv[c].double()).sum()) for c in range(3))
    assert abs(a - b) <= 1e-5 * max(abs(a), abs(b))
    u2 = [2 * x for x in u]
    D2 = [torch.zeros_like(u[0]) for _ in range(3)]
    conv.Exec(D2, u2, None, 1 / MU0)
    assert max(float((D2[c] - 2 * Du[c]).abs().max()) for c in range(3)) <= 1e-6 * max(float(Du[c].abs().max()) for c in range(3))
    del conv, Du, Dv, D2, u, v, u2
    s32.set_solver_params(fixdt=1e-14)
    s32.steps(3)
    mm = s32.get_m()
    assert np.abs(np.linalg.norm(mm, axis=0) - 1).max() < 1e-6


# ---- HIP graph replay and ensembles ----------------------------------------------------------
def _graph_pair(gpu, N, solver, prec, temp, seed=7, stream=0):
    dt = np.float32 if prec == 0 else np.float64
    sims = []
    for _ in range(2):
        s = gpu.Sim(N, (2e-9, 2e-9, 2e-9), dtype=dt)
        setup(s, gpu, Ms=8e5, Aex=1.3e-11, Dind=1e-3, Ku1=2e5, alpha=0.05, Bext=(0.0, 0.01, 0.0),
              temp=temp, seed=seed, dtype=dt)
        s.set_thermal_seed(seed, stream)
        s.set_solver(solver)
        s.set_solver_params(dt=1e-14, fixdt=2e-14)
        s.set_m(random_unit((N[2], N[1], N[0]), 31))
        sims.append(s)
    sims[1].set_graph(1)
    return sims


def _same_state(a, b):
    sa, sb = a.state(), b.state()
    for key in ("t", "nsteps", "nevals", "lasterr", "peakerr", "lasttorque"):
        if key in sa:
            assert sa[key] == sb[key], (key, sa[key], sb[key])
    assert np.array_equal(a.get_m(), b.get_m())


@pytest.mark.parametrize("solver", [5, 2, 4, 3])
@pytest.mark.parametrize("prec", [0, 1])
def test_graph_steps_match_eager(gpu, solver, prec):
    """Fixed-dt steps replayed from a captured two-step graph are bit-identical to
    eager steps (m and the step statistics), across chunks of odd length (the
    pointer parity changes, forcing a re-capture), a parameter change between
    chunks (stale graph) and T > 0 (the Philox step counter read on the device)."""
    _say("graph", solver, prec)
    eager, graph = _graph_pair(gpu, (32, 16, 2), solver, prec, temp=300.0)
    for chunk in (5, 4, 7, 2):
        eager.steps(chunk)
        graph.steps(chunk)
        _same_state(eager, graph)
    for s in (eager, graph):
        s.set_scalar(gpu.P_ALPHA, 0.2)
    eager.steps(6)
    graph.steps(6)
    _same_state(eager, graph)


def test_sims_steps_ensemble(gpu):
    """um_sims_steps over members of different sizes (Heun, T = 300 K, one thermal
    stream per member) equals stepping each member on its own."""
    sizes = [(32, 32, 1), (48, 32, 3), (64, 16, 2), (16, 48, 4)]
    batch, single = [], []
    for i, N in enumerate(sizes):
        a, b = _graph_pair(gpu, N, 2, 0, temp=300.0, seed=11, stream=i)
        single.append(a)
        batch.append(b)
    for chunk in (9, 6):
        gpu.sims_steps(batch, chunk)
        for s in single:
            s.steps(chunk)
        for a, b in zip(single, batch):
            _same_state(a, b)


def test_generic_graph_capture(gpu):
    """um_graph_begin/end/launch on the library stream: a captured Madd chain,
    replayed three times with eager work and synchronisation in between, equals
    the same chain run eagerly each time."""
    import ctypes as C
    import torch
    s = torch.cuda.Stream()
    gpu._call("um_set_stream", C.c_void_p(s.cuda_stream))
    try:
        n = 4099
        x = torch.rand(n, device="cuda", dtype=torch.float32)
        y = torch.rand(n, device="cuda", dtype=torch.float32)
        acc = torch.zeros(n, device="cuda", dtype=torch.float32)
        ref = acc.clone()
        torch.cuda.synchronize()
        with gpu.Graph() as g:
            gpu.Madd(acc, [acc, x, y], [1.0, 0.5, -0.25])
            gpu.Madd(acc, [acc, y], [0.75, 0.125])
        for _ in range(3):
            g.launch()
            gpu._call("um_sync")
            gpu.Madd(ref, [ref, x, y], [1.0, 0.5, -0.25])
            gpu.Madd(ref, [ref, y], [0.75, 0.125])
            gpu._call("um_sync")
            assert torch.equal(acc, ref)
            x.mul_(0.5)   # eager work between replays
            torch.cuda.synchronize()
    finally:
        gpu._call("um_set_stream", C.c_void_p(torch.cuda.current_stream().cuda_stream))


@pytest.mark.parametrize("prec", [1, 0])
def test_exchange_inhomo_gpu(gpu, prec):
    """test/exchange_inhomo.mx3 on the HIP path: OVF-loaded m, four regions with
    harmonic-mean Aex; E_exch equals OOMMF's value to 1e-5, for both Msat maps."""
    from test_ovf import E_EXCH_OOMMF, run_exchange_inhomo
    dt = np.float32 if prec == 0 else np.float64
    for e in run_exchange_inhomo(gpu.Sim, gpu, dt):
        assert abs(e - E_EXCH_OOMMF) <= 1e-5 * E_EXCH_OOMMF, e


# ---- RK23, relax and energies -------------------------------------------------------------------
@pytest.mark.parametrize("prec", [1, 0])
def test_energy_terms_vs_oracle(gpu, oracle, prec):
    """All energy terms (demag, exchange+DMI, Zeeman, K1, K2, Kc1..3, thermal, total)
    on a mixed configuration, vs the FP64 oracle."""
    N, cs = (12, 10, 3), (2e-9, 3e-9, 2.5e-9)
    params = dict(Ms=7e5, Aex=1.1e-11, Dind=1e-3, alpha=0.1, Ku1=3e5, Ku2=-1e5, anisU=(0.3, -0.5, 0.8),
                  Kc=(2e5, -3e4, 1e4), c1=(1, 1, 0), c2=(-1, 1, 0), Bext=(0.01, -0.02, 0.03), temp=250.0, seed=4)
    so, sg = _both(oracle, gpu, N, cs, prec, **params)
    for s in (so, sg):
        s.set_solver_params(dt=1e-14)
    m = random_unit((N[2], N[1], N[0]), 12)
    so.set_m(m)
    sg.set_m(m)
    tol = 1e-11 if prec == 1 else 1e-5
    for t in range(10):
        eo, eg = so.energy(t), sg.energy(t)
        assert abs(eg - eo) <= tol * max(abs(eo), 1e-30), (t, eg, eo)


@pytest.mark.parametrize("prec", [1, 0])
def test_rk23_trajectory_vs_oracle(gpu, oracle, prec):
    """Adaptive RK23 (engine/rk23.go) on the 128x32x1 strip's physics: <m>(t) vs the FP64 oracle."""
    N, cs = (64, 16, 1), (3.90625e-9, 3.90625e-9, 3e-9)
    so, sg = _both(oracle, gpu, N, cs, prec, Ms=8e5, Aex=1.3e-11, alpha=0.02, Bext=(-24.6e-3, 4.3e-3, 0.0))
    for s in (so, sg):
        s.set_solver(gpu.SOLVER_RK23)
    m = strip_m(1)[:, :, :16, :64]
    so.set_m(m)
    sg.set_m(m)
    for _ in range(3):
        so.run(0.02e-9)
        sg.run(0.02e-9)
        d = np.abs(sg.average_m() - so.average_m()).max()
        assert d < (1e-9 if prec == 1 else 1e-4), d


@pytest.mark.parametrize("prec", [1, 0])
def test_standard_problem4_relax_and_run(gpu, prec):
    """test/standardproblem4.mx3: relax() then B_ext = (-24.6, 4.3, 0) mT for 1 ns;
    <m> within the reference's TOL = 1e-5 after each."""
    _say("sp4", prec)
    from cases import uniform_m
    dt = np.float32 if prec == 0 else np.float64
    N, cs = (128, 32, 1), (500e-9 / 128, 125e-9 / 32, 3e-9)
    s = gpu.Sim(N, cs, dtype=dt)
    setup(s, gpu, Ms=800e3, Aex=13e-12, alpha=0.02, dtype=dt)
    s.set_m(uniform_m(N, (1, 0.1, 0)))
    s.relax()
    assert np.abs(s.average_m() - np.array([0.9669684171676636, 0.1252832127904892, 0.0])).max() < 1e-5
    s.set_vector(gpu.P_BEXT, (-24.6e-3, 4.3e-3, 0.0))
    s.run(1e-9)
    got = s.average_m()
    assert np.abs(got - np.array([-0.9846124053001404, 0.12604689081287384, 0.04327124357223511])).max() < 1e-5, got


def test_cubic_anisotropy_relax(gpu):
    """test/cubicanisotropy.mx3: relaxed <m> in a cylinder with Kc1 = 0.2e6 against an
    in-plane field, vs OOMMF (tol 1e-3)."""
    from cases import cylinder_geom, uniform_m
    N, cs = (64, 64, 1), (4e-9, 4e-9, 2e-9)
    s = gpu.Sim(N, cs, dtype=np.float32)
    setup(s, gpu, Ms=1100e3, Aex=13e-12, alpha=1.0, Kc=(0.2e6, 0.0, 0.0), c1=(100, 0, 0), c2=(0, 1, 0),
          vol=cylinder_geom(N, cs, 64 * 4e-9, 64 * 4e-9), dtype=np.float32)
    s.set_m(uniform_m(N, (1, 0, 0)))
    for b, want in [(0.0, (1, 0, 0)), (0.01, (0.9994, 0.0270, 0)), (0.03, (0.9967, 0.0776, 0)),
                    (0.10, (0.9711, 0.2367, 0)), (0.30, (0.7040, 0.7040, 0))]:
        s.set_vector(gpu.P_BEXT, (b, b, 0.0))
        s.relax()
        got = s.average_m()
        assert np.abs(got - np.array(want)).max() < 1e-3, (b, got)


@pytest.mark.parametrize("prec", [1, 0])
def test_energy_mx3_oommf(gpu, prec):
    """test/energy.mx3: energy differences between the relaxed vortex state of an
    ellipse and its relaxed state in 100 mT, vs OOMMF (the reference's tolerances)."""
    _say("energy.mx3", prec)
    from cases import ellipse_geom, vortex
    dt = np.float32 if prec == 0 else np.float64
    l, w, d, c = 512e-9, 256e-9, 40e-9, 4e-9
    vol = l * w * d * (math.pi / 4)
    N, cs = (128, 64, 1), (c, c, d)
    s = gpu.Sim(N, cs, dtype=dt)
    setup(s, gpu, Ms=860e3, Aex=13e-12, Ku1=50.0, anisU=(1, 0, 0), alpha=3.0,
          vol=ellipse_geom(N, cs, l, w), dtype=dt)
    s.set_m(vortex(N, cs, 1, 1))
    terms = {"tot": 9, "exc": 1, "dem": 0, "zee": 2}
    s.relax()
    E1 = {k: s.energy(t) for k, t in terms.items()}
    E1["ani"] = s.energy(3) + s.energy(4)
    s.set_vector(gpu.P_BEXT, (100e-3, 0.0, 0.0))
    s.relax()
    E2 = {k: s.energy(t) for k, t in terms.items()}
    E2["ani"] = s.energy(3) + s.energy(4)
    oommf = {"tot": (12274.249330 - (-56507.645615)) * vol, "exc": (5038.693945 - 71.438905) * vol,
             "ani": (19.517039 - 0.271210) * vol, "dem": (7216.038345 - 29186.031229) * vol,
             "zee": (0 - (-85765.386961)) * vol}
    tols = {"ani": 1e-21, "zee": 1e-18, "exc": 1e-18, "dem": 1e-18, "tot": 1e-18}
    for k in ("ani", "zee", "exc", "dem", "tot"):
        assert abs((E1[k] - E2[k]) - oommf[k]) < tols[k], (k, E1[k] - E2[k], oommf[k])


def test_dmi_bloch_to_neel_relax(gpu):
    """test/dmi.mx3: interfacial DMI turns a Bloch wall into a Neel wall; relaxed
    <m> = (0.059, 0.035) within 0.01."""
    from cases import two_domain
    N, cs = (128, 128, 1), (250e-9 / 128, 250e-9 / 128, 0.6e-9)
    s = gpu.Sim(N, cs, dtype=np.float32)
    setup(s, gpu, Ms=1100e3, Aex=16e-12, Dind=0.1e-3, Ku1=1.27e6, anisU=(0, 0, 1), alpha=3.0, dtype=np.float32)
    s.set_m(two_domain(N, cs, (0, 0, -1), (1, 1, 0), (0, 0, 1)))
    s.relax()
    got = s.average_m()
    assert abs(got[0] - 0.059) < 0.01 and abs(got[1] - 0.035) < 0.01, got


# ---- device-controlled adaptive steps -------------------------------------------------------
@pytest.mark.parametrize("prec", [1, 0])
@pytest.mark.parametrize("case", ["strip", "film_rejects"])
def test_device_controlled_run_matches_host(gpu, oracle, prec, case):
    """Adaptive RK45DP run() with the acceptance, the next dt and the FSAL copy
    decided on the device (graph mode: 8-step graph replays, one readback per
    replay) against the same run controlled from the host (one readback per
    step): same accepted and rejected step counts, same final time, <m> and the
    per-cell m within rounding (the controller's pow() is the device's, not
    the host's), and <m> within 1e-4 of the FP64 oracle.  film_rejects starts
    from a dt far too large, so the first steps are rejected."""
    dtype = np.float32 if prec == 0 else np.float64
    if case == "strip":
        N, cs = (128, 32, 1), (3.90625e-9, 3.90625e-9, 3e-9)
        p = dict(Ms=8e5, Aex=1.3e-11, alpha=0.02, Bext=(-24.6e-3, 4.3e-3, 0.0))
        m, dt0, T = strip_m(1), 1e-15, 0.05e-9
    else:
        N, cs = (32, 24, 6), (2e-9,) * 3
        p = dict(Ms=1.1e6, Aex=1.6e-11, Dind=3e-3, Ku1=8e5, alpha=0.1)
        m, dt0, T = random_unit((6, 24, 32), 31), 2e-12, 5e-12
    runs = {}
    for graph in (0, 1):
        s = gpu.Sim(N, cs, dtype=dtype)
        setup(s, gpu, dtype=dtype, **p)
        s.set_m(m)
        s.set_solver_params(dt=dt0)
        s.set_graph(graph)
        s.run(T)
        runs[graph] = (s.state(), s.get_m(), s.average_m())
        _say(case, prec, "graph", graph, runs[graph][0])
    (s0, m0, a0), (s1, m1, a1) = runs[0], runs[1]
    assert s1["nsteps"] == s0["nsteps"] and s1["nundone"] == s0["nundone"] and s1["nevals"] == s0["nevals"]
    if case == "film_rejects":
        assert s0["nundone"] > 0
    assert abs(s1["t"] - s0["t"]) <= 1e-12 * s0["t"] and abs(s1["t"] - T) <= 1e-9 * T
    tol = 1e-12 if prec == 1 else 1e-6
    assert np.abs(a1 - a0).max() <= tol, (a0, a1)
    assert np.abs(m1.astype(np.float64) - m0).max() <= (1e-10 if prec == 1 else 1e-5)
    for key in ("lasterr", "peakerr", "lasttorque"):
        assert abs(s1[key] - s0[key]) <= 1e-6 * max(abs(s0[key]), 1e-300), key
    so = oracle.Sim(N, cs, dtype=np.float64)
    setup(so, oracle, dtype=np.float64, **p)
    so.set_m(m)
    so.set_solver_params(dt=dt0)
    so.run(T)
    d = float(np.abs(a1 - so.average_m()).max())
    _say("device-controlled vs oracle <m>", d)
    assert d <= 1e-4
