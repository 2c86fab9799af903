"""Newell demag tensor on the GPU (mag/newellDemagKernel.go) through the C ABI:
element-wise parity with the oracle, the convolution built on it, and the
reference's Newell standard-problem goldens.

Tolerance of the tensor: |gpu - oracle| <= 1e-6 |oracle| + 1e-12 max|oracle|.
Both evaluate the same closed forms in FP64; they differ by the rounding of
the transcendental functions (different libm) amplified by the cancellation
of the 27-term sums, which grows like R^6 and reaches ~1e-6 of an element at
the default asymptotic radius of 32 cells.
"""
import sys

import numpy as np
import pytest

from cases import MU0, random_unit, setup, uniform_m

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
    return umhip


def _t(a, dtype):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype)).cuda()


def _h(t):
    return t.detach().cpu().numpy()


SP4_CS3 = (500e-9 / 128, 125e-9 / 32, 1.5e-9)
TENSOR_CASES = [
    # (N, cellsize, asymptotic_radius, zero_self_demag)
    ((16, 8, 2), SP4_CS3, 32, 0),
    ((40, 24, 5), (2e-9, 3e-9, 2.5e-9), 8, 0),           # near and far field, refined sub-cells
    ((48, 40, 1), (5e-9, 5e-9, 3e-9), 6, 0),             # thin film
    ((24, 20, 12), (1e-9, 1e-9, 1e-9), 6, 1),            # cubic cells, self term zeroed
    ((10, 10, 4), (1e-9, 1.41421356e-9, 1e-9), -1, 0),   # irrational ratio, no far field
    ((80, 70, 3), (4e-9, 4e-9, 1.5e-9), 32, 0),          # default radius reached in-plane
]


@pytest.mark.parametrize("N,cs,arad,zsd", TENSOR_CASES)
def test_newell_tensor_vs_oracle(gpu, oracle, N, cs, arad, zsd):
    _say("newell tensor", N, cs, arad, zsd)
    Kg = gpu.demag_kernel_newell(N, cs, asymptotic_radius=arad, zero_self_demag=zsd)
    Ko = oracle.demag_kernel_newell(N, cs, asymptotic_radius=arad, zero_self_demag=zsd)
    for c in range(6):
        assert Kg[c].shape == Ko[c].shape
        scale = max(np.abs(Ko[c]).max(), 1e-300)
        err = np.abs(Kg[c] - Ko[c]) - 1e-6 * np.abs(Ko[c])
        assert err.max() <= 1e-12 * scale, (c, float(err.max() / scale))


@pytest.mark.parametrize("prec", [1, 0])
@pytest.mark.parametrize("N", [(16, 8, 2), (24, 12, 1), (20, 10, 6)])
def test_newell_demag_field_vs_oracle(gpu, oracle, N, prec):
    """B_demag through the Newell plan vs the oracle's convolution of its own
    Newell tensor (FP64: 1e-10 normwise; FP32: 1e-5)."""
    _say("newell conv", N, prec)
    cs = (3e-9, 2e-9, 2.5e-9)
    rng = np.random.default_rng(sum(N))
    m = random_unit((N[2], N[1], N[0]), sum(N) + 3)
    Ms = rng.uniform(5e5, 1e6, (N[2], N[1], N[0]))
    Ko = oracle.demag_kernel_newell(N, cs)
    Bo = oracle.Demag(N, cs, dtype=np.float64, K=Ko).exec(m, Ms)
    dt = np.float32 if prec == 0 else np.float64
    conv = gpu.DemagConvolution(N, cs, precision=prec, newell=True)
    B = [_t(np.zeros_like(m[0]), dt) for _ in range(3)]
    conv.Exec(B, [_t(m[k], dt) for k in range(3)], None, _t(Ms, dt))
    Bg = np.stack([_h(b) for b in B]).astype(np.float64)
    err = np.abs(Bg - Bo).max() / np.abs(Bo).max()
    assert err < (1e-10 if prec == 1 else 1e-5), err


def _sp4(gpu, N, cs, prec, newell):
    dt = np.float32 if prec == 0 else np.float64
    s = gpu.Sim(N, cs, dtype=dt)
    setup(s, gpu, Ms=800e3, Aex=13e-12, alpha=0.02, dtype=dt)
    if newell:
        s.set_demag_newell(1)
    s.set_m(uniform_m(N, (1, 0.1, 0)))
    return s


@pytest.mark.parametrize("prec", [0, 1])
def test_standard_problem4_newell_3d(gpu, prec):
    """test/standardproblem4_newell-3d.mx3: 128x32x2 cells, EnableNewellDemag;
    relax() then B_ext = (-24.6, 4.3, 0) mT for 1 ns; TOL 1e-4."""
    _say("sp4 newell 3d", prec)
    s = _sp4(gpu, (128, 32, 2), SP4_CS3, prec, True)
    s.relax()
    got = s.average_m()
    _say("relaxed", got)
    assert np.abs(got - np.array([0.9669952392578125, 0.12521563470363617, 0])).max() < 1e-4, got
    s.set_vector(gpu.P_BEXT, (-24.6e-3, 4.3e-3, 0.0))
    s.run(1e-9)
    got = s.average_m()
    _say("reversed", got)
    assert np.abs(got - np.array([-0.9841473698616028, 0.13057094812393188, 0.043032702058553696])).max() < 1e-4, got


@pytest.mark.parametrize("prec", [0, 1])
def test_standard_problem4b_newell(gpu, prec):
    """test/standardproblem4b_newell.mx3: 128x32x1 cells, relax() then
    B_ext = (-35.5, -6.3, 0) mT for 1 ns; TOL 1e-3."""
    _say("sp4b newell", prec)
    s = _sp4(gpu, (128, 32, 1), (500e-9 / 128, 125e-9 / 32, 3e-9), prec, True)
    s.relax()
    got = s.average_m()
    _say("relaxed", got)
    assert np.abs(got - np.array([0.9669625163078308, 0.12528464198112488, 0])).max() < 1e-3, got
    s.set_vector(gpu.P_BEXT, (-35.5e-3, -6.3e-3, 0.0))
    s.run(1e-9)
    got = s.average_m()
    _say("reversed", got)
    assert np.abs(got - np.array([-0.9702987670898438, -0.12229042601585388, -0.004722175188362598])).max() < 1e-3, got


@pytest.mark.parametrize("prec", [0])
def test_standard_problem4_3d_bruteforce(gpu, prec):
    """test/standardproblem4-3d.mx3: the same 3D problem with the brute-force
    tensor; TOL 1e-4."""
    _say("sp4 3d bf", prec)
    s = _sp4(gpu, (128, 32, 2), SP4_CS3, prec, False)
    s.relax()
    got = s.average_m()
    _say("relaxed", got)
    assert np.abs(got - np.array([0.9669952392578125, 0.12521563470363617, 0])).max() < 1e-4, got
    s.set_vector(gpu.P_BEXT, (-24.6e-3, 4.3e-3, 0.0))
    s.run(1e-9)
    got = s.average_m()
    _say("reversed", got)
    assert np.abs(got - np.array([-0.9846296310424805, 0.1256464719772339, 0.04335508495569229])).max() < 1e-4, got


def test_sim_newell_switch(gpu, oracle):
    """set_demag_newell swaps the tensor of a live simulation and back (the plan
    is rebuilt); B_demag follows the oracle's Newell / brute-force fields."""
    N, cs = (12, 10, 3), (2e-9, 3e-9, 2.5e-9)
    m = random_unit((N[2], N[1], N[0]), 5)
    s = gpu.Sim(N, cs, dtype=np.float64)
    setup(s, gpu, Ms=7e5, Aex=1e-11, alpha=0.1, dtype=np.float64)
    s.set_m(m)
    so = oracle.Sim(N, cs, dtype=np.float64)
    setup(so, oracle, Ms=7e5, Aex=1e-11, alpha=0.1, dtype=np.float64)
    so.set_m(m)
    for newell in (0, 1, 0):
        s.set_demag_newell(newell)
        so.set_demag_newell(newell)
        Bg, Bo = s.demag_field(), so.demag_field()
        assert np.abs(Bg - Bo).max() <= 1e-10 * np.abs(Bo).max(), newell


@pytest.mark.parametrize("prec", [0, 1])
def test_voronoi_grains_newell_vs_oracle(gpu, oracle, prec):
    """The precision workload's physics at 96 x 80 x 8: 4 Voronoi regions with
    their own Ms and Aex, Newell tensor, alpha 0.1, random m: B_eff vs the FP64
    oracle, then 1 ps of adaptive RK45DP (<m> and the total energy)."""
    import bench
    _say("voronoi", prec)
    N = (96, 80, 8)
    make = bench.precision_workload(N, seed=3)
    dt = np.float32 if prec == 0 else np.float64
    sg, so = make(gpu, dt), make(oracle, np.float64)
    Bg, Bo = sg.beff(), so.beff()
    assert np.abs(Bg - Bo).max() <= (1e-5 if prec == 0 else 1e-11) * np.abs(Bo).max()
    for s in (sg, so):
        s.run(1e-12)
    assert np.abs(sg.average_m() - so.average_m()).max() <= (1e-4 if prec == 0 else 1e-9)
    eg, eo = sg.energy(9), so.energy(9)
    assert abs(eg - eo) <= (1e-4 if prec == 0 else 1e-9) * abs(eo), (eg, eo)
