"""Spin-transfer torques (zhangli2.cl, slonczewski2.cl) and bulk DMI
(dmibulk.cl) on the GPU through the C ABI: the unfused entry points and the
simulation handle vs the oracle, and the reference's goldens
test/standardproblem5.mx3 (Zhang-Li) and test/dmibulksymm.mx3 (bulk DMI).

Tolerances: FP64 <= 1e-12 (unfused) / 1e-11 (simulation) normwise against the
FP64 oracle; FP32 <= 2e-6 / 1e-5.
"""
import sys

import numpy as np
import pytest

from cases import random_unit, setup, uniform_m, vortex

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


def _err(got, want):
    return float(np.abs(np.asarray(got, np.float64) - want).max() / max(np.abs(want).max(), 1e-300))


@pytest.mark.parametrize("prec", [1, 0])
@pytest.mark.parametrize("N,pbc,openbc", [((12, 10, 1), 0, 0), ((10, 8, 6), 0, 0), ((10, 8, 6), 0, 1),
                                          ((8, 6, 5), 7, 0), ((9, 7, 4), 3, 0)])
def test_add_dmi_bulk_vs_oracle(gpu, oracle, N, pbc, openbc, prec):
    from umagnus_amd.params import exchange_lut
    _say("dmibulk", N, pbc, openbc, prec)
    dt = np.float32 if prec == 0 else np.float64
    rng = np.random.default_rng(sum(N) + pbc)
    shape = (N[2], N[1], N[0])
    m = random_unit(shape, 7)
    m[:, 0, 0, 1] = 0                 # a vacuum cell: extrapolated neighbour
    regions = rng.integers(0, 3, shape).astype(np.uint8)
    Ms = rng.uniform(5e5, 1e6, shape)
    A = exchange_lut({0: 1.1e-11, 1: 1.4e-11, 2: 0.9e-11}, dtype=dt)
    D = exchange_lut({0: 2e-3, 1: -1e-3, 2: 3e-3}, dtype=dt)
    cs = (2e-9, 3e-9, 2.5e-9)
    Bo = np.zeros((3,) + shape, dt)
    mo = np.ascontiguousarray(m, dt)
    R = oracle._rt(dt)
    Mo = np.ascontiguousarray(Ms, dt)
    oracle.fn("adddmibulk", dt)(*[oracle.ptr(x) for x in (Bo[0], Bo[1], Bo[2], mo[0], mo[1], mo[2])],
                                oracle.ptr(Mo), R(1.0), oracle.ptr(A), oracle.ptr(D), oracle.ptr(regions),
                                R(cs[0]), R(cs[1]), R(cs[2]), N[0], N[1], N[2], pbc, openbc)
    B = [_t(np.zeros(shape), dt) for _ in range(3)]
    gpu.AddDMIBulk(B, [_t(m[c], dt) for c in range(3)], _t(A, dt), _t(D, dt), _t(Ms, dt), _t(regions, np.uint8),
                   cs, pbc=pbc, OpenBC=bool(openbc))
    assert _err(np.stack([_h(b) for b in B]), Bo) <= (1e-12 if prec == 1 else 2e-6)


@pytest.mark.parametrize("prec", [1, 0])
@pytest.mark.parametrize("pbc", [0, 5])
def test_add_stt_vs_oracle(gpu, oracle, pbc, prec):
    _say("stt", pbc, prec)
    dt = np.float32 if prec == 0 else np.float64
    N, cs = (14, 9, 5), (2e-9, 3e-9, 2.5e-9)
    shape = (N[2], N[1], N[0])
    rng = np.random.default_rng(11 + pbc)
    m = random_unit(shape, 3)
    Ms = rng.uniform(5e5, 1e6, shape)
    Ms[0, 0, 0] = 0                   # no Slonczewski torque there
    alpha = rng.uniform(0.01, 0.3, shape)
    J = [rng.uniform(-2e12, 2e12, shape) for _ in range(3)]
    J[2][1, 2, 3] = 0
    xi, pol, lam, eps = 0.07, 0.6, 1.7, 0.02
    p = (0.2, -0.4, 0.9)
    R = oracle._rt(dt)
    to = np.zeros((3,) + shape, dt)
    mo = np.ascontiguousarray(m, dt)
    a = [np.ascontiguousarray(x, dt) for x in (Ms, J[0], J[1], J[2], alpha)]
    oracle.fn("addzhangli", dt)(*[oracle.ptr(x) for x in (to[0], to[1], to[2], mo[0], mo[1], mo[2])],
                                oracle.ptr(a[0]), R(1), oracle.ptr(a[1]), R(1), oracle.ptr(a[2]), R(1),
                                oracle.ptr(a[3]), R(1), oracle.ptr(a[4]), R(1), None, R(xi), None, R(pol),
                                R(cs[0]), R(cs[1]), R(cs[2]), N[0], N[1], N[2], pbc)
    oracle.fn("addslonczewski", dt)(*[oracle.ptr(x) for x in (to[0], to[1], to[2], mo[0], mo[1], mo[2])],
                                    oracle.ptr(a[0]), R(1), oracle.ptr(a[3]), R(1), None, R(p[0]), None, R(p[1]),
                                    None, R(p[2]), oracle.ptr(a[4]), R(1), None, R(pol), None, R(lam), None,
                                    R(eps), None, R(0.0), R(N[2] * cs[2]), R(-1.0), int(np.prod(N)))
    T = [_t(np.zeros(shape), dt) for _ in range(3)]
    mt = [_t(m[c], dt) for c in range(3)]
    Mst, at = _t(Ms, dt), _t(alpha, dt)
    gpu.AddZhangLiTorque(T, mt, Mst, [_t(j, dt) for j in J], at, xi, pol, cs, pbc=pbc)
    gpu.AddSlonczewskiTorque2(T, mt, Mst, _t(J[2], dt), p, at, pol, lam, eps, 0.0, N[2] * cs[2], flp=-1.0)
    assert _err(np.stack([_h(t) for t in T]), to) <= (1e-12 if prec == 1 else 2e-6)


def _stt_sim(mod, N, cs, dtype, bulk=True):
    from umagnus_amd.params import exchange_lut
    s = mod.Sim(N, cs, dtype=dtype)
    setup(s, mod, Ms=8e5, Aex=1.3e-11, alpha=0.1, Ku1=1e5, anisU=(0.3, 0, 1), Bext=(0.01, 0, 0.02), dtype=dtype)
    if bulk:
        s.set_dmi_bulk(exchange_lut(1.5e-3, dtype=dtype))
    s.set_vector(mod.P_J, (1.2e12, -4e11, 3e11))
    s.set_scalar(mod.P_XI, 0.05)
    s.set_scalar(mod.P_POL, 0.7)
    s.set_vector(mod.P_FIXEDLAYER, (1.0, 1.0, 0.3))
    s.set_scalar(mod.P_LAMBDA, 1.5)
    s.set_scalar(mod.P_EPSPRIME, 0.01)
    return s


@pytest.mark.parametrize("prec", [1, 0])
def test_sim_torque_stt_bulk_vs_oracle(gpu, oracle, prec):
    """Fused field + torque with bulk DMI, Zhang-Li and Slonczewski torques, the
    fixed layer below, vs the oracle; then 20 RK45DP steps."""
    _say("sim stt", prec)
    N, cs = (16, 12, 3), (2e-9, 2e-9, 1.5e-9)
    dt = np.float32 if prec == 0 else np.float64
    sg, so = _stt_sim(gpu, N, cs, dt), _stt_sim(oracle, N, cs, np.float64)
    for s in (sg, so):
        s.set_stt_flags(0, 0, fixed_layer_top=False)
    m = random_unit((N[2], N[1], N[0]), 5)
    sg.set_m(m)
    so.set_m(m)
    tol = 1e-11 if prec == 1 else 1e-5
    assert _err(sg.beff(), so.beff()) <= tol
    assert _err(sg.torque(), so.torque()) <= tol
    for s in (sg, so):
        s.set_solver_params(dt=1e-14, fixdt=1e-14)
        s.steps(20)
    assert np.abs(sg.get_m() - so.get_m()).max() <= (1e-9 if prec == 1 else 1e-4)
    # switching each torque off
    for zl, sl in ((1, 0), (0, 1)):
        for s in (sg, so):
            s.set_stt_flags(zl, sl)
        assert _err(sg.torque(), so.torque()) <= tol


def test_standard_problem5(gpu):
    """test/standardproblem5.mx3: 32x32x4 permalloy square, vortex(1, 1), relax(),
    then Zhang-Li with xi = 0.05, J = (1e12, 0, 0), Pol = 1 for 1 ns; TOL 1e-4."""
    _say("sp5")
    N, cs = (32, 32, 4), (100e-9 / 32, 100e-9 / 32, 10e-9 / 4)
    s = gpu.Sim(N, cs, dtype=np.float32)
    setup(s, gpu, Ms=800e3, Aex=13e-12, alpha=0.1, dtype=np.float32)
    s.set_m(vortex(N, cs, 1, 1))
    s.relax()
    s.set_scalar(gpu.P_XI, 0.05)
    s.set_vector(gpu.P_J, (1e12, 0, 0))
    s.set_scalar(gpu.P_POL, 1.0)
    s.run(1e-9)
    got = s.average_m()
    _say("m", got)
    assert np.abs(got - np.array([-0.23479773, -0.09453578, 0.02296375])).max() < 1e-4, got


def test_dmi_bulk_rotation_symmetry(gpu):
    """test/dmibulksymm.mx3 at full size: 16^3 cube, Dbulk 1e-3, 0.1 ns from
    uniform states; rotating the start rotates the result (TOL 1e-4)."""
    from umagnus_amd.params import exchange_lut
    N, cs = (16, 16, 16), (2e-9,) * 3

    def run(v):
        s = gpu.Sim(N, cs, dtype=np.float32)
        setup(s, gpu, Ms=1e6, Aex=10e-12, alpha=1.0, dtype=np.float32)
        s.set_dmi_bulk(exchange_lut(1e-3, dtype=np.float32))
        s.set_m(uniform_m(N, v))
        s.run(1e-10)
        return s.average_m()
    mx = run((1, 0, 0))
    assert abs(run((0, 1, 0))[1] - mx[0]) < 1e-4
    assert abs(run((0, 0, 1))[2] - mx[0]) < 1e-4
    myz = run((0, 1, 1))
    assert abs(run((1, 0, 1))[1] - myz[0]) < 1e-4
    assert abs(run((1, 1, 0))[2] - myz[0]) < 1e-4


def test_interfacial_and_bulk_dmi_refused(gpu):
    from umagnus_amd.params import exchange_lut
    s = gpu.Sim((8, 8, 1), (2e-9,) * 3, dtype=np.float32)
    setup(s, gpu, Ms=8e5, Aex=1.3e-11, alpha=0.1, Dind=1e-3, dtype=np.float32)
    s.set_dmi_bulk(exchange_lut(1e-3, dtype=np.float32))
    s.set_m(uniform_m((8, 8, 1), (1, 0, 0)))
    with pytest.raises(gpu.UmError):
        s.torque()


@pytest.mark.parametrize("dmi", ["bulk", "interfacial"])
def test_four_cell_kernel_stt_vs_oracle(gpu, oracle, dmi):
    """512 x 256 x 4 cells (the four-cell field kernel's size range, FP32):
    B_eff and torque with bulk or interfacial DMI, Zhang-Li and Slonczewski
    torques and per-cell regions vs the FP64 oracle (1e-5 normwise), then 5
    RK45DP steps (<m> within 1e-6)."""
    from umagnus_amd.params import exchange_lut
    _say("four-cell stt", dmi)
    N, cs = (512, 256, 4), (2e-9, 2e-9, 1.5e-9)
    shape = (N[2], N[1], N[0])
    regions = (np.arange(N[0])[None, None, :] // 128 + np.zeros(shape, int)).astype(np.uint8)

    def make(mod, dt):
        bulk = dmi == "bulk"
        s = _stt_sim(mod, N, cs, dt, bulk=bulk)
        if not bulk:
            s.set_exchange_luts(exchange_lut({0: 1.3e-11, 1: 1.1e-11, 2: 1.5e-11, 3: 1.2e-11}, dtype=dt),
                                exchange_lut({0: 1e-3, 1: -2e-3, 2: 1.5e-3, 3: 5e-4}, dtype=dt))
        s.set_regions(regions)
        s.set_stt_flags(0, 0, fixed_layer_top=False)
        return s
    sg, so = make(gpu, np.float32), make(oracle, np.float64)
    m = random_unit(shape, 12)
    sg.set_m(m)
    so.set_m(m)
    assert _err(sg.beff(), so.beff()) <= 1e-5
    assert _err(sg.torque(), so.torque()) <= 1e-5
    for s in (sg, so):
        s.set_solver_params(dt=1e-14, fixdt=1e-14)
        s.steps(5)
    assert np.abs(sg.average_m() - so.average_m()).max() <= 1e-6
