"""Periodic boundaries on the GPU through the C ABI: the periodic demag
tensors (mag/demagkernel.go, mag/newellDemagKernel.go with SetPBC) and the
convolution built on them, including odd Nx (a full-length X transform), vs
the oracle; the reference's PBC goldens test/demag2Dpbc.mx3, demag2Dpbc2.mx3,
demagrodpbc.mx3, zhangliPBC.mx3, pbc1.mx3 and dmipbc.mx3.

Tolerances: the brute-force tensor 2e-12 of max|K| (same FP64 operations;
the offsets next to the source are summed as wavefront trees, the oracle's
sequential sums carry ~n eps of their thousands of cancelling terms); Newell 1e-6 relative + 1e-12 of max|K| (libm rounding
amplified by the 27-term cancellation); fields FP64 1e-10, FP32 1e-5 normwise.
"""
import sys

import numpy as np
import pytest

from cases import MU0, random_unit, setup, two_domain, coords, uniform_m

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


TENSOR_CASES = [
    ((4, 3, 2), (2e-9, 3e-9, 2.5e-9), (2, 1, 0)),
    ((3, 4, 2), (2e-9, 3e-9, 2.5e-9), (0, 0, 3)),
    ((6, 5, 1), (1e-9, 1.5e-9, 2e-9), (3, 2, 0)),
    ((4, 4, 3), (1e-9, 1e-9, 1e-9), (1, 1, 1)),
    ((5, 2, 1), (2e-9, 2e-9, 1e-9), (0, 4, 2)),
    ((32, 8, 1), (1e-9, 1e-9, 0.5e-9), (0, 8, 0)),
]


@pytest.mark.parametrize("N,cs,pbc", TENSOR_CASES)
def test_bf_tensor_pbc_vs_oracle(gpu, oracle, N, cs, pbc):
    _say("bf pbc tensor", N, pbc)
    Kg = gpu.demag_kernel_bf(N, cs, pbc=pbc)
    Ko = oracle.demag_kernel_bf(N, cs, pbc=pbc)
    # near offsets: tree vs sequential sums; images can cancel an off-diagonal
    # component to rounding noise (yz at 1e-15 with Pz = 3), so the bound is
    # scaled by the whole tensor
    scale = max(np.abs(k).max() for k in Ko)
    for c in range(6):
        assert Kg[c].shape == Ko[c].shape
        assert np.abs(Kg[c] - Ko[c]).max() <= 2e-12 * scale, c


@pytest.mark.parametrize("N,cs,pbc", TENSOR_CASES)
def test_newell_tensor_pbc_vs_oracle(gpu, oracle, N, cs, pbc):
    _say("newell pbc tensor", N, pbc)
    for arad in (4, 32):
        Kg = gpu.demag_kernel_newell(N, cs, pbc=pbc, asymptotic_radius=arad)
        Ko = oracle.demag_kernel_newell(N, cs, pbc=pbc, asymptotic_radius=arad)
        # images can cancel an odd component to rounding noise (P = 2 along a
        # periodic axis): the absolute part is scaled by the whole tensor
        scale = max(np.abs(k).max() for k in Ko)
        for c in range(6):
            err = np.abs(Kg[c] - Ko[c]) - 1e-6 * np.abs(Ko[c])
            assert err.max() <= 1e-12 * scale, (arad, c, float(err.max() / scale))


CONV_CASES = [
    ((16, 6, 1), (0, 4, 0)),
    ((6, 16, 1), (4, 0, 0)),
    ((3, 16, 1), (5, 0, 0)),          # odd Nx: full-length X transform
    ((5, 6, 4), (0, 0, 3)),
    ((8, 6, 4), (2, 2, 2)),
    ((7, 6, 3), (1, 0, 0)),           # odd Nx in 3D
    ((2, 2, 2), (0, 0, 16)),
    ((32, 16, 1), (2, 2, 0)),
]


@pytest.mark.parametrize("prec", [1, 0])
@pytest.mark.parametrize("N,pbc", CONV_CASES)
def test_demag_field_pbc_vs_oracle(gpu, oracle, N, pbc, prec):
    _say("pbc conv", N, pbc, prec)
    cs = (2e-9, 3e-9, 2.5e-9)
    rng = np.random.default_rng(sum(N) + sum(pbc))
    m = random_unit((N[2], N[1], N[0]), sum(N) + 11)
    Ms = rng.uniform(5e5, 1e6, (N[2], N[1], N[0]))
    Bo = oracle.Demag(N, cs, pbc=pbc, dtype=np.float64).exec(m, Ms)
    dt = np.float32 if prec == 0 else np.float64
    conv = gpu.DemagConvolution(N, cs, pbc=pbc, precision=prec)
    B = [_t(np.zeros_like(m[0]), dt) for _ in range(3)]
    conv.Exec(B, [_t(m[k], dt) for k in range(3)], None, _t(Ms, dt))
    Bg = np.stack([_h(b) for b in B]).astype(np.float64)
    err = np.abs(Bg - Bo).max() / np.abs(Bo).max()
    assert err < (1e-10 if prec == 1 else 1e-5), err


def _avg_b(gpu, N, cs, pbc, prec=0):
    dt = np.float32 if prec == 0 else np.float64
    conv = gpu.DemagConvolution(N, cs, pbc=pbc, precision=prec)
    out = []
    for v in np.eye(3):
        B = [_t(np.zeros((N[2], N[1], N[0])), dt) for _ in range(3)]
        m = [_t(np.full((N[2], N[1], N[0]), v[c]), dt) for c in range(3)]
        conv.Exec(B, m, None, _t(np.full((N[2], N[1], N[0]), 1 / MU0), dt))
        out.append([float(_h(b).astype(np.float64).mean()) for b in B])
    return np.array(out)


def _expect_film(k):
    off = ~np.eye(3, dtype=bool)
    assert np.abs(k[off]).max() < 1e-9, k
    assert abs(k[0, 0]) < 1e-2 and abs(k[1, 1]) < 1e-2, k
    assert abs(k[2, 2] + 1) < 2e-2, k


def test_demag2d_pbc_goldens(gpu):
    """test/demag2Dpbc.mx3 and demag2Dpbc2.mx3 (even and odd Nx)."""
    cs = (1e-9, 1e-9, 0.5e-9)
    _expect_film(_avg_b(gpu, (128, 2, 1), cs, (0, 32, 0)))
    _expect_film(_avg_b(gpu, (2, 128, 1), cs, (32, 0, 0)))
    _expect_film(_avg_b(gpu, (3, 128, 1), cs, (32, 0, 0)))


def test_demag_rod_pbc_golden(gpu):
    """test/demagrodpbc.mx3: <B> = (-0.496, -0.496, 0) diagonal (TOL 1e-2)."""
    k = _avg_b(gpu, (2, 2, 2), (1e-9, 1e-9, 2e-9), (0, 0, 16))
    assert np.abs(np.diag(k) - np.array([-0.496, -0.496, 0])).max() < 1e-2, k
    assert np.abs(k[~np.eye(3, dtype=bool)]).max() < 1e-2, k


@pytest.mark.parametrize("prec", [1, 0])
def test_sim_pbc_vs_oracle(gpu, oracle, prec):
    """Effective field, torque and 20 RK45DP steps of a film periodic in x and
    y (exchange, interfacial DMI, anisotropy, demag with images) vs the oracle."""
    _say("sim pbc", prec)
    N, cs, pbc = (16, 12, 2), (2e-9, 2e-9, 1.5e-9), (2, 1, 0)
    dt = np.float32 if prec == 0 else np.float64

    def make(mod, d):
        s = mod.Sim(N, cs, pbc=pbc, dtype=d)
        setup(s, mod, Ms=8e5, Aex=1.3e-11, Dind=1e-3, alpha=0.1, Ku1=1e5, anisU=(0.3, 0, 1),
              Bext=(0.01, 0, 0.02), dtype=d)
        return s
    sg, so = make(gpu, dt), make(oracle, np.float64)
    m = random_unit((N[2], N[1], N[0]), 9)
    sg.set_m(m)
    so.set_m(m)
    tol = 1e-10 if prec == 1 else 1e-5
    Bo = so.beff()
    assert np.abs(sg.beff() - Bo).max() <= tol * np.abs(Bo).max()
    for s in (sg, so):
        s.set_solver_params(dt=1e-14, fixdt=1e-14)
        s.steps(20)
    assert np.abs(sg.get_m() - so.get_m()).max() <= (1e-9 if prec == 1 else 1e-4)


def test_zhangli_pbc(gpu):
    """test/zhangliPBC.mx3: 256x32x1 strip periodic along x, head-to-head wall
    (twodomain, the left 8 columns along -y), alpha 3 for 1 ns, then Zhang-Li
    (xi 0.1, J = 1e12 A/m^2 along x, Pol 1) at alpha 0.01 for 1 ns; TOL 1e-4."""
    _say("zhangli pbc")
    N, c = (256, 32, 1), 5e-9
    cs = (c, c, c)
    s = gpu.Sim(N, cs, pbc=(1, 0, 0), dtype=np.float32)
    setup(s, gpu, Ms=800e3, Aex=13e-12, alpha=3.0, dtype=np.float32)
    m = two_domain(N, cs, (1, 0, 0), (0, 1, 0), (-1, 0, 0))
    X, _, _ = coords(N, cs)
    m[:, X < -120 * c] = np.array([0.0, -1.0, 0.0])[:, None]
    s.set_m(m)
    s.run(1e-9)
    s.set_scalar(gpu.P_ALPHA, 0.01)
    s.set_scalar(gpu.P_XI, 0.1)
    s.set_vector(gpu.P_J, (1e12, 0, 0))
    s.set_scalar(gpu.P_POL, 1.0)
    s.run(1e-9)
    got = s.average_m()
    _say("m", got)
    assert np.abs(got - np.array([-0.081425920, -0.003434650, -0.015030215])).max() < 1e-4, got


def _fmod(a, b):
    """fmod of engine/shape.go:222-231 (period b, result in [-b/2, b/2])."""
    return np.where(np.abs(a) > b / 2, np.sign(a) * (np.mod(np.abs(a + b / 2), b) - b / 2), a)


def test_pbc1_geometry(gpu):
    """test/pbc1.mx3: 128x64x1 film periodic (2, 2, 0) holding a 64x32 rectangle,
    once split over the four corners (repeated and translated by half the box)
    and once centred: the same <m> after 1 ns at alpha 1 (TOL 1e-3)."""
    _say("pbc1")
    Nx, Ny, c = 128, 64, 5e-9
    N, cs = (Nx, Ny, 1), (c, c, c)
    X, Y, _ = coords(N, cs)
    want = np.array([0.89947968, 0.23352228, -0.00010287])
    for shift in (1, 0):
        dx, dy = (Nx / 2 * c, Ny / 2 * c) if shift else (0.0, 0.0)
        xr, yr = _fmod(X - dx, Nx * c), _fmod(Y - dy, Ny * c)
        vol = ((xr < Nx / 4 * c) & (xr > -Nx / 4 * c) & (yr < Ny / 4 * c) & (yr > -Ny / 4 * c)).astype(np.float64)
        assert vol.sum() == Nx * Ny / 4
        s = gpu.Sim(N, cs, pbc=(2, 2, 0), dtype=np.float32)
        setup(s, gpu, Ms=1000e3, Aex=10e-12, alpha=1.0, vol=vol, dtype=np.float32)
        s.set_m(uniform_m(N, (1, 0.1, 0.01)))
        s.run(1e-9)
        got = s.average_m()
        _say("m", shift, got)
        assert np.abs(got - want).max() < 1e-3, (shift, got)


def test_dmi_pbc(gpu):
    """test/dmipbc.mx3: 128x64x1 film periodic along x with interfacial DMI and
    perpendicular anisotropy, up/down domains relaxed 0.1 ns at alpha 3, then
    Zhang-Li (xi 0.4, Pol 1, J = -1e12 A/m^2 along x) at alpha 0.05 for
    0.7 ns; <m> = (0, -0.1537975966930, 0), TOL 1e-4."""
    _say("dmipbc")
    n = 64
    N, cs = (2 * n, n, 1), (2.3e-9, 2.3e-9, 1.15e-9)
    s = gpu.Sim(N, cs, pbc=(1, 0, 0), dtype=np.float32)
    setup(s, gpu, Ms=6e5, Aex=20e-12, Dind=1e-3, Ku1=5.9e5, anisU=(0, 0, 1), alpha=3.0, dtype=np.float32)
    s.set_m(two_domain(N, cs, (0, 0, 1), (0, 0, 0), (0, 0, -1)))
    s.run(0.1e-9)
    s.set_scalar(gpu.P_ALPHA, 0.05)
    s.set_scalar(gpu.P_XI, 0.4)
    s.set_scalar(gpu.P_POL, 1.0)
    s.set_vector(gpu.P_J, (-1e12, 0, 0))
    s.run(0.7e-9)
    got = s.average_m()
    _say("m", got)
    assert np.abs(got - np.array([0, -0.1537975966930, 0])).max() < 1e-4, got
