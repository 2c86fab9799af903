"""More of the reference's own test scripts (test/*.mx3) run through the
simulation handle on the GPU, with the reference's expectations and
tolerances: demag factors of films and rods, exchange field region averages,
3D and rotated DMI walls, energy conservation and dissipation, Zeeman energy,
frozen spins, standard problem 4b with the brute-force tensor, the
Slonczewski / Zhang-Li standard problems 5b, 5c (2D and 3D) and their Newell
variants, inter-region exchange scaling (ext_ScaleExchange, a LUT entry here),
uniaxial and cubic anisotropy energies of one cell, the OpenBC edge canting
of a DMI chain, the minimizer, the isotropy of a sphere's energy, a thermal
equilibrium under RK4 and demag with regions.  Scripts that need other extension functions (ext_*)
or the script RNG are not included.
"""
import math
import sys

import numpy as np
import pytest

from cases import MU0, coords, setup, two_domain, uniform_m, vortex

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


def _demag_factors(gpu, N, cs, prec=0):
    dt = np.float32 if prec == 0 else np.float64
    s = gpu.Sim(N, cs, dtype=dt)
    setup(s, gpu, Ms=1 / MU0, Aex=1e-11, alpha=0.1, dtype=dt)
    out = []
    for v in np.eye(3):
        s.set_m(uniform_m(N, v))
        out.append(s.demag_field().reshape(3, -1).astype(np.float64).mean(axis=1))
    return np.array(out)       # row: m direction, column: B component


@pytest.mark.parametrize("prec", [0, 1])
def test_demag2d_film(gpu, prec):
    """test/demag2D.mx3: 128 x 128 x 1 flat cells, kxx, kyy ~ 0 (1e-2), kzz ~ -1 (2e-2), off-diagonal 1e-9."""
    k = _demag_factors(gpu, (128, 128, 1), (1e-9, 1e-9, 0.5e-9), prec)
    assert np.abs(k[~np.eye(3, dtype=bool)]).max() < 1e-9, k
    assert abs(k[0, 0]) < 1e-2 and abs(k[1, 1]) < 1e-2 and abs(k[2, 2] + 1) < 2e-2, k


def test_demag3d_film(gpu):
    """test/demag3Dfilm.mx3: 64 x 48 x 2; kxx, kyy ~ 0 (3e-2), kzz ~ -1 (6e-2)."""
    k = _demag_factors(gpu, (64, 48, 2), (1e-9, 1.2e-9, 0.5e-9))
    assert np.abs(k[~np.eye(3, dtype=bool)]).max() < 1e-9, k
    assert abs(k[0, 0]) < 3e-2 and abs(k[1, 1]) < 3e-2 and abs(k[2, 2] + 1) < 6e-2, k


def test_demag_rod(gpu):
    """test/demagrod.mx3: 2 x 2 x 64 cells of 1 x 1 x 2 nm: diag (-0.496, -0.496, 0), 1e-2."""
    k = _demag_factors(gpu, (2, 2, 64), (1e-9, 1e-9, 2e-9))
    assert np.abs(np.diag(k) - np.array([-0.496, -0.496, 0])).max() < 1e-2, k
    assert np.abs(k[~np.eye(3, dtype=bool)]).max() < 1e-2, k


@pytest.mark.parametrize("prec", [0, 1])
def test_exchange_region_averages(gpu, prec):
    """test/exchange.mx3: 300 x 200 x 4 cells of 2 x 3 x 4 nm, twodomain(1,2,3,
    -1,-3,5, 3,-1,2), demag off; <B_exch> = 0 and its averages over x < 0 and
    x >= 0 (tolerance 1.2e-8 T)."""
    _say("exchange.mx3", prec)
    N, cs = (300, 200, 4), (2e-9, 3e-9, 4e-9)
    dt = np.float32 if prec == 0 else np.float64
    s = gpu.Sim(N, cs, dtype=dt)
    setup(s, gpu, Ms=800e3, Aex=13e-12, alpha=0.1, demag=False, dtype=dt)
    s.set_m(two_domain(N, cs, (1, 2, 3), (-1, -3, 5), (3, -1, 2)))
    B = s.beff().astype(np.float64)
    X, _, _ = coords(N, cs)
    tol = 1.2e-8
    assert np.abs(B.reshape(3, -1).mean(axis=1)).max() < tol
    b0 = B[:, X < 0].mean(axis=1)
    b1 = B[:, X >= 0].mean(axis=1)
    assert np.abs(b0 - np.array([0.00118888104756673, -0.00164631487528483, -0.00074591267903645])).max() < tol, b0
    assert np.abs(b1 - np.array([-0.00118888041178385, 0.00164631512959798, 0.00074591242472330])).max() < tol, b1


def test_exchange3d_vortex_relax(gpu):
    """test/exchange3d.mx3: 32 x 32 x 8 cubes of 100/32 nm, vortex(1, 1), relax():
    <mz> = 0.02785073 (1e-5)."""
    N = (32, 32, 8)
    c = 100e-9 / 32
    s = gpu.Sim(N, (c, c, c), dtype=np.float32)
    setup(s, gpu, Ms=800e3, Aex=13e-12, alpha=3.0, dtype=np.float32)
    s.set_m(vortex(N, (c, c, c), 1, 1))
    s.relax()
    mz = s.average_m()[2]
    _say("mz", mz)
    assert abs(mz - 0.02785073) < 1e-5, mz


def test_dmi3d_wall(gpu):
    """test/dmi3d.mx3: 64 x 64 x 2 cells, interfacial DMI turns a Bloch wall
    towards Neel; m at cell (32, 32, 0) after 0.5 ns = (0.8979, 0.3989, 0.1859), 1e-3."""
    N, c = (64, 64, 2), 0.6e-9
    cs = (c, c, c / 2)
    s = gpu.Sim(N, cs, dtype=np.float32)
    setup(s, gpu, Ms=1100e3, Aex=16e-12, alpha=3.0, Ku1=1.27e6, anisU=(0, 0, 1), Dind=0.1e-3, dtype=np.float32)
    s.set_m(two_domain(N, cs, (0, 0, -1), (1, 1, 0), (0, 0, 1)))
    s.run(0.5e-9)
    got = s.get_m()[:, 0, 32, 32]
    _say("m(32,32,0)", got)
    assert np.abs(got - np.array([0.8979, 0.3989, 0.1859])).max() < 1e-3, got


def _rotz(N, cs, cfg, theta):
    """Config.RotZ (engine/config.go:207-218) of a two-domain configuration."""
    X, Y, Z = coords(N, cs)
    co, si = math.cos(theta), math.sin(theta)
    Xr, Yr = X * co + Y * si, -X * si + Y * co
    m1, mw, m2 = cfg
    ww = 2 * cs[0]
    gauss = np.exp(-(Xr / ww) ** 2)
    m = np.empty((3,) + X.shape)
    for c in range(3):
        base = np.where(Xr < 0, m1[c], m2[c])
        m[c] = (1 - gauss) * base + gauss * mw[c]
    mx, my = m[0] * co - m[1] * si, m[0] * si + m[1] * co
    return np.stack([mx, my, m[2]])


def test_dmi_y_wall(gpu):
    """test/dmiY.mx3: test/dmi.mx3 rotated by 90 degrees (y derivatives): 128 x 128 x 1,
    MaxDt 5e-13, 5 ns; <m> = (-0.035, 0.059) within 0.01."""
    N = (128, 128, 1)
    cs = (250e-9 / 128, 250e-9 / 128, 0.6e-9)
    s = gpu.Sim(N, cs, dtype=np.float32)
    setup(s, gpu, Ms=1100e3, Aex=16e-12, alpha=3.0, Ku1=1.27e6, anisU=(0, 0, 1), Dind=0.1e-3, dtype=np.float32)
    s.set_solver_params(maxdt=5e-13)
    s.set_m(_rotz(N, cs, ((0, 0, -1), (1, 1, 0), (0, 0, 1)), math.pi / 2))
    s.run(5e-9)
    got = s.average_m()
    _say("m", got)
    assert abs(got[0] + 0.035) < 0.01 and abs(got[1] - 0.059) < 0.01, got


@pytest.mark.parametrize("prec", [0, 1])
def test_anisotropy_energy_conservation(gpu, prec):
    """test/anisenergy.mx3: 32 x 10 x 2, no demag, Ku1 along z: the energy change
    without damping over 1 ns is below 1e-6 of the damped one."""
    _say("anisenergy", prec)
    N, c = (32, 10, 2), 1e-9
    dt = np.float32 if prec == 0 else np.float64
    s = gpu.Sim(N, (c, 2 * c, 3 * c), dtype=dt)
    setup(s, gpu, Ms=1000e3, Aex=10e-12, alpha=1.0, Ku1=1e6, anisU=(0, 0, 1), demag=False, dtype=dt)
    s.set_m(uniform_m(N, (1, 0, 0.1)))
    e0 = s.energy(9)
    s.run(1e-9)
    d1 = s.energy(9) - e0
    s.set_m(uniform_m(N, (1, 0, 0.1)))
    e0 = s.energy(9)
    s.set_scalar(gpu.P_ALPHA, 0.0)
    s.run(1e-9)
    d2 = s.energy(9) - e0
    _say("dE", d1, d2)
    assert abs(d2 / d1) < 1e-6, (d1, d2)


def test_zeeman_energy(gpu):
    """test/zeemanenergy.mx3: E_zeeman = -B Ms V for B along m, 0 across, +B Ms V against (tol B Ms V / 1e5)."""
    N, cs = (128, 32, 2), (5e-9, 4e-9, 3e-9)
    V = np.prod(N) * np.prod(cs)
    Ms, B = 100e3, 1e-3
    s = gpu.Sim(N, cs, dtype=np.float32)
    setup(s, gpu, Ms=Ms, Aex=1e-11, alpha=0.1, dtype=np.float32)
    s.set_m(uniform_m(N, (1, 0, 0)))
    tol = B * Ms * V / 1e5
    for b, want in (((B, 0, 0), -B * Ms * V), ((0, B, 0), 0.0), ((-B, 0, 0), B * Ms * V)):
        s.set_vector(gpu.P_BEXT, b)
        assert abs(s.energy(2) - want) < tol, (b, s.energy(2), want)


@pytest.mark.parametrize("prec", [0, 1])
def test_dmi_energy_dissipation(gpu, prec):
    """test/dmienergy.mx3: 64 x 64 x 1 film with DMI 3e-3, anisotropy and demag,
    alpha 0.01 from uniform(1,1,1): E_total after 10 and 20 ps minus the start,
    -2.62692723e-20 and -4.93337113e-20 J (tolerance 1e-23 J)."""
    _say("dmienergy", prec)
    N = (64, 64, 1)
    cs = (1.15e-9 / 2, 1.15e-9 / 2, 1.15e-9)
    dt = np.float32 if prec == 0 else np.float64
    s = gpu.Sim(N, cs, dtype=dt)
    setup(s, gpu, Ms=6e5, Aex=20e-12, alpha=0.01, Ku1=5.9e5, anisU=(0, 0, 1), Dind=3e-3, dtype=dt)
    s.set_m(uniform_m(N, (1, 1, 1)))
    e0 = s.energy(9)
    s.run(10e-12)
    e1 = s.energy(9)
    s.run(10e-12)
    e2 = s.energy(9)
    _say("dE", e1 - e0, e2 - e0)
    assert abs((e1 - e0) - -2.62692723e-20) < 0.001e-20, e1 - e0
    assert abs((e2 - e0) - -4.93337113e-20) < 0.001e-20, e2 - e0


def test_frozen_spins(gpu):
    """test/frozenspins.mx3: standard problem 4 (3D) with every spin frozen:
    relax() and 1 ns in the reversal field leave <m> at its start (1e-4)."""
    N = (128, 32, 2)
    cs = (500e-9 / 128, 125e-9 / 32, 3e-9 / 2)
    s = gpu.Sim(N, cs, dtype=np.float32)
    setup(s, gpu, Ms=800e3, Aex=13e-12, alpha=0.02, frozen=np.ones(256, np.float32), dtype=np.float32)
    s.set_m(uniform_m(N, (1, 0.1, 0)))
    want = np.array([0.9950371980667114, 0.09950374066829681, 0])
    s.relax()
    assert np.abs(s.average_m() - want).max() < 1e-4, s.average_m()
    s.set_vector(gpu.P_BEXT, (-24.6e-3, 4.3e-3, 0))
    s.run(1e-9)
    assert np.abs(s.average_m() - want).max() < 1e-4, s.average_m()


@pytest.mark.parametrize("prec", [0, 1])
def test_standard_problem4b(gpu, prec):
    """test/standardproblem4b.mx3: 128 x 32 x 1 (brute-force tensor), relax() then
    B_ext = (-35.5, -6.3, 0) mT for 1 ns; TOL 1e-3."""
    _say("sp4b", prec)
    N = (128, 32, 1)
    dt = np.float32 if prec == 0 else np.float64
    s = gpu.Sim(N, (500e-9 / 128, 125e-9 / 32, 3e-9), dtype=dt)
    setup(s, gpu, Ms=800e3, Aex=13e-12, alpha=0.02, dtype=dt)
    s.set_m(uniform_m(N, (1, 0.1, 0)))
    s.relax()
    got = s.average_m()
    assert np.abs(got - np.array([0.9669684171676636, 0.1252732276916504, 0])).max() < 1e-3, got
    s.set_vector(gpu.P_BEXT, (-35.5e-3, -6.3e-3, 0))
    s.run(1e-9)
    got = s.average_m()
    _say("m", got)
    assert np.abs(got - np.array([-0.9692331552505493, -0.11889542198181152, -0.0053076110780239105])).max() < 1e-3


def _sp5bc(gpu, N, dtype, epsp, lam=2.0):
    """The Slonczewski film of test/std5b.mx3 / std5c.mx3: 160 x 80 x 5 nm,
    I = -6 mA through the film area, fixed layer at 20 degrees in-plane."""
    length, width, thick = 160e-9, 80e-9, 5e-9
    s = gpu.Sim(N, (length / N[0], width / N[1], thick / N[2]), dtype=dtype)
    setup(s, gpu, Ms=800e3, Aex=13e-12, alpha=0.01, dtype=dtype)
    s.set_scalar(gpu.P_POL, 0.5669)
    s.set_scalar(gpu.P_LAMBDA, lam)
    s.set_scalar(gpu.P_EPSPRIME, epsp)
    th = math.pi * 20 / 180
    s.set_vector(gpu.P_FIXEDLAYER, (math.cos(th), math.sin(th), 0))
    s.set_vector(gpu.P_J, (0, 0, -0.006 / (length * width)))
    return s


@pytest.mark.parametrize("prec", [0, 1])
def test_std5b_epsilon_prime(gpu, prec):
    """test/std5b.mx3: RK45DP with MaxErr 1e-6, MinDt 1e-18, MaxDt 1e-12,
    gammaLL 2.211e5/mu0, DemagAccuracy 29, epsilon' = 1: <m> after 1 ns vs OOMMF
    (-0.953323544827031, -0.279948071263437, 0.00528093273512820), TOL 1e-5."""
    _say("std5b", prec)
    N = (64, 32, 1)
    dt = np.float32 if prec == 0 else np.float64
    s = _sp5bc(gpu, N, dt, 1.0)
    s.set_demag_accuracy(29)
    s.set_solver_params(maxerr=1e-6, mindt=1e-18, maxdt=1e-12, gammaLL=2.211e5 / MU0)
    s.set_m(uniform_m(N, (1, 0.1, 0)))
    s.run(1e-9)
    got = s.average_m()
    _say("m", got)
    assert np.abs(got - np.array([-0.953323544827031, -0.279948071263437, 0.00528093273512820])).max() < 1e-5, got


@pytest.mark.parametrize("N,tol", [((64, 32, 1), 1e-5), ((64, 32, 2), 1e-2)])
def test_std5c(gpu, N, tol):
    """test/std5c.mx3 (2D, TOL 1e-5) and std5c3d.mx3 (3D, TOL 1e-2): epsilon' = 0,
    0.5 ns from uniform(1, 0, 0) vs OOMMF (0.9056, -0.2861, -0.0110)."""
    _say("std5c", N)
    s = _sp5bc(gpu, N, np.float32, 0.0)
    s.set_m(uniform_m(N, (1, 0, 0)))
    s.run(0.5e-9)
    got = s.average_m()
    _say("m", got)
    assert np.abs(got - np.array([0.905612587928772, -0.2860856354236603, -0.011005667969584465])).max() < tol, got


def test_standard_problem5_newell(gpu):
    """test/standardproblem5_newell.mx3: standard problem 5 with the Newell tensor; TOL 1e-4."""
    N, cs = (32, 32, 4), (100e-9 / 32, 100e-9 / 32, 10e-9 / 4)
    s = gpu.Sim(N, cs, dtype=np.float32)
    setup(s, gpu, Ms=800e3, Aex=13e-12, alpha=0.1, dtype=np.float32)
    s.set_demag_newell(1)
    s.set_m(vortex(N, cs, 1, 1))
    s.relax()
    s.set_scalar(gpu.P_XI, 0.05)
    s.set_vector(gpu.P_J, (1e12, 0, 0))
    s.set_scalar(gpu.P_POL, 1.0)
    s.run(1e-9)
    got = s.average_m()
    _say("m", got)
    assert np.abs(got - np.array([-0.234637409, -0.09450896, 0.02294345])).max() < 1e-4, got


def test_standard_problem5a_newell(gpu):
    """test/standardproblem5a_newell.mx3: minimize() with the Newell tensor, then
    Slonczewski reversal for 1 ns; TOL 1e-3."""
    N = (64, 32, 1)
    sX, sY, sZ = 160e-9, 80e-9, 5e-9
    s = gpu.Sim(N, (sX / N[0], sY / N[1], sZ / N[2]), dtype=np.float32)
    setup(s, gpu, Ms=800e3, Aex=13e-12, alpha=3.0, dtype=np.float32)
    s.set_demag_newell(1)
    s.set_m(uniform_m(N, (1, 1, 0.001)))
    s.minimize()
    got = s.average_m()
    assert np.abs(got - np.array([0.9585264325141907, 0.20422574877738953, 0])).max() < 1e-3, got
    s.set_scalar(gpu.P_ALPHA, 0.01)
    s.set_scalar(gpu.P_LAMBDA, 1.0)
    s.set_scalar(gpu.P_POL, 0.5669)
    s.set_scalar(gpu.P_EPSPRIME, 0.0)
    prad = math.pi / 180
    s.set_vector(gpu.P_FIXEDLAYER, (math.cos(prad), math.sin(prad), 0))
    s.set_vector(gpu.P_J, (0, 0, -0.008 / (sX * sY)))
    s.run(1e-9)
    got = s.average_m()
    _say("m", got)
    assert np.abs(got - np.array([0.6427677195465802, 0.5114272236824036, -0.1541234254837036])).max() < 1e-3, got


def _sp4_regions(gpu, scale, alpha):
    """128 x 32 x 1 standard-problem-4 film split into regions 1 (y >= 0) and 2
    (y < 0) with the inter-region exchange scaled (ext_ScaleExchange)."""
    from umagnus_amd.params import exchange_lut
    N, cs = (128, 32, 1), (500e-9 / 128, 125e-9 / 32, 3e-9)
    s = gpu.Sim(N, cs, dtype=np.float32)
    _, Y, _ = coords(N, cs)
    reg = np.where(Y >= 0, 1, 2).astype(np.uint8)
    setup(s, gpu, Ms=800e3, Aex=13e-12, alpha=alpha, regions=reg, dtype=np.float32)
    s.set_exchange_luts(exchange_lut(13e-12, scale={(1, 2): scale}, dtype=np.float32), None)
    s.set_m(uniform_m(N, (1, 0.1, 0)))
    return s


def test_interexchange_afm(gpu):
    """test/interexchange.mx3: the two halves coupled antiferromagnetically
    (scale -1), alpha 3, 10 ns: <m> = (0, 0.114170685410499, 0), 1e-5."""
    s = _sp4_regions(gpu, -1.0, 3.0)
    s.run(10e-9)
    got = s.average_m()
    _say("m", got)
    assert np.abs(got - np.array([0, 0.114170685410499, 0])).max() < 1e-5, got


def test_interexchange_scale_one(gpu):
    """test/interexchange2.mx3: scale 1 reproduces standard problem 4; TOL 1e-3."""
    s = _sp4_regions(gpu, 1.0, 0.02)
    s.relax()
    assert np.abs(s.average_m() - np.array([0.96696, 0.12528, 0])).max() < 1e-3, s.average_m()
    s.set_vector(gpu.P_BEXT, (-24.6e-3, 4.3e-3, 0))
    s.run(1e-9)
    got = s.average_m()
    _say("m", got)
    assert np.abs(got - np.array([-0.9845, 0.1268, 0.0432])).max() < 1e-3, got


def test_uniaxial_anisotropy_relax(gpu):
    """test/uniaxialanisotropy.mx3: easy axis along x against a field along y,
    relax() at each field, <my> vs OOMMF (TOL 1e-3); regions 2 and 3 are
    defined but carry the same parameters."""
    N, cs = (64, 64, 1), (4e-9, 4e-9, 2e-9)
    s = gpu.Sim(N, cs, dtype=np.float32)
    X, Y, _ = coords(N, cs)
    reg = np.zeros(X.shape, np.uint8)
    reg[(X / 100e-9) ** 2 + (Y / 100e-9) ** 2 <= 0.25] = 2
    reg[(np.abs(X) < 50e-9) & (np.abs(Y) < 10e-9)] = 3
    setup(s, gpu, Ms=1100e3, Aex=13e-12, alpha=1.0, Ku1=0.5e6, anisU=(1, 0, 0), regions=reg, dtype=np.float32)
    s.set_m(uniform_m(N, (1, 1, 0)))
    for B, my in ((0.0, 0.0), (0.01, 0.011), (0.03, 0.033), (0.10, 0.110), (0.30, 0.331)):
        s.set_vector(gpu.P_BEXT, (0, B, 0))
        s.relax()
        assert abs(s.average_m()[1] - my) < 1e-3, (B, s.average_m())


def _one_cell(gpu, m):
    s = gpu.Sim((1, 1, 1), (1e-9,) * 3, dtype=np.float32)
    setup(s, gpu, Ms=1000e3, Aex=0.0, alpha=0.1, dtype=np.float32)
    s.set_m(uniform_m((1, 1, 1), m))
    return s


def test_uniaxial_energy_one_cell(gpu):
    """test/uniaxial_full.mx3: one cell at 17 degrees from the x axis; E_anis / V
    for Ku2 = +-1e6 and Ku1 = +-1e6 (TOL 1 J/m^3)."""
    V = 1e-27
    th = 17 * math.pi / 180
    s = _one_cell(gpu, (math.cos(th), math.sin(th), 0))
    s.set_vector(gpu.P_ANISU, (1, 0, 0))
    ean = lambda: sum(s.energy(t) for t in (3, 4, 5, 6, 7)) / V
    for k1, k2, want in ((0, 1e6, -836344.9375), (0, -1e6, 836344.9375), (1e6, 0, -914519), (-1e6, 0, 914519)):
        s.set_scalar(gpu.P_KU1, k1)
        s.set_scalar(gpu.P_KU2, k2)
        assert abs(ean() - want) < 1, (k1, k2, ean())


def test_cubic_energy_one_cell(gpu):
    """test/cubic_full.mx3: one cell at (theta, phi) = (23, 11) degrees; E_anis / V
    for each of Kc1, Kc2, Kc3 = +-1e6 (TOL 1e-2 J/m^3)."""
    V = 1e-27
    th, ph = 23 * math.pi / 180, 11 * math.pi / 180
    s = _one_cell(gpu, (math.sin(ph) * math.cos(th), math.sin(ph) * math.sin(th), math.cos(ph)))
    s.set_vector(gpu.P_ANISC1, (1, 0, 0))
    s.set_vector(gpu.P_ANISC2, (0, 1, 0))
    ean = lambda: sum(s.energy(t) for t in (3, 4, 5, 6, 7)) / V
    cases = (((1e6, 0, 0), 35254.007), ((-1e6, 0, 0), -35254.007), ((0, 1e6, 0), 165.23309326171875),
             ((0, -1e6, 0), -165.23309326171875), ((0, 0, 1e6), 912.3794555664062))
    for (k1, k2, k3), want in cases:
        s.set_scalar(gpu.P_KC1, k1)
        s.set_scalar(gpu.P_KC2, k2)
        s.set_scalar(gpu.P_KC3, k3)
        assert abs(ean() - want) < 1e-2, ((k1, k2, k3), ean())


@pytest.mark.parametrize("axis", [0, 1])
def test_open_bc_edge_canting(gpu, axis):
    """test/openbc.mx3: a chain of 1024 cells (along x, then along y) with DMI at
    90% of the critical strength, minimized (MinimizerStop 1e-7): with OpenBC the
    edge canting matches the 1D analytic value (1e-3) and the periodic chain's
    (1e-5), in the reduced units of the script."""
    from umagnus_amd.params import exchange_lut
    ncell, c = 1024, 0.05
    N = (ncell, 1, 1) if axis == 0 else (1, ncell, 1)
    D = 0.9 * 4 / math.pi

    def make(pbc, openbc):
        s = gpu.Sim(N, (c, c, c), pbc=pbc, dtype=np.float64)
        setup(s, gpu, Ms=1.0, Aex=1.0, alpha=0.1, Ku1=1.0, anisU=(0, 0, 1), Dind=D, demag=False, openbc=openbc,
              dtype=np.float64)
        return s

    def cant(s):
        m = s.get_m()[:, 0, 0, 0]
        return math.atan2(m[axis], m[2])
    theta0 = math.asin(D / 2)
    analytic = 2 * math.atan(math.exp(-c / 2) * math.tan(theta0 / 2))
    s = make((0, 0, 0), True)
    s.set_m(uniform_m(N, (0, 0, 1)))
    s.minimize(stop=1e-7)
    c_open = cant(s)
    sp = make((0, 1, 0) if axis == 0 else (1, 0, 0), False)
    sp.set_m(s.get_m())
    sp.minimize(stop=1e-7)
    c_per = cant(sp)
    _say("cant", c_open, c_per, analytic)
    assert abs(c_open - analytic) < 1e-3
    assert abs(c_open - c_per) < 1e-5


def test_minimizer_sp4(gpu):
    """test/minimizer.mx3: standard problem 4 (2D) minimized from uniform(1, .1, 0)
    after a Msat change (the exchange LUT must not keep the old Msat):
    <m> = (0.9669684171676636, 0.1252817150004892, 0), TOL 1e-5."""
    N = (128, 32, 1)
    s = gpu.Sim(N, (500e-9 / 128, 125e-9 / 32, 3e-9), dtype=np.float32)
    setup(s, gpu, Ms=1600e3, Aex=13e-12, alpha=0.02, dtype=np.float32)
    s.set_m(uniform_m(N, (1, 0.1, 0)))
    s.energy(9)
    s.set_scalar(gpu.P_MSAT, 800e3)
    s.minimize()
    got = s.average_m()
    _say("m", got, s.state()["nsteps"])
    assert np.abs(got - np.array([0.9669684171676636, 0.1252817150004892, 0])).max() < 1e-5, got


@pytest.mark.parametrize("prec", [0, 1])
def test_isotropic_energy_sphere(gpu, prec):
    """test/isotropicEnergy.mx3: a sphere of 32^3 cubes (Ellipsoid geometry): the
    total energy of uniform states does not depend on their direction (1e-6)."""
    _say("isotropic", prec)
    N, c = (32, 32, 32), 2e-9
    cs = (c, c, c)
    dt = np.float32 if prec == 0 else np.float64
    X, Y, Z = coords(N, cs)
    vol = ((X / (N[0] * c)) ** 2 + (Y / (N[1] * c)) ** 2 + (Z / (N[2] * c)) ** 2 <= 0.25).astype(np.float64)
    s = gpu.Sim(N, cs, dtype=dt)
    setup(s, gpu, Ms=860e3, Aex=13e-12, alpha=1e-5, vol=vol, dtype=dt)
    s.set_m(uniform_m(N, (1, 0, 0)))
    e0 = s.energy(9)
    phi = 0.0
    worst = 0.0
    while phi < 4 * math.pi:
        th = phi * 1.234
        s.set_m(uniform_m(N, (math.sin(phi) * math.cos(th), math.cos(phi) * math.cos(th), math.sin(th))))
        worst = max(worst, abs((s.energy(9) - e0) / e0))
        phi += 0.1
    _say("worst", worst)
    assert worst < 1e-6, worst


def test_rk4_temperature(gpu):
    """test/rk4temperature.mx3: 768 x 768 x 1 cells of 10 nm, no exchange or
    demag, Ku1 = 2500 along z, T = 100 K, RK4 with FixDt 2e-12 for 5 ns:
    <m> = (0, 0, 0.8366) within 1e-3 (a statistical equilibrium: the Philox
    stream differs from the reference's generator)."""
    N, c = (768, 768, 1), 10e-9
    s = gpu.Sim(N, (c, c, c), dtype=np.float32)
    setup(s, gpu, Ms=1e6, Aex=0.0, alpha=0.1, Ku1=1e4 / 4, anisU=(0, 0, 1), temp=100.0, seed=7, demag=False,
          dtype=np.float32)
    s.set_solver(gpu.SOLVER_RK4)
    s.set_solver_params(dt=2e-12, fixdt=2e-12)
    s.set_m(uniform_m(N, (0, 0, 1)))
    s.run(5e-9)
    got = s.average_m()
    _say("m", got)
    assert np.abs(got - np.array([0, 0, 0.8366])).max() < 1e-3, got


def test_demag_regions(gpu):
    """test/demagRegion.mx3: 64 x 64 x 2 with three regions of Msat 2/mu0:
    kxx, kyy ~ 0 and kzz ~ -2 (0.1), off-diagonal 1e-8; Msat 1/mu0: kzz ~ -1."""
    N, cs = (64, 64, 2), (1e-9, 1e-9, 0.5e-9)
    X, Y, _ = coords(N, cs)
    reg = np.where(X >= 0, 1, 2).astype(np.uint8)
    reg[(X / 10e-9) ** 2 + (Y / 20e-9) ** 2 <= 0.25] = 3
    s = gpu.Sim(N, cs, dtype=np.float32)
    setup(s, gpu, Ms=2 / MU0, Aex=1e-11, alpha=0.1, regions=reg, dtype=np.float32)
    k = []
    for v in np.eye(3):
        s.set_m(uniform_m(N, v))
        k.append(s.demag_field().reshape(3, -1).astype(np.float64).mean(axis=1))
    k = np.array(k)
    assert np.abs(k[~np.eye(3, dtype=bool)]).max() < 1e-8, k
    assert abs(k[0, 0]) < 0.1 and abs(k[1, 1]) < 0.1 and abs(k[2, 2] + 2) < 0.1, k
    s.set_scalar(gpu.P_MSAT, 1 / MU0)
    assert abs(s.demag_field()[2].astype(np.float64).mean() + 1) < 0.1


@pytest.mark.parametrize("prec", [0, 1])
def test_nodemagspins(gpu, oracle, prec):
    """test/nodemagspins.mx3: NoDemagSpins per region (flagged cells neither
    generate nor receive the demag field), with and without an air gap; then
    B_eff and torque of a random state with regions, geometry and flags vs the
    oracle (FP64 1e-11, FP32 1e-5)."""
    from cases import nodemag_script, random_unit
    dt = np.float32 if prec == 0 else np.float64
    for label, got, want, tol in nodemag_script(gpu, dt):
        assert np.abs(got - want).max() <= tol, (label, got, want)
    N, cs = (24, 16, 3), (2e-9, 3e-9, 2e-9)
    rng = np.random.default_rng(3)
    reg = rng.integers(0, 4, (N[2], N[1], N[0])).astype(np.uint8)
    vol = (rng.uniform(size=reg.shape) > 0.1).astype(np.float64)
    flags = np.zeros(256)
    flags[[1, 3]] = 1
    m = random_unit((N[2], N[1], N[0]), 8)
    sims = []
    for mod, d in ((gpu, dt), (oracle, np.float64)):
        s = mod.Sim(N, cs, dtype=d)
        setup(s, mod, Ms=7e5, Aex=1.1e-11, alpha=0.1, Ku1=2e4, anisU=(0, 1, 1), regions=reg, vol=vol, dtype=d)
        s.set_nodemag_lut(flags)
        s.set_m(m)
        sims.append(s)
    tol = 1e-11 if prec == 1 else 1e-5
    for f in ("demag_field", "beff", "torque"):
        a, b = getattr(sims[0], f)(), getattr(sims[1], f)()
        assert np.abs(a - b).max() <= tol * np.abs(b).max(), f


def _timedep_film(gpu, vol=None):
    N, c = (32, 32, 1), 4e-9
    s = gpu.Sim(N, (c, c, c), dtype=np.float32)
    setup(s, gpu, Ms=860e3, Aex=13e-12, alpha=0.2, anisU=(0, 1, 0), vol=vol, dtype=np.float32)
    s.set_m(uniform_m(N, (1, 1, 0)))
    return s, N, (c, c, c)


@pytest.mark.parametrize("script", ["timedep", "timedepRegion"])
def test_time_dependent_anisotropy(gpu, script):
    """test/timedep.mx3 (Ku1 = 1e5 sin(2 pi f t), f = 1 GHz, 0.5 ns: <m> =
    (0, 0.9909376502037048, 0), TOL 1e-5) and timedepRegion.mx3 (the same
    function set on region 1, which covers the grid; TOL 1e-4)."""
    s, N, _ = _timedep_film(gpu)
    if script == "timedepRegion":
        s.set_regions(np.ones((N[2], N[1], N[0]), np.uint8))
    f = 1e9
    s.set_time_function(gpu.P_KU1, lambda t: 1e5 * math.sin(2 * math.pi * f * t))
    s.run(0.5e-9)
    got = s.average_m()
    _say(script, got)
    want, tol = (np.array([0, 0.9909376502037048, 0]), 1e-5) if script == "timedep" else \
        (np.array([0, 0.99090, 0]), 1e-4)
    assert np.abs(got - want).max() < tol, got


def test_time_dependent_field(gpu):
    """test/timedep3.mx3: rotating B_ext = A (sin, cos, 0)(2 pi f t) on a disc,
    0.2 ns: <m> = (0.6773565132629695, 0.7201919931496306, 0.02121575360227688), TOL 1e-5."""
    from cases import ellipse_geom
    N, c = (32, 32, 1), 4e-9
    s, N, cs = _timedep_film(gpu, vol=ellipse_geom(N, (c, c, c), 32 * c, 32 * c))
    s.set_scalar(gpu.P_KU1, 0.0)
    f, A = 1e9, 0.01
    s.set_time_function(gpu.P_BEXT, lambda t: (A * math.sin(2 * math.pi * f * t), A * math.cos(2 * math.pi * f * t),
                                               0.0), ncomp=3)
    s.run(0.2e-9)
    got = s.average_m()
    _say("timedep3", got)
    assert np.abs(got - np.array([0.6773565132629695, 0.7201919931496306, 0.02121575360227688])).max() < 1e-5, got
