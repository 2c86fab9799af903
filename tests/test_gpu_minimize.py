"""Minimize() (engine/minimizer.go) on the GPU: parity with the oracle's
restatement, and the reference's minimize goldens
(test/standardproblem4-3d-minimize.mx3, standardproblem4_newell-3d-minimize.mx3,
uniaxialanisotropy-minimize.mx3) plus test/standardproblem5a.mx3, whose
Slonczewski reversal starts from a minimized C-state.
"""
import sys

import numpy as np
import pytest

from cases import random_unit, setup, uniform_m

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


@pytest.mark.parametrize("prec", [1, 0])
def test_minimize_vs_oracle(gpu, oracle, prec):
    """Same descent on a 24 x 16 x 2 film with demag, exchange, interfacial DMI,
    uniaxial anisotropy and a field.  In FP64 the first 25 Barzilai-Borwein
    steps from a random state agree to 2e-8: the method is non-monotone and
    amplifies rounding, so that the oracle itself, started from m perturbed
    by 1e-15 relative, ends 2e-9 to 5e-9 away after these 25 steps, while
    a 1% change of Dind ends 1.2e-2 away.  From a smooth state the converged minima
    agree in <m> (1e-4, the stopping tolerance) and energy.  A random start
    is not compared in FP32: its first steps flip cells, and FP32 step
    lengths send them elsewhere."""
    _say("minimize parity", prec)
    N, cs = (24, 16, 2), (3e-9, 3e-9, 2e-9)
    dt = np.float32 if prec == 0 else np.float64
    params = dict(Ms=8e5, Aex=1.3e-11, Dind=1e-3, alpha=0.5, Ku1=2e5, anisU=(0.2, 0.1, 1.0), Bext=(0.02, -0.01, 0))

    def make(mod, d, m):
        s = mod.Sim(N, cs, dtype=d)
        setup(s, mod, dtype=d, **params)
        s.set_m(m)
        return s
    if prec == 1:   # FP32 step lengths differ from the first wild steps of a random start
        m = random_unit((N[2], N[1], N[0]), 6)
        sg, so = make(gpu, dt, m), make(oracle, np.float64, m)
        for s in (sg, so):
            s.minimize(stop=1e30, samples=25)      # exactly 25 steps
        assert sg.state()["nsteps"] == so.state()["nsteps"] == 25
        assert sg.state()["t"] == 0.0
        assert np.abs(sg.get_m() - so.get_m()).max() < 2e-8
    # converged minima (stop: max|dm| <= 1e-6 over 10 steps)
    m = uniform_m(N, (1, 0.3, 0.1))
    sg, so = make(gpu, dt, m), make(oracle, np.float64, m)
    for s in (sg, so):
        s.minimize()
    _say("steps", sg.state()["nsteps"], so.state()["nsteps"])
    assert np.abs(sg.average_m() - so.average_m()).max() < 1e-4
    assert abs(sg.energy(9) - so.energy(9)) <= (1e-6 if prec == 1 else 1e-4) * abs(so.energy(9))


SP4_CS3 = (500e-9 / 128, 125e-9 / 32, 3e-9 / 2)


@pytest.mark.parametrize("newell", [0, 1])
def test_standard_problem4_3d_minimize(gpu, newell):
    """test/standardproblem4-3d-minimize.mx3 and standardproblem4_newell-3d-minimize.mx3:
    minimize() from uniform(1, .1, 0); <m> = (0.9669952, 0.12521563, 0), TOL 1e-3."""
    _say("sp4 minimize", newell)
    N = (128, 32, 2)
    s = gpu.Sim(N, SP4_CS3, dtype=np.float32)
    setup(s, gpu, Ms=800e3, Aex=13e-12, alpha=0.02, dtype=np.float32)
    if newell:
        s.set_demag_newell(1)
    s.set_m(uniform_m(N, (1, 0.1, 0)))
    s.minimize()
    got = s.average_m()
    _say("m", got, s.state()["nsteps"])
    assert np.abs(got - np.array([0.9669952, 0.12521563, 0])).max() < 1e-3, got


def test_uniaxial_anisotropy_minimize(gpu):
    """test/uniaxialanisotropy-minimize.mx3: 64 x 64 x 1 film, anisotropy against
    a field, minimized m vs OOMMF values (TOL 1e-3)."""
    N, cs = (64, 64, 1), (4e-9, 4e-9, 2e-9)
    s = gpu.Sim(N, cs, dtype=np.float32)
    setup(s, gpu, Ms=1100e3, Aex=13e-12, alpha=1.0, Ku1=0.5e6, anisU=(1, 0, 0), dtype=np.float32)
    s.set_m(uniform_m(N, (1, 1, 0)))
    # easy axis in-plane
    for B, my in ((0.0, 0.0), (0.01, 0.011), (0.03, 0.033), (0.10, 0.110), (0.30, 0.331)):
        s.set_vector(gpu.P_BEXT, (0, B, 0))
        s.minimize()
        got = s.average_m()
        _say("easy", B, got)
        assert abs(got[1] - my) < 1e-3, (B, got)
    # hard axis in-plane
    s.set_scalar(gpu.P_KU1, -0.5e6)
    s.set_m(uniform_m(N, (-1, -2, -3)))
    s.set_vector(gpu.P_BEXT, (0, 0, 0))
    s.minimize()
    assert abs(s.average_m()[1] - 1.0) < 1e-3, s.average_m()
    for B, mx in ((0.01, 0.011), (0.10, 0.110)):
        s.set_vector(gpu.P_BEXT, (B, 0, 0))
        s.minimize()
        assert abs(s.average_m()[0] - mx) < 1e-3, (B, s.average_m())
    s.set_vector(gpu.P_ANISU, (100, 0, 0))   # unnormalised axis: normalised internally
    s.minimize()
    assert abs(s.average_m()[0] - 0.110) < 1e-3
    # hard axis perpendicular
    s.set_vector(gpu.P_ANISU, (0, 0, 1))
    s.set_vector(gpu.P_BEXT, (0, 0, 0))
    s.set_scalar(gpu.P_KU1, -1e6)
    s.minimize()
    assert abs(s.average_m()[2]) < 1e-3, s.average_m()


def test_standard_problem5a(gpu):
    """test/standardproblem5a.mx3: 64 x 32 x 1 permalloy film, minimize() from
    uniform(1, 1, 0.001) with alpha 3, then Slonczewski STT (Pol 0.5669,
    Lambda 1, eps' 0, fixed layer at 1 degree in-plane, J_z = -0.008 A over
    the film area, alpha 0.01) for 1 + 1 ns; <m> at each stage, TOL 1e-3."""
    import math
    _say("sp5a")
    N = (64, 32, 1)
    sX, sY, sZ = 160e-9, 80e-9, 5e-9
    s = gpu.Sim(N, (sX / N[0], sY / N[1], sZ / N[2]), dtype=np.float32)
    setup(s, gpu, Ms=800e3, Aex=13e-12, alpha=3.0, dtype=np.float32)
    s.set_m(uniform_m(N, (1, 1, 0.001)))
    s.minimize()
    got = s.average_m()
    _say("m0", got)
    assert np.abs(got - np.array([0.9586266279220581, 0.20391345024108887, 0])).max() < 1e-3, got
    s.set_scalar(gpu.P_ALPHA, 0.01)
    s.set_scalar(gpu.P_LAMBDA, 1.0)
    s.set_scalar(gpu.P_POL, 0.5669)
    s.set_scalar(gpu.P_EPSPRIME, 0.0)
    prad = 1 * math.pi / 180
    s.set_vector(gpu.P_FIXEDLAYER, (math.cos(prad), math.sin(prad), 0))
    s.set_vector(gpu.P_J, (0, 0, -0.008 / (sX * sY)))
    s.run(1e-9)
    got = s.average_m()
    _say("m1", got)
    assert np.abs(got - np.array([0.6440994739532471, 0.5131782293319702, -0.1569230705499649])).max() < 1e-3, got
    s.run(1e-9)
    got = s.average_m()
    _say("m2", got)
    assert np.abs(got - np.array([-0.957406222820282, 0.20698121190071106, 0.009677470661699772])).max() < 1e-3, got
