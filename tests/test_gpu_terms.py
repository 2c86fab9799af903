"""Extra field and torque terms on the simulation handle (um_sim_add_field_term):
B_ext mask terms (Excitation.Add, engine/excitation.go:41-53), custom fields
(AddFieldTerm, engine/customfield.go:68) and custom torques (AddTorqueTerm,
engine/torque.go:108-121).

Checks: a constant mask term equals a per-cell B_ext of the same values
against the FP64 oracle (B_eff, torque and the Zeeman energy, 1e-12 / 1e-5
per cell); a time-dependent mask term steps bit for bit like the same per-cell
B_ext scaled by a time function; custom fields and torques add exactly (bit
for bit) and FreezeSpins still zeroes frozen cells; custom fields stay out of
the Zeeman energy; clearing removes them.
"""
import math

import numpy as np
import pytest

from cases import percell_relerr, percell_torque_relerr, random_unit, relerr, setup

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    from umagnus_amd import umhip
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    umhip.init(0)
    return umhip


N, CS = (24, 20, 3), (2e-9, 2.5e-9, 3e-9)
SHAPE = (N[2], N[1], N[0])
PARAMS = dict(Ms=8e5, Aex=1.3e-11, Dind=1e-3, Ku1=3e5, alpha=0.1)


def _mask(seed):
    rng = np.random.default_rng(seed)
    return rng.normal(0, 0.03, (3,) + SHAPE)


@pytest.mark.parametrize("prec", [1, 0])
def test_mask_term_vs_oracle_bext(gpu, oracle, prec):
    dt = np.float32 if prec == 0 else np.float64
    mask, mul = _mask(1), 0.7
    m = random_unit(SHAPE, 2)
    sg = gpu.Sim(N, CS, dtype=dt)
    setup(sg, gpu, dtype=dt, **PARAMS)
    sg.add_field_term(gpu.Sim.FIELD_TERM_BEXT, mask, mul)
    sg.set_m(m)
    so = oracle.Sim(N, CS, dtype=np.float64)
    setup(so, oracle, dtype=np.float64, **PARAMS)
    so.set_vector(oracle.P_BEXT, [mul * mask[c] for c in range(3)])
    so.set_m(m)
    tol = 1e-5 if prec == 0 else 1e-12
    b_o, t_o = so.beff(), so.torque()
    b_g, t_g = sg.beff(), sg.torque()
    assert percell_relerr(b_g, b_o) <= tol and percell_torque_relerr(t_g, t_o, b_o) <= tol
    ez_o, ez_g = so.energy(2), sg.energy(2)
    assert abs(ez_g - ez_o) <= tol * abs(ez_o), (ez_g, ez_o)
    assert abs(sg.energy(9) - so.energy(9)) <= tol * abs(so.energy(9))
    sg.clear_field_terms(gpu.Sim.FIELD_TERM_BEXT)
    so.set_vector(oracle.P_BEXT, (0.0, 0.0, 0.0))
    assert relerr(sg.beff(), so.beff()) <= tol


@pytest.mark.parametrize("prec", [1, 0])
def test_time_dependent_mask_term_steps(gpu, prec):
    """Mask term with f(t) = sin(2 pi 5 GHz t) vs per-cell B_ext under the same
    time function: 30 adaptive RK45DP steps, the same bits."""
    dt = np.float32 if prec == 0 else np.float64
    mask = _mask(3)
    f = lambda t: math.sin(2 * math.pi * 5e9 * t) + 0.2
    m = random_unit(SHAPE, 4)
    a = gpu.Sim(N, CS, dtype=dt)
    b = gpu.Sim(N, CS, dtype=dt)
    for s in (a, b):
        setup(s, gpu, dtype=dt, **PARAMS)
        s.set_solver_params(dt=1e-14)
        s.set_m(m)
    a.add_field_term(gpu.Sim.FIELD_TERM_BEXT, mask, fn=f)
    b.set_vector(gpu.P_BEXT, [mask[c] for c in range(3)])
    b.set_time_function(gpu.P_BEXT, lambda t: (f(t), f(t), f(t)), 3)
    a.steps(30)
    b.steps(30)
    assert a.state()["t"] == b.state()["t"] and a.state()["nsteps"] == 30
    assert np.array_equal(a.get_m(), b.get_m())


@pytest.mark.parametrize("prec", [1, 0])
def test_custom_field_and_torque_terms_add_exactly(gpu, prec):
    dt = np.float32 if prec == 0 else np.float64
    m = random_unit(SHAPE, 5)
    regions = (np.arange(N[0])[None, None, :] // 8 % 3 * np.ones(SHAPE)).astype(np.uint8)
    frozen = np.zeros(256)
    frozen[2] = 1
    s = gpu.Sim(N, CS, dtype=dt)
    setup(s, gpu, dtype=dt, regions=regions, frozen=frozen, **PARAMS)
    s.set_m(m)
    b0, t0 = s.beff(), s.torque()
    ez0 = s.energy(2)
    bf, tf = _mask(6).astype(dt), _mask(7).astype(dt)
    s.add_field_term(gpu.Sim.FIELD_TERM_CUSTOM, bf, 1.0)
    b1 = s.beff()
    assert np.array_equal(b1, (b0 + bf).astype(dt))
    assert s.energy(2) == ez0                   # a custom field is not Zeeman energy
    t1 = s.torque()
    s.add_field_term(gpu.Sim.TORQUE_TERM_CUSTOM, tf, 2.0)
    t2 = s.torque()
    want = (t1 + (dt(2.0) * tf).astype(dt)).astype(dt)
    want[:, regions == 2] = 0
    assert np.array_equal(t2, want)
    assert np.all(t2[:, regions == 2] == 0) and not np.array_equal(t1, t0)
    s.clear_field_terms()
    assert np.array_equal(s.beff(), b0) and np.array_equal(s.torque(), t0)
