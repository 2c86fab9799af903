"""Pins the oracle's spin-transfer torques (zhangli2.cl, slonczewski2.cl) and
bulk DMI (dmibulk.cl) on the CPU.

- bulk DMI: a Bloch helix along any axis is an eigenvector of the discrete
  exchange + bulk-DMI operator, B = (2D sin(kc)/c + 4A (cos(kc) - 1)/c^2) m / Ms,
  the same coefficient for x, y and z (bulk DMI is isotropic); D -> 0 gives the
  exchange field; the reference's 90-degree rotation test
  (test/dmibulksymm.mx3) at 8^3 with open, extrapolated boundaries;
- Zhang-Li: on a periodic helix with current along it, (u.grad)m is tangent
  to m, so the torque has the closed form -g |hs| ((1 + xi a) t - (xi - a) n);
- Slonczewski: torque perpendicular to m, zero for p parallel to m, linear in
  J / thickness (test/freelayerthickness.mx3's invariance), odd in the fixed
  layer position sign, and pulling m towards p for J > 0.
"""
import math

import numpy as np
import pytest

MUB, QE, GAMMA0, HBAR = 9.2740091523E-24, 1.60217646E-19, 1.7595e11, 1.05457173E-34


def _lut(oracle, v, dt=np.float64):
    from umagnus_amd.params import exchange_lut
    return exchange_lut(v, dtype=dt)


def _bulk(oracle, m, Ms, A, D, cs, pbc=0, openbc=0, dt=np.float64):
    Nz, Ny, Nx = m.shape[1:]
    B = np.zeros_like(m, dt)
    m = np.ascontiguousarray(m, dt)
    a, d = _lut(oracle, A, dt), _lut(oracle, D, dt)
    R = oracle._rt(dt)
    oracle.fn("adddmibulk", dt)(*[oracle.ptr(x) for x in (B[0], B[1], B[2], m[0], m[1], m[2])],
                                None, R(Ms), oracle.ptr(a), oracle.ptr(d), None,
                                R(cs[0]), R(cs[1]), R(cs[2]), Nx, Ny, Nz, pbc, openbc)
    return B


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_bulk_dmi_bloch_helix(oracle, axis):
    cs = (2e-9, 2.5e-9, 1.5e-9)
    N = [4, 5, 6]
    N[axis] = 24
    Ms, A, D = 9e5, 1.2e-11, 2.5e-3
    c = cs[axis]
    k = 2 * math.pi * 3 / (N[axis] * c)
    pos = np.arange(N[axis]) * c
    shape = (N[2], N[1], N[0])
    phase = np.broadcast_to(np.expand_dims(k * pos, tuple(i for i in range(3) if i != 2 - axis)), shape)
    m = np.zeros((3,) + shape)
    a, b = (axis + 1) % 3, (axis + 2) % 3      # helix in the plane transverse to the axis
    m[a] = np.cos(phase)
    m[b] = np.sin(phase)
    B = _bulk(oracle, m, Ms, A, D, cs, pbc=7)
    A, D = float(np.float32(A)), float(np.float32(D))     # the LUTs are built in float32
    coef = (2 * D * math.sin(k * c) / c + 4 * A * (math.cos(k * c) - 1) / c ** 2) / Ms
    assert np.abs(B - coef * m).max() < 1e-10 * abs(coef)
    # D -> 0: the exchange field alone
    B0 = _bulk(oracle, m, Ms, A, 0.0, cs, pbc=7)
    assert np.abs(B0 - (4 * A * (math.cos(k * c) - 1) / c ** 2 / Ms) * m).max() < 1e-10 * abs(coef)


def test_bulk_dmi_rotation_symmetry(oracle):
    """test/dmibulksymm.mx3: a 16^3 cube (A 10 pJ/m, Ms 1e6, alpha 1, Dbulk
    1e-3) relaxed for 0.1 ns from uniform states along x, y and z; rotating the
    start by 90 degrees rotates the result (TOL 1e-4).  Sensitive to the signs
    of the boundary extrapolation.  Reduced here to 8^3 and 0.04 ns (the same
    symmetry; the full size runs on the GPU)."""
    from cases import setup, uniform_m
    N, cs = (8, 8, 8), (2e-9,) * 3

    def run(v):
        s = oracle.Sim(N, cs, dtype=np.float64)
        setup(s, oracle, Ms=1e6, Aex=10e-12, alpha=1.0, dtype=np.float64)
        s.set_dmi_bulk(_lut(oracle, 1e-3))
        s.set_m(uniform_m(N, v))
        s.run(0.4e-10)
        return s.average_m()
    mx = run((1, 0, 0))
    assert abs(run((0, 1, 0))[1] - mx[0]) < 1e-4
    assert abs(run((0, 0, 1))[2] - mx[0]) < 1e-4
    myz = run((0, 1, 1))
    assert abs(run((1, 0, 1))[1] - myz[0]) < 1e-4
    assert abs(run((1, 1, 0))[2] - myz[0]) < 1e-4


def _zhangli(oracle, m, Ms, J, alpha, xi, pol, cs, pbc):
    Nz, Ny, Nx = m.shape[1:]
    t = np.zeros_like(m)
    m = np.ascontiguousarray(m, np.float64)
    R = oracle._rt(np.float64)
    oracle.fn("addzhangli", np.float64)(*[oracle.ptr(x) for x in (t[0], t[1], t[2], m[0], m[1], m[2])],
                                        None, R(Ms), None, R(J[0]), None, R(J[1]), None, R(J[2]),
                                        None, R(alpha), None, R(xi), None, R(pol),
                                        R(cs[0]), R(cs[1]), R(cs[2]), Nx, Ny, Nz, pbc)
    return t


def test_zhangli_helix_closed_form(oracle):
    cs, Nx = (3e-9, 2e-9, 2e-9), 32
    Ms, Jx, alpha, xi, pol = 8e5, 1.3e12, 0.1, 0.05, 0.7
    k = 2 * math.pi * 2 / (Nx * cs[0])
    x = np.arange(Nx) * cs[0]
    m = np.zeros((3, 1, 2, Nx))
    m[0], m[1] = np.cos(k * x), np.sin(k * x)
    tq = _zhangli(oracle, m, Ms, (Jx, 0, 0), alpha, xi, pol, cs, pbc=1)
    b = MUB / (2 * QE * GAMMA0) / Ms / (1 + xi * xi)
    hs = (b / cs[0]) * (pol * Jx) * 2 * math.sin(k * cs[0])
    g = -1 / (1 + alpha * alpha)
    t = np.stack([-np.sin(k * x), np.cos(k * x), 0 * x])       # tangent; m x t = z
    want = g * hs * (-(1 + xi * alpha) * t + (xi - alpha) * np.array([0, 0, 1.0])[:, None])
    assert np.abs(tq[:, 0, 0] - want).max() < 1e-10 * np.abs(want).max()
    # no current, or uniform m: no torque
    assert not _zhangli(oracle, m, Ms, (0, 0, 0), alpha, xi, pol, cs, 1).any()
    mu = np.zeros_like(m)
    mu[2] = 1
    assert not _zhangli(oracle, mu, Ms, (Jx, 2e11, 0), alpha, xi, pol, cs, 1).any()


def _slonc(oracle, m, Ms, Jz, p, alpha, pol, lam, eps, thick, mesh_t, flp):
    n = m.shape[1]
    t = np.zeros_like(m)
    m = np.ascontiguousarray(m, np.float64)
    R = oracle._rt(np.float64)
    oracle.fn("addslonczewski", np.float64)(*[oracle.ptr(x) for x in (t[0], t[1], t[2], m[0], m[1], m[2])],
                                            None, R(Ms), None, R(Jz), None, R(p[0]), None, R(p[1]), None,
                                            R(p[2]), None, R(alpha), None, R(pol), None, R(lam), None,
                                            R(eps), None, R(thick), R(mesh_t), R(flp), n)
    return t


def test_slonczewski_properties(oracle):
    rng = np.random.default_rng(3)
    m = rng.standard_normal((3, 50))
    m /= np.linalg.norm(m, axis=0)
    p = (0.3, -0.2, 0.9)
    args = dict(Ms=8e5, Jz=-2e11, p=p, alpha=0.02, pol=0.5669, lam=2.0, eps=0.03)
    t = _slonc(oracle, m, thick=0.0, mesh_t=3e-9, flp=1.0, **args)
    assert np.abs((t * m).sum(axis=0)).max() < 1e-12 * np.abs(t).max()          # perpendicular to m
    # FreeLayerThickness = mesh thickness gives the same torque (freelayerthickness.mx3)
    assert np.array_equal(t, _slonc(oracle, m, thick=3e-9, mesh_t=7e-9, flp=1.0, **args))
    # proportional to J / thickness, odd in the fixed-layer side
    t2 = _slonc(oracle, m, thick=1.5e-9, mesh_t=7e-9, flp=1.0, **args)
    assert np.abs(t2 - 2 * t).max() < 1e-12 * np.abs(t).max()
    assert np.abs(_slonc(oracle, m, thick=0.0, mesh_t=3e-9, flp=-1.0, **args) + t).max() < 1e-12 * np.abs(t).max()
    # p parallel to m: no torque
    pm = np.zeros((3, 1))
    pm[:, 0] = np.array(p) / np.linalg.norm(p)
    assert np.abs(_slonc(oracle, pm, thick=0.0, mesh_t=3e-9, flp=1.0, **args)).max() < 1e-14 * np.abs(t).max()
    # J > 0 and epsilon' = 0: the damping-like part turns m towards p
    a2 = dict(args, Jz=2e11, eps=0.0, alpha=0.0)
    t3 = _slonc(oracle, m, thick=0.0, mesh_t=3e-9, flp=1.0, **a2)
    pn = np.array(p) / np.linalg.norm(p)
    toward = pn[:, None] - m * (pn @ m)
    assert ((t3 * toward).sum(axis=0) > 0).all()
