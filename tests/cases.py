"""Shared test cases: the same parameter set applied to an oracle Sim and a GPU
Sim (both expose the same methods), plus seeded initial states."""
import math

import numpy as np

MU0 = 4e-7 * math.pi
GAMMA = 1.7595e11


def random_unit(shape, seed):
    rng = np.random.default_rng(seed)
    m = rng.standard_normal((3,) + tuple(shape))
    return m / np.linalg.norm(m, axis=0)


def strip_m(seed=1):
    """128x32x1: (1,1,0)/sqrt2 plus a seeded uniform +-0.1 perturbation, renormalised."""
    rng = np.random.default_rng(seed)
    m = np.zeros((3, 1, 32, 128))
    m[0] = 1 / math.sqrt(2)
    m[1] = 1 / math.sqrt(2)
    m += rng.uniform(-0.1, 0.1, m.shape)
    return m / np.linalg.norm(m, axis=0)


def setup(sim, mod, *, Ms, Aex, alpha, Dind=0.0, Ku1=0.0, Ku2=0.0, anisU=(0, 0, 1), Kc=(0.0, 0.0, 0.0),
          c1=(1, 0, 0), c2=(0, 1, 0), Bext=(0.0, 0.0, 0.0), regions=None, bext_lut=None, frozen=None, vol=None,
          temp=0.0, seed=0, demag=True, openbc=False, dtype=np.float64):
    """Apply a parameter set to `sim` (oracle or GPU); `mod` provides the LUT helper."""
    from umagnus_amd.params import exchange_lut
    P = mod  # parameter ids are identical in both modules
    sim.set_scalar(P.P_MSAT, Ms)
    sim.set_scalar(P.P_ALPHA, alpha)
    sim.set_scalar(P.P_KU1, Ku1)
    sim.set_scalar(P.P_KU2, Ku2)
    sim.set_vector(P.P_ANISU, anisU)
    sim.set_scalar(P.P_KC1, Kc[0])
    sim.set_scalar(P.P_KC2, Kc[1])
    sim.set_scalar(P.P_KC3, Kc[2])
    sim.set_vector(P.P_ANISC1, c1)
    sim.set_vector(P.P_ANISC2, c2)
    sim.set_vector(P.P_BEXT, Bext)
    sim.set_scalar(P.P_TEMP, temp)
    if regions is not None:
        sim.set_regions(regions)
    a = exchange_lut(Aex, dtype=dtype)
    has_dmi = (not np.isscalar(Dind)) or Dind != 0
    d = exchange_lut(Dind, dtype=dtype) if has_dmi else None
    sim.set_exchange_luts(a, d)
    if bext_lut is not None:
        sim.set_bext_lut(bext_lut)
    if frozen is not None:
        sim.set_frozen_lut(frozen)
    if vol is not None:
        sim.set_geometry(vol)
    sim.set_flags(enable_demag=int(demag), openbc=int(openbc))
    sim.set_thermal_seed(seed)


def relerr(got, want):
    """max |got - want| / max |want| (normwise relative error of a field)."""
    return float(np.abs(np.asarray(got, np.float64) - want).max() / max(np.abs(want).max(), 1e-300))


def percell_relerr(got, want):
    """max over cells of |got_i - want_i| / |want_i| (vector norms per cell; cells
    where want_i = 0 must match exactly)."""
    d = np.linalg.norm(np.asarray(got, np.float64) - want, axis=0)
    w = np.linalg.norm(np.asarray(want, np.float64), axis=0)
    if np.any(d[w == 0] != 0):
        return float("inf")
    return float((d[w > 0] / w[w > 0]).max()) if np.any(w > 0) else 0.0


def percell_torque_relerr(got, want, beff):
    """max over cells of |tau_got_i - tau_want_i| / |B_eff_i|: the torque error of
    a cell against its field, which bounds |tau_i| (|m x B| <= |B| for unit m).
    Relative to |tau_i| itself the error is unbounded where m is nearly parallel
    to B and the cross product cancels."""
    d = np.linalg.norm(np.asarray(got, np.float64) - want, axis=0)
    b = np.linalg.norm(np.asarray(beff, np.float64), axis=0)
    if np.any(d[b == 0] != 0):
        return float("inf")
    return float((d[b > 0] / b[b > 0]).max()) if np.any(b > 0) else 0.0


# ---- initial states and shapes of the reference's scripts (engine/config.go, shape.go) ----------
def coords(N, cs):
    """Index2Coord (engine/util.go:132-140): cell centres, origin at the grid centre."""
    ix, iy, iz = (np.arange(n) for n in N)
    x = cs[0] * (ix - 0.5 * (N[0] - 1))
    y = cs[1] * (iy - 0.5 * (N[1] - 1))
    z = cs[2] * (iz - 0.5 * (N[2] - 1))
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X, Y, Z


def uniform_m(N, v):
    m = np.zeros((3, N[2], N[1], N[0]))
    for c in range(3):
        m[c] = v[c]
    return m


def vortex(N, cs, circ, pol):
    """Vortex(circ, pol), engine/config.go:63-73 (NaN at r = 0 -> (0, 0, pol))."""
    X, Y, _ = coords(N, cs)
    r2 = X * X + Y * Y
    r = np.sqrt(r2)
    with np.errstate(divide="ignore", invalid="ignore"):
        m = np.stack([-Y * circ / r, X * circ / r, 1.5 * pol * np.exp(-r2 / (2 * cs[0] ** 2))])
    bad = ~np.isfinite(m).all(axis=0)
    m[:, bad] = np.array([0.0, 0.0, float(pol)])[:, None]
    return m


def two_domain(N, cs, m1, mwall, m2):
    """TwoDomain, engine/config.go:144-159: m1 for x < 0, m2 for x >= 0, Gaussian wall."""
    X, _, _ = coords(N, cs)
    ww = 2 * cs[0]
    gauss = np.exp(-(X / ww) ** 2)
    m = np.empty((3,) + X.shape)
    for c in range(3):
        base = np.where(X < 0, m1[c], m2[c])
        m[c] = (1 - gauss) * base + gauss * mwall[c]
    return m


def ellipse_geom(N, cs, dx, dy):
    """SetGeom(Ellipse(dx, dy)) without edge smoothing (geom.go:74-140): 1 inside, 0 outside."""
    X, Y, _ = coords(N, cs)
    return ((X / dx) ** 2 + (Y / dy) ** 2 <= 0.25).astype(np.float64)


def cylinder_geom(N, cs, diam, height):
    """SetGeom(Cylinder(diam, height)), engine/shape.go:61-66."""
    X, Y, Z = coords(N, cs)
    return ((np.abs(Z) <= height / 2) & ((X / diam) ** 2 + (Y / diam) ** 2 <= 0.25)).astype(np.float64)


def nodemag_script(mod, dtype):
    """test/nodemagspins.mx3 on a simulation module (the oracle or the GPU one):
    128 x 32 x 1 standard-problem-4 film, m = uniform(1, 1, 1), regions LEFT
    (x < 0) and RIGHT (x >= 0).  Returns the checks as (label, got, want, tol)."""
    N, cs = (128, 32, 1), (500e-9 / 128, 125e-9 / 32, 3e-9)
    s = mod.Sim(N, cs, dtype=dtype)
    X, Y, _ = coords(N, cs)
    reg = np.where(X < 0, 0, 1).astype(np.uint8)
    setup(s, mod, Ms=800e3, Aex=13e-12, alpha=0.02, regions=reg, dtype=dtype)
    s.set_m(uniform_m(N, (1, 1, 1)))
    out = []

    def avg(r):
        B = s.demag_field().astype(np.float64)
        return B[:, reg == r].mean(axis=1)

    def lut(left, right):
        v = np.zeros(256)
        v[0], v[1] = left, right
        return v
    full = np.array([-0.005299, -0.022127, -0.552248])
    out += [("L", avg(0), full, 1e-4), ("R", avg(1), full, 1e-4)]
    s.set_nodemag_lut(lut(0, 1))
    out += [("L", avg(0), np.array([-0.0103589, -0.02122757, -0.5480939]), 1e-4), ("R", avg(1), np.zeros(3), 0.0)]
    s.set_nodemag_lut(lut(1, 1))
    out += [("L", avg(0), np.zeros(3), 0.0), ("R", avg(1), np.zeros(3), 0.0)]
    # the air gap: geometry y < 0, region 2 for y >= 0
    reg = np.where(Y >= 0, 2, reg).astype(np.uint8)
    s.set_regions(reg)
    s.set_geometry((Y < 0).astype(np.float64))
    out += [("L", avg(0), np.zeros(3), 0.0), ("R", avg(1), np.zeros(3), 0.0), ("A", avg(2), np.zeros(3), 0.0)]
    s.set_nodemag_lut(lut(0, 0))
    half = np.array([-0.0046074, -0.0391400, -0.53594726])
    out += [("L", avg(0), half, 1e-4), ("R", avg(1), half, 1e-4)]
    return out
