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


def config1_m(seed=1):
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
