"""Host-side parameter tables of the engine (engine/exchange.go, regions.go).

The simulation handle takes the inter-region exchange / DMI tables the way the
reference uploads them (exchParam.update, engine/exchange.go:153-167): a
symmetric 256x256 table stored as its lower triangle (symmidx), filled with
the harmonic mean of the two regions' values (exchAverage, :193-202), times a
per-pair scale plus a per-pair override.
"""
import math

import numpy as np

NREGION = 256
SYMLUT = NREGION * (NREGION + 1) // 2


def symmidx(i, j):
    """engine/exchange.go:181-187"""
    if j <= i:
        return i * (i + 1) // 2 + j
    return j * (j + 1) // 2 + i


def exch_average(exi, exj):
    """engine/exchange.go:193-202, in float32 like the reference."""
    exi, exj = np.float32(exi), np.float32(exj)
    if exi * exj >= 0:
        with np.errstate(divide="ignore"):
            return np.float32(np.float32(2) / (np.float32(1) / exi + np.float32(1) / exj))
    a, b = float(exi), float(exj)
    sign = math.copysign(1.0, a + b)
    magn = math.sqrt(math.sqrt(-a * b) * abs(a + b) / 2)
    return np.float32(sign * magn)


def exchange_lut(per_region, scale=None, inter=None, dtype=np.float32):
    """LUT for Aex or Dind.  per_region: scalar (all regions) or {region: value};
    scale/inter: optional {(r1, r2): value} (ext_ScaleExchange / ext_InterExchange)."""
    ex = np.zeros(NREGION, np.float32)
    if np.isscalar(per_region):
        ex[:] = per_region
    else:
        for r, v in per_region.items():
            ex[r] = v
    lut = np.zeros(SYMLUT, np.float32)
    # vectorised harmonic mean over the lower triangle
    ii, jj = np.tril_indices(NREGION)
    a, b = ex[ii], ex[jj]
    with np.errstate(divide="ignore", invalid="ignore"):
        same = a * b >= 0
        hm = (np.float32(2) / (np.float32(1) / a + np.float32(1) / b)).astype(np.float32)
    idx = ii * (ii + 1) // 2 + jj
    lut[idx] = np.where(same, hm, 0).astype(np.float32)
    for k in np.nonzero(~same)[0]:
        lut[idx[k]] = exch_average(a[k], b[k])
    for (r1, r2), s in (scale or {}).items():
        lut[symmidx(r1, r2)] = np.float32(s) * exch_average(ex[r1], ex[r2])
    for (r1, r2), v in (inter or {}).items():
        lut[symmidx(r1, r2)] = np.float32(v)
    return lut.astype(dtype)
