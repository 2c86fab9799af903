"""The oracle's elementwise restatements, pinned on the CPU against what fixes
them independently of their code:

- regionaddv (kernels_src/cl/regionaddv.cl:3-18): numpy fancy indexing of the
  three LUTs by the region bytes, exact in both precisions; a swapped LUT or
  an ignored region map fails;
- zeromask (kernels_src/cl/zeromask.cl:2-13): cells whose region has any
  non-zero mask value (negative ones included) become 0, the others keep their
  bits;
- normalize2 (kernels_src/cl/normalize2.cl:1-24): unit length, the direction
  of vol*v (cross product 0, positive dot product), 0 where vol = 0 and for
  the zero vector (no NaN), and independence of a positive vol;
- madd2..madd7 (kernels_src/cl/madd*.cl): exact sums on integer data with
  power-of-two factors; madd3's grouping f0 x0 + (f1 x1 + f2 x2)
  (madd3.cl:12) against left-to-right sums, told apart at 2^24 in float32;
- settemperature2 (kernels_src/cl/temperature2.cl:1-19, engine/temperature.go:
  109): B = eta sqrt(2 kB alpha T / (gamma V dt Ms)) against the closed form in
  FP64, the sample variance of normal noise against that formula, and 0 for
  T = 0 or Ms = 0;
- llnoprecess (kernels_src/cl/llnoprecess.cl:3-23): the part of B
  perpendicular to m.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as orc

DTYPES = [np.float32, np.float64]


def _p(a):
    return orc.ptr(a)


def _r(dtype, v):
    return orc._rt(dtype)(v)


@pytest.mark.parametrize("dt", DTYPES)
def test_regionaddv(dt):
    rng = np.random.default_rng(1)
    n = 5000
    regions = rng.integers(0, 256, n).astype(np.uint8)
    lut = [rng.standard_normal(256).astype(dt) for _ in range(3)]
    B = rng.standard_normal((3, n)).astype(dt)
    got = B.copy()
    orc.fn("regionaddv", dt)(_p(got[0]), _p(got[1]), _p(got[2]), _p(lut[0]), _p(lut[1]), _p(lut[2]),
                             _p(regions), n)
    want = np.stack([B[c] + lut[c][regions] for c in range(3)])
    assert np.array_equal(got, want)
    # no region map: every cell is region 0
    got = B.copy()
    orc.fn("regionaddv", dt)(_p(got[0]), _p(got[1]), _p(got[2]), _p(lut[0]), _p(lut[1]), _p(lut[2]), None, n)
    assert np.array_equal(got, np.stack([B[c] + lut[c][0] for c in range(3)]))


@pytest.mark.parametrize("dt", DTYPES)
def test_zeromask(dt):
    rng = np.random.default_rng(2)
    n = 4000
    regions = rng.integers(0, 256, n).astype(np.uint8)
    mask = np.zeros(256, dt)
    mask[[3, 17, 200]] = (1.0, -0.5, 2.0)
    x = rng.standard_normal(n).astype(dt)
    got = x.copy()
    orc.fn("zeromask", dt)(_p(got), _p(mask), _p(regions), n)
    hit = np.isin(regions, [3, 17, 200])
    assert hit.sum() > 0 and (~hit).sum() > 0
    assert np.all(got[hit] == 0)
    assert np.array_equal(got[~hit], x[~hit])
    got = x.copy()
    orc.fn("zeromask", dt)(_p(got), _p(mask), None, n)        # region 0: not masked
    assert np.array_equal(got, x)


@pytest.mark.parametrize("dt", DTYPES)
def test_normalize2(dt):
    rng = np.random.default_rng(3)
    n = 3000
    v = (rng.standard_normal((3, n)) * rng.uniform(1e-3, 1e3, n)).astype(dt)
    v[:, 5] = 0                                    # the zero vector
    vol = rng.choice(np.array([0.0, 0.25, 0.5, 1.0], dt), n)
    vol[5] = 1
    got = v.copy()
    orc.fn("normalize2", dt)(_p(got[0]), _p(got[1]), _p(got[2]), _p(vol), n)
    assert np.all(np.isfinite(got))
    zero = (vol == 0)
    assert np.all(got[:, zero] == 0) and np.all(got[:, 5] == 0)
    keep = ~zero
    keep[5] = False
    eps = np.finfo(dt).eps
    g, w = got[:, keep].astype(np.float64), v[:, keep].astype(np.float64)
    assert np.abs(np.linalg.norm(g, axis=0) - 1).max() < 4 * eps
    cr = np.linalg.norm(np.cross(g.T, w.T), axis=1) / np.linalg.norm(w, axis=0)
    assert cr.max() < 8 * eps
    assert np.all(np.einsum("ij,ij->j", g, w) > 0)
    # a positive vol does not change the direction; no vol means vol = 1
    got1 = v.copy()
    orc.fn("normalize2", dt)(_p(got1[0]), _p(got1[1]), _p(got1[2]), None, n)
    assert np.abs(got1[:, keep] - got[:, keep]).max() < 4 * eps


def _madd(dt, srcs, facs):
    n = srcs[0].size
    dst = np.zeros(n, dt)
    arr = (C.c_void_p * len(srcs))(*[s.ctypes.data for s in srcs])
    f = np.array(facs, dt)
    orc.fn("madd", dt)(_p(dst), len(srcs), arr, _p(f), n)
    return dst


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("k", [2, 3, 4, 5, 6, 7])
def test_madd_exact(dt, k):
    rng = np.random.default_rng(10 + k)
    n = 2000
    srcs = [rng.integers(-1000, 1000, n).astype(dt) for _ in range(k)]
    facs = [2.0 ** int(e) * s for e, s in zip(rng.integers(-4, 5, k), rng.choice([-1, 1], k))]
    want = sum(np.float64(f) * s.astype(np.float64) for f, s in zip(facs, srcs))
    assert np.array_equal(_madd(dt, srcs, facs).astype(np.float64), want)   # every partial sum is exact


def test_madd3_grouping():
    a = np.full(4, 2.0 ** 24, np.float32)
    one = np.ones(4, np.float32)
    # madd3: 2^24 + (1 + 1) = 2^24 + 2, exact; left to right, 2^24 + 1 rounds back to 2^24
    assert np.all(_madd(np.float32, [a, one, one], [1, 1, 1]) == np.float32(2.0 ** 24 + 2))
    zero = np.zeros(4, np.float32)
    assert np.all(_madd(np.float32, [a, one, one, zero], [1, 1, 1, 1]) == np.float32(2.0 ** 24))


@pytest.mark.parametrize("dt", DTYPES)
def test_settemperature2(dt):
    kB, gamma = 1.38064852e-23, 1.7595e11
    V, dtime = 8e-27, 1e-14
    k2 = 2 * kB / (gamma * V * dtime)
    rng = np.random.default_rng(4)
    n = 200000
    noise = rng.standard_normal(n).astype(dt)
    Ms = rng.uniform(5e5, 1e6, n).astype(dt)
    T = rng.uniform(0, 400, n).astype(dt)
    alpha = rng.uniform(0.01, 0.5, n).astype(dt)
    B = np.zeros(n, dt)
    f = orc.fn("settemperature2", dt)
    f(_p(B), _p(noise), _r(dt, k2), _p(Ms), _r(dt, 1.0), _p(T), _r(dt, 1.0), _p(alpha), _r(dt, 1.0), n)
    sigma = np.sqrt(2 * kB * alpha.astype(np.float64) * T / (gamma * V * dtime * Ms.astype(np.float64)))
    tol = 1e-6 if dt == np.float32 else 1e-13
    assert np.abs(B - noise * sigma).max() <= tol * np.abs(noise * sigma).max()
    # uniform parameters: the sample variance is the closed form (4 sigma of its spread)
    Bu = np.zeros(n, dt)
    f(_p(Bu), _p(noise), _r(dt, k2), None, _r(dt, 8e5), None, _r(dt, 300.0), None, _r(dt, 0.1), n)
    var = 2 * kB * 0.1 * 300.0 / (gamma * V * dtime * 8e5)
    assert abs(np.var(Bu.astype(np.float64)) / var - 1) < 4 * np.sqrt(2 / n)
    # T = 0 and Ms = 0 give no field (inv_Msat of 0 is 0)
    for Ms_mul, T_mul in ((8e5, 0.0), (0.0, 300.0)):
        Bz = np.ones(n, dt)
        f(_p(Bz), _p(noise), _r(dt, k2), None, _r(dt, Ms_mul), None, _r(dt, T_mul), None, _r(dt, 0.1), n)
        assert np.all(Bz == 0)


@pytest.mark.parametrize("dt", DTYPES)
def test_llnoprecess(dt):
    """llnoprecess.cl:3-23: tau = -m x (m x B), which for unit m is the part of
    B perpendicular to m (BAC-CAB), B - (m.B) m; a sign error or a missing
    cross product fails."""
    rng = np.random.default_rng(5)
    n = 3000
    m = rng.standard_normal((3, n))
    m = (m / np.linalg.norm(m, axis=0)).astype(dt)
    B = rng.standard_normal((3, n)).astype(dt)
    t = np.zeros((3, n), dt)
    orc.fn("llnoprecess", dt)(_p(t[0]), _p(t[1]), _p(t[2]), _p(m[0]), _p(m[1]), _p(m[2]),
                              _p(B[0]), _p(B[1]), _p(B[2]), n)
    m64, B64 = m.astype(np.float64), B.astype(np.float64)
    mm = np.einsum("ij,ij->j", m64, m64)
    want = B64 * mm - m64 * np.einsum("ij,ij->j", m64, B64)
    eps = np.finfo(dt).eps
    assert np.abs(t - want).max() < 16 * eps * np.abs(B64).max()
    assert np.abs(np.einsum("ij,ij->j", t.astype(np.float64), m64)).max() < 16 * eps * np.abs(B64).max()
