"""Pins the oracle's periodic demag tensors (mag/demagkernel.go:115-352,
mag/newellDemagKernel.go:1108-1324 with SetPBC) on the CPU.

What fixes them, independently of the restatement:
- images: with PBC every wrapped tensor element is the sum of the open-boundary
  tensor over the offsets of the kernelRanges box that wrap onto it.  The open
  tensor is pinned elsewhere (the reference's demag goldens, Aharoni block sums)
  and does not depend on the grid, so a larger open grid supplies the image
  values.  For Newell the origin holds the self term alone: the reference
  overwrites it after the image sums (:1300-1310);
- the reference's PBC goldens: test/demag2Dpbc.mx3 (thin film extended along
  y), demag2Dpbc2.mx3 (along x, with even and odd Nx) and demagrodpbc.mx3
  (rod extended along z): <B_demag> for uniform m with Msat = 1/mu0;
- limits: a film periodic in x and y tends to Nzz = 1, a rod periodic along its
  axis to Nxx = Nyy = 1/2, with the exact Newell tensor converging
  monotonically in the number of repetitions.
"""
import numpy as np
import pytest

MU0 = 4e-7 * np.pi


def _open_images(oracle, kind, N, cs, pbc, **kw):
    """The periodic tensor rebuilt from an open-boundary tensor large enough to
    hold every offset of the kernelRanges box."""
    P = oracle.pad_size(N, pbc)
    r = []
    for c in range(3):
        if pbc[c]:
            r.append(P[c] * pbc[c] - 1)
        else:
            r.append((P[c] - 1) // 2)
    if P[2] == 1 and pbc[2] == 0:
        r[2] = 0
    # open grid whose ranges reach r on every axis
    No = [max(N[c], r[c] + 1) for c in range(3)]
    if r[2] == 0:
        No[2] = 1
    f = oracle.demag_kernel_bf if kind == "bf" else oracle.demag_kernel_newell
    Ko = f(tuple(No), cs, **kw)
    Po = Ko[0].shape[::-1]
    out = [np.zeros((P[2], P[1], P[0])) for _ in range(6)]
    for z in range(-r[2], r[2] + 1):
        for y in range(-r[1], r[1] + 1):
            for x in range(-r[0], r[0] + 1):
                if (kind == "newell" and (x % P[0], y % P[1], z % P[2]) == (0, 0, 0)):
                    continue
                for c in range(6):
                    out[c][z % P[2], y % P[1], x % P[0]] += Ko[c][z % Po[2], y % Po[1], x % Po[0]]
    return out, P


PBC_CASES = [
    ((4, 3, 2), (2e-9, 3e-9, 2.5e-9), (2, 1, 0)),
    ((3, 4, 2), (2e-9, 3e-9, 2.5e-9), (0, 0, 3)),
    ((6, 5, 1), (1e-9, 1.5e-9, 2e-9), (3, 2, 0)),
    ((4, 4, 3), (1e-9, 1e-9, 1e-9), (1, 1, 1)),
    ((5, 2, 1), (2e-9, 2e-9, 1e-9), (0, 4, 2)),      # 2D film periodic in z as well
]


@pytest.mark.parametrize("N,cs,pbc", PBC_CASES)
def test_bf_pbc_is_sum_of_open_images(oracle, N, cs, pbc):
    K = oracle.demag_kernel_bf(N, cs, pbc=pbc)
    want, P = _open_images(oracle, "bf", N, cs, pbc)
    assert K[0].shape == (P[2], P[1], P[0])
    scale = max(np.abs(w).max() for w in want)
    for c in range(6):
        if P[2] == 1 and c in (3, 4):
            assert not K[c].any()
            continue
        # the open tensor holds negative offsets as mirror copies, the periodic
        # one evaluates them: they agree to the integration's rounding
        assert np.abs(K[c] - want[c]).max() <= 1e-12 * scale, c


@pytest.mark.parametrize("N,cs,pbc", PBC_CASES)
def test_newell_pbc_is_sum_of_open_images(oracle, N, cs, pbc):
    kw = dict(asymptotic_radius=4)
    K = oracle.demag_kernel_newell(N, cs, pbc=pbc, **kw)
    want, P = _open_images(oracle, "newell", N, cs, pbc, **kw)
    Kself = oracle.demag_kernel_newell((1, 1, 1), cs, **kw)
    scale = max(np.abs(w).max() for w in want)
    for c in range(6):
        if P[2] == 1 and c in (3, 4):
            assert not K[c].any()
            continue
        d = K[c] - want[c]
        # origin: the self term alone, whatever the images
        assert K[c][0, 0, 0] == Kself[c][0, 0, 0]
        d[0, 0, 0] = 0
        assert np.abs(d).max() <= 1e-12 * scale, c


def _avg_b(oracle, N, cs, pbc, newell=False):
    K = oracle.demag_kernel_newell(N, cs, pbc=pbc) if newell else None
    D = oracle.Demag(N, cs, pbc=pbc, dtype=np.float64, K=K)
    out = []
    for v in np.eye(3):
        m = np.zeros((3, N[2], N[1], N[0]))
        m[:] = v[:, None, None, None]
        B = D.exec(m, np.full((N[2], N[1], N[0]), 1 / MU0))
        out.append(B.reshape(3, -1).mean(axis=1))
    return np.array(out)      # row: m direction, column: B component


def _expect_film(k):
    """expectations of test/demag2Dpbc.mx3 and demag2Dpbc2.mx3"""
    off = ~np.eye(3, dtype=bool)
    assert np.abs(k[off]).max() < 1e-9, k
    assert abs(k[0, 0]) < 1e-2 and abs(k[1, 1]) < 1e-2, k
    assert abs(k[2, 2] + 1) < 2e-2, k


def test_demag2d_pbc_goldens(oracle):
    cs = (1e-9, 1e-9, 0.5e-9)
    _expect_film(_avg_b(oracle, (128, 2, 1), cs, (0, 32, 0)))
    _expect_film(_avg_b(oracle, (2, 128, 1), cs, (32, 0, 0)))
    _expect_film(_avg_b(oracle, (3, 128, 1), cs, (32, 0, 0)))     # odd Nx


def test_demag_rod_pbc_golden(oracle):
    """test/demagrodpbc.mx3: 2x2x2 cells of 1x1x2 nm, PBC 16 along z:
    <B> = (-0.496, -0.496, 0) diagonal, off-diagonal 0 (TOL 1e-2)."""
    k = _avg_b(oracle, (2, 2, 2), (1e-9, 1e-9, 2e-9), (0, 0, 16))
    assert np.abs(np.diag(k) - np.array([-0.496, -0.496, 0])).max() < 1e-2, k
    assert np.abs(k[~np.eye(3, dtype=bool)]).max() < 1e-2, k


def test_pbc_limits_newell(oracle):
    """Exact Newell tensor (no far-field expansion): the demag factor of the
    summed box approaches the infinite film / rod values monotonically."""
    cs = (2e-9, 2e-9, 1e-9)
    prev = 0.0
    for rep in (1, 2, 4, 8):
        K = oracle.demag_kernel_newell((4, 4, 1), cs, pbc=(rep, rep, 0), asymptotic_radius=-1)
        nzz = -K[2].sum()
        assert prev < nzz < 1.0
        prev = nzz
    assert nzz > 0.97
    prev = 1.0
    for rep in (1, 2, 4, 8):
        K = oracle.demag_kernel_newell((1, 1, 4), (1e-9, 1e-9, 2e-9), pbc=(0, 0, rep), asymptotic_radius=-1)
        nzz = -K[2].sum()
        assert 0.0 < nzz < prev
        assert abs(K[0].sum() - K[1].sum()) < 1e-14
        # trace: Nxx + Nyy + Nzz = 1 (far exact elements cancel to ~1e-12)
        assert abs(-K[0].sum() - (1 - nzz) / 2) < 1e-10
        prev = nzz
    assert nzz < 0.03
