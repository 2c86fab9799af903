"""OVF1/OVF2 I/O (umagnus_amd/ovf.py) against the reference's own fixtures and
expectations, plus the exchange_inhomo golden that starts from an OVF file.

Fixtures in tests/golden/ovf/ are the reference's test data files
(test/testdata/); the expected values are those of test/loadfile.mx3 and
test/exchange_inhomo.mx3."""
import os

import numpy as np
import pytest

from umagnus_amd import ovf

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ovf")


def _g(name):
    return os.path.join(GOLD, name)


@pytest.mark.parametrize("name,elem", [("binary4.ovf", -2.052794059181906e-07),
                                       ("binary8.ovf", -17100.58203125),
                                       ("oommf_ovf1_binary8.omf", -17100.58203125)])
def test_loadfile_goldens(name, elem):
    """test/loadfile.mx3:13-34: ncomp, size and s.get(2, 99, 24, 0)."""
    d, meta = ovf.read(_g(name))
    assert d.shape == (3, 1, 25, 100)
    assert meta["ncomp"] == 3 and meta["size"] == [100, 25, 1]
    assert d.dtype == np.float32
    assert d[2, 0, 24, 99] == np.float32(elem)


def test_binary8_keeps_float64_on_request():
    d8, _ = ovf.read(_g("binary8.ovf"), keep_float64=True)
    d4, _ = ovf.read(_g("binary8.ovf"))
    assert d8.dtype == np.float64 and np.array_equal(d8.astype(np.float32), d4)


def test_ovf1_text_equals_binary4():
    """The same OOMMF field saved as OVF1 text and as big-endian binary 4."""
    t, mt = ovf.read(_g("oommf_ovf1_text.omf"))
    b, mb = ovf.read(_g("oommf_ovf1_binary4.omf"))
    assert t.shape == b.shape == (3, 1, 10, 25)
    assert np.array_equal(t, b)
    assert mt["cellsize"] == mb["cellsize"] == [2e-08, 2e-08, 2e-08]
    assert mb["valuemultiplier"] == 800000.0   # parsed, not applied (as the reference)


def test_scalar_ovf2():
    d, meta = ovf.read(_g("scalarovf2.ovf"))
    assert d.shape == (1, 1, 64, 64) and meta["ncomp"] == 1 and np.isfinite(d).all()


@pytest.mark.parametrize("ver,fmt", [(2, "binary 4"), (2, "text"), (1, "binary 4"), (1, "text")])
@pytest.mark.parametrize("ncomp", [3, 1])
def test_round_trip(tmp_path, ver, fmt, ncomp):
    """test/savefile.mx3: an element survives save + load unchanged in every format."""
    if ver == 1 and ncomp == 1:
        with pytest.raises(ovf.OvfError):
            ovf.write(tmp_path / "x.ovf", np.zeros((1, 1, 2, 2)), (1e-9,) * 3, ovf=1)
        return
    rng = np.random.default_rng(ncomp + ver)
    a = rng.standard_normal((ncomp, 3, 5, 7)).astype(np.float32)
    path = tmp_path / f"f{ver}.ovf"
    ovf.write(path, a, (2e-9, 3e-9, 4e-9), name="m", unit="A/m", time=1.5e-9, fmt=fmt, ovf=ver)
    b, meta = ovf.read(path)
    assert np.array_equal(a, b)
    assert meta["size"] == [7, 5, 3] and meta["cellsize"] == [2e-9, 3e-9, 4e-9]
    assert meta["time"] == 1.5e-9
    assert meta["format"] == fmt and meta["ovf"] == ver


def test_bad_control_number(tmp_path):
    raw = open(_g("randommag4x4x1.ovf"), "rb").read()
    i = raw.index(b"# Begin: Data Binary 4\n") + len(b"# Begin: Data Binary 4\n")
    bad = raw[:i] + b"\x00\x00\x00\x00" + raw[i + 4:]
    p = tmp_path / "bad.ovf"
    p.write_bytes(bad)
    with pytest.raises(ovf.OvfError):
        ovf.read(p)


# ---- test/exchange_inhomo.mx3 -----------------------------------------------------------------
E_EXCH_OOMMF = 154.4891365


def exchange_inhomo_regions():
    """defregion(1..4, rect(2,2).transl(+-1, +-1, 0)) on the 4x4x1 grid of unit cells."""
    r = np.zeros((1, 4, 4), np.uint8)
    r[0, 2:, 2:] = 1
    r[0, 2:, :2] = 2
    r[0, :2, :2] = 3
    r[0, :2, 2:] = 4
    return r


def run_exchange_inhomo(Sim, mod, dtype):
    from umagnus_amd.params import exchange_lut
    m, _ = ovf.read(_g("randommag4x4x1.ovf"))
    regions = exchange_inhomo_regions()
    out = []
    for ms in ({1: 3.2, 2: 4.6, 3: 3.9, 4: 7.8}, {1: 1.1, 2: 1.1, 3: 1.1, 4: 1.1}):
        s = Sim((4, 4, 1), (1.0, 1.0, 1.0), dtype=dtype)
        s.set_regions(regions)
        s.set_scalar(mod.P_MSAT, np.vectorize(lambda r: ms.get(int(r), 1.0))(regions).astype(dtype))
        s.set_scalar(mod.P_ALPHA, 0.1)
        aex = {0: 1.0, 1: 1.2, 2: 1.6, 3: 4.3, 4: 4.3}
        s.set_exchange_luts(exchange_lut(aex, dtype=dtype), None)
        s.set_flags(enable_demag=0)
        s.set_m(m.astype(np.float64))
        out.append(s.energy(1))
    return out


def test_exchange_inhomo_oracle(oracle):
    """E_exch of the loaded random state with harmonic-mean inter-region Aex equals
    OOMMF's Oxc_ExchangePtwise energy (1e-5) and does not depend on Msat."""
    for e in run_exchange_inhomo(oracle.Sim, oracle, np.float64):
        assert abs(e - E_EXCH_OOMMF) <= 1e-5 * E_EXCH_OOMMF, e
