"""The demag tensor cache's host side (mag/demagkernel.go:17-109,
mag/newellDemagKernel.go:1006-1106): file names, LoadKernel and SaveKernel,
through the C ABI on the CPU (no GPU call is made).

What fixes them:
- names: Go's fmt.Sprint spelling of the key, "[128 32 1]_[0 0 0]_
  [3.90625e-09 3.90625e-09 3e-09]_6_" (the form the reference logs); the
  shortest round-trip digits of every float agree with Python's repr (David
  Gay's algorithm) for random values;
- files: the reference-written fixture test/testdata/randommag4x4x1.ovf is
  reproduced byte for byte by the Python writer (same header spelling), and
  the library's SaveKernel writes the same bytes as the Python writer for a
  scalar field; LoadKernel reads the reference's scalar fixture
  (scalarovf2.ovf) and the Python writer's binary 4 and text files.
"""
import os
import random

import numpy as np
import pytest

from umagnus_amd import ovf, umhip

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ovf")


def test_basename_spelling():
    b = umhip.kernel_cache_basename("/c", (128, 32, 1), (500e-9 / 128, 125e-9 / 32, 3e-9))
    assert b == "/c/uMagNUSkernel_[128 32 1]_[0 0 0]_[3.90625e-09 3.90625e-09 3e-09]_6_"
    b = umhip.kernel_cache_basename("/c", (256, 32, 1), (5e-9, 5e-9, 5e-9), pbc=(1, 0, 0), accuracy=6.5)
    assert b == "/c/uMagNUSkernel_[256 32 1]_[1 0 0]_[5e-09 5e-09 5e-09]_6.5_"
    # Newell: zero_self_demag ends the key, the asymptotic radius is not in it
    b = umhip.kernel_cache_basename("/c", (64, 64, 2), (2.3e-9, 2.3e-9, 1.15e-9), newell=True, zero_self_demag=1)
    assert b == "/c/uMagNUSkernel_[64 64 2]_[0 0 0]_[2.3e-09 2.3e-09 1.15e-09]_1_"
    for acc, s in ((10.0, "10"), (0.5, "0.5"), (1234567.0, "1.234567e+06"), (123456.0, "123456"),
                   (1e-5, "1e-05"), (0.0001, "0.0001")):
        assert umhip.kernel_cache_basename("", (1, 1, 1), (1, 1, 1), accuracy=acc).endswith("_" + s + "_")


def test_float_spelling_matches_repr_digits():
    rng = random.Random(5)
    for _ in range(300):
        v = rng.uniform(0.5, 9.5) * 10.0 ** rng.randint(-12, 12)
        got = umhip.kernel_cache_basename("", (1, 1, 1), (v, 1.0, 1.0)).split("_")[3][1:].split(" ")[0]
        assert float(got) == v
        assert got == ovf.go_float(v)


def test_python_writer_reproduces_reference_file(tmp_path):
    src = os.path.join(GOLD, "randommag4x4x1.ovf")
    d, meta = ovf.read(src)
    out = tmp_path / "m.ovf"
    ovf.write(str(out), d, meta["cellsize"], name=meta["title"], unit=meta["unit"], time=meta["time"])
    assert out.read_bytes() == open(src, "rb").read()


def test_save_kernel_bytes_and_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    K = rng.standard_normal((3, 5, 8)) * 1e-3
    cs = (2e-9, 3e-9, 2.5e-9)
    a, b = tmp_path / "lib.ovf", tmp_path / "py.ovf"
    umhip.save_kernel(a, K, cs, "N_01")
    ovf.write(str(b), K.astype(np.float32), cs, name="N_01", unit="1", time=0.0)
    assert a.read_bytes() == b.read_bytes()
    back = umhip.load_kernel(a, (8, 5, 3))
    assert np.array_equal(back, K.astype(np.float32).astype(np.float64))
    d, meta = ovf.read(str(a))
    assert meta["title"] == "N_01" and meta["size"] == [8, 5, 3] and meta["ncomp"] == 1
    assert meta["cellsize"] == list(cs)


def test_load_kernel_fixture_and_formats(tmp_path):
    d, meta = ovf.read(os.path.join(GOLD, "scalarovf2.ovf"))
    K = umhip.load_kernel(os.path.join(GOLD, "scalarovf2.ovf"), meta["size"])
    assert np.array_equal(K, d[0].astype(np.float64))
    t = tmp_path / "t.ovf"
    X = np.arange(24, dtype=np.float64).reshape(2, 3, 4) / 7
    ovf.write(str(t), X, (1e-9, 1e-9, 1e-9), name="N_00", fmt="text")
    assert np.allclose(umhip.load_kernel(t, (4, 3, 2)), X, rtol=1e-6)


def test_load_kernel_rejects(tmp_path):
    p = tmp_path / "k.ovf"
    umhip.save_kernel(p, np.ones((1, 2, 2)), (1e-9,) * 3, "N_00")
    with pytest.raises(umhip.UmError):
        umhip.load_kernel(p, (2, 2, 2))                    # wrong size
    with pytest.raises(umhip.UmError):
        umhip.load_kernel(tmp_path / "missing.ovf", (2, 2, 1))
    bad = tmp_path / "bad.ovf"
    bad.write_bytes(p.read_bytes().replace(b"# meshtype", b"# meshtipe"))
    with pytest.raises(umhip.UmError):
        umhip.load_kernel(bad, (2, 2, 1))                  # unknown key, as readHeader panics
    v = tmp_path / "v.ovf"
    ovf.write(str(v), np.ones((3, 1, 2, 2)), (1e-9,) * 3, name="m")
    with pytest.raises(umhip.UmError):
        umhip.load_kernel(v, (2, 2, 1))                    # not a scalar field
