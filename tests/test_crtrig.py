"""The exact right-hand sides' sin and cos (gargantua_amd/csrc/gg_crtrig.hpp),
compiled for the host from the device source, against correctly rounded
values (mpmath at 40 digits, rounded to the nearest double) on polar angles,
points near multiples of pi/2 and the table's knots; and the host libm's
agreement with the same values, for scale."""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("crtrig") / "libhostcrtrig.so")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-Wno-unknown-pragmas", "-fPIC", "-shared",
                    "-I", os.path.join(ROOT, "gargantua_amd", "csrc"), "-o", so,
                    os.path.join(ROOT, "tests", "host_crtrig.cpp")], check=True)
    lb = C.CDLL(so)
    for f in (lb.host_sincos_cr, lb.host_libm_sincos):
        f.argtypes = [C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p]
        f.restype = None
    return lb


def _run(fn, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    s, c = np.empty_like(x), np.empty_like(x)
    fn(len(x), x.ctypes.data, s.ctypes.data, c.ctypes.data)
    return s, c


def _cr(x):
    import mpmath
    mpmath.mp.dps = 40
    s = np.array([float(mpmath.sin(mpmath.mpf(float(v)))) for v in x])
    c = np.array([float(mpmath.cos(mpmath.mpf(float(v)))) for v in x])
    return s, c


def _arguments():
    rng = np.random.default_rng(11)
    x = [rng.uniform(-0.5, math.pi + 0.5, 6000)]
    for k in range(-2, 5):                          # within a few ulps and a few 1e-12 of k pi/2
        c = k * math.pi / 2
        x.append(np.array([np.nextafter(c, c + d * 1.0) for d in (-1, 1)] + [c]))
        x.append(c + rng.uniform(-1e-12, 1e-12, 50))
        x.append(c * (1.0 + rng.integers(-8, 9, 17) * 2.0 ** -52))
    knots = np.arange(0, 202) / 256.0                # the table's knots and the half-way points
    x += [knots, knots + 1.0 / 512.0, math.pi / 2 - knots, math.pi - knots]
    x.append(np.array([0.0, 1e-300, 1e-20, 1e-8, 2.0 ** -27]))
    return np.concatenate(x)


def test_sincos_cr_is_correctly_rounded(lib):
    x = _arguments()
    s, c = _run(lib.host_sincos_cr, x)
    rs, rc = _cr(x)
    bad_s = np.count_nonzero(s != rs)
    bad_c = np.count_nonzero(c != rc)
    ls, lc = _run(lib.host_libm_sincos, x)
    print(f"{len(x)} arguments: sincos_cr differs from the correctly rounded value on {bad_s} (sin) and {bad_c} "
          f"(cos); the host libm on {np.count_nonzero(ls != rs)} and {np.count_nonzero(lc != rc)}")
    assert bad_s == 0 and bad_c == 0


def test_sincos_cr_large_and_special_arguments(lib):
    x = np.array([2.0 ** 21, -1e6 * 3.0, 1e300, np.inf, -np.inf, np.nan])
    s, c = _run(lib.host_sincos_cr, x)
    ls, lc = _run(lib.host_libm_sincos, x)
    np.testing.assert_array_equal(s[:3], ls[:3])
    np.testing.assert_array_equal(c[:3], lc[:3])
    assert np.all(np.isnan(s[3:])) and np.all(np.isnan(c[3:]))
