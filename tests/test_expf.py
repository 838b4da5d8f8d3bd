"""The product's exp (vhm_math.h expf_glibc, the function every Gaussian
weight in the HIP kernels calls) against the host glibc expf that the
reference's std::exp(float) resolves to.

Pins: (1) the 32-entry table equals 2^(i/32) recomputed here from first
principles with 60-digit decimal arithmetic; (2) a strided sweep over the
whole domain the fusion path uses (x <= 0 down to the underflow bound) and
over x > 0, plus the special values.  An exhaustive run over all 2.24e9
floats in [-103, 88] found no mismatch.
(3) The octahedral dequantisation's reciprocal-FMA quotient against the
division it replaces, for all 65,536 codes.
"""
import ctypes as C
import math
import struct
from decimal import Decimal, getcontext
from pathlib import Path

import numpy as np
import pytest

LIB = Path(__file__).resolve().parents[1] / "oracle" / "libexpf_check.so"


@pytest.fixture(scope="module")
def chk():
    if not LIB.exists():
        pytest.fail(f"{LIB} missing: run `make -C oracle`")
    L = C.CDLL(str(LIB))
    L.vcheck_expf_sweep.restype = C.c_uint64
    L.vcheck_expf_sweep.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32)]
    L.vcheck_exp_table.argtypes = [C.c_void_p]
    L.vcheck_expf_product.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    L.vcheck_expf_libm.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    return L


def test_table_is_two_to_the_i_over_32(chk):
    getcontext().prec = 60
    got = np.zeros(32, np.uint64)
    chk.vcheck_exp_table(got.ctypes.data)
    for i in range(32):
        d = float(Decimal(2) ** (Decimal(i) / Decimal(32)))  # correctly rounded to double
        bits = struct.unpack("<Q", struct.pack("<d", d))[0]
        assert int(got[i]) == (bits - (i << 47)) % (1 << 64), i


@pytest.mark.parametrize("lo,hi,stride", [
    (0x80000000, 0xC2D00000, 97),     # -0 ... -104 (every Gaussian argument is <= 0)
    (0xC2AE0000, 0xC2D00000, 1),      # -87 ... -104: subnormal results, dense
    (0x00000000, 0x42B20000, 101),    # 0 ... 89 (overflow bound included)
    (0xBF800000, 0xBF900000, 1),      # around -1, dense
])
def test_sweep_matches_glibc(chk, lo, hi, stride):
    first = C.c_uint32()
    bad = chk.vcheck_expf_sweep(lo, hi, stride, C.byref(first))
    assert bad == 0, f"{bad} mismatches, first at bits 0x{first.value:08x}"


def test_special_values(chk):
    x = np.array([0.0, -0.0, -np.inf, np.inf, np.nan, -1e-45, -103.97, -103.972, -200.0, 88.72, 88.73,
                  -0.5, -1.0, -2.0, -4.5], np.float32)
    a, g = np.zeros_like(x), np.zeros_like(x)
    chk.vcheck_expf_product(x.ctypes.data, a.ctypes.data, len(x))
    chk.vcheck_expf_libm(x.ctypes.data, g.ctypes.data, len(x))
    assert np.array_equal(a.view(np.uint32)[~np.isnan(g)], g.view(np.uint32)[~np.isnan(g)])
    assert np.isnan(a[np.isnan(g)]).all()
    assert a[0] == 1.0 and a[2] == 0.0 and math.isinf(a[3])


def test_dequant_equals_division_for_every_code(chk):
    chk.vcheck_dequant_all.argtypes = [C.c_void_p, C.c_void_p]
    prod = np.zeros(65536, np.float32)
    div = np.zeros(65536, np.float32)
    chk.vcheck_dequant_all(prod.ctypes.data, div.ctypes.data)
    ref = (np.arange(65536, dtype=np.float32) / np.float32(65535.0)) * np.float32(2.0) - np.float32(1.0)
    assert np.array_equal(div.view(np.uint32), ref.view(np.uint32))  # the plain division, as numpy rounds it
    assert np.array_equal(prod.view(np.uint32), div.view(np.uint32))


def test_norm_threshold_is_the_sqrt_boundary():
    """norm3(v) <= 1e-6f  <=>  dot(v, v) <= 0x1.19799ap-40 (the constant in
    vhm_math.h): sqrtf is correctly rounded and monotone, so the boundary
    and its neighbourhood decide it."""
    T = np.float32(float.fromhex("0x1.19799ap-40"))
    lim = np.float32(1e-6)
    bits = T.view(np.uint32).astype(np.int64)
    s = (np.arange(bits - 100000, bits + 100000)).astype(np.uint32).view(np.float32)
    assert np.array_equal(np.sqrt(s) <= lim, s <= T)
