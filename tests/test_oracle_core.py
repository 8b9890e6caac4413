"""Pins the oracle's numeric core without an executable reference.

Each check uses something other than the oracle's own formula: the x86
crc32 instruction and the CRC-32C check value, Appleby's byte-oriented
MurmurHash2, the full period of xorshift128+ (23,18,5) over GF(2), libm
sin/cos, closed-form log sums, and distribution tests (KS / chi-square).
"""
import ctypes as C
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest
from scipy import stats

SALT_Z = 0x9B83516C63E249C3
M64 = (1 << 64) - 1


# ---------------------------------------------------------------- CRC32C
def crc32c_bitwise(prev, data: bytes):
    """textbook bitwise CRC-32C (Castagnoli, reflected 0x82F63B78), raw update"""
    c = prev
    for byte in data:
        c ^= byte
        for _ in range(8):
            c = (c >> 1) ^ (0x82F63B78 if c & 1 else 0)
    return c


def test_crc32c_textbook_check_value():
    # the published CRC-32C check value of "123456789"
    assert crc32c_bitwise(0xFFFFFFFF, b"123456789") ^ 0xFFFFFFFF == 0xE3069283


def test_crc32c_matches_textbook(orc):
    rng = np.random.default_rng(1)
    lib = orc.load()
    for v in rng.integers(0, 2**63, 200, dtype=np.int64).tolist() + [0, 1, M64]:
        v &= M64
        assert lib.or_crc32c_u64(0xB385FDA9, v) == crc32c_bitwise(0xB385FDA9, v.to_bytes(8, "little"))


def test_crc32c_matches_sse42_instruction(orc):
    """the reference computes HashULong with _mm_crc32_u64 (src/hash.cpp:9-16)"""
    src = r"""
    #include <stdint.h>
    uint32_t hw_crc(uint32_t p, uint64_t v) { return (uint32_t)__builtin_ia32_crc32di(p, v); }
    """
    with tempfile.TemporaryDirectory() as d:
        c, so = os.path.join(d, "h.c"), os.path.join(d, "h.so")
        open(c, "w").write(src)
        r = subprocess.run(["gcc", "-O1", "-msse4.2", "-shared", "-fPIC", "-o", so, c], capture_output=True)
        if r.returncode != 0:
            pytest.skip("no SSE4.2 toolchain")
        hw = C.CDLL(so)
        hw.hw_crc.argtypes = [C.c_uint32, C.c_uint64]
        hw.hw_crc.restype = C.c_uint32
        lib = orc.load()
        rng = np.random.default_rng(2)
        for v in rng.integers(0, 2**63, 500, dtype=np.int64).tolist():
            assert lib.or_hash_ulong(v) == hw.hw_crc(0xB385FDA9, v)


# ---------------------------------------------------------------- Murmur
def murmur2_appleby(key: bytes, seed):
    """Austin Appleby's MurmurHash2 (32-bit), byte-oriented published form"""
    m, r = 0x5BD1E995, 24
    h = (seed ^ len(key)) & 0xFFFFFFFF
    n = len(key)
    i = 0
    while n >= 4:
        k = int.from_bytes(key[i:i + 4], "little")
        k = (k * m) & 0xFFFFFFFF
        k ^= k >> r
        k = (k * m) & 0xFFFFFFFF
        h = (h * m) & 0xFFFFFFFF
        h ^= k
        i += 4
        n -= 4
    if n == 3:
        h ^= key[i + 2] << 16
    if n >= 2:
        h ^= key[i + 1] << 8
    if n >= 1:
        h ^= key[i]
        h = (h * m) & 0xFFFFFFFF
    h ^= h >> 13
    h = (h * m) & 0xFFFFFFFF
    h ^= h >> 15
    return h


def test_murmur2_is_appleby_4byte(orc):
    lib = orc.load()
    rng = np.random.default_rng(3)
    for _ in range(300):
        d, s = (int(x) for x in rng.integers(0, 2**32, 2, dtype=np.uint64))
        assert lib.or_murmur2(d, s) == murmur2_appleby(d.to_bytes(4, "little"), s)


def test_murmur64_composition(orc):
    """MurmurHash64 = (M2(~lo ^ hi) << 32) | M2(lo), src/hash.cpp:450-460"""
    lib = orc.load()
    rng = np.random.default_rng(4)
    for _ in range(200):
        d, s = (int(x) & M64 for x in rng.integers(0, 2**63, 2, dtype=np.int64))
        d1 = (~d ^ (d >> 32)) & 0xFFFFFFFF
        s1 = (~s ^ (s >> 32)) & 0xFFFFFFFF
        hi = murmur2_appleby(d1.to_bytes(4, "little"), s1)
        lo = murmur2_appleby((d & 0xFFFFFFFF).to_bytes(4, "little"), s & 0xFFFFFFFF)
        assert lib.or_murmur64(d, s) == (hi << 32) | lo


# ---------------------------------------------------------------- xorshift128+
def _step_matrix(orc):
    """columns of the one-step map on the 128-bit state, taken from the oracle"""
    cols = np.zeros((128, 2), dtype=np.uint64)
    for j in range(128):
        r = orc.Rng(0, 0)
        r.x, r.y = (1 << j if j < 64 else 0), (1 << (j - 64) if j >= 64 else 0)
        r.xorshift()
        cols[j] = (r.x, r.y)
    return cols


def _mat_mul(A, B):
    """GF(2) product of column-represented 128x128 matrices: (A B) e_j = A (B e_j)"""
    C_ = np.zeros_like(A)
    for b in range(128):
        word, bit = (0, b) if b < 64 else (1, b - 64)
        sel = ((B[:, word] >> np.uint64(bit)) & np.uint64(1)).astype(bool)
        C_[sel] ^= A[b]
    return C_


def _mat_pow(T, e):
    R = np.zeros_like(T)
    for j in range(128):
        R[j] = (1 << j, 0) if j < 64 else (0, 1 << (j - 64))
    P = T.copy()
    while e:
        if e & 1:
            R = _mat_mul(P, R)
        P = _mat_mul(P, P)
        e >>= 1
    return R


def _is_identity(M):
    for j in range(128):
        exp = (1 << j, 0) if j < 64 else (0, 1 << (j - 64))
        if int(M[j, 0]) != exp[0] or int(M[j, 1]) != exp[1]:
            return False
    return True


def test_xorshift128p_full_period(orc):
    """xorshift128+ with shifts (23,18,5) has period 2^128 - 1 (Vigna): the step
    matrix T satisfies T^(2^128-1) = I and T^((2^128-1)/p) != I for every prime
    p | 2^128 - 1.  A wrong shift amount or direction breaks this."""
    T = _step_matrix(orc)
    n = (1 << 128) - 1
    assert _is_identity(_mat_pow(T, n))
    for p in (3, 5, 17, 257, 641, 65537, 274177, 6700417, 67280421310721):
        assert n % p == 0
        assert not _is_identity(_mat_pow(T, n // p)), p


def test_xorshift_output_and_seed(orc):
    """RNG(s, salt): x = C1 ^ h, y = C2 ^ (h << 6), h = MurmurHash64(s, salt);
    XorShift returns new_x + new_y (src/statistics.cpp:62-86)"""
    lib = orc.load()
    r = orc.Rng(12345, SALT_Z)
    h = lib.or_murmur64(12345, SALT_Z)
    assert r.x == 0x159A55E5075BCD15 ^ h
    assert r.y == (0x054913331F123BB5 ^ (h << 6)) & M64
    x, y = r.x, r.y
    out = r.xorshift()
    assert r.x == y
    assert out == (r.x + r.y) & M64


def test_jump_table_equals_stepping(vps, orc):
    for D in (1, 2, 7, 64, 1000, 123457):
        r = orc.Rng(99, SALT_Z)
        x0, y0 = r.x, r.y
        for _ in range(D):
            r.xorshift()
        assert vps.debug_jump(D, x0, y0) == (r.x, r.y)


def test_jump96_table_equals_stepping(vps, orc):
    """the xorshift96 (RNG<float>) chunk jumps of the fast single-precision Z pass"""
    for D in (1, 2, 7, 64, 1000, 123457):
        r = orc.RngF(99, SALT_Z)
        x0, y0, z0 = r.x, r.y, r.z
        for _ in range(D):
            r.xorshift()
        assert vps.debug_jump96(D, x0, y0, z0) == (r.x, r.y, r.z)


def test_uniform_bits(orc):
    r1, r2 = orc.Rng(5, 7), orc.Rng(5, 7)
    for _ in range(100):
        u = r1.uniform(0.0, 1.0)
        v = r2.xorshift()
        expect = C.c_double.from_buffer_copy(((v & ((1 << 52) - 1)) | 0x3FF0000000000000).to_bytes(8, "little")).value - 1.0
        assert u == expect


# ---------------------------------------------------------------- math
def test_mysin_mycos_against_libm(orc):
    """The reference's minimax polynomials (src/math2.h:1087-1142) are accurate
    to ~4e-12 on the Box-Muller range [0, 2pi + pi/2]; a wrong quadrant, sign or
    reduction constant gives O(1) or >=1e-9 errors."""
    lib = orc.load()
    xs = np.linspace(-12.0, 12.0, 20001)
    e1 = max(abs(lib.or_mysin(x) - math.sin(x)) for x in xs)
    e2 = max(abs(lib.or_mycos(x) - math.cos(x)) for x in xs)
    assert e1 < 1e-11 and e2 < 1e-11


def test_mylog_bounds(orc):
    lib = orc.load()
    assert lib.or_mylog(0.0) == -6.90775527898214000e2
    assert lib.or_mylog(1e301) == 6.90775527898214000e2
    assert lib.or_mylog(2.5) == math.log(2.5)


def test_log_accumulator_closed_form(orc):
    rng = np.random.default_rng(5)
    for scale in (1.0, 1e-120, 1e150):
        v = rng.uniform(0.01, 1.0, 5000) * scale
        got = orc.charge_log(v)
        exact = float(np.sum(np.log(v)))
        assert abs(got - exact) <= 1e-12 * abs(exact) + 1e-9


def test_logprod_avx_closed_form(orc):
    lib = orc.load()
    rng = np.random.default_rng(6)
    for n in (5, 31, 32, 33, 1000, 4097):
        v = np.ascontiguousarray(rng.uniform(1e-3, 1.0, n))
        got = lib.or_logprod_avx(v.ctypes.data, n)
        assert abs(got - np.sum(np.log(v))) <= 1e-12 * abs(np.sum(np.log(v))) + 1e-10


def test_unify_avx(orc):
    lib = orc.load()
    for n in (2, 3, 4, 5, 9):
        a = np.ascontiguousarray(np.arange(1, n + 1, dtype=np.float64))
        lib.or_unify_avx(a.ctypes.data, n)
        expect = (np.arange(1, n + 1) + 1e-10) / (n * (n + 1) / 2 + n * 1e-10)
        assert np.allclose(a, expect, rtol=1e-15, atol=0)


# ---------------------------------------------------------------- distributions
def test_normal_ks_and_cache(orc):
    r = orc.Rng(11, 22)
    x = np.array([r.normal() for _ in range(40000)])
    assert stats.kstest(x, "norm").pvalue > 1e-3
    # the second variate of each pair comes from the cache without a draw
    r1 = orc.Rng(3, 4)
    r1.normal()
    x1 = r1.x
    r1.normal()
    assert r1.x == x1


@pytest.mark.parametrize("alpha", [0.3, 1.0, 2.5, 11.0])
def test_gamma_ks(orc, alpha):
    r = orc.Rng(int(alpha * 1000), 9)
    x = np.array([r.gamma(alpha) for _ in range(20000)])
    assert stats.kstest(x, "gamma", args=(alpha,)).pvalue > 1e-3


def test_dirichlet_mean(orc):
    r = orc.Rng(1, 2)
    a = np.array([0.5, 2.0, 3.0])
    b = np.array([1, 0, 4])
    x = np.array([r.dirichlet_ab(a, b) for _ in range(20000)])
    assert np.allclose(x.mean(0), (a + b) / (a + b).sum(), atol=0.01)
    assert np.allclose(x.sum(1), 1.0, atol=1e-12)


def test_poly_avx_exact_rule(orc):
    """k* = first k with t < w_k after subtracting w_0..w_{k-1}, t = (U-1) s"""
    rng = np.random.default_rng(8)
    for _ in range(2000):
        K = int(rng.integers(1, 9))
        w = rng.uniform(0, 1, K) + 1e-10
        r1, r2 = orc.Rng(int(rng.integers(1 << 40)), 1), None
        r2 = orc.Rng(0, 0)
        r2.x, r2.y = r1.x, r1.y
        got = r1.poly_avx(w)
        u = r2.xorshift()
        s = 0.0
        for wk in w:
            s += wk
        t = (C.c_double.from_buffer_copy(((u & ((1 << 52) - 1)) | 0x3FF0000000000000).to_bytes(8, "little")).value - 1.0) * s
        k = K - 1
        for j, wk in enumerate(w):
            if t < wk:
                k = j
                break
            t -= wk
        assert got == k


def test_poly_avx_distribution(orc):
    w = np.array([0.1, 0.5, 0.15, 0.25]) + 1e-10
    r = orc.Rng(77, 1)
    cnt = np.bincount([r.poly_avx(w) for _ in range(40000)], minlength=4)
    assert stats.chisquare(cnt, w / w.sum() * cnt.sum()).pvalue > 1e-3


def test_polylog_distribution(orc):
    a = np.log(np.array([0.2, 0.3, 0.5]))
    r = orc.Rng(5, 5)
    cnt = np.bincount([r.polylog(a) for _ in range(30000)], minlength=3)
    assert stats.chisquare(cnt, np.array([0.2, 0.3, 0.5]) * cnt.sum()).pvalue > 1e-3


def test_next_avoid(orc):
    r = orc.Rng(1, 1)
    for _ in range(500):
        av = int(r.next(6))
        x = r.next_avoid(6, av)
        assert x != av and 0 <= x < 6
