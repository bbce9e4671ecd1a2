"""The MSM's GLV path: the device scalar split against a decomposition derived here from first
principles (a cube root of unity lambda mod r, the reduced lattice basis from the extended
Euclidean algorithm, Babai rounding), and MSM results with the split on and off against the
oracle."""
import math
import os

import numpy as np
import pytest

import oracle as O
from conftest import uniform_fr_words

pytestmark = pytest.mark.gpu

R = O.R
P = 0x30644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD47


def _lambda():
    """The cube root of unity mod r whose endomorphism (x, y) -> (beta x, y) is checked below."""
    roots = sorted({pow(g, (R - 1) // 3, R) for g in range(2, 20)} - {1})
    return roots[0]


def _basis(lam):
    """Reduced basis of {(a, b) : a + b lam = 0 mod r}: extended Euclid on (r, lam) stopped at
    the first remainder below sqrt(r)."""
    r0, r1, t0, t1 = R, lam, 0, 1
    seq = [(r0, t0), (r1, t1)]
    while r1:
        q = r0 // r1
        r0, r1 = r1, r0 - q * r1
        t0, t1 = t1, t0 - q * t1
        seq.append((r1, t1))
    s = math.isqrt(R)
    k = max(i for i, (ri, _) in enumerate(seq) if ri >= s)
    v1 = (seq[k + 1][0], -seq[k + 1][1])
    v2 = min([(seq[k][0], -seq[k][1]), (seq[k + 2][0], -seq[k + 2][1])], key=lambda v: v[0] ** 2 + v[1] ** 2)
    return v1, v2


def _split(s, lam, v1, v2):
    """Babai rounding with 2^256-scaled ratios and floor, as the device computes it."""
    (a1, b1), (a2, b2) = v1, v2
    det = a1 * b2 - a2 * b1
    assert det == R

    def rdiv(x, d):
        return (2 * x + d) // (2 * d)

    g1, g2 = rdiv(b2 << 256, det), rdiv((-b1) << 256, det)
    s %= R
    c1, c2 = (s * g1) >> 256, (s * g2) >> 256
    return s - c1 * a1 - c2 * a2, -c1 * b1 - c2 * b2


def _decode(words4):
    w = [int(x) for x in words4]
    mag = w[0] | (w[1] << 32) | (w[2] << 64) | ((w[3] & 0x7FFFFFFF) << 96)
    return -mag if w[3] >> 31 else mag


def test_lambda_is_the_endomorphism_eigenvalue():
    lam = _lambda()
    assert lam == 0xB3C4D79D41A917585BFC41088D8DAAA78B17EA66B99C90DD
    beta = 0x59E26BCEA0D48BACD4F263F1ACDB5C4F5763473177FFFFFE
    assert pow(beta, 3, P) == 1 and beta != 1
    g = O.py_g1_mul(O.G1, 1)
    assert O.py_g1_mul(O.G1, lam) == (g[0] * beta % P, g[1])
    beta2 = P - 1 - beta
    assert O.py_g2_mul(O.G2, lam) == (O.py_f2_mul(O.G2[0], (beta2, 0)), O.G2[1])


def test_device_split_matches_decomposition(gpu):
    from zkgpu import _lib

    lam = _lambda()
    v1, v2 = _basis(lam)
    # the bound the 31-bit sign layout relies on: |k| <= 1.125 (|a1| + |a2|) < 2^127
    assert 9 * (abs(v1[0]) + abs(v2[0])) < 8 << 127 and 9 * (abs(v1[1]) + abs(v2[1])) < 8 << 127
    rng = np.random.default_rng(900)
    special = [0, 1, 2, R - 1, R - 2, lam, lam + 1, R - lam, (1 << 127) - 1, 1 << 127, (1 << 253) - 1,
               R, R + 1, 2 * R - 1, (1 << 256) - 1, 1 << 255, 5 * R + 7]
    words = [O.int_to_words(x % (1 << 256)) for x in special]
    rand = uniform_fr_words(rng, 20000)
    sc = np.concatenate([np.concatenate(words), rand])
    n = sc.size // 8
    out = np.zeros(8 * n, np.uint32)
    _lib.check(_lib.load().zkg_test_glv_split(gpu.handle, _lib.u32_ptr(sc), n, _lib.u32_ptr(out)))
    k1s, k2s = out[: 4 * n].reshape(n, 4), out[4 * n:].reshape(n, 4)
    for i in range(n):
        s = O.words_to_int(sc[8 * i: 8 * i + 8])
        k1, k2 = _decode(k1s[i]), _decode(k2s[i])
        assert (k1, k2) == _split(s, lam, v1, v2), i
        assert (k1 + k2 * lam - s) % R == 0
        assert abs(k1) < 1 << 127 and abs(k2) < 1 << 127


@pytest.mark.parametrize("group,n,c", [(1, 3000, 9), (1, 5000, 13), (1, 70000, 16), (2, 1500, 11), (2, 4000, 12),
                                       (1, 2000, 17), (1, 1200, 20), (2, 600, 18)])
def test_msm_glv_on_and_off_match_oracle(gpu, group, n, c):
    """Both the GLV split and the plain digit path, at windows covering the LDS and the atomic
    sort, agree with the oracle; the scalars include 0, 1, r - 1 and non-canonical values."""
    from test_gpu_msm import affine
    from zkgpu import Group, Msm

    rng = np.random.default_rng(910 + n + c)
    ks = uniform_fr_words(rng, n)
    sc = uniform_fr_words(rng, n)
    sc[0:8] = 0
    sc[8:16] = O.int_to_words(1)
    sc[16:24] = O.int_to_words(R - 1)
    sc[24:32] = O.int_to_words((1 << 256) - 1)  # non-canonical: the MSM uses it as is, i.e. mod r
    sc[32:40] = O.int_to_words(_lambda())
    bases = O.multiples(group, O.generator_words(group), ks)
    exp = O.msm(group, bases, sc)
    bb = Msm.upload_bases(gpu, bases)
    # the plain digit path, like the reference, covers ceil(255 / c) windows: canonical scalars only
    sc_plain = sc.copy()
    sc_plain[24:32] = O.int_to_words(((1 << 256) - 1) % R)
    old = os.environ.get("ZKGPU_MSM_GLV")
    try:
        for flag in ("1", "0"):
            os.environ["ZKGPU_MSM_GLV"] = flag
            sb = gpu.buffer(sc if flag == "1" else sc_plain)
            m = Msm.for_group(gpu, Group(group), n, c)
            assert np.array_equal(affine(gpu, m.run(gpu, bb, sb), group), exp), (flag, group, n, c)
    finally:
        if old is None:
            os.environ.pop("ZKGPU_MSM_GLV", None)
        else:
            os.environ["ZKGPU_MSM_GLV"] = old
