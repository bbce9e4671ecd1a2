"""The G2 MSM's GLS path (four coefficients below 2^64 of Q, phi(Q), psi(Q), phi(psi(Q)), where
psi is the Frobenius carried to the twist): the device split against its defining properties and
the kernel's integer steps restated in Python, and G2 MSMs with GLS on and off (ZKGPU_MSM_GLS=0,
the GLV split alone) against the oracle at every window (c = 9..16: 8, 7, 6 or 5 windows over 65 bits), with
special scalars, infinity and repeated bases, and witness-like scalars."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import uniform_fr_words, witness_scalars
from test_gls_constants import LAM, _source_consts, _split, gls_check

pytestmark = pytest.mark.gpu

R = O.R


def _gls(flag, run):
    old = os.environ.get("ZKGPU_MSM_GLS")
    try:
        os.environ["ZKGPU_MSM_GLS"] = flag
        return run()
    finally:
        if old is None:
            os.environ.pop("ZKGPU_MSM_GLS", None)
        else:
            os.environ["ZKGPU_MSM_GLS"] = old


def test_device_split_properties(gpu):
    from zkgpu import _lib

    x, g, _, _ = _source_consts()
    rng = np.random.default_rng(1700)
    special = [0, 1, 2, R - 1, R - 2, R, R + 1, 2 * R - 1, 5 * R + 7, (1 << 256) - 1, 1 << 255, (1 << 64) - 1,
               1 << 64, (1 << 128) - 1, 1 << 128, LAM, O.P % R, (LAM * (O.P % R)) % R]
    words = [O.int_to_words(v % (1 << 256)) for v in special]
    sc = np.concatenate([np.concatenate(words), uniform_fr_words(rng, 20000),
                         rng.integers(0, 1 << 32, 8 * 2000, dtype=np.uint64).astype(np.uint32)])
    n = sc.size // 8
    out = np.zeros(12 * n, np.uint32)
    _lib.check(_lib.load().zkg_test_gls_split(gpu.handle, _lib.u32_ptr(sc), n, _lib.u32_ptr(out)))
    o = out.reshape(n, 4, 3).astype(np.uint64)
    for i in range(n):
        s = O.words_to_int(sc[8 * i: 8 * i + 8])
        k = [int(o[i, t, 0] | (o[i, t, 1] << np.uint64(32))) * (-1 if o[i, t, 2] else 1) for t in range(4)]
        assert all(int(o[i, t, 2]) in (0, 1) for t in range(4))
        assert k == _split(s, x, g), i
        gls_check(s % R, k)


@pytest.mark.parametrize("n,c", [(1, 11), (2, 13), (3, 14), (255, 11), (1000, 13), (4000, 14), (7000, 11),
                                 (20000, 13), (3000, 9), (3000, 10), (5000, 12), (5000, 15), (6000, 16)])
def test_g2_msm_gls_on_and_off_match_oracle(gpu, n, c):
    from test_gpu_msm import affine
    from zkgpu import Group, Msm

    rng = np.random.default_rng(1710 + n + c)
    ks = uniform_fr_words(rng, n)
    sc = uniform_fr_words(rng, n)
    special = [0, 1, R - 1, (1 << 256) - 1, LAM, O.P % R, 1 << 64, (1 << 64) - 1]
    for j, v in enumerate(special[:n]):
        sc[8 * j: 8 * j + 8] = O.int_to_words(v)
    bases = O.multiples(2, O.generator_words(2), ks)
    exp = O.msm(2, bases, sc)
    bb, sb = Msm.upload_bases(gpu, bases), gpu.buffer(sc)
    for flag in ("1", "0"):
        got = _gls(flag, lambda: affine(gpu, Msm.for_group(gpu, Group.G2, n, c).run(gpu, bb, sb), 2))
        assert np.array_equal(got, exp), (flag, n, c)


def test_g2_msm_gls_infinity_and_repeated_bases(gpu):
    """Infinity bases ((0, 0): psi and phi keep them), one base repeated with opposite scalars."""
    from test_gpu_msm import affine
    from zkgpu import Group, Msm

    n = 3000
    rng = np.random.default_rng(1720)
    ks = uniform_fr_words(rng, n)
    ks[0:8 * 100] = 0  # 0 * G = infinity
    ks[8 * 100: 8 * 400] = np.tile(ks[8 * 400: 8 * 401], 300)
    sc = uniform_fr_words(rng, n)
    sc[8 * 100: 8 * 250] = np.tile(O.int_to_words(12345), 150)
    sc[8 * 250: 8 * 400] = np.tile(O.int_to_words(R - 12345), 150)
    bases = O.multiples(2, O.generator_words(2), ks)
    exp = O.msm(2, bases, sc)
    bb, sb = Msm.upload_bases(gpu, bases), gpu.buffer(sc)
    for c in (11, 13):
        got = _gls("1", lambda: affine(gpu, Msm.for_group(gpu, Group.G2, n, c).run(gpu, bb, sb), 2))
        assert np.array_equal(got, exp), c


def test_g2_msm_gls_witness_scalars(gpu):
    from test_gpu_msm import affine
    from zkgpu import Group, Msm

    n = 6000
    rng = np.random.default_rng(1730)
    ks = uniform_fr_words(rng, n)
    sc = witness_scalars(rng, n)
    bases = O.multiples(2, O.generator_words(2), ks)
    exp = O.msm(2, bases, sc)
    bb, sb = Msm.upload_bases(gpu, bases), gpu.buffer(sc)
    got = _gls("1", lambda: affine(gpu, Msm.for_group(gpu, Group.G2, n, 13).run(gpu, bb, sb), 2))
    assert np.array_equal(got, exp)



def test_g2_msm_gls_with_poisoned_workspace(gpu):
    """Every workspace array (the psi table and the phi table of both halves included) filled with
    a poison value before the run: the result still equals the oracle's, so no stage reads what
    it has not written."""
    from test_gpu_msm import affine
    from zkgpu import Group, Msm

    n = 2500
    rng = np.random.default_rng(1740)
    ks, sc = uniform_fr_words(rng, n), uniform_fr_words(rng, n)
    bases = O.multiples(2, O.generator_words(2), ks)
    exp = O.msm(2, bases, sc)
    bb, sb = Msm.upload_bases(gpu, bases), gpu.buffer(sc)
    old = os.environ.get("ZKGPU_POISON")
    try:
        os.environ["ZKGPU_POISON"] = "1"
        for c in (11, 13):
            got = _gls("1", lambda: affine(gpu, Msm.for_group(gpu, Group.G2, n, c).run(gpu, bb, sb), 2))
            assert np.array_equal(got, exp), c
    finally:
        if old is None:
            os.environ.pop("ZKGPU_POISON", None)
        else:
            os.environ["ZKGPU_POISON"] = old


def test_g2_msm_gls_seeded_fuzz(gpu):
    """Seeded random G2 MSMs on the GLS path: sizes 1..6000, every window that selects it, scalar
    mixes (uniform, small, witness-like, multiples of lambda and mu, values near r) and repeated
    bases, each against the oracle."""
    from test_gpu_msm import affine
    from zkgpu import Group, Msm

    rng = np.random.default_rng(1750)
    mu = O.P % R
    for case in range(24):
        n = int(rng.choice([1, 2, 5, 63, 64, 65, 511, 1500, 3000, 6000]))
        c = int(rng.choice([9, 10, 11, 12, 13, 14, 15, 16]))
        ks = uniform_fr_words(rng, n)
        kind = case % 4
        if kind == 0:
            sc = uniform_fr_words(rng, n)
        elif kind == 1:
            sc = witness_scalars(rng, n)
        elif kind == 2:  # multiples of the eigenvalues and values near r
            vals = [(int(rng.integers(0, 1 << 62)) * int(rng.choice([1, LAM, mu, LAM * mu % R])) +
                     int(rng.integers(-3, 4))) % R for _ in range(n)]
            sc = np.concatenate([O.int_to_words(v) for v in vals])
        else:  # repeated bases with random scalars
            ks = np.tile(ks[:8], n)
            sc = uniform_fr_words(rng, n)
        bases = O.multiples(2, O.generator_words(2), ks)
        exp = O.msm(2, bases, sc)
        bb, sb = Msm.upload_bases(gpu, bases), gpu.buffer(sc)
        got = _gls("1", lambda: affine(gpu, Msm.for_group(gpu, Group.G2, n, c).run(gpu, bb, sb), 2))
        assert np.array_equal(got, exp), (case, n, c, kind)
