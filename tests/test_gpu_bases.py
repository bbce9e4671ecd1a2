"""Base points at the boundary: Montgomery-2^256 words at or above p give the same bases and the
same MSM as canonical words (the reference's upload reduces any 256-bit word with its scaling
product, src/msm.rs:190-197); the device base check accepts infinity and curve (G2: subgroup)
points and rejects the rest; a G2 base on the twist but outside the order-r subgroup is flagged
by the check and summed exactly by a plan without the GLV split; .zkey sections with swapped
coordinates or plain (non-Montgomery) words are rejected at upload."""
import contextlib
import os

import numpy as np
import pytest

import oracle as O
import zkey_file
from conftest import uniform_fr_words

pytestmark = pytest.mark.gpu

MONT = 1 << 256


@contextlib.contextmanager
def _env(name, value):
    old = os.environ.get(name)
    os.environ[name] = value
    try:
        yield
    finally:
        if old is None:
            os.environ.pop(name, None)
        else:
            os.environ[name] = old


def _affine(gpu, res, group):
    from test_gpu_batch import _affine as aff

    return aff(gpu, res, group)


def _mont_words(plain: np.ndarray) -> np.ndarray:
    rows = np.asarray(plain, np.uint32).reshape(-1, 8)
    return np.concatenate([O.int_to_words(O.words_to_int(r) * MONT % O.P) for r in rows])


def _f2_pow(a, e):
    acc = (1, 0)
    while e:
        if e & 1:
            acc = O.py_f2_mul(acc, a)
        a = O.py_f2_mul(a, a)
        e >>= 1
    return acc


def _fq2_sqrt(a):
    """A square root of a in Fq2 = Fq[u]/(u^2 + 1), or None. p = 3 mod 4: with n = sqrt(N(a)) in Fq,
    x0^2 = (a0 +- n) / 2 and x1 = a1 / (2 x0) give (x0 + x1 u)^2 = a."""
    a0, a1 = a
    norm = (a0 * a0 + a1 * a1) % O.P
    n = pow(norm, (O.P + 1) // 4, O.P)
    if n * n % O.P != norm:
        return None
    inv2 = (O.P + 1) // 2
    for t in ((a0 + n) * inv2 % O.P, (a0 - n) * inv2 % O.P):
        x0 = pow(t, (O.P + 1) // 4, O.P)
        if x0 * x0 % O.P != t or x0 == 0:
            continue
        x = (x0, a1 * pow(2 * x0, O.P - 2, O.P) % O.P)
        if O.py_f2_mul(x, x) == (a0 % O.P, a1 % O.P):
            return x
    return None


def twist_points_outside_subgroup(rng, count):
    """Points of the twist y^2 = x^3 + 3/(9+u) with r P != O (random x; the cofactor 2p - r leaves
    almost every such point outside the order-r subgroup)."""
    out = []
    while len(out) < count:
        x = (int(rng.integers(1, 1 << 62)) * 0x9E3779B97F4A7C15 % O.P, int(rng.integers(1, 1 << 62)))
        y = _fq2_sqrt(O.py_f2_add(O.py_f2_mul(O.py_f2_mul(x, x), x), O.B2))
        if y is None:
            continue
        pt = (x, y)
        assert O.py_g2_on_curve(pt)
        if O.py_g2_mul(pt, O.R) is not None:
            out.append(pt)
    return out


@pytest.mark.parametrize("group", [1, 2])
@pytest.mark.parametrize("f29", ["1", "0"])
def test_montgomery_256_words_at_or_above_p(gpu, group, f29):
    """Each coordinate stored as m + k p (m its canonical Montgomery word, the largest k with
    m + k p < 2^256, so up to 5): the upload leaves the canonical words and the MSM equals the
    oracle's, on the 29-bit and 32-bit pipelines, with scalars whose digits take both signs."""
    from zkgpu import Group, Msm

    rng = np.random.default_rng(1700 + group)
    n = 3000
    ks = uniform_fr_words(rng, n)
    ks[8 * 11: 8 * 12] = 0  # one base at infinity: its words stay zero
    plain = O.multiples(group, O.generator_words(group), ks)
    canon = _mont_words(plain)
    rows = canon.reshape(-1, 8)
    big = np.empty_like(rows)
    ks_used = []
    for i, r in enumerate(rows):
        m = O.words_to_int(r)
        k = 0 if m == 0 else min((MONT - 1 - m) // O.P, int(rng.integers(0, 6)))
        ks_used.append(k)
        big[i] = O.int_to_words(m + k * O.P)
    assert max(ks_used) == 5 and min(ks_used) == 0
    sc = uniform_fr_words(rng, n)
    with _env("ZKGPU_MSM_F29", f29):
        buf = Msm.upload_bases_montgomery_256(gpu, big.reshape(-1))
        assert np.array_equal(gpu.read(buf, canon.size), canon)
        m = Msm.for_group(gpu, Group(group), n, 12)
        got = _affine(gpu, m.run(gpu, buf, gpu.buffer(sc)), group)
    assert np.array_equal(got, O.msm(group, plain, sc))


def test_check_bases_accepts_curve_points_and_rejects_the_rest(gpu):
    from zkgpu import Group, Msm

    rng = np.random.default_rng(1710)
    # G1: generator multiples and infinity pass
    ks = uniform_fr_words(rng, 1000)
    ks[8 * 3: 8 * 4] = 0
    g1 = O.multiples(1, O.generator_words(1), ks)
    assert Msm.check_bases(gpu, Group.G1, Msm.upload_bases(gpu, g1), 1000) == (0, 0)
    bad = g1.copy().reshape(-1, 16)
    bad[17, 8] ^= 1  # y changed: off the curve
    bad[500] = np.concatenate([bad[500, 8:], bad[500, :8]])  # x and y swapped
    assert Msm.check_bases(gpu, Group.G1, Msm.upload_bases(gpu, bad.reshape(-1)), 1000) == (2, 17)
    assert Msm.check_bases(gpu, Group.G1, Msm.upload_bases(gpu, bad.reshape(-1)), 17) == (0, 0)
    # G2: subgroup points and infinity pass; an off-twist point and a twist point outside the
    # order-r subgroup fail
    ks2 = uniform_fr_words(rng, 300)
    ks2[0:8] = 0
    g2 = O.multiples(2, O.generator_words(2), ks2)
    assert Msm.check_bases(gpu, Group.G2, Msm.upload_bases(gpu, g2), 300) == (0, 0)
    bad2 = g2.copy().reshape(-1, 32)
    bad2[123] = O.g2_words(twist_points_outside_subgroup(rng, 1)[0])
    bad2[200, 16] ^= 4  # y.a changed: off the twist
    assert Msm.check_bases(gpu, Group.G2, Msm.upload_bases(gpu, bad2.reshape(-1)), 300) == (2, 123)


def test_g2_bases_outside_the_subgroup(gpu):
    """The oracle's sum over twist points outside the subgroup is first checked against the
    textbook law; a plan without the GLV split gives it exactly; the check flags those bases, and a
    GLV plan (whose endomorphism acts as lambda only on the subgroup) does not give it."""
    from zkgpu import Group, Msm

    rng = np.random.default_rng(1720)
    n = 24
    pts = twist_points_outside_subgroup(rng, 6)
    ks = uniform_fr_words(rng, n)
    g2 = O.multiples(2, O.generator_words(2), ks).reshape(-1, 32)
    for j, pt in enumerate(pts):
        g2[4 * j + 1] = O.g2_words(pt)
    bases = g2.reshape(-1)
    sc = uniform_fr_words(rng, n)
    exp = O.msm(2, bases, sc)
    acc = None
    for i in range(n):
        acc = O.py_g2_add(acc, O.py_g2_mul(O.words_to_g2(g2[i]), O.words_to_int(sc[8 * i: 8 * i + 8])))
    assert np.array_equal(exp, O.g2_words(acc))
    buf = Msm.upload_bases(gpu, bases)
    assert Msm.check_bases(gpu, Group.G2, buf, n) == (6, 1)
    with _env("ZKGPU_MSM_GLV", "0"):
        plain = Msm.for_group(gpu, Group.G2, n, 10)
        assert np.array_equal(_affine(gpu, plain.run(gpu, buf, gpu.buffer(sc)), 2), exp)
    with _env("ZKGPU_MSM_GLV", "1"):
        glv = Msm.for_group(gpu, Group.G2, n, 10)
        assert not np.array_equal(_affine(gpu, glv.run(gpu, buf, gpu.buffer(sc)), 2), exp)


def test_zkey_sections_with_swapped_or_plain_words_are_rejected(gpu, tmp_path):
    from zkgpu import Server, Zkey, ZkgError, zkey

    rng = np.random.default_rng(1730)
    n_vars, n_public, domain = 300, 3, 1 << 10
    points, _, vk = zkey_file.make_key(rng, n_vars, n_public, domain, infinity=(2,))
    good = tmp_path / "good.zkey"
    zkey_file.write(good, n_vars, n_public, domain, points, vk)
    z = Zkey(good)
    for sid in (zkey.A, zkey.B1, zkey.B2, zkey.C, zkey.H):
        z.upload(gpu, sid)
    z.close()
    # A with x and y swapped in every point
    sw = points[5].reshape(-1, 16)
    swapped = np.concatenate([sw[:, 8:], sw[:, :8]], axis=1).reshape(-1)
    bad_a = tmp_path / "swapped.zkey"
    zkey_file.write(bad_a, n_vars, n_public, domain, points, vk, override={5: zkey_file.mont_bytes(swapped)})
    z = Zkey(bad_a)
    with pytest.raises(ZkgError, match="the .zkey section 5 holds 299 points off the curve \\(first at index 0\\)"):
        z.upload(gpu, zkey.A)
    z.upload(gpu, zkey.B1)  # the other sections are fine
    z.close()
    # B2 stored as plain words instead of Montgomery form
    bad_b2 = tmp_path / "plain.zkey"
    plain_b2 = np.asarray(points[7], np.uint32).astype("<u4").tobytes()
    zkey_file.write(bad_b2, n_vars, n_public, domain, points, vk, override={7: plain_b2})
    z = Zkey(bad_b2)
    with pytest.raises(ZkgError, match="the .zkey section 7 holds 299 points off the twist or outside the order-r"):
        z.upload(gpu, zkey.B2)
    server = Server(gpu, 4 << 30, max_inflight=1)
    with pytest.raises(ZkgError, match="section 7"):
        server.add_zkey(z)
    z.close()
