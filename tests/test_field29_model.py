"""The 29-bit-limb field arithmetic of the G1 bucket accumulation (zk-gpu_amd/csrc/field29.hpp,
msm_accumulate29_kernel), restated limb for limb in Python: the constants are re-derived from p,
every intermediate is checked against the ranges the kernel relies on (limbs in [0, 2^32),
column sums below 2^64) on worst-case and random inputs, the value bounds the kernel states
(X < 9p, Y < 5p, ZZ, ZZZ < 2p) are shown to be a fixed point, and the point law it computes equals
the textbook affine law. CPU only."""
import os
import random
import re

import pytest

import oracle as O

P = O.P
MASK = (1 << 29) - 1
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = open(os.path.join(ROOT, "zk-gpu_amd", "csrc", "field29.hpp")).read()


def limbs(v):
    return [(v >> (29 * i)) & MASK for i in range(8)] + [v >> 232]


def value(l):
    return sum(x << (29 * i) for i, x in enumerate(l))


def header_table(name):
    """The nine (or twelve) constants of one `name(int i)` function in field29.hpp."""
    m = re.search(r"constexpr uint32_t " + re.escape(name) + r"\(int [ik]\) \{(.*?)\}", HDR, re.S)
    return [int(x, 16) for x in re.findall(r"0x([0-9a-f]{8})u", m.group(1))] if m else None


def header_k(c):
    m = re.search(r"K<%d>\(int i\) \{(.*?)\}" % c, HDR, re.S)
    return [int(x, 16) for x in re.findall(r"0x([0-9a-f]{8})u", m.group(1))]


def test_constants_follow_from_p():
    assert header_table("P") == limbs(P)
    assert int(re.search(r"PINV = 0x([0-9a-f]+)u", HDR).group(1), 16) == (-pow(P, -1, 1 << 29)) % (1 << 29)
    assert header_table("ONE") == limbs(pow(2, 261, P))
    assert header_table("C256") == limbs(pow(2, 256, P))
    kp0 = header_table("KP0")
    assert [0] + kp0 == [limbs(k * P)[0] for k in range(12)]  # the k = 0 entry is written as 0u
    for c in (2, 3, 4, 6, 7, 8, 10, 12):
        k = header_k(c)
        assert value(k) == c * P
        assert k[0] >= 1 << 31 and all(x >= (1 << 31) - 4 for x in k[1:8]) and all(x < 1 << 32 for x in k)


# ---- limb-exact restatement with range checks

def mul(a, b, mod=P):
    assert all(0 <= x < 1 << 30 for x in a + b), "multiplication input limb out of range"
    pinv = (-pow(mod, -1, 1 << 29)) % (1 << 29)
    pl = limbs(mod)
    acc, m, r = 0, [0] * 9, [0] * 9
    for k in range(9):
        for i in range(k + 1):
            acc += a[i] * b[k - i]
        for j in range(k):
            acc += m[j] * pl[k - j]
        assert acc < 1 << 64
        m[k] = ((acc & 0xFFFFFFFF) * pinv) & MASK
        acc += m[k] * pl[0]
        assert acc < 1 << 64 and acc & MASK == 0
        acc >>= 29
    for k in range(9, 17):
        for i in range(k - 8, 9):
            acc += a[i] * b[k - i]
        for j in range(k - 8, 9):
            acc += m[j] * pl[k - j]
        assert acc < 1 << 64
        r[k - 9] = acc & MASK
        acc >>= 29
    r[8] = acc
    assert r[8] < 1 << 32
    assert value(r) % mod == value(a) * value(b) * pow(2, -261, mod) % mod
    return r


def add(a, b):
    return [x + y for x, y in zip(a, b)]


def sub(c, a, b):
    k = header_k(c)
    r = [x + y - z for x, y, z in zip(a, k, b)]
    assert all(0 <= x < 1 << 32 for x in r), "subtraction limb out of range"
    for i in range(8):
        r[i + 1] += r[i] >> 29
        r[i] &= MASK
    assert value(r) == value(a) + c * P - value(b)
    return r


def worst(bound_p):
    """A normalised value just below bound_p * p with every low limb at its maximum."""
    v = bound_p * P - 1
    l = limbs(v)
    return [MASK] * 8 + [l[8] - 1] if l[8] > 0 else l


def test_column_sums_fit_for_worst_case_limbs():
    # all limbs at 2^30 - 1: the largest inputs the multiplication accepts
    a = [(1 << 30) - 1] * 9
    pinv = (-pow(P, -1, 1 << 29)) % (1 << 29)
    pl = limbs(P)
    acc, m = 0, [MASK] * 9  # m at its maximum too
    worst_col = 0
    for k in range(17):
        col = sum(a[i] * a[k - i] for i in range(max(0, k - 8), min(k, 8) + 1))
        col += sum(m[j] * pl[k - j] for j in range(max(0, k - 8), min(k, 8) + 1))
        worst_col = max(worst_col, col + ((1 << 64) >> 29))
    assert worst_col < 1 << 64


def test_value_bounds_are_a_fixed_point():
    """Interval check of the kernel's stated bounds (units of p): products need a * b <= 169."""
    lim = (1 << 261) / P
    assert 169 < lim < 170
    X, Y, ZZ, ZZZ, xs = 9, 5, 2, 2, 32

    def prod(a, b):
        assert a * b < lim, (a, b)
        return a * b / lim + 1  # output < a b p / 2^261 + p

    u2 = prod(xs, ZZ)
    s2 = 3  # neg: 3p - s2
    assert prod(xs, ZZZ) <= 2
    pp = u2 + 10
    r = s2 + 6
    assert X < 10 - 1 + 1e-9 and Y < 6 - 1 + 1e-9 and s2 <= 3
    ppsq, ppp, q, rr = prod(pp, pp), prod(pp, 2), prod(X, 2), prod(r, r)
    assert max(ppsq, ppp, q, rr) <= 2
    t = 2 + 2 * 2
    assert t <= 7 - 1
    x3 = rr + 7
    d = q + 10
    assert x3 <= 9 and x3 <= 10 - 1
    # Y3 = R D - Y PPP with one reduction (f29_mul2<true>): below (r d + Y ppp) p / 2^261 + 2p
    y3 = (r * d + Y * 2) / lim + 2
    assert r * d < lim and y3 <= 2.8 and y3 <= 5
    assert prod(ZZ, 2) <= 2 and prod(ZZZ, 2) <= 2
    assert prod(X, 1) <= 2  # the final conversion by 2^256 mod p < p


def test_kernel_point_law_matches_textbook_on_random_and_edge_chains():
    """The kernel's bucket loop (madd with lazy bounds), restated, against affine addition."""
    rng = random.Random(29)
    one = limbs(pow(2, 261, P))
    c256 = limbs(pow(2, 256, P))

    def load(x256):  # R = 2^256 words -> 32x (R = 2^261 form, unreduced)
        return limbs(x256 << 5)

    def to_affine(Xl, Yl, ZZl, ZZZl):
        x, y, zz, zzz = (value(mul(v, c256)) % P * pow(2, -256, P) % P for v in (Xl, Yl, ZZl, ZZZl))
        return (x * pow(zz, -1, P) % P, y * pow(zzz, -1, P) % P)

    hits = set()
    for trial in range(6):
        n = 12
        ks = [rng.randrange(1, O.R) for _ in range(n)]
        negs = [rng.random() < 0.5 for _ in range(n)]
        if trial in (1, 2):
            ks[1] = ks[0]  # the accumulator equals the next point: doubling (1) or P + (-P) (2)
            negs[1] = negs[0] if trial == 1 else not negs[0]
        if trial == 3:
            ks[5] = O.R - ks[4]  # a point and its negative inside a chain
        pts = [O.py_g1_mul(O.G1, k) for k in ks]
        exp = None
        acc = None
        for (x, y), neg, k in zip(pts, negs, ks):
            xs, ys = load(x * 2 ** 256 % P), load(y * 2 ** 256 % P)
            yy = (P - y) % P if neg else y
            exp = O.py_g1_add(exp, (x, yy)) if exp is not None else (x, yy)
            if acc is None:
                Y = mul(ys, one)
                acc = [mul(xs, one), sub(3, [0] * 9, Y) if neg else Y, one, one]
                continue
            X, Y, ZZ, ZZZ = acc
            u2 = mul(xs, ZZ)
            s2 = mul(ys, ZZZ)
            if neg:
                s2 = sub(3, [0] * 9, s2)
            pp = sub(10, u2, X)
            r = sub(6, s2, Y)
            if value(pp) % P == 0:
                assert value(pp) < 12 * P
                hits.add("doubling" if value(r) % P == 0 else "cancel")
                if value(r) % P == 0:
                    # doubling in the 32-bit field, then back (a fresh, reduced representation)
                    ax, ay = to_affine(X, Y, ZZ, ZZZ)
                    dx, dy = O.py_g1_add((ax, ay), (ax, ay))
                    acc = [mul(load(dx * 2 ** 256 % P), one), mul(load(dy * 2 ** 256 % P), one), one, one]
                else:
                    acc = None
                continue
            ppsq = mul(pp, pp)
            ppp, q, rr = mul(pp, ppsq), mul(X, ppsq), mul(r, r)
            t = add(ppp, add(q, q))
            x3 = sub(7, rr, t)
            d = sub(10, q, x3)
            y3 = mul2(r, d, Y, ppp, True, wide=True)
            acc = [x3, y3, mul(ZZ, ppsq), mul(ZZZ, ppp)]
            assert value(x3) < 9 * P and value(y3) < 2.8 * P
            assert value(acc[2]) < 2 * P and value(acc[3]) < 2 * P
        if exp is None or acc is None:
            assert acc is None and exp is None
            continue
        assert to_affine(*acc) == exp, trial
    assert hits == {"doubling", "cancel"}


@pytest.mark.parametrize("bound", [2, 9, 12])
def test_worst_case_normalised_inputs_multiply_in_range(bound):
    a = worst(bound)
    assert value(a) < bound * P
    b = worst(12 if bound <= 12 else 2)
    if bound * 12 < (1 << 261) / P:
        r = mul(a, b)
        assert value(r) < 2 * P


# ---- Fq29: the reduced variant used by the generic G1 stages (every result normalised, below 2p)

def two_p_limbs():
    return header_table("TWO_P")


def fq29_reduce(s):
    """field29.hpp fq29_reduce: limbs in [0, 2^32), value < 4p -> normalised, value < 2p."""
    assert all(0 <= x < 1 << 32 for x in s) and value(s) < 4 * P
    s = list(s)
    for i in range(8):
        s[i + 1] = (s[i + 1] + (s[i] >> 29)) & 0xFFFFFFFF
        s[i] &= MASK
    assert value(s) < 4 * P
    t, br = [0] * 9, 0
    for i, q in enumerate(two_p_limbs()):
        d = s[i] - q - br
        assert -(1 << 30) < d < 1 << 29 if i < 8 else -(1 << 30) < d < 1 << 30
        br = 1 if d < 0 else 0
        t[i] = d & MASK if i < 8 else d & 0xFFFFFFFF
    r = s if br else t
    assert all(x < 1 << 29 for x in r[:8]) and value(r) < 2 * P
    return r


def fq29_add(a, b):
    return fq29_reduce([x + y for x, y in zip(a, b)])


def fq29_sub(a, b):
    """a + 2p - b with 2p in borrowed limbs; the top limb may wrap (mod 2^32) and is restored."""
    k = header_k(2)
    s = [(x + y - z) & 0xFFFFFFFF for x, y, z in zip(a, k, b)]
    assert all(0 <= x + y - z < 1 << 32 for x, y, z in zip(a[:8], k[:8], b[:8]))
    assert k[8] - b[8] >= -4
    r = fq29_reduce_wrapped(s)
    assert value(r) % P == (value(a) - value(b)) % P
    return r


def fq29_reduce_wrapped(s):
    # as fq29_reduce, but the top limb arrives modulo 2^32 (it may be a small negative number)
    s = list(s)
    for i in range(8):
        s[i + 1] = (s[i + 1] + (s[i] >> 29)) & 0xFFFFFFFF
        s[i] &= MASK
    return fq29_reduce(s)


def test_fq29_add_sub_at_the_edges():
    edge = [0, 1, P - 1, P, P + 1, 2 * P - 1]
    rng = random.Random(7)
    vals = edge + [rng.randrange(2 * P) for _ in range(40)]
    for x in vals:
        for y in vals[:12] + vals[-3:]:
            a, b = limbs(x), limbs(y)
            r = fq29_add(a, b)
            assert value(r) % P == (x + y) % P
            d = fq29_sub(a, b)
            assert value(d) < 2 * P
    # the product of two reduced values stays below 2p
    for x in edge:
        assert value(mul(limbs(x), limbs(2 * P - 1))) < 2 * P


def test_fq29_zero_test_and_storage_round_trip():
    for x in (0, P, 1, P - 1, 2 * P - 1):
        l = limbs(x)
        assert (all(v == 0 for v in l) or l == limbs(P)) == (x % P == 0)
        words = [(x >> (32 * j)) & 0xFFFFFFFF for j in range(8)]
        # f29_unpack: limb i = bits [29 i, 29 i + 29) of the eight words
        unpacked = [(sum(w << (32 * j) for j, w in enumerate(words)) >> (29 * i)) & MASK for i in range(8)]
        unpacked.append(words[7] >> 8)
        assert unpacked == l


# ---- the NTT's twiddle product (zk-gpu_amd/csrc/ntt.hip, mul_lazy): the same 29-bit product
# modulo r, with the element's words unpacked as they are and the twiddle's R = 2^256 words shifted
# left by five bits (its R = 2^261 form, unreduced)

R_FR = O.R


def fr_mod_table():
    m = re.search(r"struct Fr29Mod \{.*?P\(int i\) \{(.*?)\}", HDR, re.S)
    return [int(x, 16) for x in re.findall(r"0x([0-9a-f]{8})u", m.group(1))]


def words(v):
    return [(v >> (32 * j)) & 0xFFFFFFFF for j in range(8)]


def from_r256_words(w):  # f29_from_r256_words: limbs of 32 x
    x = sum(v << (32 * j) for j, v in enumerate(w))
    return limbs(x << 5)


def unpack(w):  # f29_unpack: limbs of x
    return limbs(sum(v << (32 * j) for j, v in enumerate(w)))


def ntt_mul(a, t):
    """a < 4r plain, t < r in R = 2^256 form (t = w 2^256 mod r): returns a w mod r, below 2r."""
    r = mul(unpack(words(a)), from_r256_words(words(t)), R_FR)
    assert all(x < 1 << 29 for x in r[:8]) and value(r) < 2 * R_FR
    return value(r)


def test_fr_constants_follow_from_r():
    assert fr_mod_table() == limbs(R_FR)
    m = re.search(r"struct Fr29Mod \{.*?PINV = 0x([0-9a-f]+)u", HDR, re.S)
    assert int(m.group(1), 16) == (-pow(R_FR, -1, 1 << 29)) % (1 << 29)
    assert 128 * R_FR * R_FR < R_FR << 261  # v(a) v(32 t) bound the product needs


def test_ntt_product_is_the_twiddle_product_below_2r():
    rng = random.Random(2412)
    cases = [(4 * R_FR - 1, R_FR - 1), (0, R_FR - 1), (4 * R_FR - 1, 0), (2 * R_FR, 1), (R_FR, R_FR - 1),
             ((1 << 232) - 1, (1 << 253) - 1), (4 * R_FR - 1, (R_FR - 1) // 2)]
    cases += [(rng.randrange(4 * R_FR), rng.randrange(R_FR)) for _ in range(300)]
    for a, t in cases:
        w = t * pow(2, -256, R_FR) % R_FR  # the twiddle value whose R = 2^256 form is t
        assert ntt_mul(a, t) % R_FR == a * w % R_FR, (a, t)
    # a twiddle of 1 (t = 2^256 mod r) maps a to itself mod r: the product by w^0
    one = pow(2, 256, R_FR)
    for a in (0, 1, R_FR - 1, R_FR, 4 * R_FR - 1):
        assert ntt_mul(a, one) % R_FR == a % R_FR


# ---- the one-reduction sum of two products (field29.hpp: f29_mul2), used by the G2 stages'
# Fq2 product on lane pairs: (a b - c d) or (a b + c d) times 2^-261, p 2^257 added in borrowed
# 64-bit columns for the difference

def m2col_table():
    m = re.search(r"constexpr uint64_t M2COL\(int k\) \{(.*?)\}", HDR, re.S)
    return [int(x, 16) for x in re.findall(r"0x([0-9a-f]{16})ull", m.group(1))]


def m2col_wide():
    """M2COL_WIDE as the header builds it from p's limbs: p 2^261 with the same borrows."""
    pl = limbs(P)
    return [((1 << 62) if k <= 15 else 0) - ((1 << 33) if k >= 1 else 0) + (pl[k - 9] if 9 <= k <= 15 else 0)
            + ((pl[7] + (pl[8] << 29)) if k == 16 else 0) for k in range(17)]


def mul2(a, b, c, d, sub, wide=False):
    mc = m2col_wide() if wide else m2col_table()
    pl, pinv = limbs(P), (-pow(P, -1, 1 << 29)) % (1 << 29)
    acc, m, r = 0, [0] * 9, [0] * 9
    for k in range(17):
        lo, hi = max(0, k - 8), min(k, 8)
        s = 0
        for i in range(lo, hi + 1):
            acc += a[i] * b[k - i]
            s += c[i] * d[k - i]
        assert s <= mc[k] or not sub, "a column of c d above the borrowed constant"
        acc += mc[k] - s if sub else s
        if k < 9:
            acc += sum(m[j] * pl[k - j] for j in range(k))
            m[k] = ((acc & 0xFFFFFFFF) * pinv) & MASK
            acc += m[k] * pl[0]
            assert acc & MASK == 0
            acc >>= 29
        else:
            acc += sum(m[j] * pl[k - j] for j in range(k - 8, 9))
            r[k - 9] = acc & MASK
            acc >>= 29
        assert 0 <= acc < 1 << 64
    r[8] = acc
    return r


def test_mul2_sum_of_products_with_one_reduction():
    mc = m2col_table()
    assert len(mc) == 17 and value(mc) == P << 257  # a multiple of p: the reduction is unchanged mod p
    assert value(m2col_wide()) == P << 261 and all(v > 0 for v in m2col_wide())
    rng = random.Random(2625)
    rinv = pow(2, -261, P)
    edge = [0, 1, P - 1, P, 2 * P - 1, (1 << 232) - 1]
    cases = [tuple(rng.choice(edge) for _ in range(4)) for _ in range(60)]
    cases += [tuple(rng.randrange(2 * P) for _ in range(4)) for _ in range(400)]
    for a, b, c, d in cases:
        for sub in (True, False):
            r = mul2(limbs(a), limbs(b), limbs(c), limbs(d), sub)
            v = value(r)
            assert v % P == ((a * b - c * d) if sub else (a * b + c * d)) * rinv % P
            assert all(x < 1 << 29 for x in r[:8]) and v < (a * b + c * d) // (1 << 261) + P + P // 16 + 1



def test_mul2_wide_at_the_accumulation_bounds():
    """f29_mul2<true> as the bucket accumulation uses it: R D - Y PPP with R < 9p, D < 12p, Y < 5p,
    PPP < 2p (normalised limbs), below (R D + Y PPP) / 2^261 + 2p < 2.8p."""
    rng = random.Random(2626)
    rinv = pow(2, -261, P)
    tops = (9 * P - 1, 12 * P - 1, 5 * P - 1, 2 * P - 1)
    cases = [tops, (0, 0, 5 * P - 1, 2 * P - 1), (9 * P - 1, 12 * P - 1, 0, 0)]
    cases += [tuple(rng.randrange(t + 1) for t in tops) for _ in range(400)]
    for a, b, c, d in cases:
        r = mul2(limbs(a), limbs(b), limbs(c), limbs(d), True, wide=True)
        v = value(r)
        assert v % P == (a * b - c * d) * rinv % P
        assert all(x < 1 << 29 for x in r[:8]) and v < 2.8 * P
