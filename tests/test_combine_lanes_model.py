"""The combines' Horner chains with one 29-bit limb per lane (zk-gpu_amd/csrc/msm_kernels.hpp,
namespace lpl: the product, the doubling and the addition for G1, and for G2 the Fq2 product
combined on the 64-bit columns, its doubling and addition), restated lane for lane: a wave of 64
lanes in four rows of 16, the DPP row moves (row_shr, row_shl, row_newbcast, zero for lanes with
no source) and the gfx950 row swaps (v_permlane16_swap, v_permlane32_swap) as the ISA defines
them. Every 32-bit lane value and 64-bit column is checked against its register width, the
product against Montgomery multiplication by 2^-261 mod p (random, lazy and worst-case limbs),
the value bounds of the doubling (X < 9p in, X < 8p, Y < 5p, ZZ, ZZZ < 2p out) and of the
addition (the same, with window sums below 2p) for a fixed point, and chains of doublings and
additions, with the addition's doubling and cancellation cases, against the textbook affine law;
for G2 the same over Fq2 inside the bound box the chain keeps, with no 64-bit column going
negative where the real part subtracts a1 b1. CPU only."""
import os
import random
import re

import pytest

import oracle as O

P = O.P
R = 1 << 261
MASK = (1 << 29) - 1
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KERNELS = open(os.path.join(ROOT, "zk-gpu_amd", "csrc", "msm_kernels.hpp")).read()
FIELD29 = open(os.path.join(ROOT, "zk-gpu_amd", "csrc", "field29.hpp")).read()
LANES = 64


def _table(src, pattern):
    m = re.search(pattern, src, re.S)
    return [int(x, 16) for x in re.findall(r"0x([0-9a-f]{8})u", m.group(1))]


PINV = _table(KERNELS, r"namespace lpl \{.*?constexpr uint32_t PINV\(int i\) \{(.*?)\}")
PL = _table(FIELD29, r"constexpr uint32_t P\(int i\) \{  // p(.*?)\}")


def _kp(c):
    return _table(FIELD29, r"K<%d>\(int i\) \{(.*?)\}" % c)


def test_constants():
    assert sum(v << (29 * i) for i, v in enumerate(PL)) == P
    assert sum(v << (29 * i) for i, v in enumerate(PINV)) == (-pow(P, -1, R)) % R
    for c in (3, 6, 7, 10):
        assert sum(v << (29 * i) for i, v in enumerate(_kp(c))) == c * P


# ---- the wave: lists of 64 lane values
def rl(l):
    return l & 15


def w32(v):
    assert all(0 <= x < 1 << 32 for x in v), hex(max(v))
    return v


def w64(v):
    assert all(0 <= x < 1 << 64 for x in v), hex(max(v))
    return v


def shr(v, k):
    return [v[l - k] if rl(l) >= k else 0 for l in range(LANES)]


def shl(v, k):
    return [v[l + k] if rl(l) + k <= 15 else 0 for l in range(LANES)]


def bc(v, k):
    return [v[(l & ~15) + k] for l in range(LANES)]


def swap16(a, b):  # odd rows of a <-> even rows of b
    a2, b2 = a[:], b[:]
    for r in (0, 2):
        for i in range(16):
            a2[(r + 1) * 16 + i], b2[r * 16 + i] = b[r * 16 + i], a[(r + 1) * 16 + i]
    return a2, b2


def swap32(a, b):  # rows 2, 3 of a <-> rows 0, 1 of b
    a2, b2 = a[:], b[:]
    for i in range(32):
        a2[32 + i], b2[i] = b[i], a[32 + i]
    return a2, b2


def rows(v):  # lpl::rows: each row's value over all four rows
    s0, s1 = swap16(v, v)
    e0, e1 = swap32(s0, s0)
    o0, o1 = swap32(s1, s1)
    return e0, o0, e1, o1


def columns(acc, x, ys):
    for k in range(9):
        xk = shr(x, k)
        acc = w64([acc[l] + xk[l] * ys[k][l] for l in range(LANES)])
    return acc


def carry2(c):
    hi = [x >> 29 for x in c]
    c1 = w64([(x & MASK) + y for x, y in zip(c, shr(hi, 1))])
    return w32([(x & MASK) + y for x, y in zip(c1, shr([x >> 29 for x in c1], 1))])


def mul(a, b, stats):
    """lpl::mul, step for step."""
    w32(a), w32(b)
    t = columns([0] * LANES, a, [bc(b, k) for k in range(9)])
    h = w64([x * y for x, y in zip(bc(a, 8), bc(b, 8))])
    tn = carry2(t)
    q = carry2(columns([0] * LANES, tn, [[PINV[k]] * LANES for k in range(9)]))
    q = [q[l] if rl(l) < 8 else (q[l] & MASK if rl(l) == 8 else 0) for l in range(LANES)]
    u = columns(t, q, [[PL[k]] * LANES for k in range(9)])
    h = w64([x + y * PL[8] for x, y in zip(h, bc(q, 8))])
    hi = [x >> 29 for x in u]
    c1 = w64([(x & MASK) + y for x, y in zip(u, shr(hi, 1))])
    h = w64([x + y for x, y in zip(h, bc(hi, 15))])
    hi2 = w32([x >> 29 for x in c1])
    uu = w32([(x & MASK) + y for x, y in zip(c1, shr(hi2, 1))])
    h = w64([x + y for x, y in zip(h, bc(hi2, 15))])
    stats["h"] = max(stats.get("h", 0), max(h))
    nz = [any(uu[r * 16 + i] for i in range(9)) for r in range(4)]
    r = shl(uu, 9)
    out = []
    for l in range(LANES):
        v = r[l] + (1 if rl(l) == 0 and nz[l >> 4] else 0)
        if rl(l) == 7:
            v = h[l] & MASK
        if rl(l) == 8:
            v = h[l] >> 29
        out.append(v)
    return w32(out)


def carry1(v):
    out = w32([(x & MASK) + y for x, y in zip(v, shr([x >> 29 for x in v], 1))])
    assert all(out[l] == 0 for l in range(LANES) if rl(l) >= 9)  # no carry leaves lane 8
    return out


def ksub(a, b, c):  # a + c p - b in borrowed limbs (b's limbs below 2^31 - 4)
    k = _kp(c)
    assert all(x < (1 << 31) - 4 for x in b)
    return carry1(w32([x + (k[rl(l)] if rl(l) < 9 else 0) - y for l, (x, y) in enumerate(zip(a, b))]))


def add(a, b):
    return w32([x + y for x, y in zip(a, b)])


def pick(*vs):
    return [vs[l >> 4][l] for l in range(LANES)]


def dbl(p, stats):
    """lpl::dbl: three rounds of four products."""
    x, y, zz, zzz = p
    u = add(y, y)
    v, x2, uy, uz = rows(mul(pick(u, x, u, u), pick(u, x, y, zzz), stats))
    m = carry1(add(add(x2, x2), x2))
    s, zz3, m2, wy = rows(mul(pick(x, v, m, v), pick(v, zz, m, uy), stats))
    x3 = ksub(m2, add(s, s), 6)
    d = ksub(s, x3, 10)
    md, zzz3, _, _ = rows(mul(pick(m, v, m, v), pick(d, uz, d, uz), stats))
    return x3, ksub(md, wy, 3), zz3, zzz3


def row_value(v, r):
    return sum(v[r * 16 + i] << (29 * i) for i in range(16))


def replicated(x):
    l9 = [(x >> (29 * i)) & MASK for i in range(8)] + [x >> 232]
    return [l9[rl(l)] if rl(l) < 9 else 0 for l in range(LANES)]


def lazy_limbs(x, rng, top=(1 << 29) + (1 << 7)):
    """Limbs 0..7 in [0, top), value at most x and close to it."""
    while True:
        l = [rng.randrange(top) for _ in range(8)]
        low = sum(v << (29 * i) for i, v in enumerate(l))
        if low <= x:
            return l + [(x - low) >> 232]


def test_product_random_and_lazy():
    rng = random.Random(261)
    stats = {}
    for it in range(120):
        a_l, b_l, vals = [0] * LANES, [0] * LANES, []
        for r in range(4):
            ka, kb = rng.choice([(10, 10), (9, 9), (10, 5), (6, 12), (12, 12), (2, 2)])
            a, b = (ka * P - 1, kb * P - 1) if it % 5 == 0 else (rng.randrange(ka * P), rng.randrange(kb * P))
            la, lb = lazy_limbs(a, rng), lazy_limbs(b, rng)
            for i in range(9):
                a_l[r * 16 + i], b_l[r * 16 + i] = la[i], lb[i]
            vals.append((row_value(a_l, r), row_value(b_l, r)))
        out = mul(a_l, b_l, stats)
        for r, (a, b) in enumerate(vals):
            v = row_value(out, r)
            assert v % P == a * b * pow(R, -1, P) % P
            assert v < a * b // R + P + (P >> 40)
            assert all(out[r * 16 + i] < (1 << 29) + (1 << 7) + 2 for i in range(8))
    assert stats["h"] < 1 << 53


@pytest.mark.parametrize("bound", [10, 12])
def test_product_worst_case_limbs(bound):
    # U = 2Y: limbs up to 2 (2^29 + 2^7 + 1), every lane at its maximum
    top = 2 * ((1 << 29) + (1 << 7) + 1)
    a_l = [0] * LANES
    for r in range(4):
        low = sum(top << (29 * i) for i in range(8))
        for i in range(8):
            a_l[r * 16 + i] = top
        a_l[r * 16 + 8] = (bound * P - 1 - low) >> 232
    a = row_value(a_l, 0)
    assert a < bound * P
    out = mul(a_l, a_l, {})
    assert row_value(out, 0) % P == a * a * pow(R, -1, P) % P and row_value(out, 0) < 2 * P


def _xyzz(pt, extra):
    Rm = R % P
    return [replicated(v + e * P) for v, e in zip((pt[0] * Rm % P, pt[1] * Rm % P, Rm, Rm), extra)]


def _affine(p):
    X, Y, ZZ, ZZZ = (row_value(v, 0) % P for v in p)
    return X * pow(ZZ, -1, P) % P, Y * pow(ZZZ, -1, P) % P


def test_doubling_chain_matches_textbook_at_the_bounds():
    rng = random.Random(2610)
    for trial in range(4):
        a = O.py_g1_mul(O.G1, rng.randrange(1, O.R))
        # the running point as an addition leaves it (X < 9p) with Y, ZZ, ZZZ at their bounds
        p = _xyzz(a, (8, 4, 1, 1))
        exp = a
        for k in range(15):
            p = dbl(p, {})
            exp = O.py_g1_add(exp, exp)
            for r in range(1, 4):  # every coordinate the same on all four rows
                assert all(row_value(c, r) == row_value(c, 0) for c in p)
            for c, bound in zip(p, (8, 5, 2, 2)):
                assert row_value(c, 0) < bound * P
            assert _affine(p) == exp


def is_zero(v):  # the kernel's test: the value (below 12p) is 0 mod p
    x = row_value(v, 0)
    assert x < 12 * P
    return x % P == 0


def add_pts(p, q, stats):
    """lpl::add: (0, sum), (1, None) for infinity, (2, None) when Q = P."""
    x1, y1, zz1, zzz1 = p
    x2, y2, zz2, zzz2 = q
    u1, u2, s1, s2 = rows(mul(pick(x1, x2, y1, y2), pick(zz2, zz1, zzz2, zzz1), stats))
    pp = ksub(u2, u1, 3)
    r = ksub(s2, s1, 3)
    if is_zero(pp):
        return (2, None) if is_zero(r) else (1, None)
    pp2, r2, zz12, zzz12 = rows(mul(pick(pp, r, zz1, zzz1), pick(pp, r, zz2, zzz2), stats))
    ppp, qq, zz3, _ = rows(mul(pick(pp, u1, zz12, pp), pp2, stats))
    x3 = ksub(r2, add(add(ppp, qq), qq), 7)
    d = ksub(qq, x3, 10)
    rd, s1ppp, zzz3, _ = rows(mul(pick(r, s1, zzz12, r), pick(d, ppp, ppp, d), stats))
    return 0, (x3, ksub(rd, s1ppp, 3), zz3, zzz3)


def test_horner_chain_with_additions_matches_textbook():
    rng = random.Random(2611)
    hits = set()
    for trial in range(6):
        a = O.py_g1_mul(O.G1, rng.randrange(1, O.R))
        p, exp = _xyzz(a, (1, 1, 1, 1)), a
        for step in range(3):
            for _ in range(rng.choice((1, 3))):
                p = dbl(p, {})
                exp = O.py_g1_add(exp, exp)
            if trial == 4 and step == 1:
                qpt = exp                          # Q = P: the caller doubles
            elif trial == 5 and step == 1:
                qpt = (exp[0], (P - exp[1]) % P)   # Q = -P: infinity
            else:
                qpt = O.py_g1_mul(O.G1, rng.randrange(1, O.R))
            # window sums as stored: below 2p, some raised by p
            q = _xyzz(qpt, (1, 1, 1, 1) if rng.random() < 0.5 else (0, 0, 0, 0))
            how, s = add_pts(p, q, {})
            exp = O.py_g1_add(exp, qpt)
            if how == 1:
                hits.add("inf")
                assert exp is None
                break
            if how == 2:
                hits.add("doubling")
                s = dbl(p, {})
            for c, bound in zip(s, (9, 5, 2, 2)):
                assert row_value(c, 0) < bound * P
                assert all(row_value(c, r) == row_value(c, 0) for r in range(1, 4))
            p = s
            assert _affine(p) == exp, (trial, step)
    assert hits == {"inf", "doubling"}



def test_addition_of_two_running_points_for_the_window_sums():
    """msm_window_sum29_lanes_kernel adds sums to sums: both operands at the running bounds."""
    rng = random.Random(2612)
    hits = set()
    for trial in range(8):
        a = O.py_g1_mul(O.G1, rng.randrange(1, O.R))
        b = a if trial == 6 else (a[0], (P - a[1]) % P) if trial == 7 else O.py_g1_mul(O.G1, rng.randrange(1, O.R))
        p, q = _xyzz(a, (8, 4, 1, 1)), _xyzz(b, (8, 4, 1, 1))
        how, s = add_pts(p, q, {})
        exp = O.py_g1_add(a, b)
        if how == 1:
            hits.add("inf")
            assert exp is None
            continue
        if how == 2:
            hits.add("doubling")
            s = dbl(p, {})
        for c, bound in zip(s, (9, 5, 2, 2)):
            assert row_value(c, 0) < bound * P
        assert _affine(s) == exp
    assert hits == {"inf", "doubling"}

# ---- G2 (msm_combine_g2_lanes_kernel): Fq2 products combined on the 64-bit columns
# p 2^261 in borrowed columns, as lpl::mcol and lpl::M16 build it from p's limbs
MC = [(1 << 62) - ((1 << 33) if j else 0) + (PL[j - 9] if 9 <= j <= 15 else 0) for j in range(16)]
M16 = PL[7] + (PL[8] << 29) - (1 << 33)


def test_g2_column_constant():
    assert sum(v << (29 * j) for j, v in enumerate(MC + [M16])) == P * R
    assert all(v >= (1 << 62) - (1 << 33) for v in MC) and M16 > 0


def reduce_cols(u, h, stats):
    """lpl::reduce: mul without its first step."""
    tn = carry2(u)
    q = carry2(columns([0] * LANES, tn, [[PINV[k]] * LANES for k in range(9)]))
    q = [q[l] if rl(l) < 8 else (q[l] & MASK if rl(l) == 8 else 0) for l in range(LANES)]
    u = columns(u, q, [[PL[k]] * LANES for k in range(9)])
    h = w64([x + y * PL[8] for x, y in zip(h, bc(q, 8))])
    hi = [x >> 29 for x in u]
    c1 = w64([(x & MASK) + y for x, y in zip(u, shr(hi, 1))])
    h = w64([x + y for x, y in zip(h, bc(hi, 15))])
    hi2 = w32([x >> 29 for x in c1])
    uu = w32([(x & MASK) + y for x, y in zip(c1, shr(hi2, 1))])
    h = w64([x + y for x, y in zip(h, bc(hi2, 15))])
    stats["h"] = max(stats.get("h", 0), max(h))
    nz = [any(uu[r * 16 + i] for i in range(9)) for r in range(4)]
    r = shl(uu, 9)
    out = [r[l] + (1 if rl(l) == 0 and nz[l >> 4] else 0) for l in range(LANES)]
    return w32([h[l] & MASK if rl(l) == 7 else h[l] >> 29 if rl(l) == 8 else out[l] for l in range(LANES)])


def mul2(x, y, stats):
    """lpl::mul2: rows a0 b0, a1 b1, a0 b1, a1 b0; rows 0, 1 -> a0 b0 + p 2^261 - a1 b1, rows 2, 3 ->
    a0 b1 + a1 b0 on the columns; one reduction; each component replicated over the rows."""
    for v in x + y:  # inputs normalised
        assert all(v[l] < (1 << 29) + (1 << 8) for l in range(LANES) if rl(l) < 8)
    a, b = pick(x[0], x[1], x[0], x[1]), pick(y[0], y[1], y[1], y[0])
    t = columns([0] * LANES, a, [bc(b, k) for k in range(9)])
    h = w64([p * q for p, q in zip(bc(a, 8), bc(b, 8))])
    t0, t1 = swap16(t, t)
    h0, h1 = swap16(h, h)
    u, hu = [], []
    for l in range(LANES):
        if l >> 5:
            u.append(t0[l] + t1[l])
            hu.append(h0[l] + h1[l])
        else:
            assert t1[l] <= t0[l] + MC[rl(l)] and h1[l] <= h0[l] + M16  # no column goes negative
            u.append(t0[l] + MC[rl(l)] - t1[l])
            hu.append(h0[l] + M16 - h1[l])
    r = reduce_cols(w64(u), w64(hu), stats)
    e0, e1 = swap32(r, r)
    return e0, e1


def add2(x, y):
    return add(x[0], y[0]), add(x[1], y[1])


def norm2(x):
    return carry1(x[0]), carry1(x[1])


def sub2(x, y, c):
    return ksub(x[0], y[0], c), ksub(x[1], y[1], c)


def dbl_g2(p, st):
    x, y, zz, zzz = p
    u = norm2(add2(y, y))
    v, x2, uy, uz = mul2(u, u, st), mul2(x, x, st), mul2(u, y, st), mul2(u, zzz, st)
    m = norm2(add2(add2(x2, x2), x2))
    s, zz3, m2, wy = mul2(x, v, st), mul2(v, zz, st), mul2(m, m, st), mul2(v, uy, st)
    x3 = sub2(m2, add2(s, s), 6)
    d = sub2(s, x3, 10)
    md, zzz3 = mul2(m, d, st), mul2(v, uz, st)
    return x3, sub2(md, wy, 4), zz3, zzz3


def is_zero2(x):
    return is_zero(x[0]) and is_zero(x[1])


def add_g2(p, q, st):
    x1, y1, zz1, zzz1 = p
    x2, y2, zz2, zzz2 = q
    u1, u2, s1, s2 = mul2(x1, zz2, st), mul2(x2, zz1, st), mul2(y1, zzz2, st), mul2(y2, zzz1, st)
    pp, r = sub2(u2, u1, 4), sub2(s2, s1, 4)
    if is_zero2(pp):
        return (2, None) if is_zero2(r) else (1, None)
    pp2, r2, zz12, zzz12 = mul2(pp, pp, st), mul2(r, r, st), mul2(zz1, zz2, st), mul2(zzz1, zzz2, st)
    ppp, qq, zz3 = mul2(pp, pp2, st), mul2(u1, pp2, st), mul2(zz12, pp2, st)
    x3 = sub2(r2, add2(ppp, add2(qq, qq)), 8)
    d = sub2(qq, x3, 12)
    rd, s1ppp, zzz3 = mul2(r, d, st), mul2(s1, ppp, st), mul2(zzz12, ppp, st)
    return 0, (x3, sub2(rd, s1ppp, 4), zz3, zzz3)


RINV = pow(R, -1, P)


def fq2_val(x):
    return row_value(x[0], 0), row_value(x[1], 0)


def test_fq2_product_at_the_chain_bounds():
    rng = random.Random(2620)
    st = {}
    # operand bounds (real, imaginary) in units of p that the chain multiplies: U U, X X, M M, M D, R D
    pairs = [((13.2, 12), (13.2, 12)), ((10.3, 9.4), (10.3, 9.4)), ((7.9, 6.4), (7.9, 6.4)),
             ((7.9, 6.4), (12.2, 11.4)), ((6.1, 5.1), (14.1, 13.1)), ((2, 2), (2, 2))]
    for it in range(90):
        ba, bb = pairs[it % len(pairs)]
        val = (lambda b: int(b * P) - 1 - rng.randrange(1 << 30)) if it % 3 == 0 else (lambda b: rng.randrange(int(b * P)))
        a = (val(ba[0]), val(ba[1]))
        b = (val(bb[0]), val(bb[1]))
        c = fq2_val(mul2((replicated(a[0]), replicated(a[1])), (replicated(b[0]), replicated(b[1])), st))
        assert c[0] % P == (a[0] * b[0] - a[1] * b[1]) * RINV % P
        assert c[1] % P == (a[0] * b[1] + a[1] * b[0]) * RINV % P
        assert c[0] < a[0] * b[0] // R + 2 * P + (P >> 40)
        assert c[1] < (a[0] * b[1] + a[1] * b[0]) // R + P + (P >> 40)
    assert st["h"] < 1 << 53


def _xyzz2(pt, extra):
    Rm = R % P
    x, y = pt
    coords = [(x[0] * Rm % P, x[1] * Rm % P), (y[0] * Rm % P, y[1] * Rm % P), (Rm, 0), (Rm, 0)]
    return [(replicated(c[0] + e * P), replicated(c[1] + e * P)) for c, e in zip(coords, extra)]


def _affine2(p):
    X, Y, ZZ, ZZZ = ((row_value(c[0], 0) % P, row_value(c[1], 0) % P) for c in p)
    return O.py_f2_mul(X, O.py_f2_inv(ZZ)), O.py_f2_mul(Y, O.py_f2_inv(ZZZ))


BOX = ((10.3, 9.4), (6.6, 6.0), (2.1, 2.0), (2.1, 2.0))  # the running point's bounds


def test_g2_horner_chain_matches_textbook_at_the_bounds():
    rng = random.Random(2621)
    hits = set()
    for trial in range(5):
        a = O.py_g2_mul(O.G2, rng.randrange(1, 1 << 64))
        p, exp = _xyzz2(a, (1, 1, 1, 1)), a
        for step in range(3):
            for _ in range(rng.choice((1, 2))):
                p = dbl_g2(p, {})
                exp = O.py_g2_add(exp, exp)
            if trial == 3 and step == 1:
                qpt = exp
            elif trial == 4 and step == 1:
                qpt = (exp[0], O.py_f2_sub((0, 0), exp[1]))
            else:
                qpt = O.py_g2_mul(O.G2, rng.randrange(1, 1 << 64))
            q = _xyzz2(qpt, (1, 1, 1, 1) if rng.random() < 0.5 else (0, 0, 0, 0))
            how, s = add_g2(p, q, {})
            exp = O.py_g2_add(exp, qpt)
            if how == 1:
                hits.add("inf")
                assert exp is None
                break
            if how == 2:
                hits.add("doubling")
                s = dbl_g2(p, {})
            for c, bound in zip(s, BOX):
                assert row_value(c[0], 0) < bound[0] * P and row_value(c[1], 0) < bound[1] * P
                assert all(row_value(c[i], r) == row_value(c[i], 0) for i in (0, 1) for r in range(1, 4))
            p = s
            assert _affine2(p) == exp, (trial, step)
    assert hits == {"inf", "doubling"}


def test_g2_addition_of_two_running_points_for_the_window_sums():
    """msm_window_sum_g2_lanes_kernel adds sums to sums: both operands at the G2 chain's bounds."""
    rng = random.Random(2622)
    hits = set()
    for trial in range(5):
        a = O.py_g2_mul(O.G2, rng.randrange(1, 1 << 64))
        b = a if trial == 3 else (a[0], O.py_f2_sub((0, 0), a[1])) if trial == 4 else O.py_g2_mul(O.G2, rng.randrange(1, 1 << 64))
        # both operands raised towards the box: X by 9p, Y by 5p, ZZ and ZZZ by p (real and imaginary)
        p = _xyzz2(a, (9, 5, 1, 1))
        q = _xyzz2(b, (9, 5, 1, 1))
        how, s = add_g2(p, q, {})
        exp = O.py_g2_add(a, b)
        if how == 1:
            hits.add("inf")
            assert exp is None
            continue
        if how == 2:
            hits.add("doubling")
            s = dbl_g2(p, {})
        for c, bound in zip(s, BOX):
            assert row_value(c[0], 0) < bound[0] * P and row_value(c[1], 0) < bound[1] * P
        assert _affine2(s) == exp
    assert hits == {"inf", "doubling"}
