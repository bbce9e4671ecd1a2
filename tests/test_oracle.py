"""Pins the CPU oracle (oracle/) before anything is checked against it. No GPU needed.

Every function of the oracle is checked against something independent of it:
  * the reference's golden vectors (ffjavascript 0.3.1): tests/golden/{ntt,msm}.json;
  * textbook arithmetic on Python integers (modular ops, chord-and-tangent point law, O(n^2) DFT);
  * closed forms: sum s_i (k_i G) = (sum k_i s_i) G, NTT of a delta / a constant;
  * group facts: G2 generator on the twist y^2 = x^3 + 3/(9+u), r * G2 = O.
Mutant checks plant a plausible mistake in a copy of the oracle's mathematics and show that the
golden vectors reject it.
"""
import numpy as np
import pytest

import oracle as O
from conftest import uniform_fr_words


def test_golden_ntt_cases():
    cases = O.ntt_cases()
    assert [(c["log_n"], c["inverse"]) for c in cases] == [(10, False), (10, True), (12, False)]
    for c in cases:
        assert np.array_equal(O.ntt(c["input"], c["log_n"], c["inverse"]), c["output"])


def test_golden_msm_cases():
    cases = O.msm_cases()
    assert [c["name"] for c in cases] == ["random", "edge", "small", "small_edge"]
    for c in cases:
        assert np.array_equal(O.msm(1, c["bases"], c["scalars"]), c["result"]), c["name"]
        assert np.array_equal(O.msm(1, c["bases"], c["scalars"], threads=1), c["result"]), c["name"]


def test_golden_msm_naive_small():
    for c in O.msm_cases():
        if c["scalars"].size // 8 <= 100:
            assert np.array_equal(O.msm_naive(1, c["bases"], c["scalars"]), c["result"])


@pytest.mark.parametrize("which,m", [(0, O.P), (1, O.R)])
def test_field_ops_vs_python(which, m):
    rng = np.random.default_rng(5 + which)
    vals = [0, 1, 2, m - 1, m - 2, (m - 1) // 2] + [int.from_bytes(rng.bytes(32), "little") % m for _ in range(60)]
    for a in vals[:20]:
        for b in vals[-20:]:
            A, B = O.int_to_words(a), O.int_to_words(b)
            assert O.words_to_int(O.field_op("mul", which, A, B)) == a * b % m
            assert O.words_to_int(O.field_op("add", which, A, B)) == (a + b) % m
            assert O.words_to_int(O.field_op("sub", which, A, B)) == (a - b) % m
        if a:
            assert O.words_to_int(O.field_op("inv", which, O.int_to_words(a))) == pow(a, m - 2, m)


def test_fq2_vs_python():
    rng = np.random.default_rng(9)
    for _ in range(30):
        x = tuple(int.from_bytes(rng.bytes(32), "little") % O.P for _ in range(2))
        y = tuple(int.from_bytes(rng.bytes(32), "little") % O.P for _ in range(2))
        X = np.concatenate([O.int_to_words(v) for v in x])
        Y = np.concatenate([O.int_to_words(v) for v in y])
        got = O.words_to_ints(O.fq2_mul(X, Y))
        assert tuple(got) == O.py_f2_mul(x, y)
        inv = O.words_to_ints(O.fq2_inv(X))
        assert O.py_f2_mul(tuple(inv), x) == (1, 0)


def test_root_of_unity():
    # 2-adicity of r - 1 is 28; w^(2^k) == 1 and w^(2^(k-1)) == r - 1 (primitive)
    assert ((O.R - 1) >> 28) & 1 == 1 and (O.R - 1) % (1 << 28) == 0
    for k in (1, 10, 12, 20, 28):
        w = O.words_to_int(O.root_of_unity_words(k))
        assert w == O.py_root_of_unity(k)
        assert pow(w, 1 << k, O.R) == 1 and pow(w, 1 << (k - 1), O.R) == O.R - 1


@pytest.mark.parametrize("log_n", [1, 2, 3, 5, 6])
@pytest.mark.parametrize("inverse", [False, True])
def test_ntt_vs_schoolbook_dft(log_n, inverse):
    rng = np.random.default_rng(log_n)
    x = uniform_fr_words(rng, 1 << log_n)
    exp = O.py_dft(O.words_to_ints(x), log_n, inverse)
    assert O.words_to_ints(O.ntt(x, log_n, inverse, threads=1)) == exp


def test_ntt_horner_and_roundtrip():
    rng = np.random.default_rng(3)
    log_n = 12
    x = uniform_fr_words(rng, 1 << log_n)
    y = O.ntt(x, log_n)
    w = O.py_root_of_unity(log_n)
    for j in (0, 1, 7, 2048, 4095):
        assert O.poly_eval(x, pow(w, j, O.R)) == O.words_to_int(y[8 * j: 8 * j + 8])
    assert np.array_equal(O.ntt(y, log_n, inverse=True), x)


def test_poly_eval_threaded_vs_textbook():
    """The threaded evaluation (parts scaled by x^start) equals Horner on one thread and the
    textbook sum on Python integers, for uneven splits and more threads than coefficients."""
    rng = np.random.default_rng(4)
    for n in (1, 2, 3, 7, 100, 1001):
        c = uniform_fr_words(rng, n)
        ints = O.words_to_ints(c)
        for x in (0, 1, 2, O.R - 1, int(rng.integers(1, 1 << 62)) ** 4 % O.R):
            exp = sum(ci * pow(x, i, O.R) for i, ci in enumerate(ints)) % O.R
            assert O.poly_eval(c, x, threads=1) == exp, (n, x)
            for t in (2, 3, 16):
                assert O.poly_eval(c, x, threads=t) == exp, (n, x, t)


def test_g1_point_ops_vs_textbook():
    rng = np.random.default_rng(11)
    ks = [int.from_bytes(rng.bytes(32), "little") % O.R for _ in range(6)]
    pts = [O.py_g1_mul(O.G1, k) for k in ks]
    for p in pts:
        assert O.py_g1_on_curve(p)
    g = O.g1_words
    for p, q in zip(pts, pts[1:]):
        assert O.words_to_g1(O.point_add(1, g(p), g(q))) == O.py_g1_add(p, q)
    p = pts[0]
    assert O.words_to_g1(O.point_add(1, g(p), g(p))) == O.py_g1_add(p, p)
    assert O.words_to_g1(O.point_add(1, g(p), g((p[0], O.P - p[1])))) is None
    assert O.words_to_g1(O.point_add(1, g(None), g(p))) == p
    for k in ks[:3]:
        assert O.words_to_g1(O.point_mul(1, g(O.G1), k)) == O.py_g1_mul(O.G1, k)


def test_g2_generator_and_ops():
    assert O.py_g2_on_curve(O.G2)
    gw = O.generator_words(2)
    assert not np.any(O.point_mul(2, gw, O.R))  # r * G2 = O
    assert O.words_to_g2(O.point_mul(2, gw, O.R - 1)) == (O.G2[0], ((-O.G2[1][0]) % O.P, (-O.G2[1][1]) % O.P))
    rng = np.random.default_rng(12)
    for _ in range(3):
        k1 = int.from_bytes(rng.bytes(32), "little") % O.R
        k2 = int.from_bytes(rng.bytes(32), "little") % O.R
        p, q = O.py_g2_mul(O.G2, k1), O.py_g2_mul(O.G2, k2)
        assert O.py_g2_on_curve(p)
        assert O.words_to_g2(O.point_mul(2, gw, k1)) == p
        assert O.words_to_g2(O.point_add(2, O.g2_words(p), O.g2_words(q))) == O.py_g2_add(p, q)


@pytest.mark.parametrize("group", [1, 2])
def test_multiples_and_closed_form(group):
    rng = np.random.default_rng(20 + group)
    n = 300
    ks = uniform_fr_words(rng, n)
    sc = uniform_fr_words(rng, n)
    bases = O.multiples(group, O.generator_words(group), ks)
    # every base equals k_i * G computed by double-and-add
    for i in (0, 1, n - 1):
        assert np.array_equal(bases[16 * group * i: 16 * group * (i + 1)],
                              O.point_mul(group, O.generator_words(group), O.words_to_int(ks[8 * i: 8 * i + 8])))
    cf = O.closed_form_msm(group, O.words_to_ints(ks), O.words_to_ints(sc))
    assert np.array_equal(O.msm(group, bases, sc), cf)
    assert np.array_equal(O.msm(group, bases[: 16 * group * 40], sc[: 8 * 40]), O.msm_naive(group, bases[: 16 * group * 40], sc[: 8 * 40]))
    t = O.closed_form_scalar(ks, sc)
    assert t == sum(a * b for a, b in zip(O.words_to_ints(ks), O.words_to_ints(sc))) % O.R
    assert O.closed_form_scalar_np(ks, sc) == t


def test_closed_form_scalar_threads_and_edges():
    """The threaded C dot product against exact integer arithmetic: r - 1 squared terms, zeros,
    a length that does not split evenly over the threads, and one thread vs many."""
    rng = np.random.default_rng(24)
    n = 10007
    ks = uniform_fr_words(rng, n)
    sc = uniform_fr_words(rng, n)
    ks[:8] = O.int_to_words(O.R - 1)
    sc[:8] = O.int_to_words(O.R - 1)
    sc[8:16] = 0
    want = O.closed_form_scalar_np(ks, sc)
    assert O.closed_form_scalar(ks, sc, threads=1) == want
    assert O.closed_form_scalar(ks, sc, threads=7) == want
    assert O.closed_form_scalar(ks[:8], sc[:8]) == (O.R - 1) ** 2 % O.R
    assert O.closed_form_scalar(ks[:0], sc[:0]) == 0


def test_msm_edge_cases_vs_naive():
    # zero / one / r-1 scalars, repeated bases (forces doublings), infinity bases, cancellation
    rng = np.random.default_rng(31)
    n = 64
    ks = uniform_fr_words(rng, n)
    bases = O.multiples(1, O.generator_words(1), ks).reshape(n, 16)
    bases[5] = bases[4]
    bases[7] = 0
    bases[9] = bases[8]
    bases[9, 8:] = O.int_to_words(O.P - O.words_to_int(bases[8, 8:]))  # -P8
    sc = uniform_fr_words(rng, n).reshape(n, 8)
    sc[0] = 0
    sc[1] = O.int_to_words(1)
    sc[2] = O.int_to_words(O.R - 1)
    sc[9] = sc[8]  # P8 and -P8 with the same scalar cancel
    assert np.array_equal(O.msm(1, bases, sc), O.msm_naive(1, bases, sc))
    sc[:] = 0
    assert not np.any(O.msm(1, bases, sc))


# ---------------------------------------------------------------- the checks catch planted mistakes

def _py_dft_wrong_root(values, log_n):
    # mutant: inverse root used for the forward transform
    n = 1 << log_n
    w = pow(O.py_root_of_unity(log_n), O.R - 2, O.R)
    return [sum(v * pow(w, i * j, O.R) for i, v in enumerate(values)) % O.R for j in range(n)]


def test_golden_rejects_wrong_root():
    c = O.ntt_cases()[0]
    assert c["log_n"] == 10
    first = O.words_to_ints(c["input"])
    got = _py_dft_wrong_root(first[:1024], 10)[:4]
    assert got != O.words_to_ints(c["output"])[:4]


def test_golden_rejects_dropped_digit_carry():
    # mutant: signed-digit decomposition that drops the carry, evaluated as a scalar value
    c = O.msm_cases()[3]  # small_edge, n = 5
    sc = O.words_to_ints(c["scalars"])

    def value_with_dropped_carry(s, cw=9):
        total, pos = 0, 0
        while pos < 255:
            v = (s >> pos) & ((1 << cw) - 1)
            d = v - (1 << cw) if v >= 1 << (cw - 1) else v  # carry never added
            total += d << pos
            pos += cw
        return total % O.R

    mutated = [value_with_dropped_carry(s) for s in sc]
    assert mutated != sc  # the mutant changes at least one scalar of the case
    bases = c["bases"].reshape(-1, 16)
    got = O.msm_naive(1, bases, O.ints_to_words(mutated))
    assert not np.array_equal(got, c["result"])


def test_groth16_h_is_the_quotient_for_satisfied_systems():
    """With c = a * b on the domain, A B - C vanishes there, so H is the exact quotient by
    x^n - 1: H(t) (t^n - 1) = A(t) B(t) - C(t) at random points, and deg H <= n - 2. Any coset
    gives the same H then; a different coset changes H for an unsatisfied system."""
    rng = np.random.default_rng(77)
    log_n = 10
    n = 1 << log_n
    a = uniform_fr_words(rng, n)
    b = uniform_fr_words(rng, n)
    c = O.ints_to_words([x * y % O.R for x, y in zip(O.words_to_ints(a), O.words_to_ints(b))])
    h = O.groth16_h(a, b, c, log_n)
    assert O.words_to_int(h[8 * (n - 1):]) == 0
    ca, cb, cc = (O.ntt(w, log_n, inverse=True) for w in (a, b, c))
    for t in (3, 123456789, O.R - 2):
        lhs = O.poly_eval(h, t) * (pow(t, n, O.R) - 1) % O.R
        rhs = (O.poly_eval(ca, t) * O.poly_eval(cb, t) - O.poly_eval(cc, t)) % O.R
        assert lhs == rhs, t
    assert np.array_equal(O.groth16_h(a, b, c, log_n, g=7), h)
    c2 = c.copy()
    c2[:8] = O.int_to_words((O.words_to_int(c[:8]) + 1) % O.R)  # one unsatisfied constraint
    assert not np.array_equal(O.groth16_h(a, b, c2, log_n, g=7), O.groth16_h(a, b, c2, log_n))


def test_groth16_h_odd_coset_by_horner_and_against_the_coefficient_form():
    """The odd-coset values are A B - C at t_j = w_2n w^j (w_2n^2 = w; Horner on the coefficients,
    independent of the transform), and for a satisfied system t_j^n = -1 makes them -2 H(t_j)
    with H the coefficient-form quotient."""
    rng = np.random.default_rng(78)
    log_n = 10
    n = 1 << log_n
    a = uniform_fr_words(rng, n)
    b = uniform_fr_words(rng, n)
    c = uniform_fr_words(rng, n)
    w2n = O.py_root_of_unity(log_n + 1)
    w = O.py_root_of_unity(log_n)
    assert pow(w2n, 2, O.R) == w and pow(w2n, n, O.R) == O.R - 1
    odd = O.groth16_h_odd_coset(a, b, c, log_n)
    ca, cb, cc = (O.ntt(x, log_n, inverse=True) for x in (a, b, c))
    for j in (0, 1, 5, n - 1):
        t = w2n * pow(w, j, O.R) % O.R
        exp = (O.poly_eval(ca, t) * O.poly_eval(cb, t) - O.poly_eval(cc, t)) % O.R
        assert O.words_to_int(odd[8 * j: 8 * j + 8]) == exp, j
    # satisfied system: c = a b on the domain
    c = O.ints_to_words([x * y % O.R for x, y in zip(O.words_to_ints(a), O.words_to_ints(b))])
    odd = O.groth16_h_odd_coset(a, b, c, log_n)
    h = O.groth16_h(a, b, c, log_n)
    for j in (0, 7, n - 2):
        t = w2n * pow(w, j, O.R) % O.R
        assert O.words_to_int(odd[8 * j: 8 * j + 8]) == (-2 * O.poly_eval(h, t)) % O.R, j
