"""Field and point-law kernels on the GPU vs textbook Python-integer arithmetic.

Re-creates the reference's point-op harness case list (tests/curve.rs:57-87: P+Q, 2P, (P+Q)+2P
for (P,Q), (P,P), (P,-P), (inf,Q), (P,inf), (inf,inf) and random pairs) through the C ABI's
test hooks, which run the same device code as the MSM and NTT."""
import numpy as np
import pytest

import oracle as O
from zkgpu import _lib

pytestmark = pytest.mark.gpu


def _rand(rng, n, m):
    return [int.from_bytes(rng.bytes(32), "little") % m for _ in range(n)]


@pytest.mark.parametrize("field,m", [(0, O.P), (1, O.R)])
def test_field_ops(gpu, field, m):
    rng = np.random.default_rng(field)
    n = 1024
    a, b = _rand(rng, n, m), _rand(rng, n, m)
    edge = [0, 1, 2, m - 1, m - 2, (m - 1) // 2, (m + 1) // 2, (1 << 255) % m]
    a[: len(edge) ** 2] = [x for x in edge for _ in edge]
    b[: len(edge) ** 2] = [y for _ in edge for y in edge]
    A, B = O.ints_to_words(a), O.ints_to_words(b)
    out = np.zeros_like(A)
    L = _lib.load()
    for op, f in ((0, lambda x, y: x * y % m), (1, lambda x, y: (x + y) % m), (2, lambda x, y: (x - y) % m),
                  (3, lambda x, y: x * x % m)):
        _lib.check(L.zkg_test_field_op(gpu.handle, field, op, _lib.u32_ptr(A), _lib.u32_ptr(B), _lib.u32_ptr(out), n))
        assert O.words_to_ints(out) == [f(x, y) for x, y in zip(a, b)], op


def test_fq2_ops(gpu):
    rng = np.random.default_rng(2)
    n = 512
    a = list(zip(_rand(rng, n, O.P), _rand(rng, n, O.P)))
    b = list(zip(_rand(rng, n, O.P), _rand(rng, n, O.P)))
    a[0], b[0] = (0, 0), (O.P - 1, O.P - 1)
    a[1], b[1] = (O.P - 1, 1), (O.P - 1, O.P - 1)
    A = O.ints_to_words([v for p in a for v in p])
    B = O.ints_to_words([v for p in b for v in p])
    out = np.zeros_like(A)
    L = _lib.load()
    for op, f in ((0, O.py_f2_mul), (1, O.py_f2_add), (2, O.py_f2_sub), (3, lambda x, y: O.py_f2_mul(x, x))):
        _lib.check(L.zkg_test_field_op(gpu.handle, 2, op, _lib.u32_ptr(A), _lib.u32_ptr(B), _lib.u32_ptr(out), n))
        got = O.words_to_ints(out)
        assert [(got[2 * i], got[2 * i + 1]) for i in range(n)] == [f(x, y) for x, y in zip(a, b)], op


def _neg(group, p):
    if p is None:
        return None
    if group == 1:
        return (p[0], (-p[1]) % O.P)
    return (p[0], ((-p[1][0]) % O.P, (-p[1][1]) % O.P))


@pytest.mark.parametrize("group", [1, 2])
def test_point_law(gpu, group):
    rng = np.random.default_rng(57 + group)
    mul, add = (O.py_g1_mul, O.py_g1_add) if group == 1 else (O.py_g2_mul, O.py_g2_add)
    gen = O.G1 if group == 1 else O.G2
    words = O.g1_words if group == 1 else O.g2_words
    unwords = O.words_to_g1 if group == 1 else O.words_to_g2
    rnd = [mul(gen, k) for k in _rand(rng, 2 * (20 if group == 1 else 6), O.R)]
    P0 = rnd[0]
    pairs = [(P0, rnd[1]), (P0, P0), (P0, _neg(group, P0)), (None, P0), (P0, None), (None, None)]
    pairs += list(zip(rnd[2::2], rnd[3::2]))
    ps = np.concatenate([words(p) for p, _ in pairs])
    qs = np.concatenate([words(q) for _, q in pairs])
    out = np.zeros_like(ps)
    L = _lib.load()
    aw = 16 * group
    for op in range(4):
        _lib.check(L.zkg_test_point_op(gpu.handle, group, op, _lib.u32_ptr(ps), _lib.u32_ptr(qs), _lib.u32_ptr(out),
                                       len(pairs)))
        got = [unwords(out[aw * i: aw * (i + 1)]) for i in range(len(pairs))]
        if op in (0, 2):
            exp = [add(p, q) for p, q in pairs]
        elif op == 1:
            exp = [add(p, p) for p, _ in pairs]
        else:
            exp = [add(add(p, q), add(p, p)) for p, q in pairs]
        assert got == exp, op


@pytest.mark.parametrize("group", [1, 2])
def test_generator_multiples(gpu, group):
    rng = np.random.default_rng(90 + group)
    from conftest import uniform_fr_words

    n = 777
    ks = uniform_fr_words(rng, n)
    ks[:8] = 0  # 0 * G = infinity
    kb = gpu.buffer(ks)
    ob = gpu.empty_buffer(16 * group * n)
    _lib.check(_lib.load().zkg_util_generator_multiples(gpu.handle, group, kb.handle, n, ob.handle))
    assert np.array_equal(gpu.read(ob, 16 * group * n), O.multiples(group, O.generator_words(group), ks))


def test_inversion_benchmark_checks_its_inverses(gpu):
    """The Fermat-inversion mode of the field benchmark hook (29-bit limbs, x^(p-2)) verifies
    x * x^-1 = 1 for every inversion it times and fails the call otherwise."""
    import ctypes

    ms = ctypes.c_double()
    _lib.check(_lib.load().zkg_bench_field_mul(gpu.handle, 3, 256 * 8, 3, 1, ctypes.byref(ms)))
    assert ms.value > 0
