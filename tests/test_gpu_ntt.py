"""NTT parity on the GPU (through the C ABI) — re-creates the reference's tests/ntt.rs,
tests/golden.rs (NTT half) and the NTT half of tests/properties.rs, with the CPU oracle in place
of arkworks. Bit-exact everywhere: NTT outputs are compared word for word."""
import numpy as np
import pytest

import oracle as O
from conftest import quick, uniform_fr_words

pytestmark = pytest.mark.gpu


def fwd(gpu, ntt, words):
    buf = gpu.buffer(words)
    ntt.forward(gpu, buf)
    return gpu.read(buf, words.size)


def inv(gpu, ntt, words):
    buf = gpu.buffer(words)
    ntt.inverse(gpu, buf)
    return gpu.read(buf, words.size)


def test_ntt_matches_ffjavascript(gpu):
    """tests/golden.rs:14-33: fixed ffjavascript vectors, word for word."""
    from zkgpu import Ntt

    for case in O.ntt_cases():
        ntt = Ntt(gpu, case["log_n"])
        got = (inv if case["inverse"] else fwd)(gpu, ntt, case["input"])
        assert np.array_equal(got, case["output"]), (case["log_n"], case["inverse"])


@pytest.mark.parametrize("log_n", [10, 11, 12, 13, 14, 15, 16, 18, 20])
def test_ntt_matches_oracle(gpu, log_n):
    """tests/ntt.rs:9-34: forward equals the reference transform; inverse restores the input."""
    from zkgpu import Ntt

    rng = np.random.default_rng(100 + log_n)
    x = uniform_fr_words(rng, 1 << log_n)
    ntt = Ntt(gpu, log_n)
    y = fwd(gpu, ntt, x)
    assert np.array_equal(y, O.ntt(x, log_n))
    assert np.array_equal(inv(gpu, ntt, y), x)
    assert np.array_equal(inv(gpu, ntt, x), O.ntt(x, log_n, inverse=True))


@pytest.mark.parametrize("log_n", [22, 24])
def test_ntt_large_sizes(gpu, log_n):
    """Config sizes: forward vs the oracle at 2^22; at 2^24 Horner spot checks (independent of
    any FFT structure) plus the round trip."""
    from zkgpu import Ntt

    if quick() and log_n == 24:
        pytest.skip("ZK_GPU_QUICK")
    rng = np.random.default_rng(200 + log_n)
    n = 1 << log_n
    x = uniform_fr_words(rng, n)
    ntt = Ntt(gpu, log_n)
    buf = gpu.buffer(x)
    ntt.forward(gpu, buf)
    y = gpu.read(buf, x.size)
    if log_n <= 22:
        assert np.array_equal(y, O.ntt(x, log_n))
    w = O.py_root_of_unity(log_n)
    for j in (0, 1, n // 2 + 3, n - 1):
        assert O.poly_eval(x, pow(w, j, O.R)) == O.words_to_int(y[8 * j: 8 * j + 8]), j
    ntt.inverse(gpu, buf)
    assert np.array_equal(gpu.read(buf, x.size), x)


@pytest.mark.parametrize("log_n", [25, 26, 27])
def test_ntt_largest_sizes(gpu, log_n):
    """The sizes up to the reference's limit (src/ntt.rs:27: log_n in 10..=27): Horner evaluation of
    the input at omega^j for sampled j (independent of any FFT structure) equals output j, for the
    forward and the inverse transform (inverse: (1/n) sum x_i omega^-ij), and the round trip
    restores the input. Elements are uniform below 2^252 (< r), drawn as words directly."""
    from zkgpu import Ntt

    if quick():
        pytest.skip("ZK_GPU_QUICK")
    rng = np.random.default_rng(300 + log_n)
    n = 1 << log_n
    x = rng.integers(0, 1 << 32, size=(n, 8), dtype=np.uint32)
    x[:, 7] &= 0x0FFFFFFF
    x = x.reshape(-1)
    ntt = Ntt(gpu, log_n)
    buf = gpu.buffer(x)
    ntt.forward(gpu, buf)
    y = gpu.read(buf, x.size)
    w = O.py_root_of_unity(log_n)
    js = (1, int(rng.integers(2, n - 1)), n - 1)
    for j in js:
        assert O.poly_eval(x, pow(w, j, O.R)) == O.words_to_int(y[8 * j: 8 * j + 8]), ("forward", j)
    ntt.inverse(gpu, buf)
    assert np.array_equal(gpu.read(buf, x.size), x)
    ntt.inverse(gpu, buf)
    z = gpu.read(buf, x.size)
    n_inv = pow(n, O.R - 2, O.R)
    winv = pow(w, O.R - 2, O.R)
    for j in js[:2]:
        assert O.poly_eval(x, pow(winv, j, O.R)) * n_inv % O.R == O.words_to_int(z[8 * j: 8 * j + 8]), ("inverse", j)


def test_ntt_of_a_delta_is_a_power_sequence(gpu):
    """tests/properties.rs:25-46."""
    from zkgpu import Ntt

    for log_n in (10, 11, 13):
        n = 1 << log_n
        ntt = Ntt(gpu, log_n)
        w = O.py_root_of_unity(log_n)
        for k in (0, 1, 2, 511, 512, n - 1):
            delta = np.zeros(8 * n, np.uint32)
            delta[8 * k] = 1
            got = O.words_to_ints(fwd(gpu, ntt, delta))
            assert got == [pow(w, j * k, O.R) for j in range(n)], (log_n, k)
        c = 123456789
        const = np.tile(O.int_to_words(c), n)
        got = O.words_to_ints(fwd(gpu, ntt, const))
        assert got == [c * n % O.R] + [0] * (n - 1)


def test_ntt_is_linear(gpu):
    """tests/properties.rs:48-62."""
    from zkgpu import Ntt

    rng = np.random.default_rng(1)
    for log_n in (10, 12, 14):
        n = 1 << log_n
        ntt = Ntt(gpu, log_n)
        a, b = uniform_fr_words(rng, n), uniform_fr_words(rng, n)
        c = O.words_to_int(uniform_fr_words(rng, 1))
        A, B = O.words_to_ints(a), O.words_to_ints(b)
        comb = O.ints_to_words([(x * c + y) % O.R for x, y in zip(A, B)])
        fa, fb = O.words_to_ints(fwd(gpu, ntt, a)), O.words_to_ints(fwd(gpu, ntt, b))
        assert O.words_to_ints(fwd(gpu, ntt, comb)) == [(x * c + y) % O.R for x, y in zip(fa, fb)]


def test_ntt_convolution_theorem(gpu):
    """tests/properties.rs:64-82: inverse(F(a) * F(b)) = cyclic convolution, computed the slow way."""
    from zkgpu import Ntt

    rng = np.random.default_rng(2)
    log_n = 10
    n = 1 << log_n
    ntt = Ntt(gpu, log_n)
    a, b = O.words_to_ints(uniform_fr_words(rng, n)), O.words_to_ints(uniform_fr_words(rng, n))
    expected = [0] * n
    for i, x in enumerate(a):
        if x:
            for j, y in enumerate(b):
                expected[(i + j) % n] += x * y
    expected = [v % O.R for v in expected]
    fa = O.words_to_ints(fwd(gpu, ntt, O.ints_to_words(a)))
    fb = O.words_to_ints(fwd(gpu, ntt, O.ints_to_words(b)))
    prod = O.ints_to_words([x * y % O.R for x, y in zip(fa, fb)])
    assert O.words_to_ints(inv(gpu, ntt, prod)) == expected


def test_ntt_round_trips_both_ways(gpu):
    """tests/properties.rs:84-101."""
    from zkgpu import Ntt

    rng = np.random.default_rng(4)
    for log_n in (10, 12, 15, 17):
        ntt = Ntt(gpu, log_n)
        x = uniform_fr_words(rng, 1 << log_n)
        assert np.array_equal(inv(gpu, ntt, fwd(gpu, ntt, x)), x)
        assert np.array_equal(fwd(gpu, ntt, inv(gpu, ntt, x)), x)


def test_ntt_extreme_values(gpu):
    """tests/properties.rs:103-121: 0, 1 and r-1 everywhere and mixed."""
    from zkgpu import Ntt

    log_n = 12
    n = 1 << log_n
    ntt = Ntt(gpu, log_n)
    rng = np.random.default_rng(6)
    for vals in ([0] * n, [1] * n, [O.R - 1] * n, list(rng.choice([0, 1, O.R - 1], size=n))):
        x = O.ints_to_words([int(v) for v in vals])
        assert np.array_equal(fwd(gpu, ntt, x), O.ntt(x, log_n))
        assert np.array_equal(inv(gpu, ntt, x), O.ntt(x, log_n, inverse=True))


def test_ntt_plan_reuse(gpu):
    """tests/properties.rs:123-138: one plan, many inputs and both directions, no state leaks."""
    from zkgpu import Ntt

    rng = np.random.default_rng(8)
    log_n = 16
    ntt = Ntt(gpu, log_n)
    for _ in range(3):
        x = uniform_fr_words(rng, 1 << log_n)
        assert np.array_equal(fwd(gpu, ntt, x), O.ntt(x, log_n))
        assert np.array_equal(inv(gpu, ntt, x), O.ntt(x, log_n, inverse=True))


def test_ntt_buffer_larger_than_needed(gpu):
    """tests/invalid_inputs.rs:52-59: only the first 2^log_n elements are transformed."""
    from zkgpu import Ntt

    rng = np.random.default_rng(10)
    x = uniform_fr_words(rng, 1 << 13)
    ntt = Ntt(gpu, 12)
    buf = gpu.buffer(x)
    ntt.forward(gpu, buf)
    got = gpu.read(buf, x.size)
    assert np.array_equal(got[: 8 << 12], O.ntt(x[: 8 << 12], 12))
    assert np.array_equal(got[8 << 12:], x[8 << 12:])
