"""Groth16 quotient polynomial on the device: equal to the oracle word for word (satisfied and
unsatisfied systems), and the exact quotient for satisfied systems at larger sizes."""
import numpy as np
import pytest

import oracle as O
from conftest import quick, uniform_fr_words

pytestmark = pytest.mark.gpu


def _satisfied(rng, n):
    a = uniform_fr_words(rng, n)
    b = uniform_fr_words(rng, n)
    c = O.ints_to_words([x * y % O.R for x, y in zip(O.words_to_ints(a), O.words_to_ints(b))])
    return a, b, c


@pytest.mark.parametrize("log_n", [10, 11, 13])
def test_h_matches_oracle(gpu, log_n):
    from zkgpu import Groth16H

    rng = np.random.default_rng(1300 + log_n)
    n = 1 << log_n
    plan = Groth16H(gpu, log_n)
    for satisfied in (True, False):
        if satisfied:
            a, b, c = _satisfied(rng, n)
        else:
            a, b, c = (uniform_fr_words(rng, n) for _ in range(3))
        bufs = [gpu.buffer(x) for x in (a, b, c)]
        plan.run(gpu, *bufs)
        got = gpu.read(bufs[0], 8 * n)
        assert np.array_equal(got, O.groth16_h(a, b, c, log_n)), (log_n, satisfied)
        if satisfied:
            assert not got[8 * (n - 1):].any()


@pytest.mark.parametrize("log_n", [16, 20])
def test_h_is_the_quotient_at_scale(gpu, log_n):
    """H(t) (t^n - 1) = A(t) B(t) - C(t) at random t (A, B, C coefficients from the oracle)."""
    from zkgpu import Groth16H

    if quick() and log_n > 16:
        pytest.skip("ZK_GPU_QUICK")
    rng = np.random.default_rng(1310 + log_n)
    n = 1 << log_n
    a = uniform_fr_words(rng, n)
    b = uniform_fr_words(rng, n)
    ai, bi = O.words_to_ints(a), O.words_to_ints(b)
    c = O.ints_to_words([x * y % O.R for x, y in zip(ai, bi)])
    bufs = [gpu.buffer(x) for x in (a, b, c)]
    Groth16H(gpu, log_n).run(gpu, *bufs)
    h = gpu.read(bufs[0], 8 * n)
    assert not h[8 * (n - 1):].any()
    ca, cb, cc = (O.ntt(w, log_n, inverse=True) for w in (a, b, c))
    for t in (7, 987654321):
        lhs = O.poly_eval(h, t) * (pow(t, n, O.R) - 1) % O.R
        rhs = (O.poly_eval(ca, t) * O.poly_eval(cb, t) - O.poly_eval(cc, t)) % O.R
        assert lhs == rhs, (log_n, t)


def test_h_rejects_aliased_and_short_buffers(gpu):
    from zkgpu import Groth16H, ZkgError

    plan = Groth16H(gpu, 10)
    x = gpu.buffer(np.zeros(8 << 10, np.uint32))
    y = gpu.buffer(np.zeros(8 << 10, np.uint32))
    short = gpu.buffer(np.zeros(8 << 9, np.uint32))
    with pytest.raises(ZkgError, match="distinct"):
        plan.run(gpu, x, x, y)
    with pytest.raises(ZkgError, match="data buffer holds"):
        plan.run(gpu, x, y, short)


@pytest.mark.parametrize("log_n", [10, 12])
def test_h_odd_coset_matches_oracle(gpu, log_n):
    """The snarkjs form: A B - C on the odd coset, word for word, for a random and a satisfied
    system (where it is -2 H(t) of the coefficient form, checked on the oracle side)."""
    from zkgpu import Groth16H, groth16

    rng = np.random.default_rng(1330 + log_n)
    n = 1 << log_n
    plan = Groth16H(gpu, log_n, groth16.ODD_COSET)
    for satisfied in (False, True):
        a, b, c = _satisfied(rng, n) if satisfied else (uniform_fr_words(rng, n) for _ in range(3))
        bufs = [gpu.buffer(x) for x in (a, b, c)]
        plan.run(gpu, *bufs)
        assert np.array_equal(gpu.read(bufs[0], 8 * n), O.groth16_h_odd_coset(a, b, c, log_n)), (log_n, satisfied)
