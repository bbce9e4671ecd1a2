"""MSM parity on the GPU through the C ABI — re-creates the reference's tests/msm.rs, tests/g2.rs,
tests/golden.rs (MSM half) and the MSM half of tests/properties.rs, with the CPU oracle (and the
closed form sum s_i (k_i G) = (sum k_i s_i) G) in place of arkworks. Results are compared as
affine points after normalisation, word for word."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import quick, uniform_fr_words, witness_scalars

pytestmark = pytest.mark.gpu


def affine(gpu, res_buf, group):
    import zkgpu
    from zkgpu import host

    if group == 1:
        a = zkgpu.xyzz_to_affine(gpu.read(res_buf, 32))
        if a is None:
            return np.zeros(16, np.uint32)
        return np.array(host.words32(a[0]) + host.words32(a[1]), np.uint32)
    a = zkgpu.xyzz_to_affine_g2(gpu.read(res_buf, 64))
    if a is None:
        return np.zeros(32, np.uint32)
    (xa, xb), (ya, yb) = a
    return np.array(sum((host.words32(v) for v in (xa, xb, ya, yb)), []), np.uint32)


def gpu_msm(gpu, group, bases_plain, scalars, c=None, item_points=None, mont256=False):
    from zkgpu import Group, Msm

    n = scalars.size // 8
    up = Msm.upload_bases_montgomery_256 if mont256 else Msm.upload_bases
    bb = up(gpu, bases_plain)
    sb = gpu.buffer(scalars)
    m = Msm.for_group(gpu, Group(group), n, c if c is not None else Msm.default_window(n), item_points)
    return affine(gpu, m.run(gpu, bb, sb), group)


def device_bases(gpu, group, ks):
    """k_i * G on the device (plain affine), converted in place to the library form."""
    import zkgpu
    from zkgpu import Msm, _lib

    n = ks.size // 8
    kb = gpu.buffer(ks)
    bb = gpu.empty_buffer(16 * group * n)
    _lib.check(_lib.load().zkg_util_generator_multiples(gpu.handle, group, kb.handle, n, bb.handle))
    Msm.convert_bases(bb, 16 * group * n)
    return bb


def test_msm_matches_ffjavascript(gpu):
    """tests/golden.rs:35-56: defaults, a small window with tiny items, a large window."""
    from zkgpu import Msm

    for case in O.msm_cases():
        n = case["scalars"].size // 8
        bases = Msm.upload_bases(gpu, case["bases"])
        sc = gpu.buffer(case["scalars"])
        for plan in (None, (9, 7), (12, 1024)):
            m = Msm.new(gpu, n, Msm.default_window(n)) if plan is None else Msm.with_item_points(gpu, n, *plan)
            got = affine(gpu, m.run(gpu, bases, sc), 1)
            assert np.array_equal(got, case["result"]), (case["name"], plan)


@pytest.mark.parametrize("n,c", [(1000, 9), (4096, 10), (20000, 12), (100000, 14)])
@pytest.mark.parametrize("items", [1024, 7])
def test_msm_random_vs_oracle(gpu, n, c, items):
    """tests/msm.rs:9-30."""
    rng = np.random.default_rng(n + c + items)
    ks = uniform_fr_words(rng, n)
    bases = O.multiples(1, O.generator_words(1), ks)
    sc = uniform_fr_words(rng, n)
    exp = O.msm(1, bases, sc)
    assert np.array_equal(gpu_msm(gpu, 1, bases, sc, c, items), exp)


def _edge_bases(rng, n):
    ks = uniform_fr_words(rng, n)
    return O.multiples(1, O.generator_words(1), ks).reshape(n, 16)


@pytest.mark.parametrize("kind", ["zero", "one", "r_minus_1", "same_base", "same_base_same_scalar", "tiny",
                                  "infinity_bases", "cancel"])
def test_msm_edge_cases(gpu, kind):
    """tests/msm.rs:32-60 at n = 3000, c = 10."""
    n, c = 3000, 10
    rng = np.random.default_rng(sum(map(ord, kind)))
    bases = _edge_bases(rng, n)
    sc = uniform_fr_words(rng, n).reshape(n, 8)
    if kind == "zero":
        sc[:] = 0
    elif kind == "one":
        sc[:] = 0
        sc[:, 0] = 1
    elif kind == "r_minus_1":
        sc[:] = O.int_to_words(O.R - 1)
    elif kind == "same_base":
        bases[:] = bases[0]
    elif kind == "same_base_same_scalar":
        bases[:] = bases[0]
        sc[:] = sc[0]
    elif kind == "tiny":
        sc[:] = 0
        sc[:, 0] = rng.integers(0, 4, size=n)
    elif kind == "infinity_bases":
        bases[::3] = 0
    elif kind == "cancel":
        half = n // 2
        bases[half:] = bases[:half]
        for i in range(half):
            bases[half + i, 8:] = O.int_to_words((O.P - O.words_to_int(bases[i, 8:])) % O.P)
        sc[half:] = sc[:half]
    exp = O.msm(1, bases, sc)
    got = gpu_msm(gpu, 1, bases.reshape(-1), sc.reshape(-1), c)
    assert np.array_equal(got, exp), kind
    if kind == "cancel":
        assert not np.any(got)


def test_msm_defaults_at_2_17(gpu):
    """tests/msm.rs:62-67 (library defaults at n = 2^17), via the closed form."""
    n = 1 << 17
    rng = np.random.default_rng(17)
    ks, sc = uniform_fr_words(rng, n), uniform_fr_words(rng, n)
    from zkgpu import Msm

    bb = device_bases(gpu, 1, ks)
    m = Msm.new(gpu, n, Msm.default_window(n))
    got = affine(gpu, m.run(gpu, bb, gpu.buffer(sc)), 1)
    assert np.array_equal(got, O.point_mul(1, O.generator_words(1), O.closed_form_scalar(ks, sc)))


@pytest.mark.parametrize("group,log_n", [(1, 16), (1, 20), (2, 18), (1, 21), (1, 22), (2, 22), (1, 24)])
def test_msm_config_sizes_closed_form(gpu, group, log_n):
    """G1 2^16, 2^20, G2 2^18 (the single-call sizes) and the Groth16-set / serving sizes up to
    2^24 (window 16: the largest LDS sort) at the default window, exact at full size."""
    from zkgpu import Group, Msm

    if quick() and log_n > 16:
        pytest.skip("ZK_GPU_QUICK")
    n = 1 << log_n
    rng = np.random.default_rng(1000 + 10 * group + log_n)
    ks, sc = uniform_fr_words(rng, n), uniform_fr_words(rng, n)
    bb = device_bases(gpu, group, ks)
    m = Msm.for_group(gpu, Group(group), n, Msm.default_window(n))
    got = affine(gpu, m.run(gpu, bb, gpu.buffer(sc)), group)
    assert np.array_equal(got, O.point_mul(group, O.generator_words(group), O.closed_form_scalar(ks, sc)))


@pytest.mark.parametrize("n", [1, 2, 3, 17, 255, 1000, 4000])
def test_g2_msm_matches_oracle(gpu, n):
    """tests/g2.rs:50-64: several n under (c, items) in {(13, auto), (9, 1), (10, 64)}."""
    rng = np.random.default_rng(21 + n)
    ks, sc = uniform_fr_words(rng, n), uniform_fr_words(rng, n)
    bases = O.multiples(2, O.generator_words(2), ks)
    exp = O.msm(2, bases, sc)
    for c, items in ((13, None), (9, 1), (10, 64)):
        assert np.array_equal(gpu_msm(gpu, 2, bases, sc, c, items), exp), (n, c, items)


def test_g2_special_scalars_and_points(gpu):
    """tests/g2.rs:66-87: scalars 0, 1, -1, 2^200 with infinity and repeated bases."""
    n = 64
    rng = np.random.default_rng(66)
    ks = uniform_fr_words(rng, n)
    bases = O.multiples(2, O.generator_words(2), ks).reshape(n, 32)
    bases[3] = 0
    bases[10:20] = bases[9]
    sc = uniform_fr_words(rng, n).reshape(n, 8)
    sc[0] = 0
    sc[1] = O.int_to_words(1)
    sc[2] = O.int_to_words(O.R - 1)
    sc[4] = O.int_to_words(1 << 200)
    sc[10:15] = sc[9]
    exp = O.msm(2, bases, sc)
    for c, items in ((13, None), (9, 1)):
        assert np.array_equal(gpu_msm(gpu, 2, bases.reshape(-1), sc.reshape(-1), c, items), exp)


@pytest.mark.parametrize("group", [1, 2])
def test_montgomery_256_bases(gpu, group):
    """tests/properties.rs:221-246, tests/g2.rs:89-111: bases given as x * 2^256 mod p, incl. infinity."""
    n = 500
    rng = np.random.default_rng(222 + group)
    ks, sc = uniform_fr_words(rng, n), uniform_fr_words(rng, n)
    bases = O.multiples(group, O.generator_words(group), ks).reshape(n, 16 * group)
    bases[7] = 0
    exp = O.msm(group, bases, sc)
    mont = O.ints_to_words([(v << 256) % O.P for v in O.words_to_ints(bases)])
    assert np.array_equal(gpu_msm(gpu, group, mont, sc, 11, None, mont256=True), exp)



# (group, c, width of window 0): G1 with the GLV split (128-bit halves: 8 windows of 16 bits at
# c = 16; two of 15 and seven of 14 at c = 15), G2 with the GLS split (65 bits: 5 windows of 13 at
# c = 13; five of 11 and one of 10 at c = 11) and with the GLV split at c = 17 (8 windows of 16)
@pytest.mark.parametrize("group,c,w0", [(1, 16, 16), (1, 15, 15), (2, 13, 13), (2, 11, 11), (2, 17, 16)])
def test_combine_addition_doubles_and_cancels(gpu, group, c, w0):
    """The combine's window addition at its special cases: with bases Q and 2^w0 Q and scalars 2^w0
    and 1, window 1 holds Q and window 0 holds 2^w0 Q, so after w0 doublings the chain adds a point
    to itself (the doubling case); with -2^w0 Q the two cancel (infinity)."""
    rng = np.random.default_rng(300 + c)
    q = int(rng.integers(1, 1 << 62))
    sc = O.ints_to_words([1 << w0, 1]).reshape(2, 8)
    for k2, expect_inf in ((q << w0) % O.R, False), ((O.R - (q << w0) % O.R) % O.R, True):
        ks = O.ints_to_words([q, k2])
        bases = O.multiples(group, O.generator_words(group), ks).reshape(2, 16 * group)
        exp = O.msm(group, bases, sc)
        assert (not exp.any()) == expect_inf
        assert np.array_equal(gpu_msm(gpu, group, bases.reshape(-1), sc.reshape(-1), c), exp)


@pytest.mark.parametrize("group,c", [(1, 16), (2, 13)])
def test_window_sum_partials_double_and_cancel(gpu, group, c):
    """The window sums' fold at its special cases: scalars 11 and 1001 put P1 and P2 in buckets 10
    and 1000 of window 0, which the bucket reduction gives to different workgroups (512 buckets per
    G1 workgroup, 128 per G2), so the window sum adds the partials 11 P1 and 1001 P2; with
    P2 = +-(11 / 1001) P1 they are equal (doubling) or opposite (infinity)."""
    a = 123456789
    w = 11 * pow(1001, -1, O.R) % O.R
    sc = O.ints_to_words([11, 1001]).reshape(2, 8)
    for k2, expect_inf in ((a * w % O.R, False), ((O.R - a * w % O.R) % O.R, True)):
        bases = O.multiples(group, O.generator_words(group), O.ints_to_words([a, k2])).reshape(2, 16 * group)
        exp = O.msm(group, bases, sc)
        assert (not exp.any()) == expect_inf
        assert np.array_equal(gpu_msm(gpu, group, bases.reshape(-1), sc.reshape(-1), c), exp)


def test_msm_linearity_and_scaling(gpu):
    """tests/properties.rs:150-185: MSM(a s + t) = a MSM(s) + MSM(t); MSM(k s) = k MSM(s)."""
    n = 2000
    rng = np.random.default_rng(150)
    ks = uniform_fr_words(rng, n)
    bases = O.multiples(1, O.generator_words(1), ks)
    s, t = uniform_fr_words(rng, n), uniform_fr_words(rng, n)
    a = 0xDEADBEEF1234567
    comb = O.ints_to_words([(a * x + y) % O.R for x, y in zip(O.words_to_ints(s), O.words_to_ints(t))])
    ms, mt, mc = (gpu_msm(gpu, 1, bases, v, 11) for v in (s, t, comb))
    assert np.array_equal(mc, O.point_add(1, O.point_mul(1, ms, a), mt))


def test_msm_order_independence_and_window_independence(gpu):
    """tests/properties.rs:187-202,344-355: permuting the pairs or changing c leaves the result."""
    n = 3000
    rng = np.random.default_rng(187)
    ks, sc = uniform_fr_words(rng, n), uniform_fr_words(rng, n)
    bases = O.multiples(1, O.generator_words(1), ks)
    base_res = gpu_msm(gpu, 1, bases, sc, 11)
    perm = rng.permutation(n)
    pb = bases.reshape(n, 16)[perm].reshape(-1)
    ps = sc.reshape(n, 8)[perm].reshape(-1)
    assert np.array_equal(gpu_msm(gpu, 1, pb, ps, 11), base_res)
    for c in (9, 10, 12, 13, 14, 15, 16, 17, 20):
        assert np.array_equal(gpu_msm(gpu, 1, bases, sc, c), base_res), c


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 255, 256, 257, 1023, 1025])
def test_msm_tiny_n(gpu, n):
    """tests/properties.rs:204-219: tiny n, and n = 1 gives s P."""
    rng = np.random.default_rng(204 + n)
    ks, sc = uniform_fr_words(rng, n), uniform_fr_words(rng, n)
    bases = O.multiples(1, O.generator_words(1), ks)
    assert np.array_equal(gpu_msm(gpu, 1, bases, sc), O.msm(1, bases, sc))


def test_msm_huge_buckets(gpu):
    """tests/properties.rs:248-266: most scalars equal -> one bucket per window holds most points."""
    n = 20000
    rng = np.random.default_rng(248)
    ks = uniform_fr_words(rng, n)
    bases = O.multiples(1, O.generator_words(1), ks)
    sc = uniform_fr_words(rng, n).reshape(n, 8)
    sc[: n * 9 // 10] = sc[0]
    for c, items in ((10, None), (10, 1), (13, 8)):
        assert np.array_equal(gpu_msm(gpu, 1, bases, sc.reshape(-1), c, items), O.msm(1, bases, sc.reshape(-1)))


def test_msm_window_boundary_scalars(gpu):
    """tests/properties.rs:311-342: 2^k - 1, 2^k, 2^k + 1 around every window boundary, c = 9..16."""
    vals = []
    for k in range(1, 254):
        for d in (-1, 0, 1):
            v = (1 << k) + d
            if 0 <= v < O.R:
                vals.append(v)
    n = len(vals)
    rng = np.random.default_rng(311)
    ks = uniform_fr_words(rng, n)
    bases = O.multiples(1, O.generator_words(1), ks)
    sc = O.ints_to_words(vals)
    exp = O.msm(1, bases, sc)
    for c in range(9, 17):
        assert np.array_equal(gpu_msm(gpu, 1, bases, sc, c), exp), c


def test_msm_plan_reuse(gpu):
    """tests/properties.rs:357-375: one plan, several inputs; the result buffer is overwritten."""
    from zkgpu import Msm

    n = 5000
    rng = np.random.default_rng(357)
    m = Msm.new(gpu, n, 12)
    for _ in range(3):
        ks, sc = uniform_fr_words(rng, n), uniform_fr_words(rng, n)
        bases = O.multiples(1, O.generator_words(1), ks)
        got = affine(gpu, m.run(gpu, Msm.upload_bases(gpu, bases), gpu.buffer(sc)), 1)
        assert np.array_equal(got, O.msm(1, bases, sc))


@pytest.mark.parametrize("group", [1, 2])
def test_msm_witness_scalars(gpu, group):
    """Groth16-shaped scalars (40% in {0, 1}): the skewed-bucket case of the batched config."""
    n = 1 << 15
    rng = np.random.default_rng(400 + group)
    ks, sc = uniform_fr_words(rng, n), witness_scalars(rng, n)
    from zkgpu import Group, Msm

    bb = device_bases(gpu, group, ks)
    m = Msm.for_group(gpu, Group(group), n, Msm.default_window(n))
    got = affine(gpu, m.run(gpu, bb, gpu.buffer(sc)), group)
    assert np.array_equal(got, O.point_mul(group, O.generator_words(group), O.closed_form_scalar(ks, sc)))


def test_msm_seeded_fuzz(gpu):
    """tests/properties.rs:377-431: random n, c, item size and scalar shapes."""
    import os

    seed = int(os.environ.get("ZK_GPU_FUZZ_SEED", "377"))
    iters = int(os.environ.get("ZK_GPU_FUZZ_ITERS", "6"))
    rng = np.random.default_rng(seed)
    for _ in range(iters):
        n = int(rng.integers(1, 3000))
        c = int(rng.integers(9, 17))
        items = [None, 1, 3, 8, 64][int(rng.integers(0, 5))]
        ks = uniform_fr_words(rng, n)
        bases = O.multiples(1, O.generator_words(1), ks)
        sc = witness_scalars(rng, n) if rng.integers(0, 2) else uniform_fr_words(rng, n)
        assert np.array_equal(gpu_msm(gpu, 1, bases, sc, c, items), O.msm(1, bases, sc)), (seed, n, c, items)


@pytest.mark.parametrize("c", list(range(9, 21)))
def test_msm_every_window_vs_oracle(gpu, c):
    """Every window the reference accepts (9..=20: the LDS counting sort up to 16, the atomic sort
    above), G1 at n = 3000 and G2 at n = 700, against the oracle."""
    rng = np.random.default_rng(3100 + c)
    for group, n in ((1, 3000), (2, 700)):
        ks = uniform_fr_words(rng, n)
        bases = O.multiples(group, O.generator_words(group), ks)
        sc = uniform_fr_words(rng, n)
        assert np.array_equal(gpu_msm(gpu, group, bases, sc, c, 0), O.msm(group, bases, sc)), (group, c)


@pytest.mark.parametrize("items", [1, 2, 3, 5, 8, 64])
def test_msm_item_sizes_vs_oracle(gpu, items):
    """The work-item sizes of the reference's test matrix besides 7 and 1024 (which
    test_msm_random_vs_oracle covers), at a window where buckets hold ~20 points."""
    rng = np.random.default_rng(3200 + items)
    n = 6000
    ks = uniform_fr_words(rng, n)
    bases = O.multiples(1, O.generator_words(1), ks)
    sc = uniform_fr_words(rng, n)
    assert np.array_equal(gpu_msm(gpu, 1, bases, sc, 10, items), O.msm(1, bases, sc))


def test_msm_nearly_empty_top_window(gpu):
    """tests/properties.rs:268-280: with the reference's 255-bit digits and 14-bit windows the top
    window holds 3 bits, so all points crowd into a few buckets; plans (14, 8), (14, 1), (11, 8),
    with the GLV split on and off."""
    from zkgpu import Group, Msm

    rng = np.random.default_rng(17)
    n = 20000
    ks = uniform_fr_words(rng, n)
    bases = O.multiples(1, O.generator_words(1), ks)
    sc = uniform_fr_words(rng, n)
    exp = O.msm(1, bases, sc)
    bb = Msm.upload_bases(gpu, bases)
    sb = gpu.buffer(sc)
    old = os.environ.get("ZKGPU_MSM_GLV")
    try:
        for flag in ("0", "1"):
            os.environ["ZKGPU_MSM_GLV"] = flag
            for c, items in ((14, 8), (14, 1), (11, 8)):
                m = Msm.for_group(gpu, Group.G1, n, c, items)
                assert np.array_equal(affine(gpu, m.run(gpu, bb, sb), 1), exp), (flag, c, items)
    finally:
        if old is None:
            os.environ.pop("ZKGPU_MSM_GLV", None)
        else:
            os.environ["ZKGPU_MSM_GLV"] = old


def test_msm_generator_progression_2_18(gpu):
    """tests/properties.rs:296-309: bases (i + 1) G for i < 2^18 (a progression, as the reference
    builds them), plans (default) and (15, 8), exact against (sum (i + 1) s_i) G."""
    from zkgpu import Group, Msm

    if quick():
        pytest.skip("ZK_GPU_QUICK")
    n = 1 << 18
    rng = np.random.default_rng(18)
    ks = O.ints_to_words(range(1, n + 1))
    sc = uniform_fr_words(rng, n)
    bb = device_bases(gpu, 1, ks)
    sb = gpu.buffer(sc)
    want = O.point_mul(1, O.generator_words(1), O.closed_form_scalar(ks, sc))
    for c, items in ((Msm.default_window(n), None), (15, 8)):
        m = Msm.for_group(gpu, Group.G1, n, c, items)
        assert np.array_equal(affine(gpu, m.run(gpu, bb, sb), 1), want), (c, items)


def test_result_buffer_keeps_its_plan(gpu):
    """The result buffer `run` returns is plan-owned memory: it keeps the plan alive, so the
    result of a plan that is referenced nowhere else (a temporary) can still be read."""
    import gc

    from zkgpu import Group, Msm

    n = 700
    rng = np.random.default_rng(5500)
    ks, sc = uniform_fr_words(rng, n), uniform_fr_words(rng, n)
    bases = O.multiples(1, O.generator_words(1), ks)
    bb, sb = Msm.upload_bases(gpu, bases), gpu.buffer(sc)
    res = Msm.for_group(gpu, Group.G1, n, 10).run(gpu, bb, sb)
    gc.collect()
    assert np.array_equal(affine(gpu, res, 1), O.msm(1, bases, sc))
