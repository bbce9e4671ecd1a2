"""MSMs over precomputed bases (zkg_msm_precompute_bases + zkg_msm_create_precomputed): the bases
table holds P_i and 2^64 P_i, every 128-bit GLV half-scalar splits at bit 64, and the stages run over
4n virtual points with ceil(64 / c) windows. No counterpart in the reference.

CPU: the window layout and signed digits of the split (msm_digits_split_store_kernel) are restated
and checked against what the arithmetic fixes: the digits rebuild every half-scalar exactly, with
the high half weighted by 2^64, no digit exceeds 2^(width - 1) and no carry leaves the top window,
for every c in 9..=16 on edge and random magnitudes below 2^127.
GPU: the table's second half against the oracle (a plain plan over the whole table with scalars
only on the high half), and precomputed plans against the oracle and against plain plans, in G1
and G2, with edge scalars and infinity bases; plus the argument errors."""
import random

import numpy as np
import pytest

import oracle as O
from conftest import uniform_fr_words

R = O.R


def _layout(c):
    """(W, wb, nw) of a precomputed plan: W windows over 64 bits, widths wb + (w < nw)."""
    W = -(-64 // c)
    wb = 64 // W
    return W, wb, 64 - wb * W


def _split_digits(k, c):
    """Signed digits of a magnitude k < 2^127 as the split kernel forms them: [(half, window,
    digit, width)] with the carry running on from the low half into the high half."""
    W, wb, nw = _layout(c)
    out, carry = [], 0
    for h in range(2):
        for w in range(W):
            width = wb + (1 if w < nw else 0)
            pos = 64 * h + w * wb + min(w, nw)
            v = ((k >> pos) & ((1 << width) - 1)) + carry
            half = 1 << (width - 1)
            wrap = v > half  # half-scalars may take +2^(width-1)
            d = v - (1 << width) if wrap else v
            carry = 1 if wrap else 0
            out.append((h, w, d, width, pos - 64 * h))
    return out, carry


@pytest.mark.parametrize("c", range(9, 17))
def test_split_digits_rebuild_every_half_scalar(c):
    W, wb, nw = _layout(c)
    assert wb * W + nw == 64 and 0 <= nw < W and wb <= c and wb + (1 if nw else 0) <= c
    rng = random.Random(640 + c)
    ks = [0, 1, (1 << 64) - 1, 1 << 64, (1 << 127) - 1, (1 << 126), ((1 << 127) - 1) ^ ((1 << 64) - 1),
          int("10" * 63, 2), int("01" * 63, 2)]
    # magnitudes whose every window sits at the wrap threshold (long carry chains across bit 64)
    ks.append(sum(1 << (64 * h + w * wb + min(w, nw) + wb + (1 if w < nw else 0) - 1)
                  for h in range(2) for w in range(W)) & ((1 << 127) - 1))
    ks += [rng.getrandbits(127) for _ in range(400)]
    for k in ks:
        digits, carry = _split_digits(k, c)
        assert carry == 0, (c, k)
        lo = sum(d << pos for h, w, d, width, pos in digits if h == 0)
        hi = sum(d << pos for h, w, d, width, pos in digits if h == 1)
        assert lo + (hi << 64) == k, (c, k)
        assert all(abs(d) <= 1 << (width - 1) for _, _, d, width, _ in digits)
        assert all(-(1 << (width - 1)) < d for _, _, d, width, _ in digits)  # -2^(w-1) never occurs


def _edge_scalars(rng, n):
    sc = uniform_fr_words(rng, n)
    edges = [0, 1, R - 1, (1 << 256) - 1, 1 << 64, (1 << 64) - 1, 1 << 127, (1 << 128) + 12345, R >> 1]
    for j, v in enumerate(edges):
        if 8 * (j + 1) <= sc.size:
            sc[8 * j: 8 * (j + 1)] = O.int_to_words(v)
    return sc


@pytest.mark.gpu
@pytest.mark.parametrize("group,n", [(1, 3000), (2, 700)])
def test_table_high_half_is_2_64_times_each_base(gpu, group, n):
    from test_gpu_msm import affine
    from zkgpu import Group, Msm

    rng = np.random.default_rng(650 + group)
    ks = uniform_fr_words(rng, n)
    bases = O.multiples(group, O.generator_words(group), ks)
    aw = 16 * group
    bases[aw * 5: aw * 6] = 0  # an infinity base
    bb = Msm.upload_bases(gpu, bases)
    table = Msm.precompute_bases(gpu, Group(group), bb, n)
    assert table.size() == 2 * n * 4 * aw
    sc = uniform_fr_words(rng, n)
    both = np.concatenate([np.zeros(8 * n, dtype=np.uint32), sc])  # only the 2^64 P_i half
    m = Msm.for_group(gpu, Group(group), 2 * n, 11)
    got = affine(gpu, m.run(gpu, table, gpu.buffer(both)), group)
    shifted = O.ints_to_words([(v << 64) % R for v in O.words_to_ints(sc)])
    assert np.array_equal(got, O.msm(group, bases, shifted))
    lo = np.concatenate([sc, np.zeros(8 * n, dtype=np.uint32)])  # the copied P_i half
    assert np.array_equal(affine(gpu, m.run(gpu, table, gpu.buffer(lo)), group), O.msm(group, bases, sc))


@pytest.mark.gpu
@pytest.mark.parametrize("group,n,c", [(1, 1, 9), (1, 5000, 9), (1, 20000, 13), (1, 70000, 16), (2, 1500, 11),
                                       (2, 3000, 14), (2, 600, 16)])
def test_precomputed_plan_matches_oracle_and_plain_plan(gpu, group, n, c):
    from test_gpu_msm import affine
    from zkgpu import Group, Msm

    rng = np.random.default_rng(660 + n + c)
    ks = uniform_fr_words(rng, n)
    bases = O.multiples(group, O.generator_words(group), ks)
    aw = 16 * group
    if n > 10:
        bases[aw * 3: aw * 4] = 0  # infinity
        bases[aw * 7: aw * 8] = bases[aw * 2: aw * 3]  # a repeated base
    sc = _edge_scalars(rng, n)
    exp = O.msm(group, bases, sc)
    bb = Msm.upload_bases(gpu, bases)
    sb = gpu.buffer(sc)
    table = Msm.precompute_bases(gpu, Group(group), bb, n)
    pre = Msm.precomputed(gpu, Group(group), n, c)
    assert pre.base_points() == 2 * n and pre.window() == c
    for rep in range(2):  # the plan's workspace is reused
        assert np.array_equal(affine(gpu, pre.run(gpu, table, sb), group), exp), (group, n, c, rep)
    plain = Msm.for_group(gpu, Group(group), n, c)
    assert plain.base_points() == n
    assert np.array_equal(affine(gpu, plain.run(gpu, bb, sb), group), exp)


@pytest.mark.gpu
def test_precomputed_plan_in_a_batch(gpu):
    from test_gpu_msm import affine
    from zkgpu import Batch, Group, Msm

    rng = np.random.default_rng(670)
    jobs = []
    batch = Batch(gpu)
    for group, n, c in ((1, 4000, 12), (2, 900, 10), (1, 2500, 15)):
        bases = O.multiples(group, O.generator_words(group), uniform_fr_words(rng, n))
        sc = uniform_fr_words(rng, n)
        table = Msm.precompute_bases(gpu, Group(group), Msm.upload_bases(gpu, bases), n)
        plan = Msm.precomputed(gpu, Group(group), n, c)
        sb = gpu.buffer(sc)
        batch.add_msm(plan, table, sb)
        jobs.append((group, plan, O.msm(group, bases, sc), table, sb))
    batch.run()
    gpu.wait()
    for group, plan, exp, _, _ in jobs:
        assert np.array_equal(affine(gpu, plan.result_buffer(gpu), group), exp), group


@pytest.mark.gpu
def test_precomputed_argument_errors(gpu):
    from zkgpu import Group, Msm, ZkgError

    rng = np.random.default_rng(680)
    n = 300
    bases = O.multiples(1, O.generator_words(1), uniform_fr_words(rng, n))
    bb = Msm.upload_bases(gpu, bases)
    with pytest.raises(ZkgError, match="takes c in 9..=16, got 17"):
        Msm.precomputed(gpu, Group.G1, n, 17)
    with pytest.raises(ZkgError, match="window size c must be in 9..=20, got 8"):
        Msm.precomputed(gpu, Group.G1, n, 8)
    with pytest.raises(ZkgError, match="bases buffer holds"):
        Msm.precompute_bases(gpu, Group.G1, bb, n + 1)
    pre = Msm.precomputed(gpu, Group.G1, n, 11)
    # the original bases (n points) are not a table of 2n
    with pytest.raises(ZkgError, match=f"bases buffer holds {64 * n} bytes; {2 * n} points need {128 * n}"):
        pre.run(gpu, bb, gpu.buffer(uniform_fr_words(rng, n)))


@pytest.mark.gpu
@pytest.mark.parametrize("group,log_n,c", [(1, 20, 16), (2, 18, 16), (1, 21, 15), (2, 20, 14)])
def test_precomputed_plans_at_bench_sizes_closed_form(gpu, group, log_n, c):
    """The precomputed plans the bench reports (G1 2^20 and G2 2^18 at c = 16) and two more
    (size, window) pairs, exact at full size against the closed form (sum k_i s_i mod r) G."""
    from conftest import quick
    from test_gpu_msm import affine, device_bases
    from zkgpu import Group, Msm

    if quick():
        pytest.skip("ZK_GPU_QUICK")
    n = 1 << log_n
    rng = np.random.default_rng(690 + 10 * group + log_n)
    ks, sc = uniform_fr_words(rng, n), uniform_fr_words(rng, n)
    bb = device_bases(gpu, group, ks)
    table = Msm.precompute_bases(gpu, Group(group), bb, n)
    m = Msm.precomputed(gpu, Group(group), n, c)
    got = affine(gpu, m.run(gpu, table, gpu.buffer(sc)), group)
    assert np.array_equal(got, O.point_mul(group, O.generator_words(group), O.closed_form_scalar(ks, sc)))
