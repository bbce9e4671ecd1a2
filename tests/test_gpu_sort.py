"""The MSM's two counting-sort placements against the oracle: the two-pass placement (entries
first written as runs per bucket partition, then placed per partition; the default for c <= 16)
and the one-pass placement (each entry written at its bucket's cursor; ZKGPU_MSM_PARTSORT=0).
Windows 9..16 cover every partition width (2..8 bucket bits per partition, 64 or 128 partitions
per window); one-bucket inputs put a whole window's entries into one partition; precomputed plans
place 4n virtual points."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import uniform_fr_words
from test_gpu_msm import affine

pytestmark = pytest.mark.gpu


def _both_placements(run):
    old = os.environ.get("ZKGPU_MSM_PARTSORT")
    try:
        out = []
        for flag in ("1", "0"):
            os.environ["ZKGPU_MSM_PARTSORT"] = flag
            out.append(run())
        return out
    finally:
        if old is None:
            os.environ.pop("ZKGPU_MSM_PARTSORT", None)
        else:
            os.environ["ZKGPU_MSM_PARTSORT"] = old


@pytest.mark.parametrize("c", list(range(9, 17)))
def test_placements_every_window_vs_oracle(gpu, c):
    from zkgpu import Group, Msm

    n = 3000 + 37 * c
    rng = np.random.default_rng(5100 + c)
    ks, sc = uniform_fr_words(rng, n), uniform_fr_words(rng, n)
    sc[0:8] = 0
    sc[8:16] = O.int_to_words(O.R - 1)
    bases = O.multiples(1, O.generator_words(1), ks)
    exp = O.msm(1, bases, sc)
    bb, sb = Msm.upload_bases(gpu, bases), gpu.buffer(sc)
    got = _both_placements(lambda: affine(gpu, Msm.for_group(gpu, Group.G1, n, c).run(gpu, bb, sb), 1))
    for flag, g in zip(("two-pass", "one-pass"), got):
        assert np.array_equal(g, exp), (flag, c)


@pytest.mark.parametrize("group,c", [(1, 15), (1, 10), (2, 13)])
def test_placements_one_bucket(gpu, group, c):
    """Every scalar equal: each window's entries all fall into one bucket, hence one partition."""
    from zkgpu import Group, Msm

    n = 20000 if group == 1 else 6000
    rng = np.random.default_rng(5200 + group + c)
    ks = uniform_fr_words(rng, n)
    s = uniform_fr_words(rng, 1)
    sc = np.tile(s, n)
    bases = O.multiples(group, O.generator_words(group), ks)
    exp = O.msm(group, bases, sc)
    bb, sb = Msm.upload_bases(gpu, bases), gpu.buffer(sc)
    got = _both_placements(lambda: affine(gpu, Msm.for_group(gpu, Group(group), n, c).run(gpu, bb, sb), group))
    for flag, g in zip(("two-pass", "one-pass"), got):
        assert np.array_equal(g, exp), (flag, group, c)


@pytest.mark.parametrize("group", [1, 2])
def test_placements_g2_and_witness_mix(gpu, group):
    from conftest import witness_scalars
    from zkgpu import Group, Msm

    n = 5000
    rng = np.random.default_rng(5300 + group)
    ks = uniform_fr_words(rng, n)
    sc = witness_scalars(rng, n)
    bases = O.multiples(group, O.generator_words(group), ks)
    exp = O.msm(group, bases, sc)
    bb, sb = Msm.upload_bases(gpu, bases), gpu.buffer(sc)
    got = _both_placements(
        lambda: affine(gpu, Msm.for_group(gpu, Group(group), n, Msm.default_window(n)).run(gpu, bb, sb), group))
    for flag, g in zip(("two-pass", "one-pass"), got):
        assert np.array_equal(g, exp), (flag, group)


def test_placements_precomputed_plan(gpu):
    """A plan over precomputed bases (4n virtual points: P, 2^64 P and their endomorphism images)."""
    from zkgpu import Group, Msm

    n, c = 4000, 12
    rng = np.random.default_rng(5400)
    ks, sc = uniform_fr_words(rng, n), uniform_fr_words(rng, n)
    bases = O.multiples(1, O.generator_words(1), ks)
    exp = O.msm(1, bases, sc)
    bb, sb = Msm.upload_bases(gpu, bases), gpu.buffer(sc)
    table = Msm.precompute_bases(gpu, Group.G1, bb, n)
    got = _both_placements(lambda: affine(gpu, Msm.precomputed(gpu, Group.G1, n, c).run(gpu, table, sb), 1))
    for flag, g in zip(("two-pass", "one-pass"), got):
        assert np.array_equal(g, exp), (flag, c)
