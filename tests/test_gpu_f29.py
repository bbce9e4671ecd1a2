"""The MSM pipelines in 29-bit limbs (G1: msm_accumulate29_kernel and the Fq29 stages; G2: the Fq2_29
stages) and the 32-bit-limb pipelines
both against the oracle, on inputs that drive every special case of the accumulation: equal
points in one bucket (doubling), a point and its negative (cancellation), points at infinity,
zero and r - 1 scalars, and work items long enough to run the lazy bounds through many steps."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import uniform_fr_words

pytestmark = pytest.mark.gpu


def _inputs(rng, n):
    ks = uniform_fr_words(rng, n)
    sc = uniform_fr_words(rng, n)
    for i in range(0, n - 8, 97):
        ks[8 * (i + 1): 8 * (i + 2)] = ks[8 * i: 8 * (i + 1)]  # the same point ...
        sc[8 * (i + 1): 8 * (i + 2)] = sc[8 * i: 8 * (i + 1)]  # ... with the same scalar: doubling
        k = O.words_to_int(ks[8 * (i + 3): 8 * (i + 4)])
        ks[8 * (i + 4): 8 * (i + 5)] = O.int_to_words((O.R - k) % O.R)  # -P ...
        sc[8 * (i + 4): 8 * (i + 5)] = sc[8 * (i + 3): 8 * (i + 4)]  # ... with P's scalar: cancels
        ks[8 * (i + 6): 8 * (i + 7)] = 0  # infinity
    sc[0:8] = 0
    sc[8:16] = O.int_to_words(O.R - 1)
    return ks, sc


@pytest.mark.parametrize("n,c,item_points", [(3000, 12, 0), (5000, 16, 0), (4000, 11, 64), (20000, 9, 1024)])
def test_f29_and_32bit_accumulation_match_oracle(gpu, n, c, item_points):
    from test_gpu_msm import affine
    from zkgpu import Group, Msm

    rng = np.random.default_rng(1400 + n + c)
    ks, sc = _inputs(rng, n)
    bases = O.multiples(1, O.generator_words(1), ks)
    exp = O.msm(1, bases, sc)
    bb = Msm.upload_bases(gpu, bases)
    sb = gpu.buffer(sc)
    old = os.environ.get("ZKGPU_MSM_F29")
    try:
        for flag in ("1", "0"):
            os.environ["ZKGPU_MSM_F29"] = flag
            m = Msm.for_group(gpu, Group.G1, n, c, item_points or None)
            assert np.array_equal(affine(gpu, m.run(gpu, bb, sb), 1), exp), (flag, n, c, item_points)
    finally:
        if old is None:
            os.environ.pop("ZKGPU_MSM_F29", None)
        else:
            os.environ["ZKGPU_MSM_F29"] = old


def test_f29_closed_form_at_2_20(gpu):
    """Both kernels at 2^20 points (bases k_i G): sum s_i k_i G by one Fr dot product."""
    from test_gpu_msm import affine, device_bases
    from zkgpu import Group, Msm

    n = 1 << 20
    rng = np.random.default_rng(1410)
    ks, sc = uniform_fr_words(rng, n), uniform_fr_words(rng, n)
    bb = device_bases(gpu, 1, ks)
    exp = O.point_mul(1, O.generator_words(1), O.closed_form_scalar(ks, sc))
    sb = gpu.buffer(sc)
    old = os.environ.get("ZKGPU_MSM_F29")
    try:
        for flag in ("1", "0"):
            os.environ["ZKGPU_MSM_F29"] = flag
            m = Msm.for_group(gpu, Group.G1, n, 15)
            assert np.array_equal(affine(gpu, m.run(gpu, bb, sb), 1), exp), flag
    finally:
        if old is None:
            os.environ.pop("ZKGPU_MSM_F29", None)
        else:
            os.environ["ZKGPU_MSM_F29"] = old


@pytest.mark.parametrize("n,c", [(1500, 11), (2500, 13)])
def test_g2_f29_and_32bit_stages_match_oracle(gpu, n, c):
    """G2: the 32-bit accumulation feeding the Fq2-over-29-bit stages, and the all-32-bit path."""
    from test_gpu_msm import affine
    from zkgpu import Group, Msm

    rng = np.random.default_rng(1420 + n)
    ks, sc = _inputs(rng, n)
    bases = O.multiples(2, O.generator_words(2), ks)
    exp = O.msm(2, bases, sc)
    bb = Msm.upload_bases(gpu, bases)
    sb = gpu.buffer(sc)
    old = os.environ.get("ZKGPU_MSM_F29")
    try:
        for flag in ("1", "0"):
            os.environ["ZKGPU_MSM_F29"] = flag
            m = Msm.for_group(gpu, Group.G2, n, c)
            assert np.array_equal(affine(gpu, m.run(gpu, bb, sb), 2), exp), (flag, n, c)
    finally:
        if old is None:
            os.environ.pop("ZKGPU_MSM_F29", None)
        else:
            os.environ["ZKGPU_MSM_F29"] = old


@pytest.mark.parametrize("n,c,item_points,witness", [
    (3000, 12, 0, False),    # buckets of about 3 entries, 8 per lane: lanes meet several buckets
    (20000, 15, 0, True),    # 14- and 15-bit windows (half of the 14-bit windows' slots empty), heavy buckets
    (5000, 11, 1, False),    # one entry per lane
    (4000, 9, 64, True),     # long lanes over small windows, merge rounds
    (60000, 13, 0, False),
])
def test_lane_cut_accumulation_matches_oracle(gpu, n, c, item_points, witness):
    """The G1 accumulation with lanes cut at multiples of the item size of the whole sorted list
    (ZKGPU_MSM_LANES=1; by default only for large plans) and cut at bucket starts (=0), on the
    special-case inputs above (doubling, cancellation, infinity, zero and r - 1 scalars) and
    witness-like scalars, against the oracle: partial sums stashed at bucket boundaries, lanes that
    meet three or more buckets, and empty slots between windows."""
    from conftest import witness_scalars
    from test_gpu_msm import affine
    from zkgpu import Group, Msm

    rng = np.random.default_rng(1430 + n + c)
    ks, sc = _inputs(rng, n)
    if witness:
        sc = witness_scalars(rng, n)
    bases = O.multiples(1, O.generator_words(1), ks)
    exp = O.msm(1, bases, sc)
    bb = Msm.upload_bases(gpu, bases)
    sb = gpu.buffer(sc)
    old = os.environ.get("ZKGPU_MSM_LANES")
    try:
        for flag in ("1", "0"):
            os.environ["ZKGPU_MSM_LANES"] = flag
            m = Msm.for_group(gpu, Group.G1, n, c, item_points or None)
            assert np.array_equal(affine(gpu, m.run(gpu, bb, sb), 1), exp), (flag, n, c, item_points)
    finally:
        if old is None:
            os.environ.pop("ZKGPU_MSM_LANES", None)
        else:
            os.environ["ZKGPU_MSM_LANES"] = old


@pytest.mark.parametrize("group,f29,n,c,item_points", [
    (2, "1", 1500, 11, 0),
    (2, "1", 2500, 9, 8),
    (2, "0", 2000, 10, 0),
    (1, "0", 3000, 12, 0),
    (1, "0", 20000, 15, 0),
])
def test_lane_cut_accumulation_other_fields(gpu, group, f29, n, c, item_points):
    """The lane-cut accumulation of the generic kernel: G2 (32-bit Fq2 additions, item sums stored
    for the 29-bit stages or the 32-bit ones) and the 32-bit G1 pipeline, against the oracle, with
    ZKGPU_MSM_LANES=1 and =0."""
    from test_gpu_msm import affine
    from zkgpu import Group, Msm

    rng = np.random.default_rng(1440 + n + c + group)
    ks, sc = _inputs(rng, n)
    bases = O.multiples(group, O.generator_words(group), ks)
    exp = O.msm(group, bases, sc)
    bb = Msm.upload_bases(gpu, bases)
    sb = gpu.buffer(sc)
    old = {k: os.environ.get(k) for k in ("ZKGPU_MSM_LANES", "ZKGPU_MSM_F29")}
    try:
        os.environ["ZKGPU_MSM_F29"] = f29
        for flag in ("1", "0"):
            os.environ["ZKGPU_MSM_LANES"] = flag
            m = Msm.for_group(gpu, Group(group), n, c, item_points or None)
            assert np.array_equal(affine(gpu, m.run(gpu, bb, sb), group), exp), (flag, group, f29, n, c)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
