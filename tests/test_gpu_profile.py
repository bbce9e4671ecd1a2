"""Msm::profile (src/msm.rs:321-342): one run with a wait around every stage, returning (stage,
milliseconds) in the order the stages first run, times of repeated names added up, with the stage
names of the reference's pass graph (src/msm.rs:217-263): "count", "scan", "scatter",
"accumulate", "merge 1" .. "merge 6" (later rounds counted under "merge 6"), "finalize",
"reduce chunks", "tree sum", "combine". A profiled run leaves the same result as run; Msm::group
(src/msm.rs:267-269) returns the plan's group."""
import numpy as np
import pytest

import oracle as O
from conftest import uniform_fr_words, witness_scalars

pytestmark = pytest.mark.gpu


def _names(rounds):
    merges = ["merge %d" % (r + 1) for r in range(min(rounds, 6))]
    return ["count", "scan", "scatter", "accumulate"] + merges + ["finalize", "reduce chunks", "tree sum", "combine"]


def _merge_rounds(nv, item_points):
    """Rounds of fan-in 4 until the fullest possible bucket has at most 16 partial sums left: a
    bucket of all nv virtual points meets ceil((nv - 1) / item_points) + 1 lanes of item_points
    consecutive sorted entries."""
    m, r = -(-(nv - 1) // item_points) + 1, 0
    while m > 16:
        m, r = -(-m // 4), r + 1
    return r


@pytest.mark.parametrize("group,n,c,item_points,witness", [
    (1, 1 << 16, 11, None, False),   # automatic items: a few merge rounds
    (1, 1 << 16, 11, 1, True),       # one point per item: seven rounds, the last two under "merge 6"
    (2, 3000, 10, 8, False),
    (1, 60, 9, None, False),         # no merge round
])
def test_profile_names_and_result(gpu, group, n, c, item_points, witness):
    from test_gpu_msm import affine, device_bases
    from zkgpu import Group, Msm

    rng = np.random.default_rng(1800 + n + c)
    ks = uniform_fr_words(rng, n)
    sc = witness_scalars(rng, n) if witness else uniform_fr_words(rng, n)
    bb = device_bases(gpu, group, ks)
    sb = gpu.buffer(sc)
    plan = Msm.for_group(gpu, Group(group), n, c, item_points)
    assert plan.group() == Group(group)
    prof = plan.profile(gpu, bb, sb)
    names = [name for name, _ in prof]
    # virtual points: 4n for the G2 GLS split (c <= 16, n <= 2^18), else 2n for the GLV split
    nv = 4 * n if group == 2 and c <= 16 and n <= 1 << 18 else 2 * n
    assert names == _names(_merge_rounds(nv, plan.item_points())), names
    assert all(ms >= 0 for _, ms in prof)
    if item_points == 1:
        assert _merge_rounds(nv, 1) == 7 and names.count("merge 6") == 1
    want = O.point_mul(group, O.generator_words(group), O.closed_form_scalar(ks, sc))
    profiled = affine(gpu, plan.result_buffer(gpu), group)
    assert np.array_equal(profiled, want)
    assert np.array_equal(affine(gpu, plan.run(gpu, bb, sb), group), profiled)
