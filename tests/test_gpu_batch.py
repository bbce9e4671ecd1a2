"""Batched execution: many NTTs of uneven size and direction plus several MSMs in one batch,
each result bit-identical to the oracle and to the same job run alone."""
import numpy as np
import pytest

import oracle as O
from conftest import uniform_fr_words, witness_scalars

pytestmark = pytest.mark.gpu


def _affine(gpu, res, group):
    import zkgpu
    from zkgpu import host

    if group == 1:
        a = zkgpu.xyzz_to_affine(gpu.read(res, 32))
        return np.zeros(16, np.uint32) if a is None else np.array(host.words32(a[0]) + host.words32(a[1]), np.uint32)
    a = zkgpu.xyzz_to_affine_g2(gpu.read(res, 64))
    if a is None:
        return np.zeros(32, np.uint32)
    return np.array(sum((host.words32(v) for pair in a for v in pair), []), np.uint32)


def test_batch_ntt_mixed_sizes_and_directions(gpu):
    from zkgpu import Batch

    rng = np.random.default_rng(500)
    sizes = [10, 16, 12, 11, 18, 13, 10, 15, 14, 17, 12, 20]
    inputs = [uniform_fr_words(rng, 1 << k) for k in sizes]
    bufs = [gpu.buffer(x) for x in inputs]
    batch = Batch(gpu)
    for j, (k, b) in enumerate(zip(sizes, bufs)):
        batch.add_ntt(b, k, inverse=bool(j % 2))
    batch.run()
    for j, (k, x, b) in enumerate(zip(sizes, inputs, bufs)):
        assert np.array_equal(gpu.read(b, x.size), O.ntt(x, k, inverse=bool(j % 2))), (j, k)
    # run the same batch again on fresh inputs
    for b, x in zip(bufs, inputs):
        gpu.write(b, x)
    batch.run()
    for j, (k, x, b) in enumerate(zip(sizes, inputs, bufs)):
        assert np.array_equal(gpu.read(b, x.size), O.ntt(x, k, inverse=bool(j % 2))), (j, k)


def test_batch_msms_and_ntts_match_single_calls(gpu):
    from zkgpu import Batch, Group, Msm, Ntt

    rng = np.random.default_rng(501)
    specs = [(1, 5000, 10, False), (1, 3000, 12, True), (2, 700, 11, False), (1, 20000, 13, True), (2, 2000, 9, True)]
    jobs = []
    for group, n, c, witness in specs:
        ks = uniform_fr_words(rng, n)
        bases = O.multiples(group, O.generator_words(group), ks)
        sc = witness_scalars(rng, n) if witness else uniform_fr_words(rng, n)
        plan = Msm.for_group(gpu, Group(group), n, c)
        jobs.append((group, plan, Msm.upload_bases(gpu, bases), gpu.buffer(sc), O.msm(group, bases, sc)))
    ntt_in = [uniform_fr_words(rng, 1 << k) for k in (12, 14)]
    ntt_bufs = [gpu.buffer(x) for x in ntt_in]
    batch = Batch(gpu)
    for group, plan, bb, sb, _ in jobs:
        batch.add_msm(plan, bb, sb)
    batch.add_ntt(ntt_bufs[0], 12, inverse=False)
    batch.add_ntt(ntt_bufs[1], 14, inverse=True)
    for rep in range(2):
        batch.run()
        for group, plan, bb, sb, exp in jobs:
            res = _affine(gpu, plan.result_buffer(gpu), group)
            assert np.array_equal(res, exp), (group, rep)
        assert np.array_equal(gpu.read(ntt_bufs[0], ntt_in[0].size), O.ntt(ntt_in[0], 12))
        assert np.array_equal(gpu.read(ntt_bufs[1], ntt_in[1].size), O.ntt(ntt_in[1], 14, inverse=True))
        for b, x in zip(ntt_bufs, ntt_in):
            gpu.write(b, x)
    # each MSM alone gives the same result as in the batch
    for group, plan, bb, sb, exp in jobs:
        assert np.array_equal(_affine(gpu, plan.run(gpu, bb, sb), group), exp)
    single = Ntt(gpu, 12)
    b = gpu.buffer(ntt_in[0])
    single.forward(gpu, b)
    assert np.array_equal(gpu.read(b, ntt_in[0].size), O.ntt(ntt_in[0], 12))


def test_batch_rejects_a_plan_twice(gpu):
    from zkgpu import Batch, Msm, ZkgError

    n = 100
    plan = Msm.new(gpu, n, 10)
    bb = Msm.upload_bases(gpu, np.zeros(16 * n, np.uint32))
    sb = gpu.buffer(np.zeros(8 * n, np.uint32))
    batch = Batch(gpu)
    batch.add_msm(plan, bb, sb)
    with pytest.raises(ZkgError, match="only once in a batch"):
        batch.add_msm(plan, bb, sb)


def test_batch_graph_replays_match_eager_runs(gpu):
    """A captured batch replayed several times, with eager work and synchronisation between the
    replays (fresh inputs written, unrelated NTTs and MSMs run), gives the eager result each time."""
    from zkgpu import Batch, Group, Msm, Ntt

    rng = np.random.default_rng(520)
    sizes = [10, 13, 11, 15, 12]
    inputs = [uniform_fr_words(rng, 1 << k) for k in sizes]
    bufs = [gpu.buffer(x) for x in inputs]
    specs = [(1, 3000, 11), (2, 900, 10), (1, 5000, 13)]
    jobs = []
    for group, n, c in specs:
        ks = uniform_fr_words(rng, n)
        bases = O.multiples(group, O.generator_words(group), ks)
        sc = witness_scalars(rng, n)
        jobs.append((group, Msm.for_group(gpu, Group(group), n, c), Msm.upload_bases(gpu, bases), gpu.buffer(sc),
                     O.msm(group, bases, sc)))
    batch = Batch(gpu)
    for j, (k, b) in enumerate(zip(sizes, bufs)):
        batch.add_ntt(b, k, inverse=bool(j % 2))
    for group, plan, bb, sb, _ in jobs:
        batch.add_msm(plan, bb, sb)
    expect_ntt = [O.ntt(x, k, inverse=bool(j % 2)) for j, (k, x) in enumerate(zip(sizes, inputs))]
    batch.run()  # eager
    gpu.wait()
    for j, (k, b) in enumerate(zip(sizes, bufs)):
        assert np.array_equal(gpu.read(b, 8 << k), expect_ntt[j]), ("eager", j)
    batch.capture()
    assert batch.captured()
    side = gpu.buffer(uniform_fr_words(rng, 1 << 12))
    side_plan = Ntt(gpu, 12)
    other = Msm.new(gpu, 2000, 12)
    other_bases = Msm.upload_bases(gpu, O.multiples(1, O.generator_words(1), uniform_fr_words(rng, 2000)))
    other_sc = gpu.buffer(uniform_fr_words(rng, 2000))
    for rep in range(3):
        for b, x in zip(bufs, inputs):
            gpu.write(b, x)
        side_plan.forward(gpu, side)
        other.run(gpu, other_bases, other_sc)
        gpu.wait()
        batch.run()  # graph replay
        side_plan.inverse(gpu, side)
        gpu.wait()
        for j, (k, b) in enumerate(zip(sizes, bufs)):
            assert np.array_equal(gpu.read(b, 8 << k), expect_ntt[j]), (rep, j)
        for group, plan, bb, sb, exp in jobs:
            assert np.array_equal(_affine(gpu, plan.result_buffer(gpu), group), exp), (rep, group)
    # changing the batch drops the graph; the next run is eager again
    extra = gpu.buffer(inputs[0])
    batch.add_ntt(extra, sizes[0])
    assert not batch.captured()
    for b, x in zip(bufs, inputs):
        gpu.write(b, x)
    batch.run()
    assert np.array_equal(gpu.read(extra, 8 << sizes[0]), O.ntt(inputs[0], sizes[0]))
