"""The reference's asynchronous and yielding forms (src/gpu.rs:22-40,157-190; src/msm.rs:300-318):
same results as the blocking calls; failed reads report instead of aborting."""
import asyncio

import numpy as np
import pytest

import oracle as O
from conftest import uniform_fr_words

pytestmark = pytest.mark.gpu


def test_run_yielding_equals_run(gpu):
    from test_gpu_msm import affine
    from zkgpu import Group, Msm

    rng = np.random.default_rng(1400)
    for group, n, c in ((1, 6000, 12), (2, 1500, 11)):
        bases = O.multiples(group, O.generator_words(group), uniform_fr_words(rng, n))
        sc = uniform_fr_words(rng, n)
        bb, sb = Msm.upload_bases(gpu, bases), gpu.buffer(sc)
        plan = Msm.for_group(gpu, Group(group), n, c)
        got = affine(gpu, asyncio.run(plan.run_yielding(gpu, bb, sb)), group)
        assert np.array_equal(got, O.msm(group, bases, sc))
        assert np.array_equal(affine(gpu, plan.run(gpu, bb, sb), group), got)


def test_read_async_and_try_read_async(gpu):
    from zkgpu import Ntt, ZkgError

    x = uniform_fr_words(np.random.default_rng(1401), 1 << 12)
    buf = gpu.buffer(x)
    Ntt(gpu, 12).forward(gpu, buf)

    async def go():
        await gpu.finished_async()
        words = await gpu.read_async(buf, x.size)
        ok, msg = await gpu.try_read_async(buf, x.size + 1)
        return words, ok, msg

    words, ok, msg = asyncio.run(go())
    assert np.array_equal(words, O.ntt(x, 12))
    assert not ok and "bytes from a buffer of" in msg
    with pytest.raises(ZkgError):
        asyncio.run(gpu.read_async(buf, x.size + 1))


def test_new_async_opens_the_device():
    from zkgpu import Gpu

    g = asyncio.run(Gpu.new_async())
    assert g.info()["cu_count"] > 0
