"""Autotuning on the device: the searches return valid choices, record them, and tuned MSM / NTT
plans (and the server, which reads the same table) give the oracle's results."""
import numpy as np
import pytest

import oracle as O
from conftest import uniform_fr_words, witness_scalars

pytestmark = pytest.mark.gpu


def test_tune_msm_and_tuned_plans(gpu):
    from test_gpu_msm import affine
    from zkgpu import Group, Msm, tune

    arch = gpu.info()["arch"]
    tune.clear()
    for group, n in ((1, 1 << 12), (2, 1 << 10)):
        t = tune.tune_msm(gpu, group, n)
        assert 9 <= t["c"] <= 16 and t["ms"] > 0
        assert tune.lookup_msm(arch, group, n) == {k: t[k] for k in ("c", "item_points", "sort_points")}
        rng = np.random.default_rng(1200 + group)
        ks = uniform_fr_words(rng, n)
        sc = witness_scalars(rng, n)
        bases = O.multiples(group, O.generator_words(group), ks)
        exp = O.msm(group, bases, sc)
        plan = Msm.tuned(gpu, Group(group), n)
        assert plan.window() == t["c"]
        if t["item_points"]:
            assert plan.item_points() == t["item_points"]
        got = affine(gpu, plan.run(gpu, Msm.upload_bases(gpu, bases), gpu.buffer(sc)), group)
        assert np.array_equal(got, exp)
    tune.clear()


def test_tune_ntt_and_tuned_plan(gpu):
    from zkgpu import Ntt, tune

    arch = gpu.info()["arch"]
    tune.clear()
    t = tune.tune_ntt(gpu, 14)
    assert t["tile_log"] in (10, 11, 12) and t["ms"] > 0
    assert tune.lookup_ntt(arch, 14) == {"tile_log": t["tile_log"]}
    x = uniform_fr_words(np.random.default_rng(1210), 1 << 14)
    buf = gpu.buffer(x)
    plan = Ntt(gpu, 14)
    plan.forward(gpu, buf)
    assert np.array_equal(gpu.read(buf, x.size), O.ntt(x, 14))
    plan.inverse(gpu, buf)
    assert np.array_equal(gpu.read(buf, x.size), x)
    tune.clear()


@pytest.mark.parametrize("tile", [10, 11, 12])
def test_every_ntt_tile_matches_oracle(gpu, tmp_path, tile):
    """Each tile the tuner may pick, forced through the table, at sizes with 1 to 3 passes."""
    from zkgpu import Ntt, tune

    arch = gpu.info()["arch"]
    tune.clear()
    sizes = (10, 12, 13, 17, 20)
    p = tmp_path / "t.txt"
    p.write_text("".join(f"ntt {arch} {k} {tile} 1.0\n" for k in sizes))
    tune.load(p)
    rng = np.random.default_rng(1220 + tile)
    for k in sizes:
        x = uniform_fr_words(rng, 1 << k)
        buf = gpu.buffer(x)
        plan = Ntt(gpu, k)
        plan.inverse(gpu, buf)
        assert np.array_equal(gpu.read(buf, x.size), O.ntt(x, k, inverse=True)), (tile, k)
    tune.clear()
