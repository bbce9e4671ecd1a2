"""The benchmark configurations at their own sizes, checked exactly:

- a Groth16 proof's MSM set at n = 2^22 (G1 A, B1, C of n points and H of n - 1, G2 B2 of n) with
  witness-like scalars, run as one batch of autotuned plans (eagerly and as a captured graph
  replay), each MSM equal to the closed form (sum k_i s_i) G for bases k_i G;
- the proof-job server with resident keys of log_n 14, 18 and 22 and jobs arriving while others
  run: every NTT output equal to the oracle's, every MSM equal to the closed form, and every job's
  outputs equal to the same job run alone through the plain calls;
- a batch of NTTs that includes 2^21 and 2^22 transforms (three passes of the batch's tile) next
  to small ones, each equal to the oracle."""
import time

import numpy as np
import pytest

import oracle as O
from conftest import quick, uniform_fr_words, witness_scalars

pytestmark = pytest.mark.gpu


def _affine(gpu, res, group):
    from test_gpu_batch import _affine as aff

    return aff(gpu, res, group)


def _device_bases(gpu, group, ks):
    from test_gpu_msm import device_bases

    return device_bases(gpu, group, ks)


def _closed(group, ks, scalars):
    return O.point_mul(group, O.generator_words(group), O.closed_form_scalar(ks, scalars))


def test_groth16_msm_set_2_22_batched_with_tuned_plans(gpu):
    from zkgpu import Batch, Group, Msm, tune

    if quick():
        pytest.skip("ZK_GPU_QUICK")
    n = 1 << 22
    rng = np.random.default_rng(4022)
    sets = [(1, n), (1, n), (1, n), (1, n - 1), (2, n)]  # A, B1, C, H, B2
    ks = [uniform_fr_words(rng, m) for _, m in sets]
    bases = [_device_bases(gpu, g, k) for (g, _), k in zip(sets, ks)]
    wit = witness_scalars(rng, n)  # 20% 0, 20% 1, 30% below 2^16, 30% uniform
    hsc = uniform_fr_words(rng, n - 1)  # the quotient's coefficients
    scal = [wit, wit, wit, hsc, wit]
    sbufs = {id(wit): gpu.buffer(wit), id(hsc): gpu.buffer(hsc)}
    for g in (1, 2):
        tune.tune_msm(gpu, g, n)
    plans = [Msm.tuned(gpu, Group(g), m) for g, m in sets]
    batch = Batch(gpu)
    for p, b, sc in zip(plans, bases, scal):
        batch.add_msm(p, b, sbufs[id(sc)])
    want = [_closed(g, k, sc[: 8 * m]) for (g, m), k, sc in zip(sets, ks, scal)]
    batch.run()
    gpu.wait()
    for j, ((g, _), p) in enumerate(zip(sets, plans)):
        assert np.array_equal(_affine(gpu, p.result_buffer(gpu), g), want[j]), ("eager", j)
    batch.capture()
    for rep in range(2):
        batch.run()
        gpu.wait()
        for j, ((g, _), p) in enumerate(zip(sets, plans)):
            assert np.array_equal(_affine(gpu, p.result_buffer(gpu), g), want[j]), ("graph", rep, j)


class _Key:
    """A default-shape proving key with known multipliers: A, B1, C (n points), H (n - 1), B2 (n)."""

    def __init__(self, gpu, rng, log_n):
        self.log_n = log_n
        n = 1 << log_n
        counts = [n, n, n, n - 1]
        self.ks_g1 = [uniform_fr_words(rng, m) for m in counts]
        self.ks_g2 = uniform_fr_words(rng, n)
        self.g1 = [_device_bases(gpu, 1, k) for k in self.ks_g1]
        self.g2 = _device_bases(gpu, 2, self.ks_g2)


def test_server_keys_of_log_n_14_18_22_with_arrivals(gpu):
    from test_gpu_serve import _Job, _split
    from zkgpu import Server

    if quick():
        pytest.skip("ZK_GPU_QUICK")
    rng = np.random.default_rng(5022)
    keys = [_Key(gpu, rng, ln) for ln in (14, 18, 22)]
    server = Server(gpu, 64 << 30, max_inflight=3)
    kid = [server.add_key(k.log_n, k.g1, k.g2) for k in keys]
    order = [2, 0, 1, 0, 2, 1, 0]
    masks = [0b1010101, 0b0000000, 0b1111111, 0b0110011, 0b0000001, 0b1000000, 0b0101010]
    jobs = [_Job(gpu, rng, keys[k], m) for k, m in zip(order, masks)]
    gaps = rng.exponential(0.004, size=len(jobs))  # seeded Poisson arrivals, mean 4 ms apart
    ids = []
    for job, k, gap in zip(jobs, order, gaps):
        ids.append(server.submit(kid[k], job.bufs, job.inverse_mask, job.wbuf))
        time.sleep(float(gap))
    got = [server.wait(i) for i in ids]
    st = server.stats()
    assert st["completed"] == len(jobs)
    assert st["peak_inflight"] >= 2  # later arrivals were admitted while earlier jobs ran
    for j, (job, (words, _)) in enumerate(zip(jobs, got)):
        key, log_n, n = job.key, job.key.log_n, 1 << job.key.log_n
        served_ntt = [gpu.read(b, 8 * n) for b in job.bufs]
        outs = [O.ntt(p, log_n, inverse=bool((job.inverse_mask >> k) & 1)) for k, p in enumerate(job.polys)]
        for k in range(7):
            assert np.array_equal(served_ntt[k], outs[k]), (j, log_n, k)
        served = _split(words)
        exp = [_closed(1, key.ks_g1[i], job.witness) for i in range(3)]
        exp.append(_closed(1, key.ks_g1[3], outs[6][: 8 * (n - 1)]))
        exp.append(_closed(2, key.ks_g2, job.witness))
        for m in range(5):
            assert np.array_equal(served[m], exp[m]), (j, log_n, m)
        alone_ntt, alone = job.alone(gpu)
        for k in range(7):
            assert np.array_equal(alone_ntt[k], served_ntt[k]), (j, "alone", k)
        for m in range(5):
            assert np.array_equal(alone[m], served[m]), (j, "alone", m)
    server.close()


def test_batched_ntt_with_three_pass_transforms(gpu):
    """2^21 and 2^22 transforms in a batch whose largest job sets the tile: three passes each, next
    to one- and two-pass jobs, forward and inverse."""
    from zkgpu import Batch

    rng = np.random.default_rng(3022)
    sizes = [21, 12, 22, 16, 22, 13, 21, 19]
    inputs = [uniform_fr_words(rng, 1 << k) for k in sizes]
    bufs = [gpu.buffer(x) for x in inputs]
    batch = Batch(gpu)
    for j, (k, b) in enumerate(zip(sizes, bufs)):
        batch.add_ntt(b, k, inverse=bool(j % 2))
    expect = [O.ntt(x, k, inverse=bool(j % 2)) for j, (k, x) in enumerate(zip(sizes, inputs))]
    batch.run()
    for j, (k, b) in enumerate(zip(sizes, bufs)):
        assert np.array_equal(gpu.read(b, 8 << k), expect[j]), (j, k)
    # again as a captured graph on fresh copies of the inputs
    for b, x in zip(bufs, inputs):
        gpu.write(b, x)
    batch.capture()
    batch.run()
    for j, (k, b) in enumerate(zip(sizes, bufs)):
        assert np.array_equal(gpu.read(b, 8 << k), expect[j]), ("graph", j, k)
