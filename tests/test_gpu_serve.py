"""Proof-job server: jobs of mixed proving keys submitted together (continuous batching) give
the same NTT outputs and MSM results as the same job run alone through the plain API, and the
oracle's; admission respects the HBM budget and the in-flight limit."""
import numpy as np
import pytest

import oracle as O
from conftest import uniform_fr_words, witness_scalars

pytestmark = pytest.mark.gpu


def _affine(words, group):
    import zkgpu
    from zkgpu import host

    if group == 1:
        a = zkgpu.xyzz_to_affine(words)
        return np.zeros(16, np.uint32) if a is None else np.array(host.words32(a[0]) + host.words32(a[1]), np.uint32)
    a = zkgpu.xyzz_to_affine_g2(words)
    if a is None:
        return np.zeros(32, np.uint32)
    return np.array(sum((host.words32(v) for pair in a for v in pair), []), np.uint32)


def _split(results):
    """Server result words -> affine A, B1, C, H (G1) and B2 (G2)."""
    out = [_affine(results[32 * i: 32 * i + 32], 1) for i in range(4)]
    out.append(_affine(results[128:192], 2))
    return out


class _Key:
    def __init__(self, gpu, rng, log_n):
        from zkgpu import Msm

        self.log_n = log_n
        n = 1 << log_n
        self.host_g1 = []
        for i in range(4):
            pts = n - 1 if i == 3 else n
            self.host_g1.append(O.multiples(1, O.generator_words(1), uniform_fr_words(rng, pts)))
        self.host_g2 = O.multiples(2, O.generator_words(2), uniform_fr_words(rng, n))
        self.g1 = [Msm.upload_bases(gpu, b) for b in self.host_g1]
        self.g2 = Msm.upload_bases(gpu, self.host_g2)


class _Job:
    def __init__(self, gpu, rng, key, inverse_mask):
        n = 1 << key.log_n
        self.key = key
        self.polys = [uniform_fr_words(rng, n) for _ in range(7)]
        self.witness = witness_scalars(rng, n)
        self.inverse_mask = inverse_mask
        self.bufs = [gpu.buffer(p) for p in self.polys]
        self.wbuf = gpu.buffer(self.witness)

    def reset(self, gpu):
        for b, p in zip(self.bufs, self.polys):
            gpu.write(b, p)

    def alone(self, gpu):
        """The same job through the plain Ntt / Msm calls, one at a time: (ntt outputs, affine results)."""
        from zkgpu import Group, Msm, Ntt

        log_n, n = self.key.log_n, 1 << self.key.log_n
        ntt = Ntt(gpu, log_n)
        bufs = [gpu.buffer(p) for p in self.polys]
        for k, b in enumerate(bufs):
            (ntt.inverse if (self.inverse_mask >> k) & 1 else ntt.forward)(gpu, b)
        outs = [gpu.read(b, 8 * n) for b in bufs]
        c = Msm.default_window(n)
        res = []
        for i in range(3):
            plan = Msm.for_group(gpu, Group(1), n, c)
            res.append(_affine(gpu.read(plan.run(gpu, self.key.g1[i], self.wbuf), 32), 1))
        plan = Msm.for_group(gpu, Group(1), n - 1, c)
        res.append(_affine(gpu.read(plan.run(gpu, self.key.g1[3], bufs[6]), 32), 1))
        plan = Msm.for_group(gpu, Group(2), n, c)
        res.append(_affine(gpu.read(plan.run(gpu, self.key.g2, self.wbuf), 64), 2))
        return outs, res


@pytest.fixture(scope="module")
def keys(gpu):
    rng = np.random.default_rng(700)
    return [_Key(gpu, rng, 10), _Key(gpu, rng, 12)]


def test_server_mixed_jobs_match_alone_and_oracle(gpu, keys):
    from zkgpu import Server

    rng = np.random.default_rng(701)
    server = Server(gpu, 8 << 30, max_inflight=3)
    kid = [server.add_key(k.log_n, k.g1, k.g2) for k in keys]
    masks = [0b0000000, 0b1111111, 0b0101010, 0b1010101, 0b0000001, 0b1000000]
    jobs = [_Job(gpu, rng, keys[j % 2], masks[j]) for j in range(6)]
    ids = [server.submit(kid[j % 2], job.bufs, job.inverse_mask, job.wbuf) for j, job in enumerate(jobs)]
    got = [server.wait(i) for i in ids]
    st = server.stats()
    assert st["completed"] == len(jobs) and st["submitted"] == len(jobs)
    assert 1 <= st["peak_inflight"] <= 3
    assert st["p50_ms"] > 0 and st["p99_ms"] >= st["p50_ms"]
    for j, (job, (words, lat)) in enumerate(zip(jobs, got)):
        n = 1 << job.key.log_n
        served_ntt = [gpu.read(b, 8 * n) for b in job.bufs]
        served = _split(words)
        alone_ntt, alone = job.alone(gpu)
        for k in range(7):
            assert np.array_equal(served_ntt[k], alone_ntt[k]), (j, k)
        for m in range(5):
            assert np.array_equal(served[m], alone[m]), (j, m)
        assert lat > 0
    # one job of each key against the oracle
    for job, (words, _) in zip(jobs[:2], got[:2]):
        log_n, n = job.key.log_n, 1 << job.key.log_n
        outs = [O.ntt(p, log_n, inverse=bool((job.inverse_mask >> k) & 1)) for k, p in enumerate(job.polys)]
        for k in range(7):
            assert np.array_equal(gpu.read(job.bufs[k], 8 * n), outs[k]), k
        served = _split(words)
        exp = [O.msm(1, job.key.host_g1[i], job.witness) for i in range(3)]
        exp.append(O.msm(1, job.key.host_g1[3], outs[6][: 8 * (n - 1)]))
        exp.append(O.msm(2, job.key.host_g2, job.witness))
        for m in range(5):
            assert np.array_equal(served[m], exp[m]), m
    server.close()


def test_server_budget_serialises_jobs(gpu, keys):
    """A budget with room for one job's workspace admits one job at a time; results unchanged."""
    from zkgpu import Server

    rng = np.random.default_rng(702)
    key = keys[1]
    n = 1 << key.log_n
    probe = Server(gpu, 1 << 40, max_inflight=1)
    slot = probe.slot_bytes(key.log_n)
    probe.close()
    key_bytes = sum(b.size() for b in key.g1) + key.g2.size()
    server = Server(gpu, key_bytes + slot + slot // 2, max_inflight=8)
    kid = server.add_key(key.log_n, key.g1, key.g2)
    jobs = [_Job(gpu, rng, key, m) for m in (0b0011001, 0b1100110, 0)]
    ids = [server.submit(kid, job.bufs, job.inverse_mask, job.wbuf) for job in jobs]
    got = [server.wait(i) for i in ids]
    st = server.stats()
    assert st["peak_inflight"] == 1
    assert st["peak_bytes"] <= key_bytes + slot + slot // 2
    for job, (words, _) in zip(jobs, got):
        served_ntt = [gpu.read(b, 8 * n) for b in job.bufs]
        alone_ntt, alone = job.alone(gpu)
        assert all(np.array_equal(a, b) for a, b in zip(served_ntt, alone_ntt))
        served = _split(words)
        assert all(np.array_equal(a, b) for a, b in zip(served, alone))
    server.close()


def test_server_rejects_what_cannot_fit(gpu, keys):
    from zkgpu import Server, ZkgError

    key = keys[0]
    key_bytes = sum(b.size() for b in key.g1) + key.g2.size()
    small = Server(gpu, key_bytes - 1, max_inflight=1)
    with pytest.raises(ZkgError, match="budget"):
        small.add_key(key.log_n, key.g1, key.g2)
    small.close()
    tight = Server(gpu, key_bytes + 1024, max_inflight=1)
    kid = tight.add_key(key.log_n, key.g1, key.g2)
    rng = np.random.default_rng(703)
    job = _Job(gpu, rng, key, 0)
    with pytest.raises(ZkgError, match="workspace"):
        tight.submit(kid, job.bufs, 0, job.wbuf)
    with pytest.raises(ZkgError, match="unknown proving key"):
        tight.submit(kid + 5, job.bufs, 0, job.wbuf)
    tight.close()
