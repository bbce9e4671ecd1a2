"""Quotient jobs on the proof-job server: from A, B, C's domain evaluations and a witness, the
server computes the H scalars in the key's form (coefficients for default keys, the odd coset
for snarkjs keys), runs the H MSM on them and the other four MSMs on the witness. Results equal
the oracle's quotient and MSMs, with 7-polynomial jobs of the same keys in the mix."""
import numpy as np
import pytest

import oracle as O
import zkey_file
from conftest import uniform_fr_words, witness_scalars

pytestmark = pytest.mark.gpu


def test_quotient_jobs_for_default_and_snarkjs_keys(gpu, tmp_path):
    from test_gpu_serve import _Key, _split
    from zkgpu import Server, Zkey, ZkgError

    rng = np.random.default_rng(1500)
    log_n, n_vars, n_public = 10, 1200, 3
    n = 1 << log_n
    points, _, vk = zkey_file.make_key(rng, n_vars, n_public, n)
    path = tmp_path / "q.zkey"
    zkey_file.write(path, n_vars, n_public, n, points, vk)
    z = Zkey(path)
    plain = _Key(gpu, rng, log_n)
    server = Server(gpu, 8 << 30, max_inflight=3)
    zid = server.add_zkey(z)
    pid = server.add_key(log_n, plain.g1, plain.g2)
    z.close()
    assert server.key_shape(pid)["h_form"] == 0 and server.key_shape(zid)["h_form"] == 1
    jobs = []
    for j in range(6):
        zk = j % 2 == 0
        wit = witness_scalars(rng, n_vars if zk else n)
        wbuf = gpu.buffer(wit)
        if j < 4:
            abc = [uniform_fr_words(rng, n) for _ in range(3)]
            if j == 1:  # a satisfied system: c = a b on the domain
                abc[2] = O.ints_to_words([x * y % O.R for x, y in zip(O.words_to_ints(abc[0]), O.words_to_ints(abc[1]))])
            bufs = [gpu.buffer(x) for x in abc]
            jid = server.submit_groth16(zid if zk else pid, *bufs, wbuf)
            jobs.append(("q", jid, zk, abc, bufs, wit))
        else:  # a 7-polynomial job in the same mix
            polys = [uniform_fr_words(rng, n) for _ in range(7)]
            bufs = [gpu.buffer(p) for p in polys]
            jid = server.submit(zid if zk else pid, bufs, 0b0000001, wbuf)
            jobs.append(("p", jid, zk, polys, bufs, wit))
    with pytest.raises(ZkgError, match="distinct"):
        b0 = gpu.buffer(uniform_fr_words(rng, n))
        server.submit_groth16(pid, b0, b0, gpu.buffer(uniform_fr_words(rng, n)), gpu.buffer(witness_scalars(rng, n)))
    for kind, jid, zk, inputs, bufs, wit in jobs:
        words, _ = server.wait(jid)
        if kind == "q":
            hs = O.groth16_h_odd_coset(*inputs, log_n) if zk else O.groth16_h(*inputs, log_n)
            assert np.array_equal(gpu.read(bufs[0], 8 * n), hs), jid
        else:
            hs = O.ntt(inputs[6], log_n)
        got = _split(words)
        if zk:
            exp = [O.msm(1, points[5], wit), O.msm(1, points[6], wit), O.msm(1, points[8], wit[8 * (n_public + 1):]),
                   O.msm(1, points[9], hs), O.msm(2, points[7], wit)]
        else:
            exp = [O.msm(1, plain.host_g1[i], wit) for i in range(3)]
            exp += [O.msm(1, plain.host_g1[3], hs[: 8 * (n - 1)]), O.msm(2, plain.host_g2, wit)]
        for m in range(5):
            assert np.array_equal(got[m], exp[m]), (kind, jid, m)
    server.close()
