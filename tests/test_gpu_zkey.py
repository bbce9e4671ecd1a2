"""snarkjs .zkey ingestion on the GPU: point sections uploaded as stored (Montgomery, R = 2^256)
give the oracle's MSMs over the plain points, and a server key built from a .zkey (nVars !=
domainSize, C over witness[nPublic + 1:], H over the seventh transform's outputs) serves jobs
whose results equal the oracle's, next to a default-shape key."""
import numpy as np
import pytest

import oracle as O
import zkey_file
from conftest import uniform_fr_words, witness_scalars

pytestmark = pytest.mark.gpu

N_VARS, N_PUBLIC, LOG_N = 1500, 5, 10


@pytest.fixture(scope="module")
def zkey_path(tmp_path_factory):
    rng = np.random.default_rng(1310)
    points, _, vk = zkey_file.make_key(rng, N_VARS, N_PUBLIC, 1 << LOG_N, infinity=(0, 7, 700))
    path = tmp_path_factory.mktemp("zkey") / "circuit.zkey"
    zkey_file.write(path, N_VARS, N_PUBLIC, 1 << LOG_N, points, vk)
    return path, points


def test_sections_upload_as_msm_bases(gpu, zkey_path):
    from test_gpu_serve import _affine
    from zkgpu import Group, Msm, Zkey, zkey

    path, points = zkey_path
    z = Zkey(path)
    rng = np.random.default_rng(1311)
    for sid in (zkey.A, zkey.B2, zkey.C, zkey.H, zkey.IC):
        group = 2 if sid == zkey.B2 else 1
        n = z.points(sid)
        sc = uniform_fr_words(rng, n)
        bases = z.upload(gpu, sid)
        m = Msm.for_group(gpu, Group(group), n, Msm.default_window(n) if n >= 2 ** 11 else 11)
        got = _affine(gpu.read(m.run(gpu, bases, gpu.buffer(sc)), 32 * group), group)
        assert np.array_equal(got, O.msm(group, points[sid], sc)), sid
    z.close()


def test_server_serves_a_zkey_key_next_to_a_default_key(gpu, zkey_path):
    from test_gpu_serve import _Key, _split
    from zkgpu import Server, Zkey, ZkgError

    path, points = zkey_path
    z = Zkey(path)
    rng = np.random.default_rng(1312)
    server = Server(gpu, 8 << 30, max_inflight=3)
    zid = server.add_zkey(z)
    shape = server.key_shape(zid)
    assert shape == {"log_n": LOG_N, "points": (N_VARS, N_VARS, N_VARS - N_PUBLIC - 1, 1 << LOG_N, N_VARS),
                     "c_offset": N_PUBLIC + 1, "h_form": 1}  # snarkjs's H: the odd-coset form
    assert server.key_slot_bytes(zid) > server.slot_bytes(LOG_N)  # nVars > 2^log_n
    z.close()  # the server holds its own copy of the bases
    plain = _Key(gpu, rng, LOG_N)
    pid = server.add_key(LOG_N, plain.g1, plain.g2)
    n = 1 << LOG_N
    jobs = []
    for j in range(4):
        polys = [uniform_fr_words(rng, n) for _ in range(7)]
        mask = (0b1010101, 0b0110011, 0, 0b1111111)[j]
        zk = j % 2 == 0
        wit = witness_scalars(rng, N_VARS if zk else n)
        bufs = [gpu.buffer(p) for p in polys]
        wbuf = gpu.buffer(wit)
        jid = server.submit(zid if zk else pid, bufs, mask, wbuf)
        jobs.append((jid, zk, polys, mask, wit, bufs, wbuf))
    with pytest.raises(ZkgError, match="scalars buffer holds"):
        server.submit(zid, jobs[0][5], 0, gpu.buffer(uniform_fr_words(rng, N_VARS - 1)))
    for jid, zk, polys, mask, wit, bufs, _ in jobs:
        words, _ = server.wait(jid)
        outs = [O.ntt(p, LOG_N, inverse=bool((mask >> k) & 1)) for k, p in enumerate(polys)]
        for k in range(7):
            assert np.array_equal(gpu.read(bufs[k], 8 * n), outs[k]), (jid, k)
        got = _split(words)
        if zk:
            exp = [O.msm(1, points[5], wit), O.msm(1, points[6], wit), O.msm(1, points[8], wit[8 * (N_PUBLIC + 1):]),
                   O.msm(1, points[9], outs[6]), O.msm(2, points[7], wit)]
        else:
            exp = [O.msm(1, plain.host_g1[i], wit) for i in range(3)]
            exp += [O.msm(1, plain.host_g1[3], outs[6][: 8 * (n - 1)]), O.msm(2, plain.host_g2, wit)]
        for m in range(5):
            assert np.array_equal(got[m], exp[m]), (jid, m)
    server.close()


def test_server_rejects_a_zkey_domain_below_the_ntt_range(gpu, tmp_path):
    from zkgpu import Server, Zkey, ZkgError

    rng = np.random.default_rng(1313)
    points, _, vk = zkey_file.make_key(rng, 20, 2, 16)
    path = tmp_path / "tiny.zkey"
    zkey_file.write(path, 20, 2, 16, points, vk)
    z = Zkey(path)
    server = Server(gpu, 1 << 30, max_inflight=1)
    with pytest.raises(ZkgError, match="domain is 2\\^4"):
        server.add_zkey(z)
    server.close()
    z.close()
