"""snarkjs .zkey reader (host only, no GPU): header fields, verification key, section sizes, and
the errors for files that are not BN254 Groth16 keys of the expected layout."""
import numpy as np
import pytest

import oracle as O
import zkey_file


@pytest.fixture(scope="module")
def key():
    rng = np.random.default_rng(1300)
    return zkey_file.make_key(rng, n_vars=40, n_public=3, domain_size=32, infinity=(2,))


def _open(path):
    import zkgpu

    return zkgpu.Zkey(path)


def test_reader_reports_header_sections_and_vk(tmp_path, key):
    points, _, vk = key
    path = tmp_path / "k.zkey"
    zkey_file.write(path, 40, 3, 32, points, vk)
    z = _open(path)
    assert (z.n_vars, z.n_public, z.domain_size) == (40, 3, 32)
    from zkgpu import zkey

    assert [z.points(s) for s in (zkey.IC, zkey.A, zkey.B1, zkey.B2, zkey.C, zkey.H)] == [4, 40, 40, 40, 36, 32]
    for name in ("alpha1", "beta1", "beta2", "gamma2", "delta1", "delta2"):
        assert np.array_equal(z.vk[name], vk[name]), name  # Montgomery on disk, plain here
    # section order in the file does not matter
    path2 = tmp_path / "k2.zkey"
    zkey_file.write(path2, 40, 3, 32, points, vk, order=(2, 1, 9, 8, 7, 6, 5, 4, 3, 10))
    z2 = _open(path2)
    assert (z2.n_vars, z2.domain_size) == (40, 32) and z2.points(zkey.H) == 32
    z.close()
    z2.close()


def test_infinity_is_stored_as_zeros(key):
    points, _, _ = key
    pt = points[5].reshape(-1, 16)[2]
    assert not pt.any()
    assert zkey_file.mont_bytes(pt) == bytes(64)


@pytest.mark.parametrize("case,match", [
    ("magic", "is not a .zkey file"),
    ("version", "unsupported .zkey version 2"),
    ("protocol", "only Groth16 .zkey files are supported"),
    ("curve", "not a BN254 key"),
    ("short_a", "section 5 holds 2496 bytes; expected 2560"),
    ("missing_h", "has no section 9"),
    ("domain", "domain size is not a power of two"),
    ("truncated", "truncated"),
])
def test_reader_rejects_malformed_files(tmp_path, key, case, match):
    import zkgpu

    points, _, vk = key
    path = tmp_path / f"{case}.zkey"
    kw = {}
    if case == "magic":
        kw["magic"] = b"wtns"
    elif case == "version":
        kw["version"] = 2
    elif case == "protocol":
        kw["protocol"] = 2
    elif case == "curve":
        kw["q"] = O.P + 2
    elif case == "short_a":
        kw["override"] = {5: zkey_file.mont_bytes(points[5])[:-64]}
    elif case == "missing_h":
        kw["drop"] = (9,)
    domain = 24 if case == "domain" else 32
    if case == "domain":
        pts = dict(points)
        pts[9] = points[9][: 16 * 24]
        zkey_file.write(path, 40, 3, domain, pts, vk)
    else:
        zkey_file.write(path, 40, 3, domain, points, vk, **kw)
    if case == "truncated":
        data = path.read_bytes()
        path.write_bytes(data[:-100])
    with pytest.raises(zkgpu.ZkgError, match=match):
        zkgpu.Zkey(path)


def test_missing_file_fails_loudly(tmp_path):
    import zkgpu

    with pytest.raises(zkgpu.ZkgError, match="cannot open"):
        zkgpu.Zkey(tmp_path / "absent.zkey")
