"""The autotuner's table (no device needed): text-file load / save / lookup / clear through the C
ABI, and rejection of malformed files."""
import pytest


@pytest.fixture
def tune():
    from zkgpu import tune as T

    T.clear()
    yield T
    T.clear()


def test_load_lookup_save_roundtrip(tune, tmp_path):
    src = tmp_path / "in.txt"
    src.write_text("# comment\n"
                   "msm gfx950:sramecc+:xnack- 1 20 14 64 32768 2.5\n"
                   "msm gfx950:sramecc+:xnack- 2 18 13 0 0 4.75\n"
                   "ntt gfx950:sramecc+:xnack- 24 12 1.9\n")
    tune.load(src)
    arch = "gfx950:sramecc+:xnack-"
    assert tune.lookup_msm(arch, 1, 1 << 20) == {"c": 14, "item_points": 64, "sort_points": 32768}
    # sizes share the key ceil(log2 n): 2^20 - 1 and 2^19 + 1 use the 2^20 entry
    assert tune.lookup_msm(arch, 1, (1 << 20) - 1)["c"] == 14
    assert tune.lookup_msm(arch, 1, (1 << 19) + 1)["c"] == 14
    assert tune.lookup_msm(arch, 1, 1 << 19) is None
    assert tune.lookup_msm(arch, 2, 1 << 18)["c"] == 13
    assert tune.lookup_msm("gfx942", 1, 1 << 20) is None
    assert tune.lookup_ntt(arch, 24) == {"tile_log": 12}
    assert tune.lookup_ntt(arch, 22) is None
    out = tmp_path / "out.txt"
    tune.save(out)
    tune.clear()
    assert tune.lookup_msm(arch, 1, 1 << 20) is None
    tune.load(out)
    assert tune.lookup_msm(arch, 1, 1 << 20) == {"c": 14, "item_points": 64, "sort_points": 32768}
    assert tune.lookup_ntt(arch, 24) == {"tile_log": 12}


@pytest.mark.parametrize("line", ["msm gfx950 3 20 14 0 0 1.0", "msm gfx950 1 20 8 0 0 1.0", "ntt gfx950 24 9 1.0",
                                  "fft gfx950 24 11 1.0", "msm gfx950 1 20"])
def test_malformed_files_are_rejected(tune, tmp_path, line):
    from zkgpu import ZkgError

    p = tmp_path / "bad.txt"
    p.write_text(line + "\n")
    with pytest.raises(ZkgError):
        tune.load(p)


def test_missing_file_is_an_error(tune, tmp_path):
    from zkgpu import ZkgError

    with pytest.raises(ZkgError, match="cannot read tuning file"):
        tune.load(tmp_path / "nope.txt")
