"""The C-ABI library: it loads, exports every entry point include/zkgpu.h declares, and its
host-only functions behave as the reference's. No GPU compute is called here."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "zkgpu.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(zkg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_drop_in_surface():
    names = declared_symbols()
    for must in ("zkg_device_open", "zkg_buffer_create", "zkg_buffer_empty", "zkg_buffer_read", "zkg_wait",
                 "zkg_msm_default_window", "zkg_msm_create", "zkg_msm_upload_bases",
                 "zkg_msm_upload_bases_montgomery_256", "zkg_msm_run", "zkg_msm_group", "zkg_xyzz_to_affine",
                 "zkg_xyzz_to_affine_g2", "zkg_ntt_create", "zkg_ntt_forward", "zkg_ntt_inverse", "zkg_ntt_log_n"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from zkgpu import _lib

    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    # the Python binding declares exactly the header's symbols
    assert sorted(_lib.SIGNATURES) == declared_symbols()


def test_default_window_matches_reference_rule():
    from zkgpu import Msm

    def ref(n):  # src/msm.rs:76-80: (n.max(1).ilog2().saturating_sub(5)).clamp(11, 16)
        return min(max(max(max(n, 1).bit_length() - 1 - 5, 0), 11), 16)

    for n in [0, 1, 2, 1000, 2 ** 16, 2 ** 17, 2 ** 18, 2 ** 20, 2 ** 22 - 1, 2 ** 22, 2 ** 26]:
        assert Msm.default_window(n) == ref(n), n
    assert [Msm.default_window(1 << k) for k in (16, 18, 20, 22)] == [11, 13, 15, 16]


def test_xyzz_to_affine_host_conversion():
    """xyzz_to_affine decodes the library's Montgomery form (R = 2^256) on the host."""
    import oracle as O
    import zkgpu

    R256 = 1 << 256
    x, y = O.py_g1_mul(O.G1, 123456789)
    t = 987654321  # any (zz, zzz) with zz^3 = zzz^2: zz = t^2, zzz = t^3
    zz, zzz = t * t % O.P, t * t * t % O.P
    X, Y = x * zz % O.P, y * zzz % O.P
    words = np.concatenate([O.int_to_words(v * R256 % O.P) for v in (X, Y, zz, zzz)])
    got = zkgpu.xyzz_to_affine(words)
    assert got is not None
    assert zkgpu.host.to_int(got[0]) == x and zkgpu.host.to_int(got[1]) == y
    inf = np.concatenate([O.int_to_words(v) for v in (5, 7, 0, 0)])
    assert zkgpu.xyzz_to_affine(inf) is None


def test_xyzz_to_affine_g2_host_conversion():
    import oracle as O
    import zkgpu

    R256 = 1 << 256
    (xa, xb), (ya, yb) = O.py_g2_mul(O.G2, 55555)
    t = (31337, 4242)
    zz = O.py_f2_mul(t, t)
    zzz = O.py_f2_mul(zz, t)
    X, Y = O.py_f2_mul((xa, xb), zz), O.py_f2_mul((ya, yb), zzz)
    vals = [X[0], X[1], Y[0], Y[1], zz[0], zz[1], zzz[0], zzz[1]]
    words = np.concatenate([O.int_to_words(v * R256 % O.P) for v in vals])
    got = zkgpu.xyzz_to_affine_g2(words)
    assert got is not None
    (gxa, gxb), (gya, gyb) = got
    H = zkgpu.host.to_int
    assert (H(gxa), H(gxb), H(gya), H(gyb)) == (xa, xb, ya, yb)
    assert zkgpu.xyzz_to_affine_g2(np.zeros(64, np.uint32)) is None


def test_host_module_mirror():
    from zkgpu import host

    r = host.FR
    w = host.root_of_unity(20)
    assert host.pow_mod(w, host.from_int(1 << 20), r) == host.from_int(1)
    assert host.words32(host.from_words32(list(range(8)))) == list(range(8))
    x = host.from_int(123)
    assert host.mul_mod(x, host.inverse(x, r), r) == host.from_int(1)
    # the reference's contract: R = 2^260 (src/host.rs:84-91), reached by 260 doublings there
    rr = 1
    for _ in range(260):
        rr = (rr + rr) % host.to_int(r)
    assert host.to_int(host.montgomery_r(r)) == rr
    assert host.from_montgomery(host.mul_mod(x, host.montgomery_r(r), r), r) == x
    # this library's device form: R = 2^256
    assert host.to_int(host.library_montgomery_r(r)) == (1 << 256) % host.to_int(r)
    assert host.from_library_montgomery(host.mul_mod(x, host.library_montgomery_r(r), r), r) == x


def test_device_open_without_gpu_fails_loudly():
    import torch

    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    import zkgpu

    with pytest.raises(zkgpu.ZkgError, match="GPU init failed"):
        zkgpu.Gpu()
    import asyncio

    with pytest.raises(zkgpu.ZkgError, match="GPU init failed"):
        asyncio.run(zkgpu.Gpu.new_async())


def test_kernels_contain_no_device_calls():
    """Every device helper is force-inlined; an out-of-line function needing relaxed long branches
    lost its return address (s[30:31]) to them and never returned. Checks the built objects."""
    import glob
    import subprocess
    import sys

    objs = sorted(glob.glob(os.path.join(ROOT, "zk-gpu_amd", "_build", "obj", "*.hip.o")))
    if not objs or not os.path.exists("/opt/rocm/llvm/bin/llvm-objdump"):
        pytest.skip("object files or ROCm LLVM tools not present")
    script = os.path.join(ROOT, "zk-gpu_amd", "check_code_objects.py")
    r = subprocess.run([sys.executable, script] + objs, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_plan_bounds_reject_32_bit_index_overflow():
    """Plans whose sorted (point, digit) entries W * nv would reach 2^32 are rejected on the host,
    before any allocation; the reference's own argument errors keep their text (src/msm.rs:96-98)."""
    from zkgpu import Group, Msm, ZkgError

    # accepted: the bench's plans; workspace grows with n
    small = Msm.plan_check(Group.G1, 1 << 16, 11)
    big = Msm.plan_check(Group.G1, 1 << 20, 15)
    g2 = Msm.plan_check(Group.G2, 1 << 20, 15)
    assert 0 < small < big < g2
    # GLV: nv = 2n virtual points over ceil(128 / c) windows
    assert Msm.plan_check(Group.G1, (1 << 28) - 1, 16) > 0  # 8 (2^29 - 2) < 2^32
    for n, c in [(1 << 28, 16), (1 << 29, 16), (150_000_000, 9), ((1 << 30) - 1, 20)]:
        windows = -(-128 // c)
        assert windows * 2 * n >= 1 << 32
        with pytest.raises(ZkgError, match="sorted entries; at most 2\\^32 - 1 fit one plan"):
            Msm.plan_check(Group.G1, n, c)
    # precomputed bases: 4n virtual points over ceil(64 / c) windows, n < 2^29
    assert Msm.plan_check(Group.G1, 1 << 20, 16, precomputed=True) > 0
    with pytest.raises(ZkgError, match="sorted entries"):
        Msm.plan_check(Group.G1, (1 << 29) - 1, 9, precomputed=True)  # 8 * 4 (2^29 - 1) >= 2^32
    with pytest.raises(ZkgError, match="c in 9..=16, got 17"):
        Msm.plan_check(Group.G1, 1000, 17, precomputed=True)
    with pytest.raises(ZkgError, match="an MSM needs at least one point"):
        Msm.plan_check(Group.G1, 0, 11)
    with pytest.raises(ZkgError, match="window size c must be in 9..=20, got 8"):
        Msm.plan_check(Group.G1, 10, 8)
    with pytest.raises(ZkgError, match="at most 2\\^31 - 1 points"):
        Msm.plan_check(Group.G1, 1 << 31, 16)
