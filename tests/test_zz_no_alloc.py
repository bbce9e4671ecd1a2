"""zkg_msm_run_noalloc: the MSM of a plan (zkg_msm_create + zkg_msm_run) with its workspace carved
from the library's static per-GPU scratch instead of allocated. Called through the C ABI; its
result equals the allocating entry point's on two calls in a row (the second smaller, so it runs
over what the first left in the shared scratch), for G1 and G2, with and without the poison fill
of the workspace, and equals the closed form; oversize plans are refused.

Results are compared as affine points after normalisation (the API's contract, as xyzz_to_affine
reads them): the XYZZ words themselves depend on the order of the additions inside a bucket, which
the counting sort's LDS cursor atomics leave open, so two runs of one plan may differ in Z-scaling."""
import ctypes
import os

import numpy as np
import pytest

import oracle as O
from conftest import uniform_fr_words, witness_scalars

pytestmark = pytest.mark.gpu


def _noalloc(gpu, group, n, c, bases, scalars, result):
    from zkgpu import _lib

    _lib.check(_lib.load().zkg_msm_run_noalloc(gpu.handle, group, n, c, bases.handle, scalars.handle,
                                               result.handle))


def _affine_words(group, words):
    from test_gpu_serve import _affine

    return _affine(words, group)


def _plan_affine(gpu, group, n, c, bases, scalars):
    from zkgpu import _lib

    lib = _lib.load()
    plan = ctypes.c_void_p()
    _lib.check(lib.zkg_msm_create(gpu.handle, group, n, c, 0, ctypes.byref(plan)))
    try:
        res = ctypes.c_void_p()
        _lib.check(lib.zkg_msm_run(plan, bases.handle, scalars.handle, ctypes.byref(res)))
        out = np.zeros(32 * group, np.uint32)
        _lib.check(lib.zkg_buffer_read(res, out.size, out.ctypes.data_as(_lib.c_u32p)))
        return _affine_words(group, out)
    finally:
        lib.zkg_msm_destroy(plan)


@pytest.mark.parametrize("group,calls", [
    (1, [(1 << 16, 11), (5000, 12)]),
    (2, [(1 << 15, 11), (3000, 10)]),
    (1, [(1 << 17, 12), (777, 9)]),
])
@pytest.mark.parametrize("poison", ["0", "1"])
def test_noalloc_matches_the_plan_entry_point(gpu, group, calls, poison):
    from test_gpu_msm import device_bases

    old = os.environ.get("ZKGPU_POISON")
    os.environ["ZKGPU_POISON"] = poison
    try:
        rng = np.random.default_rng(1900 + group + calls[0][0])
        result = gpu.empty_buffer(32 * group)
        for k, (n, c) in enumerate(calls):
            ks = uniform_fr_words(rng, n)
            sc = witness_scalars(rng, n) if k else uniform_fr_words(rng, n)
            bb, sb = device_bases(gpu, group, ks), gpu.buffer(sc)
            _noalloc(gpu, group, n, c, bb, sb, result)
            got = _affine_words(group, gpu.read(result, 32 * group))
            want = O.point_mul(group, O.generator_words(group), O.closed_form_scalar(ks, sc))
            assert np.array_equal(got, want), (k, n, c)
            assert np.array_equal(got, _plan_affine(gpu, group, n, c, bb, sb)), (k, n, c)
    finally:
        if old is None:
            os.environ.pop("ZKGPU_POISON", None)
        else:
            os.environ["ZKGPU_POISON"] = old


def test_noalloc_refuses_what_the_scratch_cannot_hold(gpu):
    from zkgpu import ZkgError

    n = 1 << 18
    bb = gpu.empty_buffer(16 * n)
    sb = gpu.empty_buffer(8 * n)
    result = gpu.empty_buffer(32)
    with pytest.raises(ZkgError, match="needs a workspace of [0-9]+ bytes; the static scratch holds 67108864"):
        _noalloc(gpu, 1, n, 13, bb, sb, result)
    with pytest.raises(ZkgError, match="an MSM needs at least one point"):
        _noalloc(gpu, 1, 0, 11, bb, sb, result)
    with pytest.raises(ZkgError, match="result buffer holds 128 bytes; the point needs 256"):
        _noalloc(gpu, 2, 100, 10, bb, sb, result)
    with pytest.raises(ZkgError, match="scalars buffer holds"):
        _noalloc(gpu, 1, 1000, 10, bb, gpu.empty_buffer(8 * 999), result)
