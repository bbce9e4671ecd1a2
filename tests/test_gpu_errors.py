"""Gpu::report_errors (src/gpu.rs:40-49): errors the HIP runtime returns go to the device's
callback as "GPU error: ..." as well as to the failing call; argument errors do not; the device
keeps working after a failed allocation."""
import numpy as np
import pytest

import oracle as O
from conftest import uniform_fr_words

pytestmark = pytest.mark.gpu


def test_report_errors_receives_runtime_errors(gpu):
    from zkgpu import Ntt, ZkgError

    seen = []
    gpu.report_errors(seen.append)
    try:
        with pytest.raises(ZkgError, match="hipMalloc"):
            gpu.empty_buffer(1 << 50)  # 4 PiB: the allocation fails, nothing is launched
        assert len(seen) == 1 and seen[0].startswith("GPU error: hipMalloc"), seen
        with pytest.raises(ZkgError, match="log_n must be in 10..=27"):
            Ntt(gpu, 9)
        assert len(seen) == 1  # an argument error is the caller's, not the GPU's
        # the failed allocation leaves nothing behind for later calls
        rng = np.random.default_rng(1200)
        x = uniform_fr_words(rng, 1 << 10)
        buf = gpu.buffer(x)
        Ntt(gpu, 10).forward(gpu, buf)
        assert np.array_equal(gpu.read(buf, 8 << 10), O.ntt(x, 10))
        assert len(seen) == 1
    finally:
        gpu.report_errors(None)
    with pytest.raises(ZkgError, match="hipMalloc"):
        gpu.empty_buffer(1 << 50)
    assert len(seen) == 1  # removed: no further messages
