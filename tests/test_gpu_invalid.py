"""Wrong sizes fail loudly with the reference's messages — tests/invalid_inputs.rs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 4096


def test_msm_with_no_points(gpu):
    from zkgpu import Msm, ZkgError

    with pytest.raises(ZkgError, match="at least one point"):
        Msm.new(gpu, 0, 10)


def test_msm_window_out_of_range(gpu):
    from zkgpu import Msm, ZkgError

    for c in (8, 21):
        with pytest.raises(ZkgError, match=f"window size c must be in 9..=20, got {c}"):
            Msm.new(gpu, N, c)


def test_msm_scalars_too_small(gpu):
    from zkgpu import Msm, ZkgError

    bases = Msm.upload_bases(gpu, np.zeros(16 * N, np.uint32))
    scalars = gpu.buffer(np.ones(4 * N, np.uint32))
    with pytest.raises(ZkgError, match="scalars buffer holds 65536 bytes; 4096 scalars need 131072"):
        Msm.new(gpu, N, 10).run(gpu, bases, scalars)


def test_msm_bases_too_small(gpu):
    from zkgpu import Msm, ZkgError

    bases = Msm.upload_bases(gpu, np.zeros(8 * N, np.uint32))
    scalars = gpu.buffer(np.ones(8 * N, np.uint32))
    with pytest.raises(ZkgError, match="bases buffer holds 131072 bytes; 4096 points need 262144"):
        Msm.new(gpu, N, 10).run(gpu, bases, scalars)


def test_ntt_size_out_of_range(gpu):
    from zkgpu import Ntt, ZkgError

    for log_n in (9, 28):
        with pytest.raises(ZkgError, match=r"log_n must be in 10\.\.=27"):
            Ntt(gpu, log_n)


def test_ntt_data_too_small(gpu):
    from zkgpu import Ntt, ZkgError

    data = gpu.buffer(np.ones(8 << 11, np.uint32))
    with pytest.raises(ZkgError, match=r"data buffer holds 65536 bytes; a 2\^12 transform needs 131072"):
        Ntt(gpu, 12).forward(gpu, data)


def test_buffers_larger_than_needed_are_accepted(gpu):
    from zkgpu import Ntt

    data = gpu.buffer(np.ones(8 << 13, np.uint32))
    ntt = Ntt(gpu, 12)
    ntt.forward(gpu, data)
    ntt.inverse(gpu, data)
    gpu.wait()
