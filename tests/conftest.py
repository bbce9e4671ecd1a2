"""Shared test setup: import paths, the `gpu` marker, seeded input generators."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "zk-gpu_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


@pytest.fixture(scope="session")
def gpu():
    import zkgpu

    g = zkgpu.Gpu()
    yield g


def quick() -> bool:
    """ZK_GPU_QUICK=1 trims the largest sizes (as the reference's tests/common/mod.rs:20-23)."""
    return os.environ.get("ZK_GPU_QUICK") == "1"


R = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
_R_WORDS = np.array([(R >> (32 * i)) & 0xFFFFFFFF for i in range(8)], dtype=np.uint32)


def uniform_fr_words(rng: np.random.Generator, n: int) -> np.ndarray:
    """n elements uniform in [0, r) as 8 LE u32 words each: 254 random bits, reject >= r."""
    out = np.empty((n, 8), dtype=np.uint32)
    filled = 0
    while filled < n:
        m = (n - filled) * 5 // 4 + 16
        cand = rng.integers(0, 1 << 32, size=(m, 8), dtype=np.uint64).astype(np.uint32)
        cand[:, 7] &= 0x3FFFFFFF
        # lexicographic compare from the top word: keep cand < r
        less = np.zeros(m, dtype=bool)
        undecided = np.ones(m, dtype=bool)
        for i in range(7, -1, -1):
            lt = cand[:, i] < _R_WORDS[i]
            gt = cand[:, i] > _R_WORDS[i]
            less |= undecided & lt
            undecided &= ~(lt | gt)
        keep = cand[less][: n - filled]
        out[filled: filled + keep.shape[0]] = keep
        filled += keep.shape[0]
    return out.reshape(-1)


def witness_scalars(rng: np.random.Generator, n: int) -> np.ndarray:
    """Groth16-witness-like mix: 20% zero, 20% one, 30% below 2^16, 30% uniform in [0, r)."""
    u = uniform_fr_words(rng, n).reshape(n, 8)
    kind = rng.integers(0, 10, size=n)
    out = np.zeros((n, 8), dtype=np.uint32)
    out[kind == 2, 0] = 1
    out[kind == 3, 0] = 1
    small = (kind >= 4) & (kind < 7)
    out[small, 0] = rng.integers(0, 1 << 16, size=int(small.sum()), dtype=np.uint64).astype(np.uint32)
    big = kind >= 7
    out[big] = u[big]
    return out.reshape(-1)
