"""zkgpu — BN254 MSM (G1, G2) and Fr NTT on AMD Instinct MI355X (gfx950).

Python mirror of zerotruss/zk-gpu's public API (src/lib.rs:1-6: field, gpu, group, host, msm,
ntt) over libzkgpu.so's C ABI (include/zkgpu.h). All compute runs in hand-written HIP kernels;
importing this package does not touch the GPU, and every call fails loudly if the library is
missing.
"""
from . import host, tune
from ._lib import ZkgError
from .gpu import Buffer, Gpu
from .group import Group
from .msm import Msm, xyzz_to_affine, xyzz_to_affine_g2
from .ntt import Ntt
from .batch import Batch
from .serve import KeyShape, Server
from .groth16 import Groth16H
from .zkey import Zkey

__all__ = ["Batch", "Buffer", "Gpu", "Groth16H", "Group", "KeyShape", "Msm", "Ntt", "Server", "Zkey", "ZkgError", "host",
           "tune", "xyzz_to_affine", "xyzz_to_affine_g2"]
