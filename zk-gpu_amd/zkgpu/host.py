"""256-bit host arithmetic — mirror of the reference's `host` module (src/host.rs:1-115).

U256 is a tuple of four little-endian u64 limbs, as in the reference ([u64; 4]).

`montgomery_r` and `from_montgomery` keep the reference's contract unchanged: R = 2^260, the radix
of the reference's WGSL kernels (src/host.rs:84-91,112-115), so callers that decode 2^260-form
words with them get the same values as before. This library's kernels use R = 2^256: its device
words (uploaded bases, MSM results, .zkey coordinates) are decoded with `library_montgomery_r` and
`from_library_montgomery`.
"""
from __future__ import annotations

from typing import Sequence, Tuple

U256 = Tuple[int, int, int, int]

FQ: U256 = (0x3C208C16D87CFD47, 0x97816A916871CA8D, 0xB85045B68181585D, 0x30644E72E131A029)
FR: U256 = (0x43E1F593F0000001, 0x2833E84879B97091, 0xB85045B68181585D, 0x30644E72E131A029)
MONTGOMERY_BITS = 260  # the reference's radix (src/host.rs:84)
LIBRARY_MONTGOMERY_BITS = 256  # this library's device form


def to_int(a: Sequence[int]) -> int:
    return sum(int(a[i]) << (64 * i) for i in range(4))


def from_int(x: int) -> U256:
    return tuple((x >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(4))


def add_mod(a: U256, b: U256, m: U256) -> U256:
    return from_int((to_int(a) + to_int(b)) % to_int(m))


def sub_mod(a: U256, b: U256, m: U256) -> U256:
    return from_int((to_int(a) - to_int(b)) % to_int(m))


def mul_mod(a: U256, b: U256, m: U256) -> U256:
    return from_int(to_int(a) * to_int(b) % to_int(m))


def pow_mod(base: U256, exp: U256, m: U256) -> U256:
    return from_int(pow(to_int(base), to_int(exp), to_int(m)))


def shr(a: U256, k: int) -> U256:
    return from_int(to_int(a) >> k)


def montgomery_r(m: U256) -> U256:
    """R mod m for the reference kernels' Montgomery radix R = 2^260 (src/host.rs:84-91)."""
    return from_int((1 << MONTGOMERY_BITS) % to_int(m))


def library_montgomery_r(m: U256) -> U256:
    """R mod m for this library's Montgomery radix R = 2^256 (its device words)."""
    return from_int((1 << LIBRARY_MONTGOMERY_BITS) % to_int(m))


def root_of_unity(log_n: int) -> U256:
    """Primitive 2^log_n-th root of unity of Fr (generator 5)."""
    r = to_int(FR)
    return from_int(pow(5, (r - 1) >> log_n, r))


def inverse(x: U256, m: U256) -> U256:
    mm = to_int(m)
    return from_int(pow(to_int(x), mm - 2, mm))


def words32(x: U256) -> list:
    return [(int(x[i // 2]) >> (32 * (i % 2))) & 0xFFFFFFFF for i in range(8)]


def from_words32(w) -> U256:
    return tuple(int(w[2 * i]) | (int(w[2 * i + 1]) << 32) for i in range(4))


def from_montgomery(x: U256, m: U256) -> U256:
    """x * R^-1 mod m with R = 2^260: a value in the reference kernels' Montgomery form back to its
    plain integer (src/host.rs:112-115)."""
    return mul_mod(x, inverse(montgomery_r(m), m), m)


def from_library_montgomery(x: U256, m: U256) -> U256:
    """x * 2^-256 mod m: a Montgomery-form word of this library's device buffers back to its plain
    integer."""
    return mul_mod(x, inverse(library_montgomery_r(m), m), m)
