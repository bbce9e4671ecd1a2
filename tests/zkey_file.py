"""Writes snarkjs Groth16 proving keys (.zkey, version 1) for the tests.

The layout is the one snarkjs writes (its zkey_utils module on the iden3 binary-file container):
"zkey", u32 version 1, u32 section count, then sections (u32 id, u64 byte size, data); section 1 =
u32 protocol (1 = Groth16); section 2 = n8q, q, n8r, r, nVars, nPublic, domainSize and the
verification key (alpha1, beta1, beta2, gamma2, delta1, delta2); sections 3 IC, 4 coefficients,
5 A, 6 B1, 7 B2, 8 C, 9 H, 10 contributions. Coordinates are 32-byte little-endian Montgomery
form (x * 2^256 mod q); infinity is all zeros. No snarkjs key file ships with the reference, so
this writer is the only source of the files: the tests pin the reader against it, and the MSM
results over uploaded sections against the oracle.
"""
import struct

import numpy as np

import oracle as O

MONT = 1 << 256


def mont_bytes(words: np.ndarray) -> bytes:
    """Plain affine words (8 per coordinate; zeros = infinity) -> Montgomery little-endian bytes."""
    w = np.asarray(words, dtype=np.uint32).reshape(-1, 8)
    out = bytearray()
    for row in w:
        v = O.words_to_int(row)
        out += (v * MONT % O.P).to_bytes(32, "little")
    return bytes(out)


def section(sid: int, data: bytes) -> bytes:
    return struct.pack("<IQ", sid, len(data)) + data


def groth16_header(n_vars, n_public, domain_size, vk, q=O.P, r=O.R, n8=32) -> bytes:
    h = struct.pack("<I", n8) + q.to_bytes(n8, "little") + struct.pack("<I", n8) + r.to_bytes(n8, "little")
    h += struct.pack("<III", n_vars, n_public, domain_size)
    for name in ("alpha1", "beta1", "beta2", "gamma2", "delta1", "delta2"):
        h += mont_bytes(vk[name])
    return h


def make_key(rng, n_vars: int, n_public: int, domain_size: int, infinity=()):
    """Seeded key contents: plain bases k*G for each point section and a verification key.
    Returns (points, ks, vk): points[sid] = plain words, ks[sid] = the multipliers (8 words each)."""
    from conftest import uniform_fr_words

    counts = {3: n_public + 1, 5: n_vars, 6: n_vars, 7: n_vars, 8: n_vars - n_public - 1, 9: domain_size}
    points, ks = {}, {}
    for sid, cnt in counts.items():
        group = 2 if sid == 7 else 1
        k = uniform_fr_words(rng, cnt)
        for i in infinity:
            if i < cnt:
                k[8 * i: 8 * i + 8] = 0  # 0 * G = infinity
        ks[sid] = k
        points[sid] = O.multiples(group, O.generator_words(group), k)
    vk = {}
    for name in ("alpha1", "beta1", "delta1"):
        vk[name] = O.multiples(1, O.generator_words(1), uniform_fr_words(rng, 1))
    for name in ("beta2", "gamma2", "delta2"):
        vk[name] = O.multiples(2, O.generator_words(2), uniform_fr_words(rng, 1))
    return points, ks, vk


def write(path, n_vars, n_public, domain_size, points, vk, *, magic=b"zkey", version=1, protocol=1,
          q=O.P, r=O.R, drop=(), override=None, order=(1, 2, 3, 4, 5, 6, 7, 8, 9, 10)):
    """Write a key file; `drop` leaves sections out, `override` replaces a section's bytes."""
    override = override or {}
    data = {
        1: struct.pack("<I", protocol),
        2: groth16_header(n_vars, n_public, domain_size, vk, q, r),
        4: struct.pack("<I", 0),  # no coefficients: the MSMs do not use them
        10: b"",
    }
    for sid in (3, 5, 6, 7, 8, 9):
        data[sid] = mont_bytes(points[sid])
    data.update(override)
    body = b"".join(section(sid, data[sid]) for sid in order if sid not in drop)
    count = len([s for s in order if s not in drop])
    with open(path, "wb") as f:
        f.write(magic + struct.pack("<II", version, count) + body)
