"""CPU oracle for the BN254 MSM and Fr NTT — TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the benchmark's CPU-baseline leg may import this module, and
only as the checker. The product library (zk-gpu_amd/) never imports it.

Two independent layers:
  * `liboracle.so` (bn254_oracle.cpp): fast C++ restatement (4x64-bit Montgomery, Jacobian
    points, unsigned-window Pippenger, radix-2 DIT NTT), multi-threaded.
  * the `py_*` functions below: textbook arithmetic on Python integers (affine chord-and-tangent
    law, schoolbook DFT, Horner), slow, used to pin the C++ layer on small inputs.

Reference semantics followed (zerotruss/zk-gpu):
  moduli                         src/field.rs:17-20
  LE u32 word layout             src/host.rs:104-110
  root of unity 5^((r-1)>>k)     src/host.rs:93-98
  NTT forward/inverse (1/n)      src/ntt.rs:26-50
  MSM result as affine, inf=0    src/msm.rs:345-387, tests/golden.rs:44-49
  Fq2 with u^2 = -1              src/group.rs:30-55
  golden-vector hex -> words     tests/golden_vectors/mod.rs:20-29
"""
from __future__ import annotations

import ctypes
import json
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

P = 0x30644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD47
R = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
G1 = (1, 2)
# BN254 G2 generator, coordinates a + b*u.
G2 = (
    (
        10857046999023057135944570762232829481370756359578518086990519993285655852781,
        11559732032986387107991004021392285783925812861821192530917403151452391805634,
    ),
    (
        8495653923123431417604973247489272438418190587263600148770280649306958101930,
        4082367875863433681332203403145435568316851327593401208105741076214120093531,
    ),
)

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "-B"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
        sz = ctypes.c_size_t
        i = ctypes.c_int
        for name in ("orc_field_mul", "orc_field_add", "orc_field_sub", "orc_field_pow"):
            getattr(L, name).argtypes = [i, u32p, u32p, u32p] + ([u32p] if name == "orc_field_pow" else [])
        L.orc_field_pow.argtypes = [i, u32p, u32p, u32p]
        L.orc_field_inv.argtypes = [i, u32p, u32p]
        L.orc_fq2_mul.argtypes = [u32p, u32p, u32p]
        L.orc_fq2_inv.argtypes = [u32p, u32p]
        L.orc_root_of_unity.argtypes = [i, u32p]
        L.orc_ntt.argtypes = [u32p, i, i, i]
        L.orc_poly_eval.argtypes = [u32p, sz, u32p, u32p]
        L.orc_msm.argtypes = [i, u32p, u32p, sz, u32p, i]
        L.orc_msm_naive.argtypes = [i, u32p, u32p, sz, u32p]
        L.orc_multiples.argtypes = [i, u32p, u32p, sz, u32p, i]
        L.orc_point_add.argtypes = [i, u32p, u32p, u32p]
        L.orc_point_mul.argtypes = [i, u32p, u32p, u32p]
        L.orc_dot_fr.argtypes = [u32p, u32p, sz, u32p, i]
        L.orc_poly_eval_threads.argtypes = [u32p, sz, u32p, u32p, i]
        for f in ("orc_ntt", "orc_poly_eval", "orc_msm", "orc_msm_naive", "orc_multiples",
                  "orc_point_add", "orc_point_mul", "orc_field_mul", "orc_field_add",
                  "orc_field_sub", "orc_field_inv", "orc_field_pow", "orc_fq2_mul",
                  "orc_fq2_inv", "orc_root_of_unity", "orc_dot_fr", "orc_poly_eval_threads"):
            getattr(L, f).restype = None
        _lib = L
    return _lib


def default_threads() -> int:
    return max(1, min(os.cpu_count() or 1, 64))


# ---------------------------------------------------------------- word conversions
def int_to_words(x: int) -> np.ndarray:
    return np.array([(x >> (32 * i)) & 0xFFFFFFFF for i in range(8)], dtype=np.uint32)


def words_to_int(w) -> int:
    w = np.asarray(w, dtype=np.uint32).reshape(-1)
    return sum(int(w[i]) << (32 * i) for i in range(8))


def ints_to_words(xs) -> np.ndarray:
    out = np.zeros((len(xs), 8), dtype=np.uint32)
    for k, x in enumerate(xs):
        for i in range(8):
            out[k, i] = (x >> (32 * i)) & 0xFFFFFFFF
    return out.reshape(-1)


def words_to_ints(w) -> list:
    w = np.asarray(w, dtype=np.uint32).reshape(-1, 8).astype(object)
    return [sum(int(row[i]) << (32 * i) for i in range(8)) for row in w]


def hex_to_words(h: str) -> np.ndarray:
    """Big-endian unpadded hex (the golden-vector format) -> 8 LE u32 words."""
    assert len(h) <= 64
    return int_to_words(int(h, 16))


def g1_words(pt) -> np.ndarray:
    if pt is None:
        return np.zeros(16, dtype=np.uint32)
    return np.concatenate([int_to_words(pt[0]), int_to_words(pt[1])])


def g2_words(pt) -> np.ndarray:
    if pt is None:
        return np.zeros(32, dtype=np.uint32)
    (xa, xb), (ya, yb) = pt
    return np.concatenate([int_to_words(v) for v in (xa, xb, ya, yb)])


def words_to_g1(w):
    x, y = words_to_int(w[:8]), words_to_int(w[8:16])
    return None if (x == 0 and y == 0) else (x, y)


def words_to_g2(w):
    v = [words_to_int(w[8 * i: 8 * i + 8]) for i in range(4)]
    return None if all(x == 0 for x in v) else ((v[0], v[1]), (v[2], v[3]))


# ---------------------------------------------------------------- golden vectors
GOLDEN_DIR = os.path.join(os.path.dirname(HERE), "tests", "golden")


def ntt_cases():
    with open(os.path.join(GOLDEN_DIR, "ntt.json")) as f:
        data = json.load(f)
    out = []
    for c in data["cases"]:
        out.append(dict(
            log_n=int(c["log_n"]),
            inverse=c["direction"] == "inverse",
            input=np.concatenate([hex_to_words(h) for h in c["input"]]),
            output=np.concatenate([hex_to_words(h) for h in c["output"]]),
        ))
    return out


def msm_cases():
    with open(os.path.join(GOLDEN_DIR, "msm.json")) as f:
        data = json.load(f)
    out = []
    for c in data["cases"]:
        out.append(dict(
            name=c["name"],
            bases=np.concatenate([np.concatenate([hex_to_words(h) for h in b]) for b in c["bases"]]),
            scalars=np.concatenate([hex_to_words(h) for h in c["scalars"]]),
            result=np.concatenate([hex_to_words(h) for h in c["result"]]),
        ))
    return out


# ---------------------------------------------------------------- C++ oracle wrappers
def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32).reshape(-1))


def field_op(op: str, which: int, a, b=None) -> np.ndarray:
    out = np.zeros(8, dtype=np.uint32)
    L = lib()
    if op == "inv":
        L.orc_field_inv(which, _u32(a), out)
    else:
        getattr(L, "orc_field_" + op)(which, _u32(a), _u32(b), out)
    return out


def fq2_mul(a, b) -> np.ndarray:
    out = np.zeros(16, dtype=np.uint32)
    lib().orc_fq2_mul(_u32(a), _u32(b), out)
    return out


def fq2_inv(a) -> np.ndarray:
    out = np.zeros(16, dtype=np.uint32)
    lib().orc_fq2_inv(_u32(a), out)
    return out


def root_of_unity_words(log_n: int) -> np.ndarray:
    out = np.zeros(8, dtype=np.uint32)
    lib().orc_root_of_unity(log_n, out)
    return out


def ntt(words, log_n: int, inverse: bool = False, threads: int | None = None) -> np.ndarray:
    """Natural-order NTT of 2^log_n Fr elements; returns a new word array."""
    data = _u32(words).copy()
    assert data.size == 8 << log_n
    lib().orc_ntt(data, log_n, int(inverse), threads or default_threads())
    return data


def groth16_h(a_words, b_words, c_words, log_n: int, g: int = 5) -> np.ndarray:
    """Coefficients of H = (A B - C) / (x^n - 1) from evaluations of A, B, C on {w^i}, computed
    on the coset g*{w^i}: inverse transforms, coefficients times g^i, forward transforms,
    (a b - c) / (g^n - 1) pointwise, inverse transform, coefficients times g^-i. (The reference
    has no such step; its README.md:229-230 names it as the work around the NTT.)"""
    n = 1 << log_n
    coeffs = [words_to_ints(ntt(w, log_n, inverse=True)) for w in (a_words, b_words, c_words)]
    evals = []
    for cs in coeffs:
        gi, out = 1, []
        for x in cs:
            out.append(x * gi % R)
            gi = gi * g % R
        evals.append(words_to_ints(ntt(ints_to_words(out), log_n)))
    zinv = pow((pow(g, n, R) - 1) % R, R - 2, R)
    h = [(x * y - z) * zinv % R for x, y, z in zip(*evals)]
    hc = words_to_ints(ntt(ints_to_words(h), log_n, inverse=True))
    ginv, gi, out = pow(g, R - 2, R), 1, []
    for x in hc:
        out.append(x * gi % R)
        gi = gi * ginv % R
    return ints_to_words(out)


def groth16_h_odd_coset(a_words, b_words, c_words, log_n: int) -> np.ndarray:
    """A B - C on the odd coset w_2n*{w^i} (w_2n a primitive 2n-th root of unity, generator 5), from
    evaluations of A, B, C on {w^i}: inverse transforms, coefficients times w_2n^i, forward
    transforms, a b - c pointwise. These are the H scalars of a prover whose H bases carry the
    division by Z (snarkjs's groth16 prover; the reference has no such step, README.md:224-225)."""
    g = py_root_of_unity(log_n + 1)
    coeffs = [words_to_ints(ntt(w, log_n, inverse=True)) for w in (a_words, b_words, c_words)]
    evals = []
    for cs in coeffs:
        gi, out = 1, []
        for x in cs:
            out.append(x * gi % R)
            gi = gi * g % R
        evals.append(words_to_ints(ntt(ints_to_words(out), log_n)))
    return ints_to_words([(x * y - z) % R for x, y, z in zip(*evals)])


def poly_eval(words, x: int, threads: int | None = None) -> int:
    """sum_i c_i x^i mod r by Horner; threads > 1 (the default: the host's cores) splits the
    coefficients into parts evaluated in parallel and scaled by x^start."""
    coeffs = _u32(words)
    out = np.zeros(8, dtype=np.uint32)
    t = threads or default_threads()
    if t == 1:
        lib().orc_poly_eval(coeffs, coeffs.size // 8, int_to_words(x), out)
    else:
        lib().orc_poly_eval_threads(coeffs, coeffs.size // 8, int_to_words(x), out, t)
    return words_to_int(out)


def msm(group: int, bases, scalars, threads: int | None = None) -> np.ndarray:
    """Affine plain words of sum s_i*P_i (16 words for G1, 32 for G2; infinity = zeros)."""
    b, s = _u32(bases), _u32(scalars)
    n = s.size // 8
    out = np.zeros(16 * group, dtype=np.uint32)
    lib().orc_msm(group, b, s, n, out, threads or default_threads())
    return out


def msm_naive(group: int, bases, scalars) -> np.ndarray:
    b, s = _u32(bases), _u32(scalars)
    out = np.zeros(16 * group, dtype=np.uint32)
    lib().orc_msm_naive(group, b, s, s.size // 8, out)
    return out


def multiples(group: int, point_words, ks_words, threads: int | None = None) -> np.ndarray:
    ks = _u32(ks_words)
    n = ks.size // 8
    out = np.zeros(n * 16 * group, dtype=np.uint32)
    lib().orc_multiples(group, _u32(point_words), ks, n, out, threads or default_threads())
    return out


def point_mul(group: int, point_words, scalar: int) -> np.ndarray:
    out = np.zeros(16 * group, dtype=np.uint32)
    lib().orc_point_mul(group, _u32(point_words), int_to_words(scalar % (1 << 256)), out)
    return out


def point_add(group: int, a, b) -> np.ndarray:
    out = np.zeros(16 * group, dtype=np.uint32)
    lib().orc_point_add(group, _u32(a), _u32(b), out)
    return out


def generator_words(group: int) -> np.ndarray:
    return g1_words(G1) if group == 1 else g2_words(G2)


def closed_form_msm(group: int, ks, scalars) -> np.ndarray:
    """sum s_i * (k_i * G) = (sum k_i s_i mod r) * G, exact at any n."""
    t = 0
    for k, s in zip(ks, scalars):
        t += k * s
    return point_mul(group, generator_words(group), t % R)


def closed_form_scalar(ks_words, scalar_words, threads: int | None = None) -> int:
    """sum k_i s_i mod r (every k_i, s_i < r), multi-threaded in the C restatement."""
    k, s = _u32(ks_words), _u32(scalar_words)
    out = np.zeros(8, dtype=np.uint32)
    lib().orc_dot_fr(k, s, k.size // 8, out, threads or default_threads())
    return words_to_int(out)


def closed_form_scalar_np(ks_words, scalar_words) -> int:
    """sum k_i s_i mod r from word arrays in exact integer arithmetic, vectorised over 16-bit
    pieces (independent of the field code: pins closed_form_scalar)."""
    k = _u32(ks_words).reshape(-1, 8)
    s = _u32(scalar_words).reshape(-1, 8)
    # Split into 16 sixteen-bit limbs; limb products are < 2^32 and sums over <= 2^31 rows fit u64.
    k16 = np.stack([(k >> 0) & 0xFFFF, k >> 16], axis=2).reshape(-1, 16).astype(np.uint64)
    s16 = np.stack([(s >> 0) & 0xFFFF, s >> 16], axis=2).reshape(-1, 16).astype(np.uint64)
    total = 0
    chunk = 1 << 16
    for lo in range(0, k16.shape[0], chunk):
        kk, ss = k16[lo:lo + chunk], s16[lo:lo + chunk]
        for i in range(16):
            col = (kk[:, i:i + 1] * ss).sum(axis=0)  # 16 sums, each < 2^48
            for j in range(16):
                total += int(col[j]) << (16 * (i + j))
    return total % R


# ---------------------------------------------------------------- pure-Python textbook layer
def py_inv(a: int, m: int) -> int:
    return pow(a, m - 2, m)


def py_root_of_unity(log_n: int) -> int:
    return pow(5, (R - 1) >> log_n, R)


def py_dft(values, log_n: int, inverse: bool = False):
    n = 1 << log_n
    w = py_root_of_unity(log_n)
    if inverse:
        w = py_inv(w, R)
    out = []
    for j in range(n):
        wj = pow(w, j, R)
        acc, x = 0, 1
        for v in values:
            acc = (acc + v * x) % R
            x = x * wj % R
        out.append(acc)
    if inverse:
        ninv = py_inv(n, R)
        out = [v * ninv % R for v in out]
    return out


# G1 affine (None = infinity), chord-and-tangent, y^2 = x^3 + 3.
def py_g1_add(p, q):
    if p is None:
        return q
    if q is None:
        return p
    (x1, y1), (x2, y2) = p, q
    if x1 == x2:
        if (y1 + y2) % P == 0:
            return None
        lam = 3 * x1 * x1 * py_inv(2 * y1, P) % P
    else:
        lam = (y2 - y1) * py_inv((x2 - x1) % P, P) % P
    x3 = (lam * lam - x1 - x2) % P
    return (x3, (lam * (x1 - x3) - y1) % P)


def py_g1_mul(p, k: int):
    acc = None
    for bit in bin(k)[2:] if k > 0 else "":
        acc = py_g1_add(acc, acc)
        if bit == "1":
            acc = py_g1_add(acc, p)
    return acc


def py_g1_on_curve(p) -> bool:
    return p is None or (p[1] * p[1] - p[0] ** 3 - 3) % P == 0


# Fq2 textbook: (a, b) = a + b*u, u^2 = -1.
def py_f2_mul(x, y):
    return ((x[0] * y[0] - x[1] * y[1]) % P, (x[0] * y[1] + x[1] * y[0]) % P)


def py_f2_add(x, y):
    return ((x[0] + y[0]) % P, (x[1] + y[1]) % P)


def py_f2_sub(x, y):
    return ((x[0] - y[0]) % P, (x[1] - y[1]) % P)


def py_f2_inv(x):
    nrm = (x[0] * x[0] + x[1] * x[1]) % P
    ni = py_inv(nrm, P)
    return (x[0] * ni % P, (-x[1]) * ni % P)


B2 = py_f2_mul((3, 0), py_f2_inv((9, 1)))  # twist coefficient 3/(9+u)


def py_g2_on_curve(p) -> bool:
    if p is None:
        return True
    x, y = p
    lhs = py_f2_mul(y, y)
    rhs = py_f2_add(py_f2_mul(py_f2_mul(x, x), x), B2)
    return lhs == rhs


def py_g2_add(p, q):
    if p is None:
        return q
    if q is None:
        return p
    (x1, y1), (x2, y2) = p, q
    if x1 == x2:
        if py_f2_add(y1, y2) == (0, 0):
            return None
        num = py_f2_mul((3, 0), py_f2_mul(x1, x1))
        lam = py_f2_mul(num, py_f2_inv(py_f2_add(y1, y1)))
    else:
        lam = py_f2_mul(py_f2_sub(y2, y1), py_f2_inv(py_f2_sub(x2, x1)))
    x3 = py_f2_sub(py_f2_sub(py_f2_mul(lam, lam), x1), x2)
    y3 = py_f2_sub(py_f2_mul(lam, py_f2_sub(x1, x3)), y1)
    return (x3, y3)


def py_g2_mul(p, k: int):
    acc = None
    for bit in bin(k)[2:] if k > 0 else "":
        acc = py_g2_add(acc, acc)
        if bit == "1":
            acc = py_g2_add(acc, p)
    return acc
