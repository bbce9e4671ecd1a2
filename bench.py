"""Benchmark of the plain single-call path on one MI355X.

Workload (the plain single-call configuration at scale): BN254 G1 MSM of 2^20 points at the
default window (c = 15), plus G2 MSM 2^18 and in-place Fr NTT 2^20 / 2^22 / 2^24 (forward, then
inverse). Inputs are synthetic and seeded: scalars and NTT elements uniform in [0, r); bases are
k_i * G with k_i uniform in [0, r), generated on the device. Everything is resident in HBM when the
timed region starts. Every timed path is first gated on an exact result.

One "step" = one G1 MSM 2^20 single call (`Msm::run` semantics: queue all stages, result in the
plan-owned buffer). `value` is its time in milliseconds, measured over `--steps` calls bracketed by
device synchronisation, after `--warmup` untimed calls.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
Prints one JSON line.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "zk-gpu_amd"))

import zkgpu  # noqa: E402
from zkgpu import Group, Msm, Ntt, _lib, host  # noqa: E402

METRIC = "G1 MSM 2^20 ms, NTT 2^24 HBM GB/s; batched vs one-call jobs/s, p99 ms"
R = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
R_WORDS = np.array([(R >> (32 * i)) & 0xFFFFFFFF for i in range(8)], dtype=np.uint32)
HBM_PEAK_GBPS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def uniform_fr(rng, n):
    """n elements uniform in [0, r): 254 random bits, reject >= r (8 LE u32 words each)."""
    out = np.empty((n, 8), dtype=np.uint32)
    filled = 0
    while filled < n:
        m = (n - filled) * 5 // 4 + 16
        cand = rng.integers(0, 1 << 32, size=(m, 8), dtype=np.uint64).astype(np.uint32)
        cand[:, 7] &= 0x3FFFFFFF
        less = np.zeros(m, dtype=bool)
        undecided = np.ones(m, dtype=bool)
        for i in range(7, -1, -1):
            lt, gt = cand[:, i] < R_WORDS[i], cand[:, i] > R_WORDS[i]
            less |= undecided & lt
            undecided &= ~(lt | gt)
        keep = cand[less][: n - filled]
        out[filled: filled + keep.shape[0]] = keep
        filled += keep.shape[0]
    return out.reshape(-1)


class HipEvents:
    """HIP events recorded on the library's own stream (the kernels' stream)."""

    def __init__(self, stream):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.stream = ctypes.c_void_p(stream)
        self.hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def event(self):
        e = ctypes.c_void_p()
        assert self.hip.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e):
        assert self.hip.hipEventRecord(e, self.stream) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        f = ctypes.c_float()
        assert self.hip.hipEventElapsedTime(ctypes.byref(f), a, b) == 0
        return float(f.value)


def time_calls(gpu, ev, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    gpu.wait()
    a, b = ev.event(), ev.event()
    t0 = time.perf_counter()
    ev.record(a)
    for _ in range(steps):
        fn()
    ev.record(b)
    gpu.wait()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    return ev.elapsed_ms(a, b) / steps, wall


def device_bases(gpu, group, ks):
    n = ks.size // 8
    kb = gpu.buffer(ks)
    bb = gpu.empty_buffer(16 * group * n)
    _lib.check(_lib.load().zkg_util_generator_multiples(gpu.handle, group, kb.handle, n, bb.handle))
    gpu.wait()
    return bb


def affine_words(words, group):
    if group == 1:
        a = zkgpu.xyzz_to_affine(words)
        return np.zeros(16, np.uint32) if a is None else np.array(host.words32(a[0]) + host.words32(a[1]), np.uint32)
    a = zkgpu.xyzz_to_affine_g2(words)
    if a is None:
        return np.zeros(32, np.uint32)
    (xa, xb), (ya, yb) = a
    return np.array(sum((host.words32(v) for v in (xa, xb, ya, yb)), []), np.uint32)


def cpu_baseline(plain_bases_g1, ks, scalars, gpu_affine, g2, ntt_input, ntt_output, log_n_ntt):
    """CPU leg: the oracle (a multi-threaded C++ restatement: Jacobian Pippenger, radix-2 NTT) is
    the checker of the GPU results and is timed on this host's cores on the same inputs."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    threads = max(1, min(len(os.sched_getaffinity(0)), 64))
    # exactness gate: closed form sum s_i (k_i G) = (sum k_i s_i) G
    cf = O.point_mul(1, O.generator_words(1), O.closed_form_scalar(ks, scalars))
    if not np.array_equal(cf, gpu_affine):
        raise SystemExit("G1 MSM 2^20 result differs from the closed form: refusing to report a time")
    ks2, sc2, g2_affine = g2
    if not np.array_equal(O.point_mul(2, O.generator_words(2), O.closed_form_scalar(ks2, sc2)), g2_affine):
        raise SystemExit("G2 MSM 2^18 result differs from the closed form: refusing to report a time")
    w = O.py_root_of_unity(log_n_ntt)
    for j in (1, 12345):
        if O.poly_eval(ntt_input, pow(w, j, O.R)) != O.words_to_int(ntt_output[8 * j: 8 * j + 8]):
            raise SystemExit(f"NTT 2^{log_n_ntt} output differs at j={j}: refusing to report a time")
    t0 = time.perf_counter()
    cpu_res = O.msm(1, plain_bases_g1, scalars, threads=threads)
    cpu_ms = (time.perf_counter() - t0) * 1e3
    if not np.array_equal(cpu_res, gpu_affine):
        raise SystemExit("CPU oracle MSM disagrees with the GPU")
    t0 = time.perf_counter()
    O.ntt(ntt_input, log_n_ntt, threads=threads)
    cpu_ntt_ms = (time.perf_counter() - t0) * 1e3
    return {
        "value": round(cpu_ms, 3),
        "unit": "ms",
        "cores": threads,
        "kind": "port",
        "sample": f"the full G1 MSM 2^20 on the same inputs (oracle: C++ Jacobian Pippenger, {threads} threads); "
                  f"NTT 2^{log_n_ntt} forward took {cpu_ntt_ms:.1f} ms on the same cores",
        "ntt_ms": round(cpu_ntt_ms, 3),
    }


def copy_calibration():
    """A plain copy kernel of known size (256 MiB read + 256 MiB write), so that rocprofv3's
    FETCH_SIZE / WRITE_SIZE for it calibrate the same counters on the NTT kernel."""
    import torch

    src = torch.zeros(64 << 20, dtype=torch.int32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.add(src, 1, out=dst)  # an elementwise kernel (a plain copy_ would go to the DMA engine)
    torch.cuda.synchronize()
    log("copy calibration: 3 x elementwise kernel reading 256 MiB and writing 256 MiB")


def ntt_only(gpu, args):
    """NTT 2^24 forward only: the kernel whose HBM traffic the roofline reports."""
    rng = np.random.default_rng(args.seed)
    n = 1 << 24
    x = uniform_fr(rng, n)
    plan = Ntt(gpu, 24)
    buf = gpu.buffer(x)
    for _ in range(args.steps):
        plan.forward(gpu, buf)
    gpu.wait()
    log(f"NTT 2^24 forward x {args.steps}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg (and its exactness gate)")
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--workload", choices=["all", "ntt24", "msm"], default="all",
                    help="ntt24/msm restrict the run to one kernel family (for rocprofv3 counter passes)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 256 MiB device-to-device copy kernel (reference for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}). "
              "An MSM would shard by points with one K-point gather; see DESIGN.md.", file=sys.stderr)
        sys.exit(2)

    gpu = zkgpu.Gpu(0)
    info = gpu.info()
    log(f"device {info}")
    if args.copy_calibration:
        copy_calibration()
    if args.workload == "ntt24":
        ntt_only(gpu, args)
        return
    ev = HipEvents(gpu.stream())
    rng = np.random.default_rng(args.seed)
    out = {}

    # ---------------- G1 MSM 2^20 (the headline)
    n1 = 1 << 20
    ks1, sc1 = uniform_fr(rng, n1), uniform_fr(rng, n1)
    bb1 = device_bases(gpu, 1, ks1)
    plain_g1 = gpu.read(bb1, 16 * n1) if not args.no_cpu else None
    Msm.convert_bases(bb1, 16 * n1)
    sb1 = gpu.buffer(sc1)
    c1 = Msm.default_window(n1)
    m1 = Msm.new(gpu, n1, c1)
    res1 = m1.run(gpu, bb1, sb1)
    g1_affine = affine_words(gpu.read(res1, 32), 1)
    log("G1 MSM 2^20 computed; timing")
    g1_ms, g1_wall = time_calls(gpu, ev, lambda: m1.run(gpu, bb1, sb1), args.steps, args.warmup)

    def e2e():
        gpu.write(sb1, sc1)
        gpu.read(m1.run(gpu, bb1, sb1), 32)

    t0 = time.perf_counter()
    for _ in range(3):
        e2e()
    g1_e2e = (time.perf_counter() - t0) * 1e3 / 3
    stages = m1.profile(gpu, bb1, sb1)
    log(f"G1 2^20: {g1_ms:.3f} ms (wall {g1_wall:.3f}, e2e {g1_e2e:.3f}); stages {stages}")

    # ---------------- G2 MSM 2^18
    n2 = 1 << 18
    ks2, sc2 = uniform_fr(rng, n2), uniform_fr(rng, n2)
    bb2 = device_bases(gpu, 2, ks2)
    Msm.convert_bases(bb2, 32 * n2)
    sb2 = gpu.buffer(sc2)
    m2 = Msm.for_group(gpu, Group.G2, n2, Msm.default_window(n2))
    g2_affine = affine_words(gpu.read(m2.run(gpu, bb2, sb2), 64), 2)
    g2_ms, _ = time_calls(gpu, ev, lambda: m2.run(gpu, bb2, sb2), max(3, args.steps // 4), 1)
    log(f"G2 2^18: {g2_ms:.3f} ms")

    # ---------------- NTT 2^20, 2^22, 2^24 forward (+ inverse round trip gate)
    ntt = {}
    ntt24_in = ntt24_out = None
    for log_n in ((20, 22, 24) if args.workload == "all" else ()):
        n = 1 << log_n
        x = uniform_fr(rng, n)
        plan = Ntt(gpu, log_n)
        buf = gpu.buffer(x)
        plan.forward(gpu, buf)
        y = gpu.read(buf, x.size)
        plan.inverse(gpu, buf)
        if not np.array_equal(gpu.read(buf, x.size), x):
            raise SystemExit(f"NTT 2^{log_n} round trip failed: refusing to report a time")
        fwd_ms, _ = time_calls(gpu, ev, lambda: plan.forward(gpu, buf), max(5, args.steps), 2)
        inv_ms, _ = time_calls(gpu, ev, lambda: plan.inverse(gpu, buf), max(5, args.steps), 2)
        gbps = 64.0 * n / (fwd_ms * 1e-3) / 1e9
        ntt[log_n] = {"forward_ms": round(fwd_ms, 4), "inverse_ms": round(inv_ms, 4), "hbm_GBps": round(gbps, 1),
                      "passes": plan.passes()}
        log(f"NTT 2^{log_n}: fwd {fwd_ms:.3f} ms inv {inv_ms:.3f} ms -> {gbps:.1f} GB/s")
        if log_n == 24:
            ntt24_in, ntt24_out = x, y
        del buf, plan

    # ---------------- field-multiply throughput (the integer-VALU ceiling of both operators)
    ms = ctypes.c_double()
    threads, chain = 256 * 2048, 1000
    _lib.check(_lib.load().zkg_bench_field_mul(gpu.handle, 0, threads, chain, 2, ctypes.byref(ms)))
    _lib.check(_lib.load().zkg_bench_field_mul(gpu.handle, 0, threads, chain, 2, ctypes.byref(ms)))
    mulmod_gps = threads * chain * 2 / (ms.value * 1e-3) / 1e9

    # Fq multiplications per G1 MSM (madd 10, add 14, dbl 9): accumulation of every nonzero digit
    W1 = (255 + c1 - 1) // c1
    fq_mults = n1 * W1 * 10  # upper bound: every digit nonzero
    msm_gmulmod = fq_mults / (g1_ms * 1e-3) / 1e9

    cpu = None
    if args.workload == "msm":
        print(json.dumps({"g1_msm_2^20_ms": round(g1_ms, 4), "g2_msm_2^18_ms": round(g2_ms, 4),
                          "stages": {k: round(v, 4) for k, v in stages}}), flush=True)
        return
    if not args.no_cpu:
        log("CPU baseline (oracle) on the same inputs")
        cpu = cpu_baseline(plain_g1, ks1, sc1, g1_affine, (ks2, sc2, g2_affine), ntt24_in, ntt24_out, 24)

    n24 = ntt[24]
    per_pass_ms = n24["forward_ms"] / n24["passes"]
    achieved = 64.0 * (1 << 24) / n24["passes"] / (per_pass_ms * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": round(g1_ms, 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(g1_ms, 4),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "u32 (exact BN254 modular arithmetic, 8x32-bit limbs)",
        "data": "synthetic, seeded (numpy PCG64 seed %d); bases k_i*G generated on the device" % args.seed,
        "config": {
            "workload": "plain single-call path: G1 MSM n=2^20 c=15 (value); G2 MSM 2^18 c=13; NTT 2^20/2^22/2^24",
            "g1_msm_2^20_ms": round(g1_ms, 4),
            "g1_msm_2^20_e2e_ms": round(g1_e2e, 3),
            "g1_msm_2^20_stages_ms": {k: round(v, 4) for k, v in stages},
            "g2_msm_2^18_ms": round(g2_ms, 4),
            "ntt": {f"2^{k}": v for k, v in ntt.items()},
            "ntt_2^24_hbm_GBps": n24["hbm_GBps"],
            "field_mul_Gmulmod_per_s": round(mulmod_gps, 1),
            "g1_msm_fq_mults_per_s_G": round(msm_gmulmod, 1),
            "device": info,
        },
        "roofline": {
            "kernel": "ntt_pass_kernel (NTT 2^24, %d passes)" % n24["passes"],
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": None,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
