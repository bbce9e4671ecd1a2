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
VALU_PEAK_LANE_OPS = 256 * 64 * 2.4e9  # 256 CUs x 64 lanes x 2.4 GHz, full-rate 32-bit integer ops
# rocprofv3 measurements of ntt_pass_kernel<11,512> on NTT 2^24 (profiles/ntt_pass_2^24_per_pass.md,
# profiles/valu_rate_msm_accumulate_and_ntt_pass.md): HBM bytes per launch, the mean of the three
# passes (FETCH_SIZE scaled by the copy-kernel calibration factor 2.0, WRITE_SIZE by 1.0), and VALU
# lane operations per launch.
NTT24_KERNEL = "ntt_pass_kernel<11,512>"  # the default tile at 2^24: 2^11 elements, 512 lanes
NTT24_TRAFFIC_BYTES_PER_LAUNCH = 1.537e9
NTT24_VALU_LANE_OPS_PER_LAUNCH = 2.044e10
# The same for msm_accumulate29_kernel on the headline plan (G1 2^20, c = 15, 589824 lanes).
MSM_ACC_VALU_LANE_OPS_PER_LAUNCH = 4.762e10


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


def glv_digit_counts(gpu, scalars, c):
    """Nonzero signed digits of the actual scalars as the plan splits them: each scalar becomes two
    GLV half-scalars (|k| < 2^127, sign in bit 127; split on the device by the library's own
    kernel, which tests/test_gpu_glv.py checks against a lattice decomposition), and each half
    takes W = ceil(128 / c) windows of balanced width with the plan's digit rule (v = bits + carry;
    v > 2^(w-1) wraps to v - 2^w with carry 1). Returns (nonzero digits, W, widths)."""
    n = scalars.size // 8
    halves = np.zeros(8 * n, np.uint32)
    _lib.check(_lib.load().zkg_test_glv_split(gpu.handle, scalars.ctypes.data_as(_lib.c_u32p), n,
                                              halves.ctypes.data_as(_lib.c_u32p)))
    h = halves.reshape(2 * n, 4).astype(np.uint64)
    lo = h[:, 0] | (h[:, 1] << np.uint64(32))
    hi = h[:, 2] | ((h[:, 3] & np.uint64(0x7FFFFFFF)) << np.uint64(32))
    W = -(-128 // c)
    wb, nw = 128 // W, 128 - (128 // W) * W
    widths = [wb + (1 if w < nw else 0) for w in range(W)]
    carry = np.zeros(2 * n, np.uint64)
    nnz, pos = 0, 0
    for width in widths:
        mask = np.uint64((1 << width) - 1)
        if pos + width <= 64:
            bits = (lo >> np.uint64(pos)) & mask
        elif pos >= 64:
            bits = (hi >> np.uint64(pos - 64)) & mask
        else:
            bits = ((lo >> np.uint64(pos)) | (hi << np.uint64(64 - pos))) & mask
        v = bits + carry
        wrap = v > np.uint64(1 << (width - 1))
        mag = np.where(wrap, np.uint64(1 << width) - v, v)
        carry = wrap.astype(np.uint64)
        nnz += int(np.count_nonzero(mag))
        pos += width
    return nnz, W, widths


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(plain_bases_g1, ks, scalars, gpu_affine, g2, ntt_input, ntt_output, log_n_ntt, groth=None):
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
    if groth is not None:  # the Groth16 set: every MSM (plain, batched, tuned) against its closed form
        for label, (group, gks, gsc, got) in groth.items():
            want = O.point_mul(group, O.generator_words(group), O.closed_form_scalar(gks, gsc))
            if not np.array_equal(want, got):
                raise SystemExit(f"Groth16 set MSM {label} differs from the closed form: refusing to report a time")
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
        "cpu_model": cpu_model(),
        "nproc": os.cpu_count(),
        "kind": "port",
        "sample": f"the full G1 MSM 2^20 on the same inputs (oracle: C++ Jacobian Pippenger, {threads} threads); "
                  f"NTT 2^{log_n_ntt} forward took {cpu_ntt_ms:.1f} ms on the same cores",
        "ntt_ms": round(cpu_ntt_ms, 3),
    }


def fast_ks(rng, n):
    """n base multipliers k_i uniform in [0, 2^253) (< r), without rejection sampling."""
    k = rng.integers(0, 1 << 32, size=(n, 8), dtype=np.uint64).astype(np.uint32)
    k[:, 7] &= 0x1FFFFFFF
    return k.reshape(-1)


def witness_like(rng, n):
    """Groth16-witness-like scalars: 40% in {0, 1}, 30% uniform below 2^16, 30% uniform in [0, r)."""
    u = uniform_fr(rng, n).reshape(n, 8)
    kind = rng.integers(0, 10, size=n)
    out = np.zeros((n, 8), dtype=np.uint32)
    out[kind < 4, 0] = rng.integers(0, 2, size=int((kind < 4).sum()), dtype=np.uint64).astype(np.uint32)
    small = (kind >= 4) & (kind < 7)
    out[small, 0] = rng.integers(0, 1 << 16, size=int(small.sum()), dtype=np.uint64).astype(np.uint32)
    out[kind >= 7] = u[kind >= 7]
    return out.reshape(-1)


def library_bases(gpu, group, ks):
    bb = device_bases(gpu, group, ks)
    Msm.convert_bases(bb, 16 * group * (ks.size // 8))
    return bb


def batched_ntt_leg(gpu, ev, rng, args):
    """64 polynomials, log2 sizes uniform in 12..22, half forward and half inverse: one batch
    (one grouped launch per pass phase) against 64 sequential plain calls."""
    from zkgpu import Batch

    sizes = [int(k) for k in rng.integers(12, 23, size=64)]
    inverse = [j % 2 == 1 for j in range(64)]
    rng.shuffle(inverse)
    pool = uniform_fr(rng, 1 << 22)
    bufs_b = [gpu.buffer(pool[: 8 << k]) for k in sizes]
    bufs_s = [gpu.buffer(pool[: 8 << k]) for k in sizes]
    plans = {k: Ntt(gpu, k) for k in set(sizes)}
    batch = Batch(gpu)
    for k, inv, b in zip(sizes, inverse, bufs_b):
        batch.add_ntt(b, k, inverse=inv)

    def sequential():
        for k, inv, b in zip(sizes, inverse, bufs_s):
            (plans[k].inverse if inv else plans[k].forward)(gpu, b)

    batch.run()
    sequential()
    gpu.wait()
    for j, (k, bb, bs) in enumerate(zip(sizes, bufs_b, bufs_s)):
        if not np.array_equal(gpu.read(bb, 8 << k), gpu.read(bs, 8 << k)):
            raise SystemExit(f"batched NTT {j} (2^{k}) differs from the plain call: refusing to report a time")
    steps = max(5, args.steps)
    seq_ms, _ = time_calls(gpu, ev, sequential, steps, 2)
    bat_ms, _ = time_calls(gpu, ev, batch.run, steps, 2)
    # the same batch captured as one hipGraph (gated: a replay on fresh inputs equals the plain calls)
    batch.capture()
    for b, k in zip(bufs_b, sizes):
        gpu.write(b, pool[: 8 << k])
    for b, k in zip(bufs_s, sizes):
        gpu.write(b, pool[: 8 << k])
    batch.run()
    sequential()
    gpu.wait()
    for j, (k, bb, bs) in enumerate(zip(sizes, bufs_b, bufs_s)):
        if not np.array_equal(gpu.read(bb, 8 << k), gpu.read(bs, 8 << k)):
            raise SystemExit(f"graph replay of NTT {j} (2^{k}) differs from the plain call: refusing to report a time")
    graph_ms, _ = time_calls(gpu, ev, batch.run, steps, 2)
    elems = sum(1 << k for k in sizes)
    res = {"polys": 64, "log_n": sizes, "elements": elems, "sequential_ms": round(seq_ms, 4),
           "batched_ms": round(bat_ms, 4), "graph_ms": round(graph_ms, 4),
           "speedup": round(seq_ms / min(bat_ms, graph_ms), 3),
           "GBps_algorithmic": round(64 * elems / (min(bat_ms, graph_ms) * 1e-3) / 1e9, 1)}
    log(f"batched NTT x64: sequential {seq_ms:.3f} ms, batched {bat_ms:.3f} ms, graph {graph_ms:.3f} ms "
        f"({seq_ms / min(bat_ms, graph_ms):.2f}x)")
    return res


def groth16_msm_leg(gpu, ev, rng, args, log_n=22):
    """A Groth16 proof's MSM set at n = 2^22: A, B1, C (G1, n), H (G1, n - 1), B2 (G2, n) with
    witness-like scalars (H: uniform), bases uploaded once. Batched (the five MSMs overlapped on
    their own streams) against five plain calls in sequence."""
    from zkgpu import Batch

    from zkgpu import tune

    n = 1 << log_n
    c = Msm.default_window(n)
    sets = [(1, n), (1, n), (1, n), (1, n - 1), (2, n)]
    names = ["A", "B1", "C", "H", "B2"]
    host_ks = [fast_ks(rng, m) for _, m in sets]
    bases = [library_bases(gpu, g, k) for (g, _), k in zip(sets, host_ks)]
    wit_h, hsc_h = witness_like(rng, n), uniform_fr(rng, n - 1)
    wit, hsc = gpu.buffer(wit_h), gpu.buffer(hsc_h)
    scal = [wit, wit, wit, hsc, wit]
    scal_h = [wit_h, wit_h, wit_h, hsc_h, wit_h]
    plain_plans = [Msm.for_group(gpu, Group(g), m, c) for g, m in sets]
    default_plans = [Msm.for_group(gpu, Group(g), m, c) for g, m in sets]
    t0 = time.perf_counter()
    tuned = {g: tune.tune_msm(gpu, g, n) for g in (1, 2)}
    tune_s = time.perf_counter() - t0
    plans = [Msm.tuned(gpu, Group(g), m) for g, m in sets]
    batch = Batch(gpu)
    for p, b, sc in zip(plans, bases, scal):
        batch.add_msm(p, b, sc)
    batch_default = Batch(gpu)  # the same batch with default-window plans: batching alone, no tuning
    for p, b, sc in zip(default_plans, bases, scal):
        batch_default.add_msm(p, b, sc)

    def plain():
        for p, b, sc in zip(plain_plans, bases, scal):
            p.run(gpu, b, sc)

    def results(ps):
        gpu.wait()
        return [affine_words(gpu.read(p.result_buffer(gpu), 16 * g * 2), g) for p, (g, _) in zip(ps, sets)]

    plain()
    want = results(plain_plans)
    batch.run()
    got = results(plans)
    batch_default.run()
    got_default = results(default_plans)
    if not all(np.array_equal(a, b) and np.array_equal(a, d) for a, b, d in zip(want, got, got_default)):
        raise SystemExit("batched Groth16 MSM set differs from the plain calls: refusing to report a time")
    # kept for the exactness gate of the CPU leg: each MSM against (sum k_i s_i) G
    gate = {names[j]: (g, host_ks[j], scal_h[j][: 8 * m], want[j]) for j, (g, m) in enumerate(sets)}
    steps = max(3, args.steps // 5)
    plain_ms, _ = time_calls(gpu, ev, plain, steps, 1)
    bat_default_ms, _ = time_calls(gpu, ev, batch_default.run, steps, 1)
    bat_ms, _ = time_calls(gpu, ev, batch.run, steps, 1)
    batch.capture()
    batch.run()
    gpu.wait()
    got = [affine_words(gpu.read(p.result_buffer(gpu), 16 * g * 2), g) for p, (g, _) in zip(plans, sets)]
    if not all(np.array_equal(a, b) for a, b in zip(want, got)):
        raise SystemExit("graph replay of the Groth16 MSM set differs from the plain calls: refusing to report a time")
    graph_ms, _ = time_calls(gpu, ev, batch.run, steps, 1)
    res = {"n": n, "msms": "A,B1,C,H(n-1) G1 + B2 G2, witness-like scalars (H uniform)",
           "plain": {"c": c, "ms": round(plain_ms, 3)},
           "batched_default_plans": {"c": c, "ms": round(bat_default_ms, 3),
                                     "speedup": round(plain_ms / bat_default_ms, 3)},
           "batched": {"ms": round(bat_ms, 3), "graph_ms": round(graph_ms, 3), "tuned_g1": tuned[1],
                       "tuned_g2": tuned[2], "tuning_seconds": round(tune_s, 2)},
           "speedup": round(plain_ms / min(bat_ms, graph_ms), 3),
           "exactness": "each MSM equal to the closed form (sum k_i s_i) G (CPU leg); batched = plain"}
    log(f"Groth16 MSM set 2^{log_n}: plain {plain_ms:.2f} ms, batched (default plans) {bat_default_ms:.2f} ms, "
        f"batched+tuned {bat_ms:.2f} ms ({plain_ms / bat_ms:.2f}x); tuned {tuned}")
    del batch, batch_default, plans, plain_plans, default_plans, bases
    return res, gate


def serving_leg(gpu, rng, args):
    """Poisson arrivals of proof jobs over 8 device-resident proving keys (log_n uniform in
    14..24): each job = 7 NTTs of size n (random directions) + the Groth16 MSM set of size n.
    The same arrival schedule is served (a) one call at a time through the plain API and (b) by
    the continuous-batching server; jobs/s at saturating load and p50/p99 latency at moderate load."""
    import queue
    import threading

    from zkgpu import Server

    n_keys, ring = 8, 3
    log_ns = [int(k) for k in rng.integers(14, 25, size=n_keys)]
    top = 1 << max(log_ns)
    ks_pool = fast_ks(rng, top + 64)
    poly_pool = uniform_fr(rng, top)
    wit_pool = witness_like(rng, top)
    keys = []
    off = 0
    for ln in log_ns:
        n = 1 << ln
        g1 = []
        for m in (n, n, n, n - 1):
            g1.append(library_bases(gpu, 1, ks_pool[8 * off: 8 * (off + m)]))
            off = (off + 7) % 64
        g2 = library_bases(gpu, 2, ks_pool[8 * off: 8 * (off + n)])
        off = (off + 7) % 64
        sets = [[gpu.buffer(poly_pool[: 8 * n]) for _ in range(7)] for _ in range(ring)]
        keys.append({"log_n": ln, "g1": g1, "g2": g2, "wit": gpu.buffer(wit_pool[: 8 * n]), "sets": sets})
    gpu.wait()
    log(f"serving: 8 keys at log_n {log_ns} resident")

    plain_plans = {}

    def plans_for(ln):
        if ln not in plain_plans:
            n = 1 << ln
            c = Msm.default_window(n)
            plain_plans[ln] = (Ntt(gpu, ln), [Msm.new(gpu, n, c) for _ in range(3)] + [Msm.new(gpu, n - 1, c)]
                               + [Msm.for_group(gpu, Group.G2, n, c)])
        return plain_plans[ln]

    def plain_job(key, mask, polys):
        ntt, msms = plans_for(key["log_n"])
        for p, b in enumerate(polys):
            (ntt.inverse if (mask >> p) & 1 else ntt.forward)(gpu, b)
        scal = [key["wit"]] * 3 + [polys[6], key["wit"]]
        bases = key["g1"] + [key["g2"]]
        res = [m.run(gpu, b, sc) for m, b, sc in zip(msms, bases, scal)]
        return [gpu.read(r, 32 if i < 4 else 64) for i, r in enumerate(res)]

    # service time of each key on the plain path (warm)
    service = []
    for key in keys:
        plain_job(key, 0, key["sets"][0])
        t0 = time.perf_counter()
        plain_job(key, 0, key["sets"][0])
        service.append(time.perf_counter() - t0)
    mean_service = float(np.mean(service))
    log(f"serving: plain service times {[round(1e3 * s, 2) for s in service]} ms, mean {1e3 * mean_service:.2f}")

    jobs_n = args.serve_jobs
    choice = [int(k) for k in rng.integers(0, n_keys, size=jobs_n)]
    masks = [int(m) for m in rng.integers(0, 128, size=jobs_n)]
    gaps = rng.exponential(1.0, size=jobs_n)

    def schedule(rate):
        t = np.cumsum(gaps / rate)
        return t - t[0]

    def run_plain(arrivals):
        lat = []
        t0 = time.perf_counter()
        for j in range(jobs_n):
            now = time.perf_counter() - t0
            if now < arrivals[j]:
                time.sleep(arrivals[j] - now)
            key = keys[choice[j]]
            plain_job(key, masks[j], key["sets"][j % ring])
            lat.append((time.perf_counter() - t0 - arrivals[j]) * 1e3)
        span = time.perf_counter() - t0
        return {"jobs_per_s": round(jobs_n / span, 3), "p50_ms": round(float(np.percentile(lat, 50)), 3),
                "p99_ms": round(float(np.percentile(lat, 99)), 3), "mean_ms": round(float(np.mean(lat)), 3)}

    from zkgpu import tune

    t0 = time.perf_counter()
    for ln in sorted(set(log_ns)):
        for g in (1, 2):
            tune.tune_msm(gpu, g, 1 << ln)
    tune_s = time.perf_counter() - t0
    log(f"serving: tuned the key sizes in {tune_s:.1f} s")
    server = Server(gpu, args.serve_budget_gib << 30, max_inflight=args.serve_inflight)
    kid = [server.add_key(k["log_n"], k["g1"], k["g2"]) for k in keys]

    def run_server(arrivals):
        server.reset_stats()
        done_ev = [threading.Event() for _ in range(jobs_n)]
        q = queue.Queue()
        errors = []

        def waiter():
            try:
                for _ in range(jobs_n):
                    j, jid = q.get()
                    server.wait(jid)
                    done_ev[j].set()
            except Exception as e:  # noqa: BLE001 - reported below
                errors.append(e)
                for e2 in done_ev:
                    e2.set()

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        last_use = {}
        for j in range(jobs_n):
            now = time.perf_counter() - t0
            if now < arrivals[j]:
                time.sleep(arrivals[j] - now)
            key = keys[choice[j]]
            slot = (choice[j], j % ring)
            if slot in last_use:  # the ring's input buffers are reused once their job is done
                done_ev[last_use[slot]].wait()
            last_use[slot] = j
            q.put((j, server.submit(kid[choice[j]], key["sets"][j % ring], masks[j], key["wit"])))
        th.join()
        if errors:
            raise errors[0]
        st = server.stats()
        return {"jobs_per_s": round(st["jobs_per_s"], 3), "p50_ms": round(st["p50_ms"], 3),
                "p99_ms": round(st["p99_ms"], 3), "mean_ms": round(st["mean_ms"], 3),
                "peak_inflight": st["peak_inflight"], "peak_GiB": round(st["peak_bytes"] / 2 ** 30, 2)}

    # warm the server's slots for every key size
    warm = [server.submit(kid[i], keys[i]["sets"][0], 0, keys[i]["wit"]) for i in range(n_keys)]
    for w in warm:
        server.wait(w)
    # gate: a job served in the mix gives the same results as the same job alone on the plain path
    key = keys[choice[0]]
    for b in key["sets"][1]:
        gpu.write(b, poly_pool[: 8 << key["log_n"]])
    for b in key["sets"][2]:
        gpu.write(b, poly_pool[: 8 << key["log_n"]])
    gpu.wait()
    mix = [server.submit(kid[i], keys[i]["sets"][0], masks[i], keys[i]["wit"]) for i in range(n_keys) if i != choice[0]]
    jid = server.submit(kid[choice[0]], key["sets"][1], masks[0], key["wit"])
    words, _ = server.wait(jid)
    for m in mix:
        server.wait(m)
    alone = plain_job(key, masks[0], key["sets"][2])
    served = [words[0:32], words[32:64], words[64:96], words[96:128], words[128:192]]
    n = 1 << key["log_n"]
    if not all(np.array_equal(affine_words(a, 1 if i < 4 else 2), affine_words(b, 1 if i < 4 else 2))
               for i, (a, b) in enumerate(zip(served, alone))) or not all(
            np.array_equal(gpu.read(x, 8 * n), gpu.read(y, 8 * n)) for x, y in zip(key["sets"][1], key["sets"][2])):
        raise SystemExit("a served job differs from the same job run alone: refusing to report a time")

    out = {"keys_log_n": log_ns, "jobs": jobs_n, "mean_plain_service_ms": round(1e3 * mean_service, 3),
           "inflight": args.serve_inflight, "budget_GiB": args.serve_budget_gib,
           "server_msm_plans": "tuned per key size", "plain_msm_plans": "default_window", "tuning_seconds": round(tune_s, 2)}
    for name, load in (("saturating", 4.0), ("moderate", 0.75)):
        arrivals = schedule(load / mean_service)
        p = run_plain(arrivals)
        sv = run_server(arrivals)
        out[name] = {"offered_jobs_per_s": round(load / mean_service, 3), "plain": p, "server": sv,
                     "throughput_x": round(sv["jobs_per_s"] / p["jobs_per_s"], 3)}
        log(f"serving {name}: plain {p}, server {sv}")
    server.close()
    plain_plans.clear()
    return out


def copy_calibration(gpu):
    """A plain copy kernel of known size (256 MiB read + 256 MiB write per launch, 3 launches), so
    that rocprofv3's FETCH_SIZE / WRITE_SIZE for it calibrate the same counters on the NTT kernel."""
    ms = ctypes.c_double()
    _lib.check(_lib.load().zkg_bench_copy(gpu.handle, 256 << 20, 3, ctypes.byref(ms)))
    log(f"copy calibration: copy_kernel 256 MiB -> 256 MiB x 3, {ms.value:.3f} ms per launch "
        f"({2 * 256 * 2 ** 20 / (ms.value * 1e-3) / 1e9:.0f} GB/s)")


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
    ap.add_argument("--serve-jobs", type=int, default=32, help="proof jobs per serving run")
    ap.add_argument("--serve-inflight", type=int, default=4, help="server: jobs on the device at once")
    ap.add_argument("--serve-budget-gib", type=int, default=96, help="server: HBM budget (keys + workspaces)")
    ap.add_argument("--workload", choices=["all", "ntt24", "msm", "batch", "serve"], default="all",
                    help="ntt24/msm restrict the run to one kernel family (for rocprofv3 counter passes)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run the library's plain 256 MiB copy kernel (reference for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}). "
              "An MSM would shard by points, each GPU summing its share, with one gather of K points.",
              file=sys.stderr)
        sys.exit(2)

    gpu = zkgpu.Gpu(0)
    info = gpu.info()
    log(f"device {info}")
    if args.copy_calibration:
        copy_calibration(gpu)
    if args.workload == "ntt24":
        ntt_only(gpu, args)
        return
    ev = HipEvents(gpu.stream())
    rng = np.random.default_rng(args.seed)
    if args.workload in ("batch", "serve"):
        res = {}
        if args.workload == "batch":
            res["batched_ntt_64"] = batched_ntt_leg(gpu, ev, rng, args)
            res["groth16_msm_set_2^22"] = groth16_msm_leg(gpu, ev, rng, args)[0]
        else:
            res["serving"] = serving_leg(gpu, rng, args)
        print(json.dumps(res), flush=True)
        return

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
    # the same single call with the autotuner's window / item / sort sizes for this size
    from zkgpu import tune

    tuned1 = tune.tune_msm(gpu, 1, n1)
    mt = Msm.tuned(gpu, Group.G1, n1)
    if not np.array_equal(affine_words(gpu.read(mt.run(gpu, bb1, sb1), 32), 1), g1_affine):
        raise SystemExit("tuned G1 MSM 2^20 differs from the default plan: refusing to report a time")
    g1_tuned_ms, _ = time_calls(gpu, ev, lambda: mt.run(gpu, bb1, sb1), args.steps, args.warmup)
    log(f"G1 2^20 tuned {tuned1}: {g1_tuned_ms:.3f} ms")
    del mt

    # ---------------- G2 MSM 2^18
    n2 = 1 << 18
    ks2, sc2 = uniform_fr(rng, n2), uniform_fr(rng, n2)
    bb2 = device_bases(gpu, 2, ks2)
    Msm.convert_bases(bb2, 32 * n2)
    sb2 = gpu.buffer(sc2)
    m2 = Msm.for_group(gpu, Group.G2, n2, Msm.default_window(n2))
    g2_affine = affine_words(gpu.read(m2.run(gpu, bb2, sb2), 64), 2)
    g2_ms, _ = time_calls(gpu, ev, lambda: m2.run(gpu, bb2, sb2), max(3, args.steps // 4), 1)
    stages2 = m2.profile(gpu, bb2, sb2)
    log(f"G2 2^18: {g2_ms:.3f} ms; stages {stages2}")

    # ---------------- the same MSMs over precomputed bases (P_i and 2^64 P_i, built once per key)
    precomputed = {}
    for label, group, n, bb, sb, exp_aff in (("g1_msm_2^20", 1, n1, bb1, sb1, g1_affine),
                                            ("g2_msm_2^18", 2, n2, bb2, sb2, g2_affine)):
        gpu.wait()
        t0 = time.perf_counter()
        table = Msm.precompute_bases(gpu, Group(group), bb, n)
        gpu.wait()
        pre_ms = (time.perf_counter() - t0) * 1e3
        mp = Msm.precomputed(gpu, Group(group), n, 16)
        if not np.array_equal(affine_words(gpu.read(mp.run(gpu, table, sb), 32 * group), group), exp_aff):
            raise SystemExit(f"precomputed {label} differs from the plain call: refusing to report a time")
        ms, _ = time_calls(gpu, ev, lambda: mp.run(gpu, table, sb), max(3, args.steps // 2), 1)
        precomputed[label] = {"ms": round(ms, 4), "c": 16, "precompute_ms": round(pre_ms, 2)}
        log(f"{label} over precomputed bases: {ms:.3f} ms (c = 16; table built in {pre_ms:.1f} ms)")
        del mp, table

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

    # ---------------- Groth16 quotient polynomial H = (A B - C) / (x^n - 1) at 2^22 (7 NTTs + pointwise)
    h_res = None
    if args.workload == "all":
        from zkgpu import Groth16H

        lh = 22
        nh = 1 << lh
        ha, hb = uniform_fr(rng, nh), uniform_fr(rng, nh)
        hplan = Groth16H(gpu, lh)
        hbufs = [gpu.buffer(ha), gpu.buffer(hb), gpu.buffer(uniform_fr(rng, nh))]
        hplan.run(gpu, *hbufs)
        h_ms, _ = time_calls(gpu, ev, lambda: hplan.run(gpu, *hbufs), max(3, args.steps // 4), 1)
        h_res = {"log_n": lh, "ms": round(h_ms, 3), "ntts": 7}
        log(f"Groth16 H 2^{lh}: {h_ms:.3f} ms")
        del hplan, hbufs

    # ---------------- field-multiply throughput (the integer-VALU ceiling of both operators)
    ms = ctypes.c_double()
    threads, chain = 256 * 2048, 1000
    _lib.check(_lib.load().zkg_bench_field_mul(gpu.handle, 0, threads, chain, 2, ctypes.byref(ms)))
    _lib.check(_lib.load().zkg_bench_field_mul(gpu.handle, 0, threads, chain, 2, ctypes.byref(ms)))
    mulmod_gps = threads * chain * 2 / (ms.value * 1e-3) / 1e9

    _lib.check(_lib.load().zkg_bench_field_mul(gpu.handle, 2, threads, chain, 2, ctypes.byref(ms)))
    _lib.check(_lib.load().zkg_bench_field_mul(gpu.handle, 2, threads, chain, 2, ctypes.byref(ms)))
    mul29_gps = threads * chain * 2 / (ms.value * 1e-3) / 1e9
    # Fermat inversions in the same limbs (362 products each), checked x * x^-1 = 1 on the device
    inv_chain = 8
    _lib.check(_lib.load().zkg_bench_field_mul(gpu.handle, 3, threads, inv_chain, 1, ctypes.byref(ms)))
    _lib.check(_lib.load().zkg_bench_field_mul(gpu.handle, 3, threads, inv_chain, 1, ctypes.byref(ms)))
    inv29_ms = ms.value
    inv29_mps = threads * inv_chain / (inv29_ms * 1e-3) / 1e6

    # Algorithmic Fq products of the G1 MSM 2^20 (madd 10, add 14, dbl 9 products per operation):
    # one mixed addition per nonzero signed digit of the actual GLV half-scalars, the bucket
    # reduction's 2 additions per bucket, and the window combine's doublings and additions.
    nnz, W1, widths = glv_digit_counts(gpu, sc1, c1)
    B1 = 1 << (c1 - 1)
    acc_products = 10 * nnz
    call_products = acc_products + 14 * 2 * W1 * B1 + 9 * sum(widths[:-1]) + 14 * (W1 - 1)
    acc_ms = dict(stages)["accumulate"]
    msm_roof = {
        "kernel": "msm_accumulate29_kernel (G1 2^20, c = %d, GLV: %d windows over 2^21 half-scalars)" % (c1, W1),
        "bound": "valu",
        "nonzero_digits": nnz,
        "fq_products_accumulate": acc_products,
        "fq_products_call": call_products,
        "accumulate_ms": round(acc_ms, 4),
        "achieved_G_products_per_s": round(acc_products / (acc_ms * 1e-3) / 1e9, 1),
        "peak_G_products_per_s": round(mul29_gps, 1),
        "frac": round(acc_products / (acc_ms * 1e-3) / 1e9 / mul29_gps, 4),
        "call_G_products_per_s": round(call_products / (g1_ms * 1e-3) / 1e9, 1),
        "call_frac": round(call_products / (g1_ms * 1e-3) / 1e9 / mul29_gps, 4),
        "valu": {"lane_ops_per_launch": MSM_ACC_VALU_LANE_OPS_PER_LAUNCH,
                 "achieved": round(MSM_ACC_VALU_LANE_OPS_PER_LAUNCH / (acc_ms * 1e-3) / 1e12, 2),
                 "peak": round(VALU_PEAK_LANE_OPS / 1e12, 2), "unit": "T lane-ops/s",
                 "frac": round(MSM_ACC_VALU_LANE_OPS_PER_LAUNCH / (acc_ms * 1e-3) / VALU_PEAK_LANE_OPS, 4)},
        "note": "peak = the measured rate of independent 29-bit Montgomery products (the accumulation's own "
                "arithmetic) over the whole GPU; accumulate_ms from Msm::profile",
    }
    log(f"MSM roofline: {msm_roof}")

    batch_res = groth = serve = groth_gate = None
    if args.workload == "all":
        batch_res = batched_ntt_leg(gpu, ev, rng, args)
        groth, groth_gate = groth16_msm_leg(gpu, ev, rng, args)
        serve = serving_leg(gpu, rng, args)

    cpu = None
    if args.workload == "msm":
        print(json.dumps({"g1_msm_2^20_ms": round(g1_ms, 4), "g1_msm_2^20_tuned_ms": round(g1_tuned_ms, 4),
                          "g2_msm_2^18_ms": round(g2_ms, 4),
                          "stages": {k: round(v, 4) for k, v in stages},
                          "g2_stages": {k: round(v, 4) for k, v in stages2}}), flush=True)
        return
    if not args.no_cpu:
        log("CPU baseline (oracle) on the same inputs")
        cpu = cpu_baseline(plain_g1, ks1, sc1, g1_affine, (ks2, sc2, g2_affine), ntt24_in, ntt24_out, 24,
                           groth_gate)

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
        "dtype": "u32 (exact BN254 modular arithmetic; MSM in 9x29-bit limbs except the G2 bucket accumulation; NTT data in 8x32-bit words, twiddle products in 9x29-bit limbs)",
        "data": "synthetic, seeded (numpy PCG64 seed %d); bases k_i*G generated on the device" % args.seed,
        "config": {
            "workload": "plain single-call path: G1 MSM n=2^20 c=15 (value); G2 MSM 2^18 c=13; NTT 2^20/2^22/2^24",
            "g1_msm_2^20_ms": round(g1_ms, 4),
            "g1_msm_2^20_e2e_ms": round(g1_e2e, 3),
            "g1_msm_2^20_stages_ms": {k: round(v, 4) for k, v in stages},
            "g1_msm_2^20_tuned": {"ms": round(g1_tuned_ms, 4), "c": tuned1["c"], "item_points": tuned1["item_points"],
                                  "sort_points": tuned1["sort_points"]},
            "g2_msm_2^18_ms": round(g2_ms, 4),
            "g2_msm_2^18_stages_ms": {k: round(v, 4) for k, v in stages2},
            "precomputed_bases": precomputed,
            "ntt": {f"2^{k}": v for k, v in ntt.items()},
            "ntt_2^24_hbm_GBps": n24["hbm_GBps"],
            "groth16_h_2^22": h_res,
            "batched_ntt_64": batch_res,
            "groth16_msm_set_2^22": groth,
            "serving": serve,
            "field_mul_Gmulmod_per_s": round(mulmod_gps, 1),
            "field_mul29_Gmulmod_per_s": round(mul29_gps, 1),
            "field_inv29": {"M_inversions_per_s": round(inv29_mps, 2),
                            "products_per_inversion": round(mul29_gps * 1e3 / inv29_mps, 1),
                            "method": "Fermat x^(p-2), 253 squarings + 109 products, checked x * x^-1 = 1"},
            "g1_msm_2^20_roofline": msm_roof,
            "device": info,
        },
        "roofline": {
            "kernel": "%s (NTT 2^24 forward, %d launches)" % (NTT24_KERNEL, n24["passes"]),
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": NTT24_TRAFFIC_BYTES_PER_LAUNCH,
            "traffic_GBps": round(NTT24_TRAFFIC_BYTES_PER_LAUNCH / (per_pass_ms * 1e-3) / 1e9, 1),
            "algorithmic_bytes_per_launch": 64 * (1 << 24) // n24["passes"],
            "launch_ms": round(per_pass_ms, 4),
            "valu": {"lane_ops_per_launch": NTT24_VALU_LANE_OPS_PER_LAUNCH,
                     "achieved": round(NTT24_VALU_LANE_OPS_PER_LAUNCH / (per_pass_ms * 1e-3) / 1e12, 2),
                     "peak": round(VALU_PEAK_LANE_OPS / 1e12, 2), "unit": "T lane-ops/s",
                     "frac": round(NTT24_VALU_LANE_OPS_PER_LAUNCH / (per_pass_ms * 1e-3) / VALU_PEAK_LANE_OPS, 4)},
            "note": "the HBM fraction is the metric asked for; the kernel is integer-VALU-bound "
                    "(Montgomery products), see valu.frac",
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
