"""Benchmark of the MI355X MSM / NTT library.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extra]

One step = one BN254 G1 MSM over n = 2^20 points (inputs resident in HBM: uniform standard-
form scalars in [0, r), points k_i G in affine Montgomery form), i.e. the workload the headline
metric "G1 MSM Mpoints/s @2^20 BN254+BLS12-381, NTT GB/s" is quoted on.  `value` is Mpoints/s
of that MSM.  The same line carries, under "extra", the BLS12-381 MSM at 2^20 and the BN254
Fr NTT/iNTT at 2^22 (GB/s = 2 * 32 bytes * n per transform), each checked against its closed
form, plus the roofline of the dominant kernel and a CPU baseline (the reference's own C,
built from its sources into oracle/_ref, timed on this host on a bounded sample; the C
restatement in oracle/ when the reference build is absent).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "zikkurat-algebra_amd"))

if __name__ == "__main__" and "--cpu-worker" in sys.argv:
    zkgpu = None  # a CPU baseline worker process: the reference C only, no GPU library
else:
    import zkgpu  # noqa: E402

BN_R = 21888242871839275222246405745257275088548364400416034343698204186575808495617
BN_P = 21888242871839275222246405745257275088696311157297823662689037894645226208583
BLS_R = 0x73eda753299d7d483339d80809a1d80553bda402fffe5bfeffffffff00000001
BLS_P = int("1a0111ea397fe69a4b1ba7b6434bacd764774b84f38512bf6730d2a0f6b0f624"
            "1eabfffeb153ffffb9feffffffffaaab", 16)
CURVE = {0: dict(name="bn254", r=BN_R, p=BN_P, lp=4, bits=254, b=3, g=(1, 2)),
         1: dict(name="bls12-381", r=BLS_R, p=BLS_P, lp=6, bits=255, b=4,
                 g=(int("17f1d3a73197d7942695638c4fa9ac0fc3688c4f9774b905a14e3a3f171bac586c55e83ff97a1aeffb3af00adb22c6bb", 16),
                    int("08b3f481e3aaa0f1a09e30ed741d8ae4fcf5e095d5d00af600db18cb2c04b3edd03cc744a2888ae40caa232946c5e7e1", 16)))}


# published-derived integer-multiply peak: MI355X's 157.3 TFLOP/s dense fp32 vector rate is
# 2 flops per FMA over 2 packed lanes, i.e. 39.3e12 32-bit lane operations per second
PEAK_IMUL_PUBLISHED = 157.3e12 / 4


def m32(limbs32):
    """32-bit multiplies of one CIOS Montgomery multiplication with L 32-bit limbs."""
    return 2 * limbs32 * limbs32 + limbs32


def a_msm(n, curve):
    """Algorithmic u32-multiply count of a signed-digit bucket MSM (best window)."""
    L32 = 8 if curve == 0 else 12
    b = CURVE[curve]["bits"]
    best = min(((b + 1 + c - 1) // c) * (10 * n + 28 * 2 ** (c - 1)) + 10 * b for c in range(1, 25))
    return m32(L32) * best


def uniform_scalars(rng, n, r):
    """Uniform in [0, r): 4 x u64 words, rejection on the top bits."""
    bits = r.bit_length()
    out = np.empty((n, 4), np.uint64)
    filled = 0
    top_mask = np.uint64((1 << (bits - 192)) - 1)
    r_top = r >> 192
    while filled < n:
        k = int((n - filled) * 1.4) + 16
        cand = rng.integers(0, 2**64, size=(k, 4), dtype=np.uint64)
        cand[:, 3] &= top_mask
        # accept if below r (exact compare only needed when the top word equals r's)
        ok = cand[:, 3] < np.uint64(r_top)
        eq = cand[:, 3] == np.uint64(r_top)
        if eq.any():
            for i in np.nonzero(eq)[0]:
                v = sum(int(cand[i, j]) << (64 * j) for j in range(4))
                ok[i] = v < r
        acc = cand[ok][: n - filled]
        out[filled:filled + len(acc)] = acc
        filled += len(acc)
    return out


def to_int(words):
    return sum(int(w) << (64 * i) for i, w in enumerate(words))


def ec_mul(k, P, p):
    """Affine double-and-add (closed-form check only)."""
    def add(A, B):
        if A is None:
            return B
        if B is None:
            return A
        (x1, y1), (x2, y2) = A, B
        if x1 == x2:
            if (y1 + y2) % p == 0:
                return None
            lam = 3 * x1 * x1 * pow(2 * y1, -1, p) % p
        else:
            lam = (y2 - y1) * pow(x2 - x1, -1, p) % p
        x3 = (lam * lam - x1 - x2) % p
        return (x3, (lam * (x1 - x3) - y1) % p)
    R = None
    while k:
        if k & 1:
            R = add(R, P)
        P = add(P, P)
        k >>= 1
    return R


def decode_affine(words, curve):
    c = CURVE[curve]
    L, p = c["lp"], c["p"]
    if all(int(w) == 0xFFFFFFFFFFFFFFFF for w in words):
        return None
    Rinv = pow(1 << (64 * L), -1, p)
    return (to_int(words[:L]) * Rinv % p, to_int(words[L:2 * L]) * Rinv % p)


class MsmCase:
    """Seeded inputs for one MSM, resident on the device."""

    def __init__(self, curve, logn, seed):
        self.curve, self.n = curve, 1 << logn
        c = CURVE[curve]
        rng = np.random.default_rng(seed)
        self.k = uniform_scalars(rng, self.n, c["r"])
        self.s = uniform_scalars(rng, self.n, c["r"])
        dk = zkgpu.DeviceArray.from_host(self.k)
        self.d_points = zkgpu.scalar_mul_gen(curve, dk, self.n)
        dk.free()
        self.d_scalars = zkgpu.DeviceArray.from_host(self.s)

    def run(self):
        return zkgpu.msm_device(self.curve, self.n, self.d_scalars, self.d_points)

    def expected(self):
        c = CURVE[self.curve]
        r = c["r"]
        K = [to_int(row) for row in self.k]
        S = [to_int(row) for row in self.s]
        e = sum(a * b for a, b in zip(K, S)) % r
        return ec_mul(e, c["g"], c["p"])


def time_steps(fn, steps, warmup):
    for _ in range(warmup):
        fn()
    zkgpu.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    zkgpu.synchronize()
    return (time.perf_counter() - t0) / steps


def kernel_ms(fn, steps, phase):
    """Average HIP-event duration of one MSM phase over `steps` calls (library stream)."""
    tot = 0.0
    for _ in range(steps):
        fn()
        tot += zkgpu.last_timing()[1][phase]
    return tot / steps


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), logical CPUs and the CPUs this process may use."""
    name = platform.machine()
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                name = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return name, os.cpu_count() or 1, avail


def _ref_msm_fn():
    ref_so = REPO / "oracle" / "_ref" / "libzkref.so"
    if ref_so.exists():
        return ctypes.CDLL(str(ref_so)).bn128_G1_proj_MSM_std_coeff_affine_out, "reference"
    lib = ctypes.CDLL(str(REPO / "oracle" / "libzkoracle.so"))

    def fn(n, e, g, t, enl):
        return lib.zko_msm_std_affine(0, n, e, g, t, enl)
    return fn, "port"


def cpu_worker(path):
    """One CPU baseline process: the reference MSM over the sample saved at `path`; prints its
    wall time and result words as JSON."""
    d = np.load(path)
    sc, pts = np.ascontiguousarray(d["scalars"]), np.ascontiguousarray(d["points"])
    fn, _ = _ref_msm_fn()
    out = np.zeros(8, np.uint64)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    t0 = time.perf_counter()
    fn(len(sc), sc.ctypes.data_as(u64p), pts.ctypes.data_as(u64p), out.ctypes.data_as(u64p), 4)
    print(json.dumps({"s": time.perf_counter() - t0, "out": [int(x) for x in out]}), flush=True)


def cpu_baseline(case, sample_log, max_procs=16):
    """The reference C (oracle/_ref) on a bounded prefix of the same inputs: one core, and k
    independent copies on k cores at once (the reference has no parallelism inside one MSM).
    The reference picks its window as round(log2 n - 3.5) (bn128_G1_proj.c:600): 15 bits at 2^18
    against 17 at 2^20, so its per-point cost at 2^20 is (ceil(256/17) (1 + 2^17/2^20)) /
    (ceil(256/15) (1 + 2^15/2^18)) of the sample's, which `per_point_2p20_factor` states."""
    import subprocess
    import tempfile
    fn, kind = _ref_msm_fn()
    n = 1 << sample_log
    pts = case.d_points.to_host((case.n, 8))[:n].copy()
    sc = np.ascontiguousarray(case.s[:n])
    out = np.zeros(8, np.uint64)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    t0 = time.perf_counter()
    fn(n, sc.ctypes.data_as(u64p), pts.ctypes.data_as(u64p), out.ctypes.data_as(u64p), 4)
    dt = time.perf_counter() - t0
    gpu = zkgpu.msm_std(0, sc, pts)  # the CPU result must equal the GPU's on the same sample
    name, ncpu, avail = cpu_model()
    c_s = int(np.floor(np.log2(n) - 3.5 + 0.5))
    c_20 = int(np.floor(20 - 3.5 + 0.5))
    cost = lambda c, lg: -(-256 // c) * (1 + 2 ** c / 2 ** lg)
    res = {"value": round(n / dt / 1e6, 5), "unit": "Mpoints/s", "cores": 1, "kind": kind,
           "sample": f"BN254 G1 MSM n=2^{sample_log} (first 2^{sample_log} points/scalars of the benchmark "
                     f"inputs), single thread, {dt:.2f} s, reference window c={c_s} (c={c_20} at 2^20)",
           "cpu": name, "host_logical_cpus": ncpu, "cpus_available": avail,
           "per_point_2p20_factor": round(cost(c_20, 20) / cost(c_s, sample_log), 4),
           "matches_gpu": bool((gpu == out).all())}
    # k copies on k cores: separate processes (no GPU use), all started together
    k = max(1, min(avail, max_procs))
    with tempfile.TemporaryDirectory() as td:
        path = Path(td) / "sample.npz"
        np.savez(path, scalars=sc, points=pts)
        t0 = time.perf_counter()
        procs = [subprocess.Popen([sys.executable, str(Path(__file__).resolve()), "--cpu-worker", str(path)],
                                  stdout=subprocess.PIPE, text=True) for _ in range(k)]
        outs = [json.loads(p.communicate()[0].strip().splitlines()[-1]) for p in procs]
        wall = time.perf_counter() - t0
    same = all(o["out"] == [int(x) for x in out] for o in outs)
    res["all_cores"] = {"value": round(k * n / wall / 1e6, 5), "unit": "Mpoints/s", "cores": k,
                        "jobs": f"{k} independent copies of the sample, one process each, {wall:.2f} s wall",
                        "matches": same}
    return res


def bench_ntt_pair(steps, warmup, peak_imul=None, m=22):
    """BN254 Fr NTT then iNTT of 2^m uniform Montgomery elements on the 2^m domain; round trip
    checked bit for bit."""
    N = 1 << m
    rng = np.random.default_rng(22)
    X = uniform_scalars(rng, N, BN_R)  # uniform field elements (any canonical value is a valid Montgomery form)
    dX, dY, dZ = zkgpu.DeviceArray.from_host(X), zkgpu.DeviceArray(4 * N), zkgpu.DeviceArray(4 * N)
    g = zkgpu.fft_generator(0, m)
    fwd = lambda: zkgpu.ntt_device(0, m, g, dX, dY)
    inv = lambda: zkgpu.ntt_device(0, m, g, dY, dZ, inverse=True)
    tf = time_steps(fwd, steps, warmup)
    ti = time_steps(inv, steps, warmup)
    rt_ok = bool((dZ.to_host((N, 4)) == X).all())
    out = {"bn254_ntt_2p22_fwd_ms": round(tf * 1e3, 3), "bn254_ntt_2p22_inv_ms": round(ti * 1e3, 3),
           "bn254_ntt_2p22_fwd_GBps": round(64 * N / tf / 1e9, 1),
           "bn254_ntt_2p22_inv_GBps": round(64 * N / ti / 1e9, 1),
           "bn254_ntt_2p22_hbm_frac": round(64 * N / tf / 8e12, 4),
           "bn254_ntt_2p22_roundtrip_ok": rt_ok}
    if peak_imul:
        out["bn254_ntt_2p22_imul_roofline_frac"] = round((N // 2) * m * m32(8) / tf / PEAK_IMUL_PUBLISHED, 4)
        out["bn254_ntt_2p22_imul_frac_of_measured_peak"] = round((N // 2) * m * m32(8) / tf / peak_imul, 4)
    return out


def bench_ntt_batch(njobs=256, seed=4):
    """256 BN254 Fr NTTs, log sizes uniform in 10..22, fair-coin forward/inverse, half on a coset
    (seeded eta): one hipGraph replay against one plain call per job; outputs must agree."""
    rng = np.random.default_rng(seed)
    ms = rng.integers(10, 23, size=njobs)
    inv = rng.integers(0, 2, size=njobs)
    coset = rng.integers(0, 2, size=njobs)
    etas = uniform_scalars(rng, njobs, BN_R)
    jobs, plain_out = [], []
    total = 0
    for j in range(njobs):
        N = 1 << int(ms[j])
        total += N
        d_src = zkgpu.fill_random(0, 1, N, 1000 + j)
        jobs.append((int(ms[j]), int(inv[j]), etas[j] if coset[j] else None, d_src, zkgpu.DeviceArray(4 * N)))
        plain_out.append(zkgpu.DeviceArray(4 * N))
    gens = {m: zkgpu.fft_generator(0, m) for m in range(10, 23)}

    def plain():
        for (m, iv, eta, s, _), t in zip(jobs, plain_out):
            zkgpu.ntt_device(0, m, gens[m], s, t, inverse=bool(iv), eta=eta)
    plain()  # warm-up (builds the shared twiddle table)
    zkgpu.synchronize()
    t0 = time.perf_counter()
    plain()
    zkgpu.synchronize()
    t_plain = time.perf_counter() - t0
    # the first plan of the process also pays HIP's one-time costs (a new hardware queue, the
    # first graph instantiation); a later plan is what a long-lived caller pays per batch
    t0 = time.perf_counter()
    plan = zkgpu.NttBatch(0, jobs)
    t_first_create = time.perf_counter() - t0
    plan.run()
    reps = 3
    t0 = time.perf_counter()
    dev_ms = [plan.run() for _ in range(reps)]
    t_batch = (time.perf_counter() - t0) / reps
    same = all(bool((j[4].to_host() == p.to_host()).all()) for j, p in zip(jobs, plain_out))
    plan.close()
    # one shot: a fresh plan built and replayed once, against the 256 plain calls
    t0 = time.perf_counter()
    plan = zkgpu.NttBatch(0, jobs)
    t_create = time.perf_counter() - t0
    plan.run()
    t_oneshot = time.perf_counter() - t0
    plan.close()
    bytes_moved = 64 * total
    return {"ntt_batch_jobs": njobs, "ntt_batch_elements": int(total),
            "ntt_batch_ms": round(t_batch * 1e3, 3), "ntt_batch_device_ms": round(min(dev_ms), 3),
            "ntt_plain_ms": round(t_plain * 1e3, 3), "ntt_batch_create_ms": round(t_create * 1e3, 2),
            "ntt_batch_first_create_ms": round(t_first_create * 1e3, 1),
            "ntt_batch_oneshot_ms": round(t_oneshot * 1e3, 3),
            "ntt_batch_oneshot_vs_plain": round(t_plain / t_oneshot, 3),
            "ntt_batch_GBps": round(bytes_moved / t_batch / 1e9, 1), "ntt_plain_GBps": round(bytes_moved / t_plain / 1e9, 1),
            "ntt_batched_vs_plain": round(t_plain / t_batch, 3), "ntt_batch_bit_identical_to_plain": same}


def bench_learned_msm(njobs=24, seed=3, nstreams=4):
    """The uneven MSM mix: BN254 and BLS12-381 jobs, log2 n uniform in [14, 22] (seeded), scalars
    uniform in [0, r), uniform 64-bit, or sparse (80% in {0, 1}).  Plain = one MSM at a time with
    the analytic window on the declared width; learned = scalar statistics + the committed cost
    model's schedule, the jobs spread over concurrent streams.  Outputs must agree bit for bit."""
    import costmodel
    stem = costmodel.DEFAULT_MODEL
    if not Path(str(stem) + ".json").exists():
        return {"learned_msm": "no trained cost model under zikkurat-algebra_amd/tuned"}
    net, man = costmodel.load(stem)
    jobs = costmodel.mix_jobs([0, 1], njobs, seed)
    pts = costmodel.PointSets(max(n for _, n, _, _ in jobs))
    dS = []
    for c, n, d, sd in jobs:
        dS.append(zkgpu.DeviceArray.from_host(costmodel.scalars(c, n, d, np.random.default_rng(sd))))
    curves = [c for c, _, _, _ in jobs]
    ns = [n for _, n, _, _ in jobs]
    dP = [pts.d[c] for c in curves]

    def plain():
        return [zkgpu.msm_device(c, n, s, p) for c, n, s, p in zip(curves, ns, dS, dP)]

    def timed(fn, reps=3):
        fn()
        best, out = None, None
        for _ in range(reps):
            zkgpu.synchronize()
            t0 = time.perf_counter()
            out = fn()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        return best, out

    t_plain, ref = timed(plain)
    net.install()
    t_learned_seq, out_seq = timed(lambda: [zkgpu.msm_device(c, n, s, p, policy=zkgpu.SCHED_LEARNED)
                                            for c, n, s, p in zip(curves, ns, dS, dP)])
    t_plain_batch, out_pb = timed(lambda: zkgpu.msm_batch(curves, ns, dS, dP, policy=zkgpu.SCHED_PLAIN,
                                                          nstreams=nstreams)[0])
    t_learned_batch, out_lb = timed(lambda: zkgpu.msm_batch(curves, ns, dS, dP, policy=zkgpu.SCHED_LEARNED,
                                                            nstreams=nstreams)[0])
    # the captured form: one hipGraph of every job's pipeline, replayed (plan built once)
    graphs = {}
    for name, pol in (("plain", zkgpu.SCHED_PLAIN), ("learned", zkgpu.SCHED_LEARNED)):
        zkgpu.synchronize()
        t0 = time.perf_counter()
        plan = zkgpu.MsmBatch(curves, ns, dS, dP, policy=pol, nstreams=nstreams)
        t_create = time.perf_counter() - t0
        t_run, out = timed(lambda: plan.run()[0])
        plan.close()
        graphs[name] = (t_create, t_run, out)
    zkgpu.Mlp.uninstall()
    same = all((a == b).all() for o in (out_seq, out_pb, out_lb, graphs["plain"][2], graphs["learned"][2])
               for a, b in zip(o, ref))
    pts_total = sum(ns)
    t_lg = graphs["learned"][1]
    return {"learned_msm_jobs": njobs, "learned_msm_points": pts_total,
            "learned_msm_mix": "bn254+bls12-381, 2^14..2^22 log-uniform, uniform/u64/sparse scalars",
            "learned_msm_plain_ms": round(t_plain * 1e3, 3),
            "learned_msm_learned_seq_ms": round(t_learned_seq * 1e3, 3),
            "learned_msm_plain_batched_ms": round(t_plain_batch * 1e3, 3),
            "learned_msm_learned_batched_ms": round(t_learned_batch * 1e3, 3),
            "learned_msm_batched_learned_vs_plain": round(t_plain / t_learned_batch, 3),
            "learned_msm_graph_plain_ms": round(graphs["plain"][1] * 1e3, 3),
            "learned_msm_graph_learned_ms": round(t_lg * 1e3, 3),
            "learned_msm_graph_create_ms": round(graphs["learned"][0] * 1e3, 1),
            "learned_msm_graph_learned_vs_plain": round(t_plain / t_lg, 3),
            "learned_msm_graph_learned_mpts": round(pts_total / t_lg / 1e6, 1),
            "learned_msm_plain_mpts": round(pts_total / t_plain / 1e6, 1),
            "learned_msm_learned_batched_mpts": round(pts_total / t_learned_batch / 1e6, 1),
            "learned_msm_streams": nstreams, "learned_msm_model": Path(str(stem)).name,
            "learned_msm_bit_identical": bool(same)}


def _g1_fft_times(curve, m, seed):
    """Forward and inverse G1 FFT of 2^m points k_i G given as projective (x, y, 1), on the 2^m
    subgroup generator: (forward s, inverse s, round trip bit-exact, the input words)."""
    N = 1 << m
    L = 4 if curve == 0 else 6
    R, Pm = (BN_R, BN_P) if curve == 0 else (BLS_R, BLS_P)
    rng = np.random.default_rng(seed)
    K = uniform_scalars(rng, N, R)
    aff = zkgpu.scalar_mul_gen(curve, zkgpu.DeviceArray.from_host(K), N).to_host((N, 2 * L))
    one = np.array([((1 << (64 * L)) % Pm >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(L)], np.uint64)
    P = np.concatenate([aff, np.tile(one, (N, 1))], axis=1)
    g = zkgpu.fft_generator(curve, m)
    dP, dY, dB = zkgpu.DeviceArray.from_host(P), zkgpu.DeviceArray(P.size), zkgpu.DeviceArray(P.size)
    tf = time_steps(lambda: zkgpu.g1_fft_device(curve, m, g, dP, dY), 2, 1)
    ti = time_steps(lambda: zkgpu.g1_fft_device(curve, m, g, dY, dB, inverse=True), 2, 1)
    return tf, ti, bool((dB.to_host(P.shape) == P).all()), P


def bench_g1_fft(m=12, cpu_m=7, seed=6, big_m=16):
    """BN254 G1 FFT (bn128_G1_proj_fft_forward / _inverse) of 2^m points; the inverse of the
    forward transform must return the input bit for bit.  Also BN254 at 2^big_m (throughput-bound
    stage plans) and BLS12-381 at 2^m.  CPU: the reference's own C on 2^cpu_m of the points."""
    N = 1 << m
    tf, ti, ok, P = _g1_fft_times(0, m, seed)
    out = {"g1_fft_bn254_log": m, "g1_fft_fwd_ms": round(tf * 1e3, 2), "g1_fft_inv_ms": round(ti * 1e3, 2),
           "g1_fft_fwd_mpts": round(N / tf / 1e6, 4), "g1_fft_roundtrip_ok": ok}
    tfb, tib, okb, _ = _g1_fft_times(0, big_m, seed + 1)
    tfl, til, okl, _ = _g1_fft_times(1, m, seed + 2)
    out.update({f"g1_fft_bn254_2p{big_m}_fwd_ms": round(tfb * 1e3, 2), f"g1_fft_bn254_2p{big_m}_inv_ms": round(tib * 1e3, 2),
                f"g1_fft_bls12_381_2p{m}_fwd_ms": round(tfl * 1e3, 2), f"g1_fft_bls12_381_2p{m}_inv_ms": round(til * 1e3, 2),
                "g1_fft_more_roundtrips_ok": okb and okl})
    ref_so = REPO / "oracle" / "_ref" / "libzkref.so"
    if ref_so.exists():
        lib = ctypes.CDLL(str(ref_so))
        n = 1 << cpu_m
        u64p = ctypes.POINTER(ctypes.c_uint64)
        src = np.ascontiguousarray(P[:n])
        tgt = np.zeros_like(src)
        gc = zkgpu.fft_generator(0, cpu_m)
        t0 = time.perf_counter()
        lib.bn128_G1_proj_fft_forward(cpu_m, gc.ctypes.data_as(u64p), src.ctypes.data_as(u64p), tgt.ctypes.data_as(u64p))
        dt = time.perf_counter() - t0
        gpu = zkgpu.g1_fft(0, cpu_m, gc, src)
        out["g1_fft_cpu_ref"] = {"log": cpu_m, "ms": round(dt * 1e3, 1), "cores": 1,
                                 "mpts": round(n / dt / 1e6, 6), "matches_gpu": bool((gpu == tgt).all())}
    return out


def _ref_lib():
    ref_so = REPO / "oracle" / "_ref" / "libzkref.so"
    return ctypes.CDLL(str(ref_so)) if ref_so.exists() else None


def bench_next(seed=7, logn=20, cpu_log=14):
    """The rows around the MSM, BN254, device-resident inputs: batch projective -> affine of 2^logn
    points; a binary MSM of 2^logn points (half selected) and 64 binary rows over 2^16 points;
    a KZG opening of a 2^logn-coefficient polynomial against a prepared SRS (evaluation, quotient
    by x - z, MSM), with the evaluation and the quotient alone at 2^22; the SRS written to and
    read back from a flat file.  CPU: the reference's C on 2^cpu_log-sized samples, one core."""
    u64p = ctypes.POINTER(ctypes.c_uint64)
    rng = np.random.default_rng(seed)
    N = 1 << logn
    out = {}
    # batch_to_affine: points (x z, y z, z) from k_i G and random z
    d_k = zkgpu.fill_random(0, 1, N, seed)
    aff = zkgpu.scalar_mul_gen(0, d_k, N).to_host((N, 8))
    Z = zkgpu.fill_random(0, 0, N, seed + 1)
    d_x, d_y = zkgpu.DeviceArray.from_host(np.ascontiguousarray(aff[:, :4])), zkgpu.DeviceArray.from_host(
        np.ascontiguousarray(aff[:, 4:]))
    d_t, d_u = zkgpu.DeviceArray(4 * N), zkgpu.DeviceArray(4 * N)
    zkgpu.field_op(0, 0, zkgpu.FOP["mul"], d_x, Z, d_t, N)
    zkgpu.field_op(0, 0, zkgpu.FOP["mul"], d_y, Z, d_u, N)
    P = np.concatenate([d_t.to_host((N, 4)), d_u.to_host((N, 4)), Z.to_host((N, 4))], axis=1)
    d_P, d_A = zkgpu.DeviceArray.from_host(P), zkgpu.DeviceArray(8 * N)
    tb = time_steps(lambda: zkgpu.batch_to_affine_device(0, N, d_P, d_A), 5, 1)
    ok = bool((d_A.to_host((N, 8)) == aff).all())
    out.update({"batch_to_affine_bn254_log": logn, "batch_to_affine_ms": round(tb * 1e3, 3),
                "batch_to_affine_mpts": round(N / tb / 1e6, 1), "batch_to_affine_ok": ok})
    # binary MSM and 64 binary rows
    d_aff = zkgpu.DeviceArray.from_host(aff)
    prep = zkgpu.PreparedPoints(0, d_aff, N)
    bits = (rng.random(N) < 0.5).astype(np.uint8)
    d_bits = zkgpu.DeviceArray.from_bytes(bits)
    tbin = time_steps(lambda: zkgpu.msm_binary_device(0, N, d_bits, prep), 5, 1)
    rows = (rng.random((64, 1 << 16)) < 0.5).astype(np.uint8)
    d_rows = zkgpu.DeviceArray.from_bytes(rows)
    prep16 = zkgpu.PreparedPoints(0, d_aff, 1 << 16)
    tmul = time_steps(lambda: zkgpu.multi_msm_binary_device(0, 64, 1 << 16, d_rows, prep16), 3, 1)
    out.update({"binary_msm_ms": round(tbin * 1e3, 3), "binary_msm_mpts": round(N / tbin / 1e6, 1),
                "multi_binary_64x2p16_ms": round(tmul * 1e3, 3)})
    # KZG opening against the SRS tau^i G (prepared), coefficients resident
    tau = int(uniform_scalars(rng, 1, BN_R)[0][0])
    d_srs = zkgpu.srs_generate(0, N, tau)
    srs = zkgpu.PreparedPoints(0, d_srs, N)
    d_c = zkgpu.fill_random(0, 1, N, seed + 2)
    z = uniform_scalars(rng, 1, BN_R)[0]
    topen = time_steps(lambda: zkgpu.kzg_open(0, N, d_c, z, srs), 5, 1)
    out.update({"kzg_open_bn254_log": logn, "kzg_open_ms": round(topen * 1e3, 3)})
    M = 1 << 22
    d_big = zkgpu.fill_random(0, 1, M, seed + 3)
    d_q = zkgpu.DeviceArray(4 * M)
    tq = time_steps(lambda: zkgpu.div_by_vanishing_device(0, M, d_big, 1, z, M - 1, d_q), 5, 1)
    te = time_steps(lambda: zkgpu.poly_eval_device(0, M, d_big, z), 5, 1)
    out.update({"quot_by_x_minus_z_2p22_ms": round(tq * 1e3, 3),
                "quot_by_x_minus_z_2p22_GBs": round(2 * 32 * M / tq / 1e9, 1),
                "poly_eval_2p22_ms": round(te * 1e3, 3), "poly_eval_2p22_GBs": round(32 * M / te / 1e9, 1)})
    # SRS as a flat file
    path = Path(os.environ.get("TMPDIR", "/tmp")) / f"zkg_srs_{os.getpid()}.bin"
    t0 = time.perf_counter()
    zkgpu.flat_write(path, d_srs, 8, N)
    tw = time.perf_counter() - t0
    t0 = time.perf_counter()
    back, _ = zkgpu.srs_load(0, path, prepare=False)
    tr = time.perf_counter() - t0
    same = bool((back.to_host((N, 8)) == d_srs.to_host((N, 8))).all())
    path.unlink()
    out.update({"srs_file_write_GBs": round(64 * N / tw / 1e9, 2), "srs_file_read_GBs": round(64 * N / tr / 1e9, 2),
                "srs_file_roundtrip_ok": same})
    lib = _ref_lib()
    if lib is not None:
        n = 1 << cpu_log
        cpu = {"log": cpu_log, "cores": 1}
        src = np.ascontiguousarray(P[:n])
        tgt = np.zeros((n, 8), np.uint64)
        t0 = time.perf_counter()
        lib.bn128_G1_proj_batch_to_affine(n, src.ctypes.data_as(u64p), tgt.ctypes.data_as(u64p))
        cpu["batch_to_affine_mpts"] = round(n / (time.perf_counter() - t0) / 1e6, 4)
        cpu["batch_to_affine_matches"] = bool((tgt == aff[:n]).all())
        b = np.ascontiguousarray(bits[:n])
        res = np.zeros(12, np.uint64)
        t0 = time.perf_counter()
        lib.bn128_G1_proj_MSM_binary_coeff_proj_out(n, b.ctypes.data_as(ctypes.c_void_p),
                                                    np.ascontiguousarray(aff[:n]).ctypes.data_as(u64p),
                                                    res.ctypes.data_as(u64p))
        cpu["binary_msm_mpts"] = round(n / (time.perf_counter() - t0) / 1e6, 4)
        # KZG opening on the CPU: eval_at, quot_by_vanishing of p - p(z) by x - z, MSM
        c = d_c.to_host((N, 4))[:n].copy()
        sr = d_srs.to_host((N, 8))[:n].copy()
        v = np.zeros(4, np.uint64)
        q = np.zeros((n, 4), np.uint64)
        proof = np.zeros(8, np.uint64)
        zc = np.ascontiguousarray(z)
        t0 = time.perf_counter()
        lib.bn128_poly_mont_eval_at(n, c.ctypes.data_as(u64p), zc.ctypes.data_as(u64p), v.ctypes.data_as(u64p))
        lib.bn128_poly_mont_quot_by_vanishing(n, c.ctypes.data_as(u64p), 1, zc.ctypes.data_as(u64p), n,
                                              q.ctypes.data_as(u64p))
        lib.bn128_G1_proj_MSM_mont_coeff_affine_out(n - 1, q.ctypes.data_as(u64p), sr.ctypes.data_as(u64p),
                                                    proof.ctypes.data_as(u64p), 4)
        cpu["kzg_open_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        d_c14 = zkgpu.DeviceArray.from_host(c)
        gv, gp = zkgpu.kzg_open(0, n, d_c14, z, d_srs)
        cpu["kzg_open_matches"] = bool((gv == v).all() and (gp == proof).all())
        out["next_rows_cpu_ref"] = cpu
    for x in (prep, prep16, srs):
        x.free()
    return out


def bench_serve(njobs=64, qps_list=(100, 400), seed=5):
    """KZG commits on BLS12-381: iNTT of n evaluations (n = 2^14..2^21, log-uniform) then an MSM
    against a resident SRS of 2^21 points tau^i G; Poisson arrivals at each stated QPS; served
    (4 workers, continuous batching) against the plain path (one job at a time), and the served
    form with the optional size rule (no job starts beside one of 2^21 points).  Before each
    trace a warm-up runs four jobs of every size at once, so that every form finds its device
    buffers of every size in the allocator's cache."""
    rng = np.random.default_rng(seed)
    srs_n = 1 << 21
    tau = int(uniform_scalars(rng, 1, BLS_R)[0][0])
    d_srs = zkgpu.srs_generate(1, srs_n, tau)
    srs = zkgpu.PreparedPoints(1, d_srs, srs_n)
    logns = [int(x) for x in rng.integers(14, 22, size=njobs)]
    values = [uniform_scalars(rng, 1 << m, BLS_R) for m in logns]  # canonical words = valid Montgomery elements
    out = {"serve_jobs": njobs, "serve_srs_log": 21, "serve_sizes": "2^14..2^21 log-uniform"}
    warm = [m for m in range(21, 13, -1) for _ in range(4)]
    warm_vals = [uniform_scalars(rng, 1 << m, BLS_R) for m in range(21, 13, -1) for _ in range(4)]
    for qps in qps_list:
        arrivals = list(np.cumsum(rng.exponential(1.0 / qps, size=njobs)) - 0.0)
        res = {}
        for name, workers, excl in (("plain", 1, 0), ("served", 4, 0), ("served_size_rule_2p21", 4, 21)):
            srv = zkgpu.Server(1, srs, 48 << 30, workers)
            srv.set_exclusive_log(excl)
            srv.run_trace(warm, [0.0] * len(warm), warm_vals)  # warm-up (twiddles, allocator)
            outs, lat, wall = srv.run_trace(logns, arrivals, values)
            peak, conc = srv.stats()
            srv.close()
            lat = np.array(lat)
            res[name] = (outs, {"p50_ms": round(float(np.percentile(lat, 50)), 2),
                                "p99_ms": round(float(np.percentile(lat, 99)), 2),
                                "commits_per_s": round(njobs / wall, 1), "max_concurrent": conc,
                                "peak_admitted_GB": round(peak / 2**30, 2)})
        same = all((a == b_).all() for a, b_ in zip(res["plain"][0], res["served"][0]))
        same = same and all((a == b_).all() for a, b_ in zip(res["plain"][0], res["served_size_rule_2p21"][0]))
        out[f"serve_qps{qps}"] = {"plain": res["plain"][1], "served": res["served"][1],
                                  "served_size_rule_2p21": res["served_size_rule_2p21"][1], "bit_identical": same}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--logn", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--cpu-sample-log", type=int, default=18)
    ap.add_argument("--leg", choices=["all", "msm", "ntt", "copy", "nttbatch", "learned", "g1fft", "next"],
                    default="all",
                    help="profiling runs: only the BN254 MSM, the 2^22 NTT pair, the HBM copy kernel, the "
                         "batched NTTs, the learned MSM mix, the G1 FFT or the rows around the MSM "
                         "(conversions, binary MSMs, KZG opening, SRS files)")
    a = ap.parse_args()
    if a.leg == "msm":
        a.no_extra = a.no_cpu = True
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    def log(*x):
        print(*x, file=sys.stderr, flush=True)

    cus, hbm = zkgpu.device_info()
    log(f"device: {cus} CUs, {hbm / 2**30:.0f} GiB")
    import costmodel
    ntt_entries = costmodel.install_ntt_tuning()  # autotuned NTT schedules, when saved
    log(f"ntt tuning entries installed: {ntt_entries}")
    if a.leg == "copy":
        bw = zkgpu.microbench(zkgpu.MB_COPY)
        log(f"copy kernel: 1 GiB read + 1 GiB written per launch, {bw / 1e12:.2f} TB/s")
        return
    if a.leg in ("ntt", "nttbatch", "learned", "g1fft", "next"):
        fn = {"ntt": lambda: bench_ntt_pair(a.steps, a.warmup), "nttbatch": bench_ntt_batch,
              "learned": bench_learned_msm, "g1fft": bench_g1_fft, "next": bench_next}[a.leg]
        for k, v in fn().items():
            log(f"{k}: {v}")
        return
    peak_imul = zkgpu.microbench(zkgpu.MB_IMUL)
    log(f"measured v_mad_u64_u32 peak: {peak_imul / 1e12:.2f} T/s")

    # ---- headline: BN254 MSM 2^20
    bn = MsmCase(0, a.logn, 20260101)
    log("inputs ready")
    res = bn.run()
    ok_bn = decode_affine(res, 0) == bn.expected()
    log(f"bn254 2^{a.logn} closed form: {ok_bn}")
    t = time_steps(bn.run, a.steps, a.warmup)
    value = bn.n / t / 1e6
    # per-phase HIP events in calls of their own (their records idle the stream between phases)
    zkgpu.set_phase_timing(True)
    t_acc = kernel_ms(bn.run, 5, 1)
    phases = zkgpu.last_timing()[1]
    zkgpu.set_phase_timing(False)
    log(f"bn254 msm: {t * 1e3:.3f} ms/step, {value:.1f} Mpts/s, phases(ms) {['%.3f' % x for x in phases]}")
    # algorithmic work of the accumulation kernel: one mixed XYZZ add (8M + 2S) per nonzero digit
    st = zkgpu.last_msm_stats()
    c_used, entries = st["window"], st["entries"]
    acc_mul = entries * 10 * m32(8)
    achieved = acc_mul / (t_acc * 1e-3)
    roof = {"bound": "valu-imul", "achieved": round(achieved / 1e12, 3),
            "peak": round(PEAK_IMUL_PUBLISHED / 1e12, 3), "unit": "Tmul32/s",
            "frac": round(achieved / PEAK_IMUL_PUBLISHED, 4), "traffic": None,
            "peak_note": "MI355X 157.3 TFLOP/s fp32 vector / 2 flops per FMA / 2 packed lanes = 39.3e12 "
                         "32-bit lane operations/s; v_mad_u64_u32 issues at that rate",
            "peak_measured": round(peak_imul / 1e12, 3), "frac_of_measured": round(achieved / peak_imul, 4),
            "kernel": "msm_accum_kernel<Bn254G1>", "kernel_ms": round(t_acc, 4),
            "units": f"{entries} mixed XYZZ adds x {10 * m32(8)} u32 multiplies"}
    traffic = REPO / "profiles" / "msm_accum_bn254_2p20_traffic.json"
    if traffic.exists():
        roof["traffic"] = json.loads(traffic.read_text()).get("bytes_per_launch")
    whole_frac = a_msm(bn.n, 0) / t / PEAK_IMUL_PUBLISHED

    extra = {"ntt_tuning_entries": ntt_entries, "bn254_msm_2p20_imul_roofline_frac": round(whole_frac, 4),
             "bn254_msm_2p20_imul_frac_of_measured_peak": round(a_msm(bn.n, 0) / t / peak_imul, 4),
             "bn254_msm_closed_form_ok": ok_bn,
             "bn254_msm_phases_ms": [round(x, 4) for x in phases]}
    if not a.no_extra:
        bls = MsmCase(1, a.logn, 20260202)
        rb = bls.run()
        extra["bls12_381_msm_closed_form_ok"] = decode_affine(rb, 1) == bls.expected()
        tb = time_steps(bls.run, max(3, a.steps // 2), 1)
        extra["bls12_381_msm_2p20_mpts"] = round(bls.n / tb / 1e6, 2)
        extra["bls12_381_msm_2p20_ms"] = round(tb * 1e3, 3)
        extra["bls12_381_msm_2p20_imul_roofline_frac"] = round(a_msm(bls.n, 1) / tb / PEAK_IMUL_PUBLISHED, 4)
        extra["bls12_381_msm_2p20_imul_frac_of_measured_peak"] = round(a_msm(bls.n, 1) / tb / peak_imul, 4)
        log(f"bls msm: {tb * 1e3:.3f} ms, closed form {extra['bls12_381_msm_closed_form_ok']}")
        del bls
        extra.update(bench_ntt_pair(a.steps, a.warmup, peak_imul))
        log(f"ntt 2^22: fwd {extra['bn254_ntt_2p22_fwd_ms']} ms inv {extra['bn254_ntt_2p22_inv_ms']} ms "
            f"roundtrip {extra['bn254_ntt_2p22_roundtrip_ok']}")
        extra.update(bench_ntt_batch())
        log(f"ntt batch: {extra['ntt_batch_ms']} ms batched vs {extra['ntt_plain_ms']} ms plain, identical {extra['ntt_batch_bit_identical_to_plain']}")
        extra.update(bench_g1_fft())
        log(f"g1 fft: {[(k, v) for k, v in extra.items() if k.startswith('g1_fft')]}")
        extra.update(bench_next())
        log(f"next rows: {[(k, v) for k, v in extra.items() if k.split('_')[0] in ('batch', 'binary', 'multi', 'kzg', 'quot', 'poly', 'srs', 'next')]}")
        extra.update(bench_learned_msm())
        log(f"learned msm mix: {[(k, v) for k, v in extra.items() if k.startswith('learned_msm')]}")
        extra.update(bench_serve())
        log(f"serve: {[(k, v) for k, v in extra.items() if k.startswith('serve_qps')]}")

    line = {"metric": "G1 MSM Mpoints/s @2^20 BN254+BLS12-381, NTT GB/s; batched vs plain, bit-exact",
            "value": round(value, 2), "unit": "Mpoints/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
            "ms_per_step": round(t * 1e3, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "u32", "data": "synthetic (seeded uniform scalars in [0,r), points k_i*G)",
            "config": {"workload": f"bn254-g1-msm-2^{a.logn}", "curve": "bn254", "n": bn.n,
                       "scalars": "standard form, uniform [0,r)", "points": "affine Montgomery k_i*G",
                       "window": c_used, "parallelism": "single-gpu"},
            "roofline": roof}
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(bn, a.cpu_sample_log)
        log(f"cpu baseline: {line['cpu_baseline']}")
    line["extra"] = extra
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    if "--cpu-worker" in sys.argv:
        cpu_worker(sys.argv[sys.argv.index("--cpu-worker") + 1])
    else:
        main()
