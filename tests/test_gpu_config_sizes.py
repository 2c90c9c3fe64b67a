"""Parity at the workload sizes.

* The reference's own outputs at full size (tests/golden/manifest.json "large", written by the
  reference C built from its sources, tests/golden/gen_golden.py) against the HIP path called
  through the reference C symbols:
  - BN254 G1 MSM at 2^14, 2^16 (the single-MSM workload) and 2^20, BLS12-381 at 2^14 and 2^20,
    `<c>_G1_proj_MSM_std_coeff_affine_out` (bn128_G1_proj.c:656-660);
  - Fr NTT forward and inverse digests at 2^16, 2^20 and 2^22 for both curves,
    `<c>_poly_mont_ntt_{forward,inverse}` (bn128_poly_mont.c:452-473, 601-622).
* Closed forms where the reference's CPU time rules out a stored output:
  - learned schedules (the committed cost model installed) at 2^18, 2^20 and 2^22 over the three
    scalar distributions, both curves: sum_i s_i (k_i G) = (sum_i s_i k_i mod r) G;
  - a served KZG trace against a 2^21-point BLS12-381 SRS with jobs of 2^14..2^21 points:
    commit(values) = f(tau) G for the f interpolating the values (examples/KZG.hs:91-97).
* 256 batched NTTs of 2^10..2^22 points (mixed direction, half on cosets) against 256 plain calls.
* The FFT domain: the generator of the 2^m subgroup is gen_max^(2^(M - m)) (Class/FFT.hs:60-68),
  gen_max = 5^((r-1)/2^28) for BN254 and 7^((r-1)/2^32) for BLS12-381 (Generate.hs:143-144).
"""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from zkmath import BLS, BN, seeded_fr_vector, to_limbs

pytestmark = pytest.mark.gpu
MANIFEST = json.loads((Path(__file__).resolve().parent / "golden" / "manifest.json").read_text())
CURVE_OF = {"bn128": BN, "bls12_381": BLS}


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.uint64).tobytes()).hexdigest()


def rows_to_ints(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return [int.from_bytes(r.tobytes(), "little") for r in a]


def gen_words(C, m):
    return np.array(to_limbs(C.fr_to_mont(C.fft_root(m)), 4), np.uint64)


def _large(kind):
    return [e for e in MANIFEST["large"] if e["kind"] == kind]


def _eid(e):
    return f"{e['curve']}-2^{e.get('logn', e.get('m'))}"


# ---------------------------------------------------------------------------------------
# reference outputs at full size

@pytest.mark.parametrize("e", _large("msm_std_affine"), ids=_eid)
def test_msm_reference_output(zk, e):
    """The reference's MSM result over points k_i G (k_i, s_i seeded uniform in [0, r)); the
    points are rebuilt on the device and pinned by the digest of the reference's inputs."""
    C = CURVE_OF[e["curve"]]
    n = 1 << e["logn"]
    K = seeded_fr_vector(e["kseed"], n, C.r)
    S = seeded_fr_vector(e["sseed"], n, C.r)
    dP = zk.scalar_mul_gen(C.id, zk.DeviceArray.from_host(K), n)
    P = dP.to_host((n, 2 * C.lp))
    assert sha(P) == e["points_sha256"]
    print(f"msm reference output {e['curve']} n=2^{e['logn']}", flush=True)
    got = zk.msm_ref_symbol(C.id, "proj", "std", "affine", S, P)
    assert [int(x) for x in got] == e["result"]
    if e["logn"] <= 16:
        exp = C.words_to_affine(np.array(e["result"], np.uint64))
        assert C.words_to_affine(zk.msm_ref_symbol(C.id, "jac", "std", "affine", S, P)) == exp
        assert C.proj_words_to_affine(zk.msm_slow_reference(C.id, "proj", S, P)) == exp
        assert C.jac_words_to_affine(zk.msm_slow_reference(C.id, "jac", S, P, header_spelling=True)) == exp


@pytest.mark.parametrize("e", _large("ntt_fwd"), ids=_eid)
def test_ntt_reference_digest(zk, e):
    """The reference NTT and iNTT outputs (SHA-256 of the words) over a seeded vector taken as
    Montgomery words; the generator of order 2^m from Fr's 2-adic generator."""
    C = CURVE_OF[e["curve"]]
    m = e["m"]
    X = seeded_fr_vector(e["seed"], 1 << m, C.r)
    g = gen_words(C, m)
    print(f"ntt reference digest {e['curve']} m={m}", flush=True)
    y = zk.ntt(C.id, m, g, X)
    assert sha(y) == e["sha256"]
    assert sha(zk.intt(C.id, m, g, X)) == e["inv_sha256"]
    assert (zk.intt(C.id, m, g, y) == X).all()


def test_manifest_holds_the_workload_sizes():
    """The digests and results above cover the workload sizes (the test ids come from the manifest,
    so a missing entry would silently drop a case)."""
    have = {(e["curve"], e["kind"], e.get("m", e.get("logn"))) for e in MANIFEST["large"]}
    for need in [("bn128", "msm_std_affine", 16), ("bn128", "msm_std_affine", 20),
                 ("bls12_381", "msm_std_affine", 20), ("bn128", "ntt_fwd", 22), ("bls12_381", "ntt_fwd", 22)]:
        assert need in have, need


# ---------------------------------------------------------------------------------------
# the FFT domain

@pytest.mark.parametrize("C,M,base", [(BN, 28, 5), (BLS, 32, 7)], ids=["bn254", "bls12-381"])
def test_fft_generator_every_size(zk, C, M, base):
    assert (C.r - 1) % (1 << M) == 0 and ((C.r - 1) >> M) % 2 == 1
    gmax = pow(base, (C.r - 1) >> M, C.r)
    assert gmax == C.fft_gen
    for m in range(M + 1):
        g = pow(gmax, 1 << (M - m), C.r)
        assert pow(g, 1 << m, C.r) == 1
        assert m == 0 or pow(g, 1 << (m - 1), C.r) == C.r - 1  # order exactly 2^m
        assert [int(w) for w in zk.fft_generator(C.id, m)] == to_limbs(C.fr_to_mont(g), 4), m
    with pytest.raises(zk.ZkgError):
        zk.fft_generator(C.id, M + 1)


# ---------------------------------------------------------------------------------------
# learned schedules with the committed cost model

@pytest.mark.parametrize("C", [BN, BLS], ids=["bn254", "bls12-381"])
def test_learned_policy_closed_form_at_size(zk, C):
    import costmodel
    net, man = costmodel.load()
    net.install()
    try:
        nmax = 1 << 22
        K = seeded_fr_vector(31000 + C.id, nmax, C.r)
        dP = zk.scalar_mul_gen(C.id, zk.DeviceArray.from_host(K), nmax)
        ks = rows_to_ints(K)
        sizes = (1 << 18, 1 << 20, 1 << 22)
        for dist in ("uniform", "u64", "sparse"):
            S = costmodel.scalars(C.id, nmax, dist, np.random.default_rng(32000 + 10 * C.id + len(dist)))
            dS = zk.DeviceArray.from_host(S)
            ss = rows_to_ints(S)
            acc, done = 0, 0
            for n in sizes:
                acc = (acc + sum(s * k for s, k in zip(ss[done:n], ks[done:n]))) % C.r
                done = n
                print(f"learned {C.name} n={n} {dist}", flush=True)
                got = zk.msm_device(C.id, n, dS, dP, policy=zk.SCHED_LEARNED)
                assert C.words_to_affine(got) == C.mul(acc, C.g), (n, dist)
                assert (zk.msm_device(C.id, n, dS, dP) == got).all(), (n, dist)
            dS.free()
        dP.free()
    finally:
        zk.Mlp.uninstall()
        net.close()


# ---------------------------------------------------------------------------------------
# serving at the workload size

def _batch_inverse(xs, r):
    pre, acc = [], 1
    for x in xs:
        pre.append(acc)
        acc = acc * x % r
    inv = pow(acc, -1, r)
    out = [0] * len(xs)
    for i in range(len(xs) - 1, -1, -1):
        out[i] = pre[i] * inv % r
        inv = inv * xs[i] % r
    return out


def barycentric(values, logn, tau, C):
    """f(tau) for the f of degree < n with f(w^k) = values[k] (standard form), w = the size-n root:
    f(tau) = (tau^n - 1) / n * sum_k values[k] w^k / (tau - w^k)."""
    n, r = 1 << logn, C.r
    w = C.fft_root(logn)
    ws, wk = [], 1
    for _ in range(n):
        ws.append(wk)
        wk = wk * w % r
    inv = _batch_inverse([(tau - x) % r for x in ws], r)
    acc = sum(v * x % r * i for v, x, i in zip(values, ws, inv)) % r
    return (pow(tau, n, r) - 1) * pow(n, -1, r) % r * acc % r


def test_served_trace_at_srs_2p21(zk):
    """BLS12-381 KZG commits against a resident 2^21-point SRS, jobs of 2^14..2^21 points: every
    served result equals the same job alone, and the largest job equals f(tau) G."""
    C = BLS
    tau = 0x1234567890ABCDEF1234567890ABCDEF % C.r
    srs_n = 1 << 21
    d_srs = zk.srs_generate(C.id, srs_n, tau)
    srs = zk.PreparedPoints(C.id, d_srs, srs_n)
    rng = np.random.default_rng(2121)
    lognes = [21, 14] + [int(x) for x in rng.integers(14, 22, size=8)]
    vals = [seeded_fr_vector(33000 + k, 1 << m, C.r) for k, m in enumerate(lognes)]  # Montgomery words
    alone = zk.Server(C.id, srs, 1 << 36, 1)
    ref = []
    for m, v in zip(lognes, vals):
        print(f"serve alone 2^{m}", flush=True)
        _, out = alone.submit(m, v)
        alone.drain()
        ref.append(out.copy())
    alone.close()
    mix = zk.Server(C.id, srs, 1 << 36, 4)
    arrivals = sorted(float(x) for x in rng.uniform(0.0, 0.02, size=len(lognes)))
    outs, lat, wall = mix.run_trace(lognes, arrivals, vals)
    mix.close()
    assert wall > 0 and all(x > 0 for x in lat)
    for a, b, m in zip(ref, outs, lognes):
        assert (a == b).all(), m
    rinv = pow(1 << 256, -1, C.r)
    std = [x * rinv % C.r for x in rows_to_ints(vals[0])]
    assert C.words_to_affine(outs[0]) == C.mul(barycentric(std, lognes[0], tau, C), C.g)
    srs.free()


# ---------------------------------------------------------------------------------------
# 256 batched NTTs at the workload sizes

def test_ntt_batch_256_jobs_vs_plain(zk):
    """256 BN254 NTTs, log2 n uniform over 10..22, fair-coin direction, half on a coset with a
    seeded shift: one batch plan against 256 plain calls, bit for bit, over two replays."""
    C = BN
    rng = np.random.default_rng(256)
    specs = [(int(rng.integers(10, 23)), int(rng.integers(0, 2)), bool(rng.integers(0, 2))) for _ in range(256)]
    specs[0] = (22, 0, False)
    specs[1] = (22, 1, True)
    jobs, plain = [], []
    for k, (m, inv, coset) in enumerate(specs):
        N = 1 << m
        dX = zk.fill_random(C.id, 1, N, 4000 + k)
        eta = np.array(to_limbs(C.fr_to_mont(int(rng.integers(2, 2**62))), 4), np.uint64) if coset else None
        dY, dP = zk.DeviceArray(4 * N), zk.DeviceArray(4 * N)
        zk.ntt_device(C.id, m, zk.fft_generator(C.id, m), dX, dP, inverse=bool(inv), eta=eta)
        jobs.append((m, inv, eta, dX, dY))
        plain.append(dP)
    print("ntt batch: 256 jobs, %d elements" % sum(1 << s[0] for s in specs), flush=True)
    plan = zk.NttBatch(C.id, jobs)
    for rep in range(2):
        plan.run()
        for k, ((m, _, _), (_, _, _, _, dY), dP) in enumerate(zip(specs, jobs, plain)):
            assert (dY.to_host() == dP.to_host()).all(), (rep, k, m)
    plan.close()
