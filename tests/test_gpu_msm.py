"""GPU MSM parity: bit-exact against the oracle (the reference's bucket method restated) on
small inputs with every exceptional case, and against the closed form
sum_i s_i (k_i G) = (sum_i s_i k_i mod r) G at the larger sizes.

Reference behaviour: lib/cbits/curves/g1/proj/bn128_G1_proj.c:506-673 (proj) and
lib/cbits/curves/g1/jac/bn128_G1_jac.c:555-745 (jac); affine infinity input (0,0)
(bn128_G1_affine.c:57-60), infinity output all-0xFF (bn128_G1_proj.c:132-137).
"""
import numpy as np
import pytest

from cref import Oracle
from zkmath import BN, BLS, ints_to_array, random_scalars, to_limbs, splitmix64

pytestmark = pytest.mark.gpu
CURVES = [BN, BLS]
IDS = ["bn254", "bls12-381"]


def make_points(C, ks):
    """k_i * G via the oracle's C scalar multiplication (affine infinity for k = 0)."""
    return Oracle(C).points_from_scalars(ks)


def edge_case_inputs(C, n, seed):
    """Scalars/points exercising zero digits, r-1, all-ones limbs, duplicates, P and -P,
    affine infinity and a top window straddling a limb."""
    ks = random_scalars(seed, n, C.r)
    ss = random_scalars(seed + 1, n, C.r)
    if n >= 8:
        ss[0] = 0
        ss[1] = C.r - 1
        ss[2] = (1 << 64) - 1
        ss[3] = (1 << 253) + 12345
        ks[4] = ks[5]                 # duplicate point in the same buckets: doubling
        ss[4] = ss[5]
        ks[6] = (C.r - ks[7]) % C.r   # P and -P
        ss[6] = ss[7]
        ks[3] = 0                     # affine infinity (0,0) input
    return ks, ss


@pytest.mark.parametrize("C", CURVES, ids=IDS)
@pytest.mark.parametrize("n", [1, 2, 3, 5, 17, 64, 255, 256, 1000])
def test_msm_vs_oracle(zk, C, n):
    o = Oracle(C)
    ks, ss = edge_case_inputs(C, n, 1000 + n)
    P = make_points(C, ks)
    S = ints_to_array(ss, 4)
    exp = o.msm_affine(S, P)
    assert (zk.msm_std(C.id, S, P) == exp).all()
    # Montgomery scalars (the Haskell msm path, G1/Proj.hs:254)
    Sm = ints_to_array([C.fr_to_mont(s % C.r) for s in ss], 4)
    assert (zk.msm(C.id, Sm, P) == o.msm_affine(Sm, P, mont=True)).all()
    # closed form
    assert C.words_to_affine(exp) == C.mul(sum(s * k for s, k in zip(ss, ks)) % C.r, C.g)


@pytest.mark.parametrize("C", CURVES, ids=IDS)
@pytest.mark.parametrize("n", [9, 4099, 20000])
def test_msm_scalars_above_r(zk, C, n):
    """Standard-form scalars are 256-bit integers, not field elements: values at and above r
    (r, r + 1, 2r - 1, 5r + 7, 2^256 - 1 and uniform 256-bit values) against the oracle, which
    multiplies by the integers themselves, and the closed form (s mod r) summed over k_i G."""
    o = Oracle(C)
    ks = random_scalars(300 + n, n, C.r)
    rng = np.random.default_rng(n)
    ss = [int.from_bytes(rng.bytes(32), "little") for _ in range(n)]
    top = (1 << 256) - 1
    for j, v in enumerate([C.r, C.r + 1, 2 * C.r - 1, min(5 * C.r + 7, top), top, 0, 1, C.r - 1]):
        ss[j] = v
    P = make_points(C, ks)
    S = ints_to_array(ss, 4)
    exp = o.msm_affine(S, P)
    assert (zk.msm_std(C.id, S, P) == exp).all()
    assert C.words_to_affine(exp) == C.mul(sum(s * k for s, k in zip(ss, ks)) % C.r, C.g)


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_all_reference_msm_symbols(zk, C):
    """Every exported MSM entry point (proj/jac x std/mont x affine/proj/jac out)."""
    o = Oracle(C)
    n = 300
    ks, ss = edge_case_inputs(C, n, 77)
    P = make_points(C, ks)
    S = ints_to_array(ss, 4)
    Sm = ints_to_array([C.fr_to_mont(s % C.r) for s in ss], 4)
    exp = C.words_to_affine(o.msm_affine(S, P))
    for coord in ("proj", "jac"):
        for kind, scal in (("std", S), ("mont", Sm)):
            a = zk.msm_ref_symbol(C.id, coord, kind, "affine", scal, P)
            assert C.words_to_affine(a) == exp
            full = zk.msm_ref_symbol(C.id, coord, kind, coord, scal, P)
            dec = C.proj_words_to_affine(full) if coord == "proj" else C.jac_words_to_affine(full)
            assert dec == exp


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_msm_infinity_and_empty(zk, C):
    o = Oracle(C)
    P = make_points(C, [5, 9])
    S = ints_to_array([0, 0], 4)
    res = zk.msm_std(C.id, S, P)
    assert (res == np.uint64(0xFFFFFFFFFFFFFFFF)).all()
    assert (res == o.msm_affine(S, P)).all()
    # P + (-P) = infinity
    P2 = make_points(C, [5, C.r - 5])
    assert (zk.msm_std(C.id, ints_to_array([3, 3], 4), P2) == np.uint64(0xFFFFFFFFFFFFFFFF)).all()
    # projective / Jacobian encodings of infinity
    pj = zk.msm_ref_symbol(C.id, "proj", "std", "proj", S, P)
    assert C.proj_words_to_affine(pj) is None
    jc = zk.msm_ref_symbol(C.id, "jac", "std", "jac", S, P)
    assert C.jac_words_to_affine(jc) is None


@pytest.mark.parametrize("C", CURVES, ids=IDS)
@pytest.mark.parametrize("enl", [1, 2, 3, 5, 8])
def test_msm_scalar_widths(zk, C, enl):
    """expo_nlimbs != 4: scalars are plain 64*enl-bit integers (bn128_G1_proj.c:510-556)."""
    o = Oracle(C)
    n = 200
    ks = random_scalars(5 + enl, n, C.r)
    rng = splitmix64(9 + enl)
    S = np.array([[next(rng) for _ in range(enl)] for _ in range(n)], dtype=np.uint64)
    S[0] = np.uint64(0xFFFFFFFFFFFFFFFF)
    P = make_points(C, ks)
    assert (zk.msm_std(C.id, S, P, expo_nlimbs=enl) == o.msm_affine(S, P, enl=enl)).all()


def device_points(zk, C, ks):
    dK = zk.DeviceArray.from_host(ints_to_array(ks, 4))
    return zk.scalar_mul_gen(C.id, dK, len(ks))


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_scalar_mul_gen_samples(zk, C):
    n = 4096
    ks = random_scalars(31, n, C.r)
    ks[10] = 0
    dP = device_points(zk, C, ks)
    P = dP.to_host((n, 2 * C.lp))
    for i in (0, 1, 10, 777, n - 1):
        exp = C.mul(ks[i], C.g) if ks[i] else None
        got = C.words_to_affine(P[i]) if ks[i] else (None if not P[i].any() else "nonzero")
        assert got == exp, i


@pytest.mark.parametrize("C", CURVES, ids=IDS)
@pytest.mark.parametrize("logn,dist", [(12, "uniform"), (16, "uniform"), (16, "u64"), (16, "sparse")])
def test_msm_closed_form(zk, C, logn, dist):
    n = 1 << logn
    ks = random_scalars(100 + logn, n, C.r)
    ss = random_scalars(200 + logn, n, C.r, dist)
    dP = device_points(zk, C, ks)
    dS = zk.DeviceArray.from_host(ints_to_array(ss, 4))
    res = zk.msm_device(C.id, n, dS, dP)
    assert C.words_to_affine(res) == C.mul(sum(s * k for s, k in zip(ss, ks)) % C.r, C.g)


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_msm_phase_timing_switch(zk, C):
    """Phase events are recorded only while phase timing is on (five device phases and the host
    combine); the switch never changes the result."""
    n = 1 << 14
    ks = random_scalars(300, n, C.r)
    ss = random_scalars(301, n, C.r)
    dP = device_points(zk, C, ks)
    dS = zk.DeviceArray.from_host(ints_to_array(ss, 4))
    off = zk.msm_device(C.id, n, dS, dP)
    assert len(zk.last_timing()[1]) == 1
    zk.set_phase_timing(True)
    try:
        on = zk.msm_device(C.id, n, dS, dP)
        phases = zk.last_timing()[1]
    finally:
        zk.set_phase_timing(False)
    assert (on == off).all()
    assert len(phases) == 6 and min(phases) >= 0
    assert C.words_to_affine(off) == C.mul(sum(s * k for s, k in zip(ss, ks)) % C.r, C.g)


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_msm_schedules_agree(zk, C):
    """Window width and segment length change the schedule, never the result."""
    n = 5000
    ks = random_scalars(55, n, C.r)
    ss = random_scalars(56, n, C.r, "sparse")
    dP = device_points(zk, C, ks)
    dS = zk.DeviceArray.from_host(ints_to_array(ss, 4))
    exp = C.mul(sum(s * k for s, k in zip(ss, ks)) % C.r, C.g)
    for c in (1, 2, 5, 8, 11, 13, 16):
        for seg in (1, 3, 5, 8):
            res = zk.msm_device(C.id, n, dS, dP, window=c, seg_log=seg)
            assert C.words_to_affine(res) == exp, (c, seg)


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_msm_wide_sort_words(zk, C):
    """2^22 + 3 points at a 20-bit window: the point index (23 bits), the sign and the low
    bucket bits no longer fit one 32-bit word, so the sort's first level writes 64-bit words.
    The result must equal the same MSM at c = 16 (32-bit words) bit for bit, and a skewed
    scalar set (every third scalar 1) must too."""
    n = (1 << 22) + 3
    dK = zk.fill_random(C.id, 1, n, 71)
    dP = zk.scalar_mul_gen(C.id, dK, n)
    dS = zk.fill_random(C.id, 1, n, 72)
    narrow = zk.msm_device(C.id, n, dS, dP, window=16)
    wide = zk.msm_device(C.id, n, dS, dP, window=20)
    assert (wide == narrow).all()
    s = dS.to_host((n, 4))
    s[::3] = 0
    s[::3, 0] = 1
    dS2 = zk.DeviceArray.from_host(s)
    assert (zk.msm_device(C.id, n, dS2, dP, window=20) == zk.msm_device(C.id, n, dS2, dP, window=16)).all()


def test_msm_bn254_glv_extremes(zk):
    """BN254 under the plain schedule splits k = k1 + lambda k2 (mod r) with |k1|, |k2| < 2^127
    and sums k1 P + k2 phi(P), phi(x, y) = (beta x, y).  Scalars chosen at the split's extremes
    (largest |k1| and |k2| of a random search, each sign, k = lambda and r - lambda where one
    half is zero, r - 1) must give the oracle's integer MSM.  The decomposition below only picks
    inputs; the check is the oracle and the closed form."""
    C = BN
    lam = 0x30644E72E131A029048B6E193FD84104CC37A73FEC2BC5E9B8CA0B2D36636F23
    assert pow(lam, 3, C.r) == 1 and lam != 1
    a1, b1 = 147946756881789319000765030803803410728, -9931322734385697763
    a2, b2 = 9931322734385697763, 147946756881789319010696353538189108491
    g1 = ((b2 << 256) + C.r // 2) // C.r
    g2 = ((-b1 << 256) + C.r // 2) // C.r

    def split(k):
        c1, c2 = (k * g1) >> 256, (k * g2) >> 256
        return k - c1 * a1 - c2 * a2, -c1 * b1 - c2 * b2

    rng = np.random.default_rng(9)
    pool = [int.from_bytes(rng.bytes(32), "little") % C.r for _ in range(20000)]
    sp = [(k,) + split(k) for k in pool]
    pick = set()
    for idx, sign in ((1, 1), (1, -1), (2, 1), (2, -1)):
        pick.add(max(sp, key=lambda t: sign * t[idx])[0])
    ss = sorted(pick) + [lam, C.r - lam, C.r - 1, 1, 0, (C.r - 1) // 2]
    for k in ss:
        k1, k2 = split(k)
        assert (k1 + lam * k2 - k) % C.r == 0 and abs(k1) < 2**127 and abs(k2) < 2**127
    n = 3000
    ss = (ss * (n // len(ss) + 1))[:n]
    ks = random_scalars(4242, n, C.r)
    P = make_points(C, ks)
    S = ints_to_array(ss, 4)
    exp = Oracle(C).msm_affine(S, P)
    assert (zk.msm_std(C.id, S, P) == exp).all()
    assert C.words_to_affine(exp) == C.mul(sum(s * k for s, k in zip(ss, ks)) % C.r, C.g)


@pytest.mark.parametrize("C", [BN, BLS], ids=["bn254", "bls12-381"])
def test_prepared_prefixes_match_affine(zk, C):
    """A prepared set of N points (BN254: P_i and phi(P_i) side by side, so the GLV split applies)
    used for MSMs over its first n <= N points, with the infinity among them, equals the MSM over
    the same affine words, and the oracle: plain and statistics schedules, Montgomery and
    standard scalars."""
    N = 5000
    ks = random_scalars(5150 + C.id, N, C.r)
    P = make_points(C, ks)
    P[17] = 0  # the affine infinity (0, 0)
    P[N - 1] = 0
    dP = zk.DeviceArray.from_host(P)
    prep = zk.PreparedPoints(C.id, dP, N)
    o = Oracle(C)
    for n in (N, 4096, 1000, 18, 1):
        S = ints_to_array(random_scalars(5160 + n, n, C.r), 4)
        dS = zk.DeviceArray.from_host(S)
        exp = o.msm_affine(S, P[:n])
        for pol in (zk.SCHED_PLAIN, zk.SCHED_STATS):
            a = zk.msm_device(C.id, n, dS, prep, policy=pol)
            b = zk.msm_device(C.id, n, dS, dP, policy=pol)
            assert (a == exp).all() and (b == exp).all(), (n, pol)
        dS.free()
    prep.free()
