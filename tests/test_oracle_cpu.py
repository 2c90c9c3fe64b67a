"""Pins the CPU oracle (oracle/zko.c) before it is trusted as the GPU checker.

* golden fixtures written by tests/golden/gen_golden.py from the reference's own C
  (lib/cbits compiled from its sources; there are no fixtures in the reference itself);
* closed forms that do not depend on any implementation: Montgomery products as Python
  integers, sum_i s_i (k_i G) = (sum_i s_i k_i) G, the O(N^2) DFT, g^(2^m) = 1;
* the live reference build (oracle/_ref/libzkref.so) on fresh random inputs when present.
"""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from cref import Oracle, Ref, ref
from zkmath import (BN, BLS, array_to_ints, dft, from_limbs, ints_to_array, random_scalars,
                    seeded_fr_vector, splitmix64, random_below, to_limbs)

GOLDEN = Path(__file__).resolve().parent / "golden"
CURVES = [BN, BLS]
IDS = ["bn254", "bls12-381"]


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN / "golden_small.npz", allow_pickle=False)


@pytest.fixture(scope="module")
def manifest():
    return json.loads((GOLDEN / "manifest.json").read_text())


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.uint64).tobytes()).hexdigest()


# ---------------------------------------------------------------- golden fixtures

@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_oracle_msm_vs_golden(golden, manifest, C):
    o = Oracle(C)
    keys = [k for k in manifest["msm"] if k.startswith(C.name + "_")]
    assert keys
    for key in keys:
        P, S, Sm = golden[key + "_points"], golden[key + "_scalars"], golden[key + "_scalars_mont"]
        assert (o.msm_affine(S, P) == golden[key + "_proj_std_affine"]).all(), key
        assert (o.msm_affine(Sm, P, mont=True) == golden[key + "_proj_mont_affine"]).all(), key
        # Jacobian entry points return the same affine point
        assert (golden[key + "_jac_std_affine"] == golden[key + "_proj_std_affine"]).all(), key


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_oracle_ntt_vs_golden(golden, manifest, C):
    o = Oracle(C)
    for key in [k for k in manifest["ntt"] if k.startswith(C.name + "_")]:
        m = int(key.rsplit("_", 1)[1])
        X, g = golden[key + "_src"], golden[key + "_gen"]
        assert (o.ntt(m, g, X) == golden[key + "_fwd"]).all(), key
        assert (o.ntt(m, g, X, inverse=True) == golden[key + "_inv"]).all(), key
        if key + "_eta" in golden:
            eta = golden[key + "_eta"]
            assert (o.ntt_shifted(eta, m, g, X) == golden[key + "_fwd_shifted"]).all(), key
            assert (o.ntt_shifted(eta, m, g, X, inverse=True) == golden[key + "_inv_shifted"]).all(), key
    for ms, mt in ((2, 4), (3, 7), (5, 6)):
        X, gs = golden[f"{C.name}_ntt_{ms}_src"], golden[f"{C.name}_ntt_{ms}_gen"]
        gt = np.array(to_limbs(C.fr_to_mont(C.fft_root(mt)), 4), np.uint64)
        assert (o.ntt_asym(ms, mt, gs, gt, X) == golden[f"{C.name}_asym_{ms}_{mt}"]).all()


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_oracle_g1_fft_vs_golden(golden, manifest, C):
    o = Oracle(C)
    keys = [k for k in manifest["g1fft"] if k.startswith(C.name + "_")]
    assert keys
    for key in keys:
        base = key.rsplit("_", 1)[0]
        m = int(base.rsplit("_", 1)[1])
        P, g = golden[base + "_src"], golden[key + "_gen"]
        assert (o.g1_fft(m, g, P) == golden[key + "_fwd"]).all(), key
        assert (o.g1_fft(m, g, P, inverse=True) == golden[key + "_inv"]).all(), key


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_oracle_g1_fft_closed_form_and_roundtrip(C):
    """FFT of k_i G is (NTT(k))_j G (linearity in the scalar), and the inverse undoes it."""
    o = Oracle(C)
    m = 3
    ks = random_scalars(77 + C.id, 1 << m, C.r)
    P = o.proj_points(ks)
    g = o.fft_generator(m)
    Y = o.g1_fft(m, g, P)
    w = C.fft_root(m)
    for j in range(1 << m):
        e = sum(k * pow(w, i * j, C.r) for i, k in enumerate(ks)) % C.r
        assert C.words_to_affine(Y[j][:2 * C.lp]) == C.mul(e, C.g)
    back = o.g1_fft(m, g, Y, inverse=True)
    for i, k in enumerate(ks):
        assert C.words_to_affine(back[i][:2 * C.lp]) == C.mul(k, C.g)


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_oracle_large_ntt_vs_golden_digest(manifest, C):
    o = Oracle(C)
    for e in manifest["large"]:
        if e["curve"] != C.name or e["kind"] != "ntt_fwd" or e["m"] > 16:
            continue
        m = e["m"]
        X = seeded_fr_vector(e["seed"], 1 << m, C.r)
        g = np.array(to_limbs(C.fr_to_mont(C.fft_root(m)), 4), np.uint64)
        assert sha(o.ntt(m, g, X)) == e["sha256"]
        assert sha(o.ntt(m, g, X, inverse=True)) == e["inv_sha256"]


def test_oracle_large_msm_vs_golden(manifest):
    e = next(x for x in manifest["large"] if x["kind"] == "msm_std_affine" and x["curve"] == "bn128"
             and x["logn"] == 14)
    n = 1 << e["logn"]
    K = seeded_fr_vector(e["kseed"], n, BN.r)
    S = seeded_fr_vector(e["sseed"], n, BN.r)
    o = Oracle(BN)
    P = o.points_from_scalars(array_to_ints(K))
    assert sha(P) == e["points_sha256"]
    assert [int(x) for x in o.msm_affine(S, P)] == e["result"]


# ---------------------------------------------------------------- closed forms

@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_oracle_field_mul_closed_form(C):
    o = Oracle(C)
    rng = splitmix64(3 + C.id)
    for field, p, L in ((0, C.p, C.lp), (1, C.r, 4)):
        Rinv = pow(1 << (64 * L), -1, p)
        vals = [0, 1, p - 1] + [random_below(rng, p, L) for _ in range(300)]
        for a, b in zip(vals, vals[::-1]):
            A, B = np.array(to_limbs(a, L), np.uint64), np.array(to_limbs(b, L), np.uint64)
            assert from_limbs(o.f_mul(field, A, B)) == a * b * Rinv % p


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_oracle_msm_closed_form_and_windows(C):
    """(sum s_i k_i) G for every window width: the bucket method's answer cannot depend on c."""
    o = Oracle(C)
    n = 150
    ks = random_scalars(21, n, C.r)
    ss = random_scalars(22, n, C.r)
    ss[0], ss[1] = 0, C.r - 1
    P = o.points_from_scalars(ks)
    S = ints_to_array(ss, 4)
    exp = C.mul(sum(a * b for a, b in zip(ss, ks)) % C.r, C.g)
    assert C.words_to_affine(o.msm_affine(S, P)) == exp
    for c in (1, 2, 3, 7, 11):
        assert C.proj_words_to_affine(o.msm_proj_variable(S, P, c)) == exp
    # a transposed operand (scalars reversed) must change the answer
    assert C.words_to_affine(o.msm_affine(np.ascontiguousarray(S[::-1]), P)) != exp


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_oracle_ntt_closed_form(C):
    o = Oracle(C)
    for m in (0, 1, 2, 4, 6):
        xs = random_scalars(30 + m, 1 << m, C.r)
        X = ints_to_array([C.fr_to_mont(x) for x in xs], 4)
        g = C.fft_root(m)
        G = np.array(to_limbs(C.fr_to_mont(g), 4), np.uint64)
        assert (o.fft_generator(m) == G).all()
        y = o.ntt(m, G, X)
        assert [C.fr_from_mont(v) for v in array_to_ints(y)] == dft(xs, g, C.r)
        assert (o.ntt(m, G, y, inverse=True) == X).all()


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_fft_domain_orders(C):
    """g_m has order exactly 2^m (Class/FFT.hs:60-68 with the codegen's 2-adic generator)."""
    for m in range(0, C.fft_log + 1):
        g = C.fft_root(m)
        assert pow(g, 1 << m, C.r) == 1
        if m:
            assert pow(g, 1 << (m - 1), C.r) == C.r - 1


# ---------------------------------------------------------------- live reference build

needs_ref = pytest.mark.skipif(ref() is None, reason="oracle/_ref not built (make -C oracle ref)")


@needs_ref
@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_oracle_vs_reference_random(C):
    o, r = Oracle(C), Ref(C)
    for n, seed in ((2, 1), (9, 2), (100, 3), (700, 4)):
        ks = random_scalars(seed, n, C.r)
        P = o.points_from_scalars(ks)
        rng = splitmix64(seed + 50)
        S = np.array([[next(rng) for _ in range(4)] for _ in range(n)], np.uint64)  # full 256-bit
        assert (o.msm_affine(S, P) == r.msm(S, P)).all()
        for c in (2, 5, 9):
            a = o.msm_proj_variable(S, P, c)
            b = r.msm_variable(S, P, c)
            assert C.proj_words_to_affine(a) == C.proj_words_to_affine(b)
    for m in (3, 8, 11):
        X = ints_to_array([C.fr_to_mont(x) for x in random_scalars(60 + m, 1 << m, C.r)], 4)
        G = o.fft_generator(m)
        assert (o.ntt(m, G, X) == r.ntt(m, G, X)).all()
        assert (o.ntt(m, G, X, inverse=True) == r.ntt(m, G, X, inverse=True)).all()
