"""Pins the oracle's binary MSMs, batch conversions and Jacobian G1 FFT (oracle/zko.c) before
they are trusted as the GPU checker:

* golden fixtures written by tests/golden/gen_golden.py from the reference's own C;
* closed forms in Python integers: a binary MSM over k_i G is (sum of the selected k_i) G, a
  converted point is the same affine point, the FFT of k_i G is NTT(k)_j G;
* the live reference build on fresh inputs where it is present.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from cref import Oracle, Ref, ref
from zkmath import BN, BLS, random_scalars, to_limbs

GOLDEN = Path(__file__).resolve().parent / "golden"
CURVES = [BN, BLS]
IDS = ["bn254", "bls12-381"]


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN / "golden_small.npz", allow_pickle=False)


@pytest.fixture(scope="module")
def manifest():
    return json.loads((GOLDEN / "manifest.json").read_text())


def affine_int(C, words):
    """Affine words -> Python point, all-0xFF (the reference's affine infinity) -> None."""
    if (np.asarray(words) == np.uint64(0xFFFFFFFFFFFFFFFF)).all():
        return None
    return C.words_to_affine(words)


# ---------------------------------------------------------------- golden fixtures

@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_oracle_binary_msm_vs_golden(golden, manifest, C):
    o = Oracle(C)
    keys = [k for k in manifest["binary"] if k.startswith(C.name + "_")]
    assert len(keys) == 3
    for key in keys:
        P, bits, rows = golden[key + "_points"], golden[key + "_bits"], golden[key + "_rows"]
        assert (o.proj_to_affine(o.msm_binary(bits, P)) == golden[key + "_affine"]).all(), key
        assert (golden[key + "_jac_affine"] == golden[key + "_affine"]).all(), key
        multi = o.multi_msm_binary(rows, P)
        for r in range(rows.shape[0]):
            assert (o.proj_to_affine(multi[r]) == golden[key + "_multi_affine"][r]).all(), (key, r)
        assert (golden[key + "_multi_jac_affine"] == golden[key + "_multi_affine"]).all(), key


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_oracle_batch_conversions_vs_golden(golden, C):
    o = Oracle(C)
    key = f"{C.name}_conv"
    assert (o.batch_to_affine(golden[key + "_proj"]) == golden[key + "_proj_to_affine"]).all()
    assert (o.batch_to_affine(golden[key + "_jac"], jac=True) == golden[key + "_jac_to_affine"]).all()
    assert (o.batch_from_affine(golden[key + "_affine"]) == golden[key + "_proj_from_affine"]).all()
    assert (o.batch_from_affine(golden[key + "_affine"], jac=True) == golden[key + "_jac_from_affine"]).all()


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_oracle_jac_g1_fft_vs_golden(golden, manifest, C):
    o = Oracle(C)
    keys = [k for k in manifest["g1fft_jac"] if k.startswith(C.name + "_")]
    assert keys
    for key in keys:
        m = int(key.rsplit("_", 1)[1])
        P, g = golden[key + "_src"], golden[key + "_gen"]
        assert (o.g1_fft(m, g, P, jac=True) == golden[key + "_fwd"]).all(), key
        assert (o.g1_fft(m, g, P, inverse=True, jac=True) == golden[key + "_inv"]).all(), key


# ---------------------------------------------------------------- closed forms

@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_oracle_binary_msm_closed_form(C):
    """sum over the selected k_i G equals (sum of the selected k_i) G, whatever non-zero byte
    selects a point; an empty selection is the infinity (0 : 1 : 0)."""
    o = Oracle(C)
    ks = random_scalars(31 + C.id, 50, C.r)
    ks[4] = 0
    ks[9] = (C.r - ks[10]) % C.r
    P = o.points_from_scalars(ks)
    rng = np.random.default_rng(5)
    for t in range(4):
        bits = (rng.random(50) < 0.5).astype(np.uint8) * rng.integers(1, 256, 50).astype(np.uint8)
        if t == 1:
            bits[:] = 0
            bits[9] = bits[10] = 200  # the cancelling pair alone
        res = o.msm_binary(bits, P)
        e = sum(k for k, b in zip(ks, bits) if b) % C.r
        if e == 0:
            L = C.lp
            assert (res[:L] == 0).all() and (res[2 * L:] == 0).all()
        else:
            assert C.words_to_affine(res[:2 * C.lp]) == C.mul(e, C.g)
    empty = o.msm_binary(np.zeros(50, np.uint8), P)
    assert (empty[2 * C.lp:] == 0).all()


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_oracle_conversions_closed_form(C):
    """to_affine of projective or Jacobian forms of k_i G is k_i G; from_affine then to_affine
    is the identity; the infinity becomes all-0xFF words."""
    o = Oracle(C)
    ks = random_scalars(41 + C.id, 20, C.r)
    ks[3] = 0
    for jac in (False, True):
        P = o.jac_points(ks) if jac else o.proj_points(ks)
        A = o.batch_to_affine(P, jac=jac)
        for k, a in zip(ks, A):
            assert affine_int(C, a) == (None if k == 0 else C.mul(k, C.g))
        aff = o.points_from_scalars(ks)
        back = o.batch_to_affine(o.batch_from_affine(aff, jac=jac), jac=jac)
        for i, k in enumerate(ks):
            if k:
                assert (back[i] == aff[i]).all()
            else:
                assert (back[i] == np.uint64(0xFFFFFFFFFFFFFFFF)).all()


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_oracle_jac_g1_fft_closed_form(C):
    o = Oracle(C)
    m = 3
    ks = random_scalars(91 + C.id, 1 << m, C.r)
    P = o.jac_points(ks)
    g = o.fft_generator(m)
    Y = o.g1_fft(m, g, P, jac=True)
    w = C.fft_root(m)
    for j in range(1 << m):
        e = sum(k * pow(w, i * j, C.r) for i, k in enumerate(ks)) % C.r
        assert C.words_to_affine(Y[j][:2 * C.lp]) == C.mul(e, C.g)
    assert (Y == o.g1_fft(m, g, o.proj_points(ks))).all()


# ---------------------------------------------------------------- live reference build

@pytest.mark.skipif(ref() is None, reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_oracle_g1extra_vs_reference_build(C):
    o, r = Oracle(C), Ref(C)
    ks = random_scalars(61 + C.id, 64, C.r)
    ks[2] = 0
    ks[5] = ks[6]
    P = o.points_from_scalars(ks)
    rng = np.random.default_rng(17)
    bits = (rng.random(64) < 0.3).astype(np.uint8)
    for coord in ("proj", "jac"):
        assert (r.msm_binary(bits, P, coord=coord) == o.proj_to_affine(o.msm_binary(bits, P))).all()
    rows = (rng.random((40, 64)) < 0.5).astype(np.uint8)
    mo = o.multi_msm_binary(rows, P)
    mr = r.multi_msm_binary(rows, P)
    for i in range(40):
        assert (mr[i] == o.proj_to_affine(mo[i])).all()
    assert (r.batch_to_affine(o.jac_points(ks), coord="jac") == o.batch_to_affine(o.jac_points(ks), jac=True)).all()
    m = 4
    kk = random_scalars(71 + C.id, 1 << m, C.r)
    g = np.array(to_limbs(C.fr_to_mont(123456789), 4), np.uint64)
    Pj = o.jac_points(kk)
    for inv in (False, True):
        assert (r.g1_fft(m, g, Pj, inverse=inv, coord="jac") == o.g1_fft(m, g, Pj, inverse=inv, jac=True)).all()
