"""GPU polynomial evaluation, division by x^n - eta and KZG openings, through the reference's
entry points (lib/cbits/curves/poly/mont/bn128_poly_mont.c:225-243, :317-413) and the device
API, against the golden fixtures of the reference's C, the oracle on the same inputs (up to
2^22 coefficients, where the scan runs through four levels of chunk carries), and the closed
form of a KZG proof ((p(tau) - p(z)) / (tau - z)) G."""
import json
from pathlib import Path

import numpy as np
import pytest

from cref import Oracle
from zkmath import BN, BLS, array_to_ints, ints_to_array, random_scalars, to_limbs

pytestmark = pytest.mark.gpu
CURVES = [BN, BLS]
IDS = ["bn254", "bls12-381"]
GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN / "golden_small.npz", allow_pickle=False)


@pytest.fixture(scope="module")
def manifest():
    return json.loads((GOLDEN / "manifest.json").read_text())


def fe(C, x):
    return np.array(to_limbs(C.fr_to_mont(x % C.r), 4), np.uint64)


def coeffs(zk, C, n, seed):
    """n canonical Fr elements (taken as Montgomery words) generated on the device."""
    if n == 0:
        return np.zeros((0, 4), np.uint64)
    return zk.fill_random(C.id, 1, n, seed).to_host((n, 4))


# ---------------------------------------------------------------- golden fixtures

@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_poly_vs_golden(zk, golden, manifest, C):
    one = fe(C, 1)
    for key in [k for k in manifest["poly"] if k.startswith(C.name + "_")]:
        a, x = golden[key + "_coeffs"], golden[key + "_x"]
        n1 = a.shape[0]
        assert (zk.poly_eval(C.id, a, x) == golden[key + "_eval"]).all(), key
        for n in (1, 3, 16, 64):
            for ename, eta in (("one", one), ("x", x)):
                q, r = zk.div_by_vanishing(C.id, a, n, eta, n1 + 2, n + 1)
                assert (q == golden[f"{key}_div_{n}_{ename}_quot"]).all(), (key, n, ename)
                assert (r == golden[f"{key}_div_{n}_{ename}_rem"]).all(), (key, n, ename)


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_kzg_open_vs_golden(zk, golden, C):
    key = f"{C.name}_kzg_16"
    a, z, srs = golden[key + "_coeffs"], golden[key + "_z"], golden[key + "_srs"]
    d_a, d_srs = zk.DeviceArray.from_host(a), zk.DeviceArray.from_host(srs)
    v, proof = zk.kzg_open(C.id, 16, d_a, z, d_srs)
    assert (v == golden[key + "_value"]).all()
    assert (proof == golden[key + "_proof"]).all()
    prep = zk.PreparedPoints(C.id, d_srs, 16)
    v2, proof2 = zk.kzg_open(C.id, 16, d_a, z, prep)
    assert (v2 == v).all() and (proof2 == proof).all()
    prep.free()


# ---------------------------------------------------------------- against the oracle

@pytest.mark.parametrize("C", CURVES, ids=IDS)
@pytest.mark.parametrize("n1", [0, 1, 2, 4095, 4096, 4097, 70001])
def test_poly_eval_vs_oracle(zk, C, n1):
    o = Oracle(C)
    a = coeffs(zk, C, n1, 100 + n1)
    for x in (fe(C, 0), fe(C, 1), fe(C, 987654321 + n1), fe(C, -1)):
        assert (zk.poly_eval(C.id, a, x) == o.poly_eval(a, x)).all(), n1


@pytest.mark.parametrize("C", CURVES, ids=IDS)
@pytest.mark.parametrize("n1,n", [(1, 1), (5, 1), (17, 1), (300, 1), (300, 2), (300, 16), (300, 17), (300, 299),
                                  (300, 300), (300, 301), (4097, 3), (70001, 1), (70001, 7), (70001, 4096),
                                  (70001, 40000)])
def test_div_by_vanishing_vs_oracle(zk, C, n1, n):
    o = Oracle(C)
    a = coeffs(zk, C, n1, 200 + n1 + n)
    if n1 > 4:
        a[-2:] = 0  # the degree is below the length
    for eta in (fe(C, 1), fe(C, 0), fe(C, 55555 + n)):
        nq, nr = n1 + 3, n + 2
        st, q, r = o.div_by_vanishing(a, n, eta, nq, nr)
        assert st == 0
        gq, gr = zk.div_by_vanishing(C.id, a, n, eta, nq, nr)
        assert (gq == q).all() and (gr == r).all(), (n1, n)


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_quot_by_vanishing_flag(zk, C):
    o = Oracle(C)
    s = coeffs(zk, C, 1000, 7)
    for n, e in ((1, 12345), (5, 1), (64, 999)):
        eta = fe(C, e)
        # p = (x^n - eta) s, built with the oracle's field products
        p = np.zeros((1000 + n, 4), np.uint64)
        p[n:] = s
        for i in range(1000):
            p[i] = o.f_sub(1, p[i], o.f_mul(1, s[i].copy(), eta))
        ok, q = zk.quot_by_vanishing(C.id, p, n, eta, 1000)
        assert ok and (q == s).all()
        p[0] = o.f_sub(1, p[0], fe(C, 1))
        ok, q = zk.quot_by_vanishing(C.id, p, n, eta, 1000)
        assert not ok and (q == s).all()


@pytest.mark.parametrize("C,logn", [(BN, 22), (BLS, 20)], ids=["bn254-2^22", "bls12-381-2^20"])
def test_div_by_vanishing_device_large(zk, C, logn):
    """The opening quotient (n = 1) and two coset divisors at full size, on device buffers."""
    o = Oracle(C)
    n1 = (1 << logn) + 5
    d_a = zk.fill_random(C.id, 1, n1, 31 + logn)
    a = d_a.to_host((n1, 4))
    assert zk.poly_degree_device(n1, d_a) == n1 - 1
    for n, eta in ((1, fe(C, 424242)), (3, fe(C, 1)), (1 << (logn - 2), fe(C, 77))):
        nq = n1 - n
        d_q, d_r = zk.DeviceArray(nq * 4), zk.DeviceArray(n * 4)
        zk.div_by_vanishing_device(C.id, n1, d_a, n, eta, nq, d_q, n, d_r)
        st, q, r = o.div_by_vanishing(a, n, eta, nq, n)
        assert st == 0
        assert (d_q.to_host((nq, 4)) == q).all(), n
        assert (d_r.to_host((n, 4)) == r).all(), n
        v = zk.poly_eval_device(C.id, n1, d_a, eta)
        assert (v == o.poly_eval(a, eta)).all()


@pytest.mark.parametrize("C,logn", [(BN, 16), (BLS, 14)], ids=["bn254-2^16", "bls12-381-2^14"])
def test_kzg_open_closed_form(zk, C, logn):
    n = 1 << logn
    tau = 0x1234567890ABCDEF1234567 + C.id
    d_srs = zk.srs_generate(C.id, n, tau)
    d_a = zk.fill_random(C.id, 1, n, 5 + logn)
    a = d_a.to_host((n, 4))
    p = [C.fr_from_mont(x) for x in array_to_ints(a)]
    zi = 0xABCDEF + logn
    value, proof = zk.kzg_open(C.id, n, d_a, fe(C, zi), d_srs)
    pz = 0
    for c in reversed(p):
        pz = (pz * zi + c) % C.r
    ptau = 0
    for c in reversed(p):
        ptau = (ptau * tau + c) % C.r
    assert C.fr_from_mont(array_to_ints(value.reshape(1, 4))[0]) == pz
    e = (ptau - pz) * pow((tau - zi) % C.r, C.r - 2, C.r) % C.r
    assert (proof == Oracle(C).points_from_scalars([e])[0]).all()


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_poly_degree_device(zk, C):
    """Degree on the device: top coefficient non-zero (the quick path), trailing zeros (the
    reduction over all coefficients), and the zero polynomial (-1)."""
    n = 300001
    a = coeffs(zk, C, n, 61)
    assert zk.poly_degree_device(n, zk.DeviceArray.from_host(a)) == n - 1
    a[-5:] = 0
    assert zk.poly_degree_device(n, zk.DeviceArray.from_host(a)) == n - 6
    a[:] = 0
    a[12345] = 1
    assert zk.poly_degree_device(n, zk.DeviceArray.from_host(a)) == 12345
    a[:] = 0
    assert zk.poly_degree_device(n, zk.DeviceArray.from_host(a)) == -1
    assert zk.poly_degree_device(0, zk.DeviceArray.from_host(a)) == -1
    # the device division follows the degree the reduction finds
    b = coeffs(zk, C, 1000, 62)
    b[-3:] = 0
    eta = fe(C, 5)
    st, q, r = Oracle(C).div_by_vanishing(b, 7, eta, 1000, 9)
    d_q, d_r = zk.DeviceArray(1000 * 4), zk.DeviceArray(9 * 4)
    zk.div_by_vanishing_device(C.id, 1000, zk.DeviceArray.from_host(b), 7, eta, 1000, d_q, 9, d_r)
    assert st == 0 and (d_q.to_host((1000, 4)) == q).all() and (d_r.to_host((9, 4)) == r).all()
