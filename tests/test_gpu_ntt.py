"""GPU NTT / iNTT parity against the oracle (bit-exact), the O(N^2) DFT and size-independent
properties at the large sizes.

Reference behaviour: lib/cbits/curves/poly/mont/bn128_poly_mont.c:418-638 (natural order in
and out; the inverse receives the forward generator).  The properties follow
test/src/ZK/Test/Poly/Properties.hs:376-475 (intt . ntt = id, ntt = pointwise evaluation,
shifted identities), re-expressed at the build's sizes.
"""
import numpy as np
import pytest

from cref import Oracle
from zkmath import BN, BLS, ints_to_array, array_to_ints, random_scalars, dft, horner, to_limbs

pytestmark = pytest.mark.gpu
CURVES = [BN, BLS]


def rand_vec(C, m, seed):
    xs = random_scalars(seed, 1 << m, C.r)
    return xs, ints_to_array([C.fr_to_mont(x) for x in xs], 4)


def gen_arr(C, m):
    return np.array(to_limbs(C.fr_to_mont(C.fft_root(m)), 4), np.uint64)


@pytest.mark.parametrize("C", CURVES, ids=["bn254", "bls12-381"])
@pytest.mark.parametrize("m", list(range(0, 13)))
def test_ntt_vs_oracle(zk, C, m):
    o = Oracle(C)
    xs, X = rand_vec(C, m, 100 + m)
    g = gen_arr(C, m)
    y = zk.ntt(C.id, m, g, X)
    assert (y == o.ntt(m, g, X)).all()
    z = zk.intt(C.id, m, g, X)
    assert (z == o.ntt(m, g, X, inverse=True)).all()
    assert (zk.intt(C.id, m, g, y) == X).all()


@pytest.mark.parametrize("C", CURVES, ids=["bn254", "bls12-381"])
def test_ntt_vs_dft(zk, C):
    """Closed form y_k = sum_i x_i g^(ik) (no implementation in between)."""
    for m in (1, 3, 6):
        xs, X = rand_vec(C, m, 7 + m)
        g = C.fft_root(m)
        y = zk.ntt(C.id, m, gen_arr(C, m), X)
        assert [C.fr_from_mont(v) for v in array_to_ints(y)] == dft(xs, g, C.r)


@pytest.mark.parametrize("C", CURVES, ids=["bn254", "bls12-381"])
@pytest.mark.parametrize("m", [2, 5, 9, 12])
def test_shifted_vs_oracle(zk, C, m):
    o = Oracle(C)
    _, X = rand_vec(C, m, 300 + m)
    g = gen_arr(C, m)
    eta = np.array(to_limbs(C.fr_to_mont(random_scalars(99 + m, 1, C.r)[0]), 4), np.uint64)
    assert (zk.shifted_ntt(C.id, m, g, eta, X) == o.ntt_shifted(eta, m, g, X)).all()
    assert (zk.shifted_intt(C.id, m, g, eta, X) == o.ntt_shifted(eta, m, g, X, inverse=True)).all()


@pytest.mark.parametrize("C", CURVES, ids=["bn254", "bls12-381"])
@pytest.mark.parametrize("ms,mt", [(0, 2), (2, 4), (4, 6), (5, 5), (3, 9)])
def test_asymmetric_vs_oracle(zk, C, ms, mt):
    o = Oracle(C)
    _, X = rand_vec(C, ms, 500 + ms)
    gs, gt = gen_arr(C, ms), gen_arr(C, mt)
    assert (zk.asymm_ntt(C.id, ms, gs, X, mt, gt) == o.ntt_asym(ms, mt, gs, gt, X)).all()


@pytest.mark.parametrize("C", CURVES, ids=["bn254", "bls12-381"])
@pytest.mark.parametrize("m", [2, 4, 7])
def test_non_root_generator_matches_reference_network(zk, C, m):
    """A generator whose order is not 2^m: the reference's butterfly network is reproduced."""
    o = Oracle(C)
    _, X = rand_vec(C, m, 600 + m)
    g = np.array(to_limbs(C.fr_to_mont(random_scalars(42 + m, 1, C.r)[0]), 4), np.uint64)
    assert (zk.ntt(C.id, m, g, X) == o.ntt(m, g, X)).all()
    assert (zk.intt(C.id, m, g, X) == o.ntt(m, g, X, inverse=True)).all()


@pytest.mark.parametrize("C,m", [(BN, 16), (BN, 20), (BN, 22), (BLS, 18), (BLS, 21)],
                         ids=["bn-16", "bn-20", "bn-22", "bls-18", "bls-21"])
def test_large_roundtrip_and_evaluation(zk, C, m):
    """At full size: intt(ntt(x)) == x bit-exactly, and sampled outputs equal f(g^k) by Horner."""
    N = 1 << m
    rng = np.random.default_rng(m)
    X = rng.integers(0, 2**63, size=(N, 4), dtype=np.uint64)
    X[:, 3] &= np.uint64((1 << 59) - 1)  # < 2^251 < r: canonical elements
    g = gen_arr(C, m)
    dX = zk.DeviceArray.from_host(X)
    dY = zk.DeviceArray(N * 4)
    dZ = zk.DeviceArray(N * 4)
    zk.ntt_device(C.id, m, g, dX, dY)
    zk.ntt_device(C.id, m, g, dY, dZ, inverse=True)
    assert (dZ.to_host((N, 4)) == X).all()
    if m > 20:
        return  # Horner over 2^21+ coefficients in Python is too slow; round trip covers it
    Y = dY.to_host((N, 4))
    # Montgomery form is linear, so y_k(mont) = sum_i x_i(mont) g^(ik) mod r
    coeffs = array_to_ints(X)
    w = C.fft_root(m)
    for k in (1, N // 2 + 3):
        assert array_to_ints(Y[k:k + 1])[0] == horner(coeffs, pow(w, k, C.r), C.r), k


@pytest.mark.parametrize("C", CURVES, ids=["bn254", "bls12-381"])
@pytest.mark.parametrize("m", [1, 3, 8, 11, 12, 17, 22])
def test_extreme_inputs_closed_form(zk, C, m):
    """Largest canonical inputs drive the butterflies' intermediate bounds hardest.  Closed forms:
    a constant vector c maps to (N c, 0, ..., 0); c on even indices only maps to N/2 c at k = 0
    and k = N/2; the inverse of (N c, 0, ..., 0) is the constant c."""
    N = 1 << m
    c = C.r - 1
    g = gen_arr(C, m)
    mont = lambda v: C.fr_to_mont(v % C.r)
    # constant
    X = np.tile(np.array(to_limbs(mont(c), 4), np.uint64), (N, 1))
    dX, dY, dZ = zk.DeviceArray.from_host(X), zk.DeviceArray(X.size), zk.DeviceArray(X.size)
    zk.ntt_device(C.id, m, g, dX, dY)
    Y = dY.to_host(X.shape)
    assert array_to_ints(Y[:1]) == [mont(N * c)] and not Y[1:].any()
    zk.ntt_device(C.id, m, g, dY, dZ, inverse=True)
    assert (dZ.to_host(X.shape) == X).all()
    # even indices only
    if m >= 1:
        X2 = X.copy()
        X2[1::2] = 0
        dX2 = zk.DeviceArray.from_host(X2)
        zk.ntt_device(C.id, m, g, dX2, dY)
        Y = dY.to_host(X.shape)
        half = mont((N // 2) * c)
        exp = np.zeros_like(X)
        exp[0] = to_limbs(half, 4)
        exp[N // 2] = to_limbs(half, 4)
        assert (Y == exp).all()
        zk.ntt_device(C.id, m, g, dY, dZ, inverse=True)
        assert (dZ.to_host(X.shape) == X2).all()
    # coset forward of the constant: sum_i c eta^i g^(ik) = c (eta^N - 1) / (eta g^k - 1)
    if m <= 12:
        eta = 5
        etam = np.array(to_limbs(mont(eta), 4), np.uint64)
        zk.ntt_device(C.id, m, g, dX, dY, eta=etam)
        got = array_to_ints(dY.to_host(X.shape))
        gr = C.fft_root(m)
        for k in (0, 1, N - 1):
            num = c * (pow(eta, N, C.r) - 1) % C.r
            den = (eta * pow(gr, k, C.r) - 1) % C.r
            assert got[k] == mont(num * pow(den, -1, C.r))


@pytest.mark.parametrize("C", CURVES, ids=["bn254", "bls12-381"])
@pytest.mark.parametrize("m", [1, 5, 9, 12, 16, 20])
def test_in_place_equals_out_of_place(zk, C, m):
    """zkg_ntt with the source buffer as target (the first pass then gathers into a scratch
    vector instead of the target) gives the out-of-place words: forward, inverse and both coset
    forms, and the no-allocation entry point likewise where it applies."""
    N = 1 << m
    _, X = rand_vec(C, m, 300 + m)
    g = gen_arr(C, m)
    eta = np.array(to_limbs(C.fr_to_mont(5), 4), np.uint64)
    for inverse in (False, True):
        for e in (None, eta):
            dX = zk.DeviceArray.from_host(X)
            dY = zk.DeviceArray(N * 4)
            zk.ntt_device(C.id, m, g, dX, dY, inverse=inverse, eta=e)
            zk.ntt_device(C.id, m, g, dX, dX, inverse=inverse, eta=e)
            assert (dX.to_host((N, 4)) == dY.to_host((N, 4))).all(), (inverse, e is not None)
            if m <= 18:
                dZ = zk.DeviceArray.from_host(X)
                zk.ntt_device_noalloc(C.id, m, g, dZ, dZ, inverse=inverse, eta=e)
                assert (dZ.to_host((N, 4)) == dY.to_host((N, 4))).all(), ("noalloc", inverse, e is not None)
