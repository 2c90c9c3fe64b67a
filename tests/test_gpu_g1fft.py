"""GPU FFT of G1 points (the reference's <c>_G1_proj_fft_{forward,inverse},
lib/cbits/curves/g1/proj/bn128_G1_proj.c:791-902) against the oracle's restatement of the
recursive networks, for the FFT-subgroup generator and for an arbitrary one, and against the
closed form FFT(k_i G)_j = (sum_i k_i w^(ij)) G at larger sizes, with the scalar NTT and the
fixed-base multiplication on the device as the checker of the latter."""
import numpy as np
import pytest

from cref import Oracle
from zkmath import BN, BLS, array_to_ints, ints_to_array, random_scalars, to_limbs

pytestmark = pytest.mark.gpu
CURVES = [BN, BLS]
IDS = ["bn254", "bls12-381"]


def proj_inputs(C, m, seed):
    ks = random_scalars(seed, 1 << m, C.r)
    if m >= 2:
        ks[1] = 0  # the infinity (0:1:0)
    return ks, Oracle(C).proj_points(ks)


@pytest.mark.parametrize("C", CURVES, ids=IDS)
@pytest.mark.parametrize("m", [0, 1, 2, 3, 4, 6])
def test_g1_fft_vs_oracle(zk, C, m):
    o = Oracle(C)
    _, P = proj_inputs(C, m, 500 + m)
    gens = [o.fft_generator(m), np.array(to_limbs(C.fr_to_mont(31337 + m), 4), np.uint64)]
    for g in gens:
        for inv in (False, True):
            exp = o.g1_fft(m, g, P, inverse=inv)
            assert (zk.g1_fft(C.id, m, g, P, inverse=inv) == exp).all(), (m, inv)


@pytest.mark.parametrize("C,m", [(BN, 10), (BLS, 9)], ids=["bn254-2^10", "bls12-381-2^9"])
def test_g1_fft_closed_form_and_roundtrip(zk, C, m):
    N = 1 << m
    ks = random_scalars(600 + m, N, C.r)
    P = Oracle(C).proj_points(ks, zs=[3 + 5 * i for i in range(N)]) if m <= 10 else None
    g = zk.fft_generator(C.id, m)
    d_src = zk.DeviceArray.from_host(P)
    d_y, d_back = zk.DeviceArray(P.size), zk.DeviceArray(P.size)
    zk.g1_fft_device(C.id, m, g, d_src, d_y)
    Y = d_y.to_host(P.shape)
    # closed form: NTT of the scalars (device NTT, Montgomery words) times G
    K = ints_to_array([C.fr_to_mont(k) for k in ks], 4)
    E = zk.ntt(C.id, m, g, K)
    e_std = ints_to_array([C.fr_from_mont(x) for x in array_to_ints(E)], 4)
    d_e = zk.DeviceArray.from_host(e_std)
    G = zk.scalar_mul_gen(C.id, d_e, N).to_host((N, 2 * C.lp))
    L = C.lp
    for j in range(N):
        if not G[j].any():
            assert not Y[j, :L].any() and not Y[j, 2 * L:].any()
        else:
            assert (Y[j, :2 * L] == G[j]).all(), j
            assert C.fp_from_mont(int(sum(int(w) << (64 * t) for t, w in enumerate(Y[j, 2 * L:])))) == 1
    zk.g1_fft_device(C.id, m, g, d_y, d_back, inverse=True)
    back = d_back.to_host(P.shape)
    ref = Oracle(C).points_from_scalars(ks)
    assert (back[:, :2 * L] == ref).all()


@pytest.mark.parametrize("C,m", [(BN, 14), (BN, 16), (BN, 18), (BLS, 12), (BLS, 15)],
                         ids=["bn254-2^14", "bn254-2^16", "bn254-2^18", "bls12-381-2^12", "bls12-381-2^15"])
def test_g1_fft_closed_form_large(zk, C, m):
    """The stage plans of larger transforms (2^14: three groups of four stages and one of two;
    2^16: five groups of three stages and one radix-2 stage, the table stride set by the
    largest group launch; 2^18: radix-2 stages only, whose threads stride over the butterflies,
    and the inverse's separate N^-1 pass; BLS12-381: the 2-bit windows on k) against the same
    closed form, with the points k_i G built on the device as (x, y, 1); the inverse must
    return them."""
    N = 1 << m
    L = C.lp
    ks = random_scalars(700 + m, N, C.r)
    d_k = zk.DeviceArray.from_host(ints_to_array(ks, 4))
    aff = zk.scalar_mul_gen(C.id, d_k, N).to_host((N, 2 * L))
    one = np.array(to_limbs(C.fp_to_mont(1), L), np.uint64)
    P = np.ascontiguousarray(np.concatenate([aff, np.tile(one, (N, 1))], axis=1))
    g = zk.fft_generator(C.id, m)
    d_src = zk.DeviceArray.from_host(P)
    d_y, d_back = zk.DeviceArray(P.size), zk.DeviceArray(P.size)
    zk.g1_fft_device(C.id, m, g, d_src, d_y)
    Y = d_y.to_host(P.shape)
    K = ints_to_array([C.fr_to_mont(k) for k in ks], 4)
    E = zk.ntt(C.id, m, g, K)
    e_std = ints_to_array([C.fr_from_mont(x) for x in array_to_ints(E)], 4)
    G = zk.scalar_mul_gen(C.id, zk.DeviceArray.from_host(e_std), N).to_host((N, 2 * L))
    inf = ~G.any(axis=1)
    assert (Y[~inf, :2 * L] == G[~inf]).all()
    assert (Y[~inf, 2 * L:] == one).all()
    assert not Y[inf, :L].any() and not Y[inf, 2 * L:].any()
    zk.g1_fft_device(C.id, m, g, d_y, d_back, inverse=True)
    assert (d_back.to_host(P.shape) == P).all()
