"""Device field layer (32-bit-limb CIOS Montgomery) against big-integer arithmetic.

Each op is checked on random elements and on the edge values 0, 1, p-1, R mod p and
2^k - 1 patterns below p, for all four fields (BN254 Fp/Fr, BLS12-381 Fp/Fr).  The expected
values come from Python integers (a*b*R^-1 mod p etc.), which pins the Montgomery convention
of lib/cbits/curves/fields/mont/bn128_Fr_mont.c:140-199 independently of any implementation.
"""
import numpy as np
import pytest

from zkmath import BN, BLS, to_limbs, from_limbs, splitmix64, random_below

pytestmark = pytest.mark.gpu

FIELDS = [(BN, 0), (BN, 1), (BLS, 0), (BLS, 1)]


def field_params(C, field):
    if field == 0:
        return C.p, C.lp
    return C.r, 4


def sample(C, field, n, seed):
    p, L = field_params(C, field)
    R = 1 << (64 * L)
    rng = splitmix64(seed)
    edge = [0, 1, p - 1, p - 2, R % p, (R * R) % p, (1 << (p.bit_length() - 1)) - 1, p // 2, p // 2 + 1]
    xs = edge + [random_below(rng, p, L) for _ in range(n - len(edge))]
    return xs


def to_arr(xs, L):
    return np.array([to_limbs(x, L) for x in xs], dtype=np.uint64)


@pytest.mark.parametrize("C,field", FIELDS, ids=["bn254-fp", "bn254-fr", "bls-fp", "bls-fr"])
def test_field_ops(zk, C, field):
    p, L = field_params(C, field)
    R = 1 << (64 * L)
    Rinv = pow(R, -1, p)
    n = 4096
    xs = sample(C, field, n, 1000 + 10 * C.id + field)
    ys = sample(C, field, n, 2000 + 10 * C.id + field)[::-1]
    A, B = to_arr(xs, L), to_arr(ys, L)
    dA, dB = zk.DeviceArray.from_host(A), zk.DeviceArray.from_host(B)
    dO = zk.DeviceArray(n * L)
    expect = {
        "add": [(x + y) % p for x, y in zip(xs, ys)],
        "sub": [(x - y) % p for x, y in zip(xs, ys)],
        "mul": [x * y * Rinv % p for x, y in zip(xs, ys)],
        "sqr": [x * x * Rinv % p for x in xs],
        "neg": [(-x) % p for x in xs],
        "half": [x * pow(2, -1, p) % p for x in xs],
        "to_std": [x * Rinv % p for x in xs],
        "from_std": [x * R % p for x in xs],
    }
    for op, exp in expect.items():
        zk.field_op(C.id, field, zk.FOP[op], dA, dB, dO, n)
        got = [from_limbs(r) for r in dO.to_host((n, L))]
        bad = [i for i in range(n) if got[i] != exp[i]]
        assert not bad, f"{op}: {len(bad)} mismatches, first at {bad[0]}"
    # inverse on a smaller sample (Fermat on device is ~L*64 multiplications)
    m = 256
    zk.field_op(C.id, field, zk.FOP["inv"], dA, None, dO, m)
    got = [from_limbs(r) for r in dO.to_host((n, L))[:m]]
    for i in range(m):
        x = xs[i]
        exp = 0 if x == 0 else pow(x * Rinv % p, -1, p) * R % p  # Montgomery form of 1/(x R^-1)
        assert got[i] == exp, i
