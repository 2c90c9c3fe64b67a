"""CPU test of the host-side field and XYZZ arithmetic (zikkurat-algebra_amd/csrc/hostfield.hpp):
the MSM's final window combination, the NTT's twiddle seeds and N^-1, and the generator checks
run on it.  Built for the host through tests/native/hostfield_harness.cpp and checked against
Python integers: a Montgomery product is a b R^-1 mod p, an inverse satisfies a a^-1 = 1, and
the XYZZ doubling and addition normalise to the affine group law.

Operands sit at the edges the branch-free reductions have to get right (0, 1, p - 1, R mod p,
pairs whose sum or difference lands exactly on p or 0) besides random values.
"""
import ctypes
import subprocess
from pathlib import Path

import numpy as np
import pytest

from zkmath import BN, BLS, splitmix64, random_below, to_limbs, from_limbs

REPO = Path(__file__).resolve().parent.parent
SRC = REPO / "tests" / "native" / "hostfield_harness.cpp"
LIBSRC = REPO / "zikkurat-algebra_amd" / "csrc" / "hostfield.cpp"
HDR = REPO / "zikkurat-algebra_amd" / "csrc" / "hostfield.hpp"
OUT = REPO / "tests" / "native" / "_build" / "libhfh.so"
u64p = ctypes.POINTER(ctypes.c_uint64)

# field id -> (prime, 64-bit limbs)
FIELDS = {0: (BN.p, 4), 1: (BN.r, 4), 2: (BLS.p, 6), 3: (BLS.r, 4)}
IDS = ["bn-fp", "bn-fr", "bls-fp", "bls-fr"]


@pytest.fixture(scope="module")
def h():
    OUT.parent.mkdir(exist_ok=True)
    if not OUT.exists() or OUT.stat().st_mtime < max(p.stat().st_mtime for p in (SRC, LIBSRC, HDR)):
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", str(OUT), str(SRC), str(LIBSRC)],
                       check=True, capture_output=True)
    return ctypes.CDLL(str(OUT))


def words(vals, nl):
    return np.ascontiguousarray(np.array([to_limbs(v, nl) for v in vals], dtype=np.uint64))


def field_op(h, field, op, a, b=None):
    p, nl = FIELDS[field]
    A = words(a, nl)
    B = words(b, nl) if b is not None else None
    out = np.zeros_like(A)
    h.hfh_field_op(field, op, len(a), A.ctypes.data_as(u64p), B.ctypes.data_as(u64p) if B is not None else None,
                   out.ctypes.data_as(u64p))
    return [from_limbs(list(map(int, row))) for row in out]


def operands(p, nl, seed, k=200):
    rng = splitmix64(seed)
    R = 1 << (64 * nl)
    edge = [0, 1, 2, p - 1, p - 2, R % p, (p + 1) // 2, (p - 1) // 2]
    a = [x for x in edge for _ in edge]
    b = [y for _ in edge for y in edge]
    while len(a) < k:
        a.append(random_below(rng, p, nl))
        b.append(random_below(rng, p, nl))
    # sums that land exactly on p, differences that land exactly on 0
    for _ in range(8):
        x = random_below(rng, p, nl)
        a += [x, x]
        b += [(p - x) % p, x]
    return a, b


@pytest.mark.parametrize("field", [0, 1, 2, 3], ids=IDS)
def test_field_ops_match_integers(h, field):
    p, nl = FIELDS[field]
    R = 1 << (64 * nl)
    Rinv = pow(R, -1, p)
    a, b = operands(p, nl, 11 + field)
    assert field_op(h, field, 0, a, b) == [x * y * Rinv % p for x, y in zip(a, b)]
    assert field_op(h, field, 7, a) == [x * x * Rinv % p for x in a]
    assert field_op(h, field, 1, a, b) == [(x + y) % p for x, y in zip(a, b)]
    assert field_op(h, field, 2, a, b) == [(x - y) % p for x, y in zip(a, b)]
    assert field_op(h, field, 3, a) == [(-x) % p for x in a]
    assert field_op(h, field, 5, a) == [x * Rinv % p for x in a]
    assert field_op(h, field, 6, a) == [x * R % p for x in a]


@pytest.mark.parametrize("field", [0, 1, 2, 3], ids=IDS)
def test_inverse(h, field):
    p, nl = FIELDS[field]
    R = 1 << (64 * nl)
    a, _ = operands(p, nl, 21 + field, k=40)
    inv = field_op(h, field, 4, a)
    for x, y in zip(a, inv):
        if x == 0:
            assert y == 0
        else:  # Montgomery inverse: (x R^-1)^-1 R = x^-1 R^2
            assert y == pow(x, -1, p) * R * R % p


def xyzz_words(curve, P, z):
    """Montgomery XYZZ words of the affine point P scaled by z (ZZ = z^2, ZZZ = z^3)."""
    L = curve.lp
    if P is None:
        return [0] * (4 * L)
    x, y = P
    p = curve.p
    vals = (x * z * z % p, y * z * z * z % p, z * z % p, z * z * z % p)
    out = []
    for v in vals:
        out += to_limbs(curve.fp_to_mont(v), L)
    return out


def g1_op(h, curve, op, A, B=None):
    L = curve.lp
    A = np.ascontiguousarray(np.array(A, dtype=np.uint64))
    B = np.ascontiguousarray(np.array(B, dtype=np.uint64)) if B is not None else None
    n = A.shape[0]
    out = np.zeros((n, 2 * L if op == 2 else 4 * L), np.uint64)
    h.hfh_g1_op(curve.id, op, n, A.ctypes.data_as(u64p), B.ctypes.data_as(u64p) if B is not None else None,
                out.ctypes.data_as(u64p))
    return out


@pytest.mark.parametrize("curve", [BN, BLS], ids=["bn254", "bls12_381"])
def test_xyzz_dbl_add_to_affine(h, curve):
    rng = splitmix64(31 + curve.id)
    G = curve.g
    ks = [1, 2, 3, curve.r - 1, curve.r - 2] + [random_below(rng, curve.r, 4) for _ in range(11)]
    pts = [curve.mul(k, G) for k in ks]
    zs = [1 + random_below(rng, curve.p - 1, curve.lp) for _ in ks]
    A = [xyzz_words(curve, P, z) for P, z in zip(pts, zs)]
    # doubling
    D = g1_op(h, curve, 0, A)
    assert [curve.words_to_affine(w) for w in g1_op(h, curve, 2, D)] == [curve.add(P, P) for P in pts]
    # addition: generic pairs, P + P (the doubling branch), P + (-P) and infinity on either side
    Q = pts[1:] + [pts[0]]
    cases = list(zip(pts, Q)) + [(pts[3], pts[3]), (pts[4], curve.neg(pts[4])), (None, pts[5]), (pts[6], None)]
    zA = [1 + random_below(rng, curve.p - 1, curve.lp) for _ in cases]
    zB = [1 + random_below(rng, curve.p - 1, curve.lp) for _ in cases]
    A2 = [xyzz_words(curve, P, z) for (P, _), z in zip(cases, zA)]
    B2 = [xyzz_words(curve, Q, z) for (_, Q), z in zip(cases, zB)]
    S = g1_op(h, curve, 1, A2, B2)
    assert [curve.words_to_affine(w) for w in g1_op(h, curve, 2, S)] == [curve.add(P, Q) for P, Q in cases]
