"""Pure-Python big-integer mathematics used to pin the oracle and the GPU library.

Nothing here restates the reference's algorithms: it is textbook modular and elliptic-curve
arithmetic (affine formulas with modular inverses), an O(N^2) DFT, Horner evaluation and the
seeded input generators shared by the tests and the benchmark.  Each closed form is chosen so
that a plausible slip in an optimised implementation (a dropped term, a wrong sign or index,
a transposed operand) changes the answer.
"""
from __future__ import annotations

import numpy as np

# Primes: lib/cbits/curves/fields/mont/bn128_Fp_mont.c:20, bn128_Fr_mont.c:19,
# bls12_381_Fp_mont.c:20, bls12_381_Fr_mont.c:20 (codegen/src/Zikkurat/Primes.hs:75-82).
BN254_P = 21888242871839275222246405745257275088696311157297823662689037894645226208583
BN254_R = 21888242871839275222246405745257275088548364400416034343698204186575808495617
BLS_P = int("1a0111ea397fe69a4b1ba7b6434bacd764774b84f38512bf6730d2a0f6b0f624"
            "1eabfffeb153ffffb9feffffffffaaab", 16)
BLS_R = 0x73eda753299d7d483339d80809a1d80553bda402fffe5bfeffffffff00000001

BN254, BLS12_381 = 0, 1
CURVE_NAMES = {BN254: "bn128", BLS12_381: "bls12_381"}


class Curve:
    def __init__(self, cid, p, r, b, gx, gy, lp, fft_gen, fft_log):
        self.id = cid
        self.name = CURVE_NAMES[cid]
        self.p, self.r, self.b = p, r, b
        self.g = (gx, gy)
        self.lp = lp            # 64-bit limbs of Fp
        self.lr = 4             # 64-bit limbs of Fr
        self.Rp = 1 << (64 * lp)
        self.Rr = 1 << 256
        self.fft_gen = fft_gen  # standard form
        self.fft_log = fft_log

    # ---- field conversions ----
    def fp_to_mont(self, x):
        return x * self.Rp % self.p

    def fp_from_mont(self, x):
        return x * pow(self.Rp, -1, self.p) % self.p

    def fr_to_mont(self, x):
        return x * self.Rr % self.r

    def fr_from_mont(self, x):
        return x * pow(self.Rr, -1, self.r) % self.r

    # ---- affine group law (None = infinity) ----
    def add(self, P, Q):
        if P is None:
            return Q
        if Q is None:
            return P
        p = self.p
        (x1, y1), (x2, y2) = P, Q
        if x1 == x2:
            if (y1 + y2) % p == 0:
                return None
            lam = 3 * x1 * x1 * pow(2 * y1, -1, p) % p
        else:
            lam = (y2 - y1) * pow(x2 - x1, -1, p) % p
        x3 = (lam * lam - x1 - x2) % p
        return (x3, (lam * (x1 - x3) - y1) % p)

    def neg(self, P):
        return None if P is None else (P[0], (-P[1]) % self.p)

    def mul(self, k, P):
        R = None
        if k < 0:
            k, P = -k, self.neg(P)
        while k:
            if k & 1:
                R = self.add(R, P)
            P = self.add(P, P)
            k >>= 1
        return R

    def on_curve(self, P):
        if P is None:
            return True
        x, y = P
        return (y * y - x * x * x - self.b) % self.p == 0

    # ---- flat-array encodings (reference layout: little-endian 64-bit limbs) ----
    def affine_to_words(self, P):
        """Montgomery affine encoding; infinity is (0,0) on input (bn128_G1_affine.c:57-66)."""
        if P is None:
            return [0] * (2 * self.lp)
        return to_limbs(self.fp_to_mont(P[0]), self.lp) + to_limbs(self.fp_to_mont(P[1]), self.lp)

    def words_to_affine(self, w):
        """Decodes a to_affine output; all-0xFF words mean infinity (bn128_G1_proj.c:132-137)."""
        w = [int(x) for x in w]
        if all(x == 0xFFFFFFFFFFFFFFFF for x in w):
            return None
        x = from_limbs(w[:self.lp])
        y = from_limbs(w[self.lp:2 * self.lp])
        return (self.fp_from_mont(x), self.fp_from_mont(y))

    def proj_words_to_affine(self, w):
        """Decodes homogeneous projective (X:Y:Z) Montgomery words."""
        w = [int(x) for x in w]
        L = self.lp
        X, Y, Z = (self.fp_from_mont(from_limbs(w[i * L:(i + 1) * L])) for i in range(3))
        if Z == 0:
            return None
        zi = pow(Z, -1, self.p)
        return (X * zi % self.p, Y * zi % self.p)

    def jac_words_to_affine(self, w):
        w = [int(x) for x in w]
        L = self.lp
        X, Y, Z = (self.fp_from_mont(from_limbs(w[i * L:(i + 1) * L])) for i in range(3))
        if Z == 0:
            return None
        zi = pow(Z, -1, self.p)
        return (X * zi * zi % self.p, Y * zi * zi * zi % self.p)

    def fft_root(self, m):
        """Generator of the size-2^m FFT subgroup (Class/FFT.hs:60-68), standard form."""
        return pow(self.fft_gen, 1 << (self.fft_log - m), self.r)


BN = Curve(BN254, BN254_P, BN254_R, 3, 1, 2, 4,
           19103219067921713944291392827692070036145651957329286315305642004821462161904, 28)
BLS = Curve(BLS12_381, BLS_P, BLS_R, 4,
            int("17f1d3a73197d7942695638c4fa9ac0fc3688c4f9774b905a14e3a3f171bac586c55e83ff97a1aeffb3af00adb22c6bb", 16),
            int("08b3f481e3aaa0f1a09e30ed741d8ae4fcf5e095d5d00af600db18cb2c04b3edd03cc744a2888ae40caa232946c5e7e1", 16),
            6,
            10238227357739495823651030575849232062558860180284477541189508159991286009131, 32)
CURVES = {BN254: BN, BLS12_381: BLS}


def to_limbs(x, n):
    return [(x >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(n)]


def from_limbs(ws):
    return sum(int(w) << (64 * i) for i, w in enumerate(ws))


def ints_to_array(xs, n):
    out = np.zeros((len(xs), n), dtype=np.uint64)
    for i, x in enumerate(xs):
        out[i] = to_limbs(x, n)
    return out


def array_to_ints(a):
    a = np.asarray(a, dtype=np.uint64).reshape(-1, a.shape[-1])
    return [from_limbs(row) for row in a]


# ---------------- seeded generators ----------------

def splitmix64(seed):
    """splitmix64 stream (Steele, Lea, Flood 2014)."""
    x = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        yield z ^ (z >> 31)


def random_below(rng, bound, nlimbs=4):
    """Uniform in [0, bound) by rejection on the bit length of bound."""
    bits = bound.bit_length()
    while True:
        v = from_limbs([next(rng) for _ in range(nlimbs)]) & ((1 << bits) - 1)
        if v < bound:
            return v


def random_scalars(seed, n, r, dist="uniform"):
    rng = splitmix64(seed)
    out = []
    for _ in range(n):
        if dist == "uniform":
            out.append(random_below(rng, r))
        elif dist == "u64":
            out.append(next(rng))
        elif dist == "sparse":
            u = next(rng) % 10
            out.append(random_below(rng, r) if u >= 8 else (u & 1))
        else:
            raise ValueError(dist)
    return out


# ---------------- NTT closed forms ----------------

def dft(xs, g, r):
    """O(N^2) DFT: y_k = sum_i x_i g^(ik) mod r (standard-form integers)."""
    N = len(xs)
    return [sum(x * pow(g, i * k, r) for i, x in enumerate(xs)) % r for k in range(N)]


def horner(coeffs, x, r):
    acc = 0
    for c in reversed(coeffs):
        acc = (acc * x + c) % r
    return acc


def seeded_fr_vector(seed, n, r):
    """n uniform elements of [0, r) as an (n, 4) uint64 array (numpy PCG64 stream, rejection
    on the top word); deterministic for a seed, so large inputs need not be stored."""
    rng = np.random.default_rng(seed)
    bits = r.bit_length()
    top_mask = np.uint64((1 << (bits - 192)) - 1)
    r_top = np.uint64(r >> 192)
    out = np.empty((n, 4), np.uint64)
    filled = 0
    while filled < n:
        k = int((n - filled) * 1.4) + 16
        cand = rng.integers(0, 2**64, size=(k, 4), dtype=np.uint64)
        cand[:, 3] &= top_mask
        ok = cand[:, 3] < r_top
        for i in np.nonzero(cand[:, 3] == r_top)[0]:
            ok[i] = from_limbs(cand[i]) < r
        acc = cand[ok][: n - filled]
        out[filled:filled + len(acc)] = acc
        filled += len(acc)
    return out
