"""The BN254 GLV constants used by the MSM digit kernel and the G1 FFT
(zikkurat-algebra_amd/csrc/glv.hpp, struct Bn254Glv), read from the source and checked against what the mathematics fixes:
beta and lambda are primitive cube roots of unity in Fp and Fr with (beta x, y) = lambda (x, y)
on G1; (a1, b1) and (a2, b2) lie in the lattice {(x, y) : x + lambda y = 0 mod r} with
determinant r; G1, G2 are the stated roundings; the internal form of beta is beta 2^261 mod p
in 29-bit limbs; and the split stays below 2^127 in magnitude over edge cases and random
scalars.  A wrong limb in any constant fails one of these."""
import random
import re
from pathlib import Path

from zkmath import BN

SRC = Path(__file__).resolve().parent.parent / "zikkurat-algebra_amd" / "csrc" / "glv.hpp"


def constants():
    text = SRC.read_text()
    body = text[text.index("struct Bn254Glv {"):]
    body = body[:body.index("\n};")]

    def arr(name):
        m = re.search(name + r"\[\d*\]\s*=\s*\{([^}]*)\}", body)
        return [int(re.sub(r"u?ll$", "", v.strip()), 0) for v in m.group(1).split(",") if v.strip()]

    def one(name):
        return int(re.search(name + r"\s*=\s*(0x[0-9a-fA-F]+)ull", body).group(1), 16)

    def join(limbs, bits):
        return sum(v << (bits * i) for i, v in enumerate(limbs))

    return dict(G1=join(arr("G1"), 64), G2=join(arr("G2"), 64), A1=join(arr("A1"), 64), A2=one("A2"),
                B2=join(arr("B2"), 64), BETA=join(arr("BETA"), 29))


def cube_roots(m):
    out = set()
    for g in range(2, 50):
        w = pow(g, (m - 1) // 3, m)
        if w != 1:
            out.update({w, w * w % m})
    return out


def test_glv_constants():
    k = constants()
    p, r = BN.p, BN.r
    a1, a2, b2, b1 = k["A1"], k["A2"], k["B2"], -k["A2"]  # |b1| = a2 for this basis
    # beta, lambda: the pair with phi(G) = lambda G
    beta = k["BETA"] * pow(1 << 261, -1, p) % p
    assert pow(beta, 3, p) == 1 and beta != 1
    gx, gy = BN.g
    lams = [lam for lam in cube_roots(r) if BN.mul(lam, BN.g) == (beta * gx % p, gy)]
    assert len(lams) == 1
    lam = lams[0]
    # the lattice basis and its determinant
    assert (a1 + b1 * lam) % r == 0 and (a2 + b2 * lam) % r == 0
    assert a1 * b2 - a2 * b1 == r
    # the roundings
    assert k["G1"] == ((b2 << 256) + r // 2) // r
    assert k["G2"] == ((-b1 << 256) + r // 2) // r

    def split(s):
        c1, c2 = (s * k["G1"]) >> 256, (s * k["G2"]) >> 256
        return s - c1 * a1 - c2 * a2, c1 * k["A2"] - c2 * b2

    rng = random.Random(17)
    for s in [0, 1, r - 1, lam, r - lam, (r - 1) // 2] + [rng.randrange(r) for _ in range(20000)]:
        k1, k2 = split(s)
        assert (k1 + lam * k2 - s) % r == 0
        assert abs(k1) < 2**127 and abs(k2) < 2**127
