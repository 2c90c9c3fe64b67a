"""ctypes loaders for the CPU checkers: the oracle restatement (oracle/libzkoracle.so) and,
where it has been built, the reference's own C (oracle/_ref/libzkref.so).

Test infrastructure only.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

from zkmath import to_limbs

REPO = Path(__file__).resolve().parent.parent
ORACLE_SO = REPO / "oracle" / "libzkoracle.so"
REF_SO = REPO / "oracle" / "_ref" / "libzkref.so"

u64p = ctypes.POINTER(ctypes.c_uint64)


def ptr(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(u64p)


_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        if not ORACLE_SO.exists():
            import subprocess
            subprocess.run(["make", "-C", str(REPO / "oracle")], check=True, capture_output=True)
        _oracle = ctypes.CDLL(str(ORACLE_SO))
    return _oracle


def ref():
    """The reference C built from its sources, or None when it has not been built."""
    global _ref
    if _ref is None and REF_SO.exists():
        _ref = ctypes.CDLL(str(REF_SO))
    return _ref


class Oracle:
    """Thin numpy-facing wrapper over oracle/zko.h."""

    def __init__(self, curve):
        self.c = curve
        self.lib = oracle()
        self.lp = curve.lp

    def f_mul(self, field, a, b):
        n = 4 if field == 1 else self.lp
        out = np.zeros(n, np.uint64)
        self.lib.zko_f_mul(self.c.id, field, ptr(a), ptr(b), ptr(out))
        return out

    def f_sub(self, field, a, b):
        n = 4 if field == 1 else self.lp
        out = np.zeros(n, np.uint64)
        self.lib.zko_f_sub(self.c.id, field, ptr(np.ascontiguousarray(a)), ptr(np.ascontiguousarray(b)), ptr(out))
        return out

    def msm_affine(self, scalars, points, enl=4, mont=False):
        n = points.shape[0]
        out = np.zeros(2 * self.lp, np.uint64)
        fn = self.lib.zko_msm_mont_affine if mont else self.lib.zko_msm_std_affine
        fn(self.c.id, n, ptr(np.ascontiguousarray(scalars)), ptr(np.ascontiguousarray(points)),
           ptr(out), enl)
        return out

    def msm_proj_variable(self, scalars, points, window, enl=4):
        out = np.zeros(3 * self.lp, np.uint64)
        self.lib.zko_msm_std_proj_variable(self.c.id, points.shape[0], ptr(scalars), ptr(points),
                                           ptr(out), enl, window)
        return out

    def fft_generator(self, m):
        out = np.zeros(4, np.uint64)
        self.lib.zko_fft_generator(self.c.id, m, ptr(out))
        return out

    def ntt(self, m, gen, src, inverse=False):
        out = np.zeros_like(src)
        fn = self.lib.zko_ntt_inverse if inverse else self.lib.zko_ntt_forward
        fn(self.c.id, m, ptr(gen), ptr(src), ptr(out))
        return out

    def ntt_shifted(self, eta, m, gen, src, inverse=False):
        out = np.zeros_like(src)
        fn = self.lib.zko_ntt_inverse_shifted if inverse else self.lib.zko_ntt_forward_shifted
        fn(self.c.id, ptr(eta), m, ptr(gen), ptr(src), ptr(out))
        return out

    def ntt_asym(self, ms, mt, gs, gt, src):
        out = np.zeros(((1 << mt), 4), np.uint64)
        self.lib.zko_ntt_forward_asymmetric(self.c.id, ms, mt, ptr(gs), ptr(gt), ptr(src), ptr(out))
        return out

    def points_from_scalars(self, ks):
        """Affine Montgomery points k_i * G (0 -> the affine infinity (0,0))."""
        L = self.lp
        out = np.zeros((len(ks), 2 * L), np.uint64)
        G = np.array(self.c.affine_to_words(self.c.g) + [0] * L, np.uint64)
        G[2 * L:3 * L] = np.array(to_limbs(self.c.fp_to_mont(1), L), np.uint64)
        P = np.zeros(3 * L, np.uint64)
        for i, k in enumerate(ks):
            if k == 0:
                continue
            K = np.array(to_limbs(k, 4), np.uint64)
            self.lib.zko_g1_proj_scl(self.c.id, ptr(K), 4, ptr(G), ptr(P))
            self.lib.zko_g1_proj_to_affine(self.c.id, ptr(P), ptr(out[i]))
        return out

    def proj_to_affine(self, P):
        out = np.zeros(2 * self.lp, np.uint64)
        self.lib.zko_g1_proj_to_affine(self.c.id, ptr(P), ptr(out))
        return out

    def proj_points(self, ks, zs=None):
        """Projective Montgomery points k_i G with a non-trivial Z (X z, Y z, z), infinity (0:1:0)."""
        L = self.lp
        aff = self.points_from_scalars(ks)
        out = np.zeros((len(ks), 3 * L), np.uint64)
        one = np.array(to_limbs(self.c.fp_to_mont(1), L), np.uint64)
        for i, k in enumerate(ks):
            if k == 0:
                out[i, L:2 * L] = one
                continue
            z = np.array(to_limbs(self.c.fp_to_mont((zs[i] if zs else 7 + i) % self.c.p), L), np.uint64)
            for j, w in enumerate((aff[i, :L], aff[i, L:], one)):
                self.lib.zko_f_mul(self.c.id, 0, ptr(np.ascontiguousarray(w)), ptr(z), ptr(out[i, j * L:(j + 1) * L]))
        return out

    def g1_fft(self, m, gen, src, inverse=False, jac=False):
        out = np.zeros_like(src)
        if jac:
            fn = self.lib.zko_g1_jac_fft_inverse if inverse else self.lib.zko_g1_jac_fft_forward
        else:
            fn = self.lib.zko_g1_fft_inverse if inverse else self.lib.zko_g1_fft_forward
        fn(self.c.id, m, ptr(np.ascontiguousarray(gen)), ptr(np.ascontiguousarray(src)), ptr(out))
        return out

    def poly_eval(self, coeffs, x):
        out = np.zeros(4, np.uint64)
        c = np.ascontiguousarray(coeffs, dtype=np.uint64)
        self.lib.zko_poly_eval(self.c.id, c.size // 4, ptr(c), ptr(np.ascontiguousarray(x)), ptr(out))
        return out

    def div_by_vanishing(self, coeffs, n, eta, nquot, nrem):
        """(status, quotient, remainder); status -1 where the reference asserts."""
        c = np.ascontiguousarray(coeffs, dtype=np.uint64)
        q = np.zeros((max(nquot, 1), 4), np.uint64)
        r = np.zeros((max(nrem, 1), 4), np.uint64)
        st = self.lib.zko_poly_div_by_vanishing(self.c.id, c.size // 4, ptr(c), n, ptr(np.ascontiguousarray(eta)),
                                                nquot, ptr(q), nrem, ptr(r) if nrem else None)
        return st, q[:nquot], r[:nrem]

    def quot_by_vanishing(self, coeffs, n, eta, nquot):
        c = np.ascontiguousarray(coeffs, dtype=np.uint64)
        q = np.zeros((max(nquot, 1), 4), np.uint64)
        st = self.lib.zko_poly_quot_by_vanishing(self.c.id, c.size // 4, ptr(c), n, ptr(np.ascontiguousarray(eta)),
                                                 nquot, ptr(q))
        return st, q[:nquot]

    def batch_to_affine(self, P, jac=False):
        P = np.ascontiguousarray(P, dtype=np.uint64)
        n = P.size // (3 * self.lp)
        out = np.zeros((n, 2 * self.lp), np.uint64)
        self.lib.zko_g1_batch_to_affine(self.c.id, int(jac), n, ptr(P), ptr(out))
        return out

    def batch_from_affine(self, A, jac=False):
        A = np.ascontiguousarray(A, dtype=np.uint64)
        n = A.size // (2 * self.lp)
        out = np.zeros((n, 3 * self.lp), np.uint64)
        self.lib.zko_g1_batch_from_affine(self.c.id, int(jac), n, ptr(A), ptr(out))
        return out

    def msm_binary(self, bits, points):
        """Normalised projective sum of the points whose byte is non-zero."""
        bits = np.ascontiguousarray(bits, dtype=np.uint8)
        out = np.zeros(3 * self.lp, np.uint64)
        self.lib.zko_msm_binary_proj(self.c.id, len(bits), bits.ctypes.data_as(ctypes.c_void_p),
                                     ptr(np.ascontiguousarray(points)), ptr(out))
        return out

    def multi_msm_binary(self, rows, points):
        rows = np.ascontiguousarray(rows, dtype=np.uint8)
        nr, n = rows.shape
        out = np.zeros((nr, 3 * self.lp), np.uint64)
        ptrs = (ctypes.c_void_p * max(nr, 1))(*[rows.ctypes.data + r * n for r in range(nr)])
        self.lib.zko_multi_msm_binary_proj(self.c.id, nr, n, ptrs, ptr(np.ascontiguousarray(points)), ptr(out))
        return out

    def jac_points(self, ks, zs=None):
        """Jacobian Montgomery points k_i G with a non-trivial Z (x z^2, y z^3, z), infinity [1:1:0]."""
        L = self.lp
        aff = self.points_from_scalars(ks)
        out = np.zeros((len(ks), 3 * L), np.uint64)
        one = np.array(to_limbs(self.c.fp_to_mont(1), L), np.uint64)
        for i, k in enumerate(ks):
            if k == 0:
                out[i, :L] = one
                out[i, L:2 * L] = one
                continue
            zi = (zs[i] if zs else 5 + 3 * i) % self.c.p
            z2, z3 = zi * zi % self.c.p, zi * zi * zi % self.c.p
            for j, (w, zz) in enumerate(((aff[i, :L], z2), (aff[i, L:], z3))):
                zm = np.array(to_limbs(self.c.fp_to_mont(zz), L), np.uint64)
                self.lib.zko_f_mul(self.c.id, 0, ptr(np.ascontiguousarray(w)), ptr(zm), ptr(out[i, j * L:(j + 1) * L]))
            out[i, 2 * L:] = np.array(to_limbs(self.c.fp_to_mont(zi), L), np.uint64)
        return out


class Ref:
    """The reference's own entry points (names from lib/cbits/.../*.h)."""

    def __init__(self, curve):
        self.c = curve
        self.lib = ref()
        self.lp = curve.lp
        self.pre = curve.name

    def fn(self, name):
        return getattr(self.lib, f"{self.pre}_{name}")

    def msm(self, scalars, points, coord="proj", mont=False, out="affine", enl=4):
        n = points.shape[0]
        size = 2 * self.lp if out == "affine" else 3 * self.lp
        res = np.zeros(size, np.uint64)
        kind = "mont" if mont else "std"
        outname = "affine" if out == "affine" else coord
        self.fn(f"G1_{coord}_MSM_{kind}_coeff_{outname}_out")(
            n, ptr(np.ascontiguousarray(scalars)), ptr(np.ascontiguousarray(points)), ptr(res), enl)
        return res

    def msm_variable(self, scalars, points, window, enl=4):
        res = np.zeros(3 * self.lp, np.uint64)
        self.fn("G1_proj_MSM_std_coeff_proj_out_variable")(
            points.shape[0], ptr(scalars), ptr(points), ptr(res), enl, window)
        return res

    def ntt(self, m, gen, src, inverse=False):
        out = np.zeros_like(src)
        self.fn("poly_mont_ntt_inverse" if inverse else "poly_mont_ntt_forward")(
            m, ptr(gen), ptr(src), ptr(out))
        return out

    def ntt_shifted(self, eta, m, gen, src, inverse=False):
        out = np.zeros_like(src)
        self.fn("poly_mont_ntt_inverse_shifted" if inverse else "poly_mont_ntt_forward_shifted")(
            ptr(eta), m, ptr(gen), ptr(src), ptr(out))
        return out

    def ntt_asym(self, ms, mt, gs, gt, src):
        out = np.zeros(((1 << mt), 4), np.uint64)
        self.fn("poly_mont_ntt_forward_asymmetric")(ms, mt, ptr(gs), ptr(gt), ptr(src), ptr(out))
        return out

    def f_mul(self, field, a, b):
        n = 4 if field == 1 else self.lp
        out = np.zeros(n, np.uint64)
        self.fn("Fr_mont_mul" if field == 1 else "Fp_mont_mul")(ptr(a), ptr(b), ptr(out))
        return out

    def g1_fft(self, m, gen, src, inverse=False, coord="proj"):
        out = np.zeros_like(src)
        self.fn(f"G1_{coord}_fft_{'inverse' if inverse else 'forward'}")(
            m, ptr(np.ascontiguousarray(gen)), ptr(np.ascontiguousarray(src)), ptr(out))
        return out

    def poly_eval(self, coeffs, x):
        out = np.zeros(4, np.uint64)
        c = np.ascontiguousarray(coeffs, dtype=np.uint64)
        self.fn("poly_mont_eval_at")(c.size // 4, ptr(c), ptr(np.ascontiguousarray(x)), ptr(out))
        return out

    def div_by_vanishing(self, coeffs, n, eta, nquot, nrem):
        c = np.ascontiguousarray(coeffs, dtype=np.uint64)
        q = np.zeros((max(nquot, 1), 4), np.uint64)
        r = np.zeros((max(nrem, 1), 4), np.uint64)
        self.fn("poly_mont_div_by_vanishing")(c.size // 4, ptr(c), n, ptr(np.ascontiguousarray(eta)), nquot, ptr(q),
                                              nrem, ptr(r) if nrem else None)
        return q[:nquot], r[:nrem]

    def quot_by_vanishing(self, coeffs, n, eta, nquot):
        c = np.ascontiguousarray(coeffs, dtype=np.uint64)
        q = np.zeros((max(nquot, 1), 4), np.uint64)
        f = self.fn("poly_mont_quot_by_vanishing")
        f.restype = ctypes.c_uint8
        ok = f(c.size // 4, ptr(c), n, ptr(np.ascontiguousarray(eta)), nquot, ptr(q))
        return int(ok), q[:nquot]

    def batch_to_affine(self, P, coord="proj"):
        P = np.ascontiguousarray(P, dtype=np.uint64)
        n = P.size // (3 * self.lp)
        out = np.zeros((n, 2 * self.lp), np.uint64)
        self.fn(f"G1_{coord}_batch_to_affine")(n, ptr(P), ptr(out))
        return out

    def batch_from_affine(self, A, coord="proj"):
        A = np.ascontiguousarray(A, dtype=np.uint64)
        n = A.size // (2 * self.lp)
        out = np.zeros((n, 3 * self.lp), np.uint64)
        self.fn(f"G1_{coord}_batch_from_affine")(n, ptr(A), ptr(out))
        return out

    def msm_binary(self, bits, points, coord="proj", out="affine"):
        bits = np.ascontiguousarray(bits, dtype=np.uint8)
        res = np.zeros(2 * self.lp if out == "affine" else 3 * self.lp, np.uint64)
        self.fn(f"G1_{coord}_MSM_binary_coeff_{out}_out")(
            len(bits), bits.ctypes.data_as(ctypes.c_void_p), ptr(np.ascontiguousarray(points)), ptr(res))
        return res

    def multi_msm_binary(self, rows, points, coord="proj", out="affine"):
        rows = np.ascontiguousarray(rows, dtype=np.uint8)
        nr, n = rows.shape
        res = np.zeros((nr, 2 * self.lp if out == "affine" else 3 * self.lp), np.uint64)
        ptrs = (ctypes.c_void_p * max(nr, 1))(*[rows.ctypes.data + r * n for r in range(nr)])
        self.fn(f"G1_{coord}_multi_MSM_binary_coeff_{out}_out")(nr, n, ptrs, ptr(np.ascontiguousarray(points)),
                                                                ptr(res))
        return res
