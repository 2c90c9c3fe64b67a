"""Python host interface of the MI355X MSM/NTT library (ctypes over include/zkgpu.h).

The names mirror the reference's polymorphic API so that parity tests read like the
reference's own properties:

* ``msm`` / ``msm_std``       -- Class/Curve.hs:92 ``msm`` (Montgomery scalars) and the
                                 standard-form ``msmStd`` (G1/Proj.hs:241-255)
* ``ntt`` / ``intt``          -- Class/Poly.hs:57-58 (BN128/Poly.hs:406-429)
* ``shifted_ntt`` / ``shifted_intt`` -- Class/Poly.hs:60-63 (BN128/Poly.hs:433-464)
* ``asymm_ntt``               -- Class/Poly.hs:65 (BN128/Poly.hs:468-485)

Arrays are numpy ``uint64`` with the reference's flat limb layout.  Size mismatches raise
``ValueError`` like the Haskell layer's ``error`` (Poly.hs:409, Proj.hs:248); device failures
raise ``ZkgError``.  There is no CPU fallback: without the built library (or a GPU) every call
fails.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libzkgpu.so"

BN254, BLS12_381 = 0, 1
PREFIX = {BN254: "bn128", BLS12_381: "bls12_381"}
FP_LIMBS = {BN254: 4, BLS12_381: 6}
FR_LIMBS = 4

MSM_MONT, MSM_PROJ_OUT, MSM_JAC_OUT, MSM_POINTS_INTERNAL = 1, 2, 4, 8
NTT_FORWARD, NTT_INVERSE = 0, 1

u64p = ctypes.POINTER(ctypes.c_uint64)
vp = ctypes.c_void_p


class ZkgError(RuntimeError):
    pass


SCHED_PLAIN, SCHED_STATS, SCHED_LEARNED = 0, 1, 2
MSM_NFEAT = 11


class ScalarStats(ctypes.Structure):
    """zkg_scalar_stats: counts and the largest bit length of a scalar vector."""
    _fields_ = [("n", ctypes.c_longlong), ("n_zero", ctypes.c_longlong), ("n_one", ctypes.c_longlong),
                ("max_bits", ctypes.c_int), ("reserved", ctypes.c_int)]

    def as_dict(self):
        return dict(n=self.n, n_zero=self.n_zero, n_one=self.n_one, max_bits=self.max_bits)


class MsmSchedule(ctypes.Structure):
    """zkg_msm_schedule."""
    _fields_ = [("window", ctypes.c_int), ("seg_log", ctypes.c_int), ("policy", ctypes.c_int),
                ("fix_limit", ctypes.c_int), ("prefetch", ctypes.c_int), ("reserved", ctypes.c_int * 3)]


_lib = None


def lib():
    """Loads libzkgpu.so (built by zikkurat-algebra_amd/build.py); fails loudly if absent."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ZkgError(f"{LIB_PATH} is not built; run python zikkurat-algebra_amd/build.py")
        L = ctypes.CDLL(str(LIB_PATH))
        L.zkg_last_error.restype = ctypes.c_char_p
        for name in ("zkg_malloc",):
            getattr(L, name).argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        L.zkg_free.argtypes = [vp]
        L.zkg_memcpy_h2d.argtypes = [vp, vp, ctypes.c_size_t]
        L.zkg_memcpy_d2h.argtypes = [vp, vp, ctypes.c_size_t]
        L.zkg_msm.argtypes = [ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, vp, u64p, ctypes.c_int, vp]
        L.zkg_ntt.argtypes = [ctypes.c_int, ctypes.c_int, u64p, u64p, ctypes.c_int, vp, vp]
        L.zkg_ntt_noalloc.argtypes = [ctypes.c_int, ctypes.c_int, u64p, u64p, ctypes.c_int, vp, vp]
        L.zkg_fft_generator.argtypes = [ctypes.c_int, ctypes.c_int, u64p]
        L.zkg_field_op.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_size_t, vp, vp, vp]
        L.zkg_g1_scalar_mul_gen.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp]
        L.zkg_microbench.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.zkg_last_timing.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                      ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.zkg_msm_phase_timing.argtypes = [ctypes.c_int]
        L.zkg_ntt_batch_create.argtypes = [ctypes.c_int, ctypes.c_int, vp, ctypes.POINTER(vp)]
        L.zkg_ntt_batch_run.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
        L.zkg_ntt_batch_destroy.argtypes = [vp]
        L.zkg_fill_random.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_size_t, ctypes.c_uint64, vp]
        L.zkg_srs_generate.argtypes = [ctypes.c_int, ctypes.c_int, u64p, vp]
        L.zkg_msm_prepare_points.argtypes = [ctypes.c_int, ctypes.c_int, vp, ctypes.POINTER(vp)]
        L.zkg_serve_create.argtypes = [ctypes.c_int, vp, ctypes.c_int, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(vp)]
        L.zkg_serve_submit.argtypes = [vp, ctypes.c_int, u64p, u64p, ctypes.POINTER(ctypes.c_longlong)]
        L.zkg_serve_drain.argtypes = [vp]
        L.zkg_serve_set_exclusive_log.argtypes = [vp, ctypes.c_int]
        L.zkg_serve_job_times.argtypes = [vp, ctypes.c_longlong] + [ctypes.POINTER(ctypes.c_double)] * 3
        L.zkg_serve_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_int)]
        L.zkg_serve_run_trace.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(u64p), ctypes.POINTER(u64p), ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double)]
        L.zkg_serve_destroy.argtypes = [vp]
        L.zkg_device_info.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_size_t)]
        ip, fp_, dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)
        L.zkg_msm_scalar_stats.argtypes = [ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(ScalarStats)]
        L.zkg_msm_features.argtypes = [ctypes.c_int, ctypes.POINTER(ScalarStats), ctypes.c_int, ctypes.c_int, fp_]
        L.zkg_msm_candidates.argtypes = [ctypes.c_int, ctypes.POINTER(ScalarStats), ctypes.c_int, ip, ip]
        L.zkg_msm_pick.argtypes = [ctypes.c_int, ctypes.POINTER(ScalarStats), ctypes.c_int, ctypes.POINTER(MsmSchedule)]
        L.zkg_msm_tune.argtypes = [ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int,
                                   ip, ip, ctypes.c_int, dp, ip]
        L.zkg_msm_batch.argtypes = [ctypes.c_int, ip, ip, ctypes.POINTER(vp), ctypes.c_int,
                                    ctypes.POINTER(vp), ctypes.POINTER(u64p), ctypes.c_int,
                                    ctypes.POINTER(MsmSchedule), ctypes.c_int, dp]
        L.zkg_msm_batch_create.argtypes = [ctypes.c_int, ip, ip, ctypes.POINTER(vp), ctypes.c_int,
                                           ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(MsmSchedule), ctypes.c_int,
                                           ctypes.POINTER(vp)]
        L.zkg_msm_batch_run.argtypes = [vp, ctypes.POINTER(u64p), dp]
        L.zkg_msm_batch_destroy.argtypes = [vp]
        L.zkg_mlp_create.argtypes = [ctypes.c_int, ip, ctypes.c_uint, ctypes.POINTER(vp)]
        L.zkg_mlp_destroy.argtypes = [vp]
        L.zkg_mlp_nparams.argtypes = [vp]
        L.zkg_mlp_nparams.restype = ctypes.c_longlong
        L.zkg_mlp_get_params.argtypes = [vp, fp_]
        L.zkg_mlp_set_params.argtypes = [vp, fp_]
        L.zkg_mlp_forward.argtypes = [vp, ctypes.c_int, fp_, fp_]
        L.zkg_mlp_gradients.argtypes = [vp, ctypes.c_int, fp_, fp_, fp_, fp_]
        L.zkg_mlp_train.argtypes = [vp, ctypes.c_int, fp_, fp_, ctypes.c_int, ctypes.c_float, fp_]
        L.zkg_msm_set_cost_model.argtypes = [vp]
        L.zkg_ntt_set_tuning.argtypes = [ctypes.c_int, ctypes.c_int]
        L.zkg_g1_fft.argtypes = [ctypes.c_int, ctypes.c_int, u64p, ctypes.c_int, vp, vp]
        L.zkg_flat_file_count.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        L.zkg_flat_read.argtypes = [ctypes.c_char_p, ctypes.c_size_t, vp, ctypes.c_size_t,
                                    ctypes.POINTER(ctypes.c_size_t)]
        L.zkg_flat_write.argtypes = [ctypes.c_char_p, vp, ctypes.c_size_t, ctypes.c_size_t]
        L.zkg_srs_load.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(vp),
                                   ctypes.POINTER(ctypes.c_int)]
        L.zkg_poly_eval.argtypes = [ctypes.c_int, ctypes.c_int, vp, u64p, u64p]
        L.zkg_poly_degree.argtypes = [ctypes.c_int, vp, ctypes.POINTER(ctypes.c_int)]
        L.zkg_poly_div_by_vanishing.argtypes = [ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, u64p, ctypes.c_int, vp,
                                                ctypes.c_int, vp]
        L.zkg_kzg_open.argtypes = [ctypes.c_int, ctypes.c_int, vp, u64p, vp, ctypes.c_int, u64p, u64p, ctypes.c_int]
        L.zkg_g1_batch_to_affine.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp]
        L.zkg_multi_msm_binary.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, u64p, ctypes.c_int]
        L.zkg_msm_binary.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, u64p, ctypes.c_int]
        L.zkg_ntt_set_tuning_for.argtypes = [ctypes.c_int] * 4
        L.zkg_ntt_tune.argtypes = [ctypes.c_int] * 5 + [ip, ip, dp, ip]
        _lib = L
    return _lib


def _check(status):
    if status != 0:
        raise ZkgError(lib().zkg_last_error().decode(errors="replace"))


def _p(a):
    return a.ctypes.data_as(u64p)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


# ---------------------------------------------------------------------------------------
# device buffers

class DeviceArray:
    """A device buffer holding uint64 words (freed into the library's caching allocator)."""

    def __init__(self, nwords: int):
        self.nwords = int(nwords)
        self.ptr = vp()
        _check(lib().zkg_malloc(ctypes.byref(self.ptr), max(8, 8 * self.nwords)))

    @classmethod
    def from_host(cls, a):
        a = _c(a)
        d = cls(a.size)
        _check(lib().zkg_memcpy_h2d(d.ptr, a.ctypes.data_as(vp), 8 * a.size))
        return d

    @classmethod
    def from_bytes(cls, b):
        """A byte array, zero-padded to whole words (for coefficient bytes of binary MSMs)."""
        b = np.ascontiguousarray(b, dtype=np.uint8).ravel()
        padded = np.zeros((b.size + 7) // 8 * 8, np.uint8)
        padded[:b.size] = b
        return cls.from_host(padded.view(np.uint64))

    def copy_from_host(self, a):
        """Overwrite the buffer's first a.size words (a plan bound to this buffer rereads them)."""
        a = _c(a)
        if a.size > self.nwords:
            raise ValueError("copy_from_host: more words than the buffer holds")
        _check(lib().zkg_memcpy_h2d(self.ptr, a.ctypes.data_as(vp), 8 * a.size))

    def to_host(self, shape=None):
        out = np.empty(self.nwords, dtype=np.uint64)
        _check(lib().zkg_memcpy_d2h(out.ctypes.data_as(vp), self.ptr, 8 * self.nwords))
        return out.reshape(shape) if shape is not None else out

    def free(self):
        if self.ptr:
            lib().zkg_free(self.ptr)
            self.ptr = vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    _check(lib().zkg_synchronize())


# ---------------------------------------------------------------------------------------
# reference entry points (host buffers)

def _ref_fn(curve, name):
    f = getattr(lib(), f"{PREFIX[curve]}_{name}")
    return f


def msm_ref_symbol(curve, coord, kind, out, scalars, points, expo_nlimbs=4):
    """Calls <c>_G1_<coord>_MSM_<kind>_coeff_<out>_out exactly as a C caller would."""
    scalars, points = _c(scalars), _c(points)
    n = points.shape[0] if points.ndim == 2 else points.size // (2 * FP_LIMBS[curve])
    size = 2 * FP_LIMBS[curve] if out == "affine" else 3 * FP_LIMBS[curve]
    res = np.zeros(size, np.uint64)
    f = _ref_fn(curve, f"G1_{coord}_MSM_{kind}_coeff_{out}_out")
    f.argtypes = [ctypes.c_int, u64p, u64p, u64p, ctypes.c_int]
    f(n, _p(scalars), _p(points), _p(res), expo_nlimbs)
    return res


def msm_slow_reference(curve, coord, scalars, points, expo_nlimbs=4, header_spelling=False):
    """<c>_G1_<coord>_MSM_std_coeff_<coord>_out_slow_reference (bn128_G1_proj.c:611-620), or the
    headers' spelling <coord>c_out (bn128_G1_proj.h:50); normalised proj/jac output words."""
    scalars, points = _c(scalars), _c(points)
    n = points.shape[0] if points.ndim == 2 else points.size // (2 * FP_LIMBS[curve])
    res = np.zeros(3 * FP_LIMBS[curve], np.uint64)
    out = coord + ("c" if header_spelling else "")
    f = _ref_fn(curve, f"G1_{coord}_MSM_std_coeff_{out}_out_slow_reference")
    f.argtypes = [ctypes.c_int, u64p, u64p, u64p, ctypes.c_int]
    f(n, _p(scalars), _p(points), _p(res), expo_nlimbs)
    return res


def msm_std(curve, scalars, points, expo_nlimbs=4):
    """Sum s_i P_i (standard-form scalars), affine output; infinity = all 0xFF words."""
    scalars, points = _c(scalars), _c(points)
    if scalars.size != points.size // (2 * FP_LIMBS[curve]) * expo_nlimbs:
        raise ValueError("msm: scalar and point counts differ")
    return msm_ref_symbol(curve, "proj", "std", "affine", scalars, points, expo_nlimbs)


def msm(curve, scalars, points):
    """Montgomery-form scalars (Class/Curve.hs:92 via G1/Proj.hs:254)."""
    scalars, points = _c(scalars), _c(points)
    if scalars.size // 4 != points.size // (2 * FP_LIMBS[curve]):
        raise ValueError("msm: scalar and point counts differ")
    return msm_ref_symbol(curve, "proj", "mont", "affine", scalars, points, 4)


def _ntt_call(name, curve, *args):
    f = _ref_fn(curve, name)
    f(*args)


def ntt(curve, m, gen, src):
    src = _c(src)
    if src.shape[0] != (1 << m):
        raise ValueError("forwardNTT: subgroup size differs from the array size")
    out = np.zeros_like(src)
    f = _ref_fn(curve, "poly_mont_ntt_forward")
    f.argtypes = [ctypes.c_int, u64p, u64p, u64p]
    f(m, _p(_c(gen)), _p(src), _p(out))
    return out


def intt(curve, m, gen, src):
    src = _c(src)
    if src.shape[0] != (1 << m):
        raise ValueError("inverseNTT: subgroup size differs from the array size")
    out = np.zeros_like(src)
    f = _ref_fn(curve, "poly_mont_ntt_inverse")
    f.argtypes = [ctypes.c_int, u64p, u64p, u64p]
    f(m, _p(_c(gen)), _p(src), _p(out))
    return out


def shifted_ntt(curve, m, gen, eta, src):
    src = _c(src)
    if src.shape[0] != (1 << m):
        raise ValueError("shiftedForwardNTT: subgroup size differs from the array size")
    out = np.zeros_like(src)
    f = _ref_fn(curve, "poly_mont_ntt_forward_shifted")
    f.argtypes = [u64p, ctypes.c_int, u64p, u64p, u64p]
    f(_p(_c(eta)), m, _p(_c(gen)), _p(src), _p(out))
    return out


def shifted_intt(curve, m, gen, eta, src):
    src = _c(src)
    if src.shape[0] != (1 << m):
        raise ValueError("shiftedInverseNTT: subgroup size differs from the array size")
    out = np.zeros_like(src)
    f = _ref_fn(curve, "poly_mont_ntt_inverse_shifted")
    f.argtypes = [u64p, ctypes.c_int, u64p, u64p, u64p]
    f(_p(_c(eta)), m, _p(_c(gen)), _p(src), _p(out))
    return out


def asymm_ntt(curve, m_src, gen_src, src, m_tgt, gen_tgt):
    src = _c(src)
    if src.shape[0] != (1 << m_src):
        raise ValueError("asymmForwardNTT: subgroup size differs from the array size")
    if m_tgt < m_src:
        raise ValueError("asymmForwardNTT: target subgroup size should be at least the source subgroup src")
    out = np.zeros(((1 << m_tgt), 4), np.uint64)
    f = _ref_fn(curve, "poly_mont_ntt_forward_asymmetric")
    f.argtypes = [ctypes.c_int, ctypes.c_int, u64p, u64p, u64p, u64p]
    f(m_src, m_tgt, _p(_c(gen_src)), _p(_c(gen_tgt)), _p(src), _p(out))
    return out


def g1_fft(curve, m, gen, src, inverse=False, coord="proj"):
    """G1 FFT through the reference entry point <c>_G1_<coord>_fft_{forward,inverse}: src is
    (2^m, 3 L_p) projective (or Jacobian) Montgomery points; outputs normalised."""
    src = _c(src)
    if src.shape[0] != (1 << m):
        raise ValueError("g1 fft: subgroup size differs from the array size")
    out = np.zeros_like(src)
    f = _ref_fn(curve, f"G1_{coord}_fft_{'inverse' if inverse else 'forward'}")
    f.argtypes = [ctypes.c_int, u64p, u64p, u64p]
    f(m, _p(_c(gen)), _p(src), _p(out))
    return out


G1_FFT_INVERSE, G1_FFT_JAC = 1, 2


def g1_fft_device(curve, m, gen, d_src, d_tgt, inverse=False, jac=False):
    mode = (G1_FFT_INVERSE if inverse else 0) | (G1_FFT_JAC if jac else 0)
    _check(lib().zkg_g1_fft(curve, m, _p(_c(gen)), mode, d_src.ptr, d_tgt.ptr))


def poly_eval(curve, coeffs, loc):
    """<c>_poly_mont_eval_at (bn128_poly_mont.c:225-243): sum a_i loc^i, Montgomery Fr words."""
    coeffs = _c(coeffs)
    out = np.zeros(4, np.uint64)
    f = _ref_fn(curve, "poly_mont_eval_at")
    f.argtypes = [ctypes.c_int, u64p, u64p, u64p]
    f(coeffs.size // 4, _p(coeffs), _p(_c(loc)), _p(out))
    return out


def div_by_vanishing(curve, coeffs, n, eta, nquot, nrem):
    """<c>_poly_mont_div_by_vanishing (bn128_poly_mont.c:317-395): quotient and remainder of the
    polynomial by x^n - eta, in buffers of nquot and nrem coefficients."""
    coeffs = _c(coeffs)
    q = np.zeros((max(nquot, 1), 4), np.uint64)
    r = np.zeros((max(nrem, 1), 4), np.uint64)
    f = _ref_fn(curve, "poly_mont_div_by_vanishing")
    f.argtypes = [ctypes.c_int, u64p, ctypes.c_int, u64p, ctypes.c_int, u64p, ctypes.c_int, u64p]
    f(coeffs.size // 4, _p(coeffs), n, _p(_c(eta)), nquot, _p(q), nrem, _p(r) if nrem else None)
    return q[:nquot], r[:nrem]


def quot_by_vanishing(curve, coeffs, n, eta, nquot):
    """<c>_poly_mont_quot_by_vanishing (bn128_poly_mont.c:400-413): (remainder is zero, quotient)."""
    coeffs = _c(coeffs)
    q = np.zeros((max(nquot, 1), 4), np.uint64)
    f = _ref_fn(curve, "poly_mont_quot_by_vanishing")
    f.argtypes = [ctypes.c_int, u64p, ctypes.c_int, u64p, ctypes.c_int, u64p]
    f.restype = ctypes.c_uint8
    ok = f(coeffs.size // 4, _p(coeffs), n, _p(_c(eta)), nquot, _p(q))
    return bool(ok), q[:nquot]


def poly_eval_device(curve, n, d_coeffs, loc):
    out = np.zeros(4, np.uint64)
    _check(lib().zkg_poly_eval(curve, n, d_coeffs.ptr, _p(_c(loc)), _p(out)))
    return out


def poly_degree_device(n, d_coeffs):
    d = ctypes.c_int()
    _check(lib().zkg_poly_degree(n, d_coeffs.ptr, ctypes.byref(d)))
    return d.value


def div_by_vanishing_device(curve, n1, d_src, n, eta, nquot, d_quot, nrem=0, d_rem=None):
    _check(lib().zkg_poly_div_by_vanishing(curve, n1, d_src.ptr, n, _p(_c(eta)), nquot, d_quot.ptr, nrem,
                                           d_rem.ptr if d_rem is not None else None))


def kzg_open(curve, n, d_coeffs, z, srs, out_flags=0):
    """KZG opening (examples/KZG.hs:108-114) of the polynomial with n Montgomery coefficients at
    z: (p(z), commitment of (p - p(z)) / (x - z)); srs is a DeviceArray of affine points or
    PreparedPoints."""
    L = FP_LIMBS[curve]
    value = np.zeros(4, np.uint64)
    proof = np.zeros(3 * L if out_flags & (MSM_PROJ_OUT | MSM_JAC_OUT) else 2 * L, np.uint64)
    flags = MSM_POINTS_INTERNAL if isinstance(srs, PreparedPoints) else 0
    _check(lib().zkg_kzg_open(curve, n, d_coeffs.ptr, _p(_c(z)), srs.ptr, flags, _p(value), _p(proof), out_flags))
    return value, proof


def batch_to_affine(curve, pts, coord="proj"):
    """<c>_G1_<coord>_batch_to_affine (bn128_G1_proj.c:157-166): (N, 3 L_p) -> (N, 2 L_p),
    infinity -> all 0xFF words."""
    pts = _c(pts)
    L = FP_LIMBS[curve]
    n = pts.size // (3 * L)
    out = np.zeros((n, 2 * L), np.uint64)
    f = _ref_fn(curve, f"G1_{coord}_batch_to_affine")
    f.argtypes = [ctypes.c_int, u64p, u64p]
    f(n, _p(pts), _p(out))
    return out


def batch_from_affine(curve, aff, coord="proj"):
    """<c>_G1_<coord>_batch_from_affine (bn128_G1_proj.c:146-155): (N, 2 L_p) -> (N, 3 L_p)."""
    aff = _c(aff)
    L = FP_LIMBS[curve]
    n = aff.size // (2 * L)
    out = np.zeros((n, 3 * L), np.uint64)
    f = _ref_fn(curve, f"G1_{coord}_batch_from_affine")
    f.argtypes = [ctypes.c_int, u64p, u64p]
    f(n, _p(aff), _p(out))
    return out


def batch_to_affine_device(curve, n, d_src, d_tgt, jac=False):
    _check(lib().zkg_g1_batch_to_affine(curve, int(jac), n, d_src.ptr, d_tgt.ptr))


def _bytes_ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def msm_binary(curve, bits, points, coord="proj", out="affine"):
    """<c>_G1_<coord>_MSM_binary_coeff_<out>_out (bn128_G1_proj.c:684-698): the sum of the
    points whose byte in `bits` is non-zero.  out is "affine" or the coordinate system."""
    bits = np.ascontiguousarray(bits, dtype=np.uint8)
    points = _c(points)
    L = FP_LIMBS[curve]
    n = points.size // (2 * L)
    if bits.size != n:
        raise ValueError("msm_binary: coefficient and point counts differ")
    res = np.zeros(2 * L if out == "affine" else 3 * L, np.uint64)
    f = _ref_fn(curve, f"G1_{coord}_MSM_binary_coeff_{out}_out")
    f.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_uint8), u64p, u64p]
    f(n, _bytes_ptr(bits), _p(points), _p(res))
    return res


def multi_msm_binary(curve, rows, points, coord="proj", out="affine"):
    """<c>_G1_<coord>_multi_MSM_binary_coeff_<out>_out (bn128_G1_proj.c:716-782): one binary
    MSM per row of the (nrows, npoints) byte matrix, all over the same points."""
    rows = np.ascontiguousarray(rows, dtype=np.uint8)
    points = _c(points)
    L = FP_LIMBS[curve]
    n = points.size // (2 * L)
    if rows.ndim != 2 or rows.shape[1] != n:
        raise ValueError("multi_msm_binary: need an (nrows, npoints) coefficient matrix")
    nrows = rows.shape[0]
    w = 2 * L if out == "affine" else 3 * L
    res = np.zeros((nrows, w), np.uint64)
    u8p = ctypes.POINTER(ctypes.c_uint8)
    ptrs = (u8p * max(nrows, 1))(*[ctypes.cast(rows.ctypes.data + r * n, u8p) for r in range(nrows)])
    f = _ref_fn(curve, f"G1_{coord}_multi_MSM_binary_coeff_{out}_out")
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(u8p), u64p, u64p]
    f(nrows, n, ptrs, _p(points), _p(res))
    return res


def multi_msm_binary_device(curve, nrows, n, d_bits, d_points, flags=0):
    """nrows binary MSMs on device buffers (d_bits: nrows x n bytes, row-major)."""
    if isinstance(d_points, PreparedPoints):
        flags |= MSM_POINTS_INTERNAL
    L = FP_LIMBS[curve]
    w = 3 * L if flags & (MSM_PROJ_OUT | MSM_JAC_OUT) else 2 * L
    res = np.zeros((max(nrows, 1), w), np.uint64)
    _check(lib().zkg_multi_msm_binary(curve, nrows, n, d_bits.ptr, d_points.ptr, _p(res), flags))
    return res[:nrows]


def msm_binary_device(curve, n, d_bits, d_points, flags=0):
    """Binary MSM on device buffers: d_bits holds n bytes; flags as msm_device (output form,
    MSM_POINTS_INTERNAL)."""
    if isinstance(d_points, PreparedPoints):
        flags |= MSM_POINTS_INTERNAL
    L = FP_LIMBS[curve]
    res = np.zeros(3 * L if flags & (MSM_PROJ_OUT | MSM_JAC_OUT) else 2 * L, np.uint64)
    _check(lib().zkg_msm_binary(curve, n, d_bits.ptr, d_points.ptr, _p(res), flags))
    return res


def fft_generator(curve, m):
    out = np.zeros(4, np.uint64)
    _check(lib().zkg_fft_generator(curve, m, _p(out)))
    return out


# ---------------------------------------------------------------------------------------
# extension API (device buffers)

def msm_device(curve, n, d_scalars, d_points, expo_nlimbs=4, flags=0, window=0, seg_log=0, policy=SCHED_PLAIN,
               fix_limit=0, prefetch=0):
    """d_points: DeviceArray of affine points, or PreparedPoints (flag added automatically)."""
    if isinstance(d_points, PreparedPoints):
        flags |= MSM_POINTS_INTERNAL
    size = 3 * FP_LIMBS[curve] if flags & (MSM_PROJ_OUT | MSM_JAC_OUT) else 2 * FP_LIMBS[curve]
    res = np.zeros(size, np.uint64)
    sched = None
    if window or seg_log or policy or fix_limit or prefetch:
        sched = MsmSchedule(window, seg_log, policy, fix_limit, prefetch)
    _check(lib().zkg_msm(curve, n, d_scalars.ptr, expo_nlimbs, d_points.ptr, _p(res), flags,
                         ctypes.cast(ctypes.byref(sched), vp) if sched is not None else None))
    return res


# ---------------------------------------------------------------------------------------
# schedule selection: statistics, candidates, autotuner, cost model, batched MSMs

def scalar_stats(curve, n, d_scalars, expo_nlimbs=4, flags=0):
    st = ScalarStats()
    _check(lib().zkg_msm_scalar_stats(curve, n, d_scalars.ptr, expo_nlimbs, flags, ctypes.byref(st)))
    return st


def stats_from(n, n_zero, n_one, max_bits):
    st = ScalarStats()
    st.n, st.n_zero, st.n_one, st.max_bits = n, n_zero, n_one, max_bits
    return st


def msm_features(curve, st, window, seg_log):
    out = np.zeros(MSM_NFEAT, np.float32)
    _check(lib().zkg_msm_features(curve, ctypes.byref(st), window, seg_log,
                                  out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
    return out


def msm_candidates(curve, st, max_count=256):
    cs, ss = (ctypes.c_int * max_count)(), (ctypes.c_int * max_count)()
    k = lib().zkg_msm_candidates(curve, ctypes.byref(st), max_count, cs, ss)
    if k < 0:
        _check(1)
    return [(cs[i], ss[i]) for i in range(k)]


def msm_pick(curve, st, policy):
    out = MsmSchedule()
    _check(lib().zkg_msm_pick(curve, ctypes.byref(st), policy, ctypes.byref(out)))
    return out.window, out.seg_log


def msm_tune(curve, n, d_scalars, d_points, cands, reps=3, expo_nlimbs=4, flags=0):
    """Times every (window, seg_log) candidate; returns (ms list, all results identical)."""
    if isinstance(d_points, PreparedPoints):
        flags |= MSM_POINTS_INTERNAL
    k = len(cands)
    cs = (ctypes.c_int * k)(*[c for c, _ in cands])
    ss = (ctypes.c_int * k)(*[s for _, s in cands])
    ms = (ctypes.c_double * k)()
    same = ctypes.c_int()
    _check(lib().zkg_msm_tune(curve, n, d_scalars.ptr, expo_nlimbs, d_points.ptr, flags, k, cs, ss, reps, ms,
                              ctypes.byref(same)))
    return list(ms), bool(same.value)


def msm_batch(curves, ns, d_scalars, d_points, policy=SCHED_PLAIN, nstreams=4, expo_nlimbs=4, flags=0):
    """Many MSMs of uneven sizes (curves: one id for all jobs, or one per job) on concurrent
    streams; returns (affine results, wall ms)."""
    if any(isinstance(p, PreparedPoints) for p in d_points):
        if not all(isinstance(p, PreparedPoints) for p in d_points):
            raise ValueError("msm_batch: mix of prepared and plain point sets")
        flags |= MSM_POINTS_INTERNAL
    k = len(ns)
    if isinstance(curves, int):
        curves = [curves] * k
    wide = flags & (MSM_PROJ_OUT | MSM_JAC_OUT)
    outs = [np.zeros((3 if wide else 2) * FP_LIMBS[c], np.uint64) for c in curves]
    C = (ctypes.c_int * k)(*curves)
    N = (ctypes.c_int * k)(*ns)
    S = (vp * k)(*[d.ptr for d in d_scalars])
    P = (vp * k)(*[d.ptr for d in d_points])
    O = (u64p * k)(*[_p(o) for o in outs])
    sched = MsmSchedule(0, 0, policy)
    wall = ctypes.c_double()
    _check(lib().zkg_msm_batch(k, C, N, S, expo_nlimbs, P, O, flags, ctypes.byref(sched), nstreams,
                               ctypes.byref(wall)))
    return outs, wall.value


class MsmBatch:
    """The msm_batch job list as one captured plan (zkg_msm_batch_*): create fits the schedules
    and captures the pipelines into a hipGraph; run() replays it and returns (results, wall ms)."""

    def __init__(self, curves, ns, d_scalars, d_points, policy=SCHED_PLAIN, nstreams=4, expo_nlimbs=4, flags=0):
        if any(isinstance(p, PreparedPoints) for p in d_points):
            if not all(isinstance(p, PreparedPoints) for p in d_points):
                raise ValueError("MsmBatch: mix of prepared and plain point sets")
            flags |= MSM_POINTS_INTERNAL
        k = len(ns)
        if isinstance(curves, int):
            curves = [curves] * k
        wide = flags & (MSM_PROJ_OUT | MSM_JAC_OUT)
        self.outs = [np.zeros((3 if wide else 2) * FP_LIMBS[c], np.uint64) for c in curves]
        self._O = (u64p * k)(*[_p(o) for o in self.outs])
        C = (ctypes.c_int * k)(*curves)
        N = (ctypes.c_int * k)(*ns)
        S = (vp * k)(*[d.ptr for d in d_scalars])
        P = (vp * k)(*[d.ptr for d in d_points])
        sched = MsmSchedule(0, 0, policy)
        self.plan = vp()
        _check(lib().zkg_msm_batch_create(k, C, N, S, expo_nlimbs, P, flags, ctypes.byref(sched), nstreams,
                                          ctypes.byref(self.plan)))

    def run(self):
        wall = ctypes.c_double()
        _check(lib().zkg_msm_batch_run(self.plan, self._O, ctypes.byref(wall)))
        return [o.copy() for o in self.outs], wall.value

    def close(self):
        if self.plan:
            lib().zkg_msm_batch_destroy(self.plan)
            self.plan = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Mlp:
    """Dense fp32 network trained and run by the library's HIP kernels (zkg_mlp_*)."""

    def __init__(self, dims, seed=1):
        self.dims = [int(d) for d in dims]
        self.h = vp()
        D = (ctypes.c_int * len(self.dims))(*self.dims)
        _check(lib().zkg_mlp_create(len(self.dims) - 1, D, seed, ctypes.byref(self.h)))
        self.nparams = int(lib().zkg_mlp_nparams(self.h))

    def get_params(self):
        out = np.zeros(self.nparams, np.float32)
        _check(lib().zkg_mlp_get_params(self.h, _fp(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.nparams:
            raise ValueError("mlp: parameter count differs")
        _check(lib().zkg_mlp_set_params(self.h, _fp(p)))

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        y = np.zeros((X.shape[0], self.dims[-1]), np.float32)
        _check(lib().zkg_mlp_forward(self.h, X.shape[0], _fp(X), _fp(y)))
        return y

    def gradients(self, X, y):
        X, y = np.ascontiguousarray(X, np.float32), np.ascontiguousarray(y, np.float32)
        g = np.zeros(self.nparams, np.float32)
        loss = ctypes.c_float()
        _check(lib().zkg_mlp_gradients(self.h, X.shape[0], _fp(X), _fp(y), _fp(g), ctypes.byref(loss)))
        return g, loss.value

    def train(self, X, y, epochs, lr):
        X, y = np.ascontiguousarray(X, np.float32), np.ascontiguousarray(y, np.float32)
        loss = ctypes.c_float()
        _check(lib().zkg_mlp_train(self.h, X.shape[0], _fp(X), _fp(y), epochs, lr, ctypes.byref(loss)))
        return loss.value

    def install(self):
        """Makes this network the cost model of SCHED_LEARNED."""
        _check(lib().zkg_msm_set_cost_model(self.h))

    @staticmethod
    def uninstall():
        _check(lib().zkg_msm_set_cost_model(None))

    def close(self):
        if self.h:
            lib().zkg_mlp_destroy(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ntt_set_tuning(smax, tile_log, curve=None, m=None):
    """Default NTT schedule, or one (curve, m) entry when both are given."""
    if curve is None:
        _check(lib().zkg_ntt_set_tuning(smax, tile_log))
    else:
        _check(lib().zkg_ntt_set_tuning_for(curve, m, smax, tile_log))


def ntt_tune(curve, m, inverse=False, reps=3, install=True):
    """Autotunes the NTT schedule for one size; returns (smax, tile_log, ms, identical)."""
    s, t, ok = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    ms = ctypes.c_double()
    _check(lib().zkg_ntt_tune(curve, m, int(inverse), reps, int(install), ctypes.byref(s), ctypes.byref(t),
                              ctypes.byref(ms), ctypes.byref(ok)))
    return s.value, t.value, ms.value, bool(ok.value)


def ntt_device(curve, m, gen, d_src, d_tgt, inverse=False, eta=None):
    _check(lib().zkg_ntt(curve, m, _p(_c(gen)), _p(_c(eta)) if eta is not None else None,
                         NTT_INVERSE if inverse else NTT_FORWARD, d_src.ptr, d_tgt.ptr))


def ntt_device_noalloc(curve, m, gen, d_src, d_tgt, inverse=False, eta=None):
    """zkg_ntt_noalloc: the same transform with no device allocation (m <= 18, order-2^m gen)."""
    gen = _c(np.asarray(gen, np.uint64))
    e = _p(_c(np.asarray(eta, np.uint64))) if eta is not None else None
    _check(lib().zkg_ntt_noalloc(curve, m, _p(gen), e, 1 if inverse else 0, d_src.ptr, d_tgt.ptr))


class _NttJob(ctypes.Structure):
    _fields_ = [("m", ctypes.c_int), ("inverse", ctypes.c_int), ("eta", u64p), ("d_src", vp), ("d_tgt", vp)]


class NttBatch:
    """Batched uneven NTTs: one hipGraph replay per run() (zkg_ntt_batch_*)."""

    def __init__(self, curve, jobs):
        """jobs: list of (m, inverse, eta_or_None, d_src DeviceArray, d_tgt DeviceArray)."""
        self._etas = [(_c(e) if e is not None else None) for (_, _, e, _, _) in jobs]
        arr = (_NttJob * len(jobs))()
        for i, (m, inv, _, s, t) in enumerate(jobs):
            arr[i].m, arr[i].inverse = m, int(inv)
            arr[i].eta = _p(self._etas[i]) if self._etas[i] is not None else None
            arr[i].d_src, arr[i].d_tgt = s.ptr, t.ptr
        self.plan = vp()
        _check(lib().zkg_ntt_batch_create(curve, len(jobs), arr, ctypes.byref(self.plan)))

    def run(self):
        ms = ctypes.c_double()
        _check(lib().zkg_ntt_batch_run(self.plan, ctypes.byref(ms)))
        return ms.value

    def close(self):
        if self.plan:
            lib().zkg_ntt_batch_destroy(self.plan)
            self.plan = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def flat_file_count(path, elem_words):
    n = ctypes.c_size_t()
    _check(lib().zkg_flat_file_count(str(path).encode(), 8 * elem_words, ctypes.byref(n)))
    return n.value


def flat_read(path, elem_words):
    """readFlatArray (Flat.hs:211-226) into device memory: a DeviceArray of count * elem_words."""
    n = flat_file_count(path, elem_words)
    d = DeviceArray(n * elem_words)
    got = ctypes.c_size_t()
    _check(lib().zkg_flat_read(str(path).encode(), 8 * elem_words, d.ptr, n, ctypes.byref(got)))
    return d


def flat_write(path, d, elem_words, count):
    """writeFlatArray (Flat.hs:229-233) of count elements from device memory."""
    _check(lib().zkg_flat_write(str(path).encode(), d.ptr, 8 * elem_words, count))


def srs_load(curve, path, prepare=True):
    """An SRS file of affine points: PreparedPoints (prepare=True) or a DeviceArray, and n."""
    ptr, n = vp(), ctypes.c_int()
    _check(lib().zkg_srs_load(curve, str(path).encode(), int(prepare), ctypes.byref(ptr), ctypes.byref(n)))
    if prepare:
        pp = PreparedPoints(curve, None, n.value)
        pp.ptr = ptr
        return pp, n.value
    d = DeviceArray.__new__(DeviceArray)
    d.nwords, d.ptr = n.value * 2 * FP_LIMBS[curve], ptr
    return d, n.value


def srs_generate(curve, n, tau):
    """[tau^i G]_{i<n}, affine Montgomery, on the device (tau: standard-form integer)."""
    t = np.array([(tau >> (64 * k)) & 0xFFFFFFFFFFFFFFFF for k in range(4)], dtype=np.uint64)
    out = DeviceArray(n * 2 * FP_LIMBS[curve])
    _check(lib().zkg_srs_generate(curve, n, _p(t), out.ptr))
    return out


class PreparedPoints:
    """Points converted once into the device-internal form (zkg_msm_prepare_points)."""

    def __init__(self, curve, d_points, n):
        self.n = n
        self.ptr = vp()
        if d_points is not None:
            _check(lib().zkg_msm_prepare_points(curve, n, d_points.ptr, ctypes.byref(self.ptr)))

    def free(self):
        if self.ptr:
            lib().zkg_free(self.ptr)
            self.ptr = vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """KZG-commit serving loop (zkg_serve_*): iNTT then MSM against a resident SRS."""

    def __init__(self, curve, srs: PreparedPoints, hbm_budget, nworkers):
        self.curve = curve
        self.h = vp()
        self._keep = []
        _check(lib().zkg_serve_create(curve, srs.ptr, srs.n, int(hbm_budget), nworkers, ctypes.byref(self.h)))

    def submit(self, logn, values):
        values = _c(values)
        out = np.zeros(2 * FP_LIMBS[self.curve], np.uint64)
        jid = ctypes.c_longlong()
        self._keep.append((values, out))
        _check(lib().zkg_serve_submit(self.h, logn, _p(values), _p(out), ctypes.byref(jid)))
        return jid.value, out

    def drain(self):
        _check(lib().zkg_serve_drain(self.h))

    def set_exclusive_log(self, logn):
        """No job starts while one of at least 2^logn points runs (0: admission by footprint only)."""
        _check(lib().zkg_serve_set_exclusive_log(self.h, logn))

    def times(self, jid):
        q, r, l = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(lib().zkg_serve_job_times(self.h, jid, ctypes.byref(q), ctypes.byref(r), ctypes.byref(l)))
        return q.value, r.value, l.value

    def stats(self):
        peak, conc = ctypes.c_size_t(), ctypes.c_int()
        _check(lib().zkg_serve_stats(self.h, ctypes.byref(peak), ctypes.byref(conc)))
        return peak.value, conc.value

    def run_trace(self, logns, arrivals, values):
        n = len(logns)
        vals = [_c(v) for v in values]
        outs = [np.zeros(2 * FP_LIMBS[self.curve], np.uint64) for _ in range(n)]
        L = (ctypes.c_int * n)(*logns)
        A = (ctypes.c_double * n)(*arrivals)
        V = (u64p * n)(*[_p(v) for v in vals])
        O = (u64p * n)(*[_p(o) for o in outs])
        lat = (ctypes.c_double * n)()
        wall = ctypes.c_double()
        _check(lib().zkg_serve_run_trace(self.h, n, L, A, V, O, lat, ctypes.byref(wall)))
        return outs, list(lat), wall.value

    def close(self):
        if self.h:
            lib().zkg_serve_destroy(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def field_op(curve, field, op, d_a, d_b, d_out, n):
    _check(lib().zkg_field_op(curve, field, op, n, d_a.ptr, d_b.ptr if d_b is not None else None, d_out.ptr))


def fill_random(curve, field, n, seed):
    words = 6 if (field == 0 and curve == BLS12_381) else 4
    d = DeviceArray(n * words)
    _check(lib().zkg_fill_random(curve, field, n, seed, d.ptr))
    return d


def scalar_mul_gen(curve, d_scalars, n):
    out = DeviceArray(n * 2 * FP_LIMBS[curve])
    _check(lib().zkg_g1_scalar_mul_gen(curve, n, d_scalars.ptr, out.ptr))
    return out


def microbench(which):
    r = ctypes.c_double()
    _check(lib().zkg_microbench(which, ctypes.byref(r)))
    return r.value


def last_timing():
    tot = ctypes.c_double()
    ph = (ctypes.c_double * 8)()
    n = ctypes.c_int()
    _check(lib().zkg_last_timing(ctypes.byref(tot), ph, 8, ctypes.byref(n)))
    return tot.value, list(ph)[:n.value]


def set_phase_timing(on):
    """Record HIP events between the phases of later MSMs (off by default: each event record
    idles the stream a few microseconds)."""
    _check(lib().zkg_msm_phase_timing(1 if on else 0))


def last_msm_stats():
    out = (ctypes.c_longlong * 4)()
    _check(lib().zkg_msm_last_stats(out))
    return dict(window=out[0], windows=out[1], entries=out[2], heavy=out[3])


def device_info():
    cus = ctypes.c_int()
    hbm = ctypes.c_size_t()
    _check(lib().zkg_device_info(ctypes.byref(cus), ctypes.byref(hbm)))
    return cus.value, hbm.value


FOP = dict(add=0, sub=1, mul=2, sqr=3, neg=4, inv=5, half=6, to_std=7, from_std=8)
MB_IMUL, MB_COPY, MB_FMUL = 0, 1, 2
