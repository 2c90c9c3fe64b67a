// C ABI: the reference's generated entry points (synchronous, host buffers) and the zkg_*
// extension API (device buffers).  Every reference entry point stages its host buffers
// through the pinned ring, runs the device kernels on the library stream and copies the
// result back before returning, like the single-threaded C it replaces.
#include <string.h>

#include <memory>
#include <mutex>
#include <vector>

#include "../../include/zkgpu.h"
#include "hostfield.hpp"
#include "mlp.hpp"
#include "msm.hpp"
#include "ntt.hpp"
#include "peak.hpp"
#include "runtime.hpp"

using namespace zk;

namespace {

int with_ctx(Context **c) { return context(c); }

void check(const char *fn, int status) {
  if (status != ZK_OK) fatal(fn, last_error());
}

// host-in / host-out NTT used by every reference NTT symbol
int ntt_host(int curve, int m, const uint64_t *gen, const uint64_t *eta, int flags, const uint64_t *src,
             uint64_t *tgt) {
  if (m < 0 || m > 30) {
    set_error("ntt: log size out of range");
    return ZK_EINVAL;
  }
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  const size_t bytes = ((size_t)1 << m) * 32;
  DevBuf a, b;
  ZK_TRY(a.alloc(bytes));
  ZK_TRY(b.alloc(bytes));
  ZK_TRY(copy_h2d(a.p, src, bytes, ctx->stream));
  ZK_TRY(ntt_device(curve, m, gen, eta, flags, a.p, b.p, ctx->stream, nullptr));
  ZK_TRY(copy_d2h(tgt, b.p, bytes, ctx->stream));
  return ZK_OK;
}

const zkh::FieldDesc &fr(int curve) { return curve == 0 ? zkh::BN254_FR : zkh::BLS381_FR; }

// binary MSMs of nrows coefficient rows (row r at d_bits + r * npoints) over one device point
// set, written to tgt in the form flags selects
int multi_binary_device(Context *ctx, int curve, int nrows, int npoints, const uint8_t *d_bits,
                        const uint64_t *d_points, bool points_internal, uint64_t *tgt, int flags) {
  const int lp = curve == 0 ? 4 : 6;
  const int outw = (flags & (ZKG_MSM_PROJ_OUT | ZKG_MSM_JAC_OUT)) ? 3 * lp : 2 * lp;
  if (nrows <= 0) return ZK_OK;
  std::vector<MsmResult> res(nrows);
  ZK_TRY(binary_sum_device(curve, nrows, npoints, d_bits, d_points, points_internal, ctx->stream, res.data()));
  for (int r = 0; r < nrows; r++) ZK_TRY(msm_write_result(curve, res[r], tgt + (size_t)r * outw, flags));
  return ZK_OK;
}

// host-in / host-out binary MSMs (bn128_G1_proj.c:684-782): the points go to the device once,
// converted to the internal form, and the coefficient rows are packed into one byte matrix
int binary_host(int curve, int nrows, int npoints, const uint8_t *const *rows, const uint64_t *grps, uint64_t *tgt,
                int flags) {
  const int lp = curve == 0 ? 4 : 6;
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  if (nrows <= 0 || npoints <= 0) return multi_binary_device(ctx, curve, nrows, npoints, nullptr, nullptr, true, tgt, flags);
  DevBuf dp, dint, bits;
  const size_t pb = (size_t)npoints * 2 * lp * 8;
  ZK_TRY(dp.alloc(pb));
  ZK_TRY(copy_h2d(dp.p, grps, pb, ctx->stream));
  ZK_TRY(dint.alloc((size_t)npoints * msm_internal_point_bytes(curve)));
  ZK_TRY(msm_prepare_points(curve, npoints, dp.as<uint64_t>(), dint.p, ctx->stream, false));
  ZK_TRY(bits.alloc((size_t)nrows * npoints));
  for (int r = 0; r < nrows; r++)
    ZK_TRY(copy_h2d(bits.as<uint8_t>() + (size_t)r * npoints, rows[r], (size_t)npoints, ctx->stream));
  return multi_binary_device(ctx, curve, nrows, npoints, bits.as<uint8_t>(), dint.as<uint64_t>(), true, tgt, flags);
}

// host-in / host-out batch conversions (bn128_G1_proj.c:146-166, bn128_G1_jac.c:139-158)
int batch_conv_host(int curve, bool jac, int n, const uint64_t *src, uint64_t *tgt, bool to_affine) {
  if (n <= 0) return ZK_OK;
  const int lp = curve == 0 ? 4 : 6;
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  const size_t sb = (size_t)n * (to_affine ? 3 : 2) * lp * 8, tb = (size_t)n * (to_affine ? 2 : 3) * lp * 8;
  DevBuf a, b;
  ZK_TRY(a.alloc(sb));
  ZK_TRY(b.alloc(tb));
  ZK_TRY(copy_h2d(a.p, src, sb, ctx->stream));
  if (to_affine) ZK_TRY(g1_batch_to_affine(curve, jac, n, a.as<uint64_t>(), b.as<uint64_t>(), ctx->stream));
  else ZK_TRY(g1_batch_from_affine(curve, jac, n, a.as<uint64_t>(), b.as<uint64_t>(), ctx->stream));
  return copy_d2h(tgt, b.p, tb, ctx->stream);
}

// degree as bn128_poly_mont_degree computes it (bn128_poly_mont.c:26-32): the highest index
// whose Montgomery words are not all zero, -1 for the zero polynomial
int host_degree(int n1, const uint64_t *src) {
  for (int i = n1 - 1; i >= 0; i--) {
    const uint64_t *c = src + (size_t)i * 4;
    if (c[0] | c[1] | c[2] | c[3]) return i;
  }
  return -1;
}

// host-in / host-out evaluation (bn128_poly_mont.c:225-243)
int eval_host(int curve, int n1, const uint64_t *src, const uint64_t *loc, uint64_t *tgt) {
  if (n1 <= 0) {
    memset(tgt, 0, 32);
    return ZK_OK;
  }
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  DevBuf a;
  ZK_TRY(a.alloc((size_t)n1 * 32));
  ZK_TRY(copy_h2d(a.p, src, (size_t)n1 * 32, ctx->stream));
  return poly_eval_device(curve, n1, a.as<uint64_t>(), loc, tgt, ctx->stream);
}

// host-in / host-out division by x^n - eta (bn128_poly_mont.c:317-395), with the reference's
// argument checks (its asserts) as errors and its zero filling of the unused tails
int vanish_host(int curve, int n1, const uint64_t *src, int n, const uint64_t *eta, int nquot, uint64_t *quot,
                int nrem, uint64_t *rem) {
  const int D = host_degree(n1, src);
  if (!quot) {
    set_error("div_by_vanishing: the quotient buffer cannot be NULL");
    return ZK_EINVAL;
  }
  if (nquot < D - n + 1 || (rem && nrem < n) || n < 1) {
    set_error("div_by_vanishing: need nquot >= deg p - n + 1, nrem >= n, n >= 1");
    return ZK_EINVAL;
  }
  if (D < n) {  // the quotient is zero, the remainder is p
    memset(quot, 0, (size_t)(nquot > 0 ? nquot : 0) * 32);
    if (rem) {
      if (nrem < D + 1) {
        set_error("div_by_vanishing: remainder buffer too small");
        return ZK_EINVAL;
      }
      memset(rem, 0, (size_t)nrem * 32);
      memcpy(rem, src, (size_t)(D + 1) * 32);
    }
    return ZK_OK;
  }
  const int M = D - n + 1;
  for (int j = M; j < nquot; j++) memset(quot + (size_t)j * 4, 0, 32);
  if (rem)
    for (int j = n; j < nrem; j++) memset(rem + (size_t)j * 4, 0, 32);
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  DevBuf a, q, r;
  ZK_TRY(a.alloc((size_t)(D + 1) * 32));
  ZK_TRY(q.alloc((size_t)M * 32));
  if (rem) ZK_TRY(r.alloc((size_t)n * 32));
  ZK_TRY(copy_h2d(a.p, src, (size_t)(D + 1) * 32, ctx->stream));
  ZK_TRY(vanish_div_device(curve, D, a.as<uint64_t>(), n, eta, q.as<uint64_t>(), rem ? r.as<uint64_t>() : nullptr,
                           ctx->stream));
  ZK_TRY(copy_d2h(quot, q.p, (size_t)M * 32, ctx->stream));
  if (rem) ZK_TRY(copy_d2h(rem, r.p, (size_t)n * 32, ctx->stream));
  return ZK_OK;
}

// bn128_poly_mont.c:400-413: the quotient, and whether the remainder is zero
int quot_vanish_host(int curve, int n1, const uint64_t *src, int n, const uint64_t *eta, int nquot, uint64_t *quot,
                     uint8_t *ok) {
  if (n < 1) {
    set_error("quot_by_vanishing: n must be at least 1");
    return ZK_EINVAL;
  }
  std::vector<uint64_t> rem((size_t)n * 4);
  ZK_TRY(vanish_host(curve, n1, src, n, eta, nquot, quot, n, rem.data()));
  *ok = 1;
  for (uint64_t w : rem)
    if (w) {
      *ok = 0;
      break;
    }
  return ZK_OK;
}

// host-in / host-out G1 FFT (bn128_G1_proj.c:854-866, :890-902): 2^m projective points
int g1_fft_host(int curve, int m, const uint64_t *gen, bool inverse, bool jac, const uint64_t *src, uint64_t *tgt) {
  if (m < 0 || m > 26) {
    set_error("g1_fft: log size out of range");
    return ZK_EINVAL;
  }
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  const size_t bytes = ((size_t)1 << m) * 3 * 8 * (curve == 0 ? 4 : 6);
  DevBuf a, b;
  ZK_TRY(a.alloc(bytes));
  ZK_TRY(b.alloc(bytes));
  ZK_TRY(copy_h2d(a.p, src, bytes, ctx->stream));
  ZK_TRY(g1_fft_device(curve, m, gen, inverse, jac, a.as<uint64_t>(), b.as<uint64_t>(), ctx->stream));
  ZK_TRY(copy_d2h(tgt, b.p, bytes, ctx->stream));
  return ZK_OK;
}

// bn128_poly_mont.c:492-557: K = 2^(m_tgt - m_src) shifted NTTs with eta = gen_tgt^k,
// interleaved as tgt[i*K + k].  When gen_src = gen_tgt^K (subgroups of one domain) this is
// one size-2^m_tgt NTT of the zero-padded input, which is what the device runs.
int ntt_asym_host(int curve, int ms, int mt, const uint64_t *gs, const uint64_t *gt, const uint64_t *src,
                  uint64_t *tgt) {
  if (ms < 0 || mt < ms || mt > 30) {
    set_error("ntt_forward_asymmetric: need 0 <= m_src <= m_tgt <= 30");
    return ZK_EINVAL;
  }
  const zkh::FieldDesc &f = fr(curve);
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  const size_t Ns = (size_t)1 << ms, Nt = (size_t)1 << mt, K = Nt / Ns;
  zkh::Fe g = zkh::load(f, gt);
  zkh::Fe gK = g;
  for (int i = 0; i < mt - ms; i++) gK = zkh::sqr(f, gK);
  const bool same_domain = zkh::is_eq(f, gK, zkh::load(f, gs)) && ntt_generator_ok(curve, mt, gt);
  DevBuf a, b;
  if (same_domain) {
    ZK_TRY(a.alloc(Nt * 32));
    ZK_TRY(b.alloc(Nt * 32));
    ZK_HIP(hipMemsetAsync(a.p, 0, Nt * 32, ctx->stream));
    ZK_TRY(copy_h2d(a.p, src, Ns * 32, ctx->stream));
    ZK_TRY(ntt_device(curve, mt, gt, nullptr, NTT_FORWARD, a.p, b.p, ctx->stream, nullptr));
    ZK_TRY(copy_d2h(tgt, b.p, Nt * 32, ctx->stream));
    return ZK_OK;
  }
  // general case: one shifted NTT per coset, interleaved on the host side of the copy
  ZK_TRY(a.alloc(Ns * 32));
  ZK_TRY(b.alloc(Ns * 32));
  ZK_TRY(copy_h2d(a.p, src, Ns * 32, ctx->stream));
  uint64_t *small = new uint64_t[Ns * 4];
  zkh::Fe eta = zkh::one(f);
  int st = ZK_OK;
  for (size_t k = 0; k < K && st == ZK_OK; k++) {
    if (k) eta = zkh::mul(f, eta, g);
    uint64_t e[4];
    zkh::store(f, e, eta);
    st = ntt_device(curve, ms, gs, k ? e : nullptr, NTT_FORWARD, a.p, b.p, ctx->stream, nullptr);
    if (st == ZK_OK) st = copy_d2h(small, b.p, Ns * 32, ctx->stream);
    if (st == ZK_OK)
      for (size_t i = 0; i < Ns; i++) memcpy(tgt + (i * K + k) * 4, small + i * 4, 32);
  }
  delete[] small;
  return st;
}

int msm_host(int curve, int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int enl,
             int flags) {
  const int lp = curve == 0 ? 4 : 6;
  if (npoints <= 0) {
    MsmResult r;
    msm_infinity(curve, &r);
    return msm_write_result(curve, r, tgt, flags);
  }
  if (enl < 1 || enl > 16 || ((flags & ZKG_MSM_MONT) && enl != 4)) {
    set_error("msm: expo_nlimbs must be 1..16 (4 for Montgomery scalars)");
    return ZK_EINVAL;
  }
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  DevBuf ds, dp;
  const size_t sb = (size_t)npoints * enl * 8, pb = (size_t)npoints * 2 * lp * 8;
  ZK_TRY(ds.alloc(sb));
  ZK_TRY(dp.alloc(pb));
  ZK_TRY(copy_h2d(ds.p, expos, sb, ctx->stream));
  ZK_TRY(copy_h2d(dp.p, grps, pb, ctx->stream));
  MsmResult r;
  ZK_TRY(msm_device(curve, npoints, ds.as<uint64_t>(), enl, flags & ZKG_MSM_MONT, dp.as<uint64_t>(), false,
                    nullptr, ctx->stream, &r));
  return msm_write_result(curve, r, tgt, flags);
}

}  // namespace

extern "C" {

// ---------------- reference NTT symbols ----------------
#define NTT_SYMBOLS(PFX, CID)                                                                          \
  void PFX##_poly_mont_ntt_forward(int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt) {   \
    check(#PFX "_poly_mont_ntt_forward", ntt_host(CID, m, gen, nullptr, NTT_FORWARD, src, tgt));      \
  }                                                                                                    \
  void PFX##_poly_mont_ntt_inverse(int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt) {   \
    check(#PFX "_poly_mont_ntt_inverse", ntt_host(CID, m, gen, nullptr, NTT_INVERSE, src, tgt));      \
  }                                                                                                    \
  void PFX##_poly_mont_ntt_forward_shifted(const uint64_t *eta, int m, const uint64_t *gen,           \
                                           const uint64_t *src, uint64_t *tgt) {                       \
    check(#PFX "_poly_mont_ntt_forward_shifted", ntt_host(CID, m, gen, eta, NTT_FORWARD, src, tgt));  \
  }                                                                                                    \
  void PFX##_poly_mont_ntt_inverse_shifted(const uint64_t *eta, int m, const uint64_t *gen,           \
                                           const uint64_t *src, uint64_t *tgt) {                       \
    check(#PFX "_poly_mont_ntt_inverse_shifted", ntt_host(CID, m, gen, eta, NTT_INVERSE, src, tgt));  \
  }                                                                                                    \
  void PFX##_poly_mont_ntt_forward_asymmetric(int m_src, int m_tgt, const uint64_t *gen_src,          \
                                              const uint64_t *gen_tgt, const uint64_t *src,           \
                                              uint64_t *tgt) {                                         \
    check(#PFX "_poly_mont_ntt_forward_asymmetric",                                                    \
          ntt_asym_host(CID, m_src, m_tgt, gen_src, gen_tgt, src, tgt));                               \
  }

NTT_SYMBOLS(bn128, 0)
NTT_SYMBOLS(bls12_381, 1)

// ---------------- reference MSM symbols ----------------
#define MSM_SYMBOL(NAME, CID, FLAGS)                                                                   \
  void NAME(int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int expo_nlimbs) { \
    check(#NAME, msm_host(CID, npoints, expos, grps, tgt, expo_nlimbs, FLAGS));                        \
  }

MSM_SYMBOL(bn128_G1_proj_MSM_std_coeff_proj_out, 0, ZKG_MSM_PROJ_OUT)
MSM_SYMBOL(bn128_G1_proj_MSM_mont_coeff_proj_out, 0, ZKG_MSM_PROJ_OUT | ZKG_MSM_MONT)
MSM_SYMBOL(bn128_G1_proj_MSM_std_coeff_affine_out, 0, 0)
MSM_SYMBOL(bn128_G1_proj_MSM_mont_coeff_affine_out, 0, ZKG_MSM_MONT)
MSM_SYMBOL(bn128_G1_jac_MSM_std_coeff_jac_out, 0, ZKG_MSM_JAC_OUT)
MSM_SYMBOL(bn128_G1_jac_MSM_mont_coeff_jac_out, 0, ZKG_MSM_JAC_OUT | ZKG_MSM_MONT)
MSM_SYMBOL(bn128_G1_jac_MSM_std_coeff_affine_out, 0, 0)
MSM_SYMBOL(bn128_G1_jac_MSM_mont_coeff_affine_out, 0, ZKG_MSM_MONT)
MSM_SYMBOL(bls12_381_G1_proj_MSM_std_coeff_proj_out, 1, ZKG_MSM_PROJ_OUT)
MSM_SYMBOL(bls12_381_G1_proj_MSM_mont_coeff_proj_out, 1, ZKG_MSM_PROJ_OUT | ZKG_MSM_MONT)
MSM_SYMBOL(bls12_381_G1_proj_MSM_std_coeff_affine_out, 1, 0)
MSM_SYMBOL(bls12_381_G1_proj_MSM_mont_coeff_affine_out, 1, ZKG_MSM_MONT)
MSM_SYMBOL(bls12_381_G1_jac_MSM_std_coeff_jac_out, 1, ZKG_MSM_JAC_OUT)
MSM_SYMBOL(bls12_381_G1_jac_MSM_mont_coeff_jac_out, 1, ZKG_MSM_JAC_OUT | ZKG_MSM_MONT)
MSM_SYMBOL(bls12_381_G1_jac_MSM_std_coeff_affine_out, 1, 0)
MSM_SYMBOL(bls12_381_G1_jac_MSM_mont_coeff_affine_out, 1, ZKG_MSM_MONT)

// The reference's own slow MSM (one windowed scalar multiplication per point, summed; its test
// oracle): same contract as the proj/jac output forms above, so it runs the same MSM here.  The
// .c files define the *_proj_out_ / *_jac_out_ names; the headers declare *_projc_out_ /
// *_jacc_out_, which the reference never defines, so both spellings are exported.
MSM_SYMBOL(bn128_G1_proj_MSM_std_coeff_proj_out_slow_reference, 0, ZKG_MSM_PROJ_OUT)
MSM_SYMBOL(bn128_G1_proj_MSM_std_coeff_projc_out_slow_reference, 0, ZKG_MSM_PROJ_OUT)
MSM_SYMBOL(bn128_G1_jac_MSM_std_coeff_jac_out_slow_reference, 0, ZKG_MSM_JAC_OUT)
MSM_SYMBOL(bn128_G1_jac_MSM_std_coeff_jacc_out_slow_reference, 0, ZKG_MSM_JAC_OUT)
MSM_SYMBOL(bls12_381_G1_proj_MSM_std_coeff_proj_out_slow_reference, 1, ZKG_MSM_PROJ_OUT)
MSM_SYMBOL(bls12_381_G1_proj_MSM_std_coeff_projc_out_slow_reference, 1, ZKG_MSM_PROJ_OUT)
MSM_SYMBOL(bls12_381_G1_jac_MSM_std_coeff_jac_out_slow_reference, 1, ZKG_MSM_JAC_OUT)
MSM_SYMBOL(bls12_381_G1_jac_MSM_std_coeff_jacc_out_slow_reference, 1, ZKG_MSM_JAC_OUT)

// ---------------- reference G1 FFT symbols ----------------
#define G1_FFT_SYMBOLS(PFX, CID, CO, JAC)                                                              \
  void PFX##_G1_##CO##_fft_forward(int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt) {   \
    check(#PFX "_G1_" #CO "_fft_forward", g1_fft_host(CID, m, gen, false, JAC, src, tgt));           \
  }                                                                                                    \
  void PFX##_G1_##CO##_fft_inverse(int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt) {   \
    check(#PFX "_G1_" #CO "_fft_inverse", g1_fft_host(CID, m, gen, true, JAC, src, tgt));            \
  }

G1_FFT_SYMBOLS(bn128, 0, proj, false)
G1_FFT_SYMBOLS(bls12_381, 1, proj, false)
G1_FFT_SYMBOLS(bn128, 0, jac, true)
G1_FFT_SYMBOLS(bls12_381, 1, jac, true)

// ---------------- reference binary MSM and batch conversion symbols ----------------
#define G1_EXTRA_SYMBOLS(PFX, CID, CO, OUTF, JAC)                                                       \
  void PFX##_G1_##CO##_MSM_binary_coeff_##CO##_out(int npoints, const uint8_t *expos, const uint64_t *grps, \
                                                   uint64_t *tgt) {                                       \
    check(#PFX "_G1_" #CO "_MSM_binary_coeff_" #CO "_out", binary_host(CID, 1, npoints, &expos, grps, tgt, OUTF)); \
  }                                                                                                      \
  void PFX##_G1_##CO##_MSM_binary_coeff_affine_out(int npoints, const uint8_t *expos, const uint64_t *grps, \
                                                   uint64_t *tgt) {                                       \
    check(#PFX "_G1_" #CO "_MSM_binary_coeff_affine_out", binary_host(CID, 1, npoints, &expos, grps, tgt, 0)); \
  }                                                                                                      \
  void PFX##_G1_##CO##_multi_MSM_binary_coeff_##CO##_out(int nrows, int npoints, const uint8_t **expos,  \
                                                         const uint64_t *grps, uint64_t *tgt) {          \
    check(#PFX "_G1_" #CO "_multi_MSM_binary_coeff_" #CO "_out",                                         \
          binary_host(CID, nrows, npoints, expos, grps, tgt, OUTF));                                      \
  }                                                                                                      \
  void PFX##_G1_##CO##_multi_MSM_binary_coeff_affine_out(int nrows, int npoints, const uint8_t **expos,  \
                                                         const uint64_t *grps, uint64_t *tgt) {          \
    check(#PFX "_G1_" #CO "_multi_MSM_binary_coeff_affine_out",                                           \
          binary_host(CID, nrows, npoints, expos, grps, tgt, 0));                                         \
  }                                                                                                      \
  void PFX##_G1_##CO##_batch_to_affine(int N, const uint64_t *src, uint64_t *tgt) {                      \
    check(#PFX "_G1_" #CO "_batch_to_affine", batch_conv_host(CID, JAC, N, src, tgt, true));             \
  }                                                                                                      \
  void PFX##_G1_##CO##_batch_from_affine(int N, const uint64_t *src, uint64_t *tgt) {                    \
    check(#PFX "_G1_" #CO "_batch_from_affine", batch_conv_host(CID, JAC, N, src, tgt, false));          \
  }

G1_EXTRA_SYMBOLS(bn128, 0, proj, ZKG_MSM_PROJ_OUT, false)
G1_EXTRA_SYMBOLS(bls12_381, 1, proj, ZKG_MSM_PROJ_OUT, false)
G1_EXTRA_SYMBOLS(bn128, 0, jac, ZKG_MSM_JAC_OUT, true)
G1_EXTRA_SYMBOLS(bls12_381, 1, jac, ZKG_MSM_JAC_OUT, true)

// ---------------- reference polynomial symbols of a KZG opening ----------------
#define POLY_SYMBOLS(PFX, CID)                                                                         \
  void PFX##_poly_mont_eval_at(int n1, const uint64_t *src1, const uint64_t *loc, uint64_t *tgt) {     \
    check(#PFX "_poly_mont_eval_at", eval_host(CID, n1, src1, loc, tgt));                             \
  }                                                                                                    \
  void PFX##_poly_mont_div_by_vanishing(int n1, const uint64_t *src1, int expo_n, const uint64_t *eta, \
                                        int nquot, uint64_t *quot, int nrem, uint64_t *rem) {          \
    check(#PFX "_poly_mont_div_by_vanishing",                                                          \
          vanish_host(CID, n1, src1, expo_n, eta, nquot, quot, nrem, rem));                            \
  }                                                                                                    \
  uint8_t PFX##_poly_mont_quot_by_vanishing(int n1, const uint64_t *src1, int expo_n, const uint64_t *eta, \
                                            int nquot, uint64_t *quot) {                               \
    uint8_t ok = 0;                                                                                    \
    check(#PFX "_poly_mont_quot_by_vanishing", quot_vanish_host(CID, n1, src1, expo_n, eta, nquot, quot, &ok)); \
    return ok;                                                                                         \
  }

POLY_SYMBOLS(bn128, 0)
POLY_SYMBOLS(bls12_381, 1)

// ---------------- extension API ----------------

const char *zkg_last_error(void) { return last_error(); }

int zkg_g1_batch_to_affine(int curve, int jac, int n, const uint64_t *d_src, uint64_t *d_tgt) {
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  return g1_batch_to_affine(curve, jac != 0, n, d_src, d_tgt, ctx->stream);
}

// ---- flat arrays on disk (lib/src/ZK/Algebra/Class/Flat.hs:211-233): raw little-endian limbs,
// no header, the file size giving the length ----

int zkg_flat_file_count(const char *path, size_t elem_bytes, size_t *count) {
  if (!path || !elem_bytes || !count) {
    set_error("flat array: bad argument");
    return ZK_EINVAL;
  }
  FILE *f = fopen(path, "rb");
  if (!f) {
    set_error("flat array: cannot open the file");
    return ZK_EINVAL;
  }
  int st = ZK_OK;
  if (fseeko(f, 0, SEEK_END) != 0) st = ZK_EINVAL;
  const off_t size = st == ZK_OK ? ftello(f) : -1;
  fclose(f);
  if (size < 0) {
    set_error("flat array: cannot size the file");
    return ZK_EINVAL;
  }
  if ((size_t)size % elem_bytes) {
    set_error("readFlatArray: the input file has invalid size");
    return ZK_EINVAL;
  }
  *count = (size_t)size / elem_bytes;
  return ZK_OK;
}

int zkg_flat_read(const char *path, size_t elem_bytes, void *d_dst, size_t capacity, size_t *count) {
  size_t n;
  ZK_TRY(zkg_flat_file_count(path, elem_bytes, &n));
  if (n > capacity) {
    set_error("flat array read: the device buffer is too small for the file");
    return ZK_EINVAL;
  }
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  FILE *f = fopen(path, "rb");
  if (!f) {
    set_error("flat array: cannot open the file");
    return ZK_EINVAL;
  }
  const int st = file_to_device(f, d_dst, n * elem_bytes, ctx->stream);
  fclose(f);
  if (st == ZK_OK && count) *count = n;
  return st;
}

int zkg_flat_write(const char *path, const void *d_src, size_t elem_bytes, size_t count) {
  if (!path || !elem_bytes) {
    set_error("flat array: bad argument");
    return ZK_EINVAL;
  }
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  FILE *f = fopen(path, "wb");
  if (!f) {
    set_error("flat array: cannot create the file");
    return ZK_EINVAL;
  }
  int st = device_to_file(f, d_src, count * elem_bytes, ctx->stream);
  if (fclose(f) != 0 && st == ZK_OK) {
    set_error("flat array write: close failed");
    st = ZK_EINVAL;
  }
  return st;
}

// an SRS file of affine G1 points (2 L_p limbs each) into device memory, optionally converted to
// the prepared MSM form; *d_out is released with zkg_free
int zkg_srs_load(int curve, const char *path, int prepare, void **d_out, int *n_out) {
  if ((curve != 0 && curve != 1) || !d_out || !n_out) {
    set_error("srs_load: bad argument");
    return ZK_EINVAL;
  }
  const size_t elem = (size_t)(curve == 0 ? 4 : 6) * 2 * 8;
  size_t n;
  ZK_TRY(zkg_flat_file_count(path, elem, &n));
  if (n > 0x7fffffff) {
    set_error("srs_load: too many points");
    return ZK_EINVAL;
  }
  void *aff = nullptr;
  ZK_TRY(dev_alloc(&aff, n ? n * elem : 256));
  int st = zkg_flat_read(path, elem, aff, n, nullptr);
  if (st == ZK_OK && prepare) {
    void *prep = nullptr;
    st = zkg_msm_prepare_points(curve, (int)n, (const uint64_t *)aff, &prep);
    dev_free(aff);
    aff = prep;
  }
  if (st != ZK_OK) {
    if (aff) dev_free(aff);
    return st;
  }
  *d_out = aff;
  *n_out = (int)n;
  return ZK_OK;
}

int zkg_poly_eval(int curve, int n, const uint64_t *d_coeffs, const uint64_t *loc, uint64_t *out_host) {
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  return poly_eval_device(curve, n, d_coeffs, loc, out_host, ctx->stream);
}

int zkg_poly_degree(int n, const uint64_t *d_coeffs, int *deg) {
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  return poly_degree_device(n, d_coeffs, deg, ctx->stream);
}

int zkg_poly_div_by_vanishing(int curve, int n1, const uint64_t *d_src, int expo_n, const uint64_t *eta,
                              int nquot, uint64_t *d_quot, int nrem, uint64_t *d_rem) {
  if (expo_n < 1 || !d_quot || (d_rem && nrem < expo_n)) {
    set_error("div_by_vanishing: need n >= 1, a quotient buffer and nrem >= n");
    return ZK_EINVAL;
  }
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  int D;
  ZK_TRY(poly_degree_device(n1, d_src, &D, ctx->stream));
  if (nquot < D - expo_n + 1) {
    set_error("div_by_vanishing: need nquot >= deg p - n + 1");
    return ZK_EINVAL;
  }
  const int M = D >= expo_n ? D - expo_n + 1 : 0;
  if (nquot > M) ZK_HIP(hipMemsetAsync(d_quot + (size_t)M * 4, 0, (size_t)(nquot - M) * 32, ctx->stream));
  if (d_rem) {
    const int filled = D >= expo_n ? expo_n : D + 1;
    if (nrem > filled) ZK_HIP(hipMemsetAsync(d_rem + (size_t)filled * 4, 0, (size_t)(nrem - filled) * 32, ctx->stream));
    if (D < expo_n && D >= 0)
      ZK_HIP(hipMemcpyAsync(d_rem, d_src, (size_t)(D + 1) * 32, hipMemcpyDeviceToDevice, ctx->stream));
  }
  if (D >= expo_n) ZK_TRY(vanish_div_device(curve, D, d_src, expo_n, eta, d_quot, d_rem, ctx->stream));
  ZK_HIP(hipStreamSynchronize(ctx->stream));
  return ZK_OK;
}

// KZG opening at z of the polynomial with n Montgomery coefficients (examples/KZG.hs:108-114):
// value p(z), and the commitment of (p - p(z)) / (x - z) to the first n - 1 SRS points
int zkg_kzg_open(int curve, int n, const uint64_t *d_coeffs, const uint64_t *z, const uint64_t *d_srs,
                 int srs_flags, uint64_t *value_host, uint64_t *proof_host, int out_flags) {
  if (curve != 0 && curve != 1) {
    set_error("kzg_open: bad curve id");
    return ZK_EINVAL;
  }
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  ZK_TRY(poly_eval_device(curve, n, d_coeffs, z, value_host, ctx->stream));
  MsmResult r;
  if (n < 2) {
    msm_infinity(curve, &r);
  } else {
    DevBuf q;
    ZK_TRY(q.alloc((size_t)(n - 1) * 32));
    ZK_TRY(vanish_div_device(curve, n - 1, d_coeffs, 1, z, q.as<uint64_t>(), nullptr, ctx->stream));
    ZK_TRY(msm_device(curve, n - 1, q.as<uint64_t>(), 4, 1, d_srs, (srs_flags & ZKG_MSM_POINTS_INTERNAL) != 0,
                      nullptr, ctx->stream, &r));
  }
  return msm_write_result(curve, r, proof_host, out_flags);
}

int zkg_multi_msm_binary(int curve, int nrows, int npoints, const uint8_t *d_bits, const uint64_t *d_grps,
                         uint64_t *tgt_host, int flags) {
  if (curve != 0 && curve != 1) {
    set_error("multi_msm_binary: bad curve id");
    return ZK_EINVAL;
  }
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  return multi_binary_device(ctx, curve, nrows, npoints, d_bits, d_grps, (flags & ZKG_MSM_POINTS_INTERNAL) != 0,
                             tgt_host, flags);
}

int zkg_msm_binary(int curve, int npoints, const uint8_t *d_bits, const uint64_t *d_grps, uint64_t *tgt_host,
                   int flags) {
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  return multi_binary_device(ctx, curve, 1, npoints, d_bits, d_grps, (flags & ZKG_MSM_POINTS_INTERNAL) != 0,
                             tgt_host, flags);
}

int zkg_g1_fft(int curve, int m, const uint64_t *gen, int mode, const uint64_t *d_src, uint64_t *d_tgt) {
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  return g1_fft_device(curve, m, gen, (mode & ZKG_G1_FFT_INVERSE) != 0, (mode & ZKG_G1_FFT_JAC) != 0, d_src, d_tgt,
                       ctx->stream);
}

int zkg_device_info(int *num_cus, size_t *hbm_bytes) {
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  if (num_cus) *num_cus = ctx->num_cus;
  if (hbm_bytes) {
    size_t fr = 0, tot = 0;
    ZK_HIP(hipMemGetInfo(&fr, &tot));
    *hbm_bytes = tot;
  }
  return ZK_OK;
}

int zkg_malloc(void **dptr, size_t bytes) {
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  return dev_alloc(dptr, bytes);
}
int zkg_free(void *dptr) {
  dev_free(dptr);
  return ZK_OK;
}
int zkg_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  return copy_h2d(dst, src, bytes, ctx->stream);
}
int zkg_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  return copy_d2h(dst, src, bytes, ctx->stream);
}
int zkg_synchronize(void) {
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  ZK_HIP(hipStreamSynchronize(ctx->stream));
  return ZK_OK;
}

int zkg_msm(int curve, int npoints, const uint64_t *d_expos, int expo_nlimbs, const uint64_t *d_grps,
            uint64_t *tgt_host, int flags, const zkg_msm_schedule *sched) {
  if (curve != 0 && curve != 1) {
    set_error("msm: bad curve id");
    return ZK_EINVAL;
  }
  if (npoints <= 0) {
    MsmResult r;
    msm_infinity(curve, &r);
    return msm_write_result(curve, r, tgt_host, flags);
  }
  if (expo_nlimbs < 1 || expo_nlimbs > 16 || ((flags & ZKG_MSM_MONT) && expo_nlimbs != 4)) {
    set_error("msm: expo_nlimbs must be 1..16 (4 for Montgomery scalars)");
    return ZK_EINVAL;
  }
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  MsmResult r;
  ZK_TRY(msm_device(curve, npoints, d_expos, expo_nlimbs, flags & ZKG_MSM_MONT, d_grps,
                    (flags & ZKG_MSM_POINTS_INTERNAL) != 0, sched, ctx->stream, &r));
  return msm_write_result(curve, r, tgt_host, flags);
}

int zkg_msm_prepare_points(int curve, int npoints, const uint64_t *d_grps, void **d_prepared) {
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  ZK_TRY(dev_alloc(d_prepared, (size_t)(npoints > 0 ? npoints : 1) * msm_internal_point_bytes(curve)));
  return msm_prepare_points(curve, npoints, d_grps, *d_prepared, ctx->stream);
}

int zkg_ntt(int curve, int m, const uint64_t *gen, const uint64_t *eta, int flags, const uint64_t *d_src,
            uint64_t *d_tgt) {
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  return ntt_device(curve, m, gen, eta, flags & 1, d_src, d_tgt, ctx->stream, nullptr);
}

int zkg_ntt_noalloc(int curve, int m, const uint64_t *gen, const uint64_t *eta, int flags, const uint64_t *d_src,
                    uint64_t *d_tgt) {
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  return ntt_device_noalloc(curve, m, gen, eta, flags & 1, d_src, d_tgt, ctx->stream);
}

int zkg_ntt_batch_create(int curve, int njobs, const zkg_ntt_job *jobs, void **plan) {
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::vector<NttBatchJob> js(njobs > 0 ? njobs : 0);
  for (int j = 0; j < njobs; j++) js[j] = NttBatchJob{jobs[j].m, jobs[j].inverse, jobs[j].eta, jobs[j].d_src, jobs[j].d_tgt};
  NttBatchPlan *pl = nullptr;
  ZK_TRY(ntt_batch_create(curve, njobs, js.data(), &pl));
  *plan = pl;
  return ZK_OK;
}
int zkg_ntt_batch_run(void *plan, double *ms) { return ntt_batch_run((NttBatchPlan *)plan, ms); }
int zkg_ntt_batch_destroy(void *plan) {
  ntt_batch_destroy((NttBatchPlan *)plan);
  return ZK_OK;
}

int zkg_fft_generator(int curve, int m, uint64_t *gen_out) {
  // 2-adic generators of Fr (lib/src/ZK/Algebra/Curves/BN128/Fr/Mont.hs:148-151, BLS12_381/Fr/Mont.hs:148-151)
  static const uint64_t bn_gen[4] = {0x636e735580d13d9cULL, 0xa22bf3742445ffd6ULL, 0x56452ac01eb203d8ULL,
                                     0x1860ef942963f9e7ULL};
  static const uint64_t bls_gen[4] = {0xb9b58d8c5f0e466aULL, 0x5b1b4c801819d7ecULL, 0x0af53ae352a31e64ULL,
                                      0x5bf3adda19e9b27bULL};
  const int M = curve == 0 ? 28 : 32;
  if ((curve != 0 && curve != 1) || m < 0 || m > M) {
    set_error("fft_generator: bad curve or size");
    return ZK_EINVAL;
  }
  const zkh::FieldDesc &f = fr(curve);
  zkh::Fe g = zkh::load(f, curve == 0 ? bn_gen : bls_gen);
  for (int i = 0; i < M - m; i++) g = zkh::sqr(f, g);
  zkh::store(f, gen_out, g);
  return ZK_OK;
}

int zkg_ntt_set_tuning(int smax, int tile_log) {
  if (smax < 1 || smax > 11 || tile_log < 4 || tile_log > 12) {
    set_error("ntt tuning out of range");
    return ZK_EINVAL;
  }
  ntt_set_tuning(NttTuning{smax, tile_log});
  return ZK_OK;
}

int zkg_ntt_set_tuning_for(int curve, int m, int smax, int tile_log) {
  if ((curve != 0 && curve != 1) || m < 0 || m > 30 || smax < 1 || smax > 11 || tile_log < 4 || tile_log > 12) {
    set_error("ntt tuning out of range");
    return ZK_EINVAL;
  }
  ntt_set_tuning_for(curve, m, NttTuning{smax, tile_log});
  return ZK_OK;
}

int zkg_ntt_tune(int curve, int m, int inverse, int reps, int install, int *smax, int *tile_log, double *ms,
                 int *identical) {
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  NttTuning best{0, 0};
  ZK_TRY(ntt_tune(curve, m, inverse, reps, &best, ms, identical, ctx->stream));
  if (install) ntt_set_tuning_for(curve, m, best);
  if (smax) *smax = best.smax;
  if (tile_log) *tile_log = best.tile_log;
  return ZK_OK;
}

int zkg_field_op(int curve, int field, int op, size_t n, const uint64_t *d_a, const uint64_t *d_b, uint64_t *d_out) {
  return field_op_device(curve, field, op, n, d_a, d_b, d_out);
}

int zkg_fill_random(int curve, int field, size_t n, uint64_t seed, uint64_t *d_out) {
  return fill_random_device(curve, field, n, seed, d_out);
}

int zkg_g1_scalar_mul_gen(int curve, int n, const uint64_t *d_scalars, uint64_t *d_out_affine) {
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  return g1_scalar_mul_gen(curve, n, d_scalars, d_out_affine, ctx->stream);
}

int zkg_microbench(int which, double *result) { return microbench(which, result); }

int zkg_srs_generate(int curve, int n, const uint64_t *tau_std, uint64_t *d_out_affine) {
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  return srs_generate(curve, n, tau_std, d_out_affine, ctx->stream);
}

int zkg_serve_create(int curve, const void *d_srs_prepared, int srs_n, size_t hbm_budget, int nworkers, void **server) {
  Server *s = nullptr;
  ZK_TRY(serve_create(curve, d_srs_prepared, srs_n, hbm_budget, nworkers, &s));
  *server = s;
  return ZK_OK;
}
int zkg_serve_submit(void *server, int logn, const uint64_t *values, uint64_t *out_affine, long long *job_id) {
  return serve_submit((Server *)server, logn, values, out_affine, job_id);
}
int zkg_serve_drain(void *server) { return serve_drain((Server *)server); }
int zkg_serve_set_exclusive_log(void *server, int logn) { return serve_set_exclusive_log((Server *)server, logn); }
int zkg_serve_job_times(void *server, long long id, double *queue_ms, double *run_ms, double *latency_ms) {
  return serve_job_times((Server *)server, id, queue_ms, run_ms, latency_ms);
}
int zkg_serve_stats(void *server, size_t *peak_bytes, int *max_concurrent) {
  return serve_stats((Server *)server, peak_bytes, max_concurrent);
}
int zkg_serve_run_trace(void *server, int njobs, const int *logn, const double *arrival_s,
                        const uint64_t *const *values, uint64_t *const *outs, double *latency_ms, double *wall_s) {
  return serve_run_trace((Server *)server, njobs, logn, arrival_s, values, outs, latency_ms, wall_s);
}
int zkg_serve_destroy(void *server) {
  serve_destroy((Server *)server);
  return ZK_OK;
}

int zkg_last_timing(double *ms_total, double *ms_phases, int max_phases, int *nphases) {
  return msm_last_timing(ms_total, ms_phases, max_phases, nphases);
}

int zkg_msm_phase_timing(int on) { return msm_set_phase_timing(on); }

int zkg_msm_last_stats(long long *out4) { return msm_last_stats(out4); }

// ---------------- schedule selection ----------------
static int check_msm_args(int curve, int npoints, int enl, int flags) {
  if (curve != 0 && curve != 1) {
    set_error("msm: bad curve id");
    return ZK_EINVAL;
  }
  if (npoints < 0 || enl < 1 || enl > 16 || ((flags & ZKG_MSM_MONT) && enl != 4)) {
    set_error("msm: bad size, or expo_nlimbs not 1..16 (4 for Montgomery scalars)");
    return ZK_EINVAL;
  }
  return ZK_OK;
}

int zkg_msm_scalar_stats(int curve, int npoints, const uint64_t *d_expos, int expo_nlimbs, int flags,
                         zkg_scalar_stats *out) {
  ZK_TRY(check_msm_args(curve, npoints, expo_nlimbs, flags));
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  return msm_scalar_stats(curve, npoints, d_expos, expo_nlimbs, flags & ZKG_MSM_MONT, ctx->stream, out);
}

int zkg_msm_features(int curve, const zkg_scalar_stats *st, int window, int seg_log, float *out) {
  if ((curve != 0 && curve != 1) || !st || window < 1 || window > 20 || seg_log < 1 || seg_log > 12) {
    set_error("msm_features: bad curve, window or seg_log");
    return ZK_EINVAL;
  }
  msm_features(curve, *st, window, seg_log, true, out);
  return ZK_OK;
}

int zkg_msm_candidates(int curve, const zkg_scalar_stats *st, int max, int *windows, int *seg_logs) {
  if ((curve != 0 && curve != 1) || !st || max < 0) {
    set_error("msm_candidates: bad arguments");
    return -1;
  }
  return msm_candidates(curve, *st, true, max, windows, seg_logs);
}

int zkg_msm_pick(int curve, const zkg_scalar_stats *st, int policy, zkg_msm_schedule *out) {
  if ((curve != 0 && curve != 1) || !st || !out) {
    set_error("msm_pick: bad arguments");
    return ZK_EINVAL;
  }
  zkg_msm_schedule u;
  memset(&u, 0, sizeof u);
  u.policy = policy;
  MsmSchedule s;
  ZK_TRY(msm_resolve(curve, (int)st->n, 4, &u, policy == ZKG_SCHED_PLAIN ? nullptr : st, &s, true));
  memset(out, 0, sizeof *out);
  out->window = s.c;
  out->seg_log = s.seg_log;
  out->policy = policy;
  return ZK_OK;
}

int zkg_msm_tune(int curve, int npoints, const uint64_t *d_expos, int expo_nlimbs, const uint64_t *d_grps, int flags,
                 int ncand, const int *windows, const int *seg_logs, int reps, double *ms_out, int *identical) {
  ZK_TRY(check_msm_args(curve, npoints, expo_nlimbs, flags));
  for (int k = 0; k < ncand; k++)
    if (windows[k] < 1 || windows[k] > 20 || seg_logs[k] < 1 || seg_logs[k] > 12) {
      set_error("msm_tune: candidate out of range");
      return ZK_EINVAL;
    }
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  return msm_tune(curve, npoints, d_expos, expo_nlimbs, flags & ZKG_MSM_MONT, d_grps,
                  (flags & ZKG_MSM_POINTS_INTERNAL) != 0, ncand, windows, seg_logs, reps, ms_out, identical,
                  ctx->stream);
}

int zkg_msm_batch(int njobs, const int *curves, const int *npoints, const uint64_t *const *d_expos, int expo_nlimbs,
                  const uint64_t *const *d_grps, uint64_t *const *outs, int flags, const zkg_msm_schedule *sched,
                  int nstreams, double *wall_ms) {
  for (int k = 0; k < njobs; k++) ZK_TRY(check_msm_args(curves[k], npoints[k], expo_nlimbs, flags));
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  return msm_batch(njobs, curves, npoints, d_expos, expo_nlimbs, flags & ZKG_MSM_MONT, d_grps,
                   (flags & ZKG_MSM_POINTS_INTERNAL) != 0, outs, flags & (ZKG_MSM_PROJ_OUT | ZKG_MSM_JAC_OUT), sched,
                   nstreams, wall_ms);
}

struct MsmBatchPlan {
  MsmGraph *g = nullptr;
  int out_flags = 0;
};
int zkg_msm_batch_create(int njobs, const int *curves, const int *npoints, const uint64_t *const *d_expos,
                         int expo_nlimbs, const uint64_t *const *d_grps, int flags, const zkg_msm_schedule *sched,
                         int nstreams, void **plan) {
  *plan = nullptr;
  for (int k = 0; k < njobs; k++) ZK_TRY(check_msm_args(curves[k], npoints[k], expo_nlimbs, flags));
  Context *ctx;
  ZK_TRY(with_ctx(&ctx));
  std::unique_ptr<MsmBatchPlan> pl(new MsmBatchPlan);
  pl->out_flags = flags & (ZKG_MSM_PROJ_OUT | ZKG_MSM_JAC_OUT);
  ZK_TRY(msm_graph_create(njobs, curves, npoints, d_expos, expo_nlimbs, flags & ZKG_MSM_MONT, d_grps,
                          (flags & ZKG_MSM_POINTS_INTERNAL) != 0, sched, nstreams, &pl->g));
  *plan = pl.release();
  return ZK_OK;
}
int zkg_msm_batch_run(void *plan, uint64_t *const *outs, double *wall_ms) {
  MsmBatchPlan *pl = (MsmBatchPlan *)plan;
  if (!pl || !pl->g) {
    set_error("msm_batch_run: no plan");
    return ZK_EINVAL;
  }
  return msm_graph_run(pl->g, outs, pl->out_flags, wall_ms);
}
int zkg_msm_batch_destroy(void *plan) {
  MsmBatchPlan *pl = (MsmBatchPlan *)plan;
  if (pl) msm_graph_destroy(pl->g);
  delete pl;
  return ZK_OK;
}

// ---------------- cost-model network ----------------
int zkg_mlp_create(int nlayers, const int *dims, unsigned seed, void **mlp) {
  Mlp *m = nullptr;
  ZK_TRY(mlp_create(nlayers, dims, seed, &m));
  *mlp = m;
  return ZK_OK;
}
int zkg_mlp_destroy(void *mlp) {
  cost_model_forget((Mlp *)mlp);
  mlp_destroy((Mlp *)mlp);
  return ZK_OK;
}
long long zkg_mlp_nparams(void *mlp) { return (long long)mlp_nparams((Mlp *)mlp); }
int zkg_mlp_get_params(void *mlp, float *host) { return mlp_get_params((Mlp *)mlp, host); }
int zkg_mlp_set_params(void *mlp, const float *host) { return mlp_set_params((Mlp *)mlp, host); }
int zkg_mlp_forward(void *mlp, int rows, const float *x_host, float *y_host) {
  return mlp_forward((Mlp *)mlp, rows, x_host, y_host);
}
int zkg_mlp_gradients(void *mlp, int rows, const float *x_host, const float *y_host, float *grad_host, float *loss) {
  return mlp_gradients((Mlp *)mlp, rows, x_host, y_host, grad_host, loss);
}
int zkg_mlp_train(void *mlp, int rows, const float *x_host, const float *y_host, int epochs, float lr,
                  float *final_loss) {
  return mlp_train((Mlp *)mlp, rows, x_host, y_host, epochs, lr, final_loss);
}
int zkg_msm_set_cost_model(void *mlp) { return cost_model_set((Mlp *)mlp); }

}  // extern "C"
