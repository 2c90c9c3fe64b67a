/*
 * zko.h -- CPU oracle for the MSM and NTT hot paths of zikkurat-algebra.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's generated C
 * (lib/cbits/...) used as the checker for the GPU library.  Only tests/, the smoke check
 * and the CPU-baseline leg of bench.py may load it; the product library never links it.
 *
 * Curves are selected by an integer id: 0 = BN254 ("bn128"), 1 = BLS12-381.
 * Fields: 0 = Fp, 1 = Fr.  All buffers use the reference's flat little-endian 64-bit limb
 * layout in Montgomery form (R = 2^(64*L)).
 */
#ifndef ZKO_H
#define ZKO_H
#include <stdint.h>

#define ZKO_BN254 0
#define ZKO_BLS12_381 1
#define ZKO_FP 0
#define ZKO_FR 1

#ifdef __cplusplus
extern "C" {
#endif

int  zko_nlimbs(int curve, int field);

/* field arithmetic (Montgomery form, canonical in [0,p)) */
void zko_f_add(int curve, int field, const uint64_t *a, const uint64_t *b, uint64_t *c);
void zko_f_sub(int curve, int field, const uint64_t *a, const uint64_t *b, uint64_t *c);
void zko_f_neg(int curve, int field, const uint64_t *a, uint64_t *c);
void zko_f_mul(int curve, int field, const uint64_t *a, const uint64_t *b, uint64_t *c);
void zko_f_inv(int curve, int field, const uint64_t *a, uint64_t *c);
void zko_f_div_by_2(int curve, int field, const uint64_t *a, uint64_t *c);
void zko_f_to_std(int curve, int field, const uint64_t *a, uint64_t *c);
void zko_f_from_std(int curve, int field, const uint64_t *a, uint64_t *c);
void zko_f_pow_u64(int curve, int field, const uint64_t *a, uint64_t e, uint64_t *c);

/* G1 in homogeneous projective coordinates (X:Y:Z), infinity = (0:1:0) */
void zko_g1_proj_add(int curve, const uint64_t *p, const uint64_t *q, uint64_t *r);
void zko_g1_proj_dbl(int curve, const uint64_t *p, uint64_t *r);
void zko_g1_proj_madd(int curve, const uint64_t *p, const uint64_t *aff, uint64_t *r);
void zko_g1_proj_to_affine(int curve, const uint64_t *p, uint64_t *aff);
void zko_g1_proj_from_affine(int curve, const uint64_t *aff, uint64_t *p);
void zko_g1_proj_normalize(int curve, const uint64_t *p, uint64_t *r);
void zko_g1_proj_scl(int curve, const uint64_t *k, int k_nlimbs, const uint64_t *p, uint64_t *r);
void zko_g1_proj_to_jac(int curve, const uint64_t *p, uint64_t *jac);
void zko_g1_jac_to_affine(int curve, const uint64_t *jac, uint64_t *aff);

/* MSM: sum_i s_i * P_i.  Scalars: expo_nlimbs words each, standard (or Montgomery) form. */
int  zko_msm_window(int npoints);
void zko_msm_std_proj_variable(int curve, int npoints, const uint64_t *expos, const uint64_t *grps,
                               uint64_t *tgt, int expo_nlimbs, int window);
void zko_msm_std_proj(int curve, int npoints, const uint64_t *expos, const uint64_t *grps,
                      uint64_t *tgt, int expo_nlimbs);
void zko_msm_std_affine(int curve, int npoints, const uint64_t *expos, const uint64_t *grps,
                        uint64_t *tgt, int expo_nlimbs);
void zko_msm_mont_affine(int curve, int npoints, const uint64_t *expos, const uint64_t *grps,
                         uint64_t *tgt, int expo_nlimbs);

/* NTT over Fr (natural order in and out) */
void zko_ntt_forward(int curve, int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt);
void zko_ntt_inverse(int curve, int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt);
void zko_ntt_forward_shifted(int curve, const uint64_t *eta, int m, const uint64_t *gen,
                             const uint64_t *src, uint64_t *tgt);
void zko_ntt_inverse_shifted(int curve, const uint64_t *eta, int m, const uint64_t *gen,
                             const uint64_t *src, uint64_t *tgt);
void zko_ntt_forward_asymmetric(int curve, int m_src, int m_tgt, const uint64_t *gen_src,
                                const uint64_t *gen_tgt, const uint64_t *src, uint64_t *tgt);
/* FFT of G1 points (projective, 2^m of them) with Fr twiddles gen^j; outputs normalised
 * (bn128_G1_proj.c:791-902) */
void zko_g1_fft_forward(int curve, int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt);
void zko_g1_fft_inverse(int curve, int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt);
/* the same over Jacobian points (bn128_G1_jac.c:840-950); outputs (x, y, 1), infinity (0, 1, 0) */
void zko_g1_jac_fft_forward(int curve, int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt);
void zko_g1_jac_fft_inverse(int curve, int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt);
/* batch conversions, projective (jac = 0) or Jacobian (jac = 1) <-> affine */
void zko_g1_batch_to_affine(int curve, int jac, int N, const uint64_t *src, uint64_t *tgt);
void zko_g1_batch_from_affine(int curve, int jac, int N, const uint64_t *src, uint64_t *tgt);
/* binary-coefficient MSMs, normalised projective outputs (bn128_G1_proj.c:684-782) */
void zko_msm_binary_proj(int curve, int npoints, const uint8_t *expos, const uint64_t *grps, uint64_t *tgt);
void zko_multi_msm_binary_proj(int curve, int nrows, int npoints, const uint8_t *const *rows,
                               const uint64_t *grps, uint64_t *tgt);
/* polynomials over Fr (flat Montgomery coefficients; bn128_poly_mont.c:26-32, :225-243, :317-413) */
int  zko_poly_degree(int curve, int n1, const uint64_t *a);
void zko_poly_eval(int curve, int n1, const uint64_t *a, const uint64_t *x, uint64_t *tgt);
int  zko_poly_div_by_vanishing(int curve, int n1, const uint64_t *a, int n, const uint64_t *eta, int nquot,
                               uint64_t *quot, int nrem, uint64_t *rem);
int  zko_poly_quot_by_vanishing(int curve, int n1, const uint64_t *a, int n, const uint64_t *eta, int nquot,
                                uint64_t *quot);
/* generator of the size-2^m FFT subgroup, Montgomery form (Class/FFT.hs:60-68) */
void zko_fft_generator(int curve, int m, uint64_t *gen);

#ifdef __cplusplus
}
#endif
#endif
