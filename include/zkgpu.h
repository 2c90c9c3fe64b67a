/*
 * zkgpu.h -- C ABI of the MI355X (gfx950) MSM and NTT library.
 *
 * Part 1 re-exports zikkurat-algebra's generated C entry points for the two hot paths with
 * identical names, signatures, limb layout (flat little-endian u64 limbs, Montgomery form
 * R = 2^(64 L)), infinity encodings and synchronous host-in/host-out semantics, so that a C
 * caller (or the Haskell FFI of lib/src/ZK/Algebra/Curves/...) links this library as a
 * drop-in.  Each declaration cites the reference declaration it replaces.  Differences that a
 * caller can observe are listed beside the entry point.
 *
 * Part 2 is the extension API (zkg_*): device-resident buffers, schedules, batches, the
 * autotuner/cost model and the serving loop.  These have no reference counterpart.
 *
 * Errors: the reference entry points return void and abort() with a message on failure, like
 * the reference's assert() (lib/cbits/curves/g1/proj/bn128_G1_proj.c:517).  zkg_* functions
 * return 0 on success and a nonzero status otherwise; zkg_last_error() describes it.
 * There is no CPU fallback: without a usable GPU every entry point fails.
 */
#ifndef ZKGPU_H
#define ZKGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* =========================================================================================
 * Part 1: reference entry points (curve prefix bn128_ = BN254, bls12_381_ = BLS12-381)
 *
 * MSM: tgt = sum_i expos[i] * grps[i].  expos: npoints scalars of expo_nlimbs u64 limbs
 * (std = standard form, processed as 64*expo_nlimbs-bit integers exactly like the reference;
 * mont = Montgomery-form Fr, converted to standard form first, which needs expo_nlimbs = 4).
 * grps: npoints affine points (x, y Montgomery Fp), affine infinity encoded as (0,0).
 * affine_out: tgt = 2 Fp elements, infinity = all 0xFF bytes.  proj_out / jac_out: tgt =
 * 3 Fp elements; the point equals the reference's but is returned normalised (Z = 1, or the
 * infinity (0:1:0) / [1:1:0]), while the reference returns an unnormalised representative.
 * npoints <= 0 yields infinity (the reference evaluates log2(0) there).
 * ========================================================================================= */

/* lib/cbits/curves/g1/proj/bn128_G1_proj.h:43 (impl bn128_G1_proj.c:597-605) */
void bn128_G1_proj_MSM_std_coeff_proj_out(int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int expo_nlimbs);
/* bn128_G1_proj.h:44 (impl bn128_G1_proj.c:631-645) */
void bn128_G1_proj_MSM_mont_coeff_proj_out(int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int expo_nlimbs);
/* bn128_G1_proj.h:46 (impl bn128_G1_proj.c:656-660) */
void bn128_G1_proj_MSM_std_coeff_affine_out(int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int expo_nlimbs);
/* bn128_G1_proj.h:47 (impl bn128_G1_proj.c:669-673) */
void bn128_G1_proj_MSM_mont_coeff_affine_out(int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int expo_nlimbs);
/* lib/cbits/curves/g1/jac/bn128_G1_jac.h:43-47 (impl bn128_G1_jac.c:555-722) */
void bn128_G1_jac_MSM_std_coeff_jac_out(int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int expo_nlimbs);
void bn128_G1_jac_MSM_mont_coeff_jac_out(int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int expo_nlimbs);
void bn128_G1_jac_MSM_std_coeff_affine_out(int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int expo_nlimbs);
void bn128_G1_jac_MSM_mont_coeff_affine_out(int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int expo_nlimbs);

/* lib/cbits/curves/g1/proj/bls12_381_G1_proj.h:43-47 (impl bls12_381_G1_proj.c:598-674) */
void bls12_381_G1_proj_MSM_std_coeff_proj_out(int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int expo_nlimbs);
void bls12_381_G1_proj_MSM_mont_coeff_proj_out(int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int expo_nlimbs);
void bls12_381_G1_proj_MSM_std_coeff_affine_out(int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int expo_nlimbs);
void bls12_381_G1_proj_MSM_mont_coeff_affine_out(int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int expo_nlimbs);
/* lib/cbits/curves/g1/jac/bls12_381_G1_jac.h:43-47 */
void bls12_381_G1_jac_MSM_std_coeff_jac_out(int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int expo_nlimbs);
void bls12_381_G1_jac_MSM_mont_coeff_jac_out(int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int expo_nlimbs);
void bls12_381_G1_jac_MSM_std_coeff_affine_out(int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int expo_nlimbs);
void bls12_381_G1_jac_MSM_mont_coeff_affine_out(int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int expo_nlimbs);

/* The reference's slow MSM, sum_i scl(expos[i], grps[i]) one point at a time, its own test
 * oracle.  Same contract as *_MSM_std_coeff_proj_out / *_jac_out (normalised output); here it
 * runs the same bucket MSM.  The .c files define the *_proj_out_ / *_jac_out_ names, the
 * headers declare *_projc_out_ / *_jacc_out_ (never defined there); both are exported. */
/* impl bn128_G1_proj.c:611-620 (declared bn128_G1_proj.h:50 as ..._projc_out_slow_reference) */
void bn128_G1_proj_MSM_std_coeff_proj_out_slow_reference(int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int expo_nlimbs);
void bn128_G1_proj_MSM_std_coeff_projc_out_slow_reference(int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int expo_nlimbs);
/* impl bn128_G1_jac.c:660-669 (declared bn128_G1_jac.h:50 as ..._jacc_out_slow_reference) */
void bn128_G1_jac_MSM_std_coeff_jac_out_slow_reference(int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int expo_nlimbs);
void bn128_G1_jac_MSM_std_coeff_jacc_out_slow_reference(int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int expo_nlimbs);
/* impl bls12_381_G1_proj.c:612-621 (declared bls12_381_G1_proj.h:50) */
void bls12_381_G1_proj_MSM_std_coeff_proj_out_slow_reference(int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int expo_nlimbs);
void bls12_381_G1_proj_MSM_std_coeff_projc_out_slow_reference(int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int expo_nlimbs);
/* impl bls12_381_G1_jac.c:660-669 (declared bls12_381_G1_jac.h:50) */
void bls12_381_G1_jac_MSM_std_coeff_jac_out_slow_reference(int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int expo_nlimbs);
void bls12_381_G1_jac_MSM_std_coeff_jacc_out_slow_reference(int npoints, const uint64_t *expos, const uint64_t *grps, uint64_t *tgt, int expo_nlimbs);

/* NTT over Fr: src and tgt are 2^m Montgomery Fr elements (4 u64 each); gen is the
 * Montgomery-form generator of the size-2^m subgroup (the *forward* one for the inverse too);
 * eta is the Montgomery-form coset shift.  src may equal tgt. */
/* lib/cbits/curves/poly/mont/bn128_poly_mont.h:27 (impl bn128_poly_mont.c:452-473) */
void bn128_poly_mont_ntt_forward(int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt);
/* bn128_poly_mont.h:28 (impl bn128_poly_mont.c:601-622) */
void bn128_poly_mont_ntt_inverse(int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt);
/* bn128_poly_mont.h:30 (impl bn128_poly_mont.c:477-489) */
void bn128_poly_mont_ntt_forward_shifted(const uint64_t *eta, int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt);
/* bn128_poly_mont.h:31 (impl bn128_poly_mont.c:626-638) */
void bn128_poly_mont_ntt_inverse_shifted(const uint64_t *eta, int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt);
/* bn128_poly_mont.h:33 (impl bn128_poly_mont.c:492-557) */
void bn128_poly_mont_ntt_forward_asymmetric(int m_src, int m_tgt, const uint64_t *gen_src, const uint64_t *gen_tgt, const uint64_t *src, uint64_t *tgt);

/* lib/cbits/curves/poly/mont/bls12_381_poly_mont.h:27-33 */
void bls12_381_poly_mont_ntt_forward(int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt);
void bls12_381_poly_mont_ntt_inverse(int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt);
void bls12_381_poly_mont_ntt_forward_shifted(const uint64_t *eta, int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt);
void bls12_381_poly_mont_ntt_inverse_shifted(const uint64_t *eta, int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt);
void bls12_381_poly_mont_ntt_forward_asymmetric(int m_src, int m_tgt, const uint64_t *gen_src, const uint64_t *gen_tgt, const uint64_t *src, uint64_t *tgt);

/* FFT of 2^m G1 points in projective Montgomery form (X:Y:Z), twiddles gen^j (Montgomery Fr);
 * the inverse is the reference's DIF network with the halvings, i.e. N^-1 sum gen^-ij.
 * Outputs are normalised (X/Z, Y/Z, 1) or (0, 1, 0), as the reference normalises them. */
/* lib/cbits/curves/g1/proj/bn128_G1_proj.h:54 (impl bn128_G1_proj.c:791-866) */
void bn128_G1_proj_fft_forward(int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt);
/* bn128_G1_proj.h:55 (impl bn128_G1_proj.c:871-902) */
void bn128_G1_proj_fft_inverse(int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt);
/* lib/cbits/curves/g1/proj/bls12_381_G1_proj.h:54-55 */
void bls12_381_G1_proj_fft_forward(int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt);
void bls12_381_G1_proj_fft_inverse(int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt);
/* Jacobian inputs (X:Y:Z) = (X/Z^2, Y/Z^3); outputs normalised [x : y : 1], infinity (0, 1, 0)
 * as bn128_G1_jac_normalize writes it (bn128_G1_jac.c:62-85).
 * lib/cbits/curves/g1/jac/bn128_G1_jac.h:54-55 (impl bn128_G1_jac.c:840-950) */
void bn128_G1_jac_fft_forward(int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt);
void bn128_G1_jac_fft_inverse(int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt);
/* lib/cbits/curves/g1/jac/bls12_381_G1_jac.h:54-55 */
void bls12_381_G1_jac_fft_forward(int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt);
void bls12_381_G1_jac_fft_inverse(int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt);

/* Binary-coefficient MSMs: the sum of the points whose byte is non-zero (one row), or one such
 * sum per row of an nrows x npoints byte matrix given as row pointers.  proj_out results are
 * normalised like the other MSM entry points. */
/* lib/cbits/curves/g1/proj/bn128_G1_proj.h:45 (impl bn128_G1_proj.c:684-692) */
void bn128_G1_proj_MSM_binary_coeff_proj_out(int npoints, const uint8_t *expos, const uint64_t *grps, uint64_t *tgt);
/* bn128_G1_proj.h:48 (impl bn128_G1_proj.c:694-698) */
void bn128_G1_proj_MSM_binary_coeff_affine_out(int npoints, const uint8_t *expos, const uint64_t *grps, uint64_t *tgt);
/* bn128_G1_proj.h:52 (impl bn128_G1_proj.c:716-770) */
void bn128_G1_proj_multi_MSM_binary_coeff_proj_out(int nrows, int npoints, const uint8_t **expos, const uint64_t *grps, uint64_t *tgt);
/* bn128_G1_proj.h:53 (impl bn128_G1_proj.c:773-782) */
void bn128_G1_proj_multi_MSM_binary_coeff_affine_out(int nrows, int npoints, const uint8_t **expos, const uint64_t *grps, uint64_t *tgt);
/* lib/cbits/curves/g1/proj/bls12_381_G1_proj.h:45-53 */
void bls12_381_G1_proj_MSM_binary_coeff_proj_out(int npoints, const uint8_t *expos, const uint64_t *grps, uint64_t *tgt);
void bls12_381_G1_proj_MSM_binary_coeff_affine_out(int npoints, const uint8_t *expos, const uint64_t *grps, uint64_t *tgt);
void bls12_381_G1_proj_multi_MSM_binary_coeff_proj_out(int nrows, int npoints, const uint8_t **expos, const uint64_t *grps, uint64_t *tgt);
void bls12_381_G1_proj_multi_MSM_binary_coeff_affine_out(int nrows, int npoints, const uint8_t **expos, const uint64_t *grps, uint64_t *tgt);

/* Jacobian twins (lib/cbits/curves/g1/jac/bn128_G1_jac.h:45-53, impl bn128_G1_jac.c:726-830);
 * jac_out results are normalised [x : y : 1], infinity [1 : 1 : 0]. */
void bn128_G1_jac_MSM_binary_coeff_jac_out(int npoints, const uint8_t *expos, const uint64_t *grps, uint64_t *tgt);
void bn128_G1_jac_MSM_binary_coeff_affine_out(int npoints, const uint8_t *expos, const uint64_t *grps, uint64_t *tgt);
void bn128_G1_jac_multi_MSM_binary_coeff_jac_out(int nrows, int npoints, const uint8_t **expos, const uint64_t *grps, uint64_t *tgt);
void bn128_G1_jac_multi_MSM_binary_coeff_affine_out(int nrows, int npoints, const uint8_t **expos, const uint64_t *grps, uint64_t *tgt);
/* lib/cbits/curves/g1/jac/bls12_381_G1_jac.h:45-53 */
void bls12_381_G1_jac_MSM_binary_coeff_jac_out(int npoints, const uint8_t *expos, const uint64_t *grps, uint64_t *tgt);
void bls12_381_G1_jac_MSM_binary_coeff_affine_out(int npoints, const uint8_t *expos, const uint64_t *grps, uint64_t *tgt);
void bls12_381_G1_jac_multi_MSM_binary_coeff_jac_out(int nrows, int npoints, const uint8_t **expos, const uint64_t *grps, uint64_t *tgt);
void bls12_381_G1_jac_multi_MSM_binary_coeff_affine_out(int nrows, int npoints, const uint8_t **expos, const uint64_t *grps, uint64_t *tgt);

/* Batch conversions: projective (X:Y:Z) -> affine (X/Z, Y/Z), Z = 0 -> all 0xFF bytes; affine
 * -> (x, y, 1), the affine infinity (0,0) -> (0, 1, 0).  Jacobian: (X/Z^2, Y/Z^3), infinity
 * [1 : 1 : 0]. */
/* bn128_G1_proj.h:10 (impl bn128_G1_proj.c:157-166) */
void bn128_G1_proj_batch_to_affine(int N, const uint64_t *src, uint64_t *tgt);
/* bn128_G1_proj.h:9 (impl bn128_G1_proj.c:146-155) */
void bn128_G1_proj_batch_from_affine(int N, const uint64_t *src, uint64_t *tgt);
/* bls12_381_G1_proj.h:9-10 */
void bls12_381_G1_proj_batch_to_affine(int N, const uint64_t *src, uint64_t *tgt);
void bls12_381_G1_proj_batch_from_affine(int N, const uint64_t *src, uint64_t *tgt);
/* bn128_G1_jac.h:9-10 (impl bn128_G1_jac.c:139-158), bls12_381_G1_jac.h:9-10 */
void bn128_G1_jac_batch_to_affine(int N, const uint64_t *src, uint64_t *tgt);
void bn128_G1_jac_batch_from_affine(int N, const uint64_t *src, uint64_t *tgt);
void bls12_381_G1_jac_batch_to_affine(int N, const uint64_t *src, uint64_t *tgt);
void bls12_381_G1_jac_batch_from_affine(int N, const uint64_t *src, uint64_t *tgt);

/* Polynomial operations of a KZG opening (coefficients: flat Montgomery Fr, 4 limbs each).
 * eval_at: sum a_i loc^i.  div_by_vanishing: quotient and remainder by x^n - eta, the quotient
 * zero-filled beyond deg p - n + 1 and the remainder beyond n; the reference's asserts
 * (nquot >= deg p - n + 1, nrem >= n, quot != NULL) abort with a message.
 * quot_by_vanishing: the quotient, returning 1 when the remainder is zero. */
/* lib/cbits/curves/poly/mont/bn128_poly_mont.h:6 (impl bn128_poly_mont.c:225-243) */
void bn128_poly_mont_eval_at(int n1, const uint64_t *src1, const uint64_t *loc, uint64_t *tgt);
/* bn128_poly_mont.h:24 (impl bn128_poly_mont.c:317-395) */
void bn128_poly_mont_div_by_vanishing(int n1, const uint64_t *src1, int expo_n, const uint64_t *eta, int nquot, uint64_t *quot, int nrem, uint64_t *rem);
/* bn128_poly_mont.h:25 (impl bn128_poly_mont.c:400-413) */
uint8_t bn128_poly_mont_quot_by_vanishing(int n1, const uint64_t *src1, int expo_n, const uint64_t *eta, int nquot, uint64_t *quot);
/* lib/cbits/curves/poly/mont/bls12_381_poly_mont.h:6, :24-25 */
void bls12_381_poly_mont_eval_at(int n1, const uint64_t *src1, const uint64_t *loc, uint64_t *tgt);
void bls12_381_poly_mont_div_by_vanishing(int n1, const uint64_t *src1, int expo_n, const uint64_t *eta, int nquot, uint64_t *quot, int nrem, uint64_t *rem);
uint8_t bls12_381_poly_mont_quot_by_vanishing(int n1, const uint64_t *src1, int expo_n, const uint64_t *eta, int nquot, uint64_t *quot);

/* =========================================================================================
 * Part 2: extension API (no reference counterpart)
 * ========================================================================================= */

#define ZKG_BN254 0
#define ZKG_BLS12_381 1

/* status and device memory */
const char *zkg_last_error(void);
int zkg_device_info(int *num_cus, size_t *hbm_bytes);
int zkg_malloc(void **dptr, size_t bytes);
int zkg_free(void *dptr);
int zkg_memcpy_h2d(void *dst, const void *src, size_t bytes);
int zkg_memcpy_d2h(void *dst, const void *src, size_t bytes);
int zkg_synchronize(void);

/* MSM schedule: window width c (1..20), and the split of the bucket-accumulation work.
 * A zero window / seg_log means "choose" by the policy.  Under ZKG_SCHED_STATS and
 * ZKG_SCHED_LEARNED a device pass first measures the scalars (largest bit length, count of
 * zeros and ones); the number of windows then covers only the largest scalar's bits, also
 * when window and seg_log are given.  None of this changes a result bit. */
#define ZKG_SCHED_PLAIN 0    /* analytic Pippenger heuristic from n and the declared width */
#define ZKG_SCHED_STATS 1    /* statistics pass, analytic heuristic on the measured scalars */
#define ZKG_SCHED_LEARNED 2  /* statistics pass, loaded cost model (else as ZKG_SCHED_STATS) */
typedef struct {
  int window;        /* c: signed-digit window width */
  int seg_log;       /* log2 of sorted entries per accumulation thread */
  int policy;        /* ZKG_SCHED_* */
  int fix_limit;     /* buckets spanning more segments are summed by workgroup folds (0 = 16, or
                        with more than 1024 buckets twice the mean segments per bucket of dense
                        scalars, clamped to 16..64) */
  int prefetch;      /* accumulation loads the next point early: 0 = default (off), 1 = off, 2 = on */
  int reserved[3];
} zkg_msm_schedule;

/* Statistics of a scalar vector (standard form after conversion). */
typedef struct {
  long long n;       /* number of scalars */
  long long n_zero;  /* scalars equal to 0 */
  long long n_one;   /* scalars equal to 1 */
  int max_bits;      /* bit length of the largest scalar */
  int reserved;
} zkg_scalar_stats;

/* MSM on device buffers.  flags: bit0 = scalars in Montgomery form, bit1 = output
 * projective (X:Y:Z, normalised) instead of affine, bit2 = output Jacobian.  tgt_host
 * receives the result (host memory).  sched may be NULL. */
#define ZKG_MSM_MONT 1
#define ZKG_MSM_PROJ_OUT 2
#define ZKG_MSM_JAC_OUT 4
#define ZKG_MSM_POINTS_INTERNAL 8   /* d_grps was written by zkg_msm_prepare_points */
int zkg_msm(int curve, int npoints, const uint64_t *d_expos, int expo_nlimbs, const uint64_t *d_grps,
            uint64_t *tgt_host, int flags, const zkg_msm_schedule *sched);

/* Converts npoints affine points once into the device-internal reduced-radix form (for
 * bases reused across many MSMs, e.g. a KZG SRS); free with zkg_free.  BN254 sets also hold
 * phi(P_i) = (beta x_i, y_i) beside each P_i (128 bytes per point), so MSMs over the set or
 * any prefix of it take the GLV split; BLS12-381 sets take 128 bytes per point. */
int zkg_msm_prepare_points(int curve, int npoints, const uint64_t *d_grps, void **d_prepared);

/* NTT on device buffers (flags: 0 forward, 1 inverse); eta NULL = no coset shift.
 * gen/eta are host pointers. */
int zkg_ntt(int curve, int m, const uint64_t *gen, const uint64_t *eta, int flags, const uint64_t *d_src,
            uint64_t *d_tgt);
/* zkg_ntt without any device allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Scratch (pass vector, twiddles rebuilt on every call,
 * coset tables) lives in one fixed file-scope device array of 12.3 MiB per GPU, so m <= 18
 * (status 1 above), and the generator must have order 2^m (status 4 otherwise).  Every
 * call on a GPU shares that array: calls that overlap on one GPU race.  Synchronous. */
int zkg_ntt_noalloc(int curve, int m, const uint64_t *gen, const uint64_t *eta, int flags, const uint64_t *d_src,
                    uint64_t *d_tgt);
/* Batched NTTs of uneven sizes over the standard FFT domains (generator of order 2^m from
 * lib/src/ZK/Algebra/Class/FFT.hs:60-68).  create() builds twiddle/coset tables and descriptor
 * tables and captures one launch per pass round into a hipGraph; run() replays it
 * synchronously (optionally returning the device time in ms).  Buffers must stay valid for
 * the plan's lifetime; src != tgt. */
typedef struct {
  int m;                  /* log2 size, 1..30 */
  int inverse;            /* 0 forward, 1 inverse */
  const uint64_t *eta;    /* host Montgomery Fr coset shift (shifted variants), or NULL */
  const uint64_t *d_src;
  uint64_t *d_tgt;
} zkg_ntt_job;
int zkg_ntt_batch_create(int curve, int njobs, const zkg_ntt_job *jobs, void **plan);
int zkg_ntt_batch_run(void *plan, double *ms);
int zkg_ntt_batch_destroy(void *plan);

/* generator of the size-2^m FFT subgroup, Montgomery form (lib/src/ZK/Algebra/Class/FFT.hs:60-68) */
int zkg_fft_generator(int curve, int m, uint64_t *gen_out);
/* NTT fast-path schedule (passes of at most 2^smax points, LDS tiles of 2^tile_log elements):
 * the default for every size (clears per-size entries), or one (curve, m) entry */
int zkg_ntt_set_tuning(int smax, int tile_log);
int zkg_ntt_set_tuning_for(int curve, int m, int smax, int tile_log);
/* autotuner: times every (smax 5..11, tile_log 7..12) on a random 2^m vector (median of reps
 * HIP-event timings), checks that all give the same output, and installs the fastest for
 * (curve, m) when install != 0 */
int zkg_ntt_tune(int curve, int m, int inverse, int reps, int install, int *smax, int *tile_log, double *ms,
                 int *identical);

/* device elementwise field ops (field: 0 = Fp, 1 = Fr; op: see zikkurat-algebra_amd/csrc/peak.hpp) */
int zkg_field_op(int curve, int field, int op, size_t n, const uint64_t *d_a, const uint64_t *d_b, uint64_t *d_out);

/* n canonical pseudo-random field elements (field 0 = Fp, 1 = Fr) from a seed, on the device
 * (synthetic benchmark inputs) */
int zkg_fill_random(int curve, int field, size_t n, uint64_t seed, uint64_t *d_out);

/* batch fixed-base scalar multiplication k_i * G of the standard generator, affine Montgomery
 * output (input generation for tests and benchmarks) */
int zkg_g1_scalar_mul_gen(int curve, int n, const uint64_t *d_scalars, uint64_t *d_out_affine);

/* polynomial evaluation, degree (-1 for zero) and division by x^n - eta on device buffers;
 * the result of zkg_poly_eval goes to host memory (one Fr element) */
int zkg_poly_eval(int curve, int n, const uint64_t *d_coeffs, const uint64_t *loc, uint64_t *out_host);
int zkg_poly_degree(int n, const uint64_t *d_coeffs, int *deg);
int zkg_poly_div_by_vanishing(int curve, int n1, const uint64_t *d_src, int expo_n, const uint64_t *eta,
                              int nquot, uint64_t *d_quot, int nrem, uint64_t *d_rem);
/* Flat arrays on disk in the reference's format (lib/src/ZK/Algebra/Class/Flat.hs:211-233:
 * raw little-endian limbs, no header, length = file size / element size, a size that is not a
 * multiple of the element being an error), streamed through pinned staging buffers. */
int zkg_flat_file_count(const char *path, size_t elem_bytes, size_t *count);
int zkg_flat_read(const char *path, size_t elem_bytes, void *d_dst, size_t capacity, size_t *count);
int zkg_flat_write(const char *path, const void *d_src, size_t elem_bytes, size_t count);
/* an SRS file of affine Montgomery G1 points into a new device buffer (released with
 * zkg_free), converted to the prepared MSM form when prepare != 0 */
int zkg_srs_load(int curve, const char *path, int prepare, void **d_out, int *n_out);
/* KZG opening (examples/KZG.hs:108-114) of the polynomial with n Montgomery coefficients at z:
 * value_host = p(z), proof_host = MSM of (p - p(z)) / (x - z) over the first n - 1 SRS points
 * (d_srs: affine points, or the prepared form with srs_flags = ZKG_MSM_POINTS_INTERNAL);
 * out_flags selects the proof's output form as for zkg_msm */
int zkg_kzg_open(int curve, int n, const uint64_t *d_coeffs, const uint64_t *z, const uint64_t *d_srs,
                 int srs_flags, uint64_t *value_host, uint64_t *proof_host, int out_flags);
/* projective (jac = 0) or Jacobian (jac = 1) -> affine for n device points (Montgomery batch
 * inversion, 16 points per thread); the two buffers must not overlap */
int zkg_g1_batch_to_affine(int curve, int jac, int n, const uint64_t *d_src, uint64_t *d_tgt);
/* binary-coefficient MSM on device buffers (d_bits: npoints bytes); flags as zkg_msm */
int zkg_msm_binary(int curve, int npoints, const uint8_t *d_bits, const uint64_t *d_grps, uint64_t *tgt_host,
                   int flags);
/* nrows binary MSMs over one device point set; d_bits: nrows x npoints bytes, row-major;
 * tgt_host: nrows results of the form flags selects */
int zkg_multi_msm_binary(int curve, int nrows, int npoints, const uint8_t *d_bits, const uint64_t *d_grps,
                         uint64_t *tgt_host, int flags);
/* G1 FFT on device buffers (2^m projective points in, normalised projective points out) */
#define ZKG_G1_FFT_INVERSE 1
#define ZKG_G1_FFT_JAC 2    /* Jacobian inputs (outputs are the same normalised words) */
int zkg_g1_fft(int curve, int m, const uint64_t *gen, int mode, const uint64_t *d_src, uint64_t *d_tgt);

/* [tau^i G]_{i<n} for tau in standard form (examples/KZG.hs:59-62), affine Montgomery output */
int zkg_srs_generate(int curve, int n, const uint64_t *tau_std, uint64_t *d_out_affine);

/* Serving loop for KZG commitments (examples/KZG.hs:91-97: iNTT of the evaluations on the
 * size-2^logn domain, then an MSM of the Montgomery coefficients against the SRS prefix).
 * nworkers host threads, each with its own stream, admit jobs FIFO by device footprint
 * against hbm_budget bytes.  values/out stay owned by the caller until the job completes. */
int zkg_serve_create(int curve, const void *d_srs_prepared, int srs_n, size_t hbm_budget, int nworkers, void **server);
int zkg_serve_submit(void *server, int logn, const uint64_t *values, uint64_t *out_affine, long long *job_id);
int zkg_serve_drain(void *server);
/* optional size rule: no job starts while a job of at least 2^logn points runs (default 0,
 * off: every job is admitted by footprint only) */
int zkg_serve_set_exclusive_log(void *server, int logn);
int zkg_serve_job_times(void *server, long long id, double *queue_ms, double *run_ms, double *latency_ms);
int zkg_serve_stats(void *server, size_t *peak_bytes, int *max_concurrent);
/* submits job k at arrival_s[k] seconds after the call starts, waits for all of them */
int zkg_serve_run_trace(void *server, int njobs, const int *logn, const double *arrival_s,
                        const uint64_t *const *values, uint64_t *const *outs, double *latency_ms, double *wall_s);
int zkg_serve_destroy(void *server);

/* ---- schedule selection: statistics, autotuner, cost model, batched MSMs ---- */

/* scalar statistics on the device (flags: ZKG_MSM_MONT when the scalars are Montgomery Fr) */
int zkg_msm_scalar_stats(int curve, int npoints, const uint64_t *d_expos, int expo_nlimbs, int flags,
                         zkg_scalar_stats *out);
/* cost-model input row for one (job, candidate schedule); ZKG_MSM_NFEAT floats.  These three
 * describe a job with 4-limb scalars, whose BN254 form takes the GLV split when its scalars are
 * wider than 128 bits (2n points, 127-bit halves). */
#define ZKG_MSM_NFEAT 11
int zkg_msm_features(int curve, const zkg_scalar_stats *st, int window, int seg_log, float *out);
/* candidate schedules searched by the autotuner and the cost model; returns the count */
int zkg_msm_candidates(int curve, const zkg_scalar_stats *st, int max, int *windows, int *seg_logs);
/* the schedule a policy picks for a job (window, seg_log filled in) */
int zkg_msm_pick(int curve, const zkg_scalar_stats *st, int policy, zkg_msm_schedule *out);
/* Exhaustive autotuning: runs the MSM under each candidate (statistics-trimmed windows),
 * one warm-up then reps timed runs, ms_out[k] = median wall-clock ms of candidate k.
 * *identical = 1 when every candidate produced the same point. */
int zkg_msm_tune(int curve, int npoints, const uint64_t *d_expos, int expo_nlimbs, const uint64_t *d_grps, int flags,
                 int ncand, const int *windows, const int *seg_logs, int reps, double *ms_out, int *identical);
/* Many MSMs of uneven sizes (either curve, curves[k]) on nstreams concurrent streams, largest
 * first; each job's schedule follows sched->policy (window/seg_log ignored).  outs[k] host,
 * flags as zkg_msm.  Results are those of one zkg_msm call per job. */
int zkg_msm_batch(int njobs, const int *curves, const int *npoints, const uint64_t *const *d_expos, int expo_nlimbs,
                  const uint64_t *const *d_grps, uint64_t *const *outs, int flags, const zkg_msm_schedule *sched,
                  int nstreams, double *wall_ms);
/* The same job list as one captured plan: create() fits each job's schedule (sched->policy, on
 * the scalars present then), allocates every job's buffers and captures every job's pipeline
 * into one hipGraph whose nstreams branches run concurrently; run() replays it once, waits, and
 * writes outs[k] (flags as at create).  The device buffers must stay valid for the plan's
 * lifetime; run() rereads their contents, and returns ZKG_EINVAL if a scalar is wider than the
 * windows fitted at create (statistics policies).  Results are those of one zkg_msm per job.
 * The reference runs such jobs one zkg_msm-like call at a time (bn128_G1_proj.c:506-586). */
int zkg_msm_batch_create(int njobs, const int *curves, const int *npoints, const uint64_t *const *d_expos,
                         int expo_nlimbs, const uint64_t *const *d_grps, int flags, const zkg_msm_schedule *sched,
                         int nstreams, void **plan);
int zkg_msm_batch_run(void *plan, uint64_t *const *outs, double *wall_ms);
int zkg_msm_batch_destroy(void *plan);

/* Dense fp32 MLP (ReLU hidden layers, linear output, MSE loss, Adam) run by the library's
 * own HIP kernels: the schedule cost model.  dims has nlayers+1 entries; parameters are
 * laid out W_0 (dims[1] x dims[0], row-major), b_0, W_1, b_1, ... */
int zkg_mlp_create(int nlayers, const int *dims, unsigned seed, void **mlp);
int zkg_mlp_destroy(void *mlp);
long long zkg_mlp_nparams(void *mlp);
int zkg_mlp_get_params(void *mlp, float *host);
int zkg_mlp_set_params(void *mlp, const float *host);
int zkg_mlp_forward(void *mlp, int rows, const float *x_host, float *y_host);
int zkg_mlp_gradients(void *mlp, int rows, const float *x_host, const float *y_host, float *grad_host, float *loss);
int zkg_mlp_train(void *mlp, int rows, const float *x_host, const float *y_host, int epochs, float lr,
                  float *final_loss);
/* installs an MLP (dims[0] == ZKG_MSM_NFEAT, one output: predicted ln ms) as the cost model
 * of ZKG_SCHED_LEARNED; NULL removes it.  The caller keeps ownership. */
int zkg_msm_set_cost_model(void *mlp);

/* microbenchmarks: 0 = v_mad_u64_u32 per second, 1 = HBM copy bytes per second,
 * 2..5 = Montgomery multiplications per second (BN254 Fp, BN254 Fr, BLS Fp, BLS Fr) */
int zkg_microbench(int which, double *result);

/* timing of the calling thread's last MSM: total host wall clock, and per phase the HIP-event
 * durations of sort, accumulation, continuation fix, pyramid and read-back (recorded only
 * while zkg_msm_phase_timing is on) followed by the host combine */
int zkg_last_timing(double *ms_total, double *ms_phases, int max_phases, int *nphases);
/* 1: later MSMs (any thread) record HIP events between their phases; 0 (the default): they
 * record none, since each event record idles the stream a few microseconds */
int zkg_msm_phase_timing(int on);
/* schedule facts of the last MSM: {window c, windows W, sorted (nonzero-digit) entries, heavy buckets} */
int zkg_msm_last_stats(long long *out4);

#ifdef __cplusplus
}
#endif

#endif /* ZKGPU_H */
