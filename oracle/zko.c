/*
 * zko.c -- CPU oracle: a plain-C restatement of zikkurat-algebra's MSM and NTT.
 *
 * TEST INFRASTRUCTURE ONLY.  It is the checker for the GPU library: only tests/, the smoke
 * check and bench.py's CPU-baseline leg may load it; nothing in the product links it.
 *
 * What it restates (reference paths are relative to the reference repository root):
 *   limb arithmetic     lib/cbits/bigint/bigint256.c:108-356 (add/sub, schoolbook mul)
 *   Montgomery fields   lib/cbits/curves/fields/mont/bn128_Fr_mont.c:44-199 (neg/add/sub,
 *                       REDC word by word with n' = -p^-1 mod 2^64, canonical outputs)
 *   div_by_2            lib/cbits/curves/fields/std/bn128_Fr_std.c:241-249
 *   G1 projective       lib/cbits/curves/g1/proj/bn128_G1_proj.c:230-263 (dbl-2007-bl),
 *                       :272-313 (add-2015-rcb), :333-373 (madd-1998-cmo), :132-144 (to_affine)
 *   MSM                 bn128_G1_proj.c:506-586 (unsigned-window bucket method),
 *                       :597-605 (window c = round(log2 n - 3.5)), :631-673 (wrappers)
 *   NTT / iNTT          lib/cbits/curves/poly/mont/bn128_poly_mont.c:418-473 (recursive DIT),
 *                       :477-489 (shifted), :492-557 (asymmetric), :565-638 (inverse)
 *   FFT domain          lib/src/ZK/Algebra/Class/FFT.hs:60-68 (g_m = gen_max^(2^(M-m)))
 *   G1 FFT              bn128_G1_proj.c:791-902 (recursive DIT / DIF over points)
 * The BLS12-381 files are the same templates with 6-limb Fp.
 *
 * Departures, each of which leaves every in-contract result unchanged:
 *   - carries use unsigned __int128 (the reference's platform.h helpers have undefined
 *     behaviour; see oracle/Makefile);
 *   - the field inverse is Fermat's a^(p-2) instead of binary Euclid (the inverse is unique);
 *   - npoints <= 0 returns infinity and m = 0, 1 are handled without the reference's
 *     one-past-the-end twiddle writes (bn128_poly_mont.c:457-461).
 */
#include "zko.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;
#define MAXL 6

static const uint64_t bn_fp_p[4] = { 0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL };
static const uint64_t bn_fp_one[4] = { 0xd35d438dc58f0d9dULL, 0x0a78eb28f5c70b3dULL, 0x666ea36f7879462cULL, 0x0e0a77c19a07df2fULL };
static const uint64_t bn_fp_r2[4] = { 0xf32cfc5b538afa89ULL, 0xb5e71911d44501fbULL, 0x47ab1eff0a417ff6ULL, 0x06d89f71cab8351fULL };
#define BN_FP_NP 0x87d20782e4866389ULL
static const uint64_t bn_fr_p[4] = { 0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL };
static const uint64_t bn_fr_one[4] = { 0xac96341c4ffffffbULL, 0x36fc76959f60cd29ULL, 0x666ea36f7879462eULL, 0x0e0a77c19a07df2fULL };
static const uint64_t bn_fr_r2[4] = { 0x1bb8e645ae216da7ULL, 0x53fe3ab1e35c59e3ULL, 0x8c49833d53bb8085ULL, 0x0216d0b17f4e44a5ULL };
#define BN_FR_NP 0xc2e1f593efffffffULL
static const uint64_t bls_fp_p[6] = { 0xb9feffffffffaaabULL, 0x1eabfffeb153ffffULL, 0x6730d2a0f6b0f624ULL, 0x64774b84f38512bfULL, 0x4b1ba7b6434bacd7ULL, 0x1a0111ea397fe69aULL };
static const uint64_t bls_fp_one[6] = { 0x760900000002fffdULL, 0xebf4000bc40c0002ULL, 0x5f48985753c758baULL, 0x77ce585370525745ULL, 0x5c071a97a256ec6dULL, 0x15f65ec3fa80e493ULL };
static const uint64_t bls_fp_r2[6] = { 0xf4df1f341c341746ULL, 0x0a76e6a609d104f1ULL, 0x8de5476c4c95b6d5ULL, 0x67eb88a9939d83c0ULL, 0x9a793e85b519952dULL, 0x11988fe592cae3aaULL };
#define BLS_FP_NP 0x89f3fffcfffcfffdULL
static const uint64_t bls_fr_p[4] = { 0xffffffff00000001ULL, 0x53bda402fffe5bfeULL, 0x3339d80809a1d805ULL, 0x73eda753299d7d48ULL };
static const uint64_t bls_fr_one[4] = { 0x00000001fffffffeULL, 0x5884b7fa00034802ULL, 0x998c4fefecbc4ff5ULL, 0x1824b159acc5056fULL };
static const uint64_t bls_fr_r2[4] = { 0xc999e990f3f29c6dULL, 0x2b6cedcb87925c23ULL, 0x05d314967254398fULL, 0x0748d9d99f59ff11ULL };
#define BLS_FR_NP 0xfffffffeffffffffULL
static const uint64_t bn_b3[4] = { 0xf60647ce410d7ff7ULL, 0x2f3d6f4dd31bd011ULL, 0x2943337e3940c6d1ULL, 0x1d9598e8a7e39857ULL };
static const uint64_t bn_gx[4] = { 0xd35d438dc58f0d9dULL, 0x0a78eb28f5c70b3dULL, 0x666ea36f7879462cULL, 0x0e0a77c19a07df2fULL };
static const uint64_t bn_gy[4] = { 0xa6ba871b8b1e1b3aULL, 0x14f1d651eb8e167bULL, 0xccdd46def0f28c58ULL, 0x1c14ef83340fbe5eULL };
static const uint64_t bls_b3[6] = { 0x447600000027552eULL, 0xdcb8009a43480020ULL, 0x6f7ee9ce4a6e8b59ULL, 0xb10330b7c0a95bc6ULL, 0x6140b1fcfb1e54b7ULL, 0x0381be097f0bb4e1ULL };
static const uint64_t bls_gx[6] = { 0x5cb38790fd530c16ULL, 0x7817fc679976fff5ULL, 0x154f95c7143ba1c1ULL, 0xf0ae6acdf3d0e747ULL, 0xedce6ecc21dbf440ULL, 0x120177419e0bfb75ULL };
static const uint64_t bls_gy[6] = { 0xbaac93d50ce72271ULL, 0x8c22631a7918fd8eULL, 0xdd595f13570725ceULL, 0x51ac582950405194ULL, 0x0e1c8c3fad0059c0ULL, 0x0bbc3efc5008a26aULL };
static const uint64_t bn_fft_gen[4] = { 0x636e735580d13d9cULL, 0xa22bf3742445ffd6ULL, 0x56452ac01eb203d8ULL, 0x1860ef942963f9e7ULL };
static const uint64_t bls_fft_gen[4] = { 0xb9b58d8c5f0e466aULL, 0x5b1b4c801819d7ecULL, 0x0af53ae352a31e64ULL, 0x5bf3adda19e9b27bULL };

typedef struct {
  int nl;
  const uint64_t *p, *one, *r2;
  uint64_t np;
} field_t;

static const field_t FIELDS[2][2] = {
  { { 4, bn_fp_p, bn_fp_one, bn_fp_r2, BN_FP_NP }, { 4, bn_fr_p, bn_fr_one, bn_fr_r2, BN_FR_NP } },
  { { 6, bls_fp_p, bls_fp_one, bls_fp_r2, BLS_FP_NP }, { 4, bls_fr_p, bls_fr_one, bls_fr_r2, BLS_FR_NP } },
};

typedef struct {
  const field_t *fp;
  const uint64_t *b3, *gx, *gy;
} curve_t;

static const curve_t CURVES[2] = {
  { &FIELDS[0][0], bn_b3, bn_gx, bn_gy },
  { &FIELDS[1][0], bls_b3, bls_gx, bls_gy },
};

static const field_t *F(int curve, int field) { return &FIELDS[curve][field]; }

int zko_nlimbs(int curve, int field) { return F(curve, field)->nl; }

/* ---------------- limb helpers ---------------- */

static int is_zero(const field_t *f, const uint64_t *a) {
  for (int i = 0; i < f->nl; i++) if (a[i]) return 0;
  return 1;
}
static void cpy(const field_t *f, uint64_t *d, const uint64_t *s) { memmove(d, s, 8 * f->nl); }
static void set0(const field_t *f, uint64_t *d) { memset(d, 0, 8 * f->nl); }

/* returns carry */
static uint64_t add_n(int n, const uint64_t *a, const uint64_t *b, uint64_t *c) {
  u128 t = 0;
  for (int i = 0; i < n; i++) { t += (u128)a[i] + b[i]; c[i] = (uint64_t)t; t >>= 64; }
  return (uint64_t)t;
}
/* returns borrow */
static uint64_t sub_n(int n, const uint64_t *a, const uint64_t *b, uint64_t *c) {
  uint64_t br = 0;
  for (int i = 0; i < n; i++) {
    uint64_t ai = a[i], bi = b[i];
    uint64_t d = ai - bi - br;
    br = (ai < bi) || (ai == bi && br) ? 1 : 0;
    c[i] = d;
  }
  return br;
}
/* a >= b ? */
static int geq_n(int n, const uint64_t *a, const uint64_t *b) {
  for (int i = n - 1; i >= 0; i--) {
    if (a[i] > b[i]) return 1;
    if (a[i] < b[i]) return 0;
  }
  return 1;
}

/* ---------------- Montgomery field (bn128_Fr_mont.c:44-199) ---------------- */

static void f_add(const field_t *f, const uint64_t *a, const uint64_t *b, uint64_t *c) {
  uint64_t t[MAXL];
  add_n(f->nl, a, b, t);              /* the reference drops the carry too (bn128_Fr_mont.c:85-88) */
  if (geq_n(f->nl, t, f->p)) sub_n(f->nl, t, f->p, t);
  cpy(f, c, t);
}
static void f_sub(const field_t *f, const uint64_t *a, const uint64_t *b, uint64_t *c) {
  uint64_t t[MAXL];
  if (sub_n(f->nl, a, b, t)) add_n(f->nl, t, f->p, t);
  cpy(f, c, t);
}
static void f_neg(const field_t *f, const uint64_t *a, uint64_t *c) {
  if (is_zero(f, a)) { set0(f, c); return; }
  uint64_t t[MAXL];
  sub_n(f->nl, f->p, a, t);
  cpy(f, c, t);
}
/* schoolbook product then word-by-word REDC (bigint256.c:267-356, bn128_Fr_mont.c:140-169) */
static void f_mul(const field_t *f, const uint64_t *a, const uint64_t *b, uint64_t *c) {
  int n = f->nl;
  uint64_t T[2 * MAXL + 1];
  memset(T, 0, sizeof(T));
  for (int i = 0; i < n; i++) {
    u128 carry = 0;
    for (int j = 0; j < n; j++) {
      u128 x = (u128)a[i] * b[j] + T[i + j] + carry;
      T[i + j] = (uint64_t)x;
      carry = x >> 64;
    }
    T[i + n] = (uint64_t)carry;
  }
  T[2 * n] = 0;
  for (int i = 0; i < n; i++) {
    uint64_t m = T[i] * f->np;
    u128 carry = 0;
    for (int j = 0; j < n; j++) {
      u128 x = (u128)m * f->p[j] + T[i + j] + carry;
      T[i + j] = (uint64_t)x;
      carry = x >> 64;
    }
    for (int j = i + n; j <= 2 * n && carry; j++) {
      u128 x = (u128)T[j] + carry;
      T[j] = (uint64_t)x;
      carry = x >> 64;
    }
  }
  uint64_t t[MAXL];
  memcpy(t, T + n, 8 * n);
  if (geq_n(n, t, f->p)) sub_n(n, t, f->p, t);
  cpy(f, c, t);
}
static void f_sqr(const field_t *f, const uint64_t *a, uint64_t *c) { f_mul(f, a, a, c); }

/* Fermat inverse a^(p-2); inverse of 0 is 0 (the reference returns 0 as well) */
static void f_inv(const field_t *f, const uint64_t *a, uint64_t *c) {
  uint64_t e[MAXL], two[MAXL] = {2, 0, 0, 0, 0, 0}, acc[MAXL], base[MAXL];
  sub_n(f->nl, f->p, two, e);
  cpy(f, acc, f->one);
  cpy(f, base, a);
  for (int i = f->nl * 64 - 1; i >= 0; i--) {
    f_sqr(f, acc, acc);
    if ((e[i / 64] >> (i % 64)) & 1) f_mul(f, acc, base, acc);
  }
  cpy(f, c, acc);
}
/* bn128_Fr_std.c:241-249: x even -> x>>1, odd -> (x>>1) + (p+1)/2 */
static void f_div2(const field_t *f, const uint64_t *a, uint64_t *c) {
  int n = f->nl;
  uint64_t t[MAXL], h[MAXL], one[MAXL] = {1, 0, 0, 0, 0, 0};
  int odd = (int)(a[0] & 1);
  for (int i = 0; i < n; i++) t[i] = (a[i] >> 1) | (i + 1 < n ? a[i + 1] << 63 : 0);
  if (odd) {
    add_n(n, f->p, one, h);
    for (int i = 0; i < n; i++) h[i] = (h[i] >> 1) | (i + 1 < n ? h[i + 1] << 63 : 0);
    add_n(n, t, h, t);
  }
  cpy(f, c, t);
}
static void f_to_std(const field_t *f, const uint64_t *a, uint64_t *c) {
  uint64_t one[MAXL] = {1, 0, 0, 0, 0, 0};
  f_mul(f, a, one, c);
}
static void f_from_std(const field_t *f, const uint64_t *a, uint64_t *c) { f_mul(f, a, f->r2, c); }

void zko_f_add(int cv, int fd, const uint64_t *a, const uint64_t *b, uint64_t *c) { f_add(F(cv, fd), a, b, c); }
void zko_f_sub(int cv, int fd, const uint64_t *a, const uint64_t *b, uint64_t *c) { f_sub(F(cv, fd), a, b, c); }
void zko_f_neg(int cv, int fd, const uint64_t *a, uint64_t *c) { f_neg(F(cv, fd), a, c); }
void zko_f_mul(int cv, int fd, const uint64_t *a, const uint64_t *b, uint64_t *c) { f_mul(F(cv, fd), a, b, c); }
void zko_f_inv(int cv, int fd, const uint64_t *a, uint64_t *c) { f_inv(F(cv, fd), a, c); }
void zko_f_div_by_2(int cv, int fd, const uint64_t *a, uint64_t *c) { f_div2(F(cv, fd), a, c); }
void zko_f_to_std(int cv, int fd, const uint64_t *a, uint64_t *c) { f_to_std(F(cv, fd), a, c); }
void zko_f_from_std(int cv, int fd, const uint64_t *a, uint64_t *c) { f_from_std(F(cv, fd), a, c); }
void zko_f_pow_u64(int cv, int fd, const uint64_t *a, uint64_t e, uint64_t *c) {
  const field_t *f = F(cv, fd);
  uint64_t acc[MAXL], base[MAXL];
  cpy(f, acc, f->one);
  cpy(f, base, a);
  while (e) {
    if (e & 1) f_mul(f, acc, base, acc);
    f_sqr(f, base, base);
    e >>= 1;
  }
  cpy(f, c, acc);
}

/* ---------------- G1, homogeneous projective (bn128_G1_proj.c) ---------------- */

#define PX(P) (P)
#define PY(P) ((P) + n)
#define PZ(P) ((P) + 2 * n)

static void proj_set_inf(const curve_t *C, uint64_t *P) {
  int n = C->fp->nl;
  set0(C->fp, PX(P)); cpy(C->fp, PY(P), C->fp->one); set0(C->fp, PZ(P));
}
/* bn128_G1_proj.c:171-175 */
static int proj_is_inf(const curve_t *C, const uint64_t *P) {
  int n = C->fp->nl;
  return is_zero(C->fp, PZ(P)) && !is_zero(C->fp, PY(P)) && is_zero(C->fp, PX(P));
}
/* bn128_G1_affine.c:57-60: affine infinity is (0,0) */
static int aff_is_inf(const curve_t *C, const uint64_t *A) {
  int n = C->fp->nl;
  return is_zero(C->fp, A) && is_zero(C->fp, A + n);
}
static void proj_from_affine(const curve_t *C, const uint64_t *A, uint64_t *P) {
  int n = C->fp->nl;
  if (aff_is_inf(C, A)) { proj_set_inf(C, P); return; }
  cpy(C->fp, PX(P), A); cpy(C->fp, PY(P), A + n); cpy(C->fp, PZ(P), C->fp->one);
}

/* dbl-2007-bl, bn128_G1_proj.c:230-263 */
static void proj_dbl(const curve_t *C, const uint64_t *P1, uint64_t *P3) {
  const field_t *f = C->fp; int n = f->nl;
  uint64_t XX[MAXL], ZZ[MAXL], w[MAXL], s[MAXL], ss[MAXL], R[MAXL], RR[MAXL], B[MAXL], h[MAXL];
  uint64_t X3[MAXL], Y3[MAXL], Z3[MAXL];
  f_sqr(f, PX(P1), XX);
  f_sqr(f, PZ(P1), ZZ);
  f_add(f, XX, XX, w); f_add(f, w, XX, w);           /* w = 3 XX (a = 0) */
  f_mul(f, PY(P1), PZ(P1), s); f_add(f, s, s, s);   /* s = 2 Y Z */
  f_sqr(f, s, ss);
  f_mul(f, PY(P1), s, R);
  f_sqr(f, R, RR);
  f_add(f, PX(P1), R, B); f_sqr(f, B, B); f_sub(f, B, XX, B); f_sub(f, B, RR, B);
  f_sqr(f, w, h); f_sub(f, h, B, h); f_sub(f, h, B, h);
  f_mul(f, s, ss, Z3);
  f_mul(f, h, s, X3);
  f_sub(f, B, h, Y3); f_mul(f, Y3, w, Y3); f_sub(f, Y3, RR, Y3); f_sub(f, Y3, RR, Y3);
  (void)ZZ;
  cpy(f, PX(P3), X3); cpy(f, PY(P3), Y3); cpy(f, PZ(P3), Z3);
}

/* add-2015-rcb for a = 0, bn128_G1_proj.c:272-313 */
static void proj_add(const curve_t *C, const uint64_t *P1, const uint64_t *P2, uint64_t *P3) {
  const field_t *f = C->fp; int n = f->nl;
  uint64_t t0[MAXL], t1[MAXL], t2[MAXL], t3[MAXL], t4[MAXL], t5[MAXL], X3[MAXL], Y3[MAXL], Z3[MAXL];
  f_mul(f, PX(P1), PX(P2), t0);
  f_mul(f, PY(P1), PY(P2), t1);
  f_mul(f, PZ(P1), PZ(P2), t2);
  f_add(f, PX(P1), PY(P1), t3); f_add(f, PX(P2), PY(P2), t4); f_mul(f, t3, t4, t3);
  f_add(f, t0, t1, t4); f_sub(f, t3, t4, t3);
  f_add(f, PX(P1), PZ(P1), t4); f_add(f, PX(P2), PZ(P2), t5); f_mul(f, t4, t5, t4);
  f_add(f, t0, t2, t5); f_sub(f, t4, t5, t4);
  f_add(f, PY(P1), PZ(P1), t5); f_add(f, PY(P2), PZ(P2), X3); f_mul(f, t5, X3, t5);
  f_add(f, t1, t2, X3); f_sub(f, t5, X3, t5);
  f_mul(f, C->b3, t2, X3);                           /* X3 = b3 t2 */
  cpy(f, Z3, X3);
  f_sub(f, t1, Z3, X3);
  f_add(f, Z3, t1, Z3);
  f_mul(f, X3, Z3, Y3);
  f_add(f, t0, t0, t1); f_add(f, t1, t0, t1);
  f_mul(f, C->b3, t4, t4);
  f_mul(f, t1, t4, t0); f_add(f, Y3, t0, Y3);
  f_mul(f, t4, t5, t0);
  f_mul(f, X3, t3, X3); f_sub(f, X3, t0, X3);
  f_mul(f, t1, t3, t0);
  f_mul(f, Z3, t5, Z3); f_add(f, Z3, t0, Z3);
  cpy(f, PX(P3), X3); cpy(f, PY(P3), Y3); cpy(f, PZ(P3), Z3);
}

/* madd-1998-cmo, bn128_G1_proj.c:333-373 (projective + affine) */
static void proj_madd(const curve_t *C, const uint64_t *P1, const uint64_t *A2, uint64_t *P3) {
  const field_t *f = C->fp; int n = f->nl;
  if (proj_is_inf(C, P1)) { proj_from_affine(C, A2, P3); return; }
  if (aff_is_inf(C, A2)) { if (P3 != P1) memmove(P3, P1, 24 * n); return; }
  uint64_t u[MAXL], uu[MAXL], v[MAXL], vv[MAXL], vvv[MAXL], R[MAXL], A[MAXL];
  uint64_t X3[MAXL], Y3[MAXL], Z3[MAXL];
  f_mul(f, A2 + n, PZ(P1), u); f_sub(f, u, PY(P1), u);
  f_mul(f, A2, PZ(P1), v); f_sub(f, v, PX(P1), v);
  if (is_zero(f, u) && is_zero(f, v)) { proj_dbl(C, P1, P3); return; }
  f_sqr(f, u, uu);
  f_sqr(f, v, vv);
  f_mul(f, v, vv, vvv);
  f_mul(f, vv, PX(P1), R);
  f_mul(f, uu, PZ(P1), A); f_sub(f, A, vvv, A); f_sub(f, A, R, A); f_sub(f, A, R, A);
  f_mul(f, v, A, X3);
  f_mul(f, PZ(P1), vvv, Z3);
  f_sub(f, R, A, R);
  f_mul(f, vvv, PY(P1), vvv);
  f_mul(f, u, R, Y3); f_sub(f, Y3, vvv, Y3);
  cpy(f, PX(P3), X3); cpy(f, PY(P3), Y3); cpy(f, PZ(P3), Z3);
}

/* bn128_G1_proj.c:132-144: infinity -> all 0xff bytes */
static void proj_to_affine(const curve_t *C, const uint64_t *P, uint64_t *A) {
  const field_t *f = C->fp; int n = f->nl;
  if (is_zero(f, PZ(P))) { memset(A, 0xff, 16 * n); return; }
  uint64_t zi[MAXL], x[MAXL], y[MAXL];
  f_inv(f, PZ(P), zi);
  f_mul(f, PX(P), zi, x);
  f_mul(f, PY(P), zi, y);
  cpy(f, A, x); cpy(f, A + n, y);
}

void zko_g1_proj_add(int cv, const uint64_t *p, const uint64_t *q, uint64_t *r) { proj_add(&CURVES[cv], p, q, r); }
void zko_g1_proj_dbl(int cv, const uint64_t *p, uint64_t *r) { proj_dbl(&CURVES[cv], p, r); }
void zko_g1_proj_madd(int cv, const uint64_t *p, const uint64_t *a, uint64_t *r) { proj_madd(&CURVES[cv], p, a, r); }
void zko_g1_proj_to_affine(int cv, const uint64_t *p, uint64_t *a) { proj_to_affine(&CURVES[cv], p, a); }
void zko_g1_proj_from_affine(int cv, const uint64_t *a, uint64_t *p) { proj_from_affine(&CURVES[cv], a, p); }

/* bn128_G1_proj.c:78-97: (X/Z, Y/Z, 1), infinity -> (0, 1, 0) */
void zko_g1_proj_normalize(int cv, const uint64_t *P, uint64_t *R) {
  const curve_t *C = &CURVES[cv]; const field_t *f = C->fp; int n = f->nl;
  if (is_zero(f, PZ(P))) { proj_set_inf(C, R); return; }
  uint64_t A[2 * MAXL];
  proj_to_affine(C, P, A);
  cpy(f, PX(R), A); cpy(f, PY(R), A + n); cpy(f, PZ(R), f->one);
}

/* double-and-add scalar multiplication (group law only; any exact method gives the same point) */
void zko_g1_proj_scl(int cv, const uint64_t *k, int knl, const uint64_t *P, uint64_t *R) {
  const curve_t *C = &CURVES[cv]; int n = C->fp->nl;
  uint64_t acc[3 * MAXL];
  proj_set_inf(C, acc);
  for (int i = 64 * knl - 1; i >= 0; i--) {
    /* dbl-2007-bl maps (0:1:0) to (0:0:0): skip it while the accumulator is infinity, as the
       reference's windowed scalar multiplication does (bn128_G1_proj.c:448-454) */
    if (!is_zero(C->fp, PZ(acc))) proj_dbl(C, acc, acc);
    if ((k[i / 64] >> (i % 64)) & 1) proj_add(C, acc, P, acc);
  }
  memcpy(R, acc, 24 * n);
}

/* Jacobian representation of the same point: (X/Z^2 ... ) normalised to Z = 1, or [1:1:0]
 * for infinity (bn128_G1_jac.c:183-188) */
void zko_g1_proj_to_jac(int cv, const uint64_t *P, uint64_t *J) {
  const curve_t *C = &CURVES[cv]; const field_t *f = C->fp; int n = f->nl;
  if (is_zero(f, PZ(P))) {
    cpy(f, J, f->one); cpy(f, J + n, f->one); set0(f, J + 2 * n);
    return;
  }
  uint64_t A[2 * MAXL];
  proj_to_affine(C, P, A);
  cpy(f, J, A); cpy(f, J + n, A + n); cpy(f, J + 2 * n, f->one);
}
/* bn128_G1_jac.c:120-137 */
void zko_g1_jac_to_affine(int cv, const uint64_t *J, uint64_t *A) {
  const curve_t *C = &CURVES[cv]; const field_t *f = C->fp; int n = f->nl;
  if (is_zero(f, J + 2 * n)) { memset(A, 0xff, 16 * n); return; }
  uint64_t zi[MAXL], zi2[MAXL], zi3[MAXL], x[MAXL], y[MAXL];
  f_inv(f, J + 2 * n, zi);
  f_mul(f, zi, zi, zi2);
  f_mul(f, zi, zi2, zi3);
  f_mul(f, J, zi2, x);
  f_mul(f, J + n, zi3, y);
  cpy(f, A, x); cpy(f, A + n, y);
}

/* ---------------- MSM (bn128_G1_proj.c:506-673) ---------------- */

int zko_msm_window(int npoints) {
  if (npoints <= 0) return 1;
  int c = (int)round(log2((double)npoints) - 3.5);
  if (c < 1) c = 1;
  if (c > 64) c = 64;
  return c;
}

void zko_msm_std_proj_variable(int cv, int npoints, const uint64_t *expos, const uint64_t *grps,
                               uint64_t *tgt, int enl, int window) {
  const curve_t *C = &CURVES[cv]; int n = C->fp->nl;
  proj_set_inf(C, tgt);
  if (npoints <= 0 || window < 1 || window > 30) return;
  int nbits = 64 * enl;
  int nwin = (nbits + window - 1) / window;
  int nbuck = 1 << window;
  uint64_t *S = (uint64_t *)malloc(sizeof(uint64_t) * 3 * n * (size_t)(nbuck - 1));
  if (!S) abort();
  for (int K = nwin - 1; K >= 0; K--) {
    int A = K * window, B = A + window;
    if (B > nbits) B = nbits;
    uint64_t mask = (1ULL << (B - A)) - 1;
    for (int b = 1; b < nbuck; b++) proj_set_inf(C, S + (size_t)(b - 1) * 3 * n);
    for (int j = 0; j < npoints; j++) {
      const uint64_t *e = expos + (size_t)enl * j;
      int ad = A >> 6, am = A & 63;
      uint64_t d = e[ad] >> am;
      if (((B - 1) >> 6) != ad) d |= e[ad + 1] << (64 * (ad + 1) - A);
      d &= mask;
      if (d) {
        uint64_t *s = S + (size_t)(d - 1) * 3 * n;
        proj_madd(C, s, grps + (size_t)2 * n * j, s);
      }
    }
    uint64_t T[3 * MAXL], R[3 * MAXL];
    proj_set_inf(C, T); proj_set_inf(C, R);
    for (int b = nbuck - 1; b > 0; b--) {
      proj_add(C, T, S + (size_t)(b - 1) * 3 * n, T);
      proj_add(C, R, T, R);
    }
    if (!proj_is_inf(C, tgt))
      for (int i = 0; i < window; i++) proj_dbl(C, tgt, tgt);
    proj_add(C, tgt, R, tgt);
  }
  free(S);
}

void zko_msm_std_proj(int cv, int np, const uint64_t *e, const uint64_t *g, uint64_t *t, int enl) {
  zko_msm_std_proj_variable(cv, np, e, g, t, enl, zko_msm_window(np));
}
void zko_msm_std_affine(int cv, int np, const uint64_t *e, const uint64_t *g, uint64_t *t, int enl) {
  uint64_t P[3 * MAXL];
  zko_msm_std_proj(cv, np, e, g, P, enl);
  proj_to_affine(&CURVES[cv], P, t);
}
/* bn128_G1_proj.c:631-645: Montgomery scalars are converted with Fr to_std first */
void zko_msm_mont_affine(int cv, int np, const uint64_t *e, const uint64_t *g, uint64_t *t, int enl) {
  const field_t *fr = F(cv, ZKO_FR);
  uint64_t *s = (uint64_t *)malloc(8 * (size_t)enl * (np > 0 ? np : 1));
  if (!s) abort();
  for (int i = 0; i < np; i++) f_to_std(fr, e + (size_t)enl * i, s + (size_t)enl * i);
  zko_msm_std_affine(cv, np, s, g, t, enl);
  free(s);
}

/* ---------------- NTT (bn128_poly_mont.c:418-638) ---------------- */

static void ntt_fwd_rec(const field_t *f, int m, int stride, const uint64_t *gp, const uint64_t *src,
                        uint64_t *buf, uint64_t *tgt) {
  int L = f->nl;
  if (m == 0) { cpy(f, tgt, src); return; }
  if (m == 1) {
    uint64_t a[MAXL];
    f_add(f, src, src + (size_t)stride * L, a);
    f_sub(f, src, src + (size_t)stride * L, tgt + L);
    cpy(f, tgt, a);
    return;
  }
  size_t N = (size_t)1 << m, h = N >> 1;
  ntt_fwd_rec(f, m - 1, stride << 1, gp, src, buf + N * L, buf);
  ntt_fwd_rec(f, m - 1, stride << 1, gp, src + (size_t)stride * L, buf + N * L, buf + h * L);
  for (size_t j = 0; j < h; j++) {
    uint64_t t[MAXL];
    f_mul(f, buf + (j + h) * L, gp + j * stride * L, t);
    f_sub(f, buf + j * L, t, tgt + (j + h) * L);
    f_add(f, buf + j * L, t, tgt + j * L);
  }
}

/* gpows[j] = g^j for j < max(N/2, 1) (bn128_poly_mont.c:456-466) */
static uint64_t *powers(const field_t *f, const uint64_t *g, const uint64_t *first, size_t cnt) {
  int L = f->nl;
  uint64_t *gp = (uint64_t *)malloc(8 * L * (cnt ? cnt : 1));
  if (!gp) abort();
  cpy(f, gp, first);
  for (size_t i = 1; i < cnt; i++) f_mul(f, gp + (i - 1) * L, g, gp + i * L);
  return gp;
}

void zko_ntt_forward(int cv, int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt) {
  const field_t *f = F(cv, ZKO_FR); int L = f->nl;
  size_t N = (size_t)1 << m;
  uint64_t *gp = powers(f, gen, f->one, N / 2);
  uint64_t *buf = (uint64_t *)malloc(8 * L * 2 * N);
  if (!buf) abort();
  ntt_fwd_rec(f, m, 1, gp, src, buf, tgt);
  free(buf); free(gp);
}

void zko_ntt_forward_shifted(int cv, const uint64_t *eta, int m, const uint64_t *gen,
                             const uint64_t *src, uint64_t *tgt) {
  const field_t *f = F(cv, ZKO_FR); int L = f->nl;
  size_t N = (size_t)1 << m;
  uint64_t *sh = (uint64_t *)malloc(8 * L * N), x[MAXL];
  if (!sh) abort();
  cpy(f, x, f->one);
  for (size_t i = 0; i < N; i++) { f_mul(f, src + i * L, x, sh + i * L); f_mul(f, x, eta, x); }
  zko_ntt_forward(cv, m, gen, sh, tgt);
  free(sh);
}

void zko_ntt_forward_asymmetric(int cv, int ms, int mt, const uint64_t *gs, const uint64_t *gt,
                                const uint64_t *src, uint64_t *tgt) {
  const field_t *f = F(cv, ZKO_FR); int L = f->nl;
  if (mt < ms) abort();
  size_t Ns = (size_t)1 << ms, K = (size_t)1 << (mt - ms);
  uint64_t *sh = (uint64_t *)malloc(8 * L * Ns), *small = (uint64_t *)malloc(8 * L * Ns);
  uint64_t eta[MAXL], x[MAXL];
  if (!sh || !small) abort();
  cpy(f, eta, f->one);
  for (size_t k = 0; k < K; k++) {
    if (k) f_mul(f, eta, gt, eta);
    cpy(f, x, f->one);
    for (size_t i = 0; i < Ns; i++) { f_mul(f, src + i * L, x, sh + i * L); f_mul(f, x, eta, x); }
    zko_ntt_forward(cv, ms, gs, sh, small);
    for (size_t i = 0; i < Ns; i++) memcpy(tgt + (i * K + k) * L, small + i * L, 8 * L);
  }
  free(sh); free(small);
}

static void ntt_inv_rec(const field_t *f, int m, int stride, const uint64_t *gp, const uint64_t *src,
                        uint64_t *buf, uint64_t *tgt) {
  int L = f->nl;
  if (m == 0) { cpy(f, tgt, src); return; }
  if (m == 1) {
    uint64_t a[MAXL], b[MAXL];
    f_add(f, src, src + L, a);
    f_sub(f, src, src + L, b);
    f_div2(f, a, tgt);
    f_div2(f, b, tgt + (size_t)stride * L);
    return;
  }
  size_t N = (size_t)1 << m, h = N >> 1;
  for (size_t j = 0; j < h; j++) {
    f_add(f, src + j * L, src + (j + h) * L, buf + j * L);
    f_sub(f, src + j * L, src + (j + h) * L, buf + (j + h) * L);
    f_div2(f, buf + j * L, buf + j * L);
    f_mul(f, buf + (j + h) * L, gp + j * stride * L, buf + (j + h) * L);
  }
  ntt_inv_rec(f, m - 1, stride << 1, gp, buf, buf + N * L, tgt);
  ntt_inv_rec(f, m - 1, stride << 1, gp, buf + h * L, buf + N * L, tgt + (size_t)stride * L);
}

/* gpows[j] = (1/2) g^-j (bn128_poly_mont.c:601-622) */
void zko_ntt_inverse(int cv, int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt) {
  const field_t *f = F(cv, ZKO_FR); int L = f->nl;
  size_t N = (size_t)1 << m;
  uint64_t ginv[MAXL], half[MAXL];
  f_inv(f, gen, ginv);
  f_div2(f, f->one, half);
  uint64_t *gp = powers(f, ginv, half, N / 2);
  uint64_t *buf = (uint64_t *)malloc(8 * L * 2 * N);
  if (!buf) abort();
  ntt_inv_rec(f, m, 1, gp, src, buf, tgt);
  free(buf); free(gp);
}

void zko_ntt_inverse_shifted(int cv, const uint64_t *eta, int m, const uint64_t *gen,
                             const uint64_t *src, uint64_t *tgt) {
  const field_t *f = F(cv, ZKO_FR); int L = f->nl;
  size_t N = (size_t)1 << m;
  uint64_t *un = (uint64_t *)malloc(8 * L * N), x[MAXL];
  if (!un) abort();
  zko_ntt_inverse(cv, m, gen, src, un);
  cpy(f, x, f->one);
  for (size_t i = 0; i < N; i++) { f_mul(f, un + i * L, x, tgt + i * L); f_mul(f, x, eta, x); }
  free(un);
}

void zko_fft_generator(int cv, int m, uint64_t *gen) {
  const field_t *f = F(cv, ZKO_FR);
  int M = cv == ZKO_BN254 ? 28 : 32;
  uint64_t g[MAXL];
  cpy(f, g, cv == ZKO_BN254 ? bn_fft_gen : bls_fft_gen);
  if (m < 0 || m > M) { memset(gen, 0, 8 * f->nl); return; }
  for (int i = 0; i < M - m; i++) f_sqr(f, g, g);
  cpy(f, gen, g);
}

/* ---------------- G1 FFT (bn128_G1_proj.c:791-902) ----------------
 * Forward: recursive radix-2 DIT over projective points with Fr twiddles gen^j (Montgomery),
 * each butterfly (u + g^j v, u - g^j v) computed through a scalar multiplication.  Inverse:
 * the mirrored DIF recursion with (x + y)/2 and (x - y) g^-j / 2 at every level.  Both
 * normalise every output (X/Z, Y/Z, 1), infinity (0, 1, 0). */
static void proj_neg(const curve_t *C, const uint64_t *P, uint64_t *R) {
  const field_t *f = C->fp; int n = f->nl;
  cpy(f, PX(R), PX(P)); f_neg(f, PY(P), PY(R)); cpy(f, PZ(R), PZ(P));
}
static void proj_sub(const curve_t *C, const uint64_t *P, const uint64_t *Q, uint64_t *R) {
  uint64_t t[3 * MAXL];
  proj_neg(C, Q, t);
  proj_add(C, P, t, R);
}
/* scalar in Montgomery Fr form (bn128_G1_proj_scl_Fr_mont) */
static void proj_scl_mont(int cv, const uint64_t *k, const uint64_t *P, uint64_t *R) {
  uint64_t ks[4];
  f_to_std(F(cv, ZKO_FR), k, ks);
  zko_g1_proj_scl(cv, ks, 4, P, R);
}

static void g1_fft_fwd_rec(int cv, int m, size_t stride, const uint64_t *gen, const uint64_t *src, uint64_t *buf,
                           uint64_t *tgt) {
  const curve_t *C = &CURVES[cv]; const field_t *fr = F(cv, ZKO_FR); int n = C->fp->nl;
  const size_t GN = 3 * (size_t)n;
  if (m == 0) { memcpy(tgt, src, 8 * GN); return; }
  if (m == 1) {
    proj_add(C, src, src + stride * GN, tgt);
    proj_sub(C, src, src + stride * GN, tgt + GN);
    return;
  }
  size_t N = (size_t)1 << m, half = N >> 1;
  uint64_t gpow[4];
  f_mul(fr, gen, gen, gpow);
  g1_fft_fwd_rec(cv, m - 1, stride << 1, gpow, src, buf + N * GN, buf);
  g1_fft_fwd_rec(cv, m - 1, stride << 1, gpow, src + stride * GN, buf + N * GN, buf + half * GN);
  cpy(fr, gpow, fr->one);
  for (size_t j = 0; j < half; j++) {
    proj_scl_mont(cv, gpow, buf + (j + half) * GN, tgt + j * GN);
    proj_neg(C, tgt + j * GN, tgt + (j + half) * GN);
    proj_add(C, tgt + j * GN, buf + j * GN, tgt + j * GN);
    proj_add(C, tgt + (j + half) * GN, buf + j * GN, tgt + (j + half) * GN);
    f_mul(fr, gpow, gen, gpow);
  }
}

static void g1_fft_inv_rec(int cv, int m, size_t tstride, const uint64_t *gen, const uint64_t *src, uint64_t *buf,
                           uint64_t *tgt) {
  const curve_t *C = &CURVES[cv]; const field_t *fr = F(cv, ZKO_FR); int n = C->fp->nl;
  const size_t GN = 3 * (size_t)n;
  uint64_t half2[4];  /* 1/2 in Montgomery form */
  f_div2(fr, fr->one, half2);
  if (m == 0) { memcpy(tgt, src, 8 * GN); return; }
  if (m == 1) {
    proj_add(C, src, src + GN, tgt);
    proj_sub(C, src, src + GN, tgt + tstride * GN);
    proj_scl_mont(cv, half2, tgt, tgt);
    proj_scl_mont(cv, half2, tgt + tstride * GN, tgt + tstride * GN);
    return;
  }
  size_t N = (size_t)1 << m, half = N >> 1;
  uint64_t ginv[4], gpow[4];
  f_inv(fr, gen, ginv);
  cpy(fr, gpow, half2);
  for (size_t j = 0; j < half; j++) {
    proj_add(C, src + j * GN, src + (j + half) * GN, buf + j * GN);
    proj_sub(C, src + j * GN, src + (j + half) * GN, buf + (j + half) * GN);
    proj_scl_mont(cv, half2, buf + j * GN, buf + j * GN);
    proj_scl_mont(cv, gpow, buf + (j + half) * GN, buf + (j + half) * GN);
    f_mul(fr, gpow, ginv, gpow);
  }
  f_mul(fr, gen, gen, gpow);
  g1_fft_inv_rec(cv, m - 1, tstride << 1, gpow, buf, buf + N * GN, tgt);
  g1_fft_inv_rec(cv, m - 1, tstride << 1, gpow, buf + half * GN, buf + N * GN, tgt + tstride * GN);
}

static void g1_fft(int cv, int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt, int inverse) {
  const curve_t *C = &CURVES[cv]; int n = C->fp->nl;
  const size_t GN = 3 * (size_t)n, N = (size_t)1 << m;
  uint64_t *buf = (uint64_t *)malloc(8 * GN * 2 * N);
  if (!buf) abort();
  if (inverse) g1_fft_inv_rec(cv, m, 1, gen, src, buf, tgt);
  else g1_fft_fwd_rec(cv, m, 1, gen, src, buf, tgt);
  free(buf);
  for (size_t i = 0; i < N; i++) zko_g1_proj_normalize(cv, tgt + i * GN, tgt + i * GN);
}

void zko_g1_fft_forward(int cv, int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt) {
  g1_fft(cv, m, gen, src, tgt, 0);
}
void zko_g1_fft_inverse(int cv, int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt) {
  g1_fft(cv, m, gen, src, tgt, 1);
}

/* Jacobian G1 FFT (bn128_G1_jac.c:840-950): the same recursions over Jacobian points.  A
 * Jacobian (X:Y:Z) is the projective (X Z : Y : Z^3) of the same affine point, and both
 * normalisations give (x, y, 1), infinity (0, 1, 0) (bn128_G1_jac.c:62-85), so the projective
 * recursions above compute the same outputs word for word. */
static void jac_to_proj(const curve_t *C, const uint64_t *J, uint64_t *P) {
  const field_t *f = C->fp; int n = f->nl;
  uint64_t z2[MAXL];
  if (is_zero(f, J + 2 * n)) { proj_set_inf(C, P); return; }
  f_mul(f, J, J + 2 * n, PX(P));
  cpy(f, PY(P), J + n);
  f_sqr(f, J + 2 * n, z2);
  f_mul(f, z2, J + 2 * n, PZ(P));
}
static void g1_jac_fft(int cv, int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt, int inverse) {
  const curve_t *C = &CURVES[cv]; int n = C->fp->nl;
  const size_t GN = 3 * (size_t)n, N = (size_t)1 << m;
  uint64_t *p = (uint64_t *)malloc(8 * GN * N);
  if (!p) abort();
  for (size_t i = 0; i < N; i++) jac_to_proj(C, src + i * GN, p + i * GN);
  g1_fft(cv, m, gen, p, tgt, inverse);
  free(p);
}
void zko_g1_jac_fft_forward(int cv, int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt) {
  g1_jac_fft(cv, m, gen, src, tgt, 0);
}
void zko_g1_jac_fft_inverse(int cv, int m, const uint64_t *gen, const uint64_t *src, uint64_t *tgt) {
  g1_jac_fft(cv, m, gen, src, tgt, 1);
}

/* ---------------- batch conversions (bn128_G1_proj.c:146-166, bn128_G1_jac.c:139-158) -------- */
void zko_g1_batch_to_affine(int cv, int jac, int N, const uint64_t *src, uint64_t *tgt) {
  const curve_t *C = &CURVES[cv]; int n = C->fp->nl;
  for (int i = 0; i < N; i++) {
    if (jac) zko_g1_jac_to_affine(cv, src + (size_t)i * 3 * n, tgt + (size_t)i * 2 * n);
    else proj_to_affine(C, src + (size_t)i * 3 * n, tgt + (size_t)i * 2 * n);
  }
}
/* affine (0,0) -> (0 : 1 : 0), Jacobian [1 : 1 : 0] (bn128_G1_jac.c:108-117, :183-187) */
void zko_g1_batch_from_affine(int cv, int jac, int N, const uint64_t *src, uint64_t *tgt) {
  const curve_t *C = &CURVES[cv]; const field_t *f = C->fp; int n = f->nl;
  for (int i = 0; i < N; i++) {
    const uint64_t *a = src + (size_t)i * 2 * n;
    uint64_t *p = tgt + (size_t)i * 3 * n;
    proj_from_affine(C, a, p);
    if (jac && aff_is_inf(C, a)) cpy(f, PX(p), f->one);
  }
}

/* ---------------- binary MSM (bn128_G1_proj.c:684-698) ----------------
 * sum of the affine points whose coefficient byte is non-zero, by mixed additions into a
 * projective accumulator (the reference's loop); the multi form (:716-782) returns the same
 * group element per row, so it is restated as one such sum per row.  Outputs are normalised
 * projective points (X/Z, Y/Z, 1), infinity (0, 1, 0). */
void zko_msm_binary_proj(int cv, int npoints, const uint8_t *expos, const uint64_t *grps, uint64_t *tgt) {
  const curve_t *C = &CURVES[cv]; int n = C->fp->nl;
  uint64_t acc[3 * MAXL];
  proj_set_inf(C, acc);
  for (int i = 0; i < npoints; i++)
    if (expos[i]) proj_madd(C, acc, grps + (size_t)i * 2 * n, acc);
  zko_g1_proj_normalize(cv, acc, tgt);
}
void zko_multi_msm_binary_proj(int cv, int nrows, int npoints, const uint8_t *const *rows, const uint64_t *grps,
                               uint64_t *tgt) {
  int n = CURVES[cv].fp->nl;
  for (int r = 0; r < nrows; r++) zko_msm_binary_proj(cv, npoints, rows[r], grps, tgt + (size_t)r * 3 * n);
}

/* ---------------- polynomials of a KZG opening (bn128_poly_mont.c) ---------------- */
static int poly_degree(int cv, int n1, const uint64_t *a) {
  const field_t *f = F(cv, ZKO_FR);
  for (int i = n1 - 1; i >= 0; i--)
    if (!is_zero(f, a + 4 * (size_t)i)) return i;
  return -1;
}
int zko_poly_degree(int cv, int n1, const uint64_t *a) { return poly_degree(cv, n1, a); }

/* bn128_poly_mont.c:225-243: sum a_i x^i with a running power of x */
void zko_poly_eval(int cv, int n1, const uint64_t *a, const uint64_t *x, uint64_t *tgt) {
  const field_t *f = F(cv, ZKO_FR);
  uint64_t run[MAXL], t[MAXL];
  set0(f, tgt);
  cpy(f, run, f->one);
  for (int i = 0; i < n1; i++) {
    f_mul(f, a + 4 * (size_t)i, run, t);
    f_add(f, tgt, t, tgt);
    f_mul(f, run, x, run);
  }
}

/* bn128_poly_mont.c:317-395: q_j = a_{j+n} + eta q_{j+n} from j = D-n down to 0 (q_j = a_{j+n}
 * at the top), r_j = a_j + eta q_j for j <= D-n and a_j above; D < n gives q = 0, r = p.
 * Unused tails are zero.  Returns -1 where the reference's asserts would fail. */
int zko_poly_div_by_vanishing(int cv, int n1, const uint64_t *a, int n, const uint64_t *eta, int nquot,
                              uint64_t *quot, int nrem, uint64_t *rem) {
  const field_t *f = F(cv, ZKO_FR);
  const int D = poly_degree(cv, n1, a);
  if (!quot || nquot < D - n + 1 || (rem && nrem < n) || n < 1) return -1;
  uint64_t t[MAXL];
  if (D < n) {
    for (int j = 0; j < nquot; j++) set0(f, quot + 4 * (size_t)j);
    if (rem) {
      if (nrem < D + 1) return -1;
      for (int j = 0; j < nrem; j++) set0(f, rem + 4 * (size_t)j);
      memcpy(rem, a, 32 * (size_t)(D + 1));
    }
    return 0;
  }
  for (int j = D - n + 1; j < nquot; j++) set0(f, quot + 4 * (size_t)j);
  if (rem)
    for (int j = n; j < nrem; j++) set0(f, rem + 4 * (size_t)j);
  for (int j = D - n; j >= 0; j--) {
    if (j + n <= D - n) {
      f_mul(f, quot + 4 * (size_t)(j + n), eta, t);
      f_add(f, a + 4 * (size_t)(j + n), t, quot + 4 * (size_t)j);
    } else {
      cpy(f, quot + 4 * (size_t)j, a + 4 * (size_t)(j + n));
    }
  }
  if (rem)
    for (int j = 0; j < n; j++) {
      if (j <= D - n) {
        f_mul(f, quot + 4 * (size_t)j, eta, t);
        f_add(f, a + 4 * (size_t)j, t, rem + 4 * (size_t)j);
      } else {
        cpy(f, rem + 4 * (size_t)j, a + 4 * (size_t)j);
      }
    }
  return 0;
}

/* bn128_poly_mont.c:400-413: 1 when the remainder is zero, 0 otherwise, -1 on bad sizes */
int zko_poly_quot_by_vanishing(int cv, int n1, const uint64_t *a, int n, const uint64_t *eta, int nquot,
                               uint64_t *quot) {
  if (n < 1) return -1;
  uint64_t *rem = (uint64_t *)malloc(32 * (size_t)n);
  if (!rem) abort();
  int st = zko_poly_div_by_vanishing(cv, n1, a, n, eta, nquot, quot, n, rem);
  int ok = 1;
  for (int j = 0; st == 0 && j < 4 * n; j++)
    if (rem[j]) { ok = 0; break; }
  free(rem);
  return st ? st : ok;
}
