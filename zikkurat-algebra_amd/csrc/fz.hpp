// Reduced-radix Montgomery arithmetic for the hot kernels (MSM buckets, NTT butterflies).
//
// Why: on gfx950 v_mad_u64_u32 / v_mad_i64_i32, v_mul_lo_u32 and 64-bit adds all issue at
// full VALU rate, so a field multiplication costs its instruction count, not its multiply
// count.  With 32-bit limbs every partial product needs carry-flag adds and zero-extension
// moves (~420 instructions per 256-bit multiplication).  Here an element is N signed 32-bit
// limbs in radix 2^BITS (BITS = 29 for the 254/255-bit fields, N = 9; BITS = 28 for the
// 381-bit BLS12-381 base field, N = 14).  Partial products then leave room in a 64-bit
// accumulator, so a Montgomery product is a product-scanning loop of plain multiply-adds
// (v_mad_i64_i32) with one shift per column and no carry chain (~250 instructions), and
// additions/subtractions are limb-wise.
//
// Representation.  value(a) = sum_i a.l[i] 2^(BITS i).  "Normalised" means a.l[i] in
// [0, 2^BITS) for i < N-1 and a small signed top limb; every zmul output is normalised and
// values stay in (-8p, 8p).  The Montgomery radix is R' = 2^(BITS N) (2^261, 2^392), chosen so
// that R' > 64 p: for |a|, |b| < 8p the product satisfies |zmul(a,b)| < 64 p^2 / R' + p < 2p.
// Column sums stay below 2^63 when both inputs are normalised, or one is normalised and the
// other is the limb-wise sum/difference of two normalised values.
//
// Relation to the reference's Montgomery form (R = 2^256 or 2^384): an external element
// x R is converted with one product by K_IN = R'^2/R mod p (giving x R') and back with
// K_OUT = R mod p.  Products of external data with internal constants (NTT twiddles w R')
// stay external: zmul(x R, w R') = x w R.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define ZK_HD __host__ __device__ __forceinline__

namespace zk {

// parameter packs: limbs of p, -p^-1 mod 2^BITS, conversion constants, R' mod p and R'^3 mod p
// (values generated from the primes of lib/cbits/curves/fields/mont/*_mont.c)
struct ZBn254Fp {
  static constexpr int BITS = 29;
  static constexpr int N = 9;
  static constexpr int PBITS = 254;  // bit length of p
  static constexpr int EXT_WORDS = 8;
  static constexpr uint32_t NP = 75916169u;
  static constexpr int32_t P[N] = {410844487, 17064118, 477274959, 47522512, 361093496, 47923392, 10936641, 240920116, 3171406};
  static constexpr int32_t K_IN[N] = {322215073, 442336424, 171859116, 268585440, 314135016, 244503300, 348886451, 68918589, 360451};
  static constexpr int32_t K_OUT[N] = {93261213, 451550318, 297979764, 299258347, 342016167, 297253948, 482187706, 406012155, 920183};
  static constexpr int32_t ONE[N] = {360500257, 337389400, 408039635, 21759001, 178483129, 490881230, 299191303, 86689704, 903222};
  static constexpr int32_t R3[N] = {237179570, 381705378, 197673865, 484389131, 296595421, 53368462, 493544588, 282381599, 2540242};  // R'^3 mod p
};
struct ZBn254Fr {
  static constexpr int BITS = 29;
  static constexpr int N = 9;
  static constexpr int PBITS = 254;  // bit length of p
  static constexpr int EXT_WORDS = 8;
  static constexpr uint32_t NP = 268435455u;
  static constexpr int32_t P[N] = {268435457, 521120927, 240919632, 131109107, 361091715, 47923392, 10936641, 240920116, 3171406};
  static constexpr int32_t K_IN[N] = {268430039, 492061940, 71535269, 62181526, 323781850, 244503300, 348886451, 68918589, 360451};
  static constexpr int32_t K_OUT[N] = {268435451, 78749922, 406014571, 418196286, 342025071, 297253948, 482187706, 406012155, 920183};
  static constexpr int32_t ONE[N] = {268435287, 514263732, 86771339, 391139145, 178784091, 490881230, 299191303, 86689704, 903222};
  static constexpr int32_t R3[N] = {2088370, 399706979, 438698222, 156269639, 345420443, 184590952, 364266987, 530565949, 1813142};  // R'^3 mod p
};
struct ZBls381Fr {
  static constexpr int BITS = 29;
  static constexpr int N = 9;
  static constexpr int PBITS = 255;  // bit length of p
  static constexpr int EXT_WORDS = 8;
  static constexpr uint32_t NP = 536870911u;
  static constexpr int32_t P[N] = {1, 536870904, 529989567, 457704959, 494949691, 201589968, 354471143, 174404527, 7597479};
  static constexpr int32_t K_IN[N] = {536868651, 18087, 526333248, 216203804, 294598241, 8484555, 89016444, 271483280, 1968430};
  static constexpr int32_t K_OUT[N] = {536870910, 15, 13762688, 158331904, 83842440, 133690974, 364799537, 188061856, 1582257};
  static constexpr int32_t ONE[N] = {536870842, 559, 481694080, 172907520, 250130850, 384216799, 419952827, 139714039, 5047359};
  static constexpr int32_t R3[N] = {433522269, 2153349, 370290243, 249241592, 265969956, 85437554, 317105442, 433409754, 619787};  // R'^3 mod p
};
struct ZBls381Fp {
  static constexpr int BITS = 28;
  static constexpr int N = 14;
  static constexpr int PBITS = 381;  // bit length of p
  static constexpr int EXT_WORDS = 12;
  static constexpr uint32_t NP = 268238845u;
  static constexpr int32_t P[N] = {268413611, 267386879, 67108793, 268430101, 103030443, 168782607, 258420946, 255349035, 74926980, 79351158, 195540547, 241804465, 32127359, 106513};
  static constexpr int32_t K_IN[N] = {135160473, 55574580, 246491222, 233514649, 211322480, 82798231, 118064104, 172912617, 65577043, 248456990, 203477795, 116089733, 75436678, 87770};
  static constexpr int32_t K_OUT[N] = {196605, 9437184, 201327222, 48192, 146467828, 91569269, 90130584, 117777780, 130963539, 91145943, 119183266, 239678912, 247724672, 89957};
  static constexpr int32_t ONE[N] = {55049400, 226492416, 176409, 13493970, 208802528, 138027152, 3630751, 228652862, 162569026, 19495856, 85318031, 1231484, 106560762, 22394};
  static constexpr int32_t R3[N] = {33011856, 43306061, 166965026, 190555500, 213237964, 202217308, 210348229, 113862836, 36074218, 152955081, 55988902, 139032520, 69516217, 82347};  // R'^3 mod p
};

template <class P>
struct Fz {
  int32_t l[P::N];
};

template <class P>
ZK_HD void zset_zero(Fz<P> &r) {
#pragma unroll
  for (int i = 0; i < P::N; i++) r.l[i] = 0;
}
template <class P>
ZK_HD void zset_one(Fz<P> &r) {
#pragma unroll
  for (int i = 0; i < P::N; i++) r.l[i] = P::ONE[i];
}
template <class P>
ZK_HD void zset_const(Fz<P> &r, const int32_t (&c)[P::N]) {
#pragma unroll
  for (int i = 0; i < P::N; i++) r.l[i] = c[i];
}
// all limbs zero (the explicit infinity marker of XYZZ points; not a test modulo p)
template <class P>
ZK_HD bool zis_raw_zero(const Fz<P> &a) {
  int32_t acc = 0;
#pragma unroll
  for (int i = 0; i < P::N; i++) acc |= a.l[i];
  return acc == 0;
}

template <class P>
ZK_HD void zadd(Fz<P> &r, const Fz<P> &a, const Fz<P> &b) {
#pragma unroll
  for (int i = 0; i < P::N; i++) r.l[i] = a.l[i] + b.l[i];
}
template <class P>
ZK_HD void zsub(Fz<P> &r, const Fz<P> &a, const Fz<P> &b) {
#pragma unroll
  for (int i = 0; i < P::N; i++) r.l[i] = a.l[i] - b.l[i];
}
template <class P>
ZK_HD void zneg(Fz<P> &r, const Fz<P> &a) {
#pragma unroll
  for (int i = 0; i < P::N; i++) r.l[i] = -a.l[i];
}

// carry propagation: limbs 0..N-2 into [0, 2^BITS), signed top limb (value unchanged)
template <class P>
ZK_HD void znorm(Fz<P> &r) {
  constexpr int32_t MASK = (1 << P::BITS) - 1;
#pragma unroll
  for (int i = 0; i < P::N - 1; i++) {
    const int32_t c = r.l[i] >> P::BITS;  // arithmetic shift: floor division
    r.l[i] &= MASK;
    r.l[i + 1] += c;
  }
}

// Montgomery product a b / R' (product scanning, two interleaved 64-bit accumulators per
// column so that consecutive multiply-adds are independent).  Output normalised.
template <class P>
ZK_HD void zmul(Fz<P> &r, const Fz<P> &a, const Fz<P> &b) {
  constexpr int N = P::N, B = P::BITS;
  constexpr uint32_t MASK = (1u << B) - 1;
  uint32_t m[N];
  int64_t carry = 0;
#pragma unroll
  for (int c = 0; c < 2 * N - 1; c++) {
    const int lo = c < N ? 0 : c - N + 1;
    const int hi = c < N ? c : N - 1;
    int64_t s0 = carry, s1 = 0;
#pragma unroll
    for (int i = lo; i <= hi; i++) {
      const int64_t pr = (int64_t)a.l[i] * (int64_t)b.l[c - i];
      if ((i - lo) & 1) s1 += pr;
      else s0 += pr;
    }
#pragma unroll
    for (int i = lo; i <= hi; i++) {
      if (i >= c) continue;  // m_c is computed below
      const int64_t pr = (int64_t)((uint64_t)m[i] * (uint64_t)(uint32_t)P::P[c - i]);
      if ((i - lo) & 1) s0 += pr;
      else s1 += pr;
    }
    int64_t s = s0 + s1;
    if (c < N) {
      const uint32_t mc = ((uint32_t)s * P::NP) & MASK;
      m[c] = mc;
      s += (int64_t)((uint64_t)mc * (uint64_t)(uint32_t)P::P[0]);
    } else {
      r.l[c - N] = (int32_t)((uint32_t)s & MASK);
    }
    carry = s >> B;
  }
  r.l[N - 1] = (int32_t)carry;
}

// (a b - c d) / R' with one reduction for four normalised operands: column sums hold 2N
// products plus the N reduction terms, below 27 * 2^58 < 2^63; for |a|, |b|, |c|, |d| < 8p the
// result lies in (-2p, 2p) for every field here (R' > 128 p).  Output normalised.
template <class P>
ZK_HD void zmul_sub(Fz<P> &r, const Fz<P> &a, const Fz<P> &b, const Fz<P> &c, const Fz<P> &d) {
  constexpr int N = P::N, B = P::BITS;
  constexpr uint32_t MASK = (1u << B) - 1;
  uint32_t m[N];
  int64_t carry = 0;
#pragma unroll
  for (int col = 0; col < 2 * N - 1; col++) {
    const int lo = col < N ? 0 : col - N + 1;
    const int hi = col < N ? col : N - 1;
    int64_t s0 = carry, s1 = 0;
#pragma unroll
    for (int i = lo; i <= hi; i++) {
      s0 += (int64_t)a.l[i] * (int64_t)b.l[col - i];
      s1 -= (int64_t)c.l[i] * (int64_t)d.l[col - i];
    }
#pragma unroll
    for (int i = lo; i <= hi; i++) {
      if (i >= col) continue;
      const int64_t pr = (int64_t)((uint64_t)m[i] * (uint64_t)(uint32_t)P::P[col - i]);
      if ((i - lo) & 1) s0 += pr;
      else s1 += pr;
    }
    int64_t s = s0 + s1;
    if (col < N) {
      const uint32_t mc = ((uint32_t)s * P::NP) & MASK;
      m[col] = mc;
      s += (int64_t)((uint64_t)mc * (uint64_t)(uint32_t)P::P[0]);
    } else {
      r.l[col - N] = (int32_t)((uint32_t)s & MASK);
    }
    carry = s >> B;
  }
  r.l[N - 1] = (int32_t)carry;
}

// a^2 / R' for a normalised a: each off-diagonal product a_i a_j (i < j) once, against the
// doubled limb 2 a_i (|2 a_i| < 2^30), so a column holds about half of zmul's multiply-adds
// with the same sum bound; the reduction half is zmul's.
template <class P>
ZK_HD void zsqr(Fz<P> &r, const Fz<P> &a) {
  constexpr int N = P::N, B = P::BITS;
  constexpr uint32_t MASK = (1u << B) - 1;
  int32_t d[N];
#pragma unroll
  for (int i = 0; i < N; i++) d[i] = a.l[i] * 2;
  uint32_t m[N];
  int64_t carry = 0;
#pragma unroll
  for (int c = 0; c < 2 * N - 1; c++) {
    const int lo = c < N ? 0 : c - N + 1;
    const int hi = c < N ? c : N - 1;
    int64_t s0 = carry, s1 = 0;
    int k = 0;
#pragma unroll
    for (int i = lo; i <= hi; i++) {
      const int j = c - i;
      if (i > j) continue;
      const int64_t pr = i == j ? (int64_t)a.l[i] * (int64_t)a.l[i] : (int64_t)d[i] * (int64_t)a.l[j];
      if (k & 1) s1 += pr;
      else s0 += pr;
      k++;
    }
#pragma unroll
    for (int i = lo; i <= hi; i++) {
      if (i >= c) continue;
      const int64_t pr = (int64_t)((uint64_t)m[i] * (uint64_t)(uint32_t)P::P[c - i]);
      if ((i - lo) & 1) s0 += pr;
      else s1 += pr;
    }
    int64_t s = s0 + s1;
    if (c < N) {
      const uint32_t mc = ((uint32_t)s * P::NP) & MASK;
      m[c] = mc;
      s += (int64_t)((uint64_t)mc * (uint64_t)(uint32_t)P::P[0]);
    } else {
      r.l[c - N] = (int32_t)((uint32_t)s & MASK);
    }
    carry = s >> B;
  }
  r.l[N - 1] = (int32_t)carry;
}

// a^-1 (internal form) by Fermat's little theorem: a^(p-2) in the Montgomery domain, with a
// left-to-right sliding window of up to 4 bits over p - 2 (radix-2^BITS limbs, the 2 taken
// with a borrow).  The odd powers a, a^3, ..., a^15 cost one squaring and 7 products; then
// every exponent bit costs a squaring and every window a product: about 254 + 50 + 8
// products for a 254-bit p, against 254 + 127 for square-and-multiply.  The window value is
// the same in every lane, and the switch keeps the table in registers.  Output normalised,
// < 2p; a == 0 gives 0.  (zinv below is the divstep form; this one stays as its cross-check.)
template <class P>
ZK_HD void zinv_fermat(Fz<P> &r, const Fz<P> &a) {
  constexpr int N = P::N, B = P::BITS;
  uint32_t e[N];  // p - 2
  int32_t borrow = 2;
#pragma unroll
  for (int i = 0; i < N; i++) {
    int32_t v = P::P[i] - borrow;
    borrow = 0;
    if (v < 0) {
      v += 1 << B;
      borrow = 1;
    }
    e[i] = (uint32_t)v;
  }
  auto bit = [&](int i) -> uint32_t { return (e[i / B] >> (i % B)) & 1u; };
  Fz<P> t0, t1, t2, t3, t4, t5, t6, t7, sq, acc;
  t0 = a;
  zsqr(sq, a);
  zmul(t1, t0, sq);
  zmul(t2, t1, sq);
  zmul(t3, t2, sq);
  zmul(t4, t3, sq);
  zmul(t5, t4, sq);
  zmul(t6, t5, sq);
  zmul(t7, t6, sq);
  int i = N * B - 1;
  while (i >= 0 && !bit(i)) i--;
  bool started = false;
  while (i >= 0) {
    if (!bit(i)) {
      zsqr(sq, acc);
      acc = sq;
      i--;
      continue;
    }
    int j = i - 3 < 0 ? 0 : i - 3;
    while (!bit(j)) j++;  // the window [j, i] ends in a set bit
    uint32_t v = 0;
    for (int k = i; k >= j; k--) v = (v << 1) | bit(k);
    if (started) {
      for (int k = i; k >= j; k--) {
        zsqr(sq, acc);
        acc = sq;
      }
      switch (v >> 1) {
        case 0: zmul(sq, acc, t0); break;
        case 1: zmul(sq, acc, t1); break;
        case 2: zmul(sq, acc, t2); break;
        case 3: zmul(sq, acc, t3); break;
        case 4: zmul(sq, acc, t4); break;
        case 5: zmul(sq, acc, t5); break;
        case 6: zmul(sq, acc, t6); break;
        default: zmul(sq, acc, t7); break;
      }
      acc = sq;
    } else {
      switch (v >> 1) {
        case 0: acc = t0; break;
        case 1: acc = t1; break;
        case 2: acc = t2; break;
        case 3: acc = t3; break;
        case 4: acc = t4; break;
        case 5: acc = t5; break;
        case 6: acc = t6; break;
        default: acc = t7; break;
      }
      started = true;
    }
    i = j - 1;
  }
  r = acc;
}

// a^-1 (internal form) by batched division steps (Bernstein and Yang, "Fast constant-time gcd
// computation and modular inversion", CHES 2019).  On f = p (odd) and g = A, the canonical
// value of a (A = x R'), a division step is
//   delta > 0 and g odd: (delta, f, g) <- (1 - delta, g, (g - f) / 2)
//   g odd:               (delta, f, g) <- (1 + delta, f, (g + f) / 2)
//   g even:              (delta, f, g) <- (1 + delta, f, g / 2)
// starting from delta = 1; for f^2 + 4 g^2 <= 5 2^(2 b) (b = PBITS) g reaches 0 within
// floor((49 b + 57) / 17) steps (the paper's Theorem 11.2), and then f = +-gcd = +-1.  d and e
// track f = d A, g = e A (mod p) from d = 0, e = 1, so A^-1 = f d.
// Batches of BITS steps: the steps depend only on delta and the parities of g, so they run on
// the low 32 bits of f and g (one bit of validity is used up per step) and yield an integer
// matrix T with |T00| + |T01|, |T10| + |T11| <= 2^BITS such that 2^BITS (f', g') = T (f, g).
// The batch then applies T to the full limbs (the low limb of the result is zero: the division
// by 2^BITS drops it) and to (d, e) modulo p, with k p added, k = -(T d)_0 / p mod 2^BITS, so
// that the division is exact (the Montgomery step).  |d|, |e| grow by at most p per batch; the
// last product by R'^3 reduces them: zmul(f d, R'^3) = A^-1 R'^2 = x^-1 R'.  A batch is about
// BITS x 20 32-bit operations plus 4 N multiply-adds per pair, roughly 20 batches for a 254-bit
// p against ~310 products for zinv_fermat: about a fifth of the instructions.  Every lane runs
// until its own g is zero (at most the theorem's bound).  Output normalised, < 2p; a == 0
// gives 0.
template <class P>
ZK_HD void zinv(Fz<P> &r, const Fz<P> &a) {
  constexpr int N = P::N, B = P::BITS;
  constexpr int32_t MASK = (1 << B) - 1;
  constexpr int kMaxBatches = ((49 * P::PBITS + 57) / 17 + B - 1) / B;
  Fz<P> f, g, d, e, t;
  zreduce_small(t, a);
  zcanon(g, t);
  zset_const(f, P::P);
  zset_zero(d);
  zset_zero(e);
  e.l[0] = 1;
  int32_t delta = 1;
  for (int batch = 0; batch < kMaxBatches && !zis_raw_zero(g); batch++) {
    // B steps on the low 32 bits of f and g
    // (unsigned words: after j steps only their low 32 - j bits are still exact, and a step
    // reads bit 0 only)
    uint32_t fl = (uint32_t)f.l[0] | ((uint32_t)f.l[1] << B);
    uint32_t gl = (uint32_t)g.l[0] | ((uint32_t)g.l[1] << B);
    int32_t u = 1, v = 0, q = 0, w = 1;
#pragma unroll 4
    for (int j = 0; j < B; j++) {
      const uint32_t odd = 0u - (gl & 1u);                    // all ones when g is odd
      const uint32_t sw = odd & (0u - (uint32_t)(delta > 0));  // all ones: swap to (g, -f)
      const int32_t ss = (int32_t)sw, so = (int32_t)odd;
      const uint32_t f0 = fl;
      const int32_t u0 = u, v0 = v;
      fl ^= (fl ^ gl) & sw;
      gl ^= (gl ^ (0u - f0)) & sw;
      u ^= (u ^ q) & ss;
      v ^= (v ^ w) & ss;
      q ^= (q ^ -u0) & ss;
      w ^= (w ^ -v0) & ss;
      delta = (delta ^ ss) - ss;  // -delta on a swap
      gl += fl & odd;
      q += u & so;
      w += v & so;
      gl >>= 1;  // exact: g + f is even
      u *= 2;
      v *= 2;
      delta += 1;
    }
    // (f, g) <- T (f, g) / 2^B and (d, e) <- (T (d, e) + (kd, ke) p) / 2^B, column by column
    int64_t cf = (int64_t)u * f.l[0] + (int64_t)v * g.l[0];
    int64_t cg = (int64_t)q * f.l[0] + (int64_t)w * g.l[0];
    int64_t cd = (int64_t)u * d.l[0] + (int64_t)v * e.l[0];
    int64_t ce = (int64_t)q * d.l[0] + (int64_t)w * e.l[0];
    const int64_t kd = (int64_t)(((uint32_t)cd * P::NP) & (uint32_t)MASK);
    const int64_t ke = (int64_t)(((uint32_t)ce * P::NP) & (uint32_t)MASK);
    cd += kd * P::P[0];
    ce += ke * P::P[0];
    cf >>= B;
    cg >>= B;
    cd >>= B;
    ce >>= B;
#pragma unroll
    for (int i = 1; i < N; i++) {
      cf += (int64_t)u * f.l[i] + (int64_t)v * g.l[i];
      cg += (int64_t)q * f.l[i] + (int64_t)w * g.l[i];
      cd += (int64_t)u * d.l[i] + (int64_t)v * e.l[i] + kd * P::P[i];
      ce += (int64_t)q * d.l[i] + (int64_t)w * e.l[i] + ke * P::P[i];
      f.l[i - 1] = (int32_t)(cf & MASK);
      g.l[i - 1] = (int32_t)(cg & MASK);
      d.l[i - 1] = (int32_t)(cd & MASK);
      e.l[i - 1] = (int32_t)(ce & MASK);
      cf >>= B;
      cg >>= B;
      cd >>= B;
      ce >>= B;
    }
    f.l[N - 1] = (int32_t)cf;
    g.l[N - 1] = (int32_t)cg;
    d.l[N - 1] = (int32_t)cd;
    e.l[N - 1] = (int32_t)ce;
  }
  if (f.l[N - 1] < 0) {  // f = -1
    zneg(d, d);
    znorm(d);
  }
  Fz<P> k3;
  zset_const(k3, P::R3);
  zmul(r, d, k3);
}

// a == 0 (mod p) for a normalised a with |a| < 8p, i.e. a = j p for some j in [-8, 8].  Then
// the low limb is j p mod 2^BITS, so (low * NP) mod 2^BITS = -j mod 2^BITS (NP = -p^-1): one
// product filters, and only a hit (|j| <= 8) runs the exact comparison with that j p.
template <class P>
ZK_HD bool zis_zero_mod(const Fz<P> &a) {
  constexpr int N = P::N, B = P::BITS;
  constexpr int64_t MASK = (1ll << B) - 1;
  const uint32_t t = ((uint32_t)a.l[0] * P::NP + 8u) & (uint32_t)MASK;  // 8 - j
  if (t > 16u) return false;
  const int64_t j = 8 - (int64_t)t;
  int64_t c = 0;
#pragma unroll
  for (int i = 0; i < N; i++) {
    const int64_t d = (int64_t)a.l[i] - j * P::P[i] + c;
    if (i < N - 1) {
      if (d & MASK) return false;
      c = d >> B;
    } else if (d != 0) {
      return false;
    }
  }
  return true;
}

// a - q p for q = round-down(a / p) estimated from the top two limbs in double precision: for a
// normalised a with |a| < 64 p the estimate is off by less than 1e-6 of p, so the result lies in
// (-p, 2p), ready for zcanon (about 40 instructions against a product by one's 250).  The
// column sums a_i - q P_i + carry stay far inside 64 bits.  Output normalised.
template <class P>
ZK_HD void zreduce_small(Fz<P> &r, const Fz<P> &a) {
  constexpr int N = P::N, B = P::BITS;
  constexpr int64_t MASK = (1ll << B) - 1;
  // p / 2^(B (N - 2)) and a's top two limbs at the same scale (the lower limbs are >= 0)
  constexpr double PTOP = (double)P::P[N - 1] * (double)(1 << B) + (double)P::P[N - 2] +
                          (double)P::P[N - 3] / (double)(1 << B) + 0.5 / (double)(1 << B);
  constexpr double INV = 1.0 / PTOP;
  const double top = (double)a.l[N - 1] * (double)(1 << B) + (double)a.l[N - 2];
  const int32_t q = (int32_t)__builtin_floor(top * INV);
  int64_t c = 0;
#pragma unroll
  for (int i = 0; i < N - 1; i++) {
    const int64_t t = (int64_t)a.l[i] - (int64_t)q * P::P[i] + c;
    r.l[i] = (int32_t)(t & MASK);
    c = t >> B;
  }
  r.l[N - 1] = (int32_t)((int64_t)a.l[N - 1] - (int64_t)q * P::P[N - 1] + c);
}

// canonical representative in [0, p) of a normalised a with -p < a < 2p
template <class P>
ZK_HD void zcanon(Fz<P> &r, const Fz<P> &a) {
  constexpr int N = P::N;
  Fz<P> t = a;
  if (t.l[N - 1] < 0) {
#pragma unroll
    for (int i = 0; i < N; i++) t.l[i] += P::P[i];
    znorm(t);
  }
  Fz<P> d;
#pragma unroll
  for (int i = 0; i < N; i++) d.l[i] = t.l[i] - P::P[i];
  znorm(d);
  r = d.l[N - 1] >= 0 ? d : t;
}

// ---- external layout: EXT_WORDS little-endian u32 words (the reference's u64 limbs)

// value-preserving unpack of a canonical external word string into normalised limbs
template <class P>
ZK_HD void zunpack(Fz<P> &r, const uint32_t *w) {
  constexpr int B = P::BITS;
  constexpr uint32_t MASK = (1u << B) - 1;
#pragma unroll
  for (int i = 0; i < P::N; i++) {
    const int bit = B * i, k = bit >> 5, o = bit & 31;
    uint64_t v = w[k];
    if (k + 1 < P::EXT_WORDS) v |= (uint64_t)w[k + 1] << 32;
    r.l[i] = (int32_t)((uint32_t)(v >> o) & MASK);
  }
}

// pack a canonical value (limbs in [0, 2^BITS), value < 2^(32 EXT_WORDS)) into words
template <class P>
ZK_HD void zpack(uint32_t *w, const Fz<P> &a) {
  constexpr int B = P::BITS;
#pragma unroll
  for (int k = 0; k < P::EXT_WORDS; k++) {
    const int bit = 32 * k;
    const int i = bit / B, o = bit % B;
    uint64_t v = (uint32_t)a.l[i];
    if (i + 1 < P::N) v |= (uint64_t)(uint32_t)a.l[i + 1] << B;
    if (i + 2 < P::N) v |= (uint64_t)(uint32_t)a.l[i + 2] << (2 * B);
    w[k] = (uint32_t)(v >> o);
  }
}

// external Montgomery (x R) -> internal (x R'), and back to a canonical external element
template <class P>
ZK_HD void zfrom_ext(Fz<P> &r, const uint32_t *w) {
  Fz<P> a, k;
  zunpack(a, w);
  zset_const(k, P::K_IN);
  zmul(r, a, k);
}
template <class P>
ZK_HD void zto_ext(uint32_t *w, const Fz<P> &a) {
  Fz<P> k, t;
  zset_const(k, P::K_OUT);
  zmul(t, a, k);
  zcanon(t, t);
  zpack(w, t);
}

// ---- device memory: N words per element

template <class P>
__device__ __forceinline__ void zload(Fz<P> &r, const int32_t *__restrict__ p) {
#pragma unroll
  for (int i = 0; i < P::N; i++) r.l[i] = p[i];
}
template <class P>
__device__ __forceinline__ void zstore(int32_t *__restrict__ p, const Fz<P> &a) {
#pragma unroll
  for (int i = 0; i < P::N; i++) p[i] = a.l[i];
}

}  // namespace zk
