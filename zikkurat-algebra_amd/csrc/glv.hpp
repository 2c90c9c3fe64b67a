// GLV endomorphism of BN254's G1: constants and the scalar split, shared by the MSM digit
// pass and the G1 FFT.
#pragma once
#include <stdint.h>

#include "curvez.hpp"

namespace zk {

// GLV split of BN254 scalars (Gallant, Lambert, Vanstone, CRYPTO 2001).  BN254's G1 has prime
// order r and the endomorphism phi(x, y) = (beta x, y) = lambda P, with beta and lambda
// primitive cube roots of unity in Fp and Fr.  A scalar k < r is split as
// k = k1 + lambda k2 (mod r): c1 = floor(k G1 / 2^256), c2 = floor(k G2 / 2^256),
// k1 = k - c1 a1 - c2 a2, k2 = c1 |b1| - c2 b2, where (a1, b1), (a2, b2) is a short basis of
// {(x, y) : x + lambda y = 0 mod r} (extended Euclid on r and lambda; det = r) and
// G1 = round(2^256 b2 / r), G2 = round(-2^256 b1 / r).  The floor errors lie in
// (-1.125, 0.125), so |k1|, |k2| < 1.125 a1 < 2^127: both fit 128-bit two's complement, and
// signed-digit windows over 128 bits cover them.
struct Bn254Glv {
  static constexpr uint64_t G1[3] = {0x5398fd0300ff6565ull, 0x4ccef014a773d2d2ull, 0x2ull};
  static constexpr uint64_t G2[2] = {0xd91d232ec7e0b3d7ull, 0x2ull};
  static constexpr uint64_t A1[2] = {0x8211bbeb7d4f1128ull, 0x6f4d8248eeb859fcull};
  static constexpr uint64_t A2 = 0x89d3256894d213e3ull;  // = |b1|
  static constexpr uint64_t B2[2] = {0x0be4e1541221250bull, 0x6f4d8248eeb859fdull};
  // beta R' mod p (R' = 2^261), the internal form of beta: zmul(x R', beta R') = beta x R'
  static constexpr int32_t BETA[9] = {416069521, 391846970, 193189602, 244240497, 310558434,
                                      69340804,  535036287, 352865610, 1478937};
};

// (k G) >> 256 for k < 2^256 and a constant G of NG 64-bit limbs, low 128 bits of the quotient
template <int NG>
ZK_HD unsigned __int128 mul_shift256(const uint64_t k[4], const uint64_t (&g)[NG]) {
  using u128 = unsigned __int128;
  uint64_t prod[4 + NG] = {};
#pragma unroll
  for (int i = 0; i < 4; i++) {
    uint64_t carry = 0;
#pragma unroll
    for (int j = 0; j < NG; j++) {
      const u128 t = (u128)k[i] * g[j] + prod[i + j] + carry;
      prod[i + j] = (uint64_t)t;
      carry = (uint64_t)(t >> 64);
    }
    prod[i + NG] = carry;
  }
  return (u128)prod[4] | ((u128)prod[5] << 64);
}

// k < r (4 limbs) -> |k1|, |k2| as two 64-bit limbs each, with their signs
ZK_HD void glv_split(const uint64_t k[4], uint64_t m1[2], bool &neg1, uint64_t m2[2], bool &neg2) {
  using G = Bn254Glv;
  using u128 = unsigned __int128;
  const u128 c1 = mul_shift256<3>(k, G::G1);
  const u128 c2 = mul_shift256<2>(k, G::G2);  // < 2^64
  const u128 a1 = (u128)G::A1[0] | ((u128)G::A1[1] << 64);
  const u128 b2 = (u128)G::B2[0] | ((u128)G::B2[1] << 64);
  u128 k1 = ((u128)k[0] | ((u128)k[1] << 64)) - c1 * a1 - c2 * (u128)G::A2;  // mod 2^128
  u128 k2 = c1 * (u128)G::A2 - c2 * b2;
  neg1 = (uint64_t)(k1 >> 127) != 0;
  neg2 = (uint64_t)(k2 >> 127) != 0;
  if (neg1) k1 = (u128)0 - k1;
  if (neg2) k2 = (u128)0 - k2;
  m1[0] = (uint64_t)k1;
  m1[1] = (uint64_t)(k1 >> 64);
  m2[0] = (uint64_t)k2;
  m2[1] = (uint64_t)(k2 >> 64);
}

// k P for a BN254 point (XYZZ) and a standard-form scalar k < r (8 words): the GLV split and one
// joint double-and-add over at most 127 bits (Straus-Shamir) with P, phi(P) and their sum, the
// signs of k1, k2 folded into them; phi(X, Y, ZZ, ZZZ) = (beta X, Y, ZZ, ZZZ) since x = X / ZZ.
// About 127 doublings and 96 additions against 254 and 127 for double-and-add on k.
template <class C>
ZK_HD void zxyzz_scalar_mul_glv(XyzzZ<C> &r, const XyzzZ<C> &p, const uint32_t (&k)[8]) {
  uint64_t k64[4];
#pragma unroll
  for (int j = 0; j < 4; j++) k64[j] = (uint64_t)k[2 * j] | ((uint64_t)k[2 * j + 1] << 32);
  uint64_t m1[2], m2[2];
  bool n1, n2;
  glv_split(k64, m1, n1, m2, n2);
  XyzzZ<C> q0 = p, q1 = p, q01;
  Fz<ZBn254Fp> beta;
  zset_const(beta, Bn254Glv::BETA);
  zmul(q1.x, p.x, beta);
  if (n1) zxyzz_neg(q0, q0);
  if (n2) zxyzz_neg(q1, q1);
  zxyzz_add(q01, q0, q1);
  auto bit = [&](const uint64_t (&m)[2], int i) -> uint32_t { return (uint32_t)(m[i >> 6] >> (i & 63)) & 1u; };
  int top = 127;
  while (top >= 0 && !(bit(m1, top) | bit(m2, top))) top--;
  zxyzz_set_inf(r);
  for (int i = top; i >= 0; i--) {
    XyzzZ<C> t;
    zxyzz_dbl(t, r);
    const uint32_t b1 = bit(m1, i), b2 = bit(m2, i);
    if (b1 | b2) zxyzz_add(r, t, b1 ? (b2 ? q01 : q0) : q1);
    else r = t;
  }
}

constexpr int kGlvW2Entries = 15;

// The same product k P with the joint digits of k1, k2 taken two bits at a time, for callers
// that can give each thread kGlvW2Entries thread-major table entries in device memory (its
// block tt, kGlvW2Entries * ZLayout<C>::JACC words).  A step is two doublings and one addition
// of E(i, j) = i Q0 + j Q1 (i, j < 4, not both zero; Q0 = +-P, Q1 = +-phi(P) with the signs of
// k1, k2; entry 4 j + i - 1): at most 63 additions after the first entry.  In a wavefront the
// 1-bit form adds at almost every one of its ~127 steps (a step adds as soon as one lane's bit
// pair is nonzero), so this saves ~60 additions for 12 on the table (one doubling, two
// additions, three phi, nine sums).  phi(j P) = phi(P) j, so E(0, j) comes from E(j, 0) by phi
// and the sign; phi(X, Y, Z) = (beta X, Y, Z) in Jacobian coordinates too.  The accumulator is
// Jacobian and the entries are cached Jacobian points (curvez.hpp).
// With [wlo, whi) only the windows in that range are added (the doublings run down to window
// 0): the sum of the results over a split of [0, 64) is k P, so two threads can take the high
// and the low windows of one product.  The high part keeps all ~128 doublings and adds only
// its own windows; the low part starts at its top window, so it is 2 wlo-window doublings
// shorter: split at window 40, they are ~128 doublings + ~24 additions and ~80 + ~40.
template <class C>
ZK_HD void zxyzz_scalar_mul_glv_w2(XyzzZ<C> &r, const XyzzZ<C> &p, const uint32_t (&k)[8],
                                   uint32_t *__restrict__ tt, int wlo = 0, int whi = 64) {
  uint64_t k64[4];
#pragma unroll
  for (int j = 0; j < 4; j++) k64[j] = (uint64_t)k[2 * j] | ((uint64_t)k[2 * j + 1] << 32);
  uint64_t m1[2], m2[2];
  bool n1, n2;
  glv_split(k64, m1, n1, m2, n2);
  // E(i, 0) = i Q0, i = 1..3; the sums read them back from the table (fewer live registers)
  {
    JacZ<C> e1, e2, e3;
    JacCZ<C> c;
    zjac_from_xyzz(e1, p);
    zjac_cache(c, e1);
    if (n1) zjac_neg(c);
    zjac_uncache(e1, c);
    zjac_tab_store<C>(tt, 0, c);
    zjac_dbl(e2, e1);
    zjac_add(e3, e2, c);
    zjac_cache(c, e2);
    zjac_tab_store<C>(tt, 1, c);
    zjac_cache(c, e3);
    zjac_tab_store<C>(tt, 2, c);
  }
  Fz<ZBn254Fp> beta;
  zset_const(beta, Bn254Glv::BETA);
#pragma unroll 1
  for (int j = 1; j <= 3; j++) {
    // E(0, j) = phi(E(j, 0)), negated when the signs of k1 and k2 differ; E(i, j) = E(i, 0) + E(0, j)
    JacCZ<C> b;
    zjac_tab_load<C>(b, tt, j - 1);
    zmul(b.x, b.x, beta);
    if (n1 != n2) zjac_neg(b);
    zjac_tab_store<C>(tt, 4 * j - 1, b);
#pragma unroll 1
    for (int i = 1; i <= 3; i++) {
      JacCZ<C> a;
      JacZ<C> s;
      zjac_tab_load<C>(a, tt, i - 1);
      zjac_uncache(s, a);
      zjac_add(s, s, b);
      zjac_cache(a, s);
      zjac_tab_store<C>(tt, 4 * j + i - 1, a);
    }
  }
  auto digit = [&](int s) -> uint32_t {  // entry + 1 of window s (0: both digits zero)
    const int b = 2 * s;
    const uint32_t d1 = (uint32_t)(m1[b >> 6] >> (b & 63)) & 3u, d2 = (uint32_t)(m2[b >> 6] >> (b & 63)) & 3u;
    return 4 * d2 + d1;
  };
  int top = whi - 1;
  while (top >= wlo && digit(top) == 0) top--;
  if (top < wlo) {
    zxyzz_set_inf(r);
    return;
  }
  JacZ<C> acc;
  {
    JacCZ<C> e;
    zjac_tab_load<C>(e, tt, (int)digit(top) - 1);
    zjac_uncache(acc, e);
  }
  for (int s = top - 1; s >= 0; s--) {
    const uint32_t d = s >= wlo ? digit(s) : 0u;
    JacCZ<C> e;
    if (d) zjac_tab_load<C>(e, tt, (int)d - 1);  // in flight during the doublings
    zjac_dbl(acc, acc);
    zjac_dbl(acc, acc);
    if (d) zjac_add(acc, acc, e);
  }
  zjac_to_xyzz(r, acc);
}

}  // namespace zk
