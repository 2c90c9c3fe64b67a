// Host-side prime-field and G1 arithmetic of the library (64-bit limbs, unsigned __int128).
//
// The device does all data-parallel work.  The host side only handles O(1)-sized steps that
// are inherently serial: checking that an NTT generator has the right order, N^-1 and the
// seeds of twiddle tables, and the final combination of MSM window sums sum_k 2^(kc) S_k
// (about 255 doublings on one point, i.e. ~0.1 ms on one core against several ms on a single
// GPU lane), followed by the single inversion of to_affine.  Same Montgomery convention as
// the device layer and the reference (R = 2^(64 L), canonical outputs).
#pragma once
#include <stdint.h>
#include <string.h>

namespace zkh {

typedef unsigned __int128 u128;

struct FieldDesc {
  int nl;                 // 64-bit limbs
  uint64_t p[6];
  uint64_t np;            // -p^-1 mod 2^64
  uint64_t one[6];        // R mod p
  uint64_t r2[6];         // R^2 mod p
};

extern const FieldDesc BN254_FP, BN254_FR, BLS381_FP, BLS381_FR;

struct Fe {
  uint64_t w[6];
};

inline bool is_zero(const FieldDesc &f, const Fe &a) {
  uint64_t acc = 0;
  for (int i = 0; i < f.nl; i++) acc |= a.w[i];
  return acc == 0;
}
inline bool is_eq(const FieldDesc &f, const Fe &a, const Fe &b) {
  for (int i = 0; i < f.nl; i++)
    if (a.w[i] != b.w[i]) return false;
  return true;
}
inline Fe zero() {
  Fe r;
  memset(&r, 0, sizeof r);
  return r;
}
inline Fe one(const FieldDesc &f) {
  Fe r = zero();
  memcpy(r.w, f.one, 8 * f.nl);
  return r;
}
inline Fe load(const FieldDesc &f, const uint64_t *p) {
  Fe r = zero();
  memcpy(r.w, p, 8 * f.nl);
  return r;
}
inline void store(const FieldDesc &f, uint64_t *p, const Fe &a) { memcpy(p, a.w, 8 * f.nl); }

template <int N>
inline bool geq_p_n(const FieldDesc &f, const uint64_t *t) {
  for (int i = N - 1; i >= 0; i--) {
    if (t[i] > f.p[i]) return true;
    if (t[i] < f.p[i]) return false;
  }
  return true;
}
// r = t - p if t >= p (t < 2p, plus a carry word hi of 0 or 1), else t; no data-dependent branch
template <int N>
inline void reduce_once_n(const FieldDesc &f, Fe &r, const uint64_t *t, uint64_t hi) {
  uint64_t s[N], br = 0;
  for (int i = 0; i < N; i++) {
    const u128 d = (u128)t[i] - f.p[i] - br;
    s[i] = (uint64_t)d;
    br = (uint64_t)(d >> 64) & 1;
  }
  const uint64_t keep = 0 - (br & (hi ^ 1));  // all ones: t < p (the borrow is not covered by hi)
  for (int i = 0; i < N; i++) r.w[i] = (t[i] & keep) | (s[i] & ~keep);
}
template <int N>
inline Fe add_n(const FieldDesc &f, const Fe &a, const Fe &b) {
  Fe r = zero();
  uint64_t t[N];
  u128 c = 0;
  for (int i = 0; i < N; i++) {
    c += (u128)a.w[i] + b.w[i];
    t[i] = (uint64_t)c;
    c >>= 64;
  }
  reduce_once_n<N>(f, r, t, (uint64_t)c);
  return r;
}
template <int N>
inline Fe sub_n(const FieldDesc &f, const Fe &a, const Fe &b) {
  Fe r = zero();
  uint64_t br = 0;
  for (int i = 0; i < N; i++) {
    const u128 d = (u128)a.w[i] - b.w[i] - br;
    r.w[i] = (uint64_t)d;
    br = (uint64_t)(d >> 64) & 1;
  }
  const uint64_t mask = 0 - br;  // a < b: add p back
  u128 c = 0;
  for (int i = 0; i < N; i++) {
    c += (u128)r.w[i] + (f.p[i] & mask);
    r.w[i] = (uint64_t)c;
    c >>= 64;
  }
  return r;
}
// Montgomery product by CIOS, the multiplication and reduction of each outer step in one
// inner loop.  Every modulus here has a top word below 2^63 - 1, so the running value stays
// below 2p in N words without a carry word ("no-carry" CIOS); one branch-free subtraction
// ends it.
template <int N>
inline Fe mul_n(const FieldDesc &f, const Fe &a, const Fe &b) {
  uint64_t t[N] = {0};
  for (int i = 0; i < N; i++) {
    u128 A = (u128)a.w[i] * b.w[0] + t[0];
    const uint64_t t0 = (uint64_t)A;
    const uint64_t m = t0 * f.np;
    u128 C = ((u128)m * f.p[0] + t0) >> 64;
    A >>= 64;
    for (int j = 1; j < N; j++) {
      A += (u128)a.w[i] * b.w[j] + t[j];
      C += (u128)m * f.p[j] + (uint64_t)A;
      t[j - 1] = (uint64_t)C;
      A >>= 64;
      C >>= 64;
    }
    t[N - 1] = (uint64_t)C + (uint64_t)A;
  }
  Fe r = zero();
  reduce_once_n<N>(f, r, t, 0);
  return r;
}

inline bool geq_p(const FieldDesc &f, const uint64_t *t) { return f.nl == 4 ? geq_p_n<4>(f, t) : geq_p_n<6>(f, t); }
inline Fe add(const FieldDesc &f, const Fe &a, const Fe &b) { return f.nl == 4 ? add_n<4>(f, a, b) : add_n<6>(f, a, b); }
inline Fe sub(const FieldDesc &f, const Fe &a, const Fe &b) { return f.nl == 4 ? sub_n<4>(f, a, b) : sub_n<6>(f, a, b); }
inline Fe neg(const FieldDesc &f, const Fe &a) { return is_zero(f, a) ? a : sub(f, zero(), a); }
inline Fe mul(const FieldDesc &f, const Fe &a, const Fe &b) { return f.nl == 4 ? mul_n<4>(f, a, b) : mul_n<6>(f, a, b); }
inline Fe sqr(const FieldDesc &f, const Fe &a) { return mul(f, a, a); }
inline Fe from_std(const FieldDesc &f, const Fe &a) {
  Fe r2 = zero();
  memcpy(r2.w, f.r2, 8 * f.nl);
  return mul(f, a, r2);
}
inline Fe to_std(const FieldDesc &f, const Fe &a) {
  Fe u = zero();
  u.w[0] = 1;
  return mul(f, a, u);
}
inline Fe from_u64(const FieldDesc &f, uint64_t x) {
  Fe a = zero();
  a.w[0] = x;
  return from_std(f, a);
}
inline Fe pow_u64(const FieldDesc &f, const Fe &a, uint64_t e) {
  Fe acc = one(f), base = a;
  while (e) {
    if (e & 1) acc = mul(f, acc, base);
    base = sqr(f, base);
    e >>= 1;
  }
  return acc;
}
// Fermat inverse a^(p-2); 0 -> 0
inline Fe inv(const FieldDesc &f, const Fe &a) {
  uint64_t e[6];
  memcpy(e, f.p, sizeof e);
  e[0] -= 2;  // p is odd and > 2, no borrow
  Fe acc = one(f);
  for (int i = 64 * f.nl - 1; i >= 0; i--) {
    acc = sqr(f, acc);
    if ((e[i / 64] >> (i % 64)) & 1) acc = mul(f, acc, a);
  }
  return acc;
}

// ---------------------------------------------------------------------------------------
// G1 in XYZZ coordinates (x = X/ZZ, y = Y/ZZZ, ZZ^3 = ZZZ^2); infinity has ZZ = ZZZ = 0.
// Formulas: EFD g1p/auto-shortw-xyzz (add-2008-s, dbl-2008-s-1) for a = 0.

struct Xyzz {
  Fe x, y, zz, zzz;
};

inline Xyzz xyzz_inf() {
  Xyzz r;
  r.x = r.y = r.zz = r.zzz = zero();
  return r;
}
inline bool xyzz_is_inf(const FieldDesc &f, const Xyzz &p) { return is_zero(f, p.zz); }

inline Xyzz xyzz_dbl(const FieldDesc &f, const Xyzz &p) {
  if (xyzz_is_inf(f, p) || is_zero(f, p.y)) return xyzz_inf();
  Fe u = add(f, p.y, p.y);
  Fe v = sqr(f, u);
  Fe w = mul(f, u, v);
  Fe s = mul(f, p.x, v);
  Fe xx = sqr(f, p.x);
  Fe m = add(f, add(f, xx, xx), xx);
  Xyzz r;
  r.x = sub(f, sub(f, sqr(f, m), s), s);
  r.y = sub(f, mul(f, m, sub(f, s, r.x)), mul(f, w, p.y));
  r.zz = mul(f, v, p.zz);
  r.zzz = mul(f, w, p.zzz);
  return r;
}

inline Xyzz xyzz_add(const FieldDesc &f, const Xyzz &p, const Xyzz &q) {
  if (xyzz_is_inf(f, p)) return q;
  if (xyzz_is_inf(f, q)) return p;
  Fe u1 = mul(f, p.x, q.zz), u2 = mul(f, q.x, p.zz);
  Fe s1 = mul(f, p.y, q.zzz), s2 = mul(f, q.y, p.zzz);
  Fe P = sub(f, u2, u1), R = sub(f, s2, s1);
  if (is_zero(f, P)) {
    if (is_zero(f, R)) return xyzz_dbl(f, p);
    return xyzz_inf();
  }
  Fe pp = sqr(f, P), ppp = mul(f, P, pp), q1 = mul(f, u1, pp);
  Xyzz r;
  r.x = sub(f, sub(f, sub(f, sqr(f, R), ppp), q1), q1);
  r.y = sub(f, mul(f, R, sub(f, q1, r.x)), mul(f, s1, ppp));
  r.zz = mul(f, mul(f, p.zz, q.zz), pp);
  r.zzz = mul(f, mul(f, p.zzz, q.zzz), ppp);
  return r;
}

// affine (Montgomery) output; infinity -> all 0xFF bytes (lib/cbits/curves/g1/proj/bn128_G1_proj.c:132-137)
inline void xyzz_to_affine(const FieldDesc &f, const Xyzz &p, uint64_t *out) {
  if (xyzz_is_inf(f, p)) {
    memset(out, 0xff, 16 * f.nl);
    return;
  }
  Fe zzzi = inv(f, p.zzz);           // 1/ZZZ = 1/Z^3
  Fe zi = mul(f, zzzi, p.zz);        // Z^2 / Z^3 = 1/Z
  Fe zzi = sqr(f, zi);               // 1/Z^2
  store(f, out, mul(f, p.x, zzi));
  store(f, out + f.nl, mul(f, p.y, zzzi));
}

}  // namespace zkh
