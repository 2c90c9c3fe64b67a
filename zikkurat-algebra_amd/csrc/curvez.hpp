// G1 bucket arithmetic on the reduced-radix field (fz.hpp): XYZZ accumulators, mixed
// XYZZ + affine additions (EFD madd-2008-s, 8M + 2S), full XYZZ additions (add-2008-s) and
// doublings (dbl/mdbl-2008-s-1), with every exceptional case exact (infinite accumulator,
// P + P, P + (-P)).
//
// Bounds (p = field prime, all values kept normalised, see fz.hpp): products of operands
// below 8p in magnitude are below 1.5p; the formulas below combine at most four products
// before normalising, so every stored coordinate stays below 6p and every multiplication
// input below 8p.  Infinity is the explicit all-zero ZZ (never produced by a product).
#pragma once
#include "curve.hpp"
#include "fz.hpp"

namespace zk {

template <class C>
struct ZP;  // reduced-radix pack of the curve's base field
template <>
struct ZP<Bn254G1> {
  using T = ZBn254Fp;
};
template <>
struct ZP<Bls381G1> {
  using T = ZBls381Fp;
};

template <class C>
using FzP = Fz<typename ZP<C>::T>;

template <class C>
struct AffZ {
  FzP<C> x, y;
};
template <class C>
struct XyzzZ {
  FzP<C> x, y, zz, zzz;
};

// words per stored element in device memory
template <class C>
struct ZLayout {
  static constexpr int N = ZP<C>::T::N;
  // affine point records, read at random indices by the MSM, never straddle a line:
  // PACKED records hold the canonical internal values bit-packed into EXT_WORDS words each
  // (BN254: 2 x 32 bytes, half a 128-byte line, unpacked by shifts on load); otherwise the
  // limbs as stored, padded to one 128-byte line (BLS12-381: 2 x 14 limbs)
  static constexpr bool PACKED = 2 * ZP<C>::T::EXT_WORDS == 16;
  static constexpr int AFF = PACKED ? 16 : 32;
  static_assert(PACKED || 2 * N <= AFF, "affine point must fit one 128-byte line");
  static constexpr int XYZZ = 4 * N;               // 36 or 56 words: multiples of 4
  static constexpr int JACC = (5 * N + 3) / 4 * 4;  // cached Jacobian table entry: 48 or 72 words
  // records per point in an internal point set: BN254 keeps P_i and phi(P_i) = (beta x_i, y_i)
  // side by side (records 2i and 2i + 1, one 128-byte line), so any prefix of a prepared set
  // is a prepared set and the GLV split applies to it; BLS12-381 keeps P_i alone
  static constexpr int STRIDE = ZP<C>::T::N == 9 ? 2 : 1;
};

template <class C>
ZK_HD void zxyzz_set_inf(XyzzZ<C> &p) {
  zset_zero(p.x);
  zset_zero(p.y);
  zset_zero(p.zz);
  zset_zero(p.zzz);
}
template <class C>
ZK_HD bool zxyzz_is_inf(const XyzzZ<C> &p) {
  return zis_raw_zero(p.zz);
}
template <class C>
ZK_HD bool zaff_is_inf(const AffZ<C> &a) {
  return zis_raw_zero(a.x) && zis_raw_zero(a.y);
}

// r = 2 (x, y) for an affine point
template <class C>
ZK_HD void zxyzz_mdbl(XyzzZ<C> &r, const AffZ<C> &a) {
  using F = FzP<C>;
  if (zis_zero_mod(a.y)) {
    zxyzz_set_inf(r);
    return;
  }
  F u, v, w, s, m, t;
  zadd(u, a.y, a.y);
  znorm(u);
  zsqr(v, u);
  zmul(w, u, v);
  zmul(s, a.x, v);
  zsqr(t, a.x);
  zadd(m, t, t);
  zadd(m, m, t);
  znorm(m);
  zsqr(t, m);
  zsub(t, t, s);
  zsub(t, t, s);
  znorm(t);  // X3 = M^2 - 2S
  r.x = t;
  zsub(t, s, r.x);
  znorm(t);
  zmul_sub(r.y, m, t, w, a.y);  // Y3 = M (S - X3) - W Y1, one reduction
  r.zz = v;
  r.zzz = w;
}

// r = 2 p (XYZZ)
template <class C>
ZK_HD void zxyzz_dbl(XyzzZ<C> &r, const XyzzZ<C> &p) {
  using F = FzP<C>;
  if (zxyzz_is_inf(p) || zis_zero_mod(p.y)) {
    zxyzz_set_inf(r);
    return;
  }
  F u, v, w, s, m, t, y1 = p.y, zz1 = p.zz, zzz1 = p.zzz;
  zadd(u, p.y, p.y);
  znorm(u);
  zsqr(v, u);
  zmul(w, u, v);
  zmul(s, p.x, v);
  zsqr(t, p.x);
  zadd(m, t, t);
  zadd(m, m, t);
  znorm(m);
  zsqr(t, m);
  zsub(t, t, s);
  zsub(t, t, s);
  znorm(t);
  r.x = t;
  zsub(t, s, r.x);
  znorm(t);
  zmul_sub(r.y, m, t, w, y1);
  zmul(r.zz, v, zz1);
  zmul(r.zzz, w, zzz1);
}

// acc += a (a affine, not infinity)
template <class C>
ZK_HD void zxyzz_madd(XyzzZ<C> &acc, const AffZ<C> &a) {
  using F = FzP<C>;
  if (zxyzz_is_inf(acc)) {
    acc.x = a.x;
    acc.y = a.y;
    zset_one(acc.zz);
    zset_one(acc.zzz);
    return;
  }
  // P and R stay limb-wise differences of two normalised values (limbs below 2^29 in
  // magnitude): valid operands of zsqr, zmul and zmul_sub, and of the zero test, which reads
  // the value, not the limb ranges
  F P, R, pp, ppp, q, t;
  zmul(P, a.x, acc.zz);
  zsub(P, P, acc.x);  // P = U2 - X1
  zmul(R, a.y, acc.zzz);
  zsub(R, R, acc.y);  // R = S2 - Y1
  if (zis_zero_mod(P)) {
    if (zis_zero_mod(R)) zxyzz_mdbl(acc, a);
    else zxyzz_set_inf(acc);
    return;
  }
  zsqr(pp, P);
  zmul(ppp, P, pp);
  zmul(q, acc.x, pp);
  zmul(acc.zz, acc.zz, pp);
  zmul(acc.zzz, acc.zzz, ppp);
  zsqr(t, R);
  zsub(t, t, ppp);
  zsub(t, t, q);
  zsub(t, t, q);
  znorm(t);
  acc.x = t;              // X3 = R^2 - PPP - 2Q
  zsub(t, q, acc.x);      // Q - X3, a difference of normalised values
  zmul_sub(q, R, t, acc.y, ppp);
  acc.y = q;              // Y3 = R (Q - X3) - Y1 PPP, one reduction
}

// r = p + q (XYZZ, complete)
template <class C>
ZK_HD void zxyzz_add(XyzzZ<C> &r, const XyzzZ<C> &p, const XyzzZ<C> &q) {
  using F = FzP<C>;
  if (zxyzz_is_inf(p)) {
    r = q;
    return;
  }
  if (zxyzz_is_inf(q)) {
    r = p;
    return;
  }
  F u1, s1, P, R, pp, ppp, t, zz, zzz;
  zmul(u1, p.x, q.zz);
  zmul(P, q.x, p.zz);
  zsub(P, P, u1);  // differences of normalised values, as in zxyzz_madd
  zmul(s1, p.y, q.zzz);
  zmul(R, q.y, p.zzz);
  zsub(R, R, s1);
  if (zis_zero_mod(P)) {
    if (zis_zero_mod(R)) {
      XyzzZ<C> pc = p;
      zxyzz_dbl(r, pc);
    } else {
      zxyzz_set_inf(r);
    }
    return;
  }
  zsqr(pp, P);
  zmul(ppp, P, pp);
  zmul(zz, p.zz, q.zz);
  zmul(zz, zz, pp);
  zmul(zzz, p.zzz, q.zzz);
  zmul(zzz, zzz, ppp);
  zmul(u1, u1, pp);  // Q
  zsqr(t, R);
  zsub(t, t, ppp);
  zsub(t, t, u1);
  zsub(t, t, u1);
  znorm(t);
  F x3 = t;
  zsub(t, u1, x3);
  zmul_sub(r.y, R, t, s1, ppp);
  r.x = x3;
  r.zz = zz;
  r.zzz = zzz;
}

// r = -p
template <class C>
ZK_HD void zxyzz_neg(XyzzZ<C> &r, const XyzzZ<C> &p) {
  r = p;
  zneg(r.y, p.y);
  znorm(r.y);
}

// ---- Jacobian chains of the windowed scalar multiplications
//
// A windowed chain k P is mostly doublings.  In Jacobian coordinates (x = X / Z^2,
// y = Y / Z^3; infinity is the all-zero Z) a doubling on y^2 = x^3 + b is 3S + 2M with the
// last two products in one zmul_sub, against 3S + 4M + zmul_sub in XYZZ: about a quarter
// fewer multiply-adds.  Table entries keep Z^2 and Z^3 beside (X, Y, Z) ("cached" entries),
// which makes adding one to a Jacobian accumulator 3S + 9M + zmul_sub, the cost of the full
// XYZZ addition.  Bounds as above: stored X below 6p, Y below 2p, Z below 1.5p.
template <class C>
struct JacZ {
  FzP<C> x, y, z;
};
template <class C>
struct JacCZ {
  FzP<C> x, y, z, zz, zzz;
};

template <class C>
ZK_HD void zjac_set_inf(JacZ<C> &p) {
  zset_zero(p.x);
  zset_zero(p.y);
  zset_zero(p.z);
}
template <class C>
ZK_HD bool zjac_is_inf(const JacZ<C> &p) {
  return zis_raw_zero(p.z);
}

// (X, Y, ZZ, ZZZ) -> (X ZZ ZZZ^2, Y ZZZ^4, ZZ ZZZ): for ZZ = Z^2, ZZZ = Z^3 the Jacobian point
// with Z' = Z^5 (X' = x Z^10, Y' = y Z^15), without an inversion
template <class C>
ZK_HD void zjac_from_xyzz(JacZ<C> &r, const XyzzZ<C> &p) {
  if (zxyzz_is_inf(p)) {
    zjac_set_inf(r);
    return;
  }
  FzP<C> t, u;
  zsqr(t, p.zzz);
  zmul(u, p.x, p.zz);
  zmul(r.x, u, t);
  zsqr(u, t);
  zmul(r.y, p.y, u);
  zmul(r.z, p.zz, p.zzz);
}
template <class C>
ZK_HD void zjac_to_xyzz(XyzzZ<C> &r, const JacZ<C> &p) {
  if (zjac_is_inf(p)) {
    zxyzz_set_inf(r);
    return;
  }
  r.x = p.x;
  r.y = p.y;
  zsqr(r.zz, p.z);
  zmul(r.zzz, r.zz, p.z);
}
template <class C>
ZK_HD void zjac_cache(JacCZ<C> &c, const JacZ<C> &p) {
  c.x = p.x;
  c.y = p.y;
  c.z = p.z;
  zsqr(c.zz, p.z);
  zmul(c.zzz, c.zz, p.z);
}
template <class C>
ZK_HD void zjac_uncache(JacZ<C> &r, const JacCZ<C> &c) {
  r.x = c.x;
  r.y = c.y;
  r.z = c.z;
}
// -c for a cached entry (only Y changes sign)
template <class C>
ZK_HD void zjac_neg(JacCZ<C> &c) {
  zneg(c.y, c.y);
  znorm(c.y);
}

// r = 2 p: A = X^2, B = Y^2, D = 4 X B, E = 3 A, X3 = E^2 - 2 D, Y3 = E (D - X3) - 8 B^2,
// Z3 = 2 Y Z (EFD dbl-2009-l with D as one product and the last two products as one zmul_sub,
// 8 B^2 = (2B)(4B)).  No point of these curves has y = 0, so finite points double to finite
// points.  r may be p.
template <class C>
ZK_HD void zjac_dbl(JacZ<C> &r, const JacZ<C> &p) {
  using F = FzP<C>;
  if (zjac_is_inf(p)) {
    zjac_set_inf(r);
    return;
  }
  F a, b, b2, b4, d, e, t;
  zsqr(a, p.x);
  zsqr(b, p.y);
  zadd(b2, b, b);
  znorm(b2);
  zadd(b4, b2, b2);
  znorm(b4);
  zmul(d, p.x, b4);
  zadd(t, p.y, p.y);
  znorm(t);
  zmul(r.z, t, p.z);  // p.x and p.y are still read below, p.z no more
  zadd(e, a, a);
  zadd(e, e, a);
  znorm(e);
  zsqr(t, e);
  zsub(t, t, d);
  zsub(t, t, d);
  znorm(t);
  r.x = t;
  zsub(t, d, r.x);
  znorm(t);
  zmul_sub(r.y, e, t, b2, b4);
}

// r = p + q for a cached entry q (EFD add-1998-cmo-2 with Z2^2, Z2^3 given): U1 = X1 ZZ2,
// U2 = X2 Z1^2, S1 = Y1 ZZZ2, S2 = Y2 Z1^3, H = U2 - U1, R = S2 - S1, X3 = R^2 - H^3 - 2 U1 H^2,
// Y3 = R (U1 H^2 - X3) - S1 H^3, Z3 = Z1 Z2 H; p = q doubles, p = -q gives infinity.  r may
// be p.
template <class C>
ZK_HD void zjac_add(JacZ<C> &r, const JacZ<C> &p, const JacCZ<C> &q) {
  using F = FzP<C>;
  if (zjac_is_inf(p)) {
    zjac_uncache(r, q);
    return;
  }
  if (zis_raw_zero(q.z)) {
    r = p;
    return;
  }
  F z1z1, z1c, u1, s1, H, R, hh, hhh, v, t, z3;
  zsqr(z1z1, p.z);
  zmul(z1c, z1z1, p.z);
  zmul(u1, p.x, q.zz);
  zmul(H, q.x, z1z1);
  zsub(H, H, u1);  // differences of normalised values, as in zxyzz_add
  zmul(s1, p.y, q.zzz);
  zmul(R, q.y, z1c);
  zsub(R, R, s1);
  if (zis_zero_mod(H)) {
    if (zis_zero_mod(R)) {
      JacZ<C> pc = p;
      zjac_dbl(r, pc);
    } else {
      zjac_set_inf(r);
    }
    return;
  }
  zmul(z3, p.z, q.z);
  zmul(z3, z3, H);
  zsqr(hh, H);
  zmul(hhh, H, hh);
  zmul(v, u1, hh);
  zsqr(t, R);
  zsub(t, t, hhh);
  zsub(t, t, v);
  zsub(t, t, v);
  znorm(t);
  r.x = t;
  zsub(t, v, r.x);
  zmul_sub(r.y, R, t, s1, hhh);
  r.z = z3;
}

// r = k p for a scalar k < 2^256 in standard form (8 little-endian u32 words): binary
// double-and-add from the top set bit (any exact method gives the same point)
template <class C>
ZK_HD void zxyzz_scalar_mul(XyzzZ<C> &r, const XyzzZ<C> &p, const uint32_t (&k)[8]) {
  zxyzz_set_inf(r);
  int top = 255;
  while (top >= 0 && !((k[top >> 5] >> (top & 31)) & 1u)) top--;
  for (int i = top; i >= 0; i--) {
    XyzzZ<C> t;
    zxyzz_dbl(t, r);
    if ((k[i >> 5] >> (i & 31)) & 1u) zxyzz_add(r, t, p);
    else r = t;
  }
}

// Per-thread point tables of the windowed scalar multiplications, in two layouts.
// Thread-major (zxyzz_tab_*): a thread's entries lie together, limbs as computed, entry e of
// its block tt at tt + e * XYZZ words; each lane's load is one contiguous 144-byte entry.  Best
// when there are many entries (the 15 of the joint GLV windows): lanes then rarely pick the
// same one.  Word-major (zxyzz_soa_*): word w of entry e of thread t at tab[(e * XYZZ + w) *
// stride + t] (stride >= the threads sharing the buffer); lanes that pick the same entry read
// consecutive words.  Best for few entries (the 3 of 2-bit windows on k).  Measured on the G1
// FFT: the thread-major table 13-16% faster at 2^16 for BN254, the word-major one 5% faster at
// 2^12 for BLS12-381.
template <class C>
ZK_HD void zxyzz_tab_store(uint32_t *__restrict__ tt, int e, const XyzzZ<C> &a) {
  constexpr int N = ZP<C>::T::N, TW = ZLayout<C>::XYZZ;
  static_assert(TW % 4 == 0, "entries are stored as 16-byte words");
  uint32_t w[TW];
#pragma unroll
  for (int l = 0; l < N; l++) {
    w[l] = (uint32_t)a.x.l[l];
    w[N + l] = (uint32_t)a.y.l[l];
    w[2 * N + l] = (uint32_t)a.zz.l[l];
    w[3 * N + l] = (uint32_t)a.zzz.l[l];
  }
  uint4 *q = reinterpret_cast<uint4 *>(tt + (size_t)e * TW);
#pragma unroll
  for (int j = 0; j < TW / 4; j++) q[j] = make_uint4(w[4 * j], w[4 * j + 1], w[4 * j + 2], w[4 * j + 3]);
}
template <class C>
ZK_HD void zxyzz_tab_load(XyzzZ<C> &a, const uint32_t *__restrict__ tt, int e) {
  constexpr int N = ZP<C>::T::N, TW = ZLayout<C>::XYZZ;
  uint32_t w[TW];
  const uint4 *q = reinterpret_cast<const uint4 *>(tt + (size_t)e * TW);
#pragma unroll
  for (int j = 0; j < TW / 4; j++) {
    const uint4 v = q[j];
    w[4 * j] = v.x;
    w[4 * j + 1] = v.y;
    w[4 * j + 2] = v.z;
    w[4 * j + 3] = v.w;
  }
#pragma unroll
  for (int l = 0; l < N; l++) {
    a.x.l[l] = (int32_t)w[l];
    a.y.l[l] = (int32_t)w[N + l];
    a.zz.l[l] = (int32_t)w[2 * N + l];
    a.zzz.l[l] = (int32_t)w[3 * N + l];
  }
}
template <class C>
ZK_HD void zxyzz_soa_store(uint32_t *__restrict__ tab, int e, uint64_t t, uint64_t stride, const XyzzZ<C> &a) {
  constexpr int N = ZP<C>::T::N, PW = ZLayout<C>::XYZZ;
  uint32_t *q = tab + (uint64_t)e * PW * stride + t;
#pragma unroll
  for (int l = 0; l < N; l++) {
    q[(uint64_t)l * stride] = (uint32_t)a.x.l[l];
    q[(uint64_t)(N + l) * stride] = (uint32_t)a.y.l[l];
    q[(uint64_t)(2 * N + l) * stride] = (uint32_t)a.zz.l[l];
    q[(uint64_t)(3 * N + l) * stride] = (uint32_t)a.zzz.l[l];
  }
}
template <class C>
ZK_HD void zxyzz_soa_load(XyzzZ<C> &a, const uint32_t *__restrict__ tab, int e, uint64_t t, uint64_t stride) {
  constexpr int N = ZP<C>::T::N, PW = ZLayout<C>::XYZZ;
  const uint32_t *q = tab + (uint64_t)e * PW * stride + t;
#pragma unroll
  for (int l = 0; l < N; l++) {
    a.x.l[l] = (int32_t)q[(uint64_t)l * stride];
    a.y.l[l] = (int32_t)q[(uint64_t)(N + l) * stride];
    a.zz.l[l] = (int32_t)q[(uint64_t)(2 * N + l) * stride];
    a.zzz.l[l] = (int32_t)q[(uint64_t)(3 * N + l) * stride];
  }
}

// The same two layouts for cached Jacobian entries (ZLayout<C>::JACC words a thread-major entry,
// the words past 5 N zero; 5 N words a word-major entry).
template <class C>
ZK_HD void zjac_tab_store(uint32_t *__restrict__ tt, int e, const JacCZ<C> &a) {
  constexpr int N = ZP<C>::T::N, TW = ZLayout<C>::JACC;
  uint32_t w[TW];
#pragma unroll
  for (int l = 0; l < N; l++) {
    w[l] = (uint32_t)a.x.l[l];
    w[N + l] = (uint32_t)a.y.l[l];
    w[2 * N + l] = (uint32_t)a.z.l[l];
    w[3 * N + l] = (uint32_t)a.zz.l[l];
    w[4 * N + l] = (uint32_t)a.zzz.l[l];
  }
#pragma unroll
  for (int j = 5 * N; j < TW; j++) w[j] = 0;
  uint4 *q = reinterpret_cast<uint4 *>(tt + (size_t)e * TW);
#pragma unroll
  for (int j = 0; j < TW / 4; j++) q[j] = make_uint4(w[4 * j], w[4 * j + 1], w[4 * j + 2], w[4 * j + 3]);
}
template <class C>
ZK_HD void zjac_tab_load(JacCZ<C> &a, const uint32_t *__restrict__ tt, int e) {
  constexpr int N = ZP<C>::T::N, TW = ZLayout<C>::JACC;
  uint32_t w[TW];
  const uint4 *q = reinterpret_cast<const uint4 *>(tt + (size_t)e * TW);
#pragma unroll
  for (int j = 0; j < TW / 4; j++) {
    const uint4 v = q[j];
    w[4 * j] = v.x;
    w[4 * j + 1] = v.y;
    w[4 * j + 2] = v.z;
    w[4 * j + 3] = v.w;
  }
#pragma unroll
  for (int l = 0; l < N; l++) {
    a.x.l[l] = (int32_t)w[l];
    a.y.l[l] = (int32_t)w[N + l];
    a.z.l[l] = (int32_t)w[2 * N + l];
    a.zz.l[l] = (int32_t)w[3 * N + l];
    a.zzz.l[l] = (int32_t)w[4 * N + l];
  }
}
template <class C>
ZK_HD void zjac_soa_store(uint32_t *__restrict__ tab, int e, uint64_t t, uint64_t stride, const JacCZ<C> &a) {
  constexpr int N = ZP<C>::T::N, PW = 5 * N;
  uint32_t *q = tab + (uint64_t)e * PW * stride + t;
#pragma unroll
  for (int l = 0; l < N; l++) {
    q[(uint64_t)l * stride] = (uint32_t)a.x.l[l];
    q[(uint64_t)(N + l) * stride] = (uint32_t)a.y.l[l];
    q[(uint64_t)(2 * N + l) * stride] = (uint32_t)a.z.l[l];
    q[(uint64_t)(3 * N + l) * stride] = (uint32_t)a.zz.l[l];
    q[(uint64_t)(4 * N + l) * stride] = (uint32_t)a.zzz.l[l];
  }
}
template <class C>
ZK_HD void zjac_soa_load(JacCZ<C> &a, const uint32_t *__restrict__ tab, int e, uint64_t t, uint64_t stride) {
  constexpr int N = ZP<C>::T::N, PW = 5 * N;
  const uint32_t *q = tab + (uint64_t)e * PW * stride + t;
#pragma unroll
  for (int l = 0; l < N; l++) {
    a.x.l[l] = (int32_t)q[(uint64_t)l * stride];
    a.y.l[l] = (int32_t)q[(uint64_t)(N + l) * stride];
    a.z.l[l] = (int32_t)q[(uint64_t)(2 * N + l) * stride];
    a.zz.l[l] = (int32_t)q[(uint64_t)(3 * N + l) * stride];
    a.zzz.l[l] = (int32_t)q[(uint64_t)(4 * N + l) * stride];
  }
}

// The same product by 2-bit windows over a 3-entry word-major table of cached Jacobian entries
// (p, 2p, 3p; 5 N words each) of thread t: a step is two doublings and one addition, ~127
// additions for a 255-bit scalar where the binary form, whose wavefronts add at almost every
// bit (one lane with a set bit suffices), pays ~255.  The chain runs in Jacobian coordinates.
// With [wlo, whi) only those windows are added (the doublings run down to window 0), so the
// results over a split of [0, 128) sum to k p.
template <class C>
ZK_HD void zxyzz_scalar_mul_w2(XyzzZ<C> &r, const XyzzZ<C> &p, const uint32_t (&k)[8], uint32_t *__restrict__ tab,
                               uint64_t t, uint64_t stride, int wlo = 0, int whi = 128) {
  {
    JacZ<C> e1, e2, e3;
    JacCZ<C> c;
    zjac_from_xyzz(e1, p);
    zjac_dbl(e2, e1);
    zjac_cache(c, e1);
    zjac_soa_store<C>(tab, 0, t, stride, c);
    zjac_add(e3, e2, c);
    zjac_cache(c, e2);
    zjac_soa_store<C>(tab, 1, t, stride, c);
    zjac_cache(c, e3);
    zjac_soa_store<C>(tab, 2, t, stride, c);
  }
  auto digit = [&](int s) -> uint32_t { return (k[s >> 4] >> (2 * (s & 15))) & 3u; };
  int top = whi - 1;
  while (top >= wlo && digit(top) == 0) top--;
  if (top < wlo) {
    zxyzz_set_inf(r);
    return;
  }
  JacZ<C> acc;
  {
    JacCZ<C> e;
    zjac_soa_load<C>(e, tab, (int)digit(top) - 1, t, stride);
    zjac_uncache(acc, e);
  }
  for (int s = top - 1; s >= 0; s--) {
    const uint32_t d = s >= wlo ? digit(s) : 0u;
    JacCZ<C> e;
    if (d) zjac_soa_load<C>(e, tab, (int)d - 1, t, stride);  // in flight during the doublings
    zjac_dbl(acc, acc);
    zjac_dbl(acc, acc);
    if (d) zjac_add(acc, acc, e);
  }
  zjac_to_xyzz(r, acc);
}

// ---- device memory

template <class C>
__device__ __forceinline__ void zaff_load(AffZ<C> &a, const int32_t *__restrict__ p) {
  constexpr int N = ZP<C>::T::N, E = ZP<C>::T::EXT_WORDS;
  if (ZLayout<C>::PACKED) {
    const uint4 *q = reinterpret_cast<const uint4 *>(p);
    uint32_t w[2 * E];
#pragma unroll
    for (int j = 0; j < E / 2; j++) {
      const uint4 v = q[j];
      w[4 * j] = v.x;
      w[4 * j + 1] = v.y;
      w[4 * j + 2] = v.z;
      w[4 * j + 3] = v.w;
    }
    zunpack(a.x, w);
    zunpack(a.y, w + E);
  } else {
    zload(a.x, p);
    zload(a.y, p + N);
  }
}
// a whole record (padding included); infinity is the all-zero record
template <class C>
__device__ __forceinline__ void zaff_store(int32_t *__restrict__ p, const AffZ<C> &a) {
  constexpr int N = ZP<C>::T::N, E = ZP<C>::T::EXT_WORDS;
  if (ZLayout<C>::PACKED) {
    // packed in registers, then 16-byte stores (records are 64-byte aligned): a quarter of the
    // store instructions of word stores, each lane's record in its own line
    FzP<C> t;
    uint32_t w[2 * E];
    zcanon(t, a.x);
    zpack(w, t);
    zcanon(t, a.y);
    zpack(w + E, t);
    uint4 *q = reinterpret_cast<uint4 *>(p);
#pragma unroll
    for (int j = 0; j < E / 2; j++) q[j] = make_uint4(w[4 * j], w[4 * j + 1], w[4 * j + 2], w[4 * j + 3]);
  } else {
    constexpr int AW = ZLayout<C>::AFF;
    int32_t w[AW];
#pragma unroll
    for (int j = 0; j < N; j++) {
      w[j] = a.x.l[j];
      w[N + j] = a.y.l[j];
    }
#pragma unroll
    for (int j = 2 * N; j < AW; j++) w[j] = 0;
    int4 *q = reinterpret_cast<int4 *>(p);
#pragma unroll
    for (int j = 0; j < AW / 4; j++) q[j] = make_int4(w[4 * j], w[4 * j + 1], w[4 * j + 2], w[4 * j + 3]);
  }
}
template <class C>
__device__ __forceinline__ void zxyzz_load(XyzzZ<C> &a, const int32_t *__restrict__ p) {
  constexpr int N = ZP<C>::T::N;
  zload(a.x, p);
  zload(a.y, p + N);
  zload(a.zz, p + 2 * N);
  zload(a.zzz, p + 3 * N);
}
template <class C>
__device__ __forceinline__ void zxyzz_store(int32_t *__restrict__ p, const XyzzZ<C> &a) {
  constexpr int N = ZP<C>::T::N;
  zstore(p, a.x);
  zstore(p + N, a.y);
  zstore(p + 2 * N, a.zz);
  zstore(p + 3 * N, a.zzz);
}

}  // namespace zk
