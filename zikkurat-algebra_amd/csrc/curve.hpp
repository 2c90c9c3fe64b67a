// Device G1 arithmetic for BN254 and BLS12-381 (short Weierstrass, a = 0).
//
// Bucket accumulators use XYZZ coordinates (x = X/ZZ, y = Y/ZZZ, ZZ^3 = ZZZ^2; infinity is
// ZZ = ZZZ = 0).  Mixed XYZZ + affine addition (EFD madd-2008-s) costs 8M + 2S, against
// 9M + 2S for the reference's projective madd-1998-cmo (lib/cbits/curves/g1/proj/
// bn128_G1_proj.c:333-373), and every exceptional case is handled exactly: an infinite
// accumulator takes the affine point, P + P doubles (mdbl-2008-s-1), P + (-P) gives infinity,
// and the affine infinity (0,0) (bn128_G1_affine.c:57-60) is skipped by the caller.
#pragma once
#include "field.hpp"

namespace zk {

struct Bn254G1 {
  using Fp = Bn254Fp;
  using Fr = Bn254Fr;
  static constexpr int id = 0;
  static constexpr int LP = 4;  // 64-bit limbs of Fp
  // standard generator (1, 2), Montgomery form (bn128_G1_proj.c:32)
  static constexpr uint32_t GX[8] = {0xc58f0d9du, 0xd35d438du, 0xf5c70b3du, 0x0a78eb28u,
                                     0x7879462cu, 0x666ea36fu, 0x9a07df2fu, 0x0e0a77c1u};
  static constexpr uint32_t GY[8] = {0x8b1e1b3au, 0xa6ba871bu, 0xeb8e167bu, 0x14f1d651u,
                                     0xf0f28c58u, 0xccdd46deu, 0x340fbe5eu, 0x1c14ef83u};
};

struct Bls381G1 {
  using Fp = Bls381Fp;
  using Fr = Bls381Fr;
  static constexpr int id = 1;
  static constexpr int LP = 6;
  // standard generator, Montgomery form (bls12_381_G1_proj.c:33)
  static constexpr uint32_t GX[12] = {0xfd530c16u, 0x5cb38790u, 0x9976fff5u, 0x7817fc67u,
                                      0x143ba1c1u, 0x154f95c7u, 0xf3d0e747u, 0xf0ae6acdu,
                                      0x21dbf440u, 0xedce6eccu, 0x9e0bfb75u, 0x12017741u};
  static constexpr uint32_t GY[12] = {0x0ce72271u, 0xbaac93d5u, 0x7918fd8eu, 0x8c22631au,
                                      0x570725ceu, 0xdd595f13u, 0x50405194u, 0x51ac5829u,
                                      0xad0059c0u, 0x0e1c8c3fu, 0x5008a26au, 0x0bbc3efcu};
};

template <class C>
struct Aff {
  Fe<typename C::Fp> x, y;
};

template <class C>
struct Xyzz {
  Fe<typename C::Fp> x, y, zz, zzz;
};

template <class C>
__device__ __forceinline__ void xyzz_set_inf(Xyzz<C> &p) {
  fset_zero(p.x);
  fset_zero(p.y);
  fset_zero(p.zz);
  fset_zero(p.zzz);
}
template <class C>
__device__ __forceinline__ bool xyzz_is_inf(const Xyzz<C> &p) {
  return fis_zero(p.zz);
}
template <class C>
__device__ __forceinline__ void xyzz_from_aff(Xyzz<C> &p, const Aff<C> &a) {
  p.x = a.x;
  p.y = a.y;
  fset_one(p.zz);
  fset_one(p.zzz);
}

// affine infinity encoding (0,0)
template <class C>
__device__ __forceinline__ bool aff_is_inf(const Aff<C> &a) {
  return fis_zero(a.x) && fis_zero(a.y);
}

template <class C>
__device__ __forceinline__ void aff_neg(Aff<C> &a) {
  fneg(a.y, a.y);
}

// 2 * (x, y) for affine (x, y): mdbl-2008-s-1
template <class C>
__device__ __forceinline__ void xyzz_mdbl(Xyzz<C> &r, const Aff<C> &a) {
  using F = typename C::Fp;
  if (fis_zero(a.y)) {
    xyzz_set_inf(r);
    return;
  }
  Fe<F> u, v, w, s, m, t;
  fadd(u, a.y, a.y);
  fsqr(v, u);
  fmul(w, u, v);
  fmul(s, a.x, v);
  fsqr(m, a.x);
  fadd(t, m, m);
  fadd(m, t, m);
  fsqr(t, m);
  fsub(t, t, s);
  fsub(r.x, t, s);
  fsub(t, s, r.x);
  fmul(t, m, t);
  fmul(s, w, a.y);
  fsub(r.y, t, s);
  r.zz = v;
  r.zzz = w;
}

// 2 * P for XYZZ P: dbl-2008-s-1
template <class C>
__device__ __forceinline__ void xyzz_dbl(Xyzz<C> &r, const Xyzz<C> &p) {
  using F = typename C::Fp;
  if (xyzz_is_inf(p) || fis_zero(p.y)) {
    xyzz_set_inf(r);
    return;
  }
  Fe<F> u, v, w, s, m, t, y1;
  y1 = p.y;
  fadd(u, p.y, p.y);
  fsqr(v, u);
  fmul(w, u, v);
  fmul(s, p.x, v);
  fsqr(m, p.x);
  fadd(t, m, m);
  fadd(m, t, m);
  fmul(r.zz, v, p.zz);
  fmul(r.zzz, w, p.zzz);
  fsqr(t, m);
  fsub(t, t, s);
  fsub(r.x, t, s);
  fsub(t, s, r.x);
  fmul(t, m, t);
  fmul(s, w, y1);
  fsub(r.y, t, s);
}

// acc += a  (acc XYZZ, a affine and not infinity): madd-2008-s
template <class C>
__device__ __forceinline__ void xyzz_madd(Xyzz<C> &acc, const Aff<C> &a) {
  using F = typename C::Fp;
  if (xyzz_is_inf(acc)) {
    xyzz_from_aff(acc, a);
    return;
  }
  Fe<F> P, R, pp, ppp, q, t;
  fmul(P, a.x, acc.zz);    // U2
  fsub(P, P, acc.x);       // P = U2 - X1
  fmul(R, a.y, acc.zzz);   // S2
  fsub(R, R, acc.y);       // R = S2 - Y1
  if (fis_zero(P)) {
    if (fis_zero(R)) xyzz_mdbl(acc, a);
    else xyzz_set_inf(acc);
    return;
  }
  fsqr(pp, P);
  fmul(ppp, P, pp);
  fmul(q, acc.x, pp);
  fmul(acc.zz, acc.zz, pp);
  fmul(acc.zzz, acc.zzz, ppp);
  fsqr(t, R);
  fsub(t, t, ppp);
  fsub(t, t, q);
  fsub(acc.x, t, q);       // X3 = R^2 - PPP - 2Q
  fsub(t, q, acc.x);
  fmul(t, R, t);
  fmul(q, acc.y, ppp);
  fsub(acc.y, t, q);       // Y3 = R (Q - X3) - Y1 PPP
}

// r = p + q (XYZZ, complete): add-2008-s with the exceptional cases
template <class C>
__device__ __forceinline__ void xyzz_add(Xyzz<C> &r, const Xyzz<C> &p, const Xyzz<C> &q) {
  using F = typename C::Fp;
  if (xyzz_is_inf(p)) {
    r = q;
    return;
  }
  if (xyzz_is_inf(q)) {
    r = p;
    return;
  }
  Fe<F> u1, s1, P, R, pp, ppp, t;
  fmul(u1, p.x, q.zz);
  fmul(P, q.x, p.zz);
  fsub(P, P, u1);
  fmul(s1, p.y, q.zzz);
  fmul(R, q.y, p.zzz);
  fsub(R, R, s1);
  if (fis_zero(P)) {
    if (fis_zero(R)) {
      Xyzz<C> pc = p;
      xyzz_dbl(r, pc);
    } else {
      xyzz_set_inf(r);
    }
    return;
  }
  fsqr(pp, P);
  fmul(ppp, P, pp);
  Fe<F> zz, zzz;
  fmul(zz, p.zz, q.zz);
  fmul(zz, zz, pp);
  fmul(zzz, p.zzz, q.zzz);
  fmul(zzz, zzz, ppp);
  fmul(u1, u1, pp);        // Q = U1 PP
  fsqr(t, R);
  fsub(t, t, ppp);
  fsub(t, t, u1);
  Fe<F> x3;
  fsub(x3, t, u1);
  fsub(t, u1, x3);
  fmul(t, R, t);
  fmul(s1, s1, ppp);
  fsub(r.y, t, s1);
  r.x = x3;
  r.zz = zz;
  r.zzz = zzz;
}

template <class C>
__device__ __forceinline__ void aff_load(Aff<C> &a, const uint32_t *p) {
  using F = typename C::Fp;
  fload(a.x, p);
  fload(a.y, p + F::N);
}
template <class C>
__device__ __forceinline__ void xyzz_load(Xyzz<C> &a, const uint32_t *p) {
  using F = typename C::Fp;
  fload(a.x, p);
  fload(a.y, p + F::N);
  fload(a.zz, p + 2 * F::N);
  fload(a.zzz, p + 3 * F::N);
}
template <class C>
__device__ __forceinline__ void xyzz_store(uint32_t *p, const Xyzz<C> &a) {
  using F = typename C::Fp;
  fstore(p, a.x);
  fstore(p + F::N, a.y);
  fstore(p + 2 * F::N, a.zz);
  fstore(p + 3 * F::N, a.zzz);
}

}  // namespace zk
