// Device prime-field layer: Montgomery arithmetic with 32-bit limbs on the CDNA4 VALU.
//
// Representation.  An element of a field with L 64-bit limbs in the reference's flat layout
// (little-endian u64 limbs, Montgomery form with R = 2^(64 L), canonical in [0,p)) is read
// here as N = 2L little-endian u32 limbs.  R = 2^(32 N) is the same number, so the Montgomery
// form is shared bit for bit and device kernels read the caller's buffers with no conversion
// (lib/cbits/curves/fields/mont/bn128_Fr_mont.c:129-193 defines the host-side convention).
//
// Multiplication is CIOS (coarsely integrated operand scanning).  Each row i adds a_i * b with
// v_mad_u64_u32 (t_j + a_i b_j always fits 64 bits) and merges the high halves with one
// v_addc_co_u32 carry chain; the reduction row adds m * p, m = t_0 * (-p^-1 mod 2^32), the
// same way and shifts by one limb.  The loop invariant t < 2p gives a canonical result after
// one conditional subtraction, exactly as the reference's REDC (bn128_Fr_mont.c:140-169).
// No floating point is used anywhere.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zk {

// ---------------------------------------------------------------------------------------
// field parameter packs (32-bit limbs, little endian); values equal the reference's
// lib/cbits/curves/fields/mont/*_mont.c constants (prime, R mod p, R^2 mod p)

struct Bn254Fp {
  static constexpr int N = 8;
  static constexpr uint32_t NP = 0xe4866389u;
  static constexpr uint32_t P[N] = {0xd87cfd47u, 0x3c208c16u, 0x6871ca8du, 0x97816a91u,
                                    0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
  static constexpr uint32_t ONE[N] = {0xc58f0d9du, 0xd35d438du, 0xf5c70b3du, 0x0a78eb28u,
                                      0x7879462cu, 0x666ea36fu, 0x9a07df2fu, 0x0e0a77c1u};
  static constexpr uint32_t R2[N] = {0x538afa89u, 0xf32cfc5bu, 0xd44501fbu, 0xb5e71911u,
                                     0x0a417ff6u, 0x47ab1effu, 0xcab8351fu, 0x06d89f71u};
};
struct Bn254Fr {
  static constexpr int N = 8;
  static constexpr uint32_t NP = 0xefffffffu;
  static constexpr uint32_t P[N] = {0xf0000001u, 0x43e1f593u, 0x79b97091u, 0x2833e848u,
                                    0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
  static constexpr uint32_t ONE[N] = {0x4ffffffbu, 0xac96341cu, 0x9f60cd29u, 0x36fc7695u,
                                      0x7879462eu, 0x666ea36fu, 0x9a07df2fu, 0x0e0a77c1u};
  static constexpr uint32_t R2[N] = {0xae216da7u, 0x1bb8e645u, 0xe35c59e3u, 0x53fe3ab1u,
                                     0x53bb8085u, 0x8c49833du, 0x7f4e44a5u, 0x0216d0b1u};
};
struct Bls381Fp {
  static constexpr int N = 12;
  static constexpr uint32_t NP = 0xfffcfffdu;
  static constexpr uint32_t P[N] = {0xffffaaabu, 0xb9feffffu, 0xb153ffffu, 0x1eabfffeu,
                                    0xf6b0f624u, 0x6730d2a0u, 0xf38512bfu, 0x64774b84u,
                                    0x434bacd7u, 0x4b1ba7b6u, 0x397fe69au, 0x1a0111eau};
  static constexpr uint32_t ONE[N] = {0x0002fffdu, 0x76090000u, 0xc40c0002u, 0xebf4000bu,
                                      0x53c758bau, 0x5f489857u, 0x70525745u, 0x77ce5853u,
                                      0xa256ec6du, 0x5c071a97u, 0xfa80e493u, 0x15f65ec3u};
  static constexpr uint32_t R2[N] = {0x1c341746u, 0xf4df1f34u, 0x09d104f1u, 0x0a76e6a6u,
                                     0x4c95b6d5u, 0x8de5476cu, 0x939d83c0u, 0x67eb88a9u,
                                     0xb519952du, 0x9a793e85u, 0x92cae3aau, 0x11988fe5u};
};
struct Bls381Fr {
  static constexpr int N = 8;
  static constexpr uint32_t NP = 0xffffffffu;
  static constexpr uint32_t P[N] = {0x00000001u, 0xffffffffu, 0xfffe5bfeu, 0x53bda402u,
                                    0x09a1d805u, 0x3339d808u, 0x299d7d48u, 0x73eda753u};
  static constexpr uint32_t ONE[N] = {0xfffffffeu, 0x00000001u, 0x00034802u, 0x5884b7fau,
                                      0xecbc4ff5u, 0x998c4fefu, 0xacc5056fu, 0x1824b159u};
  static constexpr uint32_t R2[N] = {0xf3f29c6du, 0xc999e990u, 0x87925c23u, 0x2b6cedcbu,
                                     0x7254398fu, 0x05d31496u, 0x9f59ff11u, 0x0748d9d9u};
};

template <class F>
struct Fe {
  uint32_t w[F::N];
};

// ---------------------------------------------------------------------------------------
// carry primitives (lower to v_add_co_u32 / v_addc_co_u32 / v_sub_co_u32 / v_subb_co_u32)

__device__ __forceinline__ uint32_t addc(uint32_t a, uint32_t b, uint32_t &c) {
  uint32_t co;
  uint32_t r = __builtin_addc(a, b, c, &co);
  c = co;
  return r;
}
__device__ __forceinline__ uint32_t subb(uint32_t a, uint32_t b, uint32_t &br) {
  uint32_t bo;
  uint32_t r = __builtin_subc(a, b, br, &bo);
  br = bo;
  return r;
}

template <class F>
__device__ __forceinline__ void fset_zero(Fe<F> &r) {
#pragma unroll
  for (int j = 0; j < F::N; j++) r.w[j] = 0;
}
template <class F>
__device__ __forceinline__ void fset_one(Fe<F> &r) {
#pragma unroll
  for (int j = 0; j < F::N; j++) r.w[j] = F::ONE[j];
}
template <class F>
__device__ __forceinline__ bool fis_zero(const Fe<F> &a) {
  uint32_t acc = 0;
#pragma unroll
  for (int j = 0; j < F::N; j++) acc |= a.w[j];
  return acc == 0;
}
template <class F>
__device__ __forceinline__ bool fis_eq(const Fe<F> &a, const Fe<F> &b) {
  uint32_t acc = 0;
#pragma unroll
  for (int j = 0; j < F::N; j++) acc |= a.w[j] ^ b.w[j];
  return acc == 0;
}

// r = t - p if t >= p (t has an optional carry word `hi`), else t
template <class F>
__device__ __forceinline__ void reduce_once(Fe<F> &r, const uint32_t *t, uint32_t hi) {
  uint32_t s[F::N];
  uint32_t br = 0;
#pragma unroll
  for (int j = 0; j < F::N; j++) s[j] = subb(t[j], F::P[j], br);
  const bool take = (hi != 0) || (br == 0);
#pragma unroll
  for (int j = 0; j < F::N; j++) r.w[j] = take ? s[j] : t[j];
}

// r = a + b mod p  (bn128_Fr_mont.c:84-88: add then subtract p when >= p)
template <class F>
__device__ __forceinline__ void fadd(Fe<F> &r, const Fe<F> &a, const Fe<F> &b) {
  uint32_t t[F::N];
  uint32_t c = 0;
#pragma unroll
  for (int j = 0; j < F::N; j++) t[j] = addc(a.w[j], b.w[j], c);
  reduce_once<F>(r, t, c);
}

// r = a - b mod p  (bn128_Fr_mont.c:97-101: subtract, add p back on borrow)
template <class F>
__device__ __forceinline__ void fsub(Fe<F> &r, const Fe<F> &a, const Fe<F> &b) {
  uint32_t t[F::N];
  uint32_t br = 0;
#pragma unroll
  for (int j = 0; j < F::N; j++) t[j] = subb(a.w[j], b.w[j], br);
  const uint32_t mask = 0u - br;
  uint32_t c = 0;
#pragma unroll
  for (int j = 0; j < F::N; j++) r.w[j] = addc(t[j], F::P[j] & mask, c);
}

// r = -a mod p (neg(0) = 0, bn128_Fr_mont.c:44-58)
template <class F>
__device__ __forceinline__ void fneg(Fe<F> &r, const Fe<F> &a) {
  const uint32_t mask = fis_zero(a) ? 0u : 0xffffffffu;
  uint32_t br = 0;
#pragma unroll
  for (int j = 0; j < F::N; j++) r.w[j] = subb(F::P[j] & mask, a.w[j], br);
}

template <class F>
__device__ __forceinline__ void fdbl(Fe<F> &r, const Fe<F> &a) {
  fadd(r, a, a);
}

// r = a * b * R^-1 mod p  (CIOS, see the file header)
template <class F>
__device__ __forceinline__ void fmul(Fe<F> &r, const Fe<F> &a, const Fe<F> &b) {
  constexpr int N = F::N;
  uint32_t t[N + 2];
#pragma unroll
  for (int j = 0; j < N + 2; j++) t[j] = 0;
#pragma unroll
  for (int i = 0; i < N; i++) {
    uint64_t pr[N];
#pragma unroll
    for (int j = 0; j < N; j++) pr[j] = (uint64_t)a.w[i] * b.w[j] + t[j];
    uint32_t c = 0;
    t[0] = (uint32_t)pr[0];
#pragma unroll
    for (int j = 1; j < N; j++) t[j] = addc((uint32_t)pr[j], (uint32_t)(pr[j - 1] >> 32), c);
    t[N] = addc(t[N], (uint32_t)(pr[N - 1] >> 32), c);
    t[N + 1] = c;
    const uint32_t m = t[0] * F::NP;
    uint64_t q[N];
#pragma unroll
    for (int j = 0; j < N; j++) q[j] = (uint64_t)m * F::P[j] + t[j];
    c = 0;
#pragma unroll
    for (int j = 1; j < N; j++) t[j - 1] = addc((uint32_t)q[j], (uint32_t)(q[j - 1] >> 32), c);
    t[N - 1] = addc(t[N], (uint32_t)(q[N - 1] >> 32), c);
    t[N] = t[N + 1] + c;
  }
  reduce_once<F>(r, t, t[N]);
}

template <class F>
__device__ __forceinline__ void fsqr(Fe<F> &r, const Fe<F> &a) {
  fmul(r, a, a);
}

// Montgomery -> standard (REDC of a) and standard -> Montgomery (times R^2)
template <class F>
__device__ __forceinline__ void fto_std(Fe<F> &r, const Fe<F> &a) {
  Fe<F> one;
#pragma unroll
  for (int j = 0; j < F::N; j++) one.w[j] = (j == 0);
  fmul(r, a, one);
}
template <class F>
__device__ __forceinline__ void ffrom_std(Fe<F> &r, const Fe<F> &a) {
  Fe<F> r2;
#pragma unroll
  for (int j = 0; j < F::N; j++) r2.w[j] = F::R2[j];
  fmul(r, a, r2);
}

// a^e for a small public exponent held in a u64
template <class F>
__device__ __forceinline__ void fpow_u64(Fe<F> &r, const Fe<F> &a, uint64_t e) {
  Fe<F> acc, base = a;
  fset_one(acc);
  while (e) {
    if (e & 1) fmul(acc, acc, base);
    fsqr(base, base);
    e >>= 1;
  }
  r = acc;
}

// a^(p-2): Fermat inverse (0 maps to 0)
template <class F>
__device__ __noinline__ void finv(Fe<F> &r, const Fe<F> &a) {
  Fe<F> acc;
  fset_one(acc);
  uint32_t e[F::N];
  uint32_t br = 0;
#pragma unroll
  for (int j = 0; j < F::N; j++) e[j] = subb(F::P[j], j == 0 ? 2u : 0u, br);
  for (int i = 32 * F::N - 1; i >= 0; i--) {
    fsqr(acc, acc);
    if ((e[i >> 5] >> (i & 31)) & 1) fmul(acc, acc, a);
  }
  r = acc;
}

// halving: (x even) x/2, (x odd) (x+p)/2  (lib/cbits/curves/fields/std/bn128_Fr_std.c:241-249)
template <class F>
__device__ __forceinline__ void fhalf(Fe<F> &r, const Fe<F> &a) {
  const uint32_t mask = 0u - (a.w[0] & 1u);
  uint32_t t[F::N];
  uint32_t c = 0;
#pragma unroll
  for (int j = 0; j < F::N; j++) t[j] = addc(a.w[j], F::P[j] & mask, c);
#pragma unroll
  for (int j = 0; j < F::N - 1; j++) r.w[j] = (t[j] >> 1) | (t[j + 1] << 31);
  r.w[F::N - 1] = (t[F::N - 1] >> 1) | (c << 31);
}

// ---------------------------------------------------------------------------------------
// global-memory access in the reference layout (u64 limbs == pairs of u32 limbs)

template <class F>
__device__ __forceinline__ void fload(Fe<F> &r, const uint32_t *__restrict__ p) {
  static_assert(F::N % 4 == 0, "vector loads assume a multiple of 4 limbs");
  const uint4 *q = reinterpret_cast<const uint4 *>(p);
#pragma unroll
  for (int j = 0; j < F::N / 4; j++) {
    uint4 v = q[j];
    r.w[4 * j + 0] = v.x;
    r.w[4 * j + 1] = v.y;
    r.w[4 * j + 2] = v.z;
    r.w[4 * j + 3] = v.w;
  }
}
template <class F>
__device__ __forceinline__ void fstore(uint32_t *__restrict__ p, const Fe<F> &a) {
  uint4 *q = reinterpret_cast<uint4 *>(p);
#pragma unroll
  for (int j = 0; j < F::N / 4; j++) q[j] = make_uint4(a.w[4 * j], a.w[4 * j + 1], a.w[4 * j + 2], a.w[4 * j + 3]);
}

}  // namespace zk
