// FFT of G1 points over the scalar field's multiplicative subgroups (BN254 / BLS12-381).
//
// Contract (lib/cbits/curves/g1/proj/bn128_G1_proj.c:791-902): src and tgt hold 2^m
// projective Montgomery points (X:Y:Z); the forward transform is y_k = sum_i gen^(ik) x_i
// computed by the recursive radix-2 DIT network with twiddles gen^j, the inverse is the
// mirrored DIF network with (x + y)/2 and (x - y) gen^-j / 2 at every level; every output is
// normalised (X/Z, Y/Z, 1), infinity (0, 1, 0).  For any gen these networks are the same
// group-linear maps as the iterative forms below (bit-reversed DIT with gen_s = gen^(N/2^s),
// DIF with gen^-j, and the m halvings of every output path applied once as N^-1), and the
// normalised point is unique, so the outputs match bit for bit.
//
// Device form: points in XYZZ reduced-radix coordinates (one record per point, in HBM),
// twiddles as standard-form scalars; each butterfly multiplies one point by its twiddle
// (double-and-add over the scalar's bits) and adds / subtracts.  One launch per stage.
#include <string.h>

#include "curve.hpp"
#include "curvez.hpp"
#include "glv.hpp"
#include "field.hpp"
#include "hostfield.hpp"
#include "msm.hpp"
#include "runtime.hpp"

namespace zk {

static constexpr int kG1Threads = 128;

__device__ __forceinline__ uint32_t g1_bitrev(uint32_t x, int bits) {
  return bits ? (__builtin_bitreverse32(x) >> (32 - bits)) : 0u;
}

// projective external (X:Y:Z) -> XYZZ (X Z, Y Z^2, Z^2, Z^3); Jacobian (X:Y:Z), whose affine
// point is (X/Z^2, Y/Z^3), -> (X, Y, Z^2, Z^3); Z = 0 is the infinity
template <class C, bool JAC>
__global__ __launch_bounds__(kG1Threads) void g1_proj_in_kernel(const uint32_t *__restrict__ src, uint32_t N, int m,
                                                               int bitrev_out, int32_t *__restrict__ dst) {
  using F = typename C::Fp;
  constexpr int LW = F::N, PW = ZLayout<C>::XYZZ;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const uint32_t *p = src + (size_t)i * 3 * LW;
  uint32_t zor = 0;
#pragma unroll
  for (int j = 0; j < LW; j++) zor |= p[2 * LW + j];
  XyzzZ<C> q;
  if (zor == 0) {
    zxyzz_set_inf(q);
  } else {
    FzP<C> x, y, z;
    zfrom_ext(x, p);
    zfrom_ext(y, p + LW);
    zfrom_ext(z, p + 2 * LW);
    zsqr(q.zz, z);
    zmul(q.zzz, q.zz, z);
    if (JAC) {
      q.x = x;
      q.y = y;
    } else {
      zmul(q.x, x, z);
      zmul(q.y, y, q.zz);
    }
  }
  const uint32_t j = bitrev_out ? g1_bitrev(i, m) : i;
  zxyzz_store(dst + (size_t)j * PW, q);
}

// XYZZ -> normalised external words (x, y, 1), infinity (0, 1, 0); the projective and the
// Jacobian normalisations agree on these words (bn128_G1_proj.c:78-97, bn128_G1_jac.c:62-85)
template <class C>
__global__ __launch_bounds__(kG1Threads) void g1_proj_out_kernel(const int32_t *__restrict__ pts, uint32_t N, int m,
                                                                int bitrev_in, uint32_t *__restrict__ dst) {
  using F = typename C::Fp;
  constexpr int LW = F::N, PW = ZLayout<C>::XYZZ;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const uint32_t j = bitrev_in ? g1_bitrev(i, m) : i;
  XyzzZ<C> q;
  zxyzz_load(q, pts + (size_t)j * PW);
  uint32_t *o = dst + (size_t)i * 3 * LW;
  FzP<C> one;
  zset_one(one);
  if (zis_zero_mod(q.zz)) {
#pragma unroll
    for (int k = 0; k < LW; k++) o[k] = 0;
    zto_ext(o + LW, one);
#pragma unroll
    for (int k = 0; k < LW; k++) o[2 * LW + k] = 0;
    return;
  }
  FzP<C> t, inv, a, x, y;
  zmul(t, q.zz, q.zzz);
  zinv(inv, t);            // 1 / (ZZ ZZZ)
  zmul(a, inv, q.zzz);     // 1 / ZZ
  zmul(x, q.x, a);
  zmul(a, inv, q.zz);      // 1 / ZZZ
  zmul(y, q.y, a);
  zto_ext(o, x);
  zto_ext(o + LW, y);
  zto_ext(o + 2 * LW, one);
}

// tab[j] = g^j in standard form (8 u32 words), j < cnt
template <class C>
__global__ __launch_bounds__(kG1Threads) void g1_twiddle_kernel(uint32_t *__restrict__ tab, Fe<typename C::Fr> g,
                                                               uint32_t cnt) {
  using FR = typename C::Fr;
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cnt) return;
  Fe<FR> r;
  fpow_u64(r, g, j);
  fto_std(r, r);
  fstore(tab + (size_t)j * FR::N, r);
}

// k P: BN254 through the GLV endomorphism (exact for every point of its prime-order G1);
// BLS12-381, whose G1 has a cofactor, by double-and-add on k
template <class C>
__device__ __forceinline__ void scalar_mul(XyzzZ<C> &r, const XyzzZ<C> &p, const uint32_t (&k)[8]) {
  if constexpr (C::id == 0) zxyzz_scalar_mul_glv(r, p, k);
  else zxyzz_scalar_mul(r, p, k);
}

// one radix-2 stage over blocks of 2^s points (DIT: u + w v, u - w v; DIF: u + v, w (u - v))
template <class C, bool DIF>
__global__ __launch_bounds__(kG1Threads) void g1_fft_stage_kernel(int32_t *__restrict__ pts,
                                                                 const uint32_t *__restrict__ tw, int m, int s) {
  constexpr int PW = ZLayout<C>::XYZZ;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (1u << (m - 1))) return;
  const uint32_t half = 1u << (s - 1), pos = t & (half - 1), grp = t >> (s - 1);
  const uint32_t i0 = (grp << s) | pos, i1 = i0 + half;
  XyzzZ<C> u, v, w, a;
  zxyzz_load(u, pts + (size_t)i0 * PW);
  zxyzz_load(v, pts + (size_t)i1 * PW);
  uint32_t k[8];
  const uint32_t *kp = tw + ((size_t)pos << (m - s)) * 8;
#pragma unroll
  for (int j = 0; j < 8; j++) k[j] = kp[j];
  if (!DIF) {
    if (pos == 0) w = v;  // twiddle 1
    else scalar_mul<C>(w, v, k);
    zxyzz_add(a, u, w);
    zxyzz_neg(w, w);
    zxyzz_add(v, u, w);
    zxyzz_store(pts + (size_t)i0 * PW, a);
    zxyzz_store(pts + (size_t)i1 * PW, v);
  } else {
    zxyzz_add(a, u, v);
    zxyzz_neg(v, v);
    zxyzz_add(w, u, v);
    if (pos != 0) {
      XyzzZ<C> r;
      scalar_mul<C>(r, w, k);
      w = r;
    }
    zxyzz_store(pts + (size_t)i0 * PW, a);
    zxyzz_store(pts + (size_t)i1 * PW, w);
  }
}

// pts[i] = k pts[i]
template <class C>
__global__ __launch_bounds__(kG1Threads) void g1_scale_kernel(int32_t *__restrict__ pts, uint32_t N, Fe<typename C::Fr> kst) {
  constexpr int PW = ZLayout<C>::XYZZ;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  uint32_t k[8];
#pragma unroll
  for (int j = 0; j < 8; j++) k[j] = kst.w[j];
  XyzzZ<C> p, r;
  zxyzz_load(p, pts + (size_t)i * PW);
  scalar_mul<C>(r, p, k);
  zxyzz_store(pts + (size_t)i * PW, r);
}

template <class C>
static int g1_fft_run(int m, const uint64_t *gen, bool inverse, bool jac, const uint64_t *d_src, uint64_t *d_dst,
                      hipStream_t st) {
  using FR = typename C::Fr;
  constexpr int ZW = ZLayout<C>::XYZZ;
  const uint32_t N = 1u << m, halfN = N >> 1;
  const zkh::FieldDesc &fr = C::id == 0 ? zkh::BN254_FR : zkh::BLS381_FR;
  zkh::Fe g = zkh::load(fr, gen);
  if (inverse) {
    if (zkh::is_zero(fr, g)) {
      set_error("g1_fft: the generator must be invertible");
      return ZK_EINVAL;
    }
    g = zkh::inv(fr, g);
  }
  DevBuf pts, tw;
  ZK_TRY(pts.alloc((size_t)N * ZW * 4));
  ZK_TRY(tw.alloc((size_t)(halfN ? halfN : 1) * 32));
  auto to_fr = [&](const zkh::Fe &x) {
    Fe<FR> d;
    for (int j = 0; j < 8; j++) d.w[j] = (uint32_t)(x.w[j / 2] >> (32 * (j & 1)));
    return d;
  };
  const dim3 blk(kG1Threads), gN((N + kG1Threads - 1) / kG1Threads), gH((halfN + kG1Threads - 1) / kG1Threads);
  if (halfN) hipLaunchKernelGGL(g1_twiddle_kernel<C>, gH, blk, 0, st, tw.as<uint32_t>(), to_fr(g), halfN);
  if (jac)
    hipLaunchKernelGGL((g1_proj_in_kernel<C, true>), gN, blk, 0, st, (const uint32_t *)d_src, N, m, inverse ? 0 : 1,
                       pts.as<int32_t>());
  else
    hipLaunchKernelGGL((g1_proj_in_kernel<C, false>), gN, blk, 0, st, (const uint32_t *)d_src, N, m, inverse ? 0 : 1,
                       pts.as<int32_t>());
  if (!inverse) {
    for (int s = 1; s <= m; s++)
      hipLaunchKernelGGL((g1_fft_stage_kernel<C, false>), gH, blk, 0, st, pts.as<int32_t>(), tw.as<uint32_t>(), m, s);
  } else {
    for (int s = m; s >= 1; s--)
      hipLaunchKernelGGL((g1_fft_stage_kernel<C, true>), gH, blk, 0, st, pts.as<int32_t>(), tw.as<uint32_t>(), m, s);
    if (m > 0) {
      const zkh::Fe ninv = zkh::to_std(fr, zkh::inv(fr, zkh::from_u64(fr, N)));
      hipLaunchKernelGGL(g1_scale_kernel<C>, gN, blk, 0, st, pts.as<int32_t>(), N, to_fr(ninv));
    }
  }
  hipLaunchKernelGGL(g1_proj_out_kernel<C>, gN, blk, 0, st, pts.as<int32_t>(), N, m, inverse ? 1 : 0,
                     (uint32_t *)d_dst);
  ZK_HIP(hipGetLastError());
  ZK_HIP(hipStreamSynchronize(st));  // pts / tw are released on return
  return ZK_OK;
}

int g1_fft_device(int curve, int m, const uint64_t *gen, bool inverse, bool jac, const uint64_t *d_src,
                  uint64_t *d_dst, hipStream_t st) {
  if (m < 0 || m > 26) {
    set_error("g1_fft: log size out of range");
    return ZK_EINVAL;
  }
  if (curve == 0) return g1_fft_run<Bn254G1>(m, gen, inverse, jac, d_src, d_dst, st);
  if (curve == 1) return g1_fft_run<Bls381G1>(m, gen, inverse, jac, d_src, d_dst, st);
  set_error("g1_fft: bad curve id");
  return ZK_EINVAL;
}

}  // namespace zk
