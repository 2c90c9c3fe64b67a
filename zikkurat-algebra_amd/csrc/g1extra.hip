// G1 helpers around the MSM: batch conversion between projective / Jacobian and affine form,
// and binary (0/1 coefficient) MSMs.
//
// batch_to_affine (lib/cbits/curves/g1/proj/bn128_G1_proj.c:157-166; Jacobian
// lib/cbits/curves/g1/jac/bn128_G1_jac.c:150-158): (X/Z, Y/Z), resp. (X/Z^2, Y/Z^3), per
// point, Z = 0 giving the all-0xFF string (proj.c:132-144, jac.c:120-137).  The reference
// inverts every Z separately; here each thread inverts one product of 16 Z values
// (Montgomery's trick), so a point costs about a dozen field products plus 1/16 of an
// inversion.
// batch_from_affine (proj.c:146-155, jac.c:139-148): (x, y, 1); the affine infinity (0, 0)
// gives (0, 1, 0), resp. [1 : 1 : 0].
// MSM_binary (proj.c:684-782, jac.c:726-830): the sum of the points whose byte coefficient is
// non-zero; here the bytes become 0/1 scalars for the bucket MSM, whose statistics pass sees
// one-bit scalars and lays out a single window, so all selected points fall into one bucket
// summed by the load-balanced accumulation and pairwise trees.
#include <string.h>

#include <vector>

#include "curve.hpp"
#include "curvez.hpp"
#include "field.hpp"
#include "msm.hpp"
#include "runtime.hpp"

namespace zk {

static constexpr int kBatchThreads = 128;
static constexpr int kInvChunk = 16;

template <class C, bool JAC>
__global__ __launch_bounds__(kBatchThreads) void g1_batch_to_affine_kernel(const uint32_t *__restrict__ src, uint32_t n,
                                                                          uint32_t *__restrict__ dst) {
  using F = typename C::Fp;
  constexpr int LW = F::N;
  const uint32_t i0 = (blockIdx.x * blockDim.x + threadIdx.x) * kInvChunk;
  if (i0 >= n) return;
  const int cnt = (int)(n - i0 < (uint32_t)kInvChunk ? n - i0 : (uint32_t)kInvChunk);
  FzP<C> pre[kInvChunk], acc, z;
  zset_one(acc);
  uint32_t inf = 0;
  auto load_z = [&](int k, FzP<C> &out) {
    const uint32_t *p = src + (size_t)(i0 + k) * 3 * LW + 2 * LW;
    uint32_t o = 0;
#pragma unroll
    for (int j = 0; j < LW; j++) o |= p[j];
    if (!o) {
      zset_one(out);
      return true;
    }
    zfrom_ext(out, p);
    return false;
  };
#pragma unroll
  for (int k = 0; k < kInvChunk; k++) {
    if (k < cnt) {
      if (load_z(k, z)) inf |= 1u << k;
      zmul(acc, acc, z);
    }
    pre[k] = acc;
  }
  FzP<C> inv;
  zinv(inv, acc);  // 1 / (Z_0 ... Z_{cnt-1})
#pragma unroll
  for (int k = kInvChunk - 1; k >= 0; k--) {
    if (k >= cnt) continue;
    FzP<C> zi;
    if (k > 0) zmul(zi, inv, pre[k - 1]);  // 1 / Z_k
    else zi = inv;
    if (k > 0) {
      load_z(k, z);
      zmul(inv, inv, z);
    }
    uint32_t *o = dst + (size_t)(i0 + k) * 2 * LW;
    if ((inf >> k) & 1u) {
#pragma unroll
      for (int j = 0; j < 2 * LW; j++) o[j] = 0xffffffffu;
      continue;
    }
    const uint32_t *p = src + (size_t)(i0 + k) * 3 * LW;
    FzP<C> x, y;
    zfrom_ext(x, p);
    zfrom_ext(y, p + LW);
    if (JAC) {
      FzP<C> zi2;
      zsqr(zi2, zi);
      zmul(x, x, zi2);
      zmul(zi2, zi2, zi);
      zmul(y, y, zi2);
    } else {
      zmul(x, x, zi);
      zmul(y, y, zi);
    }
    zto_ext(o, x);
    zto_ext(o + LW, y);
  }
}

template <class C, bool JAC>
__global__ __launch_bounds__(kBatchThreads) void g1_batch_from_affine_kernel(const uint32_t *__restrict__ src,
                                                                            uint32_t n, uint32_t *__restrict__ dst) {
  using F = typename C::Fp;
  constexpr int LW = F::N;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t *p = src + (size_t)i * 2 * LW;
  uint32_t *o = dst + (size_t)i * 3 * LW;
  uint32_t any = 0;
#pragma unroll
  for (int j = 0; j < 2 * LW; j++) any |= p[j];
  FzP<C> one;
  zset_one(one);
  if (!any) {  // infinity: (0 : 1 : 0), Jacobian [1 : 1 : 0]
    if (JAC) {
      zto_ext(o, one);
    } else {
#pragma unroll
      for (int j = 0; j < LW; j++) o[j] = 0;
    }
    zto_ext(o + LW, one);
#pragma unroll
    for (int j = 0; j < LW; j++) o[2 * LW + j] = 0;
    return;
  }
#pragma unroll
  for (int j = 0; j < 2 * LW; j++) o[j] = p[j];
  zto_ext(o + 2 * LW, one);
}

__global__ void binary_scalars_kernel(const uint8_t *__restrict__ bits, uint32_t n, uint64_t *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = bits[i] ? 1ull : 0ull;
}

template <class C, bool JAC>
static void launch_to_affine(dim3 g, dim3 b, hipStream_t st, const uint64_t *src, int n, uint64_t *dst) {
  hipLaunchKernelGGL((g1_batch_to_affine_kernel<C, JAC>), g, b, 0, st, (const uint32_t *)src, (uint32_t)n,
                     (uint32_t *)dst);
}
template <class C, bool JAC>
static void launch_from_affine(dim3 g, dim3 b, hipStream_t st, const uint64_t *src, int n, uint64_t *dst) {
  hipLaunchKernelGGL((g1_batch_from_affine_kernel<C, JAC>), g, b, 0, st, (const uint32_t *)src, (uint32_t)n,
                     (uint32_t *)dst);
}

int g1_batch_to_affine(int curve, bool jac, int n, const uint64_t *d_src, uint64_t *d_dst, hipStream_t st) {
  if (n <= 0) return ZK_OK;
  const uint32_t threads = (uint32_t)((n + kInvChunk - 1) / kInvChunk);
  const dim3 g((threads + kBatchThreads - 1) / kBatchThreads), b(kBatchThreads);
  if (curve == 0)
    (jac ? launch_to_affine<Bn254G1, true> : launch_to_affine<Bn254G1, false>)(g, b, st, d_src, n, d_dst);
  else if (curve == 1)
    (jac ? launch_to_affine<Bls381G1, true> : launch_to_affine<Bls381G1, false>)(g, b, st, d_src, n, d_dst);
  else {
    set_error("batch_to_affine: bad curve id");
    return ZK_EINVAL;
  }
  ZK_HIP(hipGetLastError());
  ZK_HIP(hipStreamSynchronize(st));
  return ZK_OK;
}

int g1_batch_from_affine(int curve, bool jac, int n, const uint64_t *d_src, uint64_t *d_dst, hipStream_t st) {
  if (n <= 0) return ZK_OK;
  const dim3 g((n + kBatchThreads - 1) / kBatchThreads), b(kBatchThreads);
  if (curve == 0)
    (jac ? launch_from_affine<Bn254G1, true> : launch_from_affine<Bn254G1, false>)(g, b, st, d_src, n, d_dst);
  else if (curve == 1)
    (jac ? launch_from_affine<Bls381G1, true> : launch_from_affine<Bls381G1, false>)(g, b, st, d_src, n, d_dst);
  else {
    set_error("batch_from_affine: bad curve id");
    return ZK_EINVAL;
  }
  ZK_HIP(hipGetLastError());
  ZK_HIP(hipStreamSynchronize(st));
  return ZK_OK;
}

int binary_scalars(int n, const uint8_t *d_bits, uint64_t *d_scalars, hipStream_t st) {
  if (n <= 0) return ZK_OK;
  hipLaunchKernelGGL(binary_scalars_kernel, dim3((n + 255) / 256), dim3(256), 0, st, d_bits, (uint32_t)n, d_scalars);
  ZK_HIP(hipGetLastError());
  return ZK_OK;
}

// sum of the points whose coefficient byte is non-zero
int msm_binary_device(int curve, int n, const uint8_t *d_bits, const uint64_t *d_points, bool points_internal,
                      hipStream_t st, MsmResult *out) {
  if (n <= 0) {
    msm_infinity(curve, out);
    return ZK_OK;
  }
  DevBuf sc;
  ZK_TRY(sc.alloc((size_t)n * 8));
  ZK_TRY(binary_scalars(n, d_bits, sc.as<uint64_t>(), st));
  zkg_msm_schedule s;
  memset(&s, 0, sizeof s);
  s.policy = ZKG_SCHED_STATS;
  return msm_device(curve, n, sc.as<uint64_t>(), 1, 0, d_points, points_internal, &s, st, out);
}

}  // namespace zk
