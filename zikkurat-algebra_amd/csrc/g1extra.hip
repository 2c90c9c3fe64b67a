// G1 helpers around the MSM: batch conversion between projective / Jacobian and affine form,
// and binary (0/1 coefficient) MSMs.
//
// batch_to_affine (lib/cbits/curves/g1/proj/bn128_G1_proj.c:157-166; Jacobian
// lib/cbits/curves/g1/jac/bn128_G1_jac.c:150-158): (X/Z, Y/Z), resp. (X/Z^2, Y/Z^3), per
// point, Z = 0 giving the all-0xFF string (proj.c:132-144, jac.c:120-137).  The reference
// inverts every Z separately; here each lane inverts one product of 8 Z values (Montgomery's
// trick), so a point costs five field products plus 1/8 of an inversion (by division steps;
// 8 points per lane measured faster than 4 or 16).
// batch_from_affine (proj.c:146-155, jac.c:139-148): (x, y, 1); the affine infinity (0, 0)
// gives (0, 1, 0), resp. [1 : 1 : 0].
// MSM_binary (proj.c:684-782, jac.c:726-830): the sum of the points whose byte coefficient is
// non-zero, one such sum per row of a byte matrix for the multi form.  A bucket method has
// nothing to share here (one bucket), so the device sums directly: every wavefront scans its
// slice of a row 64 bytes at a time, appends the selected indices to a queue in LDS (ballot +
// popcount) and, whenever 64 are queued, adds one selected point per lane into its XYZZ
// accumulator with a mixed addition, so that no lane idles on unselected points; the
// accumulators of a block, then the blocks of a row, are summed by pairwise trees.
#include <string.h>

#include <vector>

#include "curve.hpp"
#include "curvez.hpp"
#include "field.hpp"
#include "msm.hpp"
#include "runtime.hpp"

namespace zk {

static constexpr int kBatchThreads = 128;
static constexpr int kInvChunk = 8;  // points per lane and inversion

// Montgomery's trick per lane over kInvChunk points (coalesced: lane l of the wavefront takes
// points wave_base + k * 64 + l).  The external words are used as they are: unpacking the
// Montgomery words X R of X gives the internal element X R / R', and for internal factors
// f = F R' the product zmul(X R, f) = X F R is again external.  So the inverses d_k of the
// unpacked Z_k R, scaled once per lane by R / R' (the unpacked Montgomery one R), are the
// internal 1/Z_k, and x = zmul(unpack(X), d_k) needs no conversion on either side.
template <class C, bool JAC>
__global__ __launch_bounds__(kBatchThreads) void g1_batch_to_affine_kernel(const uint32_t *__restrict__ src, uint32_t n,
                                                                          uint32_t *__restrict__ dst) {
  using F = typename C::Fp;
  constexpr int LW = F::N;
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t wave_base = ((uint64_t)blockIdx.x * kBatchThreads + (threadIdx.x & ~63u)) * kInvChunk;
  if (wave_base + lane >= n) return;  // lanes without any point
  // the prefix products Z_0 ... Z_k wait in the output slot of point k (at least 2 LW words,
  // more than one internal element), which the backward pass overwrites only after reading
  FzP<C> acc, u;
  zset_one(acc);
  uint32_t inf = 0;
#pragma unroll 1
  for (int k = 0; k < kInvChunk; k++) {
    const uint64_t i = wave_base + (uint64_t)k * 64 + lane;
    if (i >= n) break;
    const uint32_t *p = src + i * 3 * LW + 2 * LW;
    uint32_t o = 0;
#pragma unroll
    for (int j = 0; j < LW; j++) o |= p[j];
    if (o) {
      zunpack(u, p);
      zmul(acc, acc, u);
    } else {
      inf |= 1u << k;
    }
    zstore((int32_t *)(dst + i * 2 * LW), acc);
  }
  FzP<C> inv, r;
  zinv(inv, acc);
  zunpack(r, F::ONE);
  zmul(inv, inv, r);  // the scale R / R'
#pragma unroll 1
  for (int k = kInvChunk - 1; k >= 0; k--) {
    const uint64_t i = wave_base + (uint64_t)k * 64 + lane;
    if (i >= n) continue;
    uint32_t *o = dst + i * 2 * LW;
    if ((inf >> k) & 1u) {
#pragma unroll
      for (int j = 0; j < 2 * LW; j++) o[j] = 0xffffffffu;
      continue;
    }
    FzP<C> zi;
    if (k > 0) {
      FzP<C> prev;
      zload(prev, (const int32_t *)(dst + (i - 64) * 2 * LW));
      zmul(zi, inv, prev);  // 1 / Z_k
      zunpack(u, src + i * 3 * LW + 2 * LW);
      zmul(inv, inv, u);
    } else {
      zi = inv;
    }
    FzP<C> x, y;
    zunpack(x, src + i * 3 * LW);
    zunpack(y, src + i * 3 * LW + LW);
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
    zcanon(x, x);
    zcanon(y, y);
    zpack(o, x);
    zpack(o + LW, y);
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

static constexpr int kSumThreads = 128;    // two wavefronts per block
static constexpr int kSumMinPerBlock = 1024;  // points of one row per block (at least)
static constexpr int kSumMaxBlocks = 256;     // blocks per row (at most): partials to reduce
static constexpr int kSumWaves = kSumThreads / 64;

template <class C>
__device__ __forceinline__ void tree_sum_lds(int32_t *red, XyzzZ<C> &acc, uint32_t tid) {
  constexpr int PW = ZLayout<C>::XYZZ;
  zxyzz_store(red + (size_t)tid * PW, acc);
  __syncthreads();
#pragma unroll 1
  for (uint32_t s = kSumThreads / 2; s > 0; s >>= 1) {
    if (tid < s) {
      XyzzZ<C> b;
      zxyzz_load(b, red + (size_t)(tid + s) * PW);
      zxyzz_add(acc, acc, b);
      zxyzz_store(red + (size_t)tid * PW, acc);
    }
    __syncthreads();
  }
}

// partial[(row, block)] = sum of the selected points of the block's range of the row;
// pts: prepared affine records (ZLayout<C>::AFF words each, STRIDE records per point)
template <class C>
__global__ __launch_bounds__(kSumThreads) void binary_sum_kernel(const uint8_t *__restrict__ bits, uint32_t n,
                                                                uint32_t per_block, const int32_t *__restrict__ pts,
                                                                int32_t *__restrict__ partial) {
  constexpr int AW = ZLayout<C>::AFF * ZLayout<C>::STRIDE, PW = ZLayout<C>::XYZZ;
  const uint32_t kPerWave = per_block / kSumWaves;  // a multiple of 64
  __shared__ uint32_t queue[kSumWaves][128];
  __shared__ int32_t red[kSumThreads * PW];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint8_t *rb = bits + (size_t)blockIdx.y * n;
  const uint64_t w0 = (uint64_t)blockIdx.x * per_block + (uint64_t)wave * kPerWave;
  const uint64_t w1 = w0 + kPerWave < n ? w0 + kPerWave : n;
  const uint64_t below = (1ull << lane) - 1;
  uint32_t *q = queue[wave];
  uint32_t count = 0;
  XyzzZ<C> acc;
  zxyzz_set_inf(acc);
  for (uint64_t base = w0; base < w1; base += 64) {
    const uint64_t i = base + lane;
    const bool sel = i < w1 && rb[i] != 0;
    const uint64_t mask = __ballot(sel);
    if (sel) q[count + __popcll(mask & below)] = (uint32_t)i;
    count += __popcll(mask);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (count >= 64) {  // uniform across the wavefront
      count -= 64;
      const uint32_t idx = q[count + lane];
      AffZ<C> a;
      zaff_load(a, pts + (size_t)idx * AW);
      if (!zaff_is_inf(a)) zxyzz_madd(acc, a);
      __builtin_amdgcn_wave_barrier();
    }
  }
  if (lane < count) {
    AffZ<C> a;
    zaff_load(a, pts + (size_t)q[lane] * AW);
    if (!zaff_is_inf(a)) zxyzz_madd(acc, a);
  }
  tree_sum_lds<C>(red, acc, tid);
  if (tid == 0) {
    XyzzZ<C> r;
    zxyzz_load(r, red);
    zxyzz_store(partial + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * PW, r);
  }
}

// out[row] = sum of the row's cnt partials, as external words (X, Y, ZZ, ZZZ), zero for infinity
template <class C>
__global__ __launch_bounds__(kSumThreads) void binary_reduce_kernel(const int32_t *__restrict__ partial, uint32_t cnt,
                                                                   uint32_t *__restrict__ out) {
  using F = typename C::Fp;
  constexpr int PW = ZLayout<C>::XYZZ;
  __shared__ int32_t red[kSumThreads * PW];
  const uint32_t tid = threadIdx.x;
  const int32_t *pr = partial + (size_t)blockIdx.x * cnt * PW;
  XyzzZ<C> acc, b;
  zxyzz_set_inf(acc);
  for (uint32_t j = tid; j < cnt; j += kSumThreads) {
    zxyzz_load(b, pr + (size_t)j * PW);
    zxyzz_add(acc, acc, b);
  }
  tree_sum_lds<C>(red, acc, tid);
  if (tid == 0) {
    zxyzz_load(acc, red);
    uint32_t *o = out + (size_t)blockIdx.x * 4 * F::N;
    if (zxyzz_is_inf(acc)) {
      for (int j = 0; j < 4 * F::N; j++) o[j] = 0;
    } else {
      zto_ext(o, acc.x);
      zto_ext(o + F::N, acc.y);
      zto_ext(o + 2 * F::N, acc.zz);
      zto_ext(o + 3 * F::N, acc.zzz);
    }
  }
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
  const int lp = curve == 0 ? 4 : 6;
  const uintptr_t s0 = (uintptr_t)d_src, s1 = s0 + (size_t)n * 3 * lp * 8;
  const uintptr_t d0 = (uintptr_t)d_dst, d1 = d0 + (size_t)n * 2 * lp * 8;
  if (s0 < d1 && d0 < s1) {  // the output slots hold intermediate products
    set_error("batch_to_affine: the source and target buffers overlap");
    return ZK_EINVAL;
  }
  const uint64_t per_block = (uint64_t)kBatchThreads * kInvChunk;  // lanes stride by 64 within a wavefront
  const dim3 g((unsigned)((n + per_block - 1) / per_block)), b(kBatchThreads);
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

template <class C>
static int binary_sum_run(int nrows, int n, const uint8_t *d_bits, const uint64_t *d_points, bool points_internal,
                          hipStream_t st, MsmResult *out) {
  using F = typename C::Fp;
  constexpr int PW = ZLayout<C>::XYZZ;
  DevBuf prep, part, res;
  StreamSyncGuard drain(st);  // the buffers outlive the queued kernels on every return
  const int32_t *pts = (const int32_t *)d_points;
  if (!points_internal) {
    ZK_TRY(prep.alloc((size_t)n * msm_internal_point_bytes(C::id)));
    ZK_TRY(msm_prepare_points(C::id, n, d_points, prep.p, st, false));
    pts = prep.as<int32_t>();
  }
  // per_block: a multiple of 64 per wavefront, at least kSumMinPerBlock, at most kSumMaxBlocks
  // blocks per row (the final tree's depth)
  uint64_t per_block = ((uint64_t)n + kSumMaxBlocks - 1) / kSumMaxBlocks;
  per_block = (per_block + 64 * kSumWaves - 1) / (64 * kSumWaves) * (64 * kSumWaves);
  if (per_block < (uint64_t)kSumMinPerBlock) per_block = kSumMinPerBlock;
  const uint32_t bpr = (uint32_t)((n + per_block - 1) / per_block);
  const int kRowsPerLaunch = 32768;
  const int rows_now = nrows < kRowsPerLaunch ? nrows : kRowsPerLaunch;
  ZK_TRY(part.alloc((size_t)rows_now * bpr * PW * 4));
  ZK_TRY(res.alloc((size_t)nrows * 4 * F::N * 4));
  for (int r0 = 0; r0 < nrows; r0 += kRowsPerLaunch) {
    const int rows = nrows - r0 < kRowsPerLaunch ? nrows - r0 : kRowsPerLaunch;
    hipLaunchKernelGGL(binary_sum_kernel<C>, dim3(bpr, rows), dim3(kSumThreads), 0, st, d_bits + (size_t)r0 * n,
                       (uint32_t)n, (uint32_t)per_block, pts, part.as<int32_t>());
    hipLaunchKernelGGL(binary_reduce_kernel<C>, dim3(rows), dim3(kSumThreads), 0, st,
                       (const int32_t *)part.as<int32_t>(), bpr, res.as<uint32_t>() + (size_t)r0 * 4 * F::N);
    ZK_HIP(hipGetLastError());
  }
  std::vector<uint32_t> w((size_t)nrows * 4 * F::N);
  ZK_HIP(hipMemcpyAsync(w.data(), res.p, w.size() * 4, hipMemcpyDeviceToHost, st));
  ZK_HIP(hipStreamSynchronize(st));
  drain.release();
  for (int r = 0; r < nrows; r++) {
    MsmResult &o = out[r];
    memset(&o, 0, sizeof o);
    o.curve = C::id;
    uint64_t *dst[4] = {o.x, o.y, o.zz, o.zzz};
    for (int e = 0; e < 4; e++)
      for (int j = 0; j < F::N; j++)
        dst[e][j / 2] |= (uint64_t)w[((size_t)r * 4 + e) * F::N + j] << (32 * (j & 1));
  }
  return ZK_OK;
}

// sums of the points selected by each of nrows byte rows (row r at d_bits + r * n)
int binary_sum_device(int curve, int nrows, int n, const uint8_t *d_bits, const uint64_t *d_points,
                      bool points_internal, hipStream_t st, MsmResult *out) {
  if (nrows <= 0) return ZK_OK;
  if (n <= 0) {
    for (int r = 0; r < nrows; r++) msm_infinity(curve, &out[r]);
    return ZK_OK;
  }
  if (curve == 0) return binary_sum_run<Bn254G1>(nrows, n, d_bits, d_points, points_internal, st, out);
  if (curve == 1) return binary_sum_run<Bls381G1>(nrows, n, d_bits, d_points, points_internal, st, out);
  set_error("binary msm: bad curve id");
  return ZK_EINVAL;
}

}  // namespace zk
