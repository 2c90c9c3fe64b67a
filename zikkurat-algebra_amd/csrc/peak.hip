// Microbenchmarks that measure the rooflines the hot paths are measured against, and
// elementwise field kernels used to check the device field layer against the oracle.
//
//   imul   : v_mad_u64_u32 throughput (32x32->64 multiply-add), the MSM/NTT bound
//   copy   : HBM streaming copy bandwidth (bytes read + written)
//   fmul   : Montgomery multiplications per second for each field (CIOS on the VALU)
#include <stdint.h>

#include "field.hpp"
#include "peak.hpp"
#include "runtime.hpp"

namespace zk {

// 16 independent multiply-add chains per thread; each step is exactly one v_mad_u64_u32
__global__ __launch_bounds__(256) void imul_peak_kernel(const uint32_t *seed, uint64_t *sink, int iters) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t acc[16];
  const uint32_t b = seed[0] | 1u;
  const uint64_t c = ((uint64_t)seed[1] << 32) | seed[2];
#pragma unroll
  for (int i = 0; i < 16; i++) acc[i] = (uint64_t)(tid * 16 + i) * 0x9e3779b9u;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int i = 0; i < 16; i++) acc[i] = (uint64_t)(uint32_t)acc[i] * b + c;
  }
  uint64_t s = 0;
#pragma unroll
  for (int i = 0; i < 16; i++) s ^= acc[i];
  if (s == 0x123456789abcdefull) sink[tid] = s;  // practically never taken; keeps the work
}

__global__ __launch_bounds__(256) void copy_kernel(const uint4 *__restrict__ in, uint4 *__restrict__ out, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = in[i];
}

template <class F>
__global__ __launch_bounds__(256) void fmul_peak_kernel(const uint32_t *in, uint32_t *out, int iters) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  Fe<F> x0, x1, c;
#pragma unroll
  for (int j = 0; j < F::N; j++) {
    x0.w[j] = F::ONE[j] ^ (uint32_t)tid;
    x1.w[j] = F::ONE[j] + (uint32_t)tid;
    c.w[j] = in[j];
  }
  x0.w[F::N - 1] &= 0x0fffffffu;
  x1.w[F::N - 1] &= 0x0fffffffu;
  for (int it = 0; it < iters; it++) {
    fmul(x0, x0, c);
    fmul(x1, x1, c);
  }
  uint32_t s = 0;
#pragma unroll
  for (int j = 0; j < F::N; j++) s ^= x0.w[j] ^ x1.w[j];
  if (s == 0x9e3779b9u) out[tid] = s;
}

static int time_launch(hipStream_t st, float *ms, void (*launch)(hipStream_t, void *), void *arg) {
  hipEvent_t a, b;
  ZK_HIP(hipEventCreate(&a));
  ZK_HIP(hipEventCreate(&b));
  launch(st, arg);  // warm-up
  ZK_HIP(hipEventRecord(a, st));
  launch(st, arg);
  ZK_HIP(hipEventRecord(b, st));
  ZK_HIP(hipEventSynchronize(b));
  ZK_HIP(hipEventElapsedTime(ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  ZK_HIP(hipGetLastError());
  return ZK_OK;
}

struct ImulArg {
  uint32_t *seed;
  uint64_t *sink;
  int blocks, iters;
};
static void launch_imul(hipStream_t st, void *p) {
  ImulArg *a = (ImulArg *)p;
  hipLaunchKernelGGL(imul_peak_kernel, dim3(a->blocks), dim3(256), 0, st, a->seed, a->sink, a->iters);
}
struct CopyArg {
  uint4 *in, *out;
  size_t n;
  int blocks;
};
static void launch_copy(hipStream_t st, void *p) {
  CopyArg *a = (CopyArg *)p;
  hipLaunchKernelGGL(copy_kernel, dim3(a->blocks), dim3(256), 0, st, a->in, a->out, a->n);
}
struct FmulArg {
  uint32_t *in, *out;
  int blocks, iters, which;
};
static void launch_fmul(hipStream_t st, void *p) {
  FmulArg *a = (FmulArg *)p;
  switch (a->which) {
    case 0: hipLaunchKernelGGL(fmul_peak_kernel<Bn254Fp>, dim3(a->blocks), dim3(256), 0, st, a->in, a->out, a->iters); break;
    case 1: hipLaunchKernelGGL(fmul_peak_kernel<Bn254Fr>, dim3(a->blocks), dim3(256), 0, st, a->in, a->out, a->iters); break;
    case 2: hipLaunchKernelGGL(fmul_peak_kernel<Bls381Fp>, dim3(a->blocks), dim3(256), 0, st, a->in, a->out, a->iters); break;
    default: hipLaunchKernelGGL(fmul_peak_kernel<Bls381Fr>, dim3(a->blocks), dim3(256), 0, st, a->in, a->out, a->iters); break;
  }
}

int microbench(int which, double *result) {
  Context *ctx;
  ZK_TRY(context(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  hipStream_t st = ctx->stream;
  float ms = 0;
  if (which == MB_IMUL) {
    DevBuf seed, sink;
    ZK_TRY(seed.alloc(64));
    ZK_TRY(sink.alloc((size_t)ctx->num_cus * 32 * 256 * 8));
    uint32_t h[4] = {0x12345679u, 0xdeadbeefu, 0x01234567u, 0};
    ZK_HIP(hipMemcpy(seed.p, h, sizeof h, hipMemcpyHostToDevice));
    ImulArg a{seed.as<uint32_t>(), sink.as<uint64_t>(), ctx->num_cus * 32, 4096};
    ZK_TRY(time_launch(st, &ms, launch_imul, &a));
    *result = (double)a.blocks * 256 * a.iters * 16 / (ms * 1e-3);
    return ZK_OK;
  }
  if (which == MB_COPY) {
    const size_t bytes = (size_t)1 << 30;
    DevBuf in, out;
    ZK_TRY(in.alloc(bytes));
    ZK_TRY(out.alloc(bytes));
    ZK_HIP(hipMemsetAsync(in.p, 1, bytes, st));
    CopyArg a{in.as<uint4>(), out.as<uint4>(), bytes / 16, ctx->num_cus * 16};
    ZK_TRY(time_launch(st, &ms, launch_copy, &a));
    *result = 2.0 * bytes / (ms * 1e-3);
    return ZK_OK;
  }
  if (which >= MB_FMUL_BN254_FP && which <= MB_FMUL_BLS381_FR) {
    DevBuf in, out;
    ZK_TRY(in.alloc(64));
    ZK_TRY(out.alloc((size_t)ctx->num_cus * 16 * 256 * 4));
    uint32_t h[16];
    for (int i = 0; i < 16; i++) h[i] = 0x01010101u * (i + 3);
    h[7] = 0x01000000u;
    h[11] = 0x01000000u;
    ZK_HIP(hipMemcpy(in.p, h, sizeof h, hipMemcpyHostToDevice));
    FmulArg a{in.as<uint32_t>(), out.as<uint32_t>(), ctx->num_cus * 16, 256, which - MB_FMUL_BN254_FP};
    ZK_TRY(time_launch(st, &ms, launch_fmul, &a));
    *result = (double)a.blocks * 256 * a.iters * 2 / (ms * 1e-3);
    return ZK_OK;
  }
  set_error("microbench: unknown benchmark id");
  return ZK_EINVAL;
}

// ---------------------------------------------------------------------------------------
// elementwise field operations on device arrays (parity checks of the field layer)

template <class F>
__global__ void field_op_kernel(int op, size_t n, const uint32_t *a, const uint32_t *b, uint32_t *out) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Fe<F> x, y, r;
  fload(x, a + i * F::N);
  if (b) fload(y, b + i * F::N);
  switch (op) {
    case FOP_ADD: fadd(r, x, y); break;
    case FOP_SUB: fsub(r, x, y); break;
    case FOP_MUL: fmul(r, x, y); break;
    case FOP_SQR: fsqr(r, x); break;
    case FOP_NEG: fneg(r, x); break;
    case FOP_INV: finv(r, x); break;
    case FOP_HALF: fhalf(r, x); break;
    case FOP_TO_STD: fto_std(r, x); break;
    case FOP_FROM_STD: ffrom_std(r, x); break;
    default: r = x; break;
  }
  fstore(out + i * F::N, r);
}

// synthetic inputs: element i = splitmix64-style hash of (seed, i, word), top word masked so
// that the value is below 2^(bits-2) < p (canonical; any canonical word string is also a valid
// Montgomery-form element)
__global__ void fill_random_kernel(uint64_t *out, size_t n, int words, uint64_t seed, uint64_t top_mask) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int k = 0; k < words; k++) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (i * words + k + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    if (k == words - 1) z &= top_mask;
    out[i * words + k] = z;
  }
}

int fill_random_device(int curve, int field, size_t n, uint64_t seed, void *out) {
  Context *ctx;
  ZK_TRY(context(&ctx));
  const int words = (field == 0 && curve == 1) ? 6 : 4;
  // top word below the prime's top word: BN254 p, r < 0x30644e72..., BLS r < 0x73eda753...,
  // BLS p < 0x1a0111ea...
  const uint64_t mask = (curve == 0) ? 0x0fffffffffffffffull : (field == 0 ? 0x0fffffffffffffffull : 0x3fffffffffffffffull);
  if (!n) return ZK_OK;
  hipLaunchKernelGGL(fill_random_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                     (uint64_t *)out, n, words, seed, mask);
  ZK_HIP(hipGetLastError());
  ZK_HIP(hipStreamSynchronize(ctx->stream));
  return ZK_OK;
}

int field_op_device(int curve, int field, int op, size_t n, const void *a, const void *b, void *out) {
  Context *ctx;
  ZK_TRY(context(&ctx));
  std::lock_guard<std::recursive_mutex> lk(ctx->lock);
  hipStream_t st = ctx->stream;
  const dim3 grid((unsigned)((n + 255) / 256)), blk(256);
  const uint32_t *pa = (const uint32_t *)a, *pb = (const uint32_t *)b;
  uint32_t *po = (uint32_t *)out;
  if (!n) return ZK_OK;
  if (curve == 0 && field == 0) hipLaunchKernelGGL(field_op_kernel<Bn254Fp>, grid, blk, 0, st, op, n, pa, pb, po);
  else if (curve == 0 && field == 1) hipLaunchKernelGGL(field_op_kernel<Bn254Fr>, grid, blk, 0, st, op, n, pa, pb, po);
  else if (curve == 1 && field == 0) hipLaunchKernelGGL(field_op_kernel<Bls381Fp>, grid, blk, 0, st, op, n, pa, pb, po);
  else if (curve == 1 && field == 1) hipLaunchKernelGGL(field_op_kernel<Bls381Fr>, grid, blk, 0, st, op, n, pa, pb, po);
  else {
    set_error("field_op: bad curve/field");
    return ZK_EINVAL;
  }
  ZK_HIP(hipGetLastError());
  ZK_HIP(hipStreamSynchronize(st));
  return ZK_OK;
}

}  // namespace zk
