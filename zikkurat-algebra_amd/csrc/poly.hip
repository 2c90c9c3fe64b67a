// Polynomial operations of a KZG opening on the device: evaluation at a point and division by
// a coset vanishing polynomial x^n - eta, over the scalar fields of BN254 / BLS12-381.
// Coefficients are the reference's flat Montgomery Fr words (lib/cbits/curves/poly/mont).
//
// eval_at (bn128_poly_mont.c:225-243): sum_i a_i x^i.  Each thread folds 16 consecutive
// coefficients by Horner's rule; a block then folds its 256 partial values pairwise, level l
// joining neighbours 16 * 2^l coefficients apart with the factor x^(16 * 2^l); the per-block
// values form the next, 4096 times shorter, level with base x^4096.  Any order of the sum
// gives the same canonical field element.
//
// div_by_vanishing (bn128_poly_mont.c:317-395): with D = deg p >= n the quotient obeys
// q_j = a_{j+n} + eta q_{j+n} (j = D-n .. 0, q_j = a_{j+n} at the top) and the remainder is
// r_j = a_j + eta q_j (j <= D-n), r_j = a_j otherwise.  The recurrence runs independently
// along each residue class j mod n; it is a linear suffix scan y_j = b_j + mu y_{j+n}, done
// here in chunks of 16 steps per thread: a first pass folds every chunk with a zero carry, the
// chunk values obey the same scan with stride n and factor mu^16 (solved recursively), and a
// final pass re-runs every chunk from its true carry.  The KZG opening quotient (p - p(z)) /
// (x - z) is the case n = 1, eta = z, where the constant term plays no part.
#include <string.h>

#include <vector>

#include "curve.hpp"
#include "field.hpp"
#include "hostfield.hpp"
#include "msm.hpp"
#include "runtime.hpp"

namespace zk {

static constexpr int kFold = 16;        // coefficients folded per thread
static constexpr int kPolyThreads = 256;
static constexpr int kEvalLevels = 8;   // log2(kPolyThreads)

template <class FR>
struct EvalPows {
  Fe<FR> mu;               // base of this level
  Fe<FR> t[kEvalLevels];   // t[l] = mu^(kFold * 2^l)
};

// out[block] = sum_{i in block} v_i mu^(i - block start), blocks of kFold * 256 values
template <class FR>
__global__ __launch_bounds__(kPolyThreads) void poly_eval_kernel(const uint32_t *__restrict__ v, uint32_t M,
                                                                EvalPows<FR> P, uint32_t *__restrict__ out) {
  __shared__ Fe<FR> sh[kPolyThreads];
  const uint32_t tid = threadIdx.x;
  const uint64_t base = ((uint64_t)blockIdx.x * kPolyThreads + tid) * kFold;
  Fe<FR> h, x;
  fset_zero(h);
#pragma unroll 4
  for (int k = kFold - 1; k >= 0; k--) {
    const uint64_t j = base + k;
    fmul(h, h, P.mu);
    if (j < M) {
      fload(x, v + j * FR::N);
      fadd(h, h, x);
    }
  }
  sh[tid] = h;
  __syncthreads();
#pragma unroll
  for (int l = 0; l < kEvalLevels; l++) {
    const uint32_t stride = 1u << l;
    if ((tid & (2 * stride - 1)) == 0) {
      fmul(x, sh[tid + stride], P.t[l]);
      fadd(h, sh[tid], x);
      sh[tid] = h;
    }
    __syncthreads();
  }
  if (tid == 0) fstore(out + (size_t)blockIdx.x * FR::N, sh[0]);
}

// chunk folds with a zero carry: agg[c n + r] = sum_{k<16} mu^k b_{r + n (16c + k)} (j < M)
template <class FR>
__global__ __launch_bounds__(kPolyThreads) void vanish_fold_kernel(const uint32_t *__restrict__ b, uint32_t M,
                                                                  uint32_t n, uint32_t C, Fe<FR> mu,
                                                                  uint32_t *__restrict__ agg) {
  const uint64_t t = (uint64_t)blockIdx.x * kPolyThreads + threadIdx.x;
  if (t >= (uint64_t)n * C) return;
  const uint64_t r = t % n, c = t / n;
  Fe<FR> y, x;
  fset_zero(y);
  for (int k = kFold - 1; k >= 0; k--) {
    const uint64_t j = r + (uint64_t)n * (c * kFold + k);
    if (j < M) {
      fload(x, b + j * FR::N);
      fmul(y, y, mu);
      fadd(y, y, x);
    }
  }
  fstore(agg + t * FR::N, y);
}

// y_j = b_j + mu y_{j+n} over each chunk, starting from the chunk above (carry[(c+1) n + r],
// the solved chunk scan) or zero at the top; out may alias b (each thread reads an element
// before it writes it)
template <class FR>
__global__ __launch_bounds__(kPolyThreads) void vanish_scan_kernel(const uint32_t *b, uint32_t M, uint32_t n,
                                                                  uint32_t C, Fe<FR> mu,
                                                                  const uint32_t *__restrict__ carry, uint32_t *out) {
  const uint64_t t = (uint64_t)blockIdx.x * kPolyThreads + threadIdx.x;
  if (t >= (uint64_t)n * C) return;
  const uint64_t r = t % n, c = t / n;
  Fe<FR> y, x;
  if (carry && c + 1 < C) fload(y, carry + ((c + 1) * n + r) * FR::N);
  else fset_zero(y);
  for (int k = kFold - 1; k >= 0; k--) {
    const uint64_t j = r + (uint64_t)n * (c * kFold + k);
    if (j < M) {
      fload(x, b + j * FR::N);
      fmul(y, y, mu);
      fadd(y, y, x);
      fstore(out + j * FR::N, y);
    }
  }
}

// r_j = a_j + eta q_j for j < M (= D - n + 1), r_j = a_j for M <= j < n
template <class FR>
__global__ __launch_bounds__(kPolyThreads) void vanish_rem_kernel(const uint32_t *__restrict__ a, uint32_t n,
                                                                 uint32_t M, const uint32_t *__restrict__ q,
                                                                 Fe<FR> eta, uint32_t *__restrict__ rem) {
  const uint32_t j = blockIdx.x * kPolyThreads + threadIdx.x;
  if (j >= n) return;
  Fe<FR> x, y;
  fload(x, a + (size_t)j * FR::N);
  if (j < M) {
    fload(y, q + (size_t)j * FR::N);
    fmul(y, y, eta);
    fadd(x, x, y);
  }
  fstore(rem + (size_t)j * FR::N, x);
}

// deg = max { i : a_i != 0 } (-1 for the zero polynomial): one atomic per block
__global__ __launch_bounds__(kPolyThreads) void poly_degree_kernel(const uint4 *__restrict__ a, uint32_t n,
                                                                  int *__restrict__ deg) {
  __shared__ int wave_best[kPolyThreads / 64];
  int best = -1;
  for (uint32_t i = blockIdx.x * kPolyThreads + threadIdx.x; i < n; i += gridDim.x * kPolyThreads) {
    const uint4 u = a[2 * (size_t)i], w = a[2 * (size_t)i + 1];
    if (u.x | u.y | u.z | u.w | w.x | w.y | w.z | w.w) best = (int)i;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(best, off);
    best = o > best ? o : best;
  }
  if ((threadIdx.x & 63) == 0) wave_best[threadIdx.x >> 6] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kPolyThreads / 64; k++) best = wave_best[k] > best ? wave_best[k] : best;
    if (best >= 0) atomicMax(deg, best);
  }
}

__global__ void poly_degree_init_kernel(int *deg) {
  if (threadIdx.x == 0) *deg = -1;
}

// ------------------------------------------------------------------ host side

static const zkh::FieldDesc &fr_desc(int curve) { return curve == 0 ? zkh::BN254_FR : zkh::BLS381_FR; }

template <class FR>
static Fe<FR> to_dev(const zkh::Fe &x) {
  Fe<FR> d;
  for (int j = 0; j < FR::N; j++) d.w[j] = (uint32_t)(x.w[j / 2] >> (32 * (j & 1)));
  return d;
}

static unsigned grid_for(uint64_t threads) { return (unsigned)((threads + kPolyThreads - 1) / kPolyThreads); }

template <class C>
static int poly_eval_run(int n, const uint64_t *d_coeffs, const uint64_t *loc, uint64_t *out_host, hipStream_t st) {
  using FR = typename C::Fr;
  const zkh::FieldDesc &f = fr_desc(C::id);
  if (n <= 0) {
    memset(out_host, 0, 32);
    return ZK_OK;
  }
  const uint64_t per_block = (uint64_t)kFold * kPolyThreads;
  const uint32_t blocks0 = (uint32_t)((n + per_block - 1) / per_block);
  DevBuf a, b;
  ZK_TRY(a.alloc((size_t)blocks0 * 32));
  ZK_TRY(b.alloc((size_t)((blocks0 + per_block - 1) / per_block) * 32 + 32));
  zkh::Fe mu = zkh::load(f, loc);
  const uint32_t *src = (const uint32_t *)d_coeffs;
  uint32_t M = (uint32_t)n;
  uint32_t *bufs[2] = {a.as<uint32_t>(), b.as<uint32_t>()};
  int k = 0;
  for (;;) {
    EvalPows<FR> P;
    P.mu = to_dev<FR>(mu);
    zkh::Fe t = zkh::pow_u64(f, mu, kFold);
    for (int l = 0; l < kEvalLevels; l++) {
      P.t[l] = to_dev<FR>(t);
      t = zkh::sqr(f, t);
    }
    const uint32_t blocks = (uint32_t)((M + per_block - 1) / per_block);
    hipLaunchKernelGGL(poly_eval_kernel<FR>, dim3(blocks), dim3(kPolyThreads), 0, st, src, M, P, bufs[k]);
    ZK_HIP(hipGetLastError());
    src = bufs[k];
    k ^= 1;
    M = blocks;
    mu = t;  // mu^(kFold * 256)
    if (M == 1) break;
  }
  ZK_HIP(hipMemcpyAsync(out_host, src, 32, hipMemcpyDeviceToHost, st));
  ZK_HIP(hipStreamSynchronize(st));
  return ZK_OK;
}

int poly_eval_device(int curve, int n, const uint64_t *d_coeffs, const uint64_t *loc, uint64_t *out_host,
                     hipStream_t st) {
  if (curve == 0) return poly_eval_run<Bn254G1>(n, d_coeffs, loc, out_host, st);
  if (curve == 1) return poly_eval_run<Bls381G1>(n, d_coeffs, loc, out_host, st);
  set_error("poly_eval: bad curve id");
  return ZK_EINVAL;
}

int poly_degree_device(int n, const uint64_t *d_coeffs, int *deg_host, hipStream_t st) {
  if (n <= 0) {
    *deg_host = -1;
    return ZK_OK;
  }
  uint64_t top[4];  // the usual case: the top coefficient is non-zero
  ZK_HIP(hipMemcpyAsync(top, d_coeffs + (size_t)(n - 1) * 4, 32, hipMemcpyDeviceToHost, st));
  ZK_HIP(hipStreamSynchronize(st));
  if (top[0] | top[1] | top[2] | top[3]) {
    *deg_host = n - 1;
    return ZK_OK;
  }
  DevBuf d;
  ZK_TRY(d.alloc(4));
  hipLaunchKernelGGL(poly_degree_init_kernel, dim3(1), dim3(64), 0, st, d.as<int>());
  const unsigned g = grid_for((uint64_t)n) < 1024 ? grid_for((uint64_t)n) : 1024;
  hipLaunchKernelGGL(poly_degree_kernel, dim3(g), dim3(kPolyThreads), 0, st, (const uint4 *)d_coeffs, (uint32_t)n,
                     d.as<int>());
  ZK_HIP(hipGetLastError());
  ZK_HIP(hipMemcpyAsync(deg_host, d.p, 4, hipMemcpyDeviceToHost, st));
  ZK_HIP(hipStreamSynchronize(st));
  return ZK_OK;
}

// y_j = b_j + mu y_{j+n} for j < M (y_j = b_j where j + n >= M), written to out (may alias b)
template <class FR>
static int strided_scan(const uint32_t *b, uint32_t M, uint32_t n, const zkh::Fe &mu, const zkh::FieldDesc &f,
                        uint32_t *out, hipStream_t st) {
  if (M == 0) return ZK_OK;
  const uint64_t len0 = ((uint64_t)M + n - 1) / n;          // longest residue class
  const uint64_t C = (len0 + kFold - 1) / kFold;            // chunks per class
  const uint64_t threads = (uint64_t)n * C;
  if (C == 1) {
    hipLaunchKernelGGL(vanish_scan_kernel<FR>, dim3(grid_for(threads)), dim3(kPolyThreads), 0, st, b, M, n,
                       (uint32_t)C, to_dev<FR>(mu), (const uint32_t *)nullptr, out);
    ZK_HIP(hipGetLastError());
    return ZK_OK;
  }
  if (threads > 0xffffffffull) {
    set_error("div_by_vanishing: polynomial too large");
    return ZK_EINVAL;
  }
  DevBuf agg;
  ZK_TRY(agg.alloc((size_t)threads * 32));
  hipLaunchKernelGGL(vanish_fold_kernel<FR>, dim3(grid_for(threads)), dim3(kPolyThreads), 0, st, b, M, n,
                     (uint32_t)C, to_dev<FR>(mu), agg.as<uint32_t>());
  ZK_HIP(hipGetLastError());
  // the chunk values obey the same scan: Y_{c n + r} = agg_{c n + r} + mu^16 Y_{(c+1) n + r}
  const zkh::Fe mu16 = zkh::pow_u64(f, mu, kFold);
  ZK_TRY(strided_scan<FR>(agg.as<uint32_t>(), (uint32_t)threads, n, mu16, f, agg.as<uint32_t>(), st));
  hipLaunchKernelGGL(vanish_scan_kernel<FR>, dim3(grid_for(threads)), dim3(kPolyThreads), 0, st, b, M, n,
                     (uint32_t)C, to_dev<FR>(mu), (const uint32_t *)agg.as<uint32_t>(), out);
  ZK_HIP(hipGetLastError());
  ZK_HIP(hipStreamSynchronize(st));  // agg is released on return
  return ZK_OK;
}

template <class C>
static int vanish_run(int D, const uint64_t *d_a, int n, const uint64_t *eta, uint64_t *d_quot, uint64_t *d_rem,
                      hipStream_t st) {
  using FR = typename C::Fr;
  const zkh::FieldDesc &f = fr_desc(C::id);
  const uint32_t M = (uint32_t)(D - n + 1);
  const zkh::Fe e = zkh::load(f, eta);
  ZK_TRY(strided_scan<FR>((const uint32_t *)(d_a + (size_t)n * 4), M, (uint32_t)n, e, f, (uint32_t *)d_quot, st));
  if (d_rem) {
    hipLaunchKernelGGL(vanish_rem_kernel<FR>, dim3(grid_for((uint64_t)n)), dim3(kPolyThreads), 0, st,
                       (const uint32_t *)d_a, (uint32_t)n, M, (const uint32_t *)d_quot, to_dev<FR>(e),
                       (uint32_t *)d_rem);
    ZK_HIP(hipGetLastError());
  }
  return ZK_OK;
}

// quotient (D - n + 1 words groups) and remainder (n, optional) of a_0..a_D by x^n - eta,
// D >= n >= 1 (the caller handles D < n, where the quotient is zero)
int vanish_div_device(int curve, int D, const uint64_t *d_a, int n, const uint64_t *eta, uint64_t *d_quot,
                      uint64_t *d_rem, hipStream_t st) {
  if (n < 1 || D < n) {
    set_error("div_by_vanishing: need deg p >= n >= 1 on the device path");
    return ZK_EINVAL;
  }
  if (curve == 0) return vanish_run<Bn254G1>(D, d_a, n, eta, d_quot, d_rem, st);
  if (curve == 1) return vanish_run<Bls381G1>(D, d_a, n, eta, d_quot, d_rem, st);
  set_error("div_by_vanishing: bad curve id");
  return ZK_EINVAL;
}

}  // namespace zk
