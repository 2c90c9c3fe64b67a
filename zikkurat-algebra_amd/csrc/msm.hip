// Pippenger MSM on BN254 / BLS12-381 G1 for gfx950.
//
// Result contract: sum_i s_i P_i exactly as lib/cbits/curves/g1/proj/bn128_G1_proj.c:506-605
// computes it (scalars are 64*expo_nlimbs-bit integers, affine (0,0) inputs contribute
// nothing, Montgomery scalars are converted with Fr to_std first, :631-645).  The group sum
// is unique, so after normalisation the output matches the reference bit for bit.
//
// Device pipeline (one MSM, window c, B = 2^(c-1) buckets per window, W windows):
//   1 digits   signed c-bit digits d in [-2^(c-1), 2^(c-1)] per (window, point); the carry
//              makes W = ceil((bits + 1) / c) windows exact for scalars below 2^bits, with
//              bits = 64 enl, or the measured largest bit length under the statistics policies
//   2 count    per-(chunk, window) bucket histograms in LDS (packed 16-bit counters)
//   3 scan     exclusive prefix sum -> bucket offsets into one sorted array of all windows
//   4 scatter  counting-sort placement of (point index | sign) words, ranks from LDS atomics
//   5 accum    load-balanced bucket accumulation: thread t owns sorted entries
//              [t 2^s, (t+1) 2^s), finds their buckets from the offsets (binary search, then
//              a walk) and folds each run with mixed XYZZ + affine adds;
//              a run that began in an earlier segment is written as a continuation partial
//   6 fix      each bucket adds its continuation partials serially, an empty bucket gets the
//              infinity the accumulation did not write; heavy buckets (many partials) are
//              folded by workgroups in LDS, several workgroups per bucket
//   7 pyramid  bucket reduction sum_b (b+1) B_b without serial running sums: level k pairs
//              neighbours (P'_i = P_2i + P_2i+1) and keeps the odd children Q^(k) = P_2i+1,
//              which are pairwise reduced alongside; then sum_b (b+1) B_b =
//              Sum(B) + sum_k 2^(k-1) U_k with U_k the reduced Q^(k) (depth c-1 additions);
//              wide levels two per launch, the narrow ones per aligned block in LDS (one
//              launch for all levels of a block), the last ones one workgroup per window
//   8 host     the window terms are combined by one Horner pass on the host
//              (sum_w 2^(start_w) (Sum_w + sum_k 2^(k-1) U_wk), ~255 doublings) and normalised.
// Window widths are balanced (c or c-1 bits each, see MsmSchedule), so every window spreads
// its digits over most of its buckets.
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "curve.hpp"
#include "curvez.hpp"
#include "glv.hpp"
#include "hostfield.hpp"
#include "msm.hpp"
#include "runtime.hpp"

namespace zk {

static constexpr int kThreads = 256;
static constexpr uint32_t kZeroDigit = 0xffffffffu;

// ---------------------------------------------------------------------------------------
// 1. signed digits

__device__ __forceinline__ uint32_t window_bits(const uint64_t *__restrict__ s, int enl, int bit, int c) {
  const int nbits = 64 * enl;
  if (bit >= nbits) return 0;
  const int limb = bit >> 6, off = bit & 63;
  uint64_t v = s[limb] >> off;
  if (off + c > 64 && limb + 1 < enl) v |= s[limb + 1] << (64 - off);
  return (uint32_t)(v & ((1ull << c) - 1));
}

// true when s has a set bit at position `from` or above
__device__ __forceinline__ bool bits_from(const uint64_t *__restrict__ s, int enl, int from) {
  uint64_t any = 0;
  for (int j = 0; j < enl; j++) {
    if (64 * (j + 1) <= from) continue;
    any |= 64 * j >= from ? s[j] : s[j] >> (from - 64 * j);
  }
  return any != 0;
}

__device__ __forceinline__ uint32_t encode_digit(uint32_t raw, int c, uint32_t &carry) {
  const uint32_t half = 1u << (c - 1);
  if (raw > half) {
    const uint32_t mag = (1u << c) - raw;  // d = raw - 2^c < 0
    carry = 1;
    return mag ? (((mag - 1) << 1) | 1u) : kZeroDigit;
  }
  carry = 0;
  return raw ? ((raw - 1) << 1) : kZeroDigit;
}

// point i (external affine words) -> its internal record (record i STRIDE of out); with phi
// (BN254) also the record of phi(P_i) next to it
template <class C>
__device__ __forceinline__ void point_to_internal(const uint32_t *__restrict__ pts, uint32_t i,
                                                  int32_t *__restrict__ out, bool phi) {
  using F = typename C::Fp;
  uint32_t w[2 * F::N];
  const uint4 *q = reinterpret_cast<const uint4 *>(pts + (size_t)i * 2 * F::N);
#pragma unroll
  for (int j = 0; j < F::N / 2; j++) {
    uint4 v = q[j];
    w[4 * j] = v.x;
    w[4 * j + 1] = v.y;
    w[4 * j + 2] = v.z;
    w[4 * j + 3] = v.w;
  }
  AffZ<C> a;
  zfrom_ext(a.x, w);
  zfrom_ext(a.y, w + F::N);
  constexpr int AW = ZLayout<C>::AFF, S = ZLayout<C>::STRIDE;
  zaff_store(out + (size_t)i * S * AW, a);
  if constexpr (C::id == 0 && S == 2) {
    if (phi) {  // phi(P) = (beta x, y); the affine infinity (0, 0) maps to itself
      Fz<ZBn254Fp> beta;
      zset_const(beta, Bn254Glv::BETA);
      zmul(a.x, a.x, beta);
      zaff_store(out + ((size_t)i * S + 1) * AW, a);
    }
  }
}

// record of entry idx (a digit position): positions below n are the points P_i, positions
// n + i (GLV) the points phi(P_i)
template <class C>
__device__ __forceinline__ size_t point_record(uint32_t idx, uint32_t n) {
  if constexpr (ZLayout<C>::STRIDE == 2) return idx < n ? 2 * (size_t)idx : 2 * (size_t)(idx - n) + 1;
  else return idx;
}

// signed digits of scalar i for every window (row w at dig + w * ns); when ext_pts is given,
// point i is converted to the internal record in the same pass (both sweeps are bound by
// memory traffic).  With glv (BN254, 4-limb scalars reduced mod r) scalar i becomes k1 for
// point i and k2 for position n + i, phi(P_i), whose record is written too (ns = 2n).  With ovf,
// a scalar the windows do not cover (a carry out of the top window or a bit above it: the
// schedule was fitted to other scalars) sets *ovf.
template <class C>
__global__ __launch_bounds__(kThreads) void msm_digits_kernel(const uint64_t *__restrict__ sc, int enl, int mont,
                                                             int reduce, int glv, uint32_t n, uint32_t ns, int c,
                                                             int W, int wide, uint32_t *__restrict__ ovf,
                                                             uint32_t *__restrict__ dig,
                                                             const uint32_t *__restrict__ ext_pts,
                                                             int32_t *__restrict__ zpts) {
  using Fr = typename C::Fr;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (ext_pts) point_to_internal<C>(ext_pts, i, zpts, glv != 0);
  uint32_t carry = 0;
  if (enl == 4) {
    uint64_t s[4];
#pragma unroll
    for (int j = 0; j < 4; j++) s[j] = sc[(size_t)i * 4 + j];
    if (mont) {
      Fe<Fr> x;
#pragma unroll
      for (int j = 0; j < 8; j++) x.w[j] = (uint32_t)(s[j >> 1] >> (32 * (j & 1)));
      fto_std(x, x);
#pragma unroll
      for (int j = 0; j < 4; j++) s[j] = (uint64_t)x.w[2 * j] | ((uint64_t)x.w[2 * j + 1] << 32);
    }
    if (reduce) {
      // s < 2^256 < 6 r: at most five subtractions of r
      uint64_t r[4];
#pragma unroll
      for (int j = 0; j < 4; j++) r[j] = (uint64_t)Fr::P[2 * j] | ((uint64_t)Fr::P[2 * j + 1] << 32);
      for (int k = 0; k < 5; k++) {
        bool ge = true;
        for (int j = 3; j >= 0; j--) {
          if (s[j] != r[j]) {
            ge = s[j] > r[j];
            break;
          }
        }
        if (!ge) break;
        uint64_t borrow = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const uint64_t d = s[j] - r[j] - borrow;
          borrow = (s[j] < r[j] || (s[j] == r[j] && borrow)) ? 1 : 0;
          s[j] = d;
        }
      }
    }
    if (C::id == 0 && glv) {
      uint64_t m1[2], m2[2];
      bool n1, n2;
      glv_split(s, m1, n1, m2, n2);
      uint32_t carry2 = 0;
      int bit = 0;
      for (int w = 0; w < W; w++) {
        const int cw = w < wide ? c : c - 1;
        uint32_t d1 = encode_digit(window_bits(m1, 2, bit, cw) + carry, cw, carry);
        uint32_t d2 = encode_digit(window_bits(m2, 2, bit, cw) + carry2, cw, carry2);
        if (n1 && d1 != kZeroDigit) d1 ^= 1u;  // digits of -|k1|: flip the sign bit
        if (n2 && d2 != kZeroDigit) d2 ^= 1u;
        dig[(size_t)w * ns + i] = d1;
        dig[(size_t)w * ns + n + i] = d2;
        bit += cw;
      }
      return;
    }
    int bit = 0;
    for (int w = 0; w < W; w++) {
      const int cw = w < wide ? c : c - 1;
      const uint32_t raw = window_bits(s, 4, bit, cw) + carry;
      dig[(size_t)w * ns + i] = encode_digit(raw, cw, carry);
      bit += cw;
    }
    if (ovf && (carry || bits_from(s, 4, bit))) atomicOr(ovf, 1u);
  } else {
    const uint64_t *s = sc + (size_t)i * enl;
    int bit = 0;
    for (int w = 0; w < W; w++) {
      const int cw = w < wide ? c : c - 1;
      const uint32_t raw = window_bits(s, enl, bit, cw) + carry;
      dig[(size_t)w * ns + i] = encode_digit(raw, cw, carry);
      bit += cw;
    }
    if (ovf && (carry || bits_from(s, enl, bit))) atomicOr(ovf, 1u);
  }
}

// scalar statistics: stats[0] = largest bit length, stats[1] = zeros, stats[2] = ones.  A
// grid-stride loop over at most kStatsBlocks blocks: each thread folds its scalars locally,
// the block folds in LDS, and one set of global atomics per block reaches the three words
// (one block per 256 scalars put ~50 000 same-address atomics on them at 2^22: 0.5 ms).
static constexpr unsigned kStatsBlocks = 1024;

template <class C>
__global__ __launch_bounds__(kThreads) void msm_stats_kernel(const uint64_t *__restrict__ sc, int enl, int mont,
                                                            uint32_t n, uint32_t *__restrict__ stats) {
  using Fr = typename C::Fr;
  __shared__ uint32_t s_max, s_zero, s_one;
  if (threadIdx.x == 0) {
    s_max = 0;
    s_zero = 0;
    s_one = 0;
  }
  __syncthreads();
  uint32_t bmax = 0, zc = 0, oc = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int bits = 0;
    if (mont && enl == 4) {
      Fe<Fr> x;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const uint64_t v = sc[(size_t)i * 4 + j];
        x.w[2 * j] = (uint32_t)v;
        x.w[2 * j + 1] = (uint32_t)(v >> 32);
      }
      fto_std(x, x);
#pragma unroll
      for (int j = 0; j < 8; j++)
        if (x.w[j]) bits = 32 * j + 32 - __clz(x.w[j]);
    } else {
      for (int j = 0; j < enl; j++) {
        const uint64_t v = sc[(size_t)i * enl + j];
        if (v) bits = 64 * j + 64 - __clzll(v);
      }
    }
    bmax = bmax > (uint32_t)bits ? bmax : (uint32_t)bits;
    zc += bits == 0;
    oc += bits == 1;
  }
  if (bmax) atomicMax(&s_max, bmax);
  if (zc) atomicAdd(&s_zero, zc);
  if (oc) atomicAdd(&s_one, oc);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (s_max) atomicMax(&stats[0], s_max);
    if (s_zero) atomicAdd(&stats[1], s_zero);
    if (s_one) atomicAdd(&stats[2], s_one);
  }
}

// ---------------------------------------------------------------------------------------
// 2-4. counting sort by (window, bucket)

static constexpr int kScanPer = 16;  // elements per thread in the block scan

template <int NT = kThreads>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *sh, uint32_t *total) {
  // Hillis-Steele over the NT threads of the workgroup in LDS
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int o = 1; o < NT; o <<= 1) {
    uint32_t x = t >= o ? sh[t - o] : 0u;
    __syncthreads();
    sh[t] += x;
    __syncthreads();
  }
  const uint32_t incl = sh[t];
  *total = sh[NT - 1];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(kThreads) void scan_block_sums_kernel(const uint32_t *__restrict__ in, size_t len,
                                                                  uint32_t *__restrict__ bsums) {
  __shared__ uint32_t sh[kThreads];
  const size_t base = (size_t)blockIdx.x * kThreads * kScanPer + (size_t)threadIdx.x * kScanPer;
  uint32_t s = 0;
  for (int k = 0; k < kScanPer; k++)
    if (base + k < len) s += in[base + k];
  uint32_t tot;
  block_exclusive_scan(s, sh, &tot);
  if (threadIdx.x == 0) bsums[blockIdx.x] = tot;
}

// out[i] = exclusive prefix of in; out[len] = total (cursor, when given, gets a copy).  bsums
// holds the per-block totals of scan_block_sums_kernel; each block adds up the totals of the
// blocks before it (a few hundred words at most), so no separate launch scans them.
__global__ __launch_bounds__(kThreads) void scan_final_kernel(const uint32_t *__restrict__ in, size_t len,
                                                             const uint32_t *__restrict__ bsums,
                                                             uint32_t *__restrict__ out, uint32_t *__restrict__ cursor) {
  __shared__ uint32_t sh[kThreads];
  const size_t base = (size_t)blockIdx.x * kThreads * kScanPer + (size_t)threadIdx.x * kScanPer;
  uint32_t v[kScanPer];
  uint32_t s = 0;
  for (int k = 0; k < kScanPer; k++) {
    v[k] = base + k < len ? in[base + k] : 0u;
    s += v[k];
  }
  uint32_t pre = 0, before, tot;
  for (uint32_t i = threadIdx.x; i < blockIdx.x; i += kThreads) pre += bsums[i];
  (void)block_exclusive_scan(pre, sh, &before);
  uint32_t run = block_exclusive_scan(s, sh, &tot) + before;
  for (int k = 0; k < kScanPer; k++) {
    if (base + k < len) {
      out[base + k] = run;
      if (cursor) cursor[base + k] = run;
    } else if (base + k == len) {
      out[len] = run;
    }
    run += v[k];
  }
}

// Two-level counting sort by (window, bucket).  A window's B = P 2^L buckets form P
// partitions of 2^L consecutive buckets.
//   level 1  workgroup (chunk, window) counts its chunk's digits per partition in LDS; a scan
//            over (window, partition, chunk) gives every chunk its run in each partition, and
//            the chunk places (entry, low bucket bits) words there: runs of ~kPChunk / P
//            consecutive 8-byte words, so HBM sees whole lines, not one 4-byte word per bucket.
//   level 2  workgroup (window, partition) counts its partition's words per low bucket, writes
//            the bucket offsets, and places the entries in bucket order.
// Positions within a bucket come from LDS atomics, so their order varies from run to run;
// the bucket sums do not depend on it.
static constexpr int kSortThreads = 1024;
static constexpr uint32_t kPChunk = 32768;  // points per level-1 workgroup
static constexpr int kPUnroll = 8;          // digits loaded ahead per thread (latency hiding)
static constexpr size_t kPartStageBytes = 16384;  // LDS for one partition's level-1 words
static constexpr uint32_t kBigPart = 16384;       // larger partitions are sorted run by run
// dynamic LDS of the staged level-1 placement: 3 counters per partition (at most 2^10
// partitions) and one chunk of words
static constexpr size_t kScatterStageLds = (3 * 1024 + kPChunk) * 4;

// LDS counter increment returning the counter's previous value plus this lane's rank.  The
// lanes that share the first active lane's key use one atomic for all of them when there are
// enough of them (skewed digit sets put most lanes of a wave on one bucket).
__device__ __forceinline__ uint32_t lds_rank(uint32_t *cnt, uint32_t key, bool active) {
  const uint64_t act = __ballot(active);
  if (act == 0) return 0;
  const int leader = __ffsll((unsigned long long)act) - 1;
  const uint32_t lkey = __shfl((int)key, leader);
  const bool mine = active && key == lkey;
  const uint64_t same = __ballot(mine);
  if (__popcll(same) >= 8) {
    const int lane = __lane_id();
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&cnt[lkey], (uint32_t)__popcll(same));
    base = (uint32_t)__shfl((int)base, leader);
    if (mine) return base + (uint32_t)__popcll(same & ((1ull << lane) - 1));
    return active ? atomicAdd(&cnt[key], 1u) : 0u;
  }
  return active ? atomicAdd(&cnt[key], 1u) : 0u;
}

// phist[(w P + p) nchunks + chunk] = digits of the chunk in partition p of window w.
// Workgroup (0, 0) also clears the heavy-bucket counters (heavy[0..1]) and the big-partition
// count (big[0]) that later kernels of the same MSM count into, in place of two fill launches.
__global__ __launch_bounds__(kSortThreads) void msm_part_hist_kernel(const uint32_t *__restrict__ dig, uint32_t n,
                                                                     int L, uint32_t P,
                                                                     uint32_t *__restrict__ phist,
                                                                     uint32_t *__restrict__ heavy,
                                                                     uint32_t *__restrict__ big) {
  extern __shared__ uint32_t cnt[];  // P counters
  const uint32_t chunk = blockIdx.x, w = blockIdx.y, nchunks = gridDim.x;
  if (chunk == 0 && w == 0 && threadIdx.x == 0) {
    heavy[0] = 0;
    heavy[1] = 0;
    big[0] = 0;
  }
  for (uint32_t j = threadIdx.x; j < P; j += kSortThreads) cnt[j] = 0;
  __syncthreads();
  const uint32_t i0 = chunk * kPChunk;
  const uint32_t i1 = n - i0 < kPChunk ? n : i0 + kPChunk;
  const uint32_t *d = dig + (size_t)w * n;
  for (uint32_t i = i0 + threadIdx.x; i < i1; i += kSortThreads * kPUnroll) {
    uint32_t code[kPUnroll];
#pragma unroll
    for (int u = 0; u < kPUnroll; u++) {
      const uint32_t k = i + u * kSortThreads;
      code[u] = k < i1 ? d[k] : kZeroDigit;
    }
#pragma unroll
    for (int u = 0; u < kPUnroll; u++) {
      const bool active = code[u] != kZeroDigit;
      lds_rank(cnt, active ? (code[u] >> 1) >> L : 0u, active);
    }
  }
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < P; p += kSortThreads) phist[((size_t)w * P + p) * nchunks + chunk] = cnt[p];
}

// Level-1 words.  Narrow (u32): point index in the low ib bits, the sign above it, the low
// bucket bits above that (for ib + 1 + L <= 32, which covers 2^20 points with up to 11 low
// bits).  Wide (u64): the final entry (point index | sign << 31) low, the low bucket bits high.
template <class T> struct PartWord;
template <> struct PartWord<uint32_t> {
  __device__ static uint32_t pack(uint32_t i, uint32_t sign, uint32_t lo, int ib) {
    return i | (sign << ib) | (lo << (ib + 1));
  }
  __device__ static uint32_t lo(uint32_t v, int ib) { return v >> (ib + 1); }
  __device__ static uint32_t entry(uint32_t v, int ib) {
    return (v & ((1u << ib) - 1)) | (((v >> ib) & 1u) << 31);
  }
};
template <> struct PartWord<uint64_t> {
  __device__ static uint64_t pack(uint32_t i, uint32_t sign, uint32_t lo, int) {
    return (uint64_t)(i | (sign << 31)) | ((uint64_t)lo << 32);
  }
  __device__ static uint32_t lo(uint64_t v, int) { return (uint32_t)(v >> 32); }
  __device__ static uint32_t entry(uint64_t v, int) { return (uint32_t)v; }
};

// tmp[ppos...] = level-1 word of every nonzero digit of the chunk
template <class T>
__global__ __launch_bounds__(kSortThreads) void msm_part_scatter_kernel(const uint32_t *__restrict__ dig, uint32_t n,
                                                                        int L, uint32_t P, int ib,
                                                                        const uint32_t *__restrict__ ppos,
                                                                        T *__restrict__ tmp) {
  extern __shared__ uint32_t cur[];  // P cursors
  const uint32_t chunk = blockIdx.x, w = blockIdx.y, nchunks = gridDim.x;
  for (uint32_t p = threadIdx.x; p < P; p += kSortThreads) cur[p] = ppos[((size_t)w * P + p) * nchunks + chunk];
  __syncthreads();
  const uint32_t i0 = chunk * kPChunk;
  const uint32_t i1 = n - i0 < kPChunk ? n : i0 + kPChunk;
  const uint32_t *d = dig + (size_t)w * n;
  const uint32_t lmask = (1u << L) - 1;
  for (uint32_t i = i0 + threadIdx.x; i < i1; i += kSortThreads * kPUnroll) {
    uint32_t code[kPUnroll];
#pragma unroll
    for (int u = 0; u < kPUnroll; u++) {
      const uint32_t k = i + u * kSortThreads;
      code[u] = k < i1 ? d[k] : kZeroDigit;
    }
#pragma unroll
    for (int u = 0; u < kPUnroll; u++) {
      const bool active = code[u] != kZeroDigit;
      const uint32_t b = code[u] >> 1;
      const uint32_t pos = lds_rank(cur, active ? b >> L : 0u, active);
      if (active) tmp[pos] = PartWord<T>::pack(i + u * kSortThreads, code[u] & 1u, b & lmask, ib);
    }
  }
}

// Narrow words with the chunk staged in LDS: the words are first placed in partition order in
// LDS, then each wave copies whole partition runs, so every global store instruction writes
// consecutive words (direct placement writes one word per run per instruction, and the runs'
// lines leave L2 half written).  LDS: 3 P counters and kPChunk words.
__global__ __launch_bounds__(kSortThreads) void msm_part_scatter_staged_kernel(
    const uint32_t *__restrict__ dig, uint32_t n, int L, uint32_t P, int ib, const uint32_t *__restrict__ phist,
    const uint32_t *__restrict__ ppos, uint32_t *__restrict__ tmp) {
  extern __shared__ uint32_t sm[];
  __shared__ uint32_t sh[kSortThreads];
  uint32_t *lcur = sm, *gbase = sm + P, *lcnt = sm + 2 * P, *stage = sm + 3 * P;
  const uint32_t chunk = blockIdx.x, w = blockIdx.y, nchunks = gridDim.x, t = threadIdx.x;
  // P <= kSortThreads: thread t owns partition t
  const size_t k = ((size_t)w * P + t) * nchunks + chunk;
  const uint32_t c = t < P ? phist[k] : 0u;
  uint32_t tot;
  const uint32_t ex = block_exclusive_scan<kSortThreads>(c, sh, &tot);
  if (t < P) {
    lcur[t] = ex;
    gbase[t] = ppos[k];
    lcnt[t] = c;
  }
  __syncthreads();
  const uint32_t i0 = chunk * kPChunk;
  const uint32_t i1 = n - i0 < kPChunk ? n : i0 + kPChunk;
  const uint32_t *d = dig + (size_t)w * n;
  const uint32_t lmask = (1u << L) - 1;
  for (uint32_t i = i0 + t; i < i1; i += kSortThreads * kPUnroll) {
    uint32_t code[kPUnroll];
#pragma unroll
    for (int u = 0; u < kPUnroll; u++) {
      const uint32_t q = i + u * kSortThreads;
      code[u] = q < i1 ? d[q] : kZeroDigit;
    }
#pragma unroll
    for (int u = 0; u < kPUnroll; u++) {
      const bool active = code[u] != kZeroDigit;
      const uint32_t b = code[u] >> 1;
      const uint32_t pos = lds_rank(lcur, active ? b >> L : 0u, active);
      if (active) stage[pos] = PartWord<uint32_t>::pack(i + u * kSortThreads, code[u] & 1u, b & lmask, ib);
    }
  }
  __syncthreads();
  // lcur[p] is now the end of run p in the stage
  const uint32_t lane = t & 63, wave = t >> 6;
  for (uint32_t p = wave; p < P; p += kSortThreads / 64) {
    const uint32_t cnt = lcnt[p], ls = lcur[p] - cnt, gs = gbase[p];
    for (uint32_t j = lane; j < cnt; j += 64) tmp[gs + j] = stage[ls + j];
  }
}

// Workgroup g = w P + p sorts tmp[ppos[g nchunks], ppos[(g + 1) nchunks]) by the low bucket
// bits into vals and writes off[g 2^L + lo] (global bucket w B + p 2^L + lo); the last
// workgroup also writes off[nparts 2^L] = the total.  A partition of at most `stage` words is
// read from HBM once into LDS and its entries are ordered in LDS, then written with
// consecutive stores; a larger one is read twice and placed directly.  A partition of more
// than kBigPart words (skewed digits: one bucket holding a large share of a window) is only
// listed in big[1..] (big[0] = count) for the run-parallel kernels below.
template <class T>
__global__ __launch_bounds__(kThreads) void msm_part_sort_kernel(const T *__restrict__ tmp,
                                                                 const uint32_t *__restrict__ ppos, uint32_t nchunks,
                                                                 int L, int ib, uint32_t stage,
                                                                 uint32_t *__restrict__ off,
                                                                 uint32_t *__restrict__ vals,
                                                                 uint32_t *__restrict__ big) {
  extern __shared__ uint32_t lc[];  // 2^L counters, then cursors; the staged words; staged entries
  __shared__ uint32_t sh[kThreads];
  const uint32_t g = blockIdx.x, NL = 1u << L, t = threadIdx.x;
  T *words = reinterpret_cast<T *>(lc + (NL < 2 ? 2 : NL));
  uint32_t *ents = reinterpret_cast<uint32_t *>(words + stage);
  const uint32_t start = ppos[(size_t)g * nchunks], end = ppos[(size_t)(g + 1) * nchunks];
  if (g == gridDim.x - 1 && t == 0) off[(size_t)gridDim.x * NL] = end;
  if (end - start > kBigPart) {  // sorted run by run by the big-partition kernels
    if (t == 0) big[1 + atomicAdd(&big[0], 1u)] = g;
    return;
  }
  const bool staged = end - start <= stage;
  for (uint32_t j = t; j < NL; j += kThreads) lc[j] = 0;
  __syncthreads();
  for (uint32_t e = start + t; e < end; e += kThreads * 4) {
    T v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const uint32_t k = e + u * kThreads;
      v[u] = k < end ? tmp[k] : T(0);
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const uint32_t k = e + u * kThreads;
      if (staged && k < end) words[k - start] = v[u];
      lds_rank(lc, PartWord<T>::lo(v[u], ib), k < end);
    }
  }
  __syncthreads();
  // exclusive scan of the counters: thread t owns `per` consecutive ones
  const uint32_t per = NL > kThreads ? NL / kThreads : 1u;
  const uint32_t j0 = t * per;
  uint32_t s = 0;
  for (uint32_t k = 0; k < per; k++)
    if (j0 + k < NL) s += lc[j0 + k];
  uint32_t tot;
  uint32_t run = start + block_exclusive_scan(s, sh, &tot);
  for (uint32_t k = 0; k < per; k++) {
    if (j0 + k < NL) {
      const uint32_t cnt = lc[j0 + k];
      lc[j0 + k] = run;
      off[(size_t)g * NL + j0 + k] = run;
      run += cnt;
    }
  }
  __syncthreads();
  for (uint32_t e = start + t; e < end; e += kThreads * 4) {
    T v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const uint32_t k = e + u * kThreads;
      v[u] = k < end ? (staged ? words[k - start] : tmp[k]) : T(0);
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const bool active = e + u * kThreads < end;
      const uint32_t pos = lds_rank(lc, PartWord<T>::lo(v[u], ib), active);
      if (active) {
        if (staged)
          ents[pos - start] = PartWord<T>::entry(v[u], ib);
        else
          vals[pos] = PartWord<T>::entry(v[u], ib);
      }
    }
  }
  if (staged) {
    __syncthreads();
    for (uint32_t e = t; e < end - start; e += kThreads) vals[start + e] = ents[e];
  }
}

// Big partitions, sorted by their level-1 runs (one per chunk) in parallel: count every run
// by its low bucket bits, scan each partition over (low bits, run) into bucket offsets and
// per-run bases, then place every run.  rcnt[(i nchunks + chunk) 2^L + lo] belongs to the
// i-th listed partition.  Persistent grids: the item count is read on the device.
template <class T>
__global__ __launch_bounds__(kThreads) void msm_big_count_kernel(const T *__restrict__ tmp,
                                                                 const uint32_t *__restrict__ ppos, uint32_t nchunks,
                                                                 int L, int ib, const uint32_t *__restrict__ big,
                                                                 uint32_t *__restrict__ rcnt) {
  extern __shared__ uint32_t lc[];
  const uint32_t NL = 1u << L, t = threadIdx.x;
  const uint32_t items = big[0] * nchunks;
  for (uint32_t it = blockIdx.x; it < items; it += gridDim.x) {
    const uint32_t bi = it / nchunks, ch = it - bi * nchunks;
    const size_t r = (size_t)big[1 + bi] * nchunks + ch;
    const uint32_t s = ppos[r], e = ppos[r + 1];
    for (uint32_t j = t; j < NL; j += kThreads) lc[j] = 0;
    __syncthreads();
    for (uint32_t k = s + t; k < e; k += kThreads * 4) {
      T v[4];
#pragma unroll
      for (int u = 0; u < 4; u++) v[u] = k + u * kThreads < e ? tmp[k + u * kThreads] : T(0);
#pragma unroll
      for (int u = 0; u < 4; u++) lds_rank(lc, PartWord<T>::lo(v[u], ib), k + u * kThreads < e);
    }
    __syncthreads();
    for (uint32_t j = t; j < NL; j += kThreads) rcnt[((size_t)bi * nchunks + ch) * NL + j] = lc[j];
    __syncthreads();
  }
}

__global__ __launch_bounds__(kThreads) void msm_big_scan_kernel(const uint32_t *__restrict__ ppos, uint32_t nchunks,
                                                                int L, const uint32_t *__restrict__ big,
                                                                uint32_t *__restrict__ rcnt,
                                                                uint32_t *__restrict__ off) {
  __shared__ uint32_t sh[kThreads];
  const uint32_t NL = 1u << L, t = threadIdx.x, nbig = big[0];
  const uint32_t per = NL > kThreads ? NL / kThreads : 1u, j0 = t * per;
  for (uint32_t bi = blockIdx.x; bi < nbig; bi += gridDim.x) {
    const uint32_t g = big[1 + bi];
    const uint32_t start = ppos[(size_t)g * nchunks];
    uint32_t *rc = rcnt + (size_t)bi * nchunks * NL;
    uint32_t s = 0;
    for (uint32_t k = 0; k < per; k++)
      if (j0 + k < NL)
        for (uint32_t ch = 0; ch < nchunks; ch++) s += rc[(size_t)ch * NL + j0 + k];
    uint32_t tot;
    uint32_t run = start + block_exclusive_scan(s, sh, &tot);
    for (uint32_t k = 0; k < per; k++) {
      if (j0 + k < NL) {
        off[(size_t)g * NL + j0 + k] = run;
        for (uint32_t ch = 0; ch < nchunks; ch++) {
          const uint32_t c = rc[(size_t)ch * NL + j0 + k];
          rc[(size_t)ch * NL + j0 + k] = run;
          run += c;
        }
      }
    }
  }
}

// Buckets whose runs span more than `limit` segments are heavy (heavy[0] = count, heavy[1] =
// most partials, heavy[2] = sorted entries, heavy[3..] = buckets).  Known from the offsets
// alone, so the last sort kernel lists them before the accumulation; the heavy-bucket kernels
// read the list on the device.
__device__ __forceinline__ void heavy_detect(const uint32_t *__restrict__ off, uint32_t nkeys, uint32_t b,
                                             int seg_log, uint32_t *__restrict__ heavy, uint32_t heavy_cap,
                                             uint32_t limit) {
  if (b == 0) heavy[2] = off[nkeys];
  const uint32_t s = off[b], t = off[b + 1];
  if (s == t) return;
  const uint64_t k = ((uint64_t)(t - 1) >> seg_log) - (s >> seg_log);
  if (k > limit) {
    const uint32_t idx = atomicAdd(&heavy[0], 1u);
    atomicMax(&heavy[1], (uint32_t)k);
    if (idx < heavy_cap) heavy[3 + idx] = b;
  }
}

// Places the big partitions' entries run by run; every bucket offset is final by now (the
// partition sort and msm_big_scan_kernel wrote them), so the grid also runs heavy_detect over
// all buckets first.
template <class T>
__global__ __launch_bounds__(kThreads) void msm_big_place_kernel(const T *__restrict__ tmp,
                                                                 const uint32_t *__restrict__ ppos, uint32_t nchunks,
                                                                 int L, int ib, const uint32_t *__restrict__ big,
                                                                 const uint32_t *__restrict__ rcnt,
                                                                 uint32_t *__restrict__ vals,
                                                                 const uint32_t *__restrict__ off, uint32_t nkeys,
                                                                 int seg_log, uint32_t *__restrict__ heavy,
                                                                 uint32_t heavy_cap, uint32_t limit) {
  extern __shared__ uint32_t lc[];
  for (uint32_t b = blockIdx.x * kThreads + threadIdx.x; b < nkeys; b += gridDim.x * kThreads)
    heavy_detect(off, nkeys, b, seg_log, heavy, heavy_cap, limit);
  const uint32_t NL = 1u << L, t = threadIdx.x;
  const uint32_t items = big[0] * nchunks;
  for (uint32_t it = blockIdx.x; it < items; it += gridDim.x) {
    const uint32_t bi = it / nchunks, ch = it - bi * nchunks;
    const size_t r = (size_t)big[1 + bi] * nchunks + ch;
    const uint32_t s = ppos[r], e = ppos[r + 1];
    for (uint32_t j = t; j < NL; j += kThreads) lc[j] = rcnt[((size_t)bi * nchunks + ch) * NL + j];
    __syncthreads();
    for (uint32_t k = s + t; k < e; k += kThreads * 4) {
      T v[4];
#pragma unroll
      for (int u = 0; u < 4; u++) v[u] = k + u * kThreads < e ? tmp[k + u * kThreads] : T(0);
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const bool active = k + u * kThreads < e;
        const uint32_t pos = lds_rank(lc, PartWord<T>::lo(v[u], ib), active);
        if (active) vals[pos] = PartWord<T>::entry(v[u], ib);
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------------------
// 5. bucket accumulation

template <class C>
__device__ __forceinline__ void flush_run(const XyzzZ<C> &acc, bool cont_run, uint64_t seg, uint32_t key,
                                          int32_t *bsum, int32_t *cont) {
  constexpr int PW = ZLayout<C>::XYZZ;
  int32_t *dst = cont_run ? cont + seg * PW : bsum + (size_t)key * PW;
  zxyzz_store(dst, acc);
}

// affine points, external Montgomery words -> internal records (prepared point sets)
template <class C>
__global__ __launch_bounds__(kThreads) void msm_points_in_kernel(const uint32_t *__restrict__ pts, uint32_t n,
                                                                int32_t *__restrict__ out, int phi) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  point_to_internal<C>(pts, i, out, phi != 0);
}

// final window terms: internal XYZZ -> external canonical words (4 field elements each)
template <class C>
__global__ __launch_bounds__(kThreads) void msm_points_out_kernel(const int32_t *__restrict__ in, uint32_t cnt,
                                                                 uint32_t *__restrict__ out,
                                                                 const uint32_t *__restrict__ info_src,
                                                                 uint32_t *__restrict__ info_dst) {
  using F = typename C::Fp;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (info_dst && i < 3) info_dst[i] = info_src[i];  // the heavy-bucket statistics words
  if (i >= cnt) return;
  XyzzZ<C> p;
  zxyzz_load(p, in + (size_t)i * ZLayout<C>::XYZZ);
  uint32_t *o = out + (size_t)i * 4 * F::N;
  if (zxyzz_is_inf(p)) {
    for (int j = 0; j < 4 * F::N; j++) o[j] = 0;
    return;
  }
  zto_ext(o, p.x);
  zto_ext(o + F::N, p.y);
  zto_ext(o + 2 * F::N, p.zz);
  zto_ext(o + 3 * F::N, p.zzz);
}

// bucket holding sorted entry e: the last b with off[b] <= e (off is non-decreasing)
__device__ __forceinline__ uint32_t bucket_of(const uint32_t *__restrict__ off, uint32_t nkeys, uint32_t e) {
  uint32_t lo = 0, hi = nkeys;  // answer in [lo, hi)
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= e) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Bucket holding sorted entry e when bucket `from` starts at or before e: the last b >= from
// with off[b] <= e.  Usually from itself; runs of empty buckets (windows narrower than the
// bucket space) are crossed by a galloping then binary search, not one bucket at a time.
__device__ __forceinline__ uint32_t bucket_from(const uint32_t *__restrict__ off, uint32_t nkeys, uint32_t from,
                                                uint32_t e) {
  if (off[from + 1] > e) return from;
  uint32_t lo = from + 1, step = 1, hi;
  for (;;) {
    hi = lo + step;
    if (hi >= nkeys || off[hi] > e) break;
    lo = hi;
    step <<= 1;
  }
  if (hi > nkeys) hi = nkeys;  // off[nkeys] = total entries > e
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= e) lo = mid;
    else hi = mid;
  }
  return lo;
}

// PF: the next entry's point is loaded before the current entry's addition, so its latency
// overlaps the arithmetic (18 more registers per thread).  Measured ~3% slower than letting
// occupancy hide the latency, so it is off unless a schedule asks for it.
template <class C, bool PF>
__global__ __launch_bounds__(kThreads) void msm_accum_kernel(const uint32_t *__restrict__ vals,
                                                            const uint32_t *__restrict__ off, uint32_t nkeys,
                                                            int seg_log, const int32_t *__restrict__ points,
                                                            uint32_t npts, int32_t *__restrict__ bsum,
                                                            int32_t *__restrict__ cont) {
  constexpr int AW = ZLayout<C>::AFF;
  const uint32_t E = off[nkeys];
  const uint64_t seg = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t e0 = seg << seg_log;
  if (e0 >= E) return;
  uint64_t e1 = e0 + (1ull << seg_log);
  if (e1 > E) e1 = E;
  uint32_t cur = bucket_of(off, nkeys, (uint32_t)e0);
  uint32_t next = off[cur + 1];
  bool cont_run = off[cur] < e0;
  XyzzZ<C> acc;
  zxyzz_set_inf(acc);
  uint32_t v = vals[e0];
  AffZ<C> a;
  if (PF) zaff_load(a, points + point_record<C>(v & 0x7fffffffu, npts) * AW);
  for (uint64_t e = e0; e < e1; e++) {
    uint32_t vn = 0;
    AffZ<C> an;
    if (PF) {
      if (e + 1 < e1) {
        vn = vals[e + 1];
        zaff_load(an, points + point_record<C>(vn & 0x7fffffffu, npts) * AW);
      }
    } else {
      v = vals[e];
      zaff_load(a, points + point_record<C>(v & 0x7fffffffu, npts) * AW);
    }
    if (e == next) {
      flush_run<C>(acc, cont_run, seg, cur, bsum, cont);
      cont_run = false;
      zxyzz_set_inf(acc);
      cur = bucket_from(off, nkeys, cur + 1, (uint32_t)e);
      next = off[cur + 1];
    }
    if (!zaff_is_inf(a)) {
      if (v >> 31) zneg(a.y, a.y);
      zxyzz_madd(acc, a);
    }
    if (PF) {
      v = vn;
      a = an;
    }
  }
  flush_run<C>(acc, cont_run, seg, cur, bsum, cont);
}

// ---------------------------------------------------------------------------------------
// 6. continuation partials

template <class C>
__global__ __launch_bounds__(kThreads) void msm_fix_kernel(const uint32_t *__restrict__ off, uint32_t nkeys,
                                                          int seg_log, int32_t *__restrict__ bsum,
                                                          const int32_t *__restrict__ cont, uint32_t limit) {
  constexpr int PW = ZLayout<C>::XYZZ;
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nkeys) return;
  const uint32_t s = off[b], t = off[b + 1];
  XyzzZ<C> acc, q;
  if (s == t) {  // no entry: the accumulation wrote nothing here
    zxyzz_set_inf(acc);
    zxyzz_store(bsum + (size_t)b * PW, acc);
    return;
  }
  const uint64_t fs = s >> seg_log, ls = (uint64_t)(t - 1) >> seg_log;
  if (ls == fs) return;
  const uint64_t k = ls - fs;
  if (k > limit) return;  // a heavy bucket
  zxyzz_load(acc, bsum + (size_t)b * PW);
  for (uint64_t j = fs + 1; j <= ls; j++) {
    zxyzz_load(q, cont + j * PW);
    zxyzz_add(acc, acc, q);
  }
  zxyzz_store(bsum + (size_t)b * PW, acc);
}

// Heavy buckets (more than fix_limit continuation partials), reduced on the device in two
// launches of G workgroups whatever their number and size, with no host round trip.  With nh
// listed buckets and nh <= G, each bucket gets Q = G / nh workgroups: workgroup (h, q) sums the
// bucket's partials q 256 + t, (q + Q) 256 + t, ... (t its thread), folds the 256 sums in LDS
// and writes one point to hsum[q nh + h]; the combine kernel then folds each bucket's Q points
// the same way and adds the total into the bucket sum.  With nh > G, one workgroup per bucket
// (grid-stride) adds straight into the bucket sum and the combine kernel does nothing.  Both
// exit at once when no bucket is heavy.  A bucket of K partials costs a chain of about
// K / (Q 256) + 8 additions, then Q / 256 + 8.
static constexpr uint32_t kHeavyBlocks = 1024;

// the sum of the workgroup's acc values, in thread 0 (part: kThreads XYZZ points of LDS)
template <class C>
__device__ __forceinline__ void wg_fold(XyzzZ<C> &acc, int32_t *part) {
  constexpr int PW = ZLayout<C>::XYZZ;
  const uint32_t t = threadIdx.x;
  zxyzz_store(part + t * PW, acc);
  __syncthreads();
  for (uint32_t s = kThreads / 2; s > 0; s >>= 1) {
    if (t < s) {
      XyzzZ<C> q;
      zxyzz_load(acc, part + t * PW);
      zxyzz_load(q, part + (t + s) * PW);
      zxyzz_add(acc, acc, q);
      zxyzz_store(part + t * PW, acc);
    }
    __syncthreads();
  }
  if (t == 0) zxyzz_load(acc, part);
  __syncthreads();  // part is reused by the caller's next item
}

template <class C>
__device__ __forceinline__ void bsum_add(int32_t *__restrict__ bsum, uint32_t b, const XyzzZ<C> &v) {
  constexpr int PW = ZLayout<C>::XYZZ;
  XyzzZ<C> a;
  zxyzz_load(a, bsum + (size_t)b * PW);
  zxyzz_add(a, a, v);
  zxyzz_store(bsum + (size_t)b * PW, a);
}

template <class C>
__global__ __launch_bounds__(kThreads) void msm_heavy_part_kernel(const uint32_t *__restrict__ off,
                                                                 const uint32_t *__restrict__ heavy,
                                                                 uint32_t heavy_cap, uint32_t G, int seg_log,
                                                                 int32_t *__restrict__ bsum,
                                                                 const int32_t *__restrict__ cont,
                                                                 int32_t *__restrict__ hsum) {
  constexpr int PW = ZLayout<C>::XYZZ;
  extern __shared__ int32_t part[];
  const uint32_t nh = min(heavy[0], heavy_cap);
  if (nh == 0) return;
  const bool direct = nh > G;
  const uint32_t Q = direct ? 1u : G / nh, items = direct ? nh : nh * Q;
  const uint64_t stride = (uint64_t)Q * kThreads;
  for (uint32_t it = blockIdx.x; it < items; it += gridDim.x) {
    const uint32_t h = it % nh, q = it / nh;
    const uint32_t b = heavy[3 + h];
    const uint64_t fs = off[b] >> seg_log, ls = (uint64_t)(off[b + 1] - 1) >> seg_log;
    XyzzZ<C> acc, x;
    zxyzz_set_inf(acc);
    for (uint64_t j = fs + 1 + (uint64_t)q * kThreads + threadIdx.x; j <= ls; j += stride) {
      zxyzz_load(x, cont + j * PW);
      zxyzz_add(acc, acc, x);
    }
    wg_fold<C>(acc, part);
    if (threadIdx.x == 0) {
      if (direct) bsum_add<C>(bsum, b, acc);
      else zxyzz_store(hsum + (size_t)it * PW, acc);
    }
  }
}

template <class C>
__global__ __launch_bounds__(kThreads) void msm_heavy_combine_kernel(const uint32_t *__restrict__ off,
                                                                    const uint32_t *__restrict__ heavy,
                                                                    uint32_t heavy_cap, uint32_t G,
                                                                    int32_t *__restrict__ bsum,
                                                                    const int32_t *__restrict__ hsum) {
  constexpr int PW = ZLayout<C>::XYZZ;
  extern __shared__ int32_t part[];
  const uint32_t nh = min(heavy[0], heavy_cap);
  if (nh == 0 || nh > G) return;
  const uint32_t Q = G / nh;
  for (uint32_t h = blockIdx.x; h < nh; h += gridDim.x) {
    XyzzZ<C> acc, x;
    zxyzz_set_inf(acc);
    for (uint32_t q = threadIdx.x; q < Q; q += kThreads) {
      zxyzz_load(x, hsum + ((size_t)q * nh + h) * PW);
      zxyzz_add(acc, acc, x);
    }
    wg_fold<C>(acc, part);
    if (threadIdx.x == 0) bsum_add<C>(bsum, heavy[3 + h], acc);
  }
}

// ---------------------------------------------------------------------------------------
// 7. bucket pyramid: input [A][W][M] -> output [A+1][W][M/2]

// waves per SIMD the pyramid kernels are compiled for.  Left free, the compiler gives a
// BLS12-381 addition 256 VGPRs plus AGPRs (one wave per SIMD); held to two waves it spills a
// little and the pyramid is 15% faster (0.57 against 0.67 ms at 2^20).  BN254 fits two waves
// without a bound, and forcing three (168 VGPRs, spills) measured slower.
template <class C> struct PyrWaves { static constexpr int v = C::id == 0 ? 1 : 2; };

template <class C>
__global__ __launch_bounds__(kThreads, PyrWaves<C>::v) void msm_pyramid_kernel(const int32_t *__restrict__ in,
                                                              int32_t *__restrict__ out, uint32_t W, uint32_t M,
                                                              uint32_t A) {
  constexpr int PW = ZLayout<C>::XYZZ;
  const uint32_t half = M >> 1;
  const uint64_t per = (uint64_t)W * half;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= per * A) return;
  const uint32_t a = (uint32_t)(t / per);
  const uint64_t rem = t % per;  // = w * half + i
  const uint64_t src = (uint64_t)a * W * M + 2 * rem;
  XyzzZ<C> p0, p1;
  zxyzz_load(p0, in + src * PW);
  zxyzz_load(p1, in + (src + 1) * PW);
  if (a == 0) zxyzz_store(out + ((uint64_t)A * per + rem) * PW, p1);  // before the sum: p1 dies in it
  zxyzz_add(p0, p0, p1);
  zxyzz_store(out + ((uint64_t)a * per + rem) * PW, p0);
}

// two pyramid levels in one pass: input [A][W][M] -> output [A+2][W][M/4] (M >= 4).  For
// a < A the output is the sum of 4 consecutive inputs (two pair levels); array A is the odd
// children of the first level reduced once more (B1 + B3), array A + 1 the odd children of the
// second level (B2 + B3).  Half the launches and half the traffic of two single levels.  The
// B1 + B3 terms have threads of their own (the last W M/4 of the grid), so no thread runs
// more than three additions, two of them independent.
template <class C>
__global__ __launch_bounds__(kThreads, PyrWaves<C>::v) void msm_pyramid2_kernel(const int32_t *__restrict__ in,
                                                               int32_t *__restrict__ out, uint32_t W, uint32_t M,
                                                               uint32_t A) {
  constexpr int PW = ZLayout<C>::XYZZ;
  const uint32_t q = M >> 2;
  const uint64_t per = (uint64_t)W * q;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= per * (A + 1)) return;
  const uint32_t a = (uint32_t)(t / per);
  const uint64_t rem = t % per;  // = w * q + i
  XyzzZ<C> x, y, p0;
  if (a == A) {  // odd children of array 0's first level, reduced once more
    const int32_t *src = in + 4 * rem * PW;
    zxyzz_load(x, src + PW);
    zxyzz_load(y, src + 3 * PW);
    zxyzz_add(x, x, y);  // B1 + B3
    zxyzz_store(out + ((uint64_t)A * per + rem) * PW, x);
    return;
  }
  const int32_t *src = in + ((uint64_t)a * W * M + 4 * rem) * PW;
  zxyzz_load(x, src);
  zxyzz_load(y, src + PW);
  zxyzz_add(p0, x, y);  // B0 + B1
  zxyzz_load(x, src + 2 * PW);
  zxyzz_load(y, src + 3 * PW);
  zxyzz_add(x, x, y);  // B2 + B3
  if (a == 0) zxyzz_store(out + ((uint64_t)(A + 1) * per + rem) * PW, x);
  zxyzz_add(p0, p0, x);
  zxyzz_store(out + ((uint64_t)a * per + rem) * PW, p0);
}

// Once a level holds a few thousand points per window, a launch per level costs a dependent
// chain of additions plus a launch for little work.  Here each workgroup owns an aligned
// block of Mb points of every array of one window (grid: M / Mb blocks x W windows) and runs
// the log2(Mb) levels that reduce the block to one point, each a single-level step (pairs of
// every array, plus the odd children of array 0): the levels of one block use only that
// block's points.  With Mb = M this is the tail of the pyramid, one workgroup per window.
// Level 0 reads `in` ([A][W][M]) and writes the workgroup's LDS; later levels update the LDS
// in place (every item is read into registers before a barrier and written after it); the
// last level writes point q of each of the A + log2(Mb) arrays of window w to `out`
// ([A + log2 Mb][W][M / Mb]).  Workgroups share no memory but `in` (read only) and disjoint
// points of `out`, and `out` must not overlap `in`.  The host guarantees at most 2 blockDim.x
// items at level 0, at most blockDim.x at every later level, and LDS for the level-0 output.
static constexpr int kTailThreads = 512;
static constexpr uint64_t kPyr2MinThreads = 65536;
static constexpr size_t kTailLdsBytes = 128 * 1024;

template <class C>
__device__ __forceinline__ void tail_item(XyzzZ<C> &r, const int32_t *src, uint32_t A, uint32_t a, uint32_t i,
                                          uint64_t stride) {
  constexpr int PW = ZLayout<C>::XYZZ;
  if (a < A) {
    XyzzZ<C> q;
    zxyzz_load(r, src + (a * stride + 2 * i) * PW);
    zxyzz_load(q, src + (a * stride + 2 * i + 1) * PW);
    zxyzz_add(r, r, q);
  } else {
    zxyzz_load(r, src + (2 * i + 1) * PW);  // odd child of array 0
  }
}

template <class C>
__global__ __launch_bounds__(kTailThreads) void msm_pyramid_tail_kernel(const int32_t *__restrict__ in,
                                                                        int32_t *__restrict__ out, uint32_t W,
                                                                        uint32_t M, uint32_t A, uint32_t Mb) {
  constexpr int PW = ZLayout<C>::XYZZ;
  extern __shared__ int32_t tl_lds[];
  const uint32_t q = blockIdx.x, w = blockIdx.y, nb = M / Mb;
  // output point q of array a (window w)
  auto dst = [&](uint32_t a) { return out + (((uint64_t)a * W + w) * nb + q) * PW; };
  // level 0: global [A][W][M] (window w's arrays are W M points apart) -> LDS [A+1][Mb/2]
  {
    const uint32_t half = Mb >> 1, items = (A + 1) * half;
    const int32_t *src = in + ((uint64_t)w * M + (uint64_t)q * Mb) * PW;
    for (uint32_t j = threadIdx.x; j < items; j += blockDim.x) {
      const uint32_t a = j / half, i = j - a * half;
      XyzzZ<C> r;
      tail_item<C>(r, src, A, a, i, (uint64_t)W * M);
      if (half == 1)
        zxyzz_store(dst(a), r);
      else
        zxyzz_store(tl_lds + (uint64_t)j * PW, r);
    }
    A += 1;
    Mb = half;
  }
  while (Mb > 1) {
    __syncthreads();  // the previous level's points are in LDS
    const uint32_t half = Mb >> 1, items = (A + 1) * half;
    const uint32_t j = threadIdx.x;
    XyzzZ<C> r;
    const uint32_t a = j / half, i = j - a * half;
    if (j < items) tail_item<C>(r, tl_lds, A, a, i, Mb);
    __syncthreads();  // every read of this level is done before the in-place writes
    if (j < items) {
      if (half == 1)
        zxyzz_store(dst(a), r);
      else
        zxyzz_store(tl_lds + (uint64_t)j * PW, r);
    }
    A += 1;
    Mb = half;
  }
}

// ---------------------------------------------------------------------------------------
// fixed-base scalar multiplication k * G (input generation), affine output, 0 -> (0,0)

template <class C>
__global__ __launch_bounds__(kThreads) void scalar_mul_gen_kernel(const uint64_t *__restrict__ sc, uint32_t n,
                                                                 uint32_t *__restrict__ out) {
  using F = typename C::Fp;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t s[4];
#pragma unroll
  for (int j = 0; j < 4; j++) s[j] = sc[(size_t)i * 4 + j];
  Aff<C> g;
#pragma unroll
  for (int j = 0; j < F::N; j++) {
    g.x.w[j] = C::GX[j];
    g.y.w[j] = C::GY[j];
  }
  Xyzz<C> acc;
  xyzz_set_inf(acc);
  for (int bit = 255; bit >= 0; bit--) {
    xyzz_dbl(acc, acc);
    if ((s[bit >> 6] >> (bit & 63)) & 1) xyzz_madd(acc, g);
  }
  Aff<C> r;
  if (xyzz_is_inf(acc)) {
    fset_zero(r.x);
    fset_zero(r.y);
  } else {
    Fe<F> zzzi, zi;
    finv(zzzi, acc.zzz);
    fmul(zi, zzzi, acc.zz);
    fsqr(zi, zi);
    fmul(r.x, acc.x, zi);
    fmul(r.y, acc.y, zzzi);
  }
  fstore(out + (size_t)i * 2 * F::N, r.x);
  fstore(out + (size_t)i * 2 * F::N + F::N, r.y);
}

// ---------------------------------------------------------------------------------------
// sparse scalar sets: compaction of the scalars outside {0, 1}

// keep[i] = scalar i is neither 0 nor 1 (u32, for the scan); one[i] = scalar i is 1
template <class C>
__global__ __launch_bounds__(kThreads) void msm_sparse_flags_kernel(const uint64_t *__restrict__ sc, int enl,
                                                                   int mont, uint32_t n,
                                                                   uint32_t *__restrict__ keep,
                                                                   uint8_t *__restrict__ one) {
  using Fr = typename C::Fr;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t lo, hi = 0;
  if (enl == 4 && mont) {
    Fe<Fr> x;
#pragma unroll
    for (int j = 0; j < 8; j++) x.w[j] = (uint32_t)(sc[(size_t)i * 4 + j / 2] >> (32 * (j & 1)));
    fto_std(x, x);
    lo = (uint64_t)x.w[0] | ((uint64_t)x.w[1] << 32);
#pragma unroll
    for (int j = 2; j < 8; j++) hi |= x.w[j];
  } else {
    lo = sc[(size_t)i * enl];
    for (int j = 1; j < enl; j++) hi |= sc[(size_t)i * enl + j];
  }
  const bool small = hi == 0 && lo <= 1;
  keep[i] = small ? 0u : 1u;
  one[i] = (hi == 0 && lo == 1) ? 1 : 0;
}

// scalar and point i move to slot pos[i] when kept (words per scalar enl, per point pw)
__global__ __launch_bounds__(kThreads) void msm_sparse_gather_kernel(const uint64_t *__restrict__ sc, int enl,
                                                                    const uint32_t *__restrict__ pts, int pw,
                                                                    uint32_t n, const uint32_t *__restrict__ keep,
                                                                    const uint32_t *__restrict__ pos,
                                                                    uint64_t *__restrict__ sc_out,
                                                                    uint32_t *__restrict__ pts_out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !keep[i]) return;
  const uint32_t k = pos[i];
  for (int j = 0; j < enl; j++) sc_out[(size_t)k * enl + j] = sc[(size_t)i * enl + j];
  const uint4 *src = reinterpret_cast<const uint4 *>(pts + (size_t)i * pw);
  uint4 *dst = reinterpret_cast<uint4 *>(pts_out + (size_t)k * pw);
  for (int j = 0; j < pw / 4; j++) dst[j] = src[j];
}

// ---------------------------------------------------------------------------------------
// host side

static const zkh::FieldDesc &fp_desc(int curve) { return curve == 0 ? zkh::BN254_FP : zkh::BLS381_FP; }

// windows up to 16 bits sort with the chunk's bucket bases in LDS; 17-bit windows look the
// bases up in HBM and wider ones fall back to global atomics, both slower than the buckets
// they save (measured: BN254 2^22 at c = 17 spends 2.9 ms sorting against 2.1 ms at c = 16)
static constexpr int kMaxAnalyticWindow = 16;

int msm_heuristic_window(int curve, long long n, int scalar_bits) {
  (void)curve;
  if (n <= 1) return 2;
  double best = 1e300;
  int bc = 2;
  for (int c = 2; c <= kMaxAnalyticWindow; c++) {
    const double W = (double)((scalar_bits + 1 + c - 1) / c);
    const double cost = W * (10.0 * (double)n + 28.0 * (double)(1ll << (c - 1)));
    if (cost < best) {
      best = cost;
      bc = c;
    }
  }
  return bc;
}

void msm_infinity(int curve, MsmResult *r) {
  memset(r, 0, sizeof *r);
  r->curve = curve;
}

// x^-1 R for the Montgomery form x R (x != 0, canonical 64-bit words) of a base-field element,
// by the kernels' division-step inversion run on the host (measured 2.7 against 8.9 us for
// the host's a^(p-2) on BN254, 4.6 against 38 us on BLS12-381): unpacked, x R is an internal
// value u, zinv gives u^-1 R'^2, and one product by R^2 / R' (= zmul(R, R), R = K_OUT) gives
// u^-1 R^2 = x^-1 R.
template <class P>
static zkh::Fe host_inv_mont(const zkh::Fe &a) {
  constexpr int E = P::EXT_WORDS;
  uint32_t w[E];
  for (int j = 0; j < E; j++) w[j] = (uint32_t)(a.w[j / 2] >> (32 * (j & 1)));
  Fz<P> u, inv, k, t;
  zunpack(u, w);
  zinv(inv, u);
  zset_const(k, P::K_OUT);
  zmul(t, k, k);
  zmul(inv, inv, t);
  zcanon(inv, inv);
  zpack(w, inv);
  zkh::Fe r = zkh::zero();
  for (int j = 0; j < E; j++) r.w[j / 2] |= (uint64_t)w[j] << (32 * (j & 1));
  return r;
}

// affine Montgomery words of an XYZZ point (all 0xFF for infinity), as zkh::xyzz_to_affine
static void msm_to_affine(int curve, const zkh::FieldDesc &f, const zkh::Xyzz &p, uint64_t *out) {
  if (zkh::xyzz_is_inf(f, p)) {
    memset(out, 0xff, 16 * f.nl);
    return;
  }
  const zkh::Fe zzzi = curve == 0 ? host_inv_mont<ZBn254Fp>(p.zzz) : host_inv_mont<ZBls381Fp>(p.zzz);
  const zkh::Fe zi = zkh::mul(f, zzzi, p.zz);  // Z^2 / Z^3
  const zkh::Fe zzi = zkh::sqr(f, zi);
  zkh::store(f, out, zkh::mul(f, p.x, zzi));
  zkh::store(f, out + f.nl, zkh::mul(f, p.y, zzzi));
}

int msm_write_result(int curve, const MsmResult &r, uint64_t *tgt, int flags) {
  const zkh::FieldDesc &f = fp_desc(curve);
  const int L = f.nl;
  zkh::Xyzz p;
  p.x = zkh::load(f, r.x);
  p.y = zkh::load(f, r.y);
  p.zz = zkh::load(f, r.zz);
  p.zzz = zkh::load(f, r.zzz);
  if (!(flags & (ZKG_MSM_PROJ_OUT | ZKG_MSM_JAC_OUT))) {
    msm_to_affine(curve, f, p, tgt);
    return ZK_OK;
  }
  const zkh::Fe one = zkh::one(f), zero = zkh::zero();
  if (zkh::xyzz_is_inf(f, p)) {
    // projective infinity (0:1:0) (bn128_G1_proj.c:178-182); Jacobian [1:1:0] (bn128_G1_jac.c:183-188)
    const bool jac = flags & ZKG_MSM_JAC_OUT;
    zkh::store(f, tgt, jac ? one : zero);
    zkh::store(f, tgt + L, one);
    zkh::store(f, tgt + 2 * L, zero);
    return ZK_OK;
  }
  uint64_t aff[12];
  msm_to_affine(curve, f, p, aff);
  memcpy(tgt, aff, 16 * L);
  zkh::store(f, tgt + 2 * L, one);
  return ZK_OK;
}

// per-phase timing of the last MSM (HIP events on the library stream + host wall clock)
// per calling thread (several host threads may each drive their own stream)
static thread_local double g_ms_total = 0, g_ms_phase[8] = {0};
static thread_local int g_nphase = 0;
static thread_local long long g_stats[4] = {0, 0, 0, 0};  // window, windows, sorted entries, heavy buckets

int msm_last_stats(long long *out4) {
  for (int i = 0; i < 4; i++) out4[i] = g_stats[i];
  return ZK_OK;
}

int msm_last_timing(double *ms_total, double *ms_phases, int max_phases, int *nphases) {
  if (ms_total) *ms_total = g_ms_total;
  const int k = g_nphase < max_phases ? g_nphase : max_phases;
  for (int i = 0; i < k; i++) ms_phases[i] = g_ms_phase[i];
  if (nphases) *nphases = g_nphase;
  return ZK_OK;
}

// Phase events are recorded only while phase timing is on: each record puts a timestamp
// barrier in the stream, which idles the GPU for a few microseconds before the next kernel.
static std::atomic<int> g_phase_timing{0};

int msm_set_phase_timing(int on) {
  g_phase_timing.store(on ? 1 : 0);
  return ZK_OK;
}

struct Events {
  hipEvent_t e[8];
  int n = 0, made = 0;
  bool ok = false;
  Events() {
    if (!g_phase_timing.load()) return;
    while (made < 8 && hipEventCreate(&e[made]) == hipSuccess) made++;
    ok = made == 8;
  }
  ~Events() {
    for (int i = 0; i < made; i++) (void)hipEventDestroy(e[i]);
  }
  void mark(hipStream_t st) {
    if (ok && n < 8) (void)hipEventRecord(e[n++], st);
  }
};

// Host-side resources of one running MSM: pinned words for the statistics read-back, pinned
// result words and a side stream.  Taken from a pool and returned after the call, so concurrent
// callers (batch workers, serving threads) each hold one and the pool never grows past the
// peak concurrency.
struct MsmStage {
  uint32_t *info = nullptr;  // heavy count, most partials, sorted entries
  uint32_t *res = nullptr;   // pinned result words
  uint32_t *info_dev = nullptr, *res_dev = nullptr;  // their device addresses (kernels write them)
  size_t res_words = 0;
  hipStream_t aux = nullptr;     // a second stream, created on first use
  hipEvent_t aux_ev = nullptr;
  int init() {
    ZK_HIP(hipHostMalloc((void **)&info, 16, hipHostMallocDefault));
    ZK_HIP(hipHostGetDevicePointer((void **)&info_dev, info, 0));
    return ZK_OK;
  }
  int ensure_aux() {
    if (aux) return ZK_OK;
    ZK_HIP(hipEventCreateWithFlags(&aux_ev, hipEventDisableTiming));
    ZK_HIP(hipStreamCreateWithFlags(&aux, hipStreamNonBlocking));
    return ZK_OK;
  }
  int reserve_res(size_t words) {
    if (words <= res_words) return ZK_OK;
    if (res) ZK_HIP(hipHostFree(res));
    res = res_dev = nullptr;
    res_words = 0;
    ZK_HIP(hipHostMalloc((void **)&res, words * 4, hipHostMallocDefault));
    ZK_HIP(hipHostGetDevicePointer((void **)&res_dev, res, 0));
    res_words = words;
    return ZK_OK;
  }
};

static std::mutex g_stage_mu;
static std::vector<MsmStage *> g_stage_free;

struct StageLease {
  MsmStage *s = nullptr;
  ~StageLease() {
    if (s) {
      std::lock_guard<std::mutex> g(g_stage_mu);
      g_stage_free.push_back(s);
    }
  }
  int acquire() {
    {
      std::lock_guard<std::mutex> g(g_stage_mu);
      if (!g_stage_free.empty()) {
        s = g_stage_free.back();
        g_stage_free.pop_back();
        return ZK_OK;
      }
    }
    MsmStage *n = new MsmStage;
    const int st = n->init();
    if (st != ZK_OK) return st;  // a half-built stage is dropped (its handles stay allocated)
    s = n;
    return ZK_OK;
  }
};

static int msm_set_lds_attributes() {
  ZK_HIP(hipFuncSetAttribute((const void *)msm_part_scatter_staged_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)kScatterStageLds));
  ZK_HIP(hipFuncSetAttribute((const void *)msm_pyramid_tail_kernel<Bn254G1>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)kTailLdsBytes));
  ZK_HIP(hipFuncSetAttribute((const void *)msm_pyramid_tail_kernel<Bls381G1>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)kTailLdsBytes));
  return ZK_OK;
}

int msm_stats_window(int curve, const zkg_scalar_stats &st) {
  (void)curve;
  const double n = (double)st.n, nfull = (double)(st.n - st.n_zero - st.n_one), ones = (double)st.n_one;
  double best = 1e300;
  int bc = 2;
  for (int c = 2; c <= kMaxAnalyticWindow; c++) {
    int w, cm, wide;
    msm_window_layout(st.max_bits, c, true, &w, &cm, &wide);
    const double W = (double)w;
    const double cost = 10.0 * (W * nfull + ones) + 3.0 * W * n + 28.0 * W * (double)(1ll << (cm - 1));
    if (cost < best) {
      best = cost;
      bc = c;
    }
  }
  return bc;
}

int msm_scalar_stats(int curve, int n, const uint64_t *d_scalars, int enl, int mont, hipStream_t st,
                     zkg_scalar_stats *out) {
  memset(out, 0, sizeof *out);
  out->n = n;
  if (n <= 0) return ZK_OK;
  DevBuf d;
  ZK_TRY(d.alloc(16));
  ZK_HIP(hipMemsetAsync(d.p, 0, 16, st));
  const dim3 g(std::min<unsigned>((unsigned)((n + kThreads - 1) / kThreads), kStatsBlocks)), b(kThreads);
  if (curve == 0)
    hipLaunchKernelGGL(msm_stats_kernel<Bn254G1>, g, b, 0, st, d_scalars, enl, mont, (uint32_t)n, d.as<uint32_t>());
  else
    hipLaunchKernelGGL(msm_stats_kernel<Bls381G1>, g, b, 0, st, d_scalars, enl, mont, (uint32_t)n, d.as<uint32_t>());
  ZK_HIP(hipGetLastError());
  uint32_t h[4];
  ZK_HIP(hipMemcpyAsync(h, d.p, 16, hipMemcpyDeviceToHost, st));
  ZK_HIP(hipStreamSynchronize(st));
  out->max_bits = (int)h[0];
  out->n_zero = h[1];
  out->n_one = h[2];
  return ZK_OK;
}

// Schedule of one MSM.  Without statistics the windows cover the declared scalar width
// (64 * enl bits, plus one for the signed-digit carry); with them they cover max_bits + 1.
int msm_resolve(int curve, int n, int enl, const zkg_msm_schedule *user, const zkg_scalar_stats *st,
                MsmSchedule *out, bool glv_ok) {
  MsmSchedule s;
  // GLV split (BN254, 4-limb scalars reduced mod r; affine words get their phi(P_i) records in
  // the digit pass, prepared sets hold them): 2n points P_i and phi(P_i) with scalars below 2^127, so windows over 128 bits; entries stay ~16 n at
  // c = 16 while the buckets and the host's doubling chain halve.  Measured scalars of at most
  // 128 bits are already short and keep the plain layout.
  const bool glv = glv_ok && curve == 0 && enl == 4 && (!st || st->max_bits > 128);
  const int fr_bits = curve == 0 ? 254 : 255;
  const int bits_for_cost = enl * 64 < fr_bits ? enl * 64 : fr_bits;
  const int policy = user ? user->policy : ZKG_SCHED_PLAIN;
  int c = (user && user->window > 0) ? user->window : 0;
  int seg = (user && user->seg_log > 0) ? user->seg_log : 0;
  if (st && (c == 0 || seg == 0) && policy == ZKG_SCHED_LEARNED) {
    int mc = 0, ms = 0;
    if (cost_model_pick(curve, *st, glv_ok && enl == 4, &mc, &ms)) {
      if (c == 0) c = mc;
      if (seg == 0) seg = ms;
    }
  }
  if (c == 0) {
    if (glv) c = msm_heuristic_window(curve, 2ll * n, 127);
    else c = st ? msm_stats_window(curve, *st) : msm_heuristic_window(curve, n, bits_for_cost);
  }
  if (seg == 0) seg = 5;
  c = c < 1 ? 1 : (c > 20 ? 20 : c);
  s.seg_log = seg < 1 ? 1 : (seg > 12 ? 12 : seg);
  s.bits = st ? st->max_bits : 64 * enl;
  // BN254's G1 has prime order r, so s P = (s mod r) P for every curve point: 4-limb scalars
  // reduced mod r need windows for 254 bits (plus the carry), not 256
  if (!st && curve == 0 && enl == 4) {
    s.bits = 254;
    s.reduce = 1;
  }
  if (glv) {
    s.bits = 127;
    s.reduce = 1;
    s.glv = 1;
  }
  msm_window_layout(s.bits, c, st != nullptr, &s.windows, &s.c, &s.wide);
  *out = s;
  return ZK_OK;
}

void msm_window_layout(int bits, int c_req, bool balanced, int *W, int *c, int *wide) {
  const int total = bits + 1;
  const int w = (total + c_req - 1) / c_req;
  *W = w;
  if (!balanced) {
    *c = c_req;
    *wide = w;
    return;
  }
  const int cm = (total + w - 1) / w;
  *c = cm;
  *wide = total - w * (cm - 1);
}

static MsmSchedule resolve_schedule(int curve, int n, int enl, const zkg_msm_schedule *user, bool glv_ok) {
  MsmSchedule s;
  msm_resolve(curve, n, enl, user, nullptr, &s, glv_ok);
  return s;
}

// sort partitions per window: about 2048 digits per partition for uniform scalars, at most
// 2^10 partitions (level-1 LDS counters) and 2^12 buckets per partition (level-2 counters)
static int sort_partition_bits(int n, int logB) {
  int logP = 0;
  while ((1ull << (logP + 11)) < (uint64_t)n) logP++;
  logP = std::min(logP, std::min(logB, 10));
  return std::max(logP, logB - 12);
}

template <class C>
static int msm_sparse_run(int n, const uint64_t *d_scalars, int enl, int mont, const uint64_t *d_points,
                          bool points_internal, const zkg_msm_schedule *user, hipStream_t st,
                          const zkg_scalar_stats &stats, MsmResult *out);

// ---------------------------------------------------------------------------------------
// One dense MSM pipeline: its schedule, launch sizes and device buffers.  msm_run queues its
// launches on a stream and waits once, for the result words; the captured batch (msm_graph_*)
// records the same launches into a graph.  No launch size depends on the data.

template <class C>
struct MsmPipe {
  using F = typename C::Fp;
  static constexpr int PW = 4 * F::N;          // u32 words per external XYZZ point (host side)
  static constexpr int ZW = ZLayout<C>::XYZZ;  // words per internal XYZZ point (device buffers)
  static constexpr int AW = ZLayout<C>::AFF;   // words per internal affine point
  int n = 0, enl = 4, mont = 0;
  const uint64_t *d_scalars = nullptr, *d_points = nullptr;
  bool points_internal = false, prefetch = false;
  uint32_t *ovf = nullptr;  // set by the digits pass when a scalar exceeds sch.bits (captured plans)
  MsmSchedule sch;
  uint32_t limit = 16;
  uint32_t np = 0, B = 0, P = 0, nchunks = 0, nsb = 0, heavy_cap = 0, A = 0, nres = 0;
  int c = 0, W = 0, logB = 0, seg_log = 0, L = 0, ib = 0;
  int pyr_final = -1;  // buffer holding the pyramid's output: -1 bsum, 0 pyrA, 1 pyrB
  bool narrow = false;
  uint64_t nkeys = 0, total = 0, nseg = 0, nph = 0, nbig_cap = 0;
  uint32_t hblocks = 0;  // workgroups of the heavy-bucket launches
  DevBuf dig, vals, tmp, phist, ppos, big, rcnt, off, bsum, cont, heavy, hsum, pyrA, pyrB, scan_b, zpts, outx;

  const int32_t *zp() const { return points_internal ? (const int32_t *)d_points : zpts.as<int32_t>(); }
  const int32_t *pyr_out() const {
    return pyr_final < 0 ? bsum.as<int32_t>() : (pyr_final == 0 ? pyrA.as<int32_t>() : pyrB.as<int32_t>());
  }

  int setup(int n_, const uint64_t *sc, int enl_, int mont_, const uint64_t *pts, bool internal,
            const MsmSchedule &s, const zkg_msm_schedule *user) {
    n = n_;
    d_scalars = sc;
    enl = enl_;
    mont = mont_;
    d_points = pts;
    points_internal = internal;
    sch = s;
    prefetch = user && user->prefetch == 2;
    np = sch.glv ? 2u * (uint32_t)n : (uint32_t)n;  // points of the bucket problem
    c = sch.c;
    W = sch.windows;
    logB = c - 1;
    B = 1u << logB;
    nkeys = (uint64_t)W * B;
    // Buckets spanning more than `limit` segments go to the heavy-bucket kernels; the others
    // add their partials one after another in the fix kernel (latency ~ one addition each).
    // The heavy kernels give each bucket one or more workgroups while there are at most
    // kHeavyBlocks of them, so with at most that many buckets in all, 16 suits every digit
    // set.  With more buckets and all of them long (narrow windows over many points: a 13-bit
    // window over 2^21 points puts ~22 segments in each of 20 000 buckets), the workgroups
    // fold bucket after bucket: 7 ms against ~0.15 ms for the serial fix.  There the limit is
    // twice the mean segments per bucket of a dense digit set, at most 64 (a longer serial
    // chain costs more than a workgroup fold).
    if (user && user->fix_limit > 0) {
      limit = (uint32_t)user->fix_limit;
    } else if (nkeys <= kHeavyBlocks) {
      limit = 16;
    } else {
      const uint64_t mean = ((uint64_t)W * np + (nkeys << sch.seg_log) - 1) / (nkeys << sch.seg_log);
      limit = (uint32_t)std::max<uint64_t>(16, std::min<uint64_t>(2 * mean, 64));
    }
    if (nkeys >= (1ull << 32) || (uint64_t)W * np >= (1ull << 32) || np >= (1u << 31)) {
      set_error("msm: problem too large for 32-bit sort keys");
      return ZK_EINVAL;
    }
    total = (uint64_t)W * np;
    seg_log = sch.seg_log;
    nseg = (total >> seg_log) + 1;
    const int logP = sort_partition_bits((int)np, logB);
    L = logB - logP;
    P = 1u << logP;
    nchunks = (uint32_t)(((uint64_t)np + kPChunk - 1) / kPChunk);
    nph = (uint64_t)W * P * nchunks;  // level-1 counts, scanned in place of positions
    ZK_TRY(dig.alloc(total * 4));
    ZK_TRY(vals.alloc(total * 4));
    // level-1 words are 32-bit when the point index, the sign and the low bucket bits fit
    ib = 0;
    while ((1ull << ib) < (uint64_t)np) ib++;
    narrow = ib < 31 && ib + 1 + L <= 32;
    ZK_TRY(tmp.alloc(total * (narrow ? 4 : 8)));
    ZK_TRY(phist.alloc(nph * 4));
    ZK_TRY(ppos.alloc((nph + 1) * 4));
    // big partitions hold more than kBigPart entries each, so there are at most total / kBigPart
    nbig_cap = std::min<uint64_t>((uint64_t)W * P, total / kBigPart + 1);
    ZK_TRY(big.alloc(((uint64_t)W * P + 1) * 4));
    ZK_TRY(rcnt.alloc(nbig_cap * nchunks * ((uint64_t)1 << L) * 4));
    ZK_TRY(off.alloc((nkeys + 1) * 4));
    ZK_TRY(bsum.alloc(nkeys * ZW * 4));
    ZK_TRY(cont.alloc(nseg * ZW * 4));
    if (!points_internal) ZK_TRY(zpts.alloc((uint64_t)n * ZLayout<C>::STRIDE * AW * 4));
    // every bucket may be heavy (small seg_log, few entries per segment)
    heavy_cap = (uint32_t)nkeys;
    ZK_TRY(heavy.alloc(((uint64_t)heavy_cap + 3) * 4));
    hblocks = (uint32_t)std::min<uint64_t>(heavy_cap, kHeavyBlocks);
    ZK_TRY(hsum.alloc((uint64_t)hblocks * ZW * 4));
    ZK_TRY(pyrA.alloc(nkeys * ZW * 4));
    ZK_TRY(pyrB.alloc(nkeys * ZW * 4));
    const uint64_t per_block = (uint64_t)kThreads * kScanPer;
    nsb = (uint32_t)((nph + 1 + per_block - 1) / per_block);
    ZK_TRY(scan_b.alloc((uint64_t)nsb * 4));
    pyramid(nullptr, false);  // A arrays of W points come out of the pyramid
    nres = A * (uint32_t)W;
    ZK_TRY(outx.alloc((uint64_t)nres * PW * 4));
    return ZK_OK;
  }

  // digits, level-1 histogram and scan, level-1 scatter, level-2 sort (bucket offsets, sorted
  // entries, heavy buckets listed)
  void launch_sort(hipStream_t st) {
    const dim3 blk(kThreads);
    const dim3 gpts((n + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(msm_digits_kernel<C>, gpts, blk, 0, st, d_scalars, enl, mont, sch.reduce, sch.glv,
                       (uint32_t)n, np, c, W, sch.wide, ovf, dig.as<uint32_t>(),
                       points_internal ? (const uint32_t *)nullptr : (const uint32_t *)d_points, zpts.as<int32_t>());
    const dim3 gch(nchunks, W);
    hipLaunchKernelGGL(msm_part_hist_kernel, gch, dim3(kSortThreads), (size_t)P * 4, st, dig.as<uint32_t>(), np, L,
                       P, phist.as<uint32_t>(), heavy.as<uint32_t>(), big.as<uint32_t>());
    hipLaunchKernelGGL(scan_block_sums_kernel, dim3(nsb), blk, 0, st, phist.as<uint32_t>(), (size_t)nph,
                       scan_b.as<uint32_t>());
    hipLaunchKernelGGL(scan_final_kernel, dim3(nsb), blk, 0, st, phist.as<uint32_t>(), (size_t)nph,
                       scan_b.as<uint32_t>(), ppos.as<uint32_t>(), (uint32_t *)nullptr);
    const size_t cnt_bytes = ((size_t)1 << std::max(L, 1)) * 4;
    // level 2: partitions of at most kBigPart words, one workgroup each; larger ones run by run
    auto level2 = [&](auto word) {
      using T = decltype(word);
      const uint32_t stage = (uint32_t)(kPartStageBytes / sizeof(T));
      hipLaunchKernelGGL(msm_part_sort_kernel<T>, dim3((unsigned)(W * P)), blk,
                         cnt_bytes + kPartStageBytes + (size_t)stage * 4, st, tmp.as<T>(), ppos.as<uint32_t>(),
                         nchunks, L, ib, stage, off.as<uint32_t>(), vals.as<uint32_t>(), big.as<uint32_t>());
      const unsigned gb = (unsigned)std::min<uint64_t>((uint64_t)nbig_cap * nchunks, 1024);
      hipLaunchKernelGGL(msm_big_count_kernel<T>, dim3(gb), blk, cnt_bytes, st, tmp.as<T>(), ppos.as<uint32_t>(),
                         nchunks, L, ib, big.as<uint32_t>(), rcnt.as<uint32_t>());
      hipLaunchKernelGGL(msm_big_scan_kernel, dim3((unsigned)std::min<uint64_t>(nbig_cap, 256)), blk, 0, st,
                         ppos.as<uint32_t>(), nchunks, L, big.as<uint32_t>(), rcnt.as<uint32_t>(), off.as<uint32_t>());
      hipLaunchKernelGGL(msm_big_place_kernel<T>, dim3(gb), blk, cnt_bytes, st, tmp.as<T>(), ppos.as<uint32_t>(),
                         nchunks, L, ib, big.as<uint32_t>(), rcnt.as<uint32_t>(), vals.as<uint32_t>(),
                         off.as<uint32_t>(), (uint32_t)nkeys, seg_log, heavy.as<uint32_t>(), heavy_cap, limit);
    };
    if (narrow) {
      hipLaunchKernelGGL(msm_part_scatter_staged_kernel, gch, dim3(kSortThreads), kScatterStageLds, st,
                         dig.as<uint32_t>(), np, L, P, ib, phist.as<uint32_t>(), ppos.as<uint32_t>(),
                         tmp.as<uint32_t>());
      level2(uint32_t(0));
    } else {
      hipLaunchKernelGGL(msm_part_scatter_kernel<uint64_t>, gch, dim3(kSortThreads), (size_t)P * 4, st,
                         dig.as<uint32_t>(), np, L, P, ib, ppos.as<uint32_t>(), tmp.as<uint64_t>());
      level2(uint64_t(0));
    }
  }

  void launch_accum(hipStream_t st) {
    const dim3 blk(kThreads), g((unsigned)((nseg + kThreads - 1) / kThreads));
    if (prefetch)
      hipLaunchKernelGGL((msm_accum_kernel<C, true>), g, blk, 0, st, vals.as<uint32_t>(), off.as<uint32_t>(),
                         (uint32_t)nkeys, seg_log, zp(), (uint32_t)n, bsum.as<int32_t>(), cont.as<int32_t>());
    else
      hipLaunchKernelGGL((msm_accum_kernel<C, false>), g, blk, 0, st, vals.as<uint32_t>(), off.as<uint32_t>(),
                         (uint32_t)nkeys, seg_log, zp(), (uint32_t)n, bsum.as<int32_t>(), cont.as<int32_t>());
  }

  void launch_fix(hipStream_t st) {
    hipLaunchKernelGGL(msm_fix_kernel<C>, dim3((unsigned)((nkeys + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                       st, off.as<uint32_t>(), (uint32_t)nkeys, seg_log, bsum.as<int32_t>(), cont.as<int32_t>(),
                       limit);
  }

  // heavy buckets: the two device-driven launches (they exit at once when none is listed)
  void launch_heavy(hipStream_t st) {
    const size_t lds = (size_t)kThreads * ZW * 4;
    hipLaunchKernelGGL(msm_heavy_part_kernel<C>, dim3(hblocks), dim3(kThreads), lds, st, off.as<uint32_t>(),
                       heavy.as<uint32_t>(), heavy_cap, hblocks, seg_log, bsum.as<int32_t>(), cont.as<int32_t>(),
                       hsum.as<int32_t>());
    hipLaunchKernelGGL(msm_heavy_combine_kernel<C>, dim3(hblocks), dim3(kThreads), lds, st, off.as<uint32_t>(),
                       heavy.as<uint32_t>(), heavy_cap, hblocks, bsum.as<int32_t>(), hsum.as<int32_t>());
  }

  // Bucket pyramid over the W windows of B bucket sums; launch = false only works out the
  // number A of output arrays and the buffer that holds them.
  void pyramid(hipStream_t st, bool launch) {
    const dim3 blk(kThreads);
    const int32_t *in = bsum.as<int32_t>();
    int32_t *bufs[2] = {pyrA.as<int32_t>(), pyrB.as<int32_t>()};
    int fin = -1;
    uint32_t a = 1, M = B;
    int k = 0;
    while (M > 1) {
      // the remaining levels in one launch once they fit one workgroup per window (level 0 at
      // most two items per thread, later levels one; level 0's output in LDS)
      const uint32_t items0 = (a + 1) * (M / 2), items1 = (a + 2) * (M / 4);
      const size_t tail_lds = (size_t)items0 * ZW * 4;
      if (items0 <= 2u * kTailThreads && items1 <= (uint32_t)kTailThreads && tail_lds <= kTailLdsBytes) {
        int levels = 0;
        while ((M >> levels) > 1) levels++;
        // one item per thread at level 0 when that fits one workgroup (each round is one more
        // addition of latency), else two
        const unsigned tb = std::min((items0 + 63) & ~63u, (unsigned)kTailThreads);
        if (launch)
          hipLaunchKernelGGL(msm_pyramid_tail_kernel<C>, dim3(1u, (unsigned)W), dim3(tb), tail_lds, st, in, bufs[k],
                             (uint32_t)W, M, a, M);
        in = bufs[k];
        fin = k;
        a += levels;
        M = 1;
        break;
      }
      // the largest block of Mb points per array whose level 0 is one item per thread
      uint32_t Mb = M / 2;
      while (Mb >= 4 && ((a + 1) * (Mb / 2) > (uint32_t)kTailThreads ||
                         (size_t)(a + 1) * (Mb / 2) * ZW * 4 > kTailLdsBytes))
        Mb >>= 1;
      // two levels per launch while that still fills the GPU (one thread per 4 points, 3-4
      // dependent additions each); below, the levels of each Mb-point block in one launch (one
      // addition per thread per level, no launch between levels), or a level per launch
      if (M >= 4 && (uint64_t)a * W * (M / 4) >= kPyr2MinThreads) {
        const uint64_t thr = (uint64_t)(a + 1) * W * (M / 4);
        if (launch)
          hipLaunchKernelGGL(msm_pyramid2_kernel<C>, dim3((unsigned)((thr + kThreads - 1) / kThreads)), blk, 0, st,
                             in, bufs[k], (uint32_t)W, M, a);
        a += 2;
        M >>= 2;
      } else if (Mb >= 4) {
        const uint32_t items = (a + 1) * (Mb / 2);
        int levels = 0;
        while ((Mb >> levels) > 1) levels++;
        if (launch)
          hipLaunchKernelGGL(msm_pyramid_tail_kernel<C>, dim3(M / Mb, (unsigned)W), dim3((items + 63) & ~63u),
                             (size_t)items * ZW * 4, st, in, bufs[k], (uint32_t)W, M, a, Mb);
        a += levels;
        M /= Mb;
      } else {
        const uint64_t thr = (uint64_t)a * W * (M / 2);
        if (launch)
          hipLaunchKernelGGL(msm_pyramid_kernel<C>, dim3((unsigned)((thr + kThreads - 1) / kThreads)), blk, 0, st,
                             in, bufs[k], (uint32_t)W, M, a);
        a += 1;
        M >>= 1;
      }
      in = bufs[k];
      fin = k;
      k ^= 1;
    }
    A = a;
    pyr_final = fin;
  }

  // the A arrays of W points ([0] = Sum_w, [j] = U_{w, j-1}) as external canonical words, into
  // outx or straight into dst (pinned host words, by their device address); with info_dst also
  // the three heavy-bucket statistics words
  void launch_out(hipStream_t st, uint32_t *dst = nullptr, uint32_t *info_dst = nullptr) {
    hipLaunchKernelGGL(msm_points_out_kernel<C>, dim3((nres + kThreads - 1) / kThreads), dim3(kThreads), 0, st,
                       pyr_out(), nres, dst ? dst : outx.as<uint32_t>(), (const uint32_t *)heavy.as<uint32_t>(),
                       info_dst);
  }

  // host Horner: sum over (w, j) of 2^(start_w + max(j-1, 0)) * term
  void combine(const uint32_t *res, MsmResult *out) const {
    const zkh::FieldDesc &f = fp_desc(C::id);
    auto get = [&](uint32_t a, uint32_t w) {
      const uint32_t *p = res + ((size_t)a * W + w) * PW;
      zkh::Xyzz q;
      zkh::Fe *dst[4] = {&q.x, &q.y, &q.zz, &q.zzz};
      for (int e = 0; e < 4; e++) {
        *dst[e] = zkh::zero();
        for (int j = 0; j < F::N; j++) dst[e]->w[j / 2] |= (uint64_t)p[e * F::N + j] << (32 * (j & 1));
      }
      return q;
    };
    const int emax = sch.start(W - 1) + (int)(A > 1 ? A - 2 : 0);
    std::vector<std::vector<uint32_t>> at(emax + 1);  // encoded (a << 16) | w
    for (uint32_t w = 0; w < (uint32_t)W; w++) {
      const int s0 = sch.start((int)w);
      at[s0].push_back((0u << 16) | w);
      for (uint32_t a = 1; a < A; a++) at[s0 + (a - 1)].push_back((a << 16) | w);
    }
    zkh::Xyzz acc = zkh::xyzz_inf();
    for (int e = emax; e >= 0; e--) {
      acc = zkh::xyzz_dbl(f, acc);
      for (uint32_t code : at[e]) acc = zkh::xyzz_add(f, acc, get(code >> 16, code & 0xffff));
    }
    memset(out, 0, sizeof *out);
    out->curve = C::id;
    zkh::store(f, out->x, acc.x);
    zkh::store(f, out->y, acc.y);
    zkh::store(f, out->zz, acc.zz);
    zkh::store(f, out->zzz, acc.zzz);
  }
};

static int msm_lds_attributes_once() {
  static std::once_flag attr_once;
  static int attr_status = ZK_OK;
  std::call_once(attr_once, [&] { attr_status = msm_set_lds_attributes(); });
  return attr_status;
}

template <class C>
static int msm_run(int n, const uint64_t *d_scalars, int enl, int mont, const uint64_t *d_points,
                   bool points_internal, const zkg_msm_schedule *user, hipStream_t st, MsmResult *out) {
  zkg_scalar_stats stats;
  const bool use_stats = user && user->policy != ZKG_SCHED_PLAIN;
  if (use_stats) ZK_TRY(msm_scalar_stats(C::id, n, d_scalars, enl, mont, st, &stats));
  // Measured sparse scalar sets (at least half the scalars 0 or 1): the other scalars and
  // their points are compacted into a dense MSM of their own, the points of the scalars
  // equal to 1 are summed directly, and the zeros drop out; the two sums are added on the host.
  // Zeros and ones cost the plain pipeline little beyond their digits, so the compaction pays
  // only while its own passes are small next to the MSM: measured faster up to 2^20 points,
  // slower from 2^21 (BN254 2^22, 80% in {0, 1}: 2.82 ms compacted against 2.52 ms).
  if (use_stats && n >= (1 << 14) && n < (1 << 21) && (uint64_t)(stats.n_zero + stats.n_one) * 2 >= (uint64_t)n)
    return msm_sparse_run<C>(n, d_scalars, enl, mont, d_points, points_internal, user, st, stats, out);
  MsmSchedule sch;
  ZK_TRY(msm_resolve(C::id, n, enl, user, use_stats ? &stats : nullptr, &sch, true));
  // every buffer the queued kernels touch lives until the stream is drained (the guard below)
  MsmPipe<C> p;
  StreamSyncGuard drain(st);
  ZK_TRY(p.setup(n, d_scalars, enl, mont, d_points, points_internal, sch, user));
  ZK_TRY(msm_lds_attributes_once());

  Events ev;
  auto t_host0 = std::chrono::steady_clock::now();
  ev.mark(st);
  p.launch_sort(st);
  ZK_HIP(hipGetLastError());
  // heavy count, most partials and sorted entries (statistics, read after the MSM)
  StageLease lease;
  ZK_TRY(lease.acquire());
  MsmStage &stage = *lease.s;
  uint32_t *info = stage.info;
  ZK_TRY(stage.reserve_res((size_t)p.nres * MsmPipe<C>::PW));
  ev.mark(st);
  p.launch_accum(st);
  ZK_HIP(hipGetLastError());
  ev.mark(st);
  p.launch_fix(st);
  p.launch_heavy(st);
  ZK_HIP(hipGetLastError());
  ev.mark(st);
  p.pyramid(st, true);
  ZK_HIP(hipGetLastError());
  ev.mark(st);
  // the result words and the statistics go straight to pinned host memory: no copy launches
  p.launch_out(st, stage.res_dev, stage.info_dev);
  ZK_HIP(hipGetLastError());
  ZK_HIP(hipStreamSynchronize(st));
  drain.release();
  ev.mark(st);
  auto t_host1 = std::chrono::steady_clock::now();
  g_stats[0] = p.c;
  g_stats[1] = p.W;
  g_stats[2] = info[2];
  g_stats[3] = std::min(info[0], p.heavy_cap);
  p.combine(stage.res, out);
  auto t_host2 = std::chrono::steady_clock::now();

  // timing: sort, accumulate, fix, pyramid, output (into pinned host memory, and the wait), host combine
  g_nphase = 0;
  if (ev.ok) {
    for (int i = 0; i + 1 < ev.n; i++) {
      float ms = 0;
      (void)hipEventElapsedTime(&ms, ev.e[i], ev.e[i + 1]);
      g_ms_phase[g_nphase++] = ms;
    }
  }
  g_ms_phase[g_nphase++] = std::chrono::duration<double, std::milli>(t_host2 - t_host1).count();
  g_ms_total = std::chrono::duration<double, std::milli>(t_host2 - t_host0).count();
  return ZK_OK;
}

template <class C>
static int msm_sparse_run(int n, const uint64_t *d_scalars, int enl, int mont, const uint64_t *d_points,
                          bool points_internal, const zkg_msm_schedule *user, hipStream_t st,
                          const zkg_scalar_stats &stats, MsmResult *out) {
  using F = typename C::Fp;
  // u32 words per point (internal: its records, P_i and phi(P_i) for BN254)
  const int pw = points_internal ? ZLayout<C>::AFF * ZLayout<C>::STRIDE : 2 * F::N;
  const uint64_t k = (uint64_t)n - stats.n_zero - stats.n_one;   // scalars outside {0, 1}
  DevBuf keep, pos, ones, scan_b, sc2, pts2;
  // drains the stream on every return: with no dense part (k = 0) nothing else waits for the
  // gather kernels that read these buffers
  StreamSyncGuard drain(st);
  ZK_TRY(keep.alloc((uint64_t)n * 4));
  ZK_TRY(pos.alloc(((uint64_t)n + 1) * 4));
  ZK_TRY(ones.alloc((uint64_t)n));
  const uint64_t per_block = (uint64_t)kThreads * kScanPer;
  const uint32_t nsb = (uint32_t)(((uint64_t)n + 1 + per_block - 1) / per_block);
  ZK_TRY(scan_b.alloc((uint64_t)nsb * 4));
  ZK_TRY(sc2.alloc((k ? k : 1) * enl * 8));
  ZK_TRY(pts2.alloc((k ? k : 1) * pw * 4));
  const dim3 blk(kThreads), g((n + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(msm_sparse_flags_kernel<C>, g, blk, 0, st, d_scalars, enl, mont, (uint32_t)n,
                     keep.as<uint32_t>(), ones.as<uint8_t>());
  hipLaunchKernelGGL(scan_block_sums_kernel, dim3(nsb), blk, 0, st, keep.as<uint32_t>(), (size_t)n,
                     scan_b.as<uint32_t>());
  hipLaunchKernelGGL(scan_final_kernel, dim3(nsb), blk, 0, st, keep.as<uint32_t>(), (size_t)n,
                     scan_b.as<uint32_t>(), pos.as<uint32_t>(), (uint32_t *)nullptr);
  hipLaunchKernelGGL(msm_sparse_gather_kernel, g, blk, 0, st, d_scalars, enl, (const uint32_t *)d_points, pw,
                     (uint32_t)n, keep.as<uint32_t>(), pos.as<uint32_t>(), sc2.as<uint64_t>(), pts2.as<uint32_t>());
  ZK_HIP(hipGetLastError());
  // the direct sum of the ones is latency-bound: it runs on a second stream, from a second
  // host thread, beside the dense MSM
  StageLease lease;
  ZK_TRY(lease.acquire());
  ZK_TRY(lease.s->ensure_aux());
  ZK_HIP(hipEventRecord(lease.s->aux_ev, st));
  ZK_HIP(hipStreamWaitEvent(lease.s->aux, lease.s->aux_ev, 0));
  MsmResult dense, ones_sum;
  int ones_status = ZK_OK;
  std::string ones_err;  // error messages are per thread
  std::thread side([&] {
    ones_status = binary_sum_device(C::id, 1, n, ones.as<uint8_t>(), d_points, points_internal, lease.s->aux,
                                    &ones_sum);
    if (ones_status != ZK_OK) ones_err = last_error();
  });
  int dense_status = ZK_OK;
  if (k) dense_status = msm_run<C>((int)k, sc2.as<uint64_t>(), enl, mont, pts2.as<uint64_t>(), points_internal, user,
                                   st, &dense);
  else msm_infinity(C::id, &dense);
  side.join();
  ZK_TRY(dense_status);
  if (ones_status != ZK_OK) {
    set_error(ones_err);
    return ones_status;
  }
  const zkh::FieldDesc &f = fp_desc(C::id);
  auto load = [&](const MsmResult &r) {
    zkh::Xyzz q;
    q.x = zkh::load(f, r.x);
    q.y = zkh::load(f, r.y);
    q.zz = zkh::load(f, r.zz);
    q.zzz = zkh::load(f, r.zzz);
    return q;
  };
  const zkh::Xyzz sum = zkh::xyzz_add(f, load(dense), load(ones_sum));
  memset(out, 0, sizeof *out);
  out->curve = C::id;
  zkh::store(f, out->x, sum.x);
  zkh::store(f, out->y, sum.y);
  zkh::store(f, out->zz, sum.zz);
  zkh::store(f, out->zzz, sum.zzz);
  return ZK_OK;
}

int msm_device(int curve, int n, const uint64_t *d_scalars, int enl, int mont, const uint64_t *d_points,
               bool points_internal,
               const zkg_msm_schedule *sched, hipStream_t st, MsmResult *out) {
  if (n <= 0) {
    msm_infinity(curve, out);
    return ZK_OK;
  }
  if (curve == 0) return msm_run<Bn254G1>(n, d_scalars, enl, mont, d_points, points_internal, sched, st, out);
  if (curve == 1) return msm_run<Bls381G1>(n, d_scalars, enl, mont, d_points, points_internal, sched, st, out);
  set_error("msm: bad curve id");
  return ZK_EINVAL;
}

template <class C>
static int msm_prepare_points_t(int n, const uint64_t *d_ext, void *d_int, bool phi, hipStream_t st) {
  hipLaunchKernelGGL(msm_points_in_kernel<C>, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, st,
                     (const uint32_t *)d_ext, (uint32_t)n, (int32_t *)d_int, (int)phi);
  ZK_HIP(hipGetLastError());
  ZK_HIP(hipStreamSynchronize(st));
  return ZK_OK;
}

size_t msm_internal_point_bytes(int curve) {
  return curve == 0 ? ZLayout<Bn254G1>::AFF * ZLayout<Bn254G1>::STRIDE * 4
                    : ZLayout<Bls381G1>::AFF * ZLayout<Bls381G1>::STRIDE * 4;
}

int msm_prepare_points(int curve, int n, const uint64_t *d_ext, void *d_int, hipStream_t st, bool phi) {
  if (n <= 0) return ZK_OK;
  if (curve == 0) return msm_prepare_points_t<Bn254G1>(n, d_ext, d_int, phi, st);
  if (curve == 1) return msm_prepare_points_t<Bls381G1>(n, d_ext, d_int, phi, st);
  set_error("msm_prepare_points: bad curve id");
  return ZK_EINVAL;
}

// device bytes one MSM of n points allocates (admission control of the serving loop)
size_t msm_footprint(int curve, int n, int enl, bool points_internal, const zkg_msm_schedule *sched) {
  const MsmSchedule s = resolve_schedule(curve, n, enl, sched, true);
  const uint64_t np = s.glv ? 2 * (uint64_t)n : (uint64_t)n;
  const uint64_t W = s.windows, B = 1ull << (s.c - 1), nk = W * B, total = W * np;
  const uint64_t zw = curve == 0 ? ZLayout<Bn254G1>::XYZZ : ZLayout<Bls381G1>::XYZZ;
  const uint64_t aw = curve == 0 ? ZLayout<Bn254G1>::AFF : ZLayout<Bls381G1>::AFF;
  const uint64_t nchunks = (np + kPChunk - 1) / kPChunk;
  const int logP = sort_partition_bits((int)np, s.c - 1), L = s.c - 1 - logP;
  const uint64_t nph = W * ((uint64_t)1 << logP) * nchunks;
  const uint64_t nbig = std::min<uint64_t>(W << logP, total / kBigPart + 1);
  // digits, entries and level-1 words; offsets and heavy list; bucket sums and two pyramid
  // buffers; continuation partials; level-1 counts and positions, big-partition list and run
  // counts
  uint64_t b = total * 16 + nk * 8 + 3 * nk * zw * 4 + ((total >> s.seg_log) + 1) * zw * 4 + nph * 8 +
               (W << logP) * 4 + nbig * nchunks * ((uint64_t)1 << L) * 4;
  if (!points_internal) b += (uint64_t)n * msm_internal_point_bytes(curve);
  return (size_t)b;
}

// SRS scalars tau^i in standard form (KZG.hs:59-62 builds [tau^i G]_{i<n})
template <class C>
__global__ __launch_bounds__(kThreads) void srs_scalars_kernel(Fe<typename C::Fr> tau, uint32_t n,
                                                              uint32_t *__restrict__ out) {
  using FR = typename C::Fr;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Fe<FR> r;
  fpow_u64(r, tau, i);
  fto_std(r, r);
  fstore(out + (size_t)i * FR::N, r);
}

int srs_generate(int curve, int n, const uint64_t *tau_std, uint64_t *d_out, hipStream_t st) {
  if (n <= 0) return ZK_OK;
  const zkh::FieldDesc &fr = curve == 0 ? zkh::BN254_FR : zkh::BLS381_FR;
  const zkh::Fe t = zkh::from_std(fr, zkh::load(fr, tau_std));
  DevBuf sc;
  ZK_TRY(sc.alloc((size_t)n * 32));
  const dim3 g((n + kThreads - 1) / kThreads), b(kThreads);
  if (curve == 0) {
    Fe<Bn254Fr> td;
    for (int j = 0; j < 8; j++) td.w[j] = (uint32_t)(t.w[j / 2] >> (32 * (j & 1)));
    hipLaunchKernelGGL(srs_scalars_kernel<Bn254G1>, g, b, 0, st, td, (uint32_t)n, sc.as<uint32_t>());
  } else if (curve == 1) {
    Fe<Bls381Fr> td;
    for (int j = 0; j < 8; j++) td.w[j] = (uint32_t)(t.w[j / 2] >> (32 * (j & 1)));
    hipLaunchKernelGGL(srs_scalars_kernel<Bls381G1>, g, b, 0, st, td, (uint32_t)n, sc.as<uint32_t>());
  } else {
    set_error("srs_generate: bad curve id");
    return ZK_EINVAL;
  }
  ZK_HIP(hipGetLastError());
  return g1_scalar_mul_gen(curve, n, sc.as<uint64_t>(), d_out, st);
}

int g1_scalar_mul_gen(int curve, int n, const uint64_t *d_scalars, uint64_t *d_out, hipStream_t st) {
  if (n <= 0) return ZK_OK;
  const dim3 g((n + kThreads - 1) / kThreads), b(kThreads);
  if (curve == 0)
    hipLaunchKernelGGL(scalar_mul_gen_kernel<Bn254G1>, g, b, 0, st, d_scalars, (uint32_t)n, (uint32_t *)d_out);
  else if (curve == 1)
    hipLaunchKernelGGL(scalar_mul_gen_kernel<Bls381G1>, g, b, 0, st, d_scalars, (uint32_t)n, (uint32_t *)d_out);
  else {
    set_error("scalar_mul_gen: bad curve id");
    return ZK_EINVAL;
  }
  ZK_HIP(hipGetLastError());
  ZK_HIP(hipStreamSynchronize(st));
  return ZK_OK;
}

// ---------------------------------------------------------------------------------------
// Captured MSM batch.  create() fits every job's schedule (its policy, on the scalars present
// then), allocates every job's buffers and queues every job's pipeline once into a hipGraph:
// the jobs are dealt, largest first, to `nstreams` streams forked from and joined back into one
// capture stream, so the graph's branches run concurrently.  run() launches the graph once,
// waits for it, and combines the jobs' window sums on the host.  Nothing inside a replay waits
// on the host: every launch size is fixed at capture.  A replay with scalars the fitted windows do not cover (a wider bit length than
// at create) is caught by the digits pass and reported as an error.

__global__ void msm_flag_clear_kernel(uint32_t *flag) { *flag = 0; }

struct MsmGraphJob {
  virtual ~MsmGraphJob() {
    if (res) (void)hipHostFree(res);
  }
  virtual int setup(int n, const uint64_t *sc, int enl, int mont, const uint64_t *pts, bool internal,
                    const zkg_msm_schedule *user, uint32_t *ovf, hipStream_t st) = 0;
  virtual void enqueue(hipStream_t st) = 0;
  virtual void combine(MsmResult *out) const = 0;
  int curve = 0;
  double work = 0;           // relative cost, for dealing jobs to streams
  uint32_t *res = nullptr;   // pinned result words
};

template <class C>
struct MsmGraphJobT : MsmGraphJob {
  MsmPipe<C> p;
  int setup(int n, const uint64_t *sc, int enl, int mont, const uint64_t *pts, bool internal,
            const zkg_msm_schedule *user, uint32_t *ovf, hipStream_t st) override {
    curve = C::id;
    zkg_scalar_stats stats;
    const bool use_stats = user && user->policy != ZKG_SCHED_PLAIN;
    if (use_stats) ZK_TRY(msm_scalar_stats(C::id, n, sc, enl, mont, st, &stats));
    MsmSchedule sch;
    ZK_TRY(msm_resolve(C::id, n, enl, user, use_stats ? &stats : nullptr, &sch, true));
    ZK_TRY(p.setup(n, sc, enl, mont, pts, internal, sch, user));
    p.ovf = use_stats ? ovf : nullptr;  // declared-width windows cover every scalar
    ZK_HIP(hipHostMalloc((void **)&res, (size_t)p.nres * MsmPipe<C>::PW * 4, hipHostMallocDefault));
    work = (double)p.total * (C::id == 1 ? 2.0 : 1.0);
    return ZK_OK;
  }
  void enqueue(hipStream_t st) override {
    p.launch_sort(st);
    p.launch_accum(st);
    p.launch_fix(st);
    p.launch_heavy(st);
    p.pyramid(st, true);
    p.launch_out(st);
    (void)hipMemcpyAsync(res, p.outx.p, (size_t)p.nres * MsmPipe<C>::PW * 4, hipMemcpyDeviceToHost, st);
  }
  void combine(MsmResult *out) const override { p.combine(res, out); }
};

// a job of no points: the point at infinity, no device work
struct MsmGraphJobEmpty : MsmGraphJob {
  int setup(int, const uint64_t *, int, int, const uint64_t *, bool, const zkg_msm_schedule *, uint32_t *,
            hipStream_t) override {
    return ZK_OK;
  }
  void enqueue(hipStream_t) override {}
  void combine(MsmResult *out) const override { msm_infinity(curve, out); }
};

struct MsmGraph {
  std::vector<MsmGraphJob *> jobs;
  hipStream_t origin = nullptr;
  std::vector<hipStream_t> lanes;
  std::vector<hipEvent_t> events;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  DevBuf flag;
  uint32_t *flag_host = nullptr;
  ~MsmGraph() {
    if (origin) (void)hipStreamSynchronize(origin);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
    for (hipStream_t s : lanes) (void)hipStreamDestroy(s);
    if (origin) (void)hipStreamDestroy(origin);
    if (flag_host) (void)hipHostFree(flag_host);
    for (MsmGraphJob *j : jobs) delete j;
  }
};

int msm_graph_create(int njobs, const int *curves, const int *ns, const uint64_t *const *d_scalars, int enl, int mont,
                     const uint64_t *const *d_points, bool points_internal, const zkg_msm_schedule *sched,
                     int nstreams, MsmGraph **out) {
  *out = nullptr;
  if (njobs <= 0) {
    set_error("msm graph: no jobs");
    return ZK_EINVAL;
  }
  std::unique_ptr<MsmGraph> g(new MsmGraph);
  ZK_HIP(hipStreamCreateWithFlags(&g->origin, hipStreamNonBlocking));
  ZK_TRY(g->flag.alloc(4));
  ZK_HIP(hipHostMalloc((void **)&g->flag_host, 4, hipHostMallocDefault));
  ZK_TRY(msm_lds_attributes_once());
  zkg_msm_schedule s;
  memset(&s, 0, sizeof s);
  s.policy = sched ? sched->policy : ZKG_SCHED_PLAIN;
  s.fix_limit = sched ? sched->fix_limit : 0;
  for (int k = 0; k < njobs; k++) {
    if (ns[k] < 0 || (curves[k] != 0 && curves[k] != 1)) {
      set_error("msm graph: every job needs a curve id 0/1 and a point count >= 0");
      return ZK_EINVAL;
    }
    MsmGraphJob *j = ns[k] == 0       ? (MsmGraphJob *)new MsmGraphJobEmpty
                     : curves[k] == 0 ? (MsmGraphJob *)new MsmGraphJobT<Bn254G1>
                                      : (MsmGraphJob *)new MsmGraphJobT<Bls381G1>;
    j->curve = curves[k];
    g->jobs.push_back(j);
    ZK_TRY(j->setup(ns[k], d_scalars[k], enl, mont, d_points[k], points_internal, &s, g->flag.as<uint32_t>(),
                    g->origin));
  }
  nstreams = std::max(1, std::min(nstreams, 16));
  g->lanes.resize(nstreams, nullptr);
  for (auto &l : g->lanes) ZK_HIP(hipStreamCreateWithFlags(&l, hipStreamNonBlocking));
  g->events.resize(nstreams + 1, nullptr);
  for (auto &e : g->events) ZK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  // largest first, each to the stream with the least work so far
  std::vector<int> order(njobs);
  for (int k = 0; k < njobs; k++) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return g->jobs[a]->work > g->jobs[b]->work; });
  std::vector<double> load(nstreams, 0.0);
  ZK_HIP(hipStreamSynchronize(g->origin));  // the statistics passes above
  ZK_HIP(hipStreamBeginCapture(g->origin, hipStreamCaptureModeThreadLocal));
  hipLaunchKernelGGL(msm_flag_clear_kernel, dim3(1), dim3(1), 0, g->origin, g->flag.as<uint32_t>());
  (void)hipEventRecord(g->events[nstreams], g->origin);
  for (hipStream_t l : g->lanes) (void)hipStreamWaitEvent(l, g->events[nstreams], 0);
  for (int k : order) {
    const int li = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    g->jobs[k]->enqueue(g->lanes[li]);
    load[li] += g->jobs[k]->work;
  }
  for (int li = 0; li < nstreams; li++) {
    (void)hipEventRecord(g->events[li], g->lanes[li]);
    (void)hipStreamWaitEvent(g->origin, g->events[li], 0);
  }
  (void)hipMemcpyAsync(g->flag_host, g->flag.p, 4, hipMemcpyDeviceToHost, g->origin);
  const hipError_t le = hipGetLastError();
  const hipError_t ce = hipStreamEndCapture(g->origin, &g->graph);
  if (le != hipSuccess || ce != hipSuccess) {
    set_error(std::string("msm graph: capture failed: ") + hipGetErrorString(le != hipSuccess ? le : ce));
    return ZK_EHIP;
  }
  ZK_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  *out = g.release();
  return ZK_OK;
}

int msm_graph_run(MsmGraph *g, uint64_t *const *outs, int out_flags, double *wall_ms) {
  const auto t0 = std::chrono::steady_clock::now();
  ZK_HIP(hipGraphLaunch(g->exec, g->origin));
  ZK_HIP(hipStreamSynchronize(g->origin));
  if (*g->flag_host) {
    set_error("msm graph: a scalar is wider than the windows fitted at create; create the plan again");
    return ZK_EINVAL;
  }
  // host Horner of every job, spread over a few threads
  const int nj = (int)g->jobs.size();
  std::vector<int> status(nj, ZK_OK);
  std::atomic<int> next{0};
  auto worker = [&] {
    for (int k; (k = next.fetch_add(1)) < nj;) {
      MsmResult r;
      g->jobs[k]->combine(&r);
      status[k] = msm_write_result(g->jobs[k]->curve, r, outs[k], out_flags);
    }
  };
  std::vector<std::thread> th;
  const int nt = std::min(nj, 8);
  for (int t = 1; t < nt; t++) th.emplace_back(worker);
  worker();
  for (auto &t : th) t.join();
  for (int k = 0; k < nj; k++) ZK_TRY(status[k]);
  if (wall_ms) *wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return ZK_OK;
}

void msm_graph_destroy(MsmGraph *g) { delete g; }

}  // namespace zk
