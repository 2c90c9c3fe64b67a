// NTT / iNTT over the scalar fields Fr of BN254 and BLS12-381 on gfx950.
//
// Contract (lib/cbits/curves/poly/mont/bn128_poly_mont.c:418-638): natural-order input and
// output, y_k = sum_i x_i g^(ik) (forward) and x_i = N^-1 sum_k y_k g^(-ik) (inverse, given
// the *forward* generator g), Montgomery-form Fr elements in the flat 4x64-bit layout.
// Shifted variants pre-multiply x_i by eta^i (forward, :477-489) or post-multiply the inverse
// output by eta^i (:626-638).  For a generator of order exactly 2^m every correct DFT
// algorithm gives the same canonical field elements, so the GPU factorisation below matches
// the reference's recursive radix-2 network bit for bit.  Other generators go through
// ntt_network_* (the reference's own butterfly network, stage by stage).
//
// Fast path: the transform is factored N = N_1 N_2 ... N_P (each N_p <= 2^11).  Pass p views
// the array as [R_p][N_p][C_p] (R_p = N_1..N_{p-1}, C_p = N_{p+1}..N_P).  The passes run in
// the transposed order (PassArgs): the first to run (p = P, C_P = 1) gathers its rows from
// the digit-reversed input offsets into the output vector; each later pass loads a tile of T
// columns, multiplies by the four-step twiddle omega^(R_p c n), runs log2(N_p) radix-2 DIT
// stages in LDS (bit-reversed placement, twiddles omega^(N/N_p * j)) and writes the tile back
// in place, so the output ends in natural order with no scratch vector.  All passes read and
// write HBM coalesced in T*32-byte runs.
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "field.hpp"
#include "fz.hpp"
#include "hostfield.hpp"
#include "ntt.hpp"
#include "peak.hpp"
#include "runtime.hpp"

namespace zk {

static constexpr int kNttThreads = 256;

__device__ __forceinline__ uint32_t bitrev(uint32_t x, int bits) {
  return bits ? (__builtin_bitreverse32(x) >> (32 - bits)) : 0u;
}

// One pass of the multi-pass transform.  With passes of s_0, ..., s_{P-1} stages (sum m), pass
// p views the vector as [R_p][N_p][C_p] (R_p = 2^(s_0 + ... + s_{p-1}), N_p = 2^(s_p)).  The
// classic decomposition runs p = 0, 1, ..., P-1, each an N_p-point DFT along the middle axis
// followed by the four-step twiddle omega^(R_p c k), and writes its output in digit-reversed
// order.  The DFT matrix is symmetric, so the transposed product runs the same passes in the
// opposite order, each twiddle before its DFT, and the digit reversal moves to the input:
//   gather pass (p = P-1, first to run)  reads x at rho + (n << logr), writes row rev(rho)
//                                        (rows of 2^s contiguous elements);
//   passes p = P-2 ... 0                 DFT, write back in place.
// A twiddle is diagonal in the element's offset, so pass p's twiddle omega^(R_p c n) is applied
// by the pass that runs before it, at its store (the offset decomposed in pass p's view): the
// product leaves a value below 2p, ready for the canonical store.  Every pass after the first
// works in place on the output vector, so an out-of-place transform needs no scratch vector.
// The last pass to run (p = 0, or the gather pass when P = 1) writes natural order and applies
// N^-1 and the coset post-scale.
struct PassArgs {
  const uint32_t *in;
  uint32_t *out;
  const uint32_t *tw;       // omega^j table
  const uint32_t *eta_pre;  // forward shifted: multiply input i by eta^i (gather pass only)
  const uint32_t *eta_post; // inverse shifted: multiply output i by eta^i (final pass only)
  uint32_t ninv[8];         // final pass factor, internal form: N^-1 R' (inverse) or R' (forward)
  int logn, s, logc, logr, logt, gather, fin, inverse, tw_shift;
  int nx_s, nx_logc, nx_logr;  // the next pass's view (not fin): its twiddle is applied at the store
  int has_ninv;             // final pass: the factor is not one (inverse, or a coset post-scale)
  int nsplit;               // gather pass: number of digits of rho (passes p < P-1)
  int split[6];             // their log sizes, s_0 ... s_{P-2}
  // gather pass of an inverse transform (or null): omega^(e) N^-1 for the next pass's twiddle
  // exponents e (multiples of 2^nx_logr below N/2), so that N^-1 rides on that product and the
  // final pass only reduces
  const uint32_t *tw_ninv;
};

// ---- reduced-radix butterflies (fz.hpp).  Data stays in the reference's Montgomery form
// x R in HBM (canonical 8-word strings); twiddles, coset powers and N^-1 are kept in the
// internal form w R' (R' = 2^261), so zmul(x R, w R') = x w R stays external.  Bounds: loads
// are canonical (< p); a twiddle product of a normalised |v| < 32 p is in (-p/2, 3p/2) for
// both fields (p / R' <= 2^-6); each radix-2 stage adds at most 2p, so after s <= 11 stages
// every value is below 22 p, inside zmul's input range; every stored value goes through a
// final product and zcanon.
template <class F> struct ZOf;
template <> struct ZOf<Bn254Fr> { using T = ZBn254Fr; };
template <> struct ZOf<Bls381Fr> { using T = ZBls381Fr; };

static constexpr int kZW = 9;  // LDS words per element (9 limbs of 29 bits)
// LDS word offset of tile element e: 9 words each, one pad word after every 32 elements.  The
// bit-reversed placement of the load and the radix-4 groups of the first stages hit 4 to 32 of
// the 32 banks' words at once with a plain 9-word stride; the pad word spreads them (at most
// two-way in every phase for tiles of 2^7 to 2^10 elements).
__device__ __forceinline__ int lds_slot(int e) { return e * kZW + (e >> 5); }
// LDS bytes of a tile of 2^log elements
static inline size_t lds_tile_bytes(int log) { return (((size_t)1 << log) * kZW + (((size_t)1 << log) >> 5)) * 4; }
// largest LDS tile: 2^12 elements of 36 bytes (144 KiB)
static constexpr int kMaxTileLog = 12;
static constexpr int kBatchTileLog = 9;  // largest LDS tile (log2 elements) of a batched job

// omega^e as an internal-form table entry; returns true when the table value must be negated
// an element of 8 words in HBM (32-byte aligned): two 16-byte loads / stores around the
// (un)packing, instead of eight word accesses
template <class Z>
__device__ __forceinline__ void zunpack_g(Fz<Z> &r, const uint32_t *__restrict__ p) {
  const uint4 *q = reinterpret_cast<const uint4 *>(p);
  const uint4 v0 = q[0], v1 = q[1];
  const uint32_t w[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
  zunpack(r, w);
}
template <class Z>
__device__ __forceinline__ void zpack_g(uint32_t *__restrict__ p, const Fz<Z> &a) {
  uint32_t w[8];
  zpack(w, a);
  uint4 *q = reinterpret_cast<uint4 *>(p);
  q[0] = make_uint4(w[0], w[1], w[2], w[3]);
  q[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

template <class Z>
__device__ __forceinline__ bool twiddle_z(Fz<Z> &w, const uint32_t *__restrict__ tw, uint64_t e, int logn,
                                          int tw_shift, int inverse) {
  const uint64_t N = 1ull << logn;
  if (inverse) e = (N - e) & (N - 1);
  const uint64_t half = N >> 1;
  const bool neg = e >= half;
  if (neg) e -= half;
  zunpack_g(w, tw + ((e << tw_shift) * 8));
  return neg;
}

// Butterfly twiddle omega^(+-e) for 0 <= e < N/2, stored value w and true value w (forward) or
// -w (inverse, INV: the outputs then swap).  The table holds omega^j for j <= N/2 (entry N/2 is
// -1), so omega^-e = -omega^(N/2 - e) for every e in [0, N/2): the sign depends only on the
// direction, and no lane selects between swapped and unswapped outputs.
template <bool INV, class Z>
__device__ __forceinline__ void twiddle_bf(Fz<Z> &w, const uint32_t *__restrict__ tw, uint64_t e, int logn,
                                           int tw_shift) {
  if (INV) e = (1ull << (logn - 1)) - e;
  zunpack_g(w, tw + ((e << tw_shift) * 8));
}

// eta^n (internal form, normalised, < 2p) from three internal-form tables
template <class Z>
__device__ __forceinline__ void eta_pow_z(Fz<Z> &r, const uint32_t *__restrict__ et, uint64_t n) {
  Fz<Z> a, b;
  zunpack_g(a, et + (n & 2047) * 8);
  zunpack_g(b, et + (2048 + ((n >> 11) & 2047)) * 8);
  zmul(r, a, b);
  if (n >> 22) {
    zunpack_g(a, et + (4096 + ((n >> 22) & 2047)) * 8);
    zmul(r, r, a);
  }
}

// canonical (p - t) mod p for canonical t
template <class Z>
__device__ __forceinline__ void zneg_canon(Fz<Z> &r, const Fz<Z> &t) {
  if (zis_raw_zero(t)) {
    r = t;
    return;
  }
#pragma unroll
  for (int i = 0; i < Z::N; i++) r.l[i] = Z::P[i] - t.l[i];
  znorm(r);
}

template <class Z>
__device__ __forceinline__ void zlds_load(Fz<Z> &r, const int32_t *p) {
  static_assert(Z::N <= kZW, "element exceeds its LDS slot");
#pragma unroll
  for (int i = 0; i < Z::N; i++) r.l[i] = p[i];
}
template <class Z>
__device__ __forceinline__ void zlds_store(int32_t *p, const Fz<Z> &a) {
#pragma unroll
  for (int i = 0; i < Z::N; i++) p[i] = a.l[i];
}

// DIT butterfly (u, v) -> (u + w v, u - w v) for the stored twiddle w, whose true value is -w
// when neg (the outputs then swap); w == nullptr is the twiddle 1.  NORM = false leaves the
// outputs as limb-wise sums / differences: the first stage of a radix-4 group takes normalised
// inputs, so its outputs are sums or differences of two normalised values, which zmul accepts
// against a normalised twiddle; the group's second stage (limbs below 4 * 2^29 < 2^31)
// normalises before the store.
template <bool NORM, bool NEG, class Z>
__device__ __forceinline__ void zbutterfly(Fz<Z> &u, Fz<Z> &v, const Fz<Z> *w) {
  Fz<Z> t, s0, s1;
  if (w) zmul(t, v, *w);
  else t = v;
  zadd(s0, u, t);
  zsub(s1, u, t);
  if (NORM) {
    znorm(s0);
    znorm(s1);
  }
  u = NEG ? s1 : s0;
  v = NEG ? s0 : s1;
}

// one workgroup's share of a pass: `blk` indexes the pass's tiles (see PassArgs); INV = a.inverse
template <class F, bool INV>
__device__ __forceinline__ void ntt_pass_body(const PassArgs &a, uint64_t blk, int32_t *lds) {
  using Z = typename ZOf<F>::T;
  using E = Fz<Z>;
  const int S = 1 << a.s, T = 1 << a.logt;
  const int elems = S << a.logt;
  const int tid = threadIdx.x;

  // tile: T consecutive columns c0.. of row r (other passes), or T consecutive rho0.. (gather)
  uint64_t r = 0, c0 = 0, rho0 = 0;
  if (!a.gather) {
    const int lg = a.logc - a.logt;
    r = blk >> lg;
    c0 = (blk & ((1ull << lg) - 1)) << a.logt;
  } else {
    rho0 = blk << a.logt;
  }

  // ---- load (coalesced over the tile's T consecutive positions), coset pre-scale (gather
  // pass), bit-reversed LDS placement
  for (int idx = tid; idx < elems; idx += kNttThreads) {
    const uint32_t n = idx >> a.logt, c = idx & (T - 1);
    const uint64_t gi = a.gather ? (rho0 + c) + ((uint64_t)n << a.logr)
                                 : ((((r << a.s) + n) << a.logc) | (c0 + c));
    E x;
    zunpack_g(x, a.in + gi * 8);
    if (a.eta_pre) {
      E e;
      eta_pow_z(e, a.eta_pre, gi);
      zmul(x, x, e);
    }
    zlds_store(lds + lds_slot((bitrev(n, a.s) << a.logt) | c), x);
  }
  __syncthreads();

  // ---- DIT stages in LDS (natural-order result).  Tiles with at least four elements per
  // thread take two stages at a time: a thread carries the four elements n0 + {0, h, 2h, 3h}
  // through stages lh and lh + 1 in registers (a radix-4 group); smaller tiles keep every
  // thread busy with one radix-2 butterfly per stage.
  const bool pairs = (elems >> 2) >= kNttThreads;
  int lh = 0;
  while (lh < a.s) {
    const int h = 1 << lh;
    if (pairs && lh + 1 < a.s) {
      for (int q = tid; q < (elems >> 2); q += kNttThreads) {
        const int c = q & (T - 1);
        const int g = q >> a.logt;
        const int j = g & (h - 1);
        const int blk = g >> lh;
        const int n0 = (blk << (lh + 2)) | j;
        const int e0 = (n0 << a.logt) | c, es = h << a.logt;
        const int p0 = lds_slot(e0), p1 = lds_slot(e0 + es), p2 = lds_slot(e0 + 2 * es), p3 = lds_slot(e0 + 3 * es);
        E x0, x1, x2, x3;
        zlds_load(x0, lds + p0);
        zlds_load(x1, lds + p1);
        zlds_load(x2, lds + p2);
        zlds_load(x3, lds + p3);
        if (lh > 0) {
          E w1;
          twiddle_bf<INV>(w1, a.tw, (uint64_t)j << (a.logn - lh - 1), a.logn, a.tw_shift);
          zbutterfly<false, INV>(x0, x1, &w1);
          zbutterfly<false, INV>(x2, x3, &w1);
        } else {
          zbutterfly<false, false>(x0, x1, (const E *)nullptr);
          zbutterfly<false, false>(x2, x3, (const E *)nullptr);
        }
        E w2;
        twiddle_bf<INV>(w2, a.tw, (uint64_t)j << (a.logn - lh - 2), a.logn, a.tw_shift);
        if (lh > 0 || j > 0) zbutterfly<true, INV>(x0, x2, &w2);
        else zbutterfly<true, false>(x0, x2, (const E *)nullptr);
        twiddle_bf<INV>(w2, a.tw, (uint64_t)(j + h) << (a.logn - lh - 2), a.logn, a.tw_shift);
        zbutterfly<true, INV>(x1, x3, &w2);
        zlds_store(lds + p0, x0);
        zlds_store(lds + p1, x1);
        zlds_store(lds + p2, x2);
        zlds_store(lds + p3, x3);
      }
      lh += 2;
    } else {
      for (int q = tid; q < (elems >> 1); q += kNttThreads) {
        const int c = q & (T - 1);
        const int bf = q >> a.logt;
        const int j = bf & (h - 1);
        const int blk = bf >> lh;
        const int eu = (((blk << (lh + 1)) | j) << a.logt) | c;
        const int pu = lds_slot(eu), pv = lds_slot(eu + (h << a.logt));
        E u, v;
        zlds_load(u, lds + pu);
        zlds_load(v, lds + pv);
        if (lh > 0) {
          E w;
          twiddle_bf<INV>(w, a.tw, (uint64_t)j << (a.logn - lh - 1), a.logn, a.tw_shift);
          zbutterfly<true, INV>(u, v, &w);
        } else {
          zbutterfly<true, false>(u, v, (const E *)nullptr);
        }
        zlds_store(lds + pu, u);
        zlds_store(lds + pv, v);
      }
      lh += 1;
    }
    __syncthreads();
  }

  // ---- store: row rev(rho) (gather pass) or in place, times the next pass's twiddle
  // omega^(c n R) of the offset in that pass's [R][N][C] view; the final pass multiplies by
  // N^-1 and the coset post-scale (or reduces: the factor is one); canonical words
  E ninv;
#pragma unroll
  for (int j = 0; j < Z::N; j++) ninv.l[j] = 0;
  if (a.fin && a.has_ninv) zunpack(ninv, a.ninv);  // internal-form factor N^-1 R' (or R' with a post-scale)
  for (int idx = tid; idx < elems; idx += kNttThreads) {
    const uint32_t k = idx >> a.logt, c = idx & (T - 1);
    E x;
    zlds_load(x, lds + lds_slot(idx));
    uint64_t go;
    if (a.gather) {
      uint64_t rho = rho0 + c, rr = 0;
      for (int i = 0; i < a.nsplit; i++) {
        rr = (rr << a.split[i]) | (rho & ((1ull << a.split[i]) - 1));
        rho >>= a.split[i];
      }
      go = (rr << a.s) + k;
    } else {
      go = (((r << a.s) + k) << a.logc) | (c0 + c);
    }
    bool neg = false;
    if (!a.fin) {
      const uint64_t col = go & ((1ull << a.nx_logc) - 1);
      const uint64_t nn = (go >> a.nx_logc) & ((1ull << a.nx_s) - 1);
      E w;
      if (a.tw_ninv) {
        const uint64_t N = 1ull << a.logn, half = N >> 1;
        uint64_t e = (N - ((col * nn) << a.nx_logr)) & (N - 1);  // inverse: omega^-e
        neg = e >= half;
        if (neg) e -= half;
        zunpack_g(w, a.tw_ninv + (e >> a.nx_logr) * 8);
      } else {
        neg = twiddle_z(w, a.tw, (col * nn) << a.nx_logr, a.logn, a.tw_shift, a.inverse);
      }
      zmul(x, x, w);
    } else if (a.has_ninv) {
      E f = ninv;
      if (a.eta_post) {
        E e;
        eta_pow_z(e, a.eta_post, go);
        zmul(f, f, e);
      }
      zmul(x, x, f);
    } else {
      zreduce_small(x, x);  // |x| < 64 p after at most 11 stages
    }
    zcanon(x, x);
    if (neg) zneg_canon(x, x);
    zpack_g(a.out + go * 8, x);
  }
}

template <class F>
__global__ __launch_bounds__(kNttThreads) void ntt_pass_kernel(PassArgs a) {
  extern __shared__ uint4 lds_raw[];
  if (a.inverse) ntt_pass_body<F, true>(a, blockIdx.x, reinterpret_cast<int32_t *>(lds_raw));
  else ntt_pass_body<F, false>(a, blockIdx.x, reinterpret_cast<int32_t *>(lds_raw));
}

// batched pass: item i owns blocks [prefix[i], prefix[i+1]) of the launch
template <class F>
__global__ __launch_bounds__(kNttThreads) void ntt_batch_pass_kernel(const PassArgs *__restrict__ items,
                                                                     const uint32_t *__restrict__ prefix,
                                                                     int nitems) {
  extern __shared__ uint4 lds_raw[];
  const uint32_t b = blockIdx.x;
  int lo = 0, hi = nitems;  // last i with prefix[i] <= b
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (__builtin_amdgcn_readfirstlane(prefix[mid]) <= b) lo = mid;
    else hi = mid;
  }
  // the item is the same for the whole workgroup: copy it through readfirstlane so that its
  // fields live in scalar registers, as kernel arguments do (vector copies would cost ~40
  // VGPRs and the occupancy with them)
  static_assert(sizeof(PassArgs) % 4 == 0, "PassArgs is copied as words");
  PassArgs a;
  const uint32_t *src = reinterpret_cast<const uint32_t *>(items + lo);
  uint32_t *dst = reinterpret_cast<uint32_t *>(&a);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(PassArgs) / 4); i++) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
  const uint32_t first = __builtin_amdgcn_readfirstlane(prefix[lo]);
  if (a.inverse) ntt_pass_body<F, true>(a, b - first, reinterpret_cast<int32_t *>(lds_raw));
  else ntt_pass_body<F, false>(a, b - first, reinterpret_cast<int32_t *>(lds_raw));
}

// ---------------------------------------------------------------------------------------
// table builders

// tab[j] = base^(j * step) for j < cnt
template <class F>
__global__ void pow_table_kernel(uint32_t *tab, Fe<F> base, uint64_t step, int cnt) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cnt) return;
  Fe<F> r;
  fpow_u64(r, base, step * (uint64_t)j);
  fstore(tab + (size_t)j * F::N, r);
}

// tw[e] = lo[e & 2047] * hi[e >> 11]
template <class F>
__global__ void twiddle_table_kernel(uint32_t *tw, const uint32_t *lo, const uint32_t *hi, uint64_t cnt) {
  uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= cnt) return;
  Fe<F> a, b;
  fload(a, lo + (e & 2047) * F::N);
  fload(b, hi + (e >> 11) * F::N);
  fmul(a, a, b);
  fstore(tw + e * F::N, a);
}

// ---------------------------------------------------------------------------------------
// reference butterfly network, for generators whose order is not 2^m
// (bn128_poly_mont.c:418-447 forward; :565-596 inverse), one launch per radix-2 stage.
//
// Forward, stage t = 1..m with s = N / 2^t:  Y_t[r][j]        = Y_{t-1}[r][j] + g^(js) Y_{t-1}[r+s][j]
//                                            Y_t[r][j+2^(t-1)] = Y_{t-1}[r][j] - g^(js) Y_{t-1}[r+s][j]
// with Y_t stored as [s][2^t] and Y_0 = x.  g^(js) comes from gp[j s] = g^(js) (powers table).
template <class F>
__global__ void ntt_net_fwd_stage(const uint32_t *in, uint32_t *out, const uint32_t *gp, int m, int t) {
  const uint64_t N = 1ull << m;
  uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;  // butterfly id < N/2
  if (q >= (N >> 1)) return;
  const int ht = t - 1;
  const uint64_t hw = 1ull << ht;          // 2^(t-1)
  const uint64_t s = N >> t;
  const uint64_t r = q >> ht, j = q & (hw - 1);
  Fe<F> u, v, w;
  fload(u, in + ((r << ht) + j) * F::N);          // Y_{t-1}[r][j], layout [2s][2^(t-1)]
  fload(v, in + (((r + s) << ht) + j) * F::N);    // Y_{t-1}[r+s][j]
  fload(w, gp + (j * s) * F::N);
  fmul(v, v, w);
  Fe<F> a;
  fadd(a, u, v);
  fsub(v, u, v);
  fstore(out + ((r << t) + j) * F::N, a);
  fstore(out + ((r << t) + j + hw) * F::N, v);
}

// Inverse, reference DIF with halving.  Level d = 0..m-1 works on blocks of size M = N/2^d
// with stride 2^d in the *output* ordering; written iteratively over a [2^d][M] layout:
//   Z_{d+1}[2b  ][j] = (Z_d[b][j] + Z_d[b][j+M/2]) / 2
//   Z_{d+1}[2b+1][j] = (Z_d[b][j] - Z_d[b][j+M/2]) * gp[j 2^d]     (gp[i] = 1/2 g^-i)
// and the final level (M = 2) uses halving for both outputs.  The output index of block b
// after all levels is the bit reversal of b, exactly as the recursion's tgt_stride doubling.
template <class F>
__global__ void ntt_net_inv_stage(const uint32_t *in, uint32_t *out, const uint32_t *gp, int m, int d) {
  const uint64_t N = 1ull << m;
  uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (N >> 1)) return;
  const int lm = m - d;               // log block size
  const uint64_t M = 1ull << lm, half = M >> 1;
  const uint64_t b = q >> (lm - 1), j = q & (half - 1);
  Fe<F> x, y;
  fload(x, in + ((b << lm) + j) * F::N);
  fload(y, in + ((b << lm) + j + half) * F::N);
  Fe<F> s, df;
  fadd(s, x, y);
  fsub(df, x, y);
  fhalf(s, s);
  if (lm == 1) {
    fhalf(df, df);
  } else {
    Fe<F> w;
    fload(w, gp + ((j << d)) * F::N);
    fmul(df, df, w);
  }
  // next layout [2^(d+1)][M/2]: block 2b and 2b+1
  fstore(out + (((2 * b) << (lm - 1)) + j) * F::N, s);
  fstore(out + (((2 * b + 1) << (lm - 1)) + j) * F::N, df);
}

// final permutation of the inverse network: out[bitrev(b)] = Z_m[b]
template <class F>
__global__ void bitrev_copy_kernel(const uint32_t *in, uint32_t *out, int m) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >> m) return;
  uint64_t j = __builtin_bitreverse64(i) >> (64 - m);
  Fe<F> x;
  fload(x, in + i * F::N);
  fstore(out + j * F::N, x);
}

// sequential powers table for the network path (host-seeded, built on device in blocks)
// gp[j] = first * g^j
template <class F>
__global__ void powers_kernel(uint32_t *gp, Fe<F> first, Fe<F> g, uint64_t cnt) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cnt) return;
  Fe<F> r;
  fpow_u64(r, g, j);
  fmul(r, r, first);
  fstore(gp + j * F::N, r);
}

// ---------------------------------------------------------------------------------------
// host side

template <class F>
static Fe<F> to_dev(const zkh::Fe &a) {
  Fe<F> r;
  for (int j = 0; j < F::N; j++) r.w[j] = (uint32_t)(a.w[j / 2] >> (32 * (j & 1)));
  return r;
}

struct TwEntry {
  int curve;
  int m;
  zkh::Fe gen;
  DevBuf buf;  // omega^j, j < 2^(m-1)
};
static std::mutex g_tw_lock;
static std::vector<TwEntry *> g_tw;

static const zkh::FieldDesc &fr_desc(int curve) { return curve == 0 ? zkh::BN254_FR : zkh::BLS381_FR; }

// returns true when g has multiplicative order exactly 2^m
bool ntt_generator_ok(int curve, int m, const uint64_t *gen) {
  const zkh::FieldDesc &f = fr_desc(curve);
  zkh::Fe g = zkh::load(f, gen);
  if (m == 0) return zkh::is_eq(f, g, zkh::one(f));
  for (int i = 0; i < m - 1; i++) g = zkh::sqr(f, g);
  zkh::Fe mone = zkh::neg(f, zkh::one(f));
  return zkh::is_eq(f, g, mone);
}

// table entries x R (canonical, 8 words) -> x R' (canonical, 8 words), in place
template <class F>
__global__ void tab_to_internal_kernel(uint32_t *tab, uint64_t cnt) {
  using Z = typename ZOf<F>::T;
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= cnt) return;
  Fz<Z> a, k;
  zunpack(a, tab + e * 8);
  zset_const(k, Z::K_IN);
  zmul(a, a, k);
  zcanon(a, a);
  zpack(tab + e * 8, a);
}

template <class F>
static int tab_to_internal(uint32_t *tab, uint64_t cnt, hipStream_t st) {
  hipLaunchKernelGGL(tab_to_internal_kernel<F>, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st, tab, cnt);
  ZK_HIP(hipGetLastError());
  return ZK_OK;
}

// internal form (x R', canonical words) of a host Montgomery element x R
template <class F>
static void internal_words(uint32_t out[8], const zkh::Fe &x) {
  using Z = typename ZOf<F>::T;
  const Fe<F> d = to_dev<F>(x);
  Fz<Z> a, k;
  zunpack(a, d.w);
  zset_const(k, Z::K_IN);
  zmul(a, a, k);
  zcanon(a, a);
  zpack(out, a);
}

// last-pass factor: N^-1 for the inverse, 1 for the forward transform (internal form)
template <class F>
static void last_factor(uint32_t out[8], int curve, int m, bool inverse) {
  const zkh::FieldDesc &f = fr_desc(curve);
  const zkh::Fe v = inverse ? zkh::inv(f, zkh::from_u64(f, (uint64_t)1 << m)) : zkh::one(f);
  internal_words<F>(out, v);
}

// tab[j] = g^j R' (internal form, canonical) for j < cnt, through lo (2048 entries) and hi
// (ceil(cnt / 2048) entries) powers
template <class F>
static int launch_twiddles(uint32_t *tab, uint32_t *lo, uint32_t *hi, uint64_t cnt, const zkh::Fe &g,
                           hipStream_t st) {
  const uint64_t hicnt = (cnt + 2047) / 2048;
  Fe<F> gd = to_dev<F>(g);
  hipLaunchKernelGGL(pow_table_kernel<F>, dim3(2048 / 256), dim3(256), 0, st, lo, gd, 1ull, 2048);
  hipLaunchKernelGGL(pow_table_kernel<F>, dim3((hicnt + 255) / 256), dim3(256), 0, st, hi, gd, 2048ull,
                     (int)hicnt);
  hipLaunchKernelGGL(twiddle_table_kernel<F>, dim3((cnt + 255) / 256), dim3(256), 0, st, tab,
                     (const uint32_t *)lo, (const uint32_t *)hi, cnt);
  ZK_HIP(hipGetLastError());
  return tab_to_internal<F>(tab, cnt, st);  // the pass kernels read w R'
}

template <class F>
static int build_twiddles(int curve, int m, const zkh::Fe &g, DevBuf &buf, hipStream_t st) {
  const uint64_t cnt = (m >= 1 ? (1ull << (m - 1)) : 1) + 1;  // omega^j, j <= N/2
  ZK_TRY(buf.alloc(cnt * F::N * 4));
  DevBuf lo, hi;
  StreamSyncGuard drain(st);  // lo/hi outlive the queued table kernels on every return
  ZK_TRY(lo.alloc(2048 * F::N * 4));
  const uint64_t hicnt = (cnt + 2047) / 2048;
  ZK_TRY(hi.alloc(hicnt * F::N * 4));
  ZK_TRY(launch_twiddles<F>(buf.as<uint32_t>(), lo.as<uint32_t>(), hi.as<uint32_t>(), cnt, g, st));
  ZK_HIP(hipStreamSynchronize(st));  // lo/hi are released on return
  drain.release();
  (void)curve;
  return ZK_OK;
}

// Finds (or builds) a table of omega^j for the domain generated by `gen` of order 2^m.
// A cached table of a larger domain whose generator squares down to `gen` is reused with a
// stride (all FFT subgroups of one field come from one 2-adic generator).
template <class F>
static int get_twiddles(int curve, int m, const uint64_t *gen, const uint32_t **tab, int *shift, hipStream_t st) {
  const zkh::FieldDesc &f = fr_desc(curve);
  zkh::Fe g = zkh::load(f, gen);
  std::lock_guard<std::mutex> lk(g_tw_lock);
  for (TwEntry *e : g_tw) {
    if (e->curve != curve || e->m < m) continue;
    zkh::Fe h = e->gen;
    for (int i = 0; i < e->m - m; i++) h = zkh::sqr(f, h);
    if (zkh::is_eq(f, h, g)) {
      *tab = e->buf.as<uint32_t>();
      *shift = e->m - m;
      return ZK_OK;
    }
  }
  TwEntry *e = new TwEntry();
  e->curve = curve;
  e->m = m;
  e->gen = g;
  int s = build_twiddles<F>(curve, m, g, e->buf, st);
  if (s != ZK_OK) {
    delete e;
    return s;
  }
  g_tw.push_back(e);
  *tab = e->buf.as<uint32_t>();
  *shift = 0;
  return ZK_OK;
}

// out[i] = omega^(i 2^step) N^-1 (internal form, canonical), i < cnt, from the twiddle table tw
// (entries omega^j R' of a table shifted by tw_shift) and N^-1 R' (fac)
template <class F>
__global__ void ninv_twiddle_kernel(uint32_t *__restrict__ out, const uint32_t *__restrict__ tw, int step_shift,
                                    uint64_t cnt, Fe<F> fac) {
  using Z = typename ZOf<F>::T;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  Fz<Z> w, f;
  zunpack_g(w, tw + ((i << step_shift) * 8));
  zunpack(f, fac.w);
  zmul(w, w, f);
  zcanon(w, w);
  zpack_g(out + i * 8, w);
}

struct NinvEntry {
  int curve, m, step;
  zkh::Fe gen;
  DevBuf buf;
};
static std::vector<NinvEntry *> g_ninv;  // under g_tw_lock
static constexpr uint64_t kNinvMaxEntries = 1ull << 18;

// the N^-1-scaled twiddles of an inverse transform's gather pass (see PassArgs::tw_ninv), cached
// per (curve, m, generator, step); null (no fold) when the table would exceed kNinvMaxEntries
template <class F>
static int get_ninv_twiddles(int curve, int m, const uint64_t *gen, int step, const uint32_t *tw, int tw_shift,
                             const uint32_t *ninv_words, const uint32_t **out, hipStream_t st) {
  *out = nullptr;
  if (m < 1 || step >= m) return ZK_OK;
  const uint64_t cnt = (1ull << (m - 1)) >> step;
  if (cnt > kNinvMaxEntries) return ZK_OK;
  const zkh::FieldDesc &f = fr_desc(curve);
  const zkh::Fe g = zkh::load(f, gen);
  std::lock_guard<std::mutex> lk(g_tw_lock);
  for (NinvEntry *e : g_ninv)
    if (e->curve == curve && e->m == m && e->step == step && zkh::is_eq(f, e->gen, g)) {
      *out = e->buf.as<uint32_t>();
      return ZK_OK;
    }
  NinvEntry *e = new NinvEntry();
  e->curve = curve;
  e->m = m;
  e->step = step;
  e->gen = g;
  if (e->buf.alloc(cnt * F::N * 4) != ZK_OK) {
    delete e;
    return ZK_ENOMEM;
  }
  Fe<F> fac;
  for (int j = 0; j < F::N; j++) fac.w[j] = ninv_words[j];
  hipLaunchKernelGGL(ninv_twiddle_kernel<F>, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st,
                     e->buf.as<uint32_t>(), tw, step + tw_shift, cnt, fac);
  hipError_t err = hipGetLastError();
  if (err == hipSuccess) err = hipStreamSynchronize(st);  // complete before other streams may use it
  if (err != hipSuccess) {
    (void)hipStreamSynchronize(st);
    delete e;
    set_error(std::string("ntt: ") + hipGetErrorString(err));
    return ZK_EHIP;
  }
  g_ninv.push_back(e);
  *out = e->buf.as<uint32_t>();
  return ZK_OK;
}

// coset tables in internal form: tab[k][j] = (eta^(2^(11 k)))^j R', k < 3, j < 2048; the three
// bases are raised on the host, so each entry is at most 11 squarings away from its base
template <class F>
__global__ void eta_tables_kernel(uint32_t *tab, Fe<F> b0, Fe<F> b1, Fe<F> b2) {
  using Z = typename ZOf<F>::T;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * 2048) return;
  const int k = t >> 11, j = t & 2047;
  Fe<F> r;
  fpow_u64(r, k == 0 ? b0 : (k == 1 ? b1 : b2), (uint64_t)j);
  Fz<Z> a, c;
  zunpack(a, r.w);
  zset_const(c, Z::K_IN);
  zmul(a, a, c);
  zcanon(a, a);
  zpack(tab + (size_t)t * 8, a);
}

// the same tables for many cosets in one launch: bases[3 i + k] are coset i's three bases, its
// tables at tab + i * 3 * 2048 * 8 words
template <class F>
__global__ void eta_tables_multi_kernel(uint32_t *tab, const Fe<F> *__restrict__ bases) {
  using Z = typename ZOf<F>::T;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;  // grid: exactly 3 * 2048 per coset
  const uint64_t coset = t / (3 * 2048);
  const int rem = (int)(t - coset * 3 * 2048), k = rem >> 11, j = rem & 2047;
  Fe<F> r;
  fpow_u64(r, bases[coset * 3 + k], (uint64_t)j);
  Fz<Z> a, c;
  zunpack(a, r.w);
  zset_const(c, Z::K_IN);
  zmul(a, a, c);
  zcanon(a, a);
  zpack(tab + t * 8, a);
}

// eta, eta^2048, eta^(2^22): the bases of a coset's three tables (eta_pow_z)
template <class F>
static void eta_bases(int curve, const zkh::Fe &eta, std::vector<Fe<F>> *out) {
  const zkh::FieldDesc &f = fr_desc(curve);
  zkh::Fe b1 = eta, b2;
  for (int i = 0; i < 11; i++) b1 = zkh::sqr(f, b1);
  b2 = b1;
  for (int i = 0; i < 11; i++) b2 = zkh::sqr(f, b2);
  out->push_back(to_dev<F>(eta));
  out->push_back(to_dev<F>(b1));
  out->push_back(to_dev<F>(b2));
}

template <class F>
static int launch_eta_tables(uint32_t *tab, int curve, const zkh::Fe &eta, hipStream_t st) {
  std::vector<Fe<F>> b;
  eta_bases<F>(curve, eta, &b);
  hipLaunchKernelGGL(eta_tables_kernel<F>, dim3(24), dim3(256), 0, st, tab, b[0], b[1], b[2]);
  ZK_HIP(hipGetLastError());
  return ZK_OK;
}

template <class F>
static int build_eta_tables(int curve, const zkh::Fe &eta, DevBuf &buf, hipStream_t st) {
  ZK_TRY(buf.alloc(3 * 2048 * F::N * 4));
  return launch_eta_tables<F>(buf.as<uint32_t>(), curve, eta, st);
}

NttPlan ntt_plan(int m, int smax, int tile_log) {
  NttPlan p;
  p.m = m;
  if (smax < 1) smax = 1;
  if (smax > 11) smax = 11;
  p.npass = (m + smax - 1) / smax;
  if (p.npass < 1) p.npass = 1;
  int rem = m;
  for (int i = 0; i < p.npass; i++) {
    int left = p.npass - i;
    p.s[i] = (rem + left - 1) / left;
    rem -= p.s[i];
  }
  p.tile_log = tile_log;
  return p;
}

// the passes of one transform in execution order (gather pass first, see PassArgs); `buf`
// receives the gather pass's output and holds the vector in place until the final pass writes
// d_out (buf == d_out for an out-of-place transform, a scratch vector when d_in == d_out).
// groups[q] = workgroups of pass q, lds_bytes[q] its dynamic LDS.
template <class F>
static int ntt_passes(const NttPlan &plan, const uint32_t *d_in, uint32_t *buf, uint32_t *d_out, const uint32_t *tw,
                      int tw_shift, int inverse, const uint32_t *eta_pre, const uint32_t *eta_post,
                      const uint32_t *last_fac, PassArgs *out, uint64_t *groups, size_t *lds_bytes,
                      const uint32_t *tw_ninv = nullptr) {
  const int m = plan.m, P = plan.npass;
  if (P > 7) {
    set_error("ntt: at most 7 passes");
    return ZK_EINVAL;
  }
  int logr_of[8];
  logr_of[0] = 0;
  for (int p = 0; p < P; p++) logr_of[p + 1] = logr_of[p] + plan.s[p];
  for (int q = 0; q < P; q++) {
    const int p = P - 1 - q;  // the classic pass this one transposes
    PassArgs a;
    memset(&a, 0, sizeof a);
    a.gather = q == 0;
    a.fin = p == 0;
    a.in = q == 0 ? d_in : buf;
    a.out = a.fin ? d_out : buf;
    a.tw = tw;
    a.tw_shift = tw_shift;
    a.logn = m;
    a.s = plan.s[p];
    a.logr = logr_of[p];
    a.logc = m - a.logr - a.s;
    a.inverse = inverse;
    a.eta_pre = a.gather ? eta_pre : nullptr;
    a.eta_post = a.fin ? eta_post : nullptr;
    // with tw_ninv (inverse, several passes) N^-1 rides on the gather pass's store twiddle and
    // last_fac is one
    const bool folded = tw_ninv && inverse && P > 1;
    if (a.gather && folded) a.tw_ninv = tw_ninv;
    if (a.fin) {
      a.has_ninv = (inverse && !folded) || eta_post;
      for (int j = 0; j < 8; j++) a.ninv[j] = last_fac[j];
    } else {
      a.nx_s = plan.s[p - 1];
      a.nx_logr = logr_of[p - 1];
      a.nx_logc = m - a.nx_logr - a.nx_s;
    }
    int tl = plan.tile_log - a.s;
    if (tl < 0) tl = 0;
    const int avail = a.gather ? a.logr : a.logc;
    if (tl > avail) tl = avail;
    a.logt = tl;
    if (a.gather) {
      a.nsplit = p;
      for (int i = 0; i < p; i++) a.split[i] = plan.s[i];
    }
    out[q] = a;
    groups[q] = a.gather ? (1ull << (a.logr - tl)) : ((1ull << a.logr) << (a.logc - tl));
    lds_bytes[q] = lds_tile_bytes(a.s + tl);
  }
  return ZK_OK;
}

// d_tmp: a scratch vector, used only when d_in == d_out and the plan has several passes
template <class F>
static int ntt_fast(const NttPlan &plan, const uint32_t *d_in, uint32_t *d_out, uint32_t *d_tmp,
                    const uint32_t *tw, int tw_shift, int inverse, const uint32_t *eta_pre,
                    const uint32_t *eta_post, const uint32_t *last_fac, hipStream_t st,
                    const uint32_t *tw_ninv = nullptr) {
  uint32_t *buf = (d_in == d_out && plan.npass > 1) ? d_tmp : d_out;
  if (!buf) {
    set_error("ntt: an in-place multi-pass transform needs a scratch vector");
    return ZK_EINVAL;
  }
  PassArgs args[8];
  uint64_t groups[8];
  size_t lds[8];
  ZK_TRY(ntt_passes<F>(plan, d_in, buf, d_out, tw, tw_shift, inverse, eta_pre, eta_post, last_fac, args, groups,
                       lds, tw_ninv));
  for (int q = 0; q < plan.npass; q++) {
    hipLaunchKernelGGL(ntt_pass_kernel<F>, dim3((unsigned)groups[q]), dim3(kNttThreads), lds[q], st, args[q]);
    ZK_HIP(hipGetLastError());
  }
  return ZK_OK;
}

template <class F>
static int ntt_network(int curve, int m, const uint64_t *gen, const uint32_t *d_in, uint32_t *d_out,
                       uint32_t *d_tmp, int inverse, hipStream_t st) {
  const zkh::FieldDesc &f = fr_desc(curve);
  const uint64_t N = 1ull << m;
  const uint64_t half = N >> 1;
  DevBuf gp;
  StreamSyncGuard drain(st);  // gp outlives the queued stages on every return
  ZK_TRY(gp.alloc((half ? half : 1) * F::N * 4));
  zkh::Fe g = zkh::load(f, gen);
  zkh::Fe first = zkh::one(f);
  if (inverse) {
    g = zkh::inv(f, g);
    zkh::Fe two = zkh::from_u64(f, 2);
    first = zkh::inv(f, two);
  }
  if (half)
    hipLaunchKernelGGL(powers_kernel<F>, dim3((half + 255) / 256), dim3(256), 0, st, gp.as<uint32_t>(),
                       to_dev<F>(first), to_dev<F>(g), half);
  const unsigned blocks = (unsigned)((half + 255) / 256);
  if (!inverse) {
    // ping-pong so that the last stage lands in d_out
    const uint32_t *src = d_in;
    for (int t = 1; t <= m; t++) {
      uint32_t *dst = ((m - t) % 2 == 0) ? d_out : d_tmp;
      hipLaunchKernelGGL(ntt_net_fwd_stage<F>, dim3(blocks), dim3(256), 0, st, src, dst, gp.as<uint32_t>(), m, t);
      src = dst;
    }
  } else {
    const uint32_t *src = d_in;
    uint32_t *bufs[2] = {d_tmp, d_out};
    int k = 0;
    for (int d = 0; d < m; d++) {
      uint32_t *dst = bufs[k];
      hipLaunchKernelGGL(ntt_net_inv_stage<F>, dim3(blocks), dim3(256), 0, st, src, dst, gp.as<uint32_t>(), m, d);
      src = dst;
      k ^= 1;
    }
    // src holds Z_m in block order; permute into d_out (via d_tmp if needed)
    uint32_t *dst = d_out;
    if (src == d_out) {
      ZK_HIP(hipMemcpyAsync(d_tmp, d_out, N * F::N * 4, hipMemcpyDeviceToDevice, st));
      src = d_tmp;
    }
    hipLaunchKernelGGL(bitrev_copy_kernel<F>, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, src, dst, m);
  }
  ZK_HIP(hipGetLastError());
  ZK_HIP(hipStreamSynchronize(st));  // gp is released on return
  drain.release();
  return ZK_OK;
}

// Default schedule (passes of at most 2^8 points, LDS tiles of 2^9 elements: 16 KiB per
// workgroup keeps several workgroups per CU and enough of them for small transforms), and a
// per-(curve, size) table filled by the autotuner.  A set entry overrides the default.
static std::mutex g_tuning_mu;
static NttTuning g_tuning = {8, 9};
static NttTuning g_tuning_for[2][31] = {};

void ntt_set_tuning(const NttTuning &t) {
  std::lock_guard<std::mutex> lk(g_tuning_mu);
  g_tuning = t;
  memset(g_tuning_for, 0, sizeof g_tuning_for);
}
NttTuning ntt_get_tuning() {
  std::lock_guard<std::mutex> lk(g_tuning_mu);
  return g_tuning;
}
void ntt_set_tuning_for(int curve, int m, const NttTuning &t) {
  std::lock_guard<std::mutex> lk(g_tuning_mu);
  if (curve >= 0 && curve < 2 && m >= 0 && m <= 30) g_tuning_for[curve][m] = t;
}
NttTuning ntt_tuning_for(int curve, int m) {
  std::lock_guard<std::mutex> lk(g_tuning_mu);
  if (curve >= 0 && curve < 2 && m >= 0 && m <= 30 && g_tuning_for[curve][m].smax > 0) return g_tuning_for[curve][m];
  return g_tuning;
}

template <class F>
static int ntt_device_t(int curve, int m, const uint64_t *gen, const uint64_t *eta, int flags,
                        const uint32_t *d_in, uint32_t *d_out, hipStream_t st, const NttTuning *tune) {
  const uint64_t N = 1ull << m;
  const bool inverse = flags & NTT_INVERSE;
  if (m == 0 && !eta) {
    if (d_in != d_out) ZK_HIP(hipMemcpyAsync(d_out, d_in, F::N * 4, hipMemcpyDeviceToDevice, st));
    return ZK_OK;
  }
  DevBuf tmp, alias, etab;
  StreamSyncGuard drain(st);  // the buffers outlive the queued kernels on every return
  const zkh::FieldDesc &f = fr_desc(curve);
  const bool ok = ntt_generator_ok(curve, m, gen);
  if (!ok) {
    ZK_TRY(tmp.alloc(N * F::N * 4));
    // reference network (exact for any generator); coset scaling through the fast kernels is
    // not available there, so shifted transforms scale with a separate elementwise pass
    if (eta) {
      set_error("shifted NTT with a generator of order != 2^m is not supported");
      return ZK_ENOTROOT;
    }
    if (d_in == d_out) {
      ZK_TRY(alias.alloc(N * F::N * 4));
      ZK_HIP(hipMemcpyAsync(alias.p, d_in, N * F::N * 4, hipMemcpyDeviceToDevice, st));
      d_in = alias.as<uint32_t>();
    }
    return ntt_network<F>(curve, m, gen, d_in, d_out, tmp.as<uint32_t>(), inverse, st);
  }
  const uint32_t *tw = nullptr;
  int tw_shift = 0;
  if (m >= 1) ZK_TRY(get_twiddles<F>(curve, m, gen, &tw, &tw_shift, st));
  if (eta) ZK_TRY(build_eta_tables<F>(curve, zkh::load(f, eta), etab, st));
  uint32_t lastf[8];
  last_factor<F>(lastf, curve, m, inverse);
  const NttTuning tu = tune ? *tune : ntt_tuning_for(curve, m);
  NttPlan plan = ntt_plan(m, tu.smax, tu.tile_log);
  if (d_in == d_out && plan.npass > 1) ZK_TRY(tmp.alloc(N * F::N * 4));  // in place: the gather pass needs it
  const uint32_t *ep = (eta && !inverse) ? etab.as<uint32_t>() : nullptr;
  const uint32_t *eo = (eta && inverse) ? etab.as<uint32_t>() : nullptr;
  // inverse over several passes: N^-1 rides on the gather pass's store twiddle (its next pass
  // views the vector with logr = s_0 + ... + s_{P-3}), and the final pass only reduces
  const uint32_t *tw_ninv = nullptr;
  if (inverse && plan.npass > 1) {
    int step = 0;
    for (int i = 0; i + 2 < plan.npass; i++) step += plan.s[i];
    ZK_TRY(get_ninv_twiddles<F>(curve, m, gen, step, tw, tw_shift, lastf, &tw_ninv, st));
    if (tw_ninv) last_factor<F>(lastf, curve, m, false);  // one: only a coset post-scale remains
  }
  ZK_TRY(ntt_fast<F>(plan, d_in, d_out, tmp.as<uint32_t>(), tw, tw_shift, inverse, ep, eo,
                     lastf, st, tw_ninv));
  ZK_HIP(hipStreamSynchronize(st));  // tmp and the eta tables are released on return
  drain.release();
  return ZK_OK;
}

// ---------------------------------------------------------------------------------------
// batched NTTs of uneven sizes: one launch per pass round across all jobs, captured once
// into a hipGraph and replayed

zkh::Fe fft_generator_host(int curve, int m) {
  // 2-adic generators of Fr (lib/src/ZK/Algebra/Curves/BN128/Fr/Mont.hs:148-151,
  // BLS12_381/Fr/Mont.hs:148-151), Montgomery form
  static const uint64_t bn_gen[4] = {0x636e735580d13d9cULL, 0xa22bf3742445ffd6ULL, 0x56452ac01eb203d8ULL,
                                     0x1860ef942963f9e7ULL};
  static const uint64_t bls_gen[4] = {0xb9b58d8c5f0e466aULL, 0x5b1b4c801819d7ecULL, 0x0af53ae352a31e64ULL,
                                      0x5bf3adda19e9b27bULL};
  const zkh::FieldDesc &f = fr_desc(curve);
  const int M = curve == 0 ? 28 : 32;
  zkh::Fe g = zkh::load(f, curve == 0 ? bn_gen : bls_gen);
  for (int i = 0; i < M - m; i++) g = zkh::sqr(f, g);
  return g;
}

struct NttBatchPlan {
  int curve = 0;
  int njobs = 0;
  hipStream_t st = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  DevBuf eta, eta_bases, items, prefix;
  uint64_t bytes_moved = 0;  // 2 * 32 * sum(N) (one read + one write of every vector)
  ~NttBatchPlan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (st) (void)hipStreamDestroy(st);
  }
};

template <class F>
static int ntt_batch_create_t(int curve, int njobs, const NttBatchJob *jobs, NttBatchPlan *pl) {
  const zkh::FieldDesc &f = fr_desc(curve);
  int mmax = 0;
  for (int j = 0; j < njobs; j++) {
    if (jobs[j].m < 1 || jobs[j].m > 30) {
      set_error("ntt batch: job log sizes must be in [1, 30]");
      return ZK_EINVAL;
    }
    if (jobs[j].d_src == jobs[j].d_tgt) {
      set_error("ntt batch: in-place jobs are not supported");
      return ZK_EINVAL;
    }
    if (jobs[j].m > mmax) mmax = jobs[j].m;
  }
  ZK_HIP(hipStreamCreateWithFlags(&pl->st, hipStreamNonBlocking));
  hipStream_t st = pl->st;
  // one twiddle table for the largest domain serves every job through a stride
  zkh::Fe gmax = fft_generator_host(curve, mmax);
  uint64_t gw[4];
  zkh::store(f, gw, gmax);
  const uint32_t *tw = nullptr;
  int shift0 = 0;
  ZK_TRY(get_twiddles<F>(curve, mmax, gw, &tw, &shift0, st));
  // coset tables (the passes need no scratch: every job is out of place)
  uint64_t ncoset = 0;
  std::vector<NttPlan> plans(njobs);
  for (int j = 0; j < njobs; j++) {
    // every item of a round shares the launch's LDS size (the largest item's): keep all tiles
    // at most 2^kBatchTileLog elements so that no job lowers the others' occupancy
    NttTuning tu = ntt_tuning_for(curve, jobs[j].m);
    tu.smax = std::min(tu.smax, kBatchTileLog);
    tu.tile_log = std::min(tu.tile_log, kBatchTileLog);
    plans[j] = ntt_plan(jobs[j].m, tu.smax, tu.tile_log);
    if (jobs[j].eta) ncoset++;
    pl->bytes_moved += 64ull << jobs[j].m;
  }
  std::vector<Fe<F>> bases;  // read by an asynchronous copy: lives until the sync below
  if (ncoset) {
    // every coset's three power tables in one launch (bases from the host)
    ZK_TRY(pl->eta.alloc(ncoset * 3 * 2048 * F::N * 4));
    bases.reserve(ncoset * 3);
    for (int j = 0; j < njobs; j++)
      if (jobs[j].eta) eta_bases<F>(curve, zkh::load(f, jobs[j].eta), &bases);
    ZK_TRY(pl->eta_bases.alloc(bases.size() * sizeof(Fe<F>)));
    ZK_HIP(hipMemcpyAsync(pl->eta_bases.p, bases.data(), bases.size() * sizeof(Fe<F>), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(eta_tables_multi_kernel<F>, dim3((unsigned)(ncoset * 3 * 2048 / 256)), dim3(256), 0, st,
                       pl->eta.as<uint32_t>(), pl->eta_bases.as<Fe<F>>());
  }
  std::vector<std::vector<PassArgs>> rounds;
  std::vector<std::vector<uint32_t>> blocks;
  uint64_t eoff = 0;
  for (int j = 0; j < njobs; j++) {
    const NttPlan &pj = plans[j];
    const int m = jobs[j].m;
    const bool inverse = jobs[j].inverse;
    const uint32_t *etab = nullptr;
    if (jobs[j].eta) {
      etab = pl->eta.as<uint32_t>() + eoff;
      eoff += 3 * 2048 * F::N;
    }
    uint32_t lastf[8];
    last_factor<F>(lastf, curve, m, inverse);
    // inverse over several passes: N^-1 on the gather pass's store twiddle, as for one transform
    const uint32_t *tw_ninv = nullptr;
    if (inverse && pj.npass > 1) {
      int step = 0;
      for (int i = 0; i + 2 < pj.npass; i++) step += pj.s[i];
      zkh::Fe gm = gmax;
      for (int i = m; i < mmax; i++) gm = zkh::sqr(f, gm);
      uint64_t gmw[4];
      zkh::store(f, gmw, gm);
      ZK_TRY(get_ninv_twiddles<F>(curve, m, gmw, step, tw, shift0 + (mmax - m), lastf, &tw_ninv, st));
      if (tw_ninv) last_factor<F>(lastf, curve, m, false);
    }
    PassArgs args[8];
    uint64_t groups[8];
    size_t lds[8];
    ZK_TRY(ntt_passes<F>(pj, (const uint32_t *)jobs[j].d_src, (uint32_t *)jobs[j].d_tgt, (uint32_t *)jobs[j].d_tgt,
                         tw, shift0 + (mmax - m), inverse, (!inverse) ? etab : nullptr, inverse ? etab : nullptr,
                         lastf, args, groups, lds, tw_ninv));
    for (int q = 0; q < pj.npass; q++) {
      if ((int)rounds.size() <= q) {
        rounds.resize(q + 1);
        blocks.resize(q + 1);
      }
      rounds[q].push_back(args[q]);
      blocks[q].push_back((uint32_t)groups[q]);
    }
  }
  ZK_HIP(hipGetLastError());
  // device tables: all rounds' items and block prefixes, back to back
  uint64_t nitems_total = 0;
  for (auto &r : rounds) nitems_total += r.size();
  ZK_TRY(pl->items.alloc(nitems_total * sizeof(PassArgs)));
  ZK_TRY(pl->prefix.alloc((nitems_total + rounds.size()) * 4));
  std::vector<PassArgs> h_items;
  std::vector<uint32_t> h_prefix;
  std::vector<uint32_t> round_blocks, round_item0, round_prefix0;
  std::vector<size_t> round_lds;
  for (size_t r = 0; r < rounds.size(); r++) {
    round_item0.push_back((uint32_t)h_items.size());
    round_prefix0.push_back((uint32_t)h_prefix.size());
    uint32_t acc = 0;
    size_t lds = 0;
    for (size_t i = 0; i < rounds[r].size(); i++) {
      h_items.push_back(rounds[r][i]);
      h_prefix.push_back(acc);
      acc += blocks[r][i];
      const size_t l = lds_tile_bytes(rounds[r][i].s + rounds[r][i].logt);
      if (l > lds) lds = l;
    }
    h_prefix.push_back(acc);
    round_blocks.push_back(acc);
    round_lds.push_back(lds);
  }
  ZK_HIP(hipMemcpyAsync(pl->items.p, h_items.data(), h_items.size() * sizeof(PassArgs), hipMemcpyHostToDevice, st));
  ZK_HIP(hipMemcpyAsync(pl->prefix.p, h_prefix.data(), h_prefix.size() * 4, hipMemcpyHostToDevice, st));
  ZK_HIP(hipStreamSynchronize(st));
  // capture the pass rounds
  ZK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  for (size_t r = 0; r < rounds.size(); r++) {
    hipLaunchKernelGGL(ntt_batch_pass_kernel<F>, dim3(round_blocks[r]), dim3(kNttThreads), round_lds[r], st,
                       pl->items.as<PassArgs>() + round_item0[r], pl->prefix.as<uint32_t>() + round_prefix0[r],
                       (int)rounds[r].size());
  }
  hipError_t ce = hipStreamEndCapture(st, &pl->graph);
  if (ce != hipSuccess) {
    set_error(std::string("ntt batch capture: ") + hipGetErrorString(ce));
    return ZK_EHIP;
  }
  ZK_HIP(hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0));
  pl->curve = curve;
  pl->njobs = njobs;
  return ZK_OK;
}

int ntt_batch_create(int curve, int njobs, const NttBatchJob *jobs, NttBatchPlan **out) {
  if ((curve != 0 && curve != 1) || njobs <= 0) {
    set_error("ntt batch: bad curve or job count");
    return ZK_EINVAL;
  }
  NttBatchPlan *pl = new NttBatchPlan();
  int s = curve == 0 ? ntt_batch_create_t<Bn254Fr>(curve, njobs, jobs, pl)
                     : ntt_batch_create_t<Bls381Fr>(curve, njobs, jobs, pl);
  if (s != ZK_OK) {
    delete pl;
    return s;
  }
  *out = pl;
  return ZK_OK;
}

int ntt_batch_run(NttBatchPlan *pl, double *ms) {
  hipEvent_t a, b;
  ZK_HIP(hipEventCreate(&a));
  ZK_HIP(hipEventCreate(&b));
  ZK_HIP(hipEventRecord(a, pl->st));
  ZK_HIP(hipGraphLaunch(pl->exec, pl->st));
  ZK_HIP(hipEventRecord(b, pl->st));
  ZK_HIP(hipEventSynchronize(b));
  float t = 0;
  ZK_HIP(hipEventElapsedTime(&t, a, b));
  if (ms) *ms = t;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ZK_OK;
}

void ntt_batch_destroy(NttBatchPlan *pl) { delete pl; }

uint64_t ntt_batch_bytes(const NttBatchPlan *pl) { return pl->bytes_moved; }

// ---------------------------------------------------------------------------------------
// autotuner: every (smax, tile_log) on one random vector, timed with HIP events; all outputs
// must agree word for word

__global__ void count_mismatch_kernel(const uint4 *__restrict__ a, const uint4 *__restrict__ b, uint64_t n,
                                      uint32_t *__restrict__ bad) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint4 x = a[i], y = b[i];
  if (x.x != y.x || x.y != y.y || x.z != y.z || x.w != y.w) atomicAdd(bad, 1u);
}

int ntt_tune(int curve, int m, int inverse, int reps, NttTuning *best, double *best_ms, int *identical,
             hipStream_t st) {
  if (m < 1 || m > 26 || (curve != 0 && curve != 1)) {
    set_error("ntt_tune: bad curve id or log size");
    return ZK_EINVAL;
  }
  const uint64_t N = 1ull << m, words = N * 8;
  DevBuf in, out, ref, bad;
  ZK_TRY(in.alloc(words * 4));
  ZK_TRY(out.alloc(words * 4));
  ZK_TRY(ref.alloc(words * 4));
  ZK_TRY(bad.alloc(16));
  ZK_TRY(fill_random_device(curve, 1, N, 0x5eed0000u + (uint64_t)m, in.p));
  const zkh::Fe g = fft_generator_host(curve, m);
  uint64_t gw[4];
  zkh::store(fr_desc(curve), gw, g);
  hipEvent_t e0, e1;
  ZK_HIP(hipEventCreate(&e0));
  ZK_HIP(hipEventCreate(&e1));
  bool first = true, same = true;
  double bms = 1e300;
  NttTuning bt = ntt_tuning_for(curve, m);
  for (int smax = 5; smax <= 11; smax++) {
    for (int tl = 7; tl <= kMaxTileLog; tl++) {
      const NttTuning t{smax, tl};
      std::vector<float> ts;
      for (int r = -1; r < std::max(1, reps); r++) {
        ZK_HIP(hipEventRecord(e0, st));
        ZK_TRY(ntt_device(curve, m, gw, nullptr, inverse ? NTT_INVERSE : NTT_FORWARD, in.p, out.p, st, &t));
        ZK_HIP(hipEventRecord(e1, st));
        ZK_HIP(hipEventSynchronize(e1));
        float ms = 0;
        ZK_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (r >= 0) ts.push_back(ms);
      }
      std::sort(ts.begin(), ts.end());
      const double med = ts[ts.size() / 2];
      if (first) {
        ZK_HIP(hipMemcpyAsync(ref.p, out.p, words * 4, hipMemcpyDeviceToDevice, st));
        first = false;
      } else {
        uint32_t nbad = 0;
        ZK_HIP(hipMemsetAsync(bad.p, 0, 4, st));
        hipLaunchKernelGGL(count_mismatch_kernel, dim3((unsigned)((words / 4 + 255) / 256)), dim3(256), 0, st,
                           out.as<uint4>(), ref.as<uint4>(), words / 4, bad.as<uint32_t>());
        ZK_HIP(hipMemcpyAsync(&nbad, bad.p, 4, hipMemcpyDeviceToHost, st));
        ZK_HIP(hipStreamSynchronize(st));
        if (nbad) same = false;
      }
      if (med < bms) {
        bms = med;
        bt = t;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (best) *best = bt;
  if (best_ms) *best_ms = bms;
  if (identical) *identical = same ? 1 : 0;
  return ZK_OK;
}

int ntt_device(int curve, int m, const uint64_t *gen, const uint64_t *eta, int flags, const void *d_in,
               void *d_out, hipStream_t st, const NttTuning *tune) {
  if (m < 0 || m > 30 || (curve != 0 && curve != 1)) {
    set_error("ntt: bad curve id or log size");
    return ZK_EINVAL;
  }
  if (curve == 0)
    return ntt_device_t<Bn254Fr>(curve, m, gen, eta, flags, (const uint32_t *)d_in, (uint32_t *)d_out, st, tune);
  return ntt_device_t<Bls381Fr>(curve, m, gen, eta, flags, (const uint32_t *)d_in, (uint32_t *)d_out, st, tune);
}

// ---------------------------------------------------------------------------------------
// NTT without allocation, for callers that cannot allocate device memory.  The pass scratch
// vector, the twiddle table (rebuilt on every call, never cached) and the coset tables live in
// one file-scope device array, `scratch`, of a fixed size: transforms of at most 2^18 elements
// (12.3 MiB, held on every GPU the library runs on until the process ends).  Every call on a
// GPU shares that GPU's array and leaves the last call's contents in it, so calls that overlap
// on one GPU race.  Generators of order 2^m only (the fast path): others would need the
// reference network's powers table beside an aliasing copy.

static constexpr int kNoAllocMaxLog = 18;
static constexpr size_t kNaWords = 8;  // u32 words per stored Fr element
static constexpr size_t kNaTmp = 0;                                                  // 2^18 elements
static constexpr size_t kNaTw = kNaTmp + (kNaWords << kNoAllocMaxLog);               // 2^17 + 1 twiddles
static constexpr size_t kNaLo = kNaTw + kNaWords * ((1u << (kNoAllocMaxLog - 1)) + 8);   // 2048 powers
static constexpr size_t kNaHi = kNaLo + kNaWords * 2048;                             // 65 powers
static constexpr size_t kNaEta = kNaHi + kNaWords * ((1u << (kNoAllocMaxLog - 1)) / 2048 + 8);  // 3 x 2048
static constexpr size_t kNaTotal = kNaEta + kNaWords * 3 * 2048;
static_assert(kNaTw % 64 == 0 && kNaLo % 64 == 0 && kNaHi % 64 == 0 && kNaEta % 64 == 0,
              "scratch regions stay 256-byte aligned");

static __device__ __align__(256) uint32_t scratch[kNaTotal];

// device address of `scratch` on each device (the host-side name is not a device address)
static std::mutex g_scratch_mu;
static std::vector<uint32_t *> g_scratch_ptr;

static int scratch_on_current_device(uint32_t **out) {
  int dev = 0;
  ZK_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((size_t)dev >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    ZK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      set_error("ntt_noalloc: no device address for the scratch array");
      return ZK_EHIP;
    }
    g_scratch_ptr[dev] = (uint32_t *)p;
  }
  *out = g_scratch_ptr[dev];
  return ZK_OK;
}

template <class F>
static int ntt_noalloc_t(int curve, int m, const uint64_t *gen, const uint64_t *eta, int flags, const uint32_t *d_in,
                         uint32_t *d_out, hipStream_t st, uint32_t *scr) {
  const bool inverse = flags & NTT_INVERSE;
  if (m == 0 && !eta) {
    if (d_in != d_out) ZK_HIP(hipMemcpyAsync(d_out, d_in, F::N * 4, hipMemcpyDeviceToDevice, st));
    ZK_HIP(hipStreamSynchronize(st));
    return ZK_OK;
  }
  if (!ntt_generator_ok(curve, m, gen)) {
    set_error("ntt_noalloc: the generator must have order 2^m");
    return ZK_ENOTROOT;
  }
  const zkh::FieldDesc &f = fr_desc(curve);
  const uint32_t *tw = nullptr;
  if (m >= 1) {
    ZK_TRY(launch_twiddles<F>(scr + kNaTw, scr + kNaLo, scr + kNaHi, (1ull << (m - 1)) + 1, zkh::load(f, gen), st));
    tw = scr + kNaTw;
  }
  if (eta) ZK_TRY(launch_eta_tables<F>(scr + kNaEta, curve, zkh::load(f, eta), st));
  uint32_t lastf[8];
  last_factor<F>(lastf, curve, m, inverse);
  const NttTuning tu = ntt_tuning_for(curve, m);
  const NttPlan plan = ntt_plan(m, tu.smax, tu.tile_log);
  const uint32_t *ep = (eta && !inverse) ? scr + kNaEta : nullptr;
  const uint32_t *eo = (eta && inverse) ? scr + kNaEta : nullptr;
  ZK_TRY(ntt_fast<F>(plan, d_in, d_out, scr + kNaTmp, tw, 0, inverse, ep, eo, lastf, st));
  ZK_HIP(hipStreamSynchronize(st));
  return ZK_OK;
}

int ntt_device_noalloc(int curve, int m, const uint64_t *gen, const uint64_t *eta, int flags, const void *d_in,
                       void *d_out, hipStream_t st) {
  if (curve != 0 && curve != 1) {
    set_error("ntt_noalloc: bad curve id");
    return ZK_EINVAL;
  }
  if (m < 0 || m > kNoAllocMaxLog) {
    set_error("ntt_noalloc: log size must be 0..18 (its fixed scratch holds 2^18 elements)");
    return ZK_EINVAL;
  }
  uint32_t *scr = nullptr;
  ZK_TRY(scratch_on_current_device(&scr));
  if (curve == 0)
    return ntt_noalloc_t<Bn254Fr>(curve, m, gen, eta, flags, (const uint32_t *)d_in, (uint32_t *)d_out, st, scr);
  return ntt_noalloc_t<Bls381Fr>(curve, m, gen, eta, flags, (const uint32_t *)d_in, (uint32_t *)d_out, st, scr);
}

}  // namespace zk
