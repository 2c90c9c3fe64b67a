// FFT of G1 points over the scalar field's multiplicative subgroups (BN254 / BLS12-381).
//
// Contract (lib/cbits/curves/g1/proj/bn128_G1_proj.c:791-902): src and tgt hold 2^m
// projective Montgomery points (X:Y:Z); the forward transform is y_k = sum_i gen^(ik) x_i
// computed by the recursive radix-2 DIT network with twiddles gen^j, the inverse is the
// mirrored DIF network with (x + y)/2 and (x - y) gen^-j / 2 at every level; every output is
// normalised (X/Z, Y/Z, 1), infinity (0, 1, 0).  For any gen these networks are the same
// group-linear maps as the iterative forms below (bit-reversed DIT with gen_s = gen^(N/2^s),
// DIF with gen^-j, and the m halvings of every output path applied once as N^-1), and the
// normalised point is unique, so the outputs match bit for bit.  When gen is a primitive 2^m-th
// root of unity both networks compute the DFT, and the inverse runs as N^-1 times the forward
// network by gen^-1 (g1_fft_run).
//
// Device form: points in XYZZ reduced-radix coordinates (one record per point, in HBM),
// twiddles as standard-form scalars; each butterfly multiplies one point by its twiddle (2-bit
// windows over a per-thread table of small multiples: the joint GLV digits for BN254, the
// scalar's own for BLS12-381; a Jacobian chain, see table_scalar_mul) and adds / subtracts.
//
// A scalar multiplication is a chain of ~127 (BN254, GLV) or ~255 dependent doublings, so a
// stage of a small transform is one chain long however few threads it has, and one launch per
// radix-2 stage makes the transform m chains long.  Grouped stages cut that depth: r radix-2
// stages s0+1 .. s0+r map each group of 2^r points (spaced 2^s0 apart) by a 2^r x 2^r matrix
// whose entries are +-(products of twiddles along the unique network path), so every scalar
// multiplication of the group can run at once, followed (DIT) or preceded and followed (DIF)
// by additions.  Entries equal up to sign are shared: (4^r - 1) / 3 scalar multiplications per
// group instead of r 2^(r-1), in one chain instead of r.  The stage plan picks the group sizes
// by a launch-cost model (g1_stage_plan), so large transforms keep the cheaper radix-2 stages.
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "curve.hpp"
#include "curvez.hpp"
#include "glv.hpp"
#include "field.hpp"
#include "hostfield.hpp"
#include "msm.hpp"
#include "runtime.hpp"

namespace zk {

static constexpr int kG1Threads = 128;
// group and radix-2 launches of at most one wavefront per SIMD (split or not) of BN254, whose
// kernels could hold two wavefronts per SIMD: workgroups of four wavefronts with LDS reserved so
// that a CU holds one (with two-wavefront workgroups the dispatcher pairs wavefronts on SIMDs:
// BN254 2^17, 1024 wavefronts a radix-2 stage, 17.2 against 28.1 ms).
// BLS12-381's hold one wavefront per SIMD by their registers and launch as the other kernels.
static constexpr int kG1SplitThreads = 256;
static constexpr size_t kG1SplitLds = 96 * 1024;
template <class C>
__host__ __device__ constexpr int g1_split_threads() {
  return C::id == 0 ? kG1SplitThreads : kG1Threads;
}
template <class C>
__host__ __device__ constexpr size_t g1_split_lds() {
  return C::id == 0 ? kG1SplitLds : 0;
}

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

// tab[j] = g^j in standard form (8 u32 words), j < cnt; with tabm, also the Montgomery form
// there (the grouped stages' path products) from the same power
template <class C>
__global__ __launch_bounds__(kG1Threads) void g1_twiddle_kernel(uint32_t *__restrict__ tab,
                                                               uint32_t *__restrict__ tabm, Fe<typename C::Fr> g,
                                                               uint32_t cnt) {
  using FR = typename C::Fr;
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cnt) return;
  Fe<FR> r;
  fpow_u64(r, g, j);
  if (tabm) fstore(tabm + (size_t)j * FR::N, r);
  fto_std(r, r);
  fstore(tab + (size_t)j * FR::N, r);
}

// k P for the twiddle products of every kernel below, where thread t of a launch (stride = the
// table's threads) owns g1_tab_entries<C>() cached Jacobian point-table entries of tab: BN254
// by 2-bit joint GLV windows over its thread-major block (at t g1_tab_words words; exact for
// every point of its prime-order G1), BLS12-381, whose G1 has a cofactor, by 2-bit windows on
// k over the word-major table.  The chains run in Jacobian coordinates.
template <class C>
__host__ __device__ constexpr int g1_tab_entries() {
  return C::id == 0 ? kGlvW2Entries : 3;
}
// u32 words of one thread's table
template <class C>
__host__ __device__ constexpr int g1_tab_words() {
  return C::id == 0 ? kGlvW2Entries * ZLayout<C>::JACC : 3 * 5 * ZLayout<C>::N;
}
// 2-bit windows of a product: 64 joint GLV windows (BN254), 128 windows on k (BLS12-381)
template <class C>
__host__ __device__ constexpr int g1_nwin() {
  return C::id == 0 ? 64 : 128;
}
// Part j of a product split over nparts threads (1 to 3): the windows [wlo, whi) it adds, part 0
// the highest.  Every part but the lowest still doubles down to window 0, so the split points
// balance the chains, a window's addition costing about three doublings (measured): for two
// parts BN254 windows [40, 64) (128 doublings, ~24 additions) and [0, 40) (80, ~40), BLS12-381
// [80, 128) and [0, 80) (1-5% faster at 2^12 than 38, 42 or 45 / 76, 84 or 90); for three
// parts [52, 64), [33, 52), [0, 33) and [104, 128), [65, 104), [0, 65) (longest chain about
// 160 doubling-equivalents for BN254, against 200 with two parts and 320 whole; at 2^12 these
// points measured 1-2% faster than 50 / 31, 54 / 35 and 52 / 30).
template <class C>
__host__ __device__ __forceinline__ void g1_part_windows(int nparts, int j, int &wlo, int &whi) {
  constexpr int W = g1_nwin<C>();
  constexpr int A2 = C::id == 0 ? 40 : 80;
  constexpr int A3 = C::id == 0 ? 52 : 104, B3 = C::id == 0 ? 33 : 65;
  if (nparts == 3) {
    wlo = j == 0 ? A3 : (j == 1 ? B3 : 0);
    whi = j == 0 ? W : (j == 1 ? A3 : B3);
  } else if (nparts == 2) {
    wlo = j == 0 ? A2 : 0;
    whi = j == 0 ? W : A2;
  } else {
    wlo = 0;
    whi = W;
  }
}
template <class C>
__device__ __forceinline__ void table_scalar_mul(XyzzZ<C> &r, const XyzzZ<C> &p, const uint32_t (&k)[8],
                                                 uint32_t *__restrict__ tab, uint64_t t, uint64_t stride,
                                                 int wlo = 0, int whi = g1_nwin<C>()) {
  if constexpr (C::id == 0)
    zxyzz_scalar_mul_glv_w2(r, p, k, tab + t * (uint64_t)g1_tab_words<C>(), wlo, whi);
  else
    zxyzz_scalar_mul_w2(r, p, k, tab, t, stride, wlo, whi);
}

// one radix-2 stage over blocks of 2^s points (DIT: u + w v, u - w v; DIF: u + v, w (u - v)),
// nthr threads striding over the N/2 butterflies (thread t owns table block t of tab, stride
// nthr)
template <class C, bool DIF>
__global__ __launch_bounds__(kG1SplitThreads) void g1_fft_stage_kernel(int32_t *__restrict__ pts,
                                                                 const uint32_t *__restrict__ tw, int m, int s,
                                                                 uint32_t *__restrict__ tab, uint32_t nthr) {
  constexpr int PW = ZLayout<C>::XYZZ;
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= nthr) return;
  for (uint32_t t = tid; t < (1u << (m - 1)); t += nthr) {
    const uint32_t half = 1u << (s - 1), pos = t & (half - 1), grp = t >> (s - 1);
    const uint32_t i0 = (grp << s) | pos, i1 = i0 + half;
    uint32_t k[8];
    const uint32_t *kp = tw + ((size_t)pos << (m - s)) * 8;
#pragma unroll
    for (int j = 0; j < 8; j++) k[j] = kp[j];
    // u is read after (DIT) or stored before (DIF) the scalar multiplication, so that no point
    // but the one multiplied is live across the chain (the register count sets the occupancy)
    if (!DIF) {
      XyzzZ<C> v, w;
      zxyzz_load(v, pts + (size_t)i1 * PW);
      if (pos == 0) w = v;  // twiddle 1
      else table_scalar_mul<C>(w, v, k, tab, tid, nthr);
      XyzzZ<C> u, a;
      zxyzz_load(u, pts + (size_t)i0 * PW);
      zxyzz_add(a, u, w);
      zxyzz_neg(w, w);
      zxyzz_add(v, u, w);
      zxyzz_store(pts + (size_t)i0 * PW, a);
      zxyzz_store(pts + (size_t)i1 * PW, v);
    } else {
      XyzzZ<C> w;
      {
        XyzzZ<C> u, v, a;
        zxyzz_load(u, pts + (size_t)i0 * PW);
        zxyzz_load(v, pts + (size_t)i1 * PW);
        zxyzz_add(a, u, v);
        zxyzz_store(pts + (size_t)i0 * PW, a);
        zxyzz_neg(v, v);
        zxyzz_add(w, u, v);
      }
      if (pos != 0) {
        XyzzZ<C> r;
        table_scalar_mul<C>(r, w, k, tab, tid, nthr);
        w = r;
      }
      zxyzz_store(pts + (size_t)i1 * PW, w);
    }
  }
}

// pts[i] = k pts[i], nthr threads striding over the N points (tables as above)
template <class C>
__global__ __launch_bounds__(kG1Threads) void g1_scale_kernel(int32_t *__restrict__ pts, uint32_t N, Fe<typename C::Fr> kst,
                                                             uint32_t *__restrict__ tab, uint32_t nthr) {
  constexpr int PW = ZLayout<C>::XYZZ;
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= nthr) return;
  uint32_t k[8];
#pragma unroll
  for (int j = 0; j < 8; j++) k[j] = kst.w[j];
  for (uint32_t i = tid; i < N; i += nthr) {
    XyzzZ<C> p, r;
    zxyzz_load(p, pts + (size_t)i * PW);
    table_scalar_mul<C>(r, p, k, tab, tid, nthr);
    zxyzz_store(pts + (size_t)i * PW, r);
  }
}

// ---- grouped stages ---------------------------------------------------------------------
//
// Group G = (ghi, p0), p0 < 2^s0: points b + q 2^s0, q < 2^r, with b = ghi 2^(s0+r) + p0.  In
// the DIT network (stages s0+1 .. s0+r, t = 0 .. r-1) input q reaches output q' through
// stage t on the multiplied side exactly when bit t of q is set, with twiddle
// gen^((p0 + (q' mod 2^t) 2^s0) 2^(m - s0 - t - 1)) and sign -1 when bit t of q' is set too:
//     y_q' = x_0 + sum_{q >= 1} (-1)^popcount(q & q') M(q, q' mod 2^h(q)) x_q,
// h(q) the top bit of q.  In the DIF network (t = r-1 .. 0, twiddles of gen^-1) the roles of
// q and q' swap:
//     y_q' = sum_{L < 2^h(q')} M'(q', L) sum_{q mod 2^h(q') = L} (-1)^popcount(q & q') x_q,
// and y_0 = sum_q x_q.  Slot k of a group, k < (4^r - 1) / 3, is the pair (index with top bit
// h, L < 2^h): k = (4^h - 1) / 3 + ((index - 2^h) << h | L); slot (4^r - 1) / 3 holds x_0
// (DIT) or y_0 before the final scale (DIF).

__host__ __device__ __forceinline__ uint32_t g1_kcount(int r) { return ((1u << (2 * r)) - 1) / 3; }

__device__ __forceinline__ void g1_slot_decode(uint32_t k, int &h, uint32_t &idx, uint32_t &L) {
  uint32_t base = 0;
  h = 0;
  while (k >= base + (1u << (2 * h))) {
    base += 1u << (2 * h);
    h++;
  }
  const uint32_t i = k - base;
  idx = (1u << h) | (i >> h);
  L = i & ((1u << h) - 1);
}
__device__ __forceinline__ uint32_t g1_slot(int h, uint32_t idx, uint32_t L) {
  return ((1u << (2 * h)) - 1) / 3 + (((idx - (1u << h)) << h) | L);
}

// M(idx, L) = product over set bits t of idx of tw[(p0 + (L mod 2^t) 2^s0) << (m - s0 - t - 1)]
// (Montgomery powers of the stage generator; every index is below N / 2), times `extra` when
// has_extra; returned in standard form for the scalar multiplication
template <class C>
__device__ __forceinline__ void g1_path_coeff(Fe<typename C::Fr> &c, const uint32_t *__restrict__ twm, int m, int s0,
                                              uint32_t p0, uint32_t idx, uint32_t L, int h, bool has_extra,
                                              const Fe<typename C::Fr> &extra) {
  using FR = typename C::Fr;
  bool have = false;
  for (int t = 0; t <= h; t++) {
    if (!((idx >> t) & 1u)) continue;
    const uint64_t pos = p0 + ((uint64_t)(L & ((1u << t) - 1)) << s0);
    Fe<FR> w;
    fload(w, twm + (pos << (m - s0 - t - 1)) * FR::N);
    if (have) fmul(c, c, w);
    else c = w;
    have = true;
  }
  if (!have) fset_one(c);
  if (has_extra) fmul(c, c, extra);
  fto_std(c, c);
}

// Unsplit, one thread per slot; slot k of group G lives at scr[G (K + 1) + k] (plane 0).
// (Splitting a BN254 slot into its two GLV halves k1 P and k2 phi(P) on two threads gives each
// the whole doubling chain and measured 1.75x slower; the window splits below shorten it.)

// Sum of the nterm terms of one output over L lanes (L a power of two <= 64): lane j adds terms
// j, j + L, ... (add_term(i, acc) adds term i into acc), then log2(L) xor-shuffle rounds combine
// the L partial sums (a chain of nterm / L + log2(L) additions instead of nterm).  All L lanes
// of an output group must call it.
template <class C, class AddTerm>
__device__ __forceinline__ void g1_sum_lanes(XyzzZ<C> &acc, uint32_t nterm, uint32_t lane, int L, AddTerm add_term) {
  constexpr int N = ZP<C>::T::N;
  XyzzZ<C> o;
  zxyzz_set_inf(acc);
  for (uint32_t i = lane; i < nterm; i += (uint32_t)L) add_term(i, acc);
  for (int off = L >> 1; off >= 1; off >>= 1) {
#pragma unroll
    for (int l = 0; l < N; l++) {
      o.x.l[l] = __shfl_xor(acc.x.l[l], off, L);
      o.y.l[l] = __shfl_xor(acc.y.l[l], off, L);
      o.zz.l[l] = __shfl_xor(acc.zz.l[l], off, L);
      o.zzz.l[l] = __shfl_xor(acc.zzz.l[l], off, L);
    }
    zxyzz_add(acc, acc, o);
  }
}

// Split launches (small transforms, where one thread per slot leaves most SIMDs idle): each
// product is split over nparts threads by windows (g1_part_windows), in separate wavefronts,
// so the longest chain is shorter; part j lands in plane j of scr and the sum kernels add the
// planes.  Plane 0 holds the K + 1 slots of every group (T0 = ngroups (K + 1) threads, thread
// = slot); planes 1.. hold only the slots that are products, k1 a group (K, or K + 1 when
// slot K is a product too: the group that takes N^-1), T1 = ngroups k1 threads each; a
// thread's index is also its record in scr.  The launch keeps one wavefront per SIMD (a
// wavefront sharing its SIMD runs its chain at most half as fast): BN254 through workgroups
// of four wavefronts and LDS reserved so that a CU holds one, BLS12-381 through its register
// count.  (With 128-thread workgroups and two per CU, BN254 2^12 measured 4.1 against 2.6 ms:
// the launch's wavefronts were paired on SIMDs.)  At 2^12 three parts take 256 x 86 + 2 x 256
// x 85 = 65536 threads: one wavefront for each of the 1024 SIMDs.

// thread -> (part, group, slot) of a split launch (nparts = 1: thread = slot of plane 0)
struct G1Slot {
  int part;
  uint32_t G, k;
};
__device__ __forceinline__ G1Slot g1_thread_slot(uint64_t tid, uint32_t K, uint32_t k1, uint64_t T0, uint64_t T1) {
  G1Slot r;
  if (tid < T0) {
    r.part = 0;
    r.G = (uint32_t)(tid / (K + 1));
    r.k = (uint32_t)(tid % (K + 1));
  } else {
    const uint64_t u = tid - T0;
    r.part = 1 + (int)(u / T1);
    const uint64_t v = u % T1;
    r.G = (uint32_t)(v / k1);
    r.k = (uint32_t)(v % k1);
  }
  return r;
}
// acc += (-1)^neg slot k of group G, one plane at a time (one point live beside acc)
template <class C>
__device__ __forceinline__ void g1_slot_add(XyzzZ<C> &acc, const int32_t *__restrict__ scr, uint32_t G, uint32_t k,
                                            uint32_t K, uint32_t k1, uint64_t T0, uint64_t T1, int nparts,
                                            bool neg) {
  constexpr int PW = ZLayout<C>::XYZZ;
  const int planes = k < k1 ? nparts : 1;
  for (int j = 0; j < planes; j++) {
    XyzzZ<C> t;
    zxyzz_load(t, scr + (j == 0 ? (uint64_t)G * (K + 1) + k : T0 + (uint64_t)(j - 1) * T1 + (uint64_t)G * k1 + k) * PW);
    if (neg) zxyzz_neg(t, t);
    zxyzz_add(acc, acc, t);
  }
}

// DIT: slot k < K of group G = M(q, L) x_q; slot K = x_0; with has_fin every slot is also
// multiplied by fin (N^-1, on the first group of an inverse run as a DIT network), slot K then
// being a product too
template <class C>
__global__ __launch_bounds__(kG1SplitThreads) void g1_group_dit_mul_kernel(const int32_t *__restrict__ pts,
                                                                     int32_t *__restrict__ scr,
                                                                     const uint32_t *__restrict__ twm, int m, int s0,
                                                                     int r, uint32_t ngroups, int has_fin,
                                                                     Fe<typename C::Fr> fin, uint32_t *__restrict__ tab,
                                                                     uint64_t stride, int nparts) {
  using FR = typename C::Fr;
  constexpr int PW = ZLayout<C>::XYZZ;
  const uint32_t K = g1_kcount(r), k1 = has_fin ? K + 1 : K;
  const uint64_t T0 = (uint64_t)ngroups * (K + 1), T1 = (uint64_t)ngroups * k1;
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= T0 + (uint64_t)(nparts - 1) * T1) return;
  const G1Slot sl = g1_thread_slot(tid, K, k1, T0, T1);
  const uint32_t G = sl.G, k = sl.k;
  const uint32_t p0 = G & ((1u << s0) - 1), ghi = G >> s0;
  const uint64_t b = ((uint64_t)ghi << (s0 + r)) | p0;
  XyzzZ<C> x;
  Fe<FR> c;
  if (k == K) {
    zxyzz_load(x, pts + b * PW);
    if (!has_fin) {  // x_0 (plane 0 only)
      zxyzz_store(scr + tid * PW, x);
      return;
    }
    c = fin;
    fto_std(c, c);
  } else {
    int h;
    uint32_t q, L;
    g1_slot_decode(k, h, q, L);
    g1_path_coeff<C>(c, twm, m, s0, p0, q, L, h, has_fin != 0, fin);
    zxyzz_load(x, pts + (b + ((uint64_t)q << s0)) * PW);
  }
  XyzzZ<C> y;
  int wlo, whi;
  g1_part_windows<C>(nparts, sl.part, wlo, whi);
  table_scalar_mul<C>(y, x, c.w, tab, tid, stride, wlo, whi);
  zxyzz_store(scr + tid * PW, y);
}

// DIT: y_q' = x_0 + sum_q (-1)^popcount(q & q') slot(q, q' mod 2^h(q)), min(Q, 8) lanes per
// output (r = 2: four lanes of one term each, against eight with half idle: BN254 2^16 2.5%
// faster; sixteen lanes for r = 4 measured 1.5% slower than eight at 2^12)
template <class C>
__global__ __launch_bounds__(kG1Threads) void g1_group_dit_sum_kernel(int32_t *__restrict__ pts,
                                                                     const int32_t *__restrict__ scr, int s0, int r,
                                                                     uint32_t ngroups, int nparts, int has_fin) {
  constexpr int PW = ZLayout<C>::XYZZ;
  const uint32_t K = g1_kcount(r), Q = 1u << r, k1 = has_fin ? K + 1 : K;
  const uint64_t T0 = (uint64_t)ngroups * (K + 1), T1 = (uint64_t)ngroups * k1;
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lg = r < 3 ? r : 3;  // log2 of the lanes per output
  const uint64_t out = tid >> lg;
  const uint32_t lane = (uint32_t)(tid & ((1u << lg) - 1));
  const bool live = out < (uint64_t)ngroups * Q;  // whole lane groups are live or not
  const uint32_t G = live ? (uint32_t)(out / Q) : 0u, qq = live ? (uint32_t)(out % Q) : 0u;
  // term 0: x_0; term q >= 1: slot (q, q' mod 2^h(q)) with its sign
  const uint32_t nterm = live ? Q : 0u;
  XyzzZ<C> acc;
  g1_sum_lanes<C>(acc, nterm, lane, 1 << lg, [&](uint32_t q, XyzzZ<C> &a) {
    if (q == 0) {
      g1_slot_add<C>(a, scr, G, K, K, k1, T0, T1, nparts, false);
      return;
    }
    const int h = 31 - __clz(q);
    g1_slot_add<C>(a, scr, G, g1_slot(h, q, qq & ((1u << h) - 1)), K, k1, T0, T1, nparts, (__popc(q & qq) & 1) != 0);
  });
  if (live && lane == 0) {
    const uint32_t p0 = G & ((1u << s0) - 1), ghi = G >> s0;
    const uint64_t b = ((uint64_t)ghi << (s0 + r)) | p0;
    zxyzz_store(pts + (b + ((uint64_t)qq << s0)) * PW, acc);
  }
}

// DIF: slot k < K = M'(q', L) (sum over q mod 2^h = L of (-1)^popcount(q & q') x_q), times
// `fin` (N^-1) on the last group; slot K = fin (sum of all x_q)
template <class C>
__global__ __launch_bounds__(kG1SplitThreads) void g1_group_dif_mul_kernel(const int32_t *__restrict__ pts,
                                                                     int32_t *__restrict__ scr,
                                                                     const uint32_t *__restrict__ twm, int m, int s0,
                                                                     int r, uint32_t ngroups, int has_fin,
                                                                     Fe<typename C::Fr> fin, uint32_t *__restrict__ tab,
                                                                     uint64_t stride, int nparts) {
  using FR = typename C::Fr;
  constexpr int PW = ZLayout<C>::XYZZ;
  const uint32_t K = g1_kcount(r), Q = 1u << r, k1 = has_fin ? K + 1 : K;
  const uint64_t T0 = (uint64_t)ngroups * (K + 1), T1 = (uint64_t)ngroups * k1;
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= T0 + (uint64_t)(nparts - 1) * T1) return;
  const G1Slot sl = g1_thread_slot(tid, K, k1, T0, T1);
  const uint32_t G = sl.G, k = sl.k;
  const uint32_t p0 = G & ((1u << s0) - 1), ghi = G >> s0;
  const uint64_t b = ((uint64_t)ghi << (s0 + r)) | p0;
  XyzzZ<C> s, t;
  int h = r;
  uint32_t qq = 0, L = 0;
  if (k < K) g1_slot_decode(k, h, qq, L);
  // the inputs q with q mod 2^h = L (all of them for y_0)
  const uint32_t nterm = k < K ? (Q >> h) : Q;
  for (uint32_t j = 0; j < nterm; j++) {
    const uint32_t q = k < K ? (L | (j << h)) : j;
    zxyzz_load(t, pts + (b + ((uint64_t)q << s0)) * PW);
    if (__popc(q & qq) & 1) zxyzz_neg(t, t);
    if (j == 0) s = t;
    else zxyzz_add(s, s, t);
  }
  Fe<FR> c;
  if (k < K) {
    g1_path_coeff<C>(c, twm, m, s0, p0, qq, L, h, has_fin != 0, fin);
  } else if (has_fin) {
    c = fin;
    fto_std(c, c);
  } else {  // the sum (plane 0 only)
    zxyzz_store(scr + tid * PW, s);
    return;
  }
  XyzzZ<C> y;
  int wlo, whi;
  g1_part_windows<C>(nparts, sl.part, wlo, whi);
  table_scalar_mul<C>(y, s, c.w, tab, tid, stride, wlo, whi);
  zxyzz_store(scr + tid * PW, y);
}

// DIF: y_0 = slot K; y_q' = sum_{L < 2^h(q')} slot(q', L); 8 lanes per output
template <class C>
__global__ __launch_bounds__(kG1Threads) void g1_group_dif_sum_kernel(int32_t *__restrict__ pts,
                                                                     const int32_t *__restrict__ scr, int s0, int r,
                                                                     uint32_t ngroups, int nparts, int has_fin) {
  constexpr int PW = ZLayout<C>::XYZZ;
  const uint32_t K = g1_kcount(r), Q = 1u << r, k1 = has_fin ? K + 1 : K;
  const uint64_t T0 = (uint64_t)ngroups * (K + 1), T1 = (uint64_t)ngroups * k1;
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t out = tid >> 3;
  const uint32_t lane = (uint32_t)(tid & 7);
  const bool live = out < (uint64_t)ngroups * Q;
  const uint32_t G = live ? (uint32_t)(out / Q) : 0u, qq = live ? (uint32_t)(out % Q) : 0u;
  const int h = qq ? 31 - __clz(qq) : 0;
  const uint32_t nterm = live ? (qq ? (1u << h) : 1u) : 0u;
  XyzzZ<C> acc;
  g1_sum_lanes<C>(acc, nterm, lane, 8, [&](uint32_t L, XyzzZ<C> &a) {
    g1_slot_add<C>(a, scr, G, qq ? g1_slot(h, qq, L) : K, K, k1, T0, T1, nparts, false);
  });
  if (live && lane == 0) {
    const uint32_t p0 = G & ((1u << s0) - 1), ghi = G >> s0;
    const uint64_t b = ((uint64_t)ghi << (s0 + r)) | p0;
    zxyzz_store(pts + (b + ((uint64_t)qq << s0)) * PW, acc);
  }
}


// Stage plan: group sizes r <= 4 summing to m (r = 1: the radix-2 kernel), the composition
// of least modelled time (dynamic programming over m; a launch's cost does not depend on where
// it sits).  A launch costs its chain (whole, or split in S parts: 1, 0.73, 0.67, measured at
// BN254 2^12) times the wavefronts on its busiest SIMD, two co-resident wavefronts (BN254's
// kernels hold two a SIMD, BLS12-381's one) costing 1.75 chains
// (the chains are latency-bound, so sharing a SIMD gains little), plus 0.08 of a chain for the
// sum launch.  Measured at BN254: 2^14 6.24 predicted / 6.2-6.4 ms, 2^15 9.15 / 9.3, 2^16 14.6 /
// 14.8 (in chains of ~1 ms).  The model holds where each wavefront that could have a SIMD of
// its own gets one: BN254's group launches of at most one wavefront a SIMD take the split
// launch shape, split or not.  group_cost(r) gives a group launch's cost, or the radix-2
// stage's for r = 1.
template <class Cost>
static void g1_stage_plan(int m, Cost group_cost, std::vector<int> &rs) {
  std::vector<double> best(m + 1, 1e300);
  std::vector<int> choice(m + 1, 1);
  best[0] = 0;
  for (int k = 1; k <= m; k++)
    for (int r = std::min(4, k); r >= 1; r--) {  // the larger r on (near) ties: fewer launches
      const double c = best[k - r] + group_cost(r);
      if (c < best[k] - 1e-9) {
        best[k] = c;
        choice[k] = r;
      }
    }
  rs.clear();
  for (int k = m; k > 0; k -= choice[k]) rs.push_back(choice[k]);
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
  // For a primitive N-th root (g^(N/2) = -1) both networks compute the DFT, so the inverse is
  // N^-1 times the forward network by g^-1: it runs the DIT kernels, whose products need no
  // input sums (the DIF products sum up to 2^r inputs first), with N^-1 folded into the first
  // group.  Any other generator keeps the reference's DIF network.
  bool dit = !inverse;
  if (inverse && m > 0) {
    const zkh::Fe h = zkh::pow_u64(fr, g, halfN), mone = zkh::neg(fr, zkh::one(fr));
    dit = memcmp(h.w, mone.w, 8 * fr.nl) == 0;
  }
  DevBuf pts, tw, twm, scr, tab;
  StreamSyncGuard drain(st);  // the buffers outlive the queued kernels on every return
  ZK_TRY(pts.alloc((size_t)N * ZW * 4));
  ZK_TRY(tw.alloc((size_t)(halfN ? halfN : 1) * 32));
  Context *ctx = nullptr;
  ZK_TRY(context(&ctx));
  std::vector<int> plan;
  // thread count cap of the radix-2 stages and the scaling pass (they stride over the rest)
  const uint64_t round_threads = (uint64_t)ctx->num_cus * 4 * 3 * 64;
  // a group launch splits each product over as many threads (up to three) as still give every
  // wavefront a SIMD (the chains are latency-bound there)
  const uint64_t split_cap = (uint64_t)ctx->num_cus * 4 * 64;
  if (g1_split_lds<C>() > 0) {
    static std::once_flag lds_once;  // the split launches reserve more than the default 64 KiB
    static hipError_t lds_err = hipSuccess;
    std::call_once(lds_once, [] {
      lds_err = hipFuncSetAttribute((const void *)g1_group_dit_mul_kernel<C>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)g1_split_lds<C>());
      if (lds_err == hipSuccess)
        lds_err = hipFuncSetAttribute((const void *)g1_group_dif_mul_kernel<C>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)g1_split_lds<C>());
      if (lds_err == hipSuccess)
        lds_err = hipFuncSetAttribute((const void *)g1_fft_stage_kernel<C, false>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)g1_split_lds<C>());
      if (lds_err == hipSuccess)
        lds_err = hipFuncSetAttribute((const void *)g1_fft_stage_kernel<C, true>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)g1_split_lds<C>());
    });
    ZK_HIP(lds_err);
  }
  // threads of a group launch split in S parts (prod_k: slot K is a product, in the group that takes N^-1)
  auto group_threads = [&](int r, int S, bool prod_k) {
    const uint64_t ng = N >> r, K = g1_kcount(r);
    return ng * (K + 1) + (uint64_t)(S - 1) * ng * (K + (prod_k ? 1 : 0));
  };
  auto group_parts = [&](int r, bool prod_k) {
    for (int S = 3; S > 1; S--)
      if (group_threads(r, S, prod_k) <= split_cap) return S;
    return 1;
  };
  {
    const uint64_t simds = (uint64_t)ctx->num_cus * 4;
    // chains on the busiest SIMD for `threads` threads, co-resident wavefronts `occ` a SIMD
    auto launch_cost = [&](uint64_t threads, double chain, int occ) {
      const uint64_t w = ((threads + 63) / 64 + simds - 1) / simds;
      return chain * (occ >= 2 ? 1.75 * (double)(w / 2) + (double)(w % 2) : (double)w);
    };
    auto group_cost = [&](int r) {
      if (r == 1) return launch_cost(halfN, 1.0, C::id == 0 ? 2 : 1);
      const int S = group_parts(r, false);
      return launch_cost(group_threads(r, S, false), S == 3 ? 0.67 : (S == 2 ? 0.73 : 1.0), C::id == 0 ? 2 : 1) + 0.08;
    };
    g1_stage_plan(m, group_cost, plan);
  }
  uint64_t scr_pts = 0;
  bool grouped = false;
  for (int r : plan)
    if (r > 1) {
      grouped = true;
      for (bool pk : {false, true})
        scr_pts = std::max<uint64_t>(scr_pts, group_threads(r, group_parts(r, pk), pk));
    }
  // g1_tab_words per thread of every launch that multiplies (BN254: 15 cached entries, 2880
  // bytes, 127 MB at 2^12 with split products; the radix-2 stages and the scaling pass stride
  // over their points with at most one round of threads, so the table stays below ~0.57 GB;
  // BLS12-381: 3 entries, 840 bytes)
  const uint32_t tab_threads = (uint32_t)std::max<uint64_t>(scr_pts, std::min<uint64_t>(N, round_threads));
  const uint32_t stage_threads = std::min<uint32_t>(halfN, tab_threads), scale_threads = std::min<uint32_t>(N, tab_threads);
  const bool stage_sp = stage_threads <= split_cap;  // one wavefront a SIMD: the split launch shape
  ZK_TRY(tab.alloc((size_t)tab_threads * g1_tab_words<C>() * 4));
  if (grouped) {
    ZK_TRY(twm.alloc((size_t)(halfN ? halfN : 1) * 32));
    ZK_TRY(scr.alloc(scr_pts * ZW * 4));
  }
  auto to_fr = [&](const zkh::Fe &x) {
    Fe<FR> d;
    for (int j = 0; j < 8; j++) d.w[j] = (uint32_t)(x.w[j / 2] >> (32 * (j & 1)));
    return d;
  };
  const dim3 blk(kG1Threads), gN((N + kG1Threads - 1) / kG1Threads), gH((halfN + kG1Threads - 1) / kG1Threads);
  // Montgomery form of g for the path products (to_fr gives the words of the Montgomery value)
  if (halfN)
    hipLaunchKernelGGL(g1_twiddle_kernel<C>, gH, blk, 0, st, tw.as<uint32_t>(),
                       grouped ? twm.as<uint32_t>() : (uint32_t *)nullptr, to_fr(g), halfN);
  if (jac)
    hipLaunchKernelGGL((g1_proj_in_kernel<C, true>), gN, blk, 0, st, (const uint32_t *)d_src, N, m, dit ? 1 : 0,
                       pts.as<int32_t>());
  else
    hipLaunchKernelGGL((g1_proj_in_kernel<C, false>), gN, blk, 0, st, (const uint32_t *)d_src, N, m, dit ? 1 : 0,
                       pts.as<int32_t>());
  auto grid = [](uint64_t threads) { return dim3((unsigned)((threads + kG1Threads - 1) / kG1Threads)); };
  auto grid_n = [](uint64_t threads, int per) { return dim3((unsigned)((threads + per - 1) / per)); };
  const zkh::Fe ninv_m = zkh::inv(fr, zkh::from_u64(fr, N));
  if (dit) {
    // an inverse run folds N^-1 into the first group when it is a grouped stage, else a scaling
    // pass follows
    int s0 = 0;
    bool folded = false;
    for (int r : plan) {
      if (r == 1) {
        hipLaunchKernelGGL((g1_fft_stage_kernel<C, false>), stage_sp ? grid_n(stage_threads, g1_split_threads<C>())
                                                                     : grid(stage_threads),
                           stage_sp ? dim3(g1_split_threads<C>()) : blk, stage_sp ? g1_split_lds<C>() : 0, st,
                           pts.as<int32_t>(), tw.as<uint32_t>(), m, s0 + 1, tab.as<uint32_t>(), stage_threads);
      } else {
        const uint32_t ng = N >> r;
        const int fin = inverse && s0 == 0 ? 1 : 0;
        folded = folded || fin;
        const int S = group_parts(r, fin != 0);
        const uint64_t nt = group_threads(r, S, fin != 0);
        const bool sp = nt <= split_cap;  // one wavefront a SIMD: the split launch shape
        hipLaunchKernelGGL(g1_group_dit_mul_kernel<C>, sp ? grid_n(nt, g1_split_threads<C>()) : grid(nt),
                           sp ? dim3(g1_split_threads<C>()) : blk, sp ? g1_split_lds<C>() : 0, st,
                           (const int32_t *)pts.as<int32_t>(), scr.as<int32_t>(), twm.as<uint32_t>(), m, s0, r, ng,
                           fin, to_fr(ninv_m), tab.as<uint32_t>(), tab_threads, S);
        hipLaunchKernelGGL(g1_group_dit_sum_kernel<C>, grid(((uint64_t)ng << r) << std::min(r, 3)), blk, 0, st,
                           pts.as<int32_t>(),
                           (const int32_t *)scr.as<int32_t>(), s0, r, ng, S, fin);
      }
      s0 += r;
    }
    if (inverse && !folded) {
      const zkh::Fe ninv = zkh::to_std(fr, ninv_m);
      hipLaunchKernelGGL(g1_scale_kernel<C>, grid(scale_threads), blk, 0, st, pts.as<int32_t>(), N, to_fr(ninv),
                         tab.as<uint32_t>(), scale_threads);
    }
  } else if (m > 0) {
    // N^-1 folds into the last group when it is a grouped stage, else a scaling pass follows
    int s0 = m;
    bool folded = false;
    for (size_t i = plan.size(); i-- > 0;) {
      const int r = plan[i];
      s0 -= r;
      if (r == 1) {
        hipLaunchKernelGGL((g1_fft_stage_kernel<C, true>), stage_sp ? grid_n(stage_threads, g1_split_threads<C>())
                                                                    : grid(stage_threads),
                           stage_sp ? dim3(g1_split_threads<C>()) : blk, stage_sp ? g1_split_lds<C>() : 0, st,
                           pts.as<int32_t>(), tw.as<uint32_t>(), m, s0 + 1, tab.as<uint32_t>(), stage_threads);
      } else {
        const uint32_t ng = N >> r;
        const int fin = s0 == 0 ? 1 : 0;
        folded = folded || fin;
        const int S = group_parts(r, fin != 0);
        const uint64_t nt = group_threads(r, S, fin != 0);
        const bool sp = nt <= split_cap;  // one wavefront a SIMD: the split launch shape
        hipLaunchKernelGGL(g1_group_dif_mul_kernel<C>, sp ? grid_n(nt, g1_split_threads<C>()) : grid(nt),
                           sp ? dim3(g1_split_threads<C>()) : blk, sp ? g1_split_lds<C>() : 0, st,
                           (const int32_t *)pts.as<int32_t>(), scr.as<int32_t>(), twm.as<uint32_t>(), m, s0, r, ng,
                           fin, to_fr(ninv_m), tab.as<uint32_t>(), tab_threads, S);
        hipLaunchKernelGGL(g1_group_dif_sum_kernel<C>, grid(((uint64_t)ng << r) * 8), blk, 0, st, pts.as<int32_t>(),
                           (const int32_t *)scr.as<int32_t>(), s0, r, ng, S, fin);
      }
    }
    if (!folded) {
      const zkh::Fe ninv = zkh::to_std(fr, ninv_m);
      hipLaunchKernelGGL(g1_scale_kernel<C>, grid(scale_threads), blk, 0, st, pts.as<int32_t>(), N, to_fr(ninv),
                         tab.as<uint32_t>(), scale_threads);
    }
  }
  hipLaunchKernelGGL(g1_proj_out_kernel<C>, gN, blk, 0, st, pts.as<int32_t>(), N, m, dit ? 0 : 1,
                     (uint32_t *)d_dst);
  ZK_HIP(hipGetLastError());
  ZK_HIP(hipStreamSynchronize(st));  // pts / tw / twm / scr / tab are released on return
  drain.release();
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
