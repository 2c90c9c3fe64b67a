// Host build of the device field/curve code in zikkurat-algebra_amd/csrc/fz.hpp and
// curvez.hpp (the same __host__ __device__ functions the GPU kernels inline), exposed to the
// CPU tests through a C ABI.  Test infrastructure only.
#include <string.h>

#include "../../zikkurat-algebra_amd/csrc/curvez.hpp"
#include "../../zikkurat-algebra_amd/csrc/glv.hpp"

using namespace zk;

template <class P>
static void run_op(int op, int n, const int32_t *a, const int32_t *b, int32_t *out) {
  for (int k = 0; k < n; k++) {
    Fz<P> x, y, r;
    memcpy(x.l, a + k * P::N, sizeof x.l);
    if (b) memcpy(y.l, b + k * P::N, sizeof y.l);
    switch (op) {
      case 0: zmul(r, x, y); break;
      case 1: zadd(r, x, y); break;
      case 2: zsub(r, x, y); break;
      case 3: r = x; znorm(r); break;
      case 4: zcanon(r, x); break;
      case 5: zset_zero(r); r.l[0] = zis_zero_mod(x) ? 1 : 0; break;
      case 6: zsqr(r, x); break;
      case 7: zmul_sub(r, x, y, y, x); break;  // x y - y x = 0 exercises the signs; see also op 8
      case 8: {  // x y - (x + 1) (y - 1) with normalised operands built from x and y
        Fz<P> x1 = x, y1 = y;
        x1.l[0] += 1;
        y1.l[0] -= 1;
        znorm(x1);
        znorm(y1);
        zmul_sub(r, x, y, x1, y1);
        break;
      }
      case 9: zinv(r, x); break;
      case 12: zinv_fermat(r, x); break;
      case 10: zmul_sub(r, x, y, y, y); break;  // x y - y y, both operands possibly unnormalised
      case 11: zreduce_small(r, x); break;
      default: r = x; break;
    }
    memcpy(out + k * P::N, r.l, sizeof r.l);
  }
}

template <class P>
static void run_conv(int dir, int n, const uint32_t *win, int32_t *lout, const int32_t *lin, uint32_t *wout) {
  for (int k = 0; k < n; k++) {
    Fz<P> r;
    if (dir == 0) {  // external words -> internal limbs
      zfrom_ext(r, win + k * P::EXT_WORDS);
      memcpy(lout + k * P::N, r.l, sizeof r.l);
    } else if (dir == 1) {  // internal limbs -> external words
      memcpy(r.l, lin + k * P::N, sizeof r.l);
      zto_ext(wout + k * P::EXT_WORDS, r);
    } else if (dir == 2) {  // unpack only
      zunpack(r, win + k * P::EXT_WORDS);
      memcpy(lout + k * P::N, r.l, sizeof r.l);
    } else {  // pack only
      memcpy(r.l, lin + k * P::N, sizeof r.l);
      zpack(wout + k * P::EXT_WORDS, r);
    }
  }
}

template <class C>
static void run_curve(int op, int n, const int32_t *a, const int32_t *b, int32_t *out) {
  constexpr int N = ZP<C>::T::N;
  for (int k = 0; k < n; k++) {
    XyzzZ<C> p, q, r;
    memcpy(&p, a + k * 4 * N, sizeof p);
    if (op == 0) {  // madd: b holds affine points
      AffZ<C> aff;
      memcpy(&aff, b + k * 2 * N, sizeof aff);
      r = p;
      zxyzz_madd(r, aff);
    } else if (op == 1) {
      memcpy(&q, b + k * 4 * N, sizeof q);
      zxyzz_add(r, p, q);
    } else if (op == 3) {  // scalar multiplication: b holds 8-word standard-form scalars
      uint32_t sc[8];
      memcpy(sc, b + k * 8, sizeof sc);
      zxyzz_scalar_mul(r, p, sc);
    } else if (op == 4 || op == 5) {  // BN254 GLV forms (scalars below r): 1-bit joint, 2-bit joint windows
      uint32_t sc[8];
      memcpy(sc, b + k * 8, sizeof sc);
      if constexpr (C::id == 0) {
        if (op == 4) {
          zxyzz_scalar_mul_glv(r, p, sc);
        } else {
          alignas(16) uint32_t tab[kGlvW2Entries * ZLayout<C>::JACC];
          zxyzz_scalar_mul_glv_w2(r, p, sc, tab);
        }
      } else {
        zxyzz_set_inf(r);
      }
    } else if (op == 6) {  // 2-bit windows on k over a 3-entry host table
      uint32_t sc[8];
      memcpy(sc, b + k * 8, sizeof sc);
      uint32_t tab[3 * 5 * ZLayout<C>::N];
      zxyzz_scalar_mul_w2(r, p, sc, tab, 0, 1);
    } else if (op == 7 || op == 8) {  // the windows above and below the G1 FFT split point, apart, added
      uint32_t sc[8];
      memcpy(sc, b + k * 8, sizeof sc);
      XyzzZ<C> hi, lo;
      if (op == 7) {
        if constexpr (C::id == 0) {
          alignas(16) uint32_t tab[kGlvW2Entries * ZLayout<C>::JACC];
          zxyzz_scalar_mul_glv_w2(hi, p, sc, tab, 40, 64);
          zxyzz_scalar_mul_glv_w2(lo, p, sc, tab, 0, 40);
        } else {
          zxyzz_set_inf(hi);
          zxyzz_set_inf(lo);
        }
      } else {
        uint32_t tab[3 * 5 * ZLayout<C>::N];
        zxyzz_scalar_mul_w2(hi, p, sc, tab, 0, 1, 80, 128);
        zxyzz_scalar_mul_w2(lo, p, sc, tab, 0, 1, 0, 80);
      }
      zxyzz_add(r, hi, lo);
    } else if (op == 12 || op == 13) {  // the three window parts of the G1 FFT's 3-way split, apart, added
      uint32_t sc[8];
      memcpy(sc, b + k * 8, sizeof sc);
      XyzzZ<C> hi, mid, lo, t;
      if (op == 12) {
        if constexpr (C::id == 0) {
          alignas(16) uint32_t tab[kGlvW2Entries * ZLayout<C>::JACC];
          zxyzz_scalar_mul_glv_w2(hi, p, sc, tab, 52, 64);
          zxyzz_scalar_mul_glv_w2(mid, p, sc, tab, 33, 52);
          zxyzz_scalar_mul_glv_w2(lo, p, sc, tab, 0, 33);
        } else {
          zxyzz_set_inf(hi);
          zxyzz_set_inf(mid);
          zxyzz_set_inf(lo);
        }
      } else {
        uint32_t tab[3 * 5 * ZLayout<C>::N];
        zxyzz_scalar_mul_w2(hi, p, sc, tab, 0, 1, 104, 128);
        zxyzz_scalar_mul_w2(mid, p, sc, tab, 0, 1, 65, 104);
        zxyzz_scalar_mul_w2(lo, p, sc, tab, 0, 1, 0, 65);
      }
      zxyzz_add(t, hi, mid);
      zxyzz_add(r, t, lo);
    } else if (op == 9 || op == 10 || op == 11) {  // Jacobian steps, XYZZ in and out: p + q (q cached), 2p, 4p + q
      JacZ<C> jp, jq;
      JacCZ<C> cq;
      zjac_from_xyzz(jp, p);
      if (op != 10) {
        memcpy(&q, b + k * 4 * N, sizeof q);
        zjac_from_xyzz(jq, q);
        zjac_cache(cq, jq);
      }
      if (op == 9) {
        zjac_add(jp, jp, cq);
      } else if (op == 10) {
        zjac_dbl(jp, jp);
      } else {
        zjac_dbl(jp, jp);
        zjac_dbl(jp, jp);
        zjac_add(jp, jp, cq);
      }
      zjac_to_xyzz(r, jp);
    } else {
      zxyzz_dbl(r, p);
    }
    memcpy(out + k * 4 * N, &r, sizeof r);
  }
}

extern "C" {
// field: 0 BN254 Fp, 1 BN254 Fr, 2 BLS Fp, 3 BLS Fr
int fzh_limbs(int field) {
  const int n[4] = {ZBn254Fp::N, ZBn254Fr::N, ZBls381Fp::N, ZBls381Fr::N};
  return n[field];
}
void fzh_op(int field, int op, int n, const int32_t *a, const int32_t *b, int32_t *out) {
  switch (field) {
    case 0: run_op<ZBn254Fp>(op, n, a, b, out); break;
    case 1: run_op<ZBn254Fr>(op, n, a, b, out); break;
    case 2: run_op<ZBls381Fp>(op, n, a, b, out); break;
    default: run_op<ZBls381Fr>(op, n, a, b, out); break;
  }
}
void fzh_conv(int field, int dir, int n, const uint32_t *win, int32_t *lout, const int32_t *lin, uint32_t *wout) {
  switch (field) {
    case 0: run_conv<ZBn254Fp>(dir, n, win, lout, lin, wout); break;
    case 1: run_conv<ZBn254Fr>(dir, n, win, lout, lin, wout); break;
    case 2: run_conv<ZBls381Fp>(dir, n, win, lout, lin, wout); break;
    default: run_conv<ZBls381Fr>(dir, n, win, lout, lin, wout); break;
  }
}
void fzh_curve(int curve, int op, int n, const int32_t *a, const int32_t *b, int32_t *out) {
  if (curve == 0) run_curve<Bn254G1>(op, n, a, b, out);
  else run_curve<Bls381G1>(op, n, a, b, out);
}
}
