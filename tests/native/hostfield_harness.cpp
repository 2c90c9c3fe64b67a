// Host build of zikkurat-algebra_amd/csrc/hostfield.hpp (the serial host-side field and XYZZ
// arithmetic: MSM window combination, NTT seeds, generator checks), exposed to the CPU tests
// through a C ABI.  Test infrastructure only.
#include <string.h>

#include "../../zikkurat-algebra_amd/csrc/hostfield.hpp"

using namespace zkh;

static const FieldDesc &desc(int field) {
  switch (field) {
    case 0: return BN254_FP;
    case 1: return BN254_FR;
    case 2: return BLS381_FP;
    default: return BLS381_FR;
  }
}

// field elements of nl 64-bit words each (Montgomery form where the operation expects it)
extern "C" void hfh_field_op(int field, int op, int n, const uint64_t *a, const uint64_t *b, uint64_t *out) {
  const FieldDesc &f = desc(field);
  for (int k = 0; k < n; k++) {
    const Fe x = load(f, a + (size_t)k * f.nl);
    const Fe y = b ? load(f, b + (size_t)k * f.nl) : zero();
    Fe r;
    switch (op) {
      case 0: r = mul(f, x, y); break;
      case 1: r = add(f, x, y); break;
      case 2: r = sub(f, x, y); break;
      case 3: r = neg(f, x); break;
      case 4: r = inv(f, x); break;
      case 5: r = to_std(f, x); break;
      case 6: r = from_std(f, x); break;
      case 7: r = sqr(f, x); break;
      default: r = x; break;
    }
    store(f, out + (size_t)k * f.nl, r);
  }
}

// G1 points in XYZZ (4 nl words each, Montgomery); op 0: dbl(a), 1: add(a, b); op 2: affine
// output of a (2 nl words, all 0xFF for infinity)
extern "C" void hfh_g1_op(int curve, int op, int n, const uint64_t *a, const uint64_t *b, uint64_t *out) {
  const FieldDesc &f = curve == 0 ? BN254_FP : BLS381_FP;
  const int L = f.nl;
  auto get = [&](const uint64_t *p) {
    Xyzz q;
    q.x = load(f, p);
    q.y = load(f, p + L);
    q.zz = load(f, p + 2 * L);
    q.zzz = load(f, p + 3 * L);
    return q;
  };
  for (int k = 0; k < n; k++) {
    const Xyzz p = get(a + (size_t)k * 4 * L);
    if (op == 2) {
      xyzz_to_affine(f, p, out + (size_t)k * 2 * L);
      continue;
    }
    const Xyzz r = op == 0 ? xyzz_dbl(f, p) : xyzz_add(f, p, get(b + (size_t)k * 4 * L));
    uint64_t *o = out + (size_t)k * 4 * L;
    store(f, o, r.x);
    store(f, o + L, r.y);
    store(f, o + 2 * L, r.zz);
    store(f, o + 3 * L, r.zzz);
  }
}
