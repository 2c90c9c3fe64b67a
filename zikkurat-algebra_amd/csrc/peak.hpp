#pragma once
#include <stddef.h>

namespace zk {

enum Microbench : int {
  MB_IMUL = 0,
  MB_COPY = 1,
  MB_FMUL_BN254_FP = 2,
  MB_FMUL_BN254_FR = 3,
  MB_FMUL_BLS381_FP = 4,
  MB_FMUL_BLS381_FR = 5,
};

enum FieldOp : int {
  FOP_ADD = 0,
  FOP_SUB = 1,
  FOP_MUL = 2,
  FOP_SQR = 3,
  FOP_NEG = 4,
  FOP_INV = 5,
  FOP_HALF = 6,
  FOP_TO_STD = 7,
  FOP_FROM_STD = 8,
};

int microbench(int which, double *result);
int fill_random_device(int curve, int field, size_t n, uint64_t seed, void *out);
int field_op_device(int curve, int field, int op, size_t n, const void *a, const void *b, void *out);

}  // namespace zk
