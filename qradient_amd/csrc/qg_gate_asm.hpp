// One Pauli rotation on one register bit of a thread's 16 amplitudes, as a
// single inline-asm block.
//
// U = cos(phi) I - i sin(phi) P acts on the 8 amplitude pairs (x0, x1) that
// differ in the bit.  Each pair is two real planes rotated in place by three
// shears (a = tan(phi/2), b = sin(phi); see shear3 in qg_pass.hpp):
//   X: (re0, im1) by -phi, (im0, re1) by +phi
//   Y: (re0, re1) by +phi, (im0, im1) by +phi
//   Z: (re0, im0) by -phi, (re1, im1) by +phi
// The gate type is a uniform (scalar) branch inside the block, so every type
// updates the same registers: no register copies at a join, which the
// compiler otherwise inserts after a C++ branch over the type.  Operands
// %0..%31 are the pairs in order (pair p: re0 = 4p, im0 = 4p+1, re1 = 4p+2,
// im1 = 4p+3), %32 the type, %33 = a, %34 = b.  Within each shear step the
// eight pairs' FMAs are independent, so consecutive instructions do not wait
// on each other.
#pragma once

#define QG_GF(op, d, s, x) op " %" #d ", " s ", %" #x ", %" #d "\n\t"
#define QG_PAIRS(M, op) \
  M(op, 0, 1, 2, 3) M(op, 4, 5, 6, 7) M(op, 8, 9, 10, 11) M(op, 12, 13, 14, 15) M(op, 16, 17, 18, 19) \
  M(op, 20, 21, 22, 23) M(op, 24, 25, 26, 27) M(op, 28, 29, 30, 31)

// X: re0 += a im1, im0 -= a re1 | im1 -= b re0, re1 += b im0 | first step again
#define QG_X1(op, r0, i0, r1, i1) QG_GF(op, r0, "%33", i1) QG_GF(op, i0, "-%33", r1)
#define QG_X2(op, r0, i0, r1, i1) QG_GF(op, i1, "-%34", r0) QG_GF(op, r1, "%34", i0)
// Y: re0 -= a re1, im0 -= a im1 | re1 += b re0, im1 += b im0 | first step again
#define QG_Y1(op, r0, i0, r1, i1) QG_GF(op, r0, "-%33", r1) QG_GF(op, i0, "-%33", i1)
#define QG_Y2(op, r0, i0, r1, i1) QG_GF(op, r1, "%34", r0) QG_GF(op, i1, "%34", i0)
// Z: re0 += a im0, re1 -= a im1 | im0 -= b re0, im1 += b re1 | first step again
#define QG_Z1(op, r0, i0, r1, i1) QG_GF(op, r0, "%33", i0) QG_GF(op, r1, "-%33", i1)
#define QG_Z2(op, r0, i0, r1, i1) QG_GF(op, i0, "-%34", r0) QG_GF(op, i1, "%34", r1)

#define QG_GATE_ASM(op)                                                                         \
  "s_cmp_eq_u32 %32, 1\n\t"                                                                     \
  "s_cbranch_scc1 1f\n\t"                                                                       \
  "s_cmp_eq_u32 %32, 2\n\t"                                                                     \
  "s_cbranch_scc1 2f\n\t"                                                                       \
  "s_cmp_eq_u32 %32, 3\n\t"                                                                     \
  "s_cbranch_scc1 3f\n\t"                                                                       \
  "s_branch 4f\n"                                                                               \
  "1:\n\t" QG_PAIRS(QG_X1, op) QG_PAIRS(QG_X2, op) QG_PAIRS(QG_X1, op) "s_branch 4f\n"          \
  "2:\n\t" QG_PAIRS(QG_Y1, op) QG_PAIRS(QG_Y2, op) QG_PAIRS(QG_Y1, op) "s_branch 4f\n"          \
  "3:\n\t" QG_PAIRS(QG_Z1, op) QG_PAIRS(QG_Z2, op) QG_PAIRS(QG_Z1, op) "\n"                     \
  "4:\n\t"

#define QG_GATE_OPERANDS(x, M)                                                                    \
  "+v"(x[M[0]].re), "+v"(x[M[0]].im), "+v"(x[M[1]].re), "+v"(x[M[1]].im), "+v"(x[M[2]].re),        \
      "+v"(x[M[2]].im), "+v"(x[M[3]].re), "+v"(x[M[3]].im), "+v"(x[M[4]].re), "+v"(x[M[4]].im),     \
      "+v"(x[M[5]].re), "+v"(x[M[5]].im), "+v"(x[M[6]].re), "+v"(x[M[6]].im), "+v"(x[M[7]].re),     \
      "+v"(x[M[7]].im), "+v"(x[M[8]].re), "+v"(x[M[8]].im), "+v"(x[M[9]].re), "+v"(x[M[9]].im),     \
      "+v"(x[M[10]].re), "+v"(x[M[10]].im), "+v"(x[M[11]].re), "+v"(x[M[11]].im), "+v"(x[M[12]].re), \
      "+v"(x[M[12]].im), "+v"(x[M[13]].re), "+v"(x[M[13]].im), "+v"(x[M[14]].re), "+v"(x[M[14]].im), \
      "+v"(x[M[15]].re), "+v"(x[M[15]].im)
