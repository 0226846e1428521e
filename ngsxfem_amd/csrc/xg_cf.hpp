// xg_cf.hpp -- coefficient functions evaluated on the device.
//
// NGSolve evaluates a CoefficientFunction at the mapped integration points
// of a rule (symboliccutbfi.cpp:229-287 multiplies the proxies by the
// coefficient's values there; python_cutint.cpp:66-98 sums cf * weight).
// Here a coefficient is a small postfix program over the physical point x
// and the element's P1 grid functions, built on the host from the
// expression tree and interpreted per point: every lane of a wave runs the
// same program, so the dispatch on the opcode is wave-uniform.
//
// The header is shared with host builds (tests compile it with g++ and
// define XG_CF_FN themselves).
#pragma once
#include <math.h>
#include <stdint.h>

#ifndef XG_CF_FN
#include <hip/hip_runtime.h>
#define XG_CF_FN __device__ __forceinline__
#define XG_CF_HD __host__ __device__ inline
#endif
#ifndef XG_CF_HD
#define XG_CF_HD inline
#endif

namespace xg {
namespace cf {

// opcodes (the values are part of the C ABI, include/xg_abi.h)
enum Op : int32_t {
  END = 0,
  CONST = 1,   // + index into the constants: push c[k]
  COORD = 2,   // + coordinate d: push x_d (physical)
  ADD = 3,     // a b -> a + b
  SUB = 4,     // a b -> a - b
  MUL = 5,     // a b -> a * b
  DIV = 6,     // a b -> a / b
  NEG = 7,     // a -> -a
  POWI = 8,    // + integer exponent n: a -> a^n (n < 0: 1 / a^-n)
  POW = 9,     // a b -> pow(a, b)
  SQRT = 10,
  SIN = 11,
  COS = 12,
  TAN = 13,
  EXP = 14,
  LOG = 15,
  ATAN = 16,
  TANH = 17,
  ABS = 18,
  GF1 = 19,    // + grid function slot s: push sum_k lam_k u_s[v_k] (P1 nodal values)
  IFPOS = 20,  // a b c -> (a > 0 ? b : c)  (NGSolve IfPos)
  MAX = 21,    // a b -> max(a, b)
  MIN = 22,    // a b -> min(a, b)
  EL0 = 23,    // + grid function slot s: push u_s[e] (piecewise constant, one value per element)
  NOPS = 24
};

constexpr int MAXCODE = 192;  // program words
constexpr int MAXCONST = 48;  // constants
constexpr int MAXGF = 4;      // P1 grid functions
constexpr int MAXSTACK = 16;  // evaluation stack depth

// a validated program, passed to the kernels by value (kernel arguments are
// read through the scalar cache: wave-uniform)
struct Prog {
  int32_t n;  // words of code
  int32_t code[MAXCODE];
  double c[MAXCONST];
  const double *gf[MAXGF];
};

// words taken by an opcode (itself and its immediate operand)
XG_CF_HD int op_words(int32_t op) {
  return (op == CONST || op == COORD || op == POWI || op == GF1 || op == EL0) ? 2 : 1;
}

// stack effect of an opcode
XG_CF_HD int op_delta(int32_t op) {
  switch (op) {
    case CONST: case COORD: case GF1: case EL0: return 1;
    case ADD: case SUB: case MUL: case DIV: case POW: case MAX: case MIN: return -1;
    case IFPOS: return -2;
    default: return 0;
  }
}

// host: checks a program (known opcodes, operand ranges, stack depth within
// MAXSTACK, exactly one value left) and copies it into p.  Returns 0 or a
// message.
inline const char *validate(int32_t ncode, const int32_t *code, int32_t nconst, const double *consts, int32_t ngf,
                            const double *const *gf, Prog &p) {
  if (ncode <= 0 || ncode > MAXCODE || !code) return "coefficient program: 1 .. 192 code words";
  if (nconst < 0 || nconst > MAXCONST || (nconst > 0 && !consts)) return "coefficient program: at most 48 constants";
  if (ngf < 0 || ngf > MAXGF || (ngf > 0 && !gf)) return "coefficient program: at most 4 grid functions";
  int depth = 0;
  for (int pc = 0; pc < ncode;) {
    const int32_t op = code[pc];
    if (op <= END || op >= NOPS) return "coefficient program: unknown opcode";
    const int w = op_words(op);
    if (pc + w > ncode) return "coefficient program: missing operand";
    const int32_t a = w == 2 ? code[pc + 1] : 0;
    if (op == CONST && (a < 0 || a >= nconst)) return "coefficient program: constant index out of range";
    if (op == COORD && (a < 0 || a > 2)) return "coefficient program: coordinate out of range";
    if ((op == GF1 || op == EL0) && (a < 0 || a >= ngf || !gf[a]))
      return "coefficient program: grid function slot out of range";
    const int need = (op == IFPOS) ? 3 : (op_delta(op) == -1 ? 2 : (op_delta(op) == 0 ? 1 : 0));
    if (depth < need) return "coefficient program: stack underflow";
    depth += op_delta(op);
    if (depth > MAXSTACK) return "coefficient program: stack deeper than 16";
    pc += w;
  }
  if (depth != 1) return "coefficient program: must leave exactly one value";
  p.n = ncode;
  for (int k = 0; k < ncode; ++k) p.code[k] = code[k];
  for (int k = ncode; k < MAXCODE; ++k) p.code[k] = END;
  for (int k = 0; k < MAXCONST; ++k) p.c[k] = k < nconst ? consts[k] : 0.0;
  for (int k = 0; k < MAXGF; ++k) p.gf[k] = k < ngf ? gf[k] : nullptr;
  return nullptr;
}

// true when the program reads piecewise constant grid functions (EL0)
inline bool uses_elements(const Prog &p) {
  for (int pc = 0; pc < p.n; pc += op_words(p.code[pc]))
    if (p.code[pc] == EL0) return true;
  return false;
}

// value of the program at the physical point x of element e with vertices
// vv (D + 1 of them) and barycentrics lam (lam_k of vertex k)
template <int D>
XG_CF_FN double eval(const Prog &P, const double *x, const double *lam, const int32_t *vv, int e) {
  double st[MAXSTACK];
  int sp = 0;
  for (int pc = 0; pc < P.n;) {
    const int32_t op = P.code[pc];
    const int32_t a = P.code[pc + 1 < MAXCODE ? pc + 1 : pc];
    pc += op_words(op);
    switch (op) {
      case CONST: st[sp++] = P.c[a]; break;
      case COORD: st[sp++] = x[a]; break;
      case GF1: {
        const double *u = P.gf[a];
        double s = 0.0;
        for (int k = 0; k <= D; ++k) s += lam[k] * u[vv[k]];
        st[sp++] = s;
        break;
      }
      case EL0: st[sp++] = P.gf[a][e]; break;
      case ADD: --sp; st[sp - 1] = st[sp - 1] + st[sp]; break;
      case SUB: --sp; st[sp - 1] = st[sp - 1] - st[sp]; break;
      case MUL: --sp; st[sp - 1] = st[sp - 1] * st[sp]; break;
      case DIV: --sp; st[sp - 1] = st[sp - 1] / st[sp]; break;
      case POW: --sp; st[sp - 1] = pow(st[sp - 1], st[sp]); break;
      case MAX: --sp; st[sp - 1] = fmax(st[sp - 1], st[sp]); break;
      case MIN: --sp; st[sp - 1] = fmin(st[sp - 1], st[sp]); break;
      case IFPOS: sp -= 2; st[sp - 1] = st[sp - 1] > 0.0 ? st[sp] : st[sp + 1]; break;
      case NEG: st[sp - 1] = -st[sp - 1]; break;
      case POWI: {
        const double b = st[sp - 1];
        int n = a < 0 ? -a : a;
        double r = 1.0, q = b;
        while (n) {  // square and multiply
          if (n & 1) r *= q;
          q *= q;
          n >>= 1;
        }
        st[sp - 1] = a < 0 ? 1.0 / r : r;
        break;
      }
      case SQRT: st[sp - 1] = sqrt(st[sp - 1]); break;
      case SIN: st[sp - 1] = sin(st[sp - 1]); break;
      case COS: st[sp - 1] = cos(st[sp - 1]); break;
      case TAN: st[sp - 1] = tan(st[sp - 1]); break;
      case EXP: st[sp - 1] = exp(st[sp - 1]); break;
      case LOG: st[sp - 1] = log(st[sp - 1]); break;
      case ATAN: st[sp - 1] = atan(st[sp - 1]); break;
      case TANH: st[sp - 1] = tanh(st[sp - 1]); break;
      case ABS: st[sp - 1] = fabs(st[sp - 1]); break;
      default: break;
    }
  }
  return st[0];
}

}  // namespace cf
}  // namespace xg
