// project.hip — per-row expression programs (bolt_project).
//
// Replaces the projection kernel of Engine::execute_projection
// (src/exec/engine.rs:4267), which the reference lowers from a KernelSpec's
// Op list (src/plan/physical_plan.rs:81) to PTX (src/jit/ptx_gen.rs).  Here
// the program is type-checked on the host, printed as one HIP kernel (one C
// statement per op, registers as typed locals, Decimal128 as __int128), and
// compiled with hipRTC; the code object is cached by the program's text, so a
// repeated query shape compiles once.  The kernel is a plain grid-stride pass:
// every input column is read once and every output written once, so it runs at
// the HBM bound of its columns.
#include <cstdint>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "bolt_internal.h"
#include "jit.h"
#include "reduce_common.h"

namespace bolt {
namespace {

struct RegInfo {
  int dtype = 0;       // 0 = not yet written
  uint32_t deps = 0;   // input columns whose values reach the register (bit per input)
  int unit = 0;        // Timestamp: the unit as digits of a second (0, 3, 6, 9)
};

const char* ctype(int dt) {
  switch (dt) {
    case BOLT_BOOL: case BOLT_I32: case BOLT_DATE32: return "int";
    case BOLT_I64: case BOLT_TIMESTAMP: return "long long";
    case BOLT_F32: return "float";
    case BOLT_F64: return "double";
    case BOLT_DEC128: return "__int128";
  }
  return "?";
}

// Storage type of a column element.
const char* stype(int dt) {
  switch (dt) {
    case BOLT_BOOL: return "unsigned char";
    case BOLT_I32: case BOLT_DATE32: return "int";
    case BOLT_I64: case BOLT_TIMESTAMP: return "long long";
    case BOLT_F32: return "float";
    case BOLT_F64: return "double";
  }
  return "?";
}

bool is_int(int dt) { return dt == BOLT_I32 || dt == BOLT_I64 || dt == BOLT_DATE32; }
bool is_float(int dt) { return dt == BOLT_F32 || dt == BOLT_F64; }
bool supported(int dt) {
  return dt == BOLT_BOOL || dt == BOLT_I32 || dt == BOLT_I64 || dt == BOLT_F32 || dt == BOLT_F64 ||
         dt == BOLT_DATE32 || dt == BOLT_DEC128 || dt == BOLT_TIMESTAMP;
}
bool valid_unit(int u) { return u == 0 || u == 3 || u == 6 || u == 9; }
constexpr int kMaxFloatScale = 22;  // 10^s exact in f64
bool is_cmp(int op) { return op >= BOLT_BIN_EQ && op <= BOLT_BIN_GE; }

// Functions the generated kernels call (guarded integer division and
// remainder, saturating float -> int conversion, ordered comparisons).
const char* kProjectPrelude = R"BOLTP(
#pragma clang fp contract(off)
typedef unsigned long long u64;
__device__ __forceinline__ int bdiv32(int a, int b) { return b == 0 ? 0 : (a == (-2147483647 - 1) && b == -1) ? a : a / b; }
__device__ __forceinline__ long long bdiv64(long long a, long long b) {
  return b == 0 ? 0 : (a == (-9223372036854775807LL - 1) && b == -1) ? a : a / b;
}
__device__ __forceinline__ int bmod32(int a, int b) { return (b == 0 || b == -1) ? 0 : a % b; }
__device__ __forceinline__ long long bmod64(long long a, long long b) { return (b == 0 || b == -1) ? 0 : a % b; }
__device__ __forceinline__ int f2i32(double x) {
  if (!(x == x)) return 0;
  if (x >= 2147483647.0) return 2147483647;
  if (x <= -2147483648.0) return -2147483647 - 1;
  return (int)x;
}
__device__ __forceinline__ long long f2i64(double x) {
  if (!(x == x)) return 0;
  if (x >= 9223372036854775807.0) return 9223372036854775807LL;
  if (x <= -9223372036854775808.0) return -9223372036854775807LL - 1;
  return (long long)x;
}
__device__ __forceinline__ int fne(double a, double b) { return a < b || a > b; }
// floor division (date_scalar.rs floor_div) and the civil calendar of
// date_scalar.rs:37-60 (days since 1970-01-01 <-> proleptic Gregorian y-m-d)
__device__ __forceinline__ long long bfdiv(long long a, long long b) {
  const long long q = a / b, r = a % b;
  return (r != 0 && ((r < 0) != (b < 0))) ? q - 1 : q;
}
__device__ __forceinline__ void bcivil(long long days, long long* y, long long* m, long long* d) {
  const long long z = days + 719468;
  const long long era = bfdiv(z, 146097);
  const long long doe = z - era * 146097;
  const long long yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
  const long long doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
  const long long mp = (5 * doy + 2) / 153;
  *d = doy - (153 * mp + 2) / 5 + 1;
  *m = mp < 10 ? mp + 3 : mp - 9;
  *y = yoe + era * 400 + (*m <= 2 ? 1 : 0);
}
__device__ __forceinline__ long long bdays(long long y, long long m, long long d) {
  y -= m <= 2 ? 1 : 0;
  const long long era = bfdiv(y, 400);
  const long long yoe = y - era * 400;
  const long long doy = (153 * (m > 2 ? m - 3 : m + 9) + 2) / 5 + d - 1;
  return era * 146097 + yoe * 365 + yoe / 4 - yoe / 100 + doy - 719468;
}
__device__ __forceinline__ long long bp10(int s) { long long r = 1; for (int i = 0; i < s; i++) r *= 10; return r; }
// EXTRACT(field FROM x): x days (Date32, ts == 0) or ticks of 10^-unit s
__device__ __forceinline__ long long bextract(long long x, int ts, int unit, int field) {
  long long days = x, sod = 0;
  if (ts) {
    const long long tps = bp10(unit), tpd = 86400LL * tps;
    days = bfdiv(x, tpd);
    const long long r = x % tpd;  // the floor remainder, without forming days * tpd
    sod = (r < 0 ? r + tpd : r) / tps;
  }
  if (field >= 4) return field == 4 ? sod / 3600 : field == 5 ? (sod / 60) % 60 : sod % 60;
  long long y, m, d;
  bcivil(days, &y, &m, &d);
  return field == 1 ? y : field == 2 ? m : d;
}
// DATE_TRUNC(field, x): the start of x's year / month / day / hour / minute / second
__device__ __forceinline__ long long btrunc(long long x, int ts, int unit, int field) {
  const long long tps = ts ? bp10(unit) : 1, tpd = ts ? 86400LL * tps : 1;
  if (field >= 4) {
    const long long q = field == 4 ? 3600 * tps : field == 5 ? 60 * tps : tps;
    return (long long)((unsigned long long)bfdiv(x, q) * (unsigned long long)q);  // wraps
  }
  const long long days = bfdiv(x, tpd);
  long long y, m, d;
  bcivil(days, &y, &m, &d);
  const long long start = field == 1 ? bdays(y, 1, 1) : field == 2 ? bdays(y, m, 1) : days;
  return (long long)((unsigned long long)start * (unsigned long long)tpd);
}
// Decimal128 <-> float (physical_plan.rs F64ToI128 / I128ToF64): round half
// away from zero as trunc(|v| + 0.5), NaN -> 0, |v| >= 2^127 saturates; the
// i128 as (double)hi * 2^64 + (double)lo
__device__ __forceinline__ __int128 bf2d(double v) {
  if (!(v == v)) return 0;
  const double a = v < 0 ? -v : v;
  if (a >= 170141183460469231731687303715884105728.0)
    return v < 0 ? (__int128)((unsigned __int128)1 << 127) : (__int128)(((unsigned __int128)1 << 127) - 1);
  const double m = __builtin_trunc(a + 0.5);
  const double hd = __builtin_floor(m * 5.42101086242752217003726400434970855712890625e-20);
  const unsigned long long hi = (unsigned long long)hd;
  const unsigned long long lo = (unsigned long long)(m - hd * 18446744073709551616.0);
  const __int128 r = (__int128)(((unsigned __int128)hi << 64) | lo);
  return v < 0 ? -r : r;
}
__device__ __forceinline__ double bd2f(__int128 x) {
  return (double)(long long)(x >> 64) * 18446744073709551616.0 + (double)(unsigned long long)x;
}
__device__ __forceinline__ int bitv(const unsigned char* v, long long i) { return (v[i >> 3] >> (i & 7)) & 1; }
struct PArgs {
  const void* in[16];
  const unsigned char* iv[16];
  void* out[16];
  unsigned char* ov[16];
  long long n;
};
)BOLTP";

std::string reg(int r) { return "r" + std::to_string(r); }

std::string literal(const bolt_expr_op_t& o) {
  char buf[96];
  switch (o.dtype) {
    case BOLT_BOOL: return o.lit_i64 ? "1" : "0";
    case BOLT_I32: case BOLT_DATE32:
      snprintf(buf, sizeof buf, "(int)%lldLL", (long long)(int32_t)o.lit_i64);
      return buf;
    case BOLT_I64: case BOLT_TIMESTAMP:
      snprintf(buf, sizeof buf, "(long long)0x%016llxULL", (unsigned long long)o.lit_i64);
      return buf;
    case BOLT_F32: {
      float f = (float)o.lit_f64;
      uint32_t b;
      memcpy(&b, &f, 4);
      snprintf(buf, sizeof buf, "__builtin_bit_cast(float, 0x%08xu)", b);
      return buf;
    }
    case BOLT_F64: {
      uint64_t b;
      memcpy(&b, &o.lit_f64, 8);
      snprintf(buf, sizeof buf, "__builtin_bit_cast(double, 0x%016llxULL)", (unsigned long long)b);
      return buf;
    }
    case BOLT_DEC128:
      snprintf(buf, sizeof buf, "(__int128)(((unsigned __int128)0x%016llxULL << 64) | 0x%016llxULL)",
               (unsigned long long)o.lit_hi, (unsigned long long)o.lit_i64);
      return buf;
  }
  return "0";
}

// 10^s as a C literal: a 128-bit integer, or a double (s <= kMaxFloatScale)
std::string pow10_i128(int s) {
  unsigned __int128 v = 1;
  for (int i = 0; i < s; i++) v *= 10;
  char buf[96];
  snprintf(buf, sizeof buf, "(__int128)(((unsigned __int128)0x%016llxULL << 64) | 0x%016llxULL)",
           (unsigned long long)(v >> 64), (unsigned long long)v);
  return buf;
}
std::string pow10_f64(int s) {
  double v = 1.0;
  for (int i = 0; i < s; i++) v *= 10.0;  // exact for s <= 22
  uint64_t b;
  memcpy(&b, &v, 8);
  char buf[64];
  snprintf(buf, sizeof buf, "__builtin_bit_cast(double, 0x%016llxULL)", (unsigned long long)b);
  return buf;
}

// Casts that involve Decimal128 (physical_plan.rs:2046-2060); `sf` / `st`
// the source / target scales.
int decimal_cast_expr(int from, int to, int sf, int st, const std::string& src, std::string* out) {
  if (sf < 0 || sf > 38 || st < 0 || st > 38)
    return set_error(BOLT_ERR_TYPE, "bolt_project: Decimal128 scale out of 0..38 (%d, %d)", sf, st);
  const std::string wrapmul = "(__int128)((unsigned __int128)";
  if (to == BOLT_DEC128 && (from == BOLT_I32 || from == BOLT_I64 || from == BOLT_DATE32)) {
    *out = wrapmul + "(__int128)" + src + " * (unsigned __int128)" + pow10_i128(st) + ")";
    return BOLT_OK;
  }
  if (to == BOLT_DEC128 && from == BOLT_DEC128) {
    if (st >= sf) *out = wrapmul + src + " * (unsigned __int128)" + pow10_i128(st - sf) + ")";
    else *out = "(" + src + " / " + pow10_i128(sf - st) + ")";  // truncating toward zero
    return BOLT_OK;
  }
  if (from == BOLT_DEC128 && (to == BOLT_I32 || to == BOLT_I64)) {
    const std::string q = sf > 0 ? "(" + src + " / " + pow10_i128(sf) + ")" : src;
    *out = to == BOLT_I64 ? "(long long)(unsigned long long)(unsigned __int128)" + q
                          : "(int)(unsigned)(unsigned long long)(unsigned __int128)" + q;
    return BOLT_OK;
  }
  if (to == BOLT_DEC128 && is_float(from)) {
    if (st > kMaxFloatScale) return set_error(BOLT_ERR_TYPE, "bolt_project: float -> Decimal128 scale %d over 22", st);
    *out = "bf2d((double)" + src + " * " + pow10_f64(st) + ")";
    return BOLT_OK;
  }
  if (from == BOLT_DEC128 && is_float(to)) {
    if (sf > kMaxFloatScale) return set_error(BOLT_ERR_TYPE, "bolt_project: Decimal128 -> float scale %d over 22", sf);
    const std::string v = "(bd2f(" + src + ") / " + pow10_f64(sf) + ")";
    *out = to == BOLT_F32 ? "(float)" + v : v;
    return BOLT_OK;
  }
  return set_error(BOLT_ERR_TYPE, "bolt_project: no cast from %s to %s", dtype_name(from), dtype_name(to));
}

// C expression for `(to) src` under ptx_gen.rs's cast rules (:1890-1990).
int cast_expr(int from, int to, const std::string& src, std::string* out) {
  if (from == to) { *out = src; return BOLT_OK; }
  if (from == BOLT_DEC128 || to == BOLT_DEC128 || from == BOLT_DATE32 || to == BOLT_DATE32 ||
      from == BOLT_TIMESTAMP || to == BOLT_TIMESTAMP)
    return set_error(BOLT_ERR_TYPE, "bolt_project: no cast from %s to %s", dtype_name(from), dtype_name(to));
  if (to == BOLT_BOOL) {
    *out = is_float(from) ? "fne((double)" + src + ", 0.0)" : "(" + src + " != 0 ? 1 : 0)";
    return BOLT_OK;
  }
  if (is_float(from) && (to == BOLT_I32 || to == BOLT_I64)) {
    *out = std::string(to == BOLT_I32 ? "f2i32" : "f2i64") + "((double)" + src + ")";
    return BOLT_OK;
  }
  *out = std::string("(") + ctype(to) + ")" + src;  // Int / Bool -> Int / Float, float widening / narrowing
  return BOLT_OK;
}

int binary_expr(int op, int dt, const std::string& a, const std::string& b, std::string* out, int* rdt) {
  if (dt == BOLT_TIMESTAMP && !is_cmp(op))
    return set_error(BOLT_ERR_TYPE, "bolt_project: binary op %d not defined on Timestamp", op);
  const bool i = is_int(dt), f = is_float(dt), d = dt == BOLT_DEC128, bl = dt == BOLT_BOOL;
  const char* u = dt == BOLT_I64 ? "u64" : d ? "unsigned __int128" : "unsigned";
  *rdt = dt;
  auto wrap = [&](const char* sym) {
    return std::string("(") + ctype(dt) + ")((" + u + ")" + a + " " + sym + " (" + u + ")" + b + ")";
  };
  switch (op) {
    case BOLT_BIN_ADD: case BOLT_BIN_SUB: case BOLT_BIN_MUL: {
      const char* sym = op == BOLT_BIN_ADD ? "+" : op == BOLT_BIN_SUB ? "-" : "*";
      if (dt == BOLT_DATE32) {
        if (op != BOLT_BIN_SUB) break;  // Date32 - Date32 -> Int32 days only
        *out = wrap(sym);
        *rdt = BOLT_I32;
        return BOLT_OK;
      }
      if (f) *out = "(" + a + " " + sym + " " + b + ")";
      else if (i || d) *out = wrap(sym);
      else break;
      return BOLT_OK;
    }
    case BOLT_BIN_DIV: case BOLT_BIN_MOD:
      if (f && op == BOLT_BIN_DIV) { *out = "(" + a + " / " + b + ")"; return BOLT_OK; }
      if (dt == BOLT_I32 || dt == BOLT_I64) {
        *out = std::string(op == BOLT_BIN_DIV ? "bdiv" : "bmod") + (dt == BOLT_I32 ? "32(" : "64(") + a + ", " + b + ")";
        return BOLT_OK;
      }
      break;
    case BOLT_BIN_AND: case BOLT_BIN_OR:
      if (!bl) break;
      *out = "(" + a + (op == BOLT_BIN_AND ? " & " : " | ") + b + ")";
      return BOLT_OK;
    case BOLT_BIN_BITAND: case BOLT_BIN_BITOR:
      if (!(dt == BOLT_I32 || dt == BOLT_I64)) break;
      *out = "(" + a + (op == BOLT_BIN_BITAND ? " & " : " | ") + b + ")";
      return BOLT_OK;
    default:
      if (!is_cmp(op)) return set_error(BOLT_ERR_PLAN, "bolt_project: unknown binary op %d", op);
      *rdt = BOLT_BOOL;
      // (Timestamp operands: comparisons only, of equal units, checked by the caller)
      if (f && op == BOLT_BIN_NE) { *out = "fne(" + a + ", " + b + ")"; return BOLT_OK; }
      {
        static const char* syms[] = {"==", "!=", "<", "<=", ">", ">="};
        *out = "(" + a + " " + syms[op - BOLT_BIN_EQ] + " " + b + " ? 1 : 0)";
      }
      return BOLT_OK;
  }
  return set_error(BOLT_ERR_TYPE, "bolt_project: binary op %d not defined on %s", op, dtype_name(dt));
}

// A type-checked program: its statements (one per op, for row `r`) and, per
// STORE, the register, dtype and input dependencies of the stored value.
struct Compiled {
  std::string body;
  std::vector<int> reg, dtype, unit;
  std::vector<uint32_t> deps;
};

// `n_out` STOREs are expected (outputs 0 .. n_out - 1, each once); with
// `out` non-null each must match its output column's dtype.
int compile_program(const bolt_column_t* in, int n_in, const bolt_expr_op_t* prog, int n_ops,
                    const bolt_column_t* out, int n_out, Compiled* cp, const std::string& sfx = "",
                    const std::string& row = "r") {
  auto reg = [&](int r) { return "r" + std::to_string(r) + sfx; };
  std::vector<RegInfo> regs(BOLT_MAX_REGS);
  cp->reg.assign(n_out, -1);
  cp->dtype.assign(n_out, 0);
  cp->unit.assign(n_out, 0);
  cp->deps.assign(n_out, 0);
  std::ostringstream body;
  auto use = [&](int r, const char* what) -> int {
    if (r < 0 || r >= BOLT_MAX_REGS || regs[r].dtype == 0)
      return set_error(BOLT_ERR_PLAN, "bolt_project: %s reads register %d before it is written", what, r);
    return BOLT_OK;
  };
  for (int k = 0; k < n_ops; k++) {
    const bolt_expr_op_t& o = prog[k];
    if (o.kind != BOLT_X_STORE && (o.dst < 0 || o.dst >= BOLT_MAX_REGS || regs[o.dst].dtype != 0))
      return set_error(BOLT_ERR_PLAN, "bolt_project: op %d writes register %d twice or out of range", k, o.dst);
    RegInfo& d = regs[o.kind == BOLT_X_STORE ? 0 : o.dst];
    std::string e;
    switch (o.kind) {
      case BOLT_X_LOAD: {
        if (o.col < 0 || o.col >= n_in) return set_error(BOLT_ERR_PLAN, "bolt_project: op %d loads column %d", k, o.col);
        const int dt = in[o.col].dtype;
        if (!supported(dt) || dt != o.dtype)
          return set_error(BOLT_ERR_TYPE, "bolt_project: op %d loads column %d (%s) as %s", k, o.col, dtype_name(dt),
                           dtype_name(o.dtype));
        const std::string c = std::to_string(o.col);
        if (dt == BOLT_DEC128)
          e = "(__int128)(((unsigned __int128)((const u64*)p.in[" + c + "])[2 * " + row + " + 1] << 64) | ((const u64*)p.in[" +
              c + "])[2 * " + row + "])";
        else if (dt == BOLT_BOOL)
          e = "(((const unsigned char*)p.in[" + c + "])[" + row + "] != 0 ? 1 : 0)";
        else
          e = std::string("((const ") + stype(dt) + "*)p.in[" + c + "])[" + row + "]";
        d.dtype = dt;
        d.deps = 1u << o.col;
        if (dt == BOLT_TIMESTAMP) {
          if (!valid_unit(in[o.col].scale))
            return set_error(BOLT_ERR_TYPE, "bolt_project: column %d has Timestamp unit %d (0, 3, 6 or 9)", o.col,
                             in[o.col].scale);
          d.unit = in[o.col].scale;
        }
        break;
      }
      case BOLT_X_CONST:
        if (!supported(o.dtype)) return set_error(BOLT_ERR_TYPE, "bolt_project: op %d constant of dtype %d", k, o.dtype);
        if (o.dtype == BOLT_TIMESTAMP && !valid_unit(o.op))
          return set_error(BOLT_ERR_TYPE, "bolt_project: op %d Timestamp constant of unit %d", k, o.op);
        e = literal(o);
        d.dtype = o.dtype;
        d.unit = o.dtype == BOLT_TIMESTAMP ? o.op : 0;
        break;
      case BOLT_X_CAST: {
        BOLT_TRY(use(o.a, "CAST"));
        if (!supported(o.dtype)) return set_error(BOLT_ERR_TYPE, "bolt_project: op %d casts to dtype %d", k, o.dtype);
        const int from = regs[o.a].dtype;
        if ((from == BOLT_DEC128 || o.dtype == BOLT_DEC128) &&
            !(from == BOLT_DEC128 && o.dtype == BOLT_DEC128 && o.c == o.op)) {
          if (from == BOLT_BOOL || from == BOLT_TIMESTAMP || o.dtype == BOLT_BOOL || o.dtype == BOLT_DATE32 ||
              o.dtype == BOLT_TIMESTAMP)
            return set_error(BOLT_ERR_TYPE, "bolt_project: no cast from %s to %s", dtype_name(from),
                             dtype_name(o.dtype));
          BOLT_TRY(decimal_cast_expr(from, o.dtype, from == BOLT_DEC128 ? o.c : 0, o.dtype == BOLT_DEC128 ? o.op : 0,
                                     reg(o.a), &e));
        } else {
          BOLT_TRY(cast_expr(from, o.dtype, reg(o.a), &e));
        }
        d.dtype = o.dtype;
        d.unit = regs[o.a].unit;
        d.deps = regs[o.a].deps;
        break;
      }
      case BOLT_X_EXTRACT:
      case BOLT_X_DATE_TRUNC: {
        const char* what = o.kind == BOLT_X_EXTRACT ? "EXTRACT" : "DATE_TRUNC";
        BOLT_TRY(use(o.a, what));
        const int dt = regs[o.a].dtype;
        if (dt != BOLT_DATE32 && dt != BOLT_TIMESTAMP)
          return set_error(BOLT_ERR_TYPE, "bolt_project: op %d %s requires Date32 / Timestamp, got %s", k, what,
                           dtype_name(dt));
        if (o.op < BOLT_DF_YEAR || o.op > BOLT_DF_SECOND)
          return set_error(BOLT_ERR_PLAN, "bolt_project: op %d %s of unknown field %d", k, what, o.op);
        if (dt == BOLT_DATE32 && o.op >= BOLT_DF_HOUR)  // date_scalar.rs:465,557
          return set_error(BOLT_ERR_TYPE, "bolt_project: op %d %s of a time-of-day field from Date32 is undefined", k,
                           what);
        const std::string args = "(long long)" + reg(o.a) + ", " + (dt == BOLT_TIMESTAMP ? "1, " : "0, ") +
                                 std::to_string(regs[o.a].unit) + ", " + std::to_string(o.op) + ")";
        if (o.kind == BOLT_X_EXTRACT) {
          e = "bextract(" + args;
          d.dtype = BOLT_I64;
        } else {
          e = std::string("(") + ctype(dt) + ")btrunc(" + args;
          d.dtype = dt;
          d.unit = regs[o.a].unit;
        }
        d.deps = regs[o.a].deps;
        break;
      }
      case BOLT_X_BINARY: {
        BOLT_TRY(use(o.a, "BINARY"));
        BOLT_TRY(use(o.b, "BINARY"));
        if (regs[o.a].dtype != o.dtype || regs[o.b].dtype != o.dtype)
          return set_error(BOLT_ERR_TYPE, "bolt_project: op %d operands are %s and %s, op says %s (cast first)", k,
                           dtype_name(regs[o.a].dtype), dtype_name(regs[o.b].dtype), dtype_name(o.dtype));
        if (o.dtype == BOLT_TIMESTAMP && regs[o.a].unit != regs[o.b].unit)
          return set_error(BOLT_ERR_TYPE, "bolt_project: op %d compares Timestamps of units %d and %d", k,
                           regs[o.a].unit, regs[o.b].unit);
        int rdt = 0;
        BOLT_TRY(binary_expr(o.op, o.dtype, reg(o.a), reg(o.b), &e, &rdt));
        d.dtype = rdt;
        d.deps = regs[o.a].deps | regs[o.b].deps;
        break;
      }
      case BOLT_X_SELECT:
        BOLT_TRY(use(o.a, "SELECT"));
        BOLT_TRY(use(o.b, "SELECT"));
        BOLT_TRY(use(o.c, "SELECT"));
        if (regs[o.a].dtype != BOLT_BOOL || regs[o.b].dtype != regs[o.c].dtype ||
            (regs[o.b].dtype == BOLT_TIMESTAMP && regs[o.b].unit != regs[o.c].unit))
          return set_error(BOLT_ERR_TYPE, "bolt_project: op %d SELECT needs a Bool condition and equal value dtypes", k);
        e = "(" + reg(o.a) + " ? " + reg(o.b) + " : " + reg(o.c) + ")";
        d.dtype = regs[o.b].dtype;
        d.unit = regs[o.b].unit;
        d.deps = regs[o.a].deps | regs[o.b].deps | regs[o.c].deps;
        break;
      case BOLT_X_NOT:
        BOLT_TRY(use(o.a, "NOT"));
        if (regs[o.a].dtype != BOLT_BOOL) return set_error(BOLT_ERR_TYPE, "bolt_project: op %d NOT of a non-Bool", k);
        e = "(" + reg(o.a) + " ? 0 : 1)";
        d.dtype = BOLT_BOOL;
        d.deps = regs[o.a].deps;
        break;
      case BOLT_X_IS_NULL:
        if (o.col < 0 || o.col >= n_in) return set_error(BOLT_ERR_PLAN, "bolt_project: op %d tests column %d", k, o.col);
        if (in[o.col].validity)
          e = std::string("(bitv(p.iv[") + std::to_string(o.col) + "], " + row + ") ? " + (o.op ? "0 : 1)" : "1 : 0)");
        else
          e = o.op ? "0" : "1";  // a column without a bitmap has no NULLs
        d.dtype = BOLT_BOOL;
        break;
      case BOLT_X_STORE: {
        BOLT_TRY(use(o.a, "STORE"));
        if (o.col < 0 || o.col >= n_out) return set_error(BOLT_ERR_PLAN, "bolt_project: op %d stores column %d", k, o.col);
        if (cp->reg[o.col] >= 0) return set_error(BOLT_ERR_PLAN, "bolt_project: output %d is stored twice", o.col);
        if (out && out[o.col].dtype != regs[o.a].dtype)
          return set_error(BOLT_ERR_TYPE, "bolt_project: output %d is %s, its value %s", o.col,
                           dtype_name(out[o.col].dtype), dtype_name(regs[o.a].dtype));
        if (out && regs[o.a].dtype == BOLT_TIMESTAMP && out[o.col].scale != regs[o.a].unit)
          return set_error(BOLT_ERR_TYPE, "bolt_project: output %d has Timestamp unit %d, its value %d", o.col,
                           out[o.col].scale, regs[o.a].unit);
        cp->reg[o.col] = o.a;
        cp->dtype[o.col] = regs[o.a].dtype;
        cp->unit[o.col] = regs[o.a].unit;
        cp->deps[o.col] = regs[o.a].deps;
        continue;
      }
      default: return set_error(BOLT_ERR_PLAN, "bolt_project: op %d has unknown kind %d", k, o.kind);
    }
    body << "    const " << ctype(d.dtype) << " " << reg(o.dst) << " = " << e << ";\n";
  }
  for (int c = 0; c < n_out; c++)
    if (cp->reg[c] < 0) return set_error(BOLT_ERR_PLAN, "bolt_project: output %d is never stored", c);
  cp->body = body.str();
  return BOLT_OK;
}

// C condition "every input column in `deps` is non-NULL at row `row`".
std::string valid_expr(const bolt_column_t* in, int n_in, uint32_t deps, const std::string& row = "r") {
  std::string v = "1";
  for (int c = 0; c < n_in; c++)
    if ((deps >> c) & 1u && in[c].validity) v += " && bitv(p.iv[" + std::to_string(c) + "], " + row + ")";
  return v;
}

// The projection kernel.
int gen_source(const bolt_column_t* in, int n_in, const bolt_expr_op_t* prog, int n_ops, const bolt_column_t* out,
               int n_out, std::string* src) {
  Compiled cp;
  BOLT_TRY(compile_program(in, n_in, prog, n_ops, out, n_out, &cp));
  std::ostringstream stores, valid;
  for (int c = 0; c < n_out; c++) {
    const int dt = cp.dtype[c];
    const std::string oc = std::to_string(c), rg = reg(cp.reg[c]);
    if (dt == BOLT_DEC128)
      stores << "      ((u64*)p.out[" << oc << "])[2 * i] = (u64)" << rg << "; ((u64*)p.out[" << oc
             << "])[2 * i + 1] = (u64)((unsigned __int128)" << rg << " >> 64);\n";
    else
      stores << "      ((" << stype(dt) << "*)p.out[" << oc << "])[i] = (" << stype(dt) << ")" << rg << ";\n";
    if (out[c].validity)
      valid << "    { const u64 m = __ballot(in && (" << valid_expr(in, n_in, cp.deps[c]) << "));\n"
            << "      if ((threadIdx.x & 63) == 0)\n"
            << "        for (int k = 0; k < 8 && (i >> 3) + k < nbytes; k++) p.ov[" << oc
            << "][(i >> 3) + k] = (unsigned char)(m >> (8 * k)); }\n";
  }
  std::ostringstream s;
  s << kProjectPrelude
    << "extern \"C\" __global__ void __launch_bounds__(256) bolt_jit_kernel(PArgs p) {\n"
    << "  const long long stride = (long long)gridDim.x * 256;\n"
    << "  const long long end = (p.n + 63) / 64 * 64;  // wave-uniform trip count for the validity ballots\n"
    << "  const long long nbytes = (p.n + 7) / 8;\n"
    << "  (void)nbytes;\n"
    << "  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < end; i += stride) {\n"
    << "    const int in = i < p.n;\n"
    << "    const long long r = in ? i : 0;  // rows past the end load row 0 and store nothing\n"
    << cp.body << "    if (in) {\n"
    << stores.str() << "    }\n"
    << valid.str() << "  }\n}\n";
  *src = s.str();
  return BOLT_OK;
}

// The fused WHERE + aggregate kernel: the program per row, the predicate
// (a Bool output, true and non-NULL), then each aggregate over its output
// (skipped where that output is NULL); per-workgroup partials as the
// precompiled scalar aggregate writes them.  Each lane runs kAggUnroll rows
// per iteration, 256 apart, whose loads the compiler issues together.
constexpr int kAggUnroll = 4;

int gen_agg_source(const bolt_column_t* in, int n_in, const bolt_expr_op_t* prog, int n_ops, int n_values,
                   int filter_col, const AggList& aggs, const bolt_agg_t* spec, std::string* src) {
  std::ostringstream s;
  s << jit_prelude_text() << "\n" << kProjectPrelude
    << "extern \"C\" __global__ void __launch_bounds__(256) bolt_jit_kernel(PArgs p, bolt::Acc* __restrict__ parts) {\n"
    << "  using namespace bolt;\n";
  for (int a = 0; a < aggs.n; a++) s << "  Acc acc" << a << "; acc_init(acc" << a << ", " << aggs.a[a].op << ");\n";
  s << "  const long long tile = 256LL * " << kAggUnroll << ";\n"
    << "  for (long long base = (long long)blockIdx.x * tile + threadIdx.x; base < p.n; base += (long long)gridDim.x * tile) {\n";
  std::vector<Compiled> cps(kAggUnroll);
  for (int u = 0; u < kAggUnroll; u++) {
    const std::string sfx = "_" + std::to_string(u), row = "r" + sfx;
    s << "    const bool in" << sfx << " = base + " << 256 * u << " < p.n;\n"
      << "    const long long " << row << " = in" << sfx << " ? base + " << 256 * u << " : 0;\n";
    BOLT_TRY(compile_program(in, n_in, prog, n_ops, nullptr, n_values, &cps[u], sfx, row));
    s << cps[u].body;
  }
  for (int u = 0; u < kAggUnroll; u++) {
    const Compiled& cp = cps[u];
    const std::string sfx = "_" + std::to_string(u), row = "r" + sfx;
    auto rg = [&](int c) { return "r" + std::to_string(cp.reg[c]) + sfx; };
    s << "    if (in" << sfx;
    if (filter_col >= 0) s << " && " << rg(filter_col) << " && (" << valid_expr(in, n_in, cp.deps[filter_col], row) << ")";
    s << ") {\n";
    for (int a = 0; a < aggs.n; a++) {
      const std::string A = "acc" + std::to_string(a);
      if (aggs.a[a].op == BOLT_AGG_COUNT_STAR) {
        s << "      " << A << ".n++;\n";
        continue;
      }
      const int c = spec[a].col, dt = cp.dtype[c];
      const std::string v = rg(c);
      std::string vi = "0", vhi = "0", vf = "0.0";
      if (dt == BOLT_DEC128) {
        vi = "(long long)(u64)" + v;
        vhi = "(long long)(u64)((unsigned __int128)" + v + " >> 64)";
      } else if (is_float(dt)) {
        vf = "(double)" + v;
      } else {
        vi = "(long long)" + v;
        vf = "(double)(long long)" + v;
      }
      s << "      if (" << valid_expr(in, n_in, cp.deps[c], row) << ") acc_update(" << A << ", " << aggs.a[a].op << ", "
        << dt << ", " << vi << ", " << vhi << ", " << vf << ");\n";
    }
    s << "    }\n";
  }
  s << "  }\n  __shared__ Acc sh[4];\n";
  for (int a = 0; a < aggs.n; a++)
    s << "  block_reduce_acc(acc" << a << ", " << aggs.a[a].op << ", " << aggs.a[a].dtype
      << ", sh, 4, parts + (long long)blockIdx.x * " << aggs.n << " + " << a << ");\n";
  s << "}\n";
  *src = s.str();
  return BOLT_OK;
}

struct PArgs {
  const void* in[16];
  const unsigned char* iv[16];
  void* out[16];
  unsigned char* ov[16];
  long long n;
};

}  // namespace
}  // namespace bolt

using namespace bolt;

extern "C" int bolt_project(bolt_stream_t stream, const bolt_column_t* inputs, int32_t n_inputs,
                            const bolt_expr_op_t* prog, int32_t n_ops, bolt_column_t* outputs, int32_t n_outputs) {
  BOLT_CHECK_DEVICE();
  clear_error();
  hipStream_t s = (hipStream_t)stream;
  if (n_outputs < 1 || n_outputs > BOLT_MAX_PROJECT_OUTPUTS || outputs == nullptr)
    return set_error(BOLT_ERR_PLAN, "bolt_project: 1 to %d outputs", BOLT_MAX_PROJECT_OUTPUTS);
  if (n_inputs < 0 || n_inputs > BOLT_MAX_PROJECT_INPUTS || (n_inputs > 0 && inputs == nullptr))
    return set_error(BOLT_ERR_PLAN, "bolt_project: 0 to %d inputs", BOLT_MAX_PROJECT_INPUTS);
  if (n_ops < 1 || prog == nullptr) return set_error(BOLT_ERR_PLAN, "bolt_project: empty program");
  const int64_t n = outputs[0].length;
  if (n < 0) return set_error(BOLT_ERR_PLAN, "bolt_project: negative length");
  for (int c = 0; c < n_inputs; c++) {
    if (inputs[c].length != n)
      return set_error(BOLT_ERR_PLAN, "bolt_project: input %d has %lld rows, output 0 has %lld", c,
                       (long long)inputs[c].length, (long long)n);
    if (n > 0 && inputs[c].data == nullptr) return set_error(BOLT_ERR_PLAN, "bolt_project: input %d has no data", c);
  }
  for (int c = 0; c < n_outputs; c++) {
    if (outputs[c].length != n)
      return set_error(BOLT_ERR_PLAN, "bolt_project: output %d has %lld rows, output 0 has %lld", c,
                       (long long)outputs[c].length, (long long)n);
    if (!supported(outputs[c].dtype))
      return set_error(BOLT_ERR_TYPE, "bolt_project: output %d dtype %d", c, outputs[c].dtype);
    if (n > 0 && outputs[c].data == nullptr) return set_error(BOLT_ERR_PLAN, "bolt_project: output %d has no data", c);
  }
  std::string src;
  BOLT_TRY(gen_source(inputs, n_inputs, prog, n_ops, outputs, n_outputs, &src));
  if (n == 0) return BOLT_OK;
  hipFunction_t fn;
  BOLT_TRY(jit_get(src, &fn));
  PArgs a;
  memset(&a, 0, sizeof a);
  for (int c = 0; c < n_inputs; c++) {
    a.in[c] = inputs[c].data;
    a.iv[c] = inputs[c].validity;
  }
  for (int c = 0; c < n_outputs; c++) {
    a.out[c] = outputs[c].data;
    a.ov[c] = const_cast<uint8_t*>(outputs[c].validity);
  }
  a.n = n;
  void* args[] = {&a};
  const unsigned grid = grid_for(n, 256, 4, 256 * 32);
  BOLT_HIP(hipModuleLaunchKernel(fn, grid, 1, 1, 256, 1, 1, 0, s, args, nullptr));
  BOLT_HIP(hipStreamSynchronize(s));
  return BOLT_OK;
}

extern "C" int bolt_aggregate_expr(bolt_stream_t stream, const bolt_column_t* inputs, int32_t n_inputs,
                                   const bolt_expr_op_t* prog, int32_t n_ops, int32_t n_values, int32_t filter_col,
                                   const bolt_agg_t* aggs, int32_t n_aggs, bolt_scalar_t* results) {
  BOLT_CHECK_DEVICE();
  clear_error();
  hipStream_t s = (hipStream_t)stream;
  if (n_inputs < 1 || n_inputs > BOLT_MAX_PROJECT_INPUTS || inputs == nullptr)
    return set_error(BOLT_ERR_PLAN, "bolt_aggregate_expr: 1 to %d inputs", BOLT_MAX_PROJECT_INPUTS);
  if (n_values < 1 || n_values > BOLT_MAX_PROJECT_OUTPUTS)
    return set_error(BOLT_ERR_PLAN, "bolt_aggregate_expr: 1 to %d program values", BOLT_MAX_PROJECT_OUTPUTS);
  if (n_ops < 1 || prog == nullptr) return set_error(BOLT_ERR_PLAN, "bolt_aggregate_expr: empty program");
  const int64_t n = inputs[0].length;
  for (int c = 0; c < n_inputs; c++) {
    if (inputs[c].length != n)
      return set_error(BOLT_ERR_PLAN, "bolt_aggregate_expr: input %d has %lld rows, input 0 has %lld", c,
                       (long long)inputs[c].length, (long long)n);
    if (n > 0 && inputs[c].data == nullptr) return set_error(BOLT_ERR_PLAN, "bolt_aggregate_expr: input %d has no data", c);
  }
  Compiled cp;
  BOLT_TRY(compile_program(inputs, n_inputs, prog, n_ops, nullptr, n_values, &cp));
  if (filter_col >= n_values || (filter_col >= 0 && cp.dtype[filter_col] != BOLT_BOOL))
    return set_error(BOLT_ERR_PLAN, "bolt_aggregate_expr: filter value %d is not a Bool value of the program", filter_col);
  // the program's values as (dataless) columns, for the aggregate checks and output dtypes
  std::vector<bolt_column_t> vcols(n_values);
  static const uint8_t kSomeBitmap = 0;
  for (int v = 0; v < n_values; v++) {
    memset(&vcols[v], 0, sizeof vcols[v]);
    vcols[v].dtype = cp.dtype[v];
    vcols[v].length = n;
    bool nullable = false;
    for (int c = 0; c < n_inputs; c++) nullable |= ((cp.deps[v] >> c) & 1u) && inputs[c].validity;
    vcols[v].validity = nullable ? &kSomeBitmap : nullptr;
  }
  AggList all;
  BOLT_TRY(make_agg_list(aggs, n_aggs, vcols.data(), n_values, 0, &all));
  if (n_aggs == 0) return BOLT_OK;
  std::string src;
  BOLT_TRY(gen_agg_source(inputs, n_inputs, prog, n_ops, n_values, filter_col, all, aggs, &src));
  hipFunction_t fn;
  BOLT_TRY(jit_get(src, &fn));
  PArgs a;
  memset(&a, 0, sizeof a);
  for (int c = 0; c < n_inputs; c++) {
    a.in[c] = inputs[c].data;
    a.iv[c] = inputs[c].validity;
  }
  a.n = n;
  // at most 1024 workgroups of 256 (the best shape for a streaming read:
  // 0.95 -> 0.91 ms for the Q6-shaped expression over 256M rows)
  const unsigned grid = stream_grid(n, 256, kAggUnroll, 1024);
  DevBuf<Acc> parts, outs;
  BOLT_TRY(parts.alloc((size_t)grid * n_aggs, s));
  BOLT_TRY(outs.alloc((size_t)n_aggs, s));
  Acc* pp = parts.ptr;
  void* args[] = {&a, &pp};
  prof_begin(s);
  BOLT_HIP(hipModuleLaunchKernel(fn, grid, 1, 1, 256, 1, 1, 0, s, args, nullptr));
  prof_end(s);
  for (int first = 0; first < n_aggs; first += 4) {
    const int na = n_aggs - first < 4 ? n_aggs - first : 4;
    BOLT_TRY(launch_finish(na, parts.ptr, (int)grid, n_aggs, first, agg_chunk(all, first, na), outs.ptr + first, s));
  }
  Acc host[kMaxAggs];
  BOLT_HIP(hipMemcpyAsync(host, outs.ptr, sizeof(Acc) * n_aggs, hipMemcpyDeviceToHost, s));
  BOLT_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < n_aggs; i++) BOLT_TRY(acc_finalize(host[i], all.a[i].op, all.a[i].dtype, &results[i]));
  return BOLT_OK;
}

// ---- programs feeding GROUP BY and the join probe ----------------------------

namespace {

// The pre stage of groupby_with_pre.rs / agg_with_pre.rs: the program's
// n_values values materialised into pooled device columns (a value gets a
// validity bitmap when an input it loads has one), `filter_col` checked to
// be a Bool value.  Timestamp values keep the program's unit; Decimal128
// values take value_scales[v] (0 without it).
struct PreValues {
  std::vector<bolt_column_t> cols;
  std::vector<DevBuf<uint8_t>> bufs;
};

int run_pre(hipStream_t s, const char* who, const bolt_column_t* inputs, int32_t n_inputs,
            const bolt_expr_op_t* prog, int32_t n_ops, int32_t n_values, const int32_t* value_scales,
            int32_t filter_col, PreValues* pv) {
  if (n_inputs < 1 || n_inputs > BOLT_MAX_PROJECT_INPUTS || inputs == nullptr)
    return set_error(BOLT_ERR_PLAN, "%s: 1 to %d inputs", who, BOLT_MAX_PROJECT_INPUTS);
  if (n_values < 1 || n_values > BOLT_MAX_PROJECT_OUTPUTS)
    return set_error(BOLT_ERR_PLAN, "%s: 1 to %d program values", who, BOLT_MAX_PROJECT_OUTPUTS);
  if (n_ops < 1 || prog == nullptr) return set_error(BOLT_ERR_PLAN, "%s: empty program", who);
  const int64_t n = inputs[0].length;
  for (int c = 0; c < n_inputs; c++)
    if (inputs[c].length != n)
      return set_error(BOLT_ERR_PLAN, "%s: input %d has %lld rows, input 0 has %lld", who, c,
                       (long long)inputs[c].length, (long long)n);
  Compiled cp;
  BOLT_TRY(compile_program(inputs, n_inputs, prog, n_ops, nullptr, n_values, &cp));
  if (filter_col >= n_values || (filter_col >= 0 && cp.dtype[filter_col] != BOLT_BOOL))
    return set_error(BOLT_ERR_PLAN, "%s: filter value %d is not a Bool value of the program", who, filter_col);
  pv->cols.assign(n_values, bolt_column_t{});
  pv->bufs.resize(2 * (size_t)n_values);
  for (int v = 0; v < n_values; v++) {
    bolt_column_t& c = pv->cols[v];
    c.dtype = cp.dtype[v];
    c.length = n;
    c.scale = c.dtype == BOLT_TIMESTAMP ? cp.unit[v] : c.dtype == BOLT_DEC128 && value_scales ? value_scales[v] : 0;
    BOLT_TRY(pv->bufs[2 * v].alloc((size_t)(n ? n : 1) * dtype_size(c.dtype), s));
    c.data = pv->bufs[2 * v].ptr;
    bool nullable = false;
    for (int i = 0; i < n_inputs; i++) nullable |= ((cp.deps[v] >> i) & 1u) && inputs[i].validity;
    if (nullable) {
      BOLT_TRY(pv->bufs[2 * v + 1].alloc((size_t)(n + 7) / 8 + 8, s));
      c.validity = pv->bufs[2 * v + 1].ptr;
    }
  }
  return bolt_project((bolt_stream_t)s, inputs, n_inputs, prog, n_ops, pv->cols.data(), n_values);
}

}  // namespace

extern "C" int bolt_groupby_expr(bolt_stream_t stream, const bolt_column_t* inputs, int32_t n_inputs,
                                 const bolt_expr_op_t* prog, int32_t n_ops, int32_t n_values,
                                 const int32_t* value_scales, int32_t filter_col, const int32_t* key_values,
                                 int32_t n_keys, const bolt_agg_t* aggs, int32_t n_aggs, uint32_t flags,
                                 bolt_table_t** out) {
  BOLT_CHECK_DEVICE();
  clear_error();
  if (!out) return set_error(BOLT_ERR_PLAN, "bolt_groupby_expr: null output");
  *out = nullptr;
  if (n_keys < 1 || n_keys > 2 || key_values == nullptr)
    return set_error(BOLT_ERR_PLAN, "bolt_groupby_expr: 1 or 2 key values");
  for (int k = 0; k < n_keys; k++)
    if (key_values[k] < 0 || key_values[k] >= n_values)
      return set_error(BOLT_ERR_PLAN, "bolt_groupby_expr: key value %d of %d", key_values[k], n_values);
  hipStream_t s = (hipStream_t)stream;
  PreValues pv;
  BOLT_TRY(run_pre(s, "bolt_groupby_expr", inputs, n_inputs, prog, n_ops, n_values, value_scales, filter_col, &pv));
  bolt_column_t keys[2];
  for (int k = 0; k < n_keys; k++) keys[k] = pv.cols[key_values[k]];
  const bolt_pred_term_t keep = {filter_col, BOLT_CMP_EQ, 1, 0.0, 0};  // true and not NULL
  const bolt_filter_t f = {&keep, 1};
  return bolt_groupby(stream, keys, n_keys, pv.cols.data(), n_values, aggs, n_aggs, filter_col >= 0 ? &f : nullptr,
                      flags, out);
}

extern "C" int bolt_join_probe_aggregate_expr(bolt_stream_t stream, const bolt_join_table_t* t,
                                              const bolt_column_t* inputs, int32_t n_inputs,
                                              const bolt_expr_op_t* prog, int32_t n_ops, int32_t n_values,
                                              const int32_t* value_scales, int32_t key_value, int32_t filter_col,
                                              const bolt_agg_t* aggs, int32_t n_aggs, bolt_scalar_t* results) {
  BOLT_CHECK_DEVICE();
  clear_error();
  if (key_value < 0 || key_value >= n_values)
    return set_error(BOLT_ERR_PLAN, "bolt_join_probe_aggregate_expr: key value %d of %d", key_value, n_values);
  hipStream_t s = (hipStream_t)stream;
  PreValues pv;
  BOLT_TRY(run_pre(s, "bolt_join_probe_aggregate_expr", inputs, n_inputs, prog, n_ops, n_values, value_scales,
                   filter_col, &pv));
  const bolt_pred_term_t keep = {filter_col, BOLT_CMP_EQ, 1, 0.0, 0};
  const bolt_filter_t f = {&keep, 1};
  return bolt_join_probe_aggregate(stream, t, pv.cols.data(), n_values, key_value, filter_col >= 0 ? &f : nullptr,
                                   aggs, n_aggs, results);
}
