// kzo_flops.cpp — FP64 operation count of the oracle's implicit solve.
//
// TEST INFRASTRUCTURE (the oracle, built differently): the oracle's C
// sources are compiled here as C++ with `double` replaced by a counting
// type, so that every +, -, *, / and sqrt executed by kzo_solve_implicit_lab
// (implicit.zig:927-1004 and everything it calls, including the kz_math.h
// transcendentals, whose internal operations are counted the same way)
// adds one to a counter, and every fused multiply-add two.  Comparisons, selects, fabs, negation and bit
// manipulation add nothing.  The count is the algorithmic FP64 work of a
// cell-solve: the GPU kernels execute the same operation sequence (they are
// checked bit for bit against this oracle), so the count times the solves
// gives the flops of an implicit pass.
//
// Exported (C ABI): kzo_count_implicit_batch — the oracle's batch entry
// point plus a per-record flop count; kzo_count_totals — running totals of
// all operations, divisions and square roots.
#include <cstdint>
#include <cstring>
#include <cmath>

namespace kzc {
extern unsigned long long nflop, ndiv, nsqrt;
struct cd {
  double v;
  cd() = default;
  constexpr cd(double x) : v(x) {}
  explicit operator double() const { return v; }
  explicit operator float() const { return (float)v; }
  explicit operator int() const { return (int)v; }
  explicit operator long() const { return (long)v; }
  explicit operator long long() const { return (long long)v; }
  explicit operator unsigned() const { return (unsigned)v; }
  explicit operator unsigned long() const { return (unsigned long)v; }
  explicit operator unsigned long long() const { return (unsigned long long)v; }
  explicit operator bool() const { return v != 0.0; }
  cd operator-() const { return cd(-v); }
  cd operator+() const { return *this; }
  cd &operator+=(cd b) { ++nflop; v += b.v; return *this; }
  cd &operator-=(cd b) { ++nflop; v -= b.v; return *this; }
  cd &operator*=(cd b) { ++nflop; v *= b.v; return *this; }
  cd &operator/=(cd b) { ++nflop; ++ndiv; v /= b.v; return *this; }
};
inline cd operator+(cd a, cd b) { ++nflop; return cd(a.v + b.v); }
inline cd operator-(cd a, cd b) { ++nflop; return cd(a.v - b.v); }
inline cd operator*(cd a, cd b) { ++nflop; return cd(a.v * b.v); }
inline cd operator/(cd a, cd b) { ++nflop; ++ndiv; return cd(a.v / b.v); }
inline bool operator<(cd a, cd b) { return a.v < b.v; }
inline bool operator>(cd a, cd b) { return a.v > b.v; }
inline bool operator<=(cd a, cd b) { return a.v <= b.v; }
inline bool operator>=(cd a, cd b) { return a.v >= b.v; }
inline bool operator==(cd a, cd b) { return a.v == b.v; }
inline bool operator!=(cd a, cd b) { return a.v != b.v; }
inline cd csqrt(cd a) { ++nflop; ++nsqrt; return cd(std::sqrt(a.v)); }
inline cd cfabs(cd a) { return cd(std::fabs(a.v)); }
inline cd cfma(cd a, cd b, cd c) { nflop += 2; return cd(std::fma(a.v, b.v, c.v)); }
static_assert(sizeof(cd) == 8, "counting double must keep the layout of double");
}  // namespace kzc

#define double kzc::cd
#define __builtin_sqrt(x) kzc::csqrt(x)
#define __builtin_fabs(x) kzc::cfabs(x)
#define KZM_FMA(a, b, c) kzc::cfma(a, b, c)
#define __builtin_inf() kzc::cd(HUGE_VAL)
#define __builtin_nan(s) kzc::cd(NAN)

#include "../kzo_cell.c"
#include "../kzo_solve.c"

#undef double

unsigned long long kzc::nflop = 0, kzc::ndiv = 0, kzc::nsqrt = 0;

// totals since load: operations, of which divisions and square roots
extern "C" void kzo_count_totals(uint64_t out[3]) {
  out[0] = kzc::nflop;
  out[1] = kzc::ndiv;
  out[2] = kzc::nsqrt;
}

extern "C" int kzo_count_implicit_batch(int64_t n, kzc::cd *uu, kzc::cd *pp, const kz_geom *g, const kzc::cd *dt,
                                        kzc::cd gam, const kz_rad_params *rp, const kz_opac_params *op,
                                        const kz_implicit_params *ip, kz_implicit_result *r, uint64_t *flops) {
  for (int64_t i = 0; i < n; ++i) {
    kzo_geo gm;
    kzo_geo_from_kz(&g[i], &gm);
    unsigned long long f0 = kzc::nflop;
    r[i] = kzo_solve_implicit_lab(uu + i * NV, pp + i * NV, &gm, dt[i], gam, rp, op, ip);
    flops[i] = kzc::nflop - f0;
  }
  return 0;
}
