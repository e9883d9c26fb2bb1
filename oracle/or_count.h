/* or_count.h -- prelude of the counting build of the oracle (liboracle_count.so).
 *
 * TEST INFRASTRUCTURE ONLY.  elegant_oracle.c is compiled a third time, as C++,
 * with every `double` replaced by or_cd: a double whose +, -, *, / and sqrt and
 * whose sin, cos, tan, asin and log calls are counted.  Values are computed
 * exactly as in liboracle_libm.so (IEEE doubles, C-library transcendentals).
 *
 * Two tallies are kept.  or_cnt_all counts every operation executed.  or_cnt is
 * the census: every operation executed, except inside the generic
 * transport-matrix loops (track_particles for DRIF's drift matrix and ROTATE's
 * rotation matrix, marked by OR_COUNT_MATRIX_BEGIN/END), which multiply through
 * every zero and unit entry of the matrix: there an addition or subtraction with
 * an exact zero operand and a multiplication by an exact 0 or +-1 leave the
 * other operand (or a constant) as it is and are not counted.  With particles
 * whose coordinates are all non-zero, such operands come only from the matrix.
 * Negation, fabs and comparisons count in neither tally; every transcendental
 * call counts in both. */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <type_traits>

extern "C" {
/* add, sub, mul, div, sqrt, transcendental */
extern uint64_t or_cnt[6], or_cnt_all[6];
extern int or_cnt_matrix; /* inside a generic matrix loop */
}

static inline bool or_trivial_mul(double x) { return x == 0 || x == 1 || x == -1; }
static inline void or_tally(int k, bool work) {
  or_cnt_all[k]++;
  if (work || !or_cnt_matrix)
    or_cnt[k]++;
}
#define OR_COUNT_MATRIX_BEGIN() (or_cnt_matrix = 1)
#define OR_COUNT_MATRIX_END() (or_cnt_matrix = 0)

struct or_cd {
  double v;
  or_cd() = default;
  constexpr or_cd(double x) : v(x) {}
  explicit operator double() const { return v; }
  explicit operator bool() const { return v != 0; }
  template <class T, class = typename std::enable_if<std::is_integral<T>::value>::type>
  explicit operator T() const {
    return (T)v;
  }
  or_cd &operator+=(or_cd o) {
    or_tally(0, v != 0 && o.v != 0);
    v += o.v;
    return *this;
  }
  or_cd &operator-=(or_cd o) {
    or_tally(1, v != 0 && o.v != 0);
    v -= o.v;
    return *this;
  }
  or_cd &operator*=(or_cd o) {
    or_tally(2, !or_trivial_mul(v) && !or_trivial_mul(o.v));
    v *= o.v;
    return *this;
  }
  or_cd &operator/=(or_cd o) {
    or_tally(3, v != 0 && o.v != 1 && o.v != -1);
    v /= o.v;
    return *this;
  }
};
static_assert(sizeof(or_cd) == sizeof(double) && std::is_trivially_copyable<or_cd>::value, "layout");

static inline or_cd operator+(or_cd a, or_cd b) {
  or_tally(0, a.v != 0 && b.v != 0);
  return or_cd(a.v + b.v);
}
static inline or_cd operator-(or_cd a, or_cd b) {
  or_tally(1, a.v != 0 && b.v != 0);
  return or_cd(a.v - b.v);
}
static inline or_cd operator*(or_cd a, or_cd b) {
  or_tally(2, !or_trivial_mul(a.v) && !or_trivial_mul(b.v));
  return or_cd(a.v * b.v);
}
static inline or_cd operator/(or_cd a, or_cd b) {
  or_tally(3, a.v != 0 && b.v != 1 && b.v != -1);
  return or_cd(a.v / b.v);
}
static inline or_cd operator-(or_cd a) { return or_cd(-a.v); }
static inline or_cd operator+(or_cd a) { return a; }
static inline bool operator<(or_cd a, or_cd b) { return a.v < b.v; }
static inline bool operator>(or_cd a, or_cd b) { return a.v > b.v; }
static inline bool operator<=(or_cd a, or_cd b) { return a.v <= b.v; }
static inline bool operator>=(or_cd a, or_cd b) { return a.v >= b.v; }
static inline bool operator==(or_cd a, or_cd b) { return a.v == b.v; }
static inline bool operator!=(or_cd a, or_cd b) { return a.v != b.v; }
static inline bool operator!(or_cd a) { return !a.v; }

static inline or_cd sqrt(or_cd a) {
  or_tally(4, a.v != 0 && a.v != 1);
  return or_cd(::sqrt(a.v));
}
static inline or_cd fabs(or_cd a) { return or_cd(::fabs(a.v)); }
static inline or_cd ceil(or_cd a) { return or_cd(::ceil(a.v)); }
static inline bool isnan(or_cd a) { return a.v != a.v; }
static inline bool isinf(or_cd a) { return ::fabs(a.v) == HUGE_VAL; }
/* the arguments of the transcendental calls, when a log is open (kind: 0 sin,
 * 1 cos, 2 tan, 3 asin, 4 log) */
extern "C" {
extern double *or_targ;
extern int32_t *or_tkind;
extern long or_targ_n, or_targ_cap;
}
static inline void or_targ_note(int kind, double x) {
  if (or_targ_n < or_targ_cap) {
    or_targ[or_targ_n] = x;
    or_tkind[or_targ_n] = kind;
    or_targ_n++;
  }
}
#define OR_COUNT_T(name, kind)                                                                                    \
  static inline or_cd name(or_cd a) {                                                                              \
    or_tally(5, true);                                                                                             \
    or_targ_note(kind, a.v);                                                                                       \
    return or_cd(::name(a.v));                                                                                     \
  }
OR_COUNT_T(sin, 0)
OR_COUNT_T(cos, 1)
OR_COUNT_T(tan, 2)
OR_COUNT_T(asin, 3)
OR_COUNT_T(log, 4)
OR_COUNT_T(exp, 5)
#undef OR_COUNT_T

#define double or_cd
