// TEST INFRASTRUCTURE ONLY: a double that counts its floating-point operations, for the
// flop-counting build of the oracle (liboracle_flops.so). The oracle's internal arithmetic type
// `Real` is this class in that build and plain double in liboracle.so.
//
// Counting convention: add, sub, mul, div, sqrt, log, exp, pow (and the other elementary
// functions) count 1 each; negation, comparisons, fabs, fmin, fmax and copies count 0. The oracle
// is compiled without FMA contraction, so no operation counts 2. Counts go to the bucket selected
// by flopcount::g_bucket (the reference's timer sections and the volume-flux and change-of-basis
// kernels); the counting build is single-threaded.
#pragma once
#include <cmath>
#include <cstdint>
#include <type_traits>

namespace flopcount {
constexpr int kBuckets = 16;
inline uint64_t g_flops[kBuckets] = {};
inline uint64_t g_calls[kBuckets] = {};
inline int g_bucket = 0;
inline void tick() { ++g_flops[g_bucket]; }
// selects a bucket for the lifetime of the object (nested sections restore the outer one)
struct Scope {
    int prev;
    explicit Scope(int b) : prev(g_bucket) {
        g_bucket = b;
        ++g_calls[b];
    }
    ~Scope() { g_bucket = prev; }
};
}  // namespace flopcount

struct Real {
    double v;
    Real() = default;
    constexpr Real(double x) : v(x) {}
    constexpr operator double() const { return v; }
    Real& operator+=(Real o) { flopcount::tick(); v += o.v; return *this; }
    Real& operator-=(Real o) { flopcount::tick(); v -= o.v; return *this; }
    Real& operator*=(Real o) { flopcount::tick(); v *= o.v; return *this; }
    Real& operator/=(Real o) { flopcount::tick(); v /= o.v; return *this; }
};
static_assert(sizeof(Real) == sizeof(double), "Real must be layout-compatible with double");

// with an integer operand the operation is a double one (counted); with a long double operand it
// stays long double (set-up code of the basis, not counted)
#define FLOPCOUNT_BINOP(op)                                                                            \
    inline Real operator op(Real a, Real b) { flopcount::tick(); return Real(a.v op b.v); }          \
    inline Real operator op(Real a, double b) { flopcount::tick(); return Real(a.v op b); }          \
    inline Real operator op(double a, Real b) { flopcount::tick(); return Real(a op b.v); }          \
    inline Real operator op(Real a, float b) { flopcount::tick(); return Real(a.v op (double)b); }   \
    inline Real operator op(float a, Real b) { flopcount::tick(); return Real((double)a op b.v); }   \
    template <typename I, typename = std::enable_if_t<std::is_integral_v<I>>>                         \
    inline Real operator op(Real a, I b) { flopcount::tick(); return Real(a.v op (double)b); }       \
    template <typename I, typename = std::enable_if_t<std::is_integral_v<I>>>                         \
    inline Real operator op(I a, Real b) { flopcount::tick(); return Real((double)a op b.v); }       \
    inline long double operator op(Real a, long double b) { return (long double)a.v op b; }          \
    inline long double operator op(long double a, Real b) { return a op (long double)b.v; }
FLOPCOUNT_BINOP(+)
FLOPCOUNT_BINOP(-)
FLOPCOUNT_BINOP(*)
FLOPCOUNT_BINOP(/)
#undef FLOPCOUNT_BINOP
inline Real operator-(Real a) { return Real(-a.v); }
inline Real operator+(Real a) { return a; }

#define FLOPCOUNT_CMP(op)                                                                   \
    inline bool operator op(Real a, Real b) { return a.v op b.v; }                         \
    inline bool operator op(Real a, double b) { return a.v op b; }                         \
    inline bool operator op(double a, Real b) { return a op b.v; }                         \
    template <typename I, typename = std::enable_if_t<std::is_integral_v<I>>>               \
    inline bool operator op(Real a, I b) { return a.v op (double)b; }                      \
    template <typename I, typename = std::enable_if_t<std::is_integral_v<I>>>               \
    inline bool operator op(I a, Real b) { return (double)a op b.v; }                      \
    inline bool operator op(Real a, long double b) { return (long double)a.v op b; }       \
    inline bool operator op(long double a, Real b) { return a op (long double)b.v; }
FLOPCOUNT_CMP(<)
FLOPCOUNT_CMP(>)
FLOPCOUNT_CMP(<=)
FLOPCOUNT_CMP(>=)
FLOPCOUNT_CMP(==)
FLOPCOUNT_CMP(!=)
#undef FLOPCOUNT_CMP

// elementary functions of Real, in Real's namespace so that the oracle's unqualified calls find them
// by argument-dependent lookup (the double ones come from its using-declarations of std's)
#define FLOPCOUNT_FN1(f)                                                     \
    inline Real f(Real a) { flopcount::tick(); return Real(::std::f(a.v)); }
FLOPCOUNT_FN1(sqrt)
FLOPCOUNT_FN1(log)
FLOPCOUNT_FN1(exp)
FLOPCOUNT_FN1(tanh)
FLOPCOUNT_FN1(sin)
FLOPCOUNT_FN1(cos)
FLOPCOUNT_FN1(tan)
#undef FLOPCOUNT_FN1
inline Real pow(Real a, Real b) { flopcount::tick(); return Real(::std::pow(a.v, b.v)); }
inline Real pow(Real a, double b) { flopcount::tick(); return Real(::std::pow(a.v, b)); }
inline Real pow(double a, Real b) { flopcount::tick(); return Real(::std::pow(a, b.v)); }
inline Real fabs(Real a) { return Real(::std::fabs(a.v)); }
