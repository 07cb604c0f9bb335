// flopcount.cpp -- the oracle recompiled with an operation-counting scalar (test infrastructure).
//
// spx_oracle.c is included unchanged with the keyword `double` replaced by `Flop`, a
// one-double struct whose arithmetic operators and math overloads increment counters.
// The result, liboracle_flops.so, exposes the oracle's C interface plus orc_flops_read /
// orc_flops_reset; oracle/flops_per_step.py runs a workload through it and divides the
// counted operations by the trajectory steps taken.  Only literal constants and integer
// arithmetic escape the count, which is what the kernels fold at compile time as well.
#include <stdint.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <stddef.h>
#include <omp.h>
#include <type_traits>

namespace flop {
enum { ADD, MUL, DIV, SQRT, CMP, NEG, EXP, LOG, POW, TANH, TRIG, ABSMINMAX, N_KIND };
static unsigned long long count[N_KIND];
}

struct Flop {
   double v;
   Flop() = default;
   constexpr Flop(double x) : v(x) {}
   template <typename I, typename std::enable_if<std::is_integral<I>::value, int>::type = 0>
   constexpr Flop(I x) : v((double)x) {}
   constexpr Flop(float x) : v((double)x) {}
   template <typename T, typename std::enable_if<std::is_arithmetic<T>::value, int>::type = 0>
   explicit operator T() const { return (T)v; }
   explicit operator bool() const { return v != 0.0; }
   Flop operator-() const { flop::count[flop::NEG]++; return Flop(-v); }
   Flop operator+() const { return *this; }
   Flop& operator+=(Flop o) { flop::count[flop::ADD]++; v += o.v; return *this; }
   Flop& operator-=(Flop o) { flop::count[flop::ADD]++; v -= o.v; return *this; }
   Flop& operator*=(Flop o) { flop::count[flop::MUL]++; v *= o.v; return *this; }
   Flop& operator/=(Flop o) { flop::count[flop::DIV]++; v /= o.v; return *this; }
};
static_assert(sizeof(Flop) == sizeof(double) && std::is_trivially_copyable<Flop>::value, "Flop must stay a plain double");

static inline Flop operator+(Flop a, Flop b) { flop::count[flop::ADD]++; return Flop(a.v + b.v); }
static inline Flop operator-(Flop a, Flop b) { flop::count[flop::ADD]++; return Flop(a.v - b.v); }
static inline Flop operator*(Flop a, Flop b) { flop::count[flop::MUL]++; return Flop(a.v * b.v); }
static inline Flop operator/(Flop a, Flop b) { flop::count[flop::DIV]++; return Flop(a.v / b.v); }
static inline bool operator<(Flop a, Flop b) { flop::count[flop::CMP]++; return a.v < b.v; }
static inline bool operator>(Flop a, Flop b) { flop::count[flop::CMP]++; return a.v > b.v; }
static inline bool operator<=(Flop a, Flop b) { flop::count[flop::CMP]++; return a.v <= b.v; }
static inline bool operator>=(Flop a, Flop b) { flop::count[flop::CMP]++; return a.v >= b.v; }
static inline bool operator==(Flop a, Flop b) { flop::count[flop::CMP]++; return a.v == b.v; }
static inline bool operator!=(Flop a, Flop b) { flop::count[flop::CMP]++; return a.v != b.v; }

static inline Flop sqrt(Flop a) { flop::count[flop::SQRT]++; return Flop(::sqrt(a.v)); }
static inline Flop exp(Flop a) { flop::count[flop::EXP]++; return Flop(::exp(a.v)); }
static inline Flop log(Flop a) { flop::count[flop::LOG]++; return Flop(::log(a.v)); }
static inline Flop log10(Flop a) { flop::count[flop::LOG]++; return Flop(::log10(a.v)); }
static inline Flop pow(Flop a, Flop b) { flop::count[flop::POW]++; return Flop(::pow(a.v, b.v)); }
static inline Flop cbrt(Flop a) { flop::count[flop::POW]++; return Flop(::cbrt(a.v)); }
static inline Flop tgamma(Flop a) { flop::count[flop::POW]++; return Flop(::tgamma(a.v)); }
static inline Flop tanh(Flop a) { flop::count[flop::TANH]++; return Flop(::tanh(a.v)); }
static inline Flop cos(Flop a) { flop::count[flop::TRIG]++; return Flop(::cos(a.v)); }
static inline Flop sin(Flop a) { flop::count[flop::TRIG]++; return Flop(::sin(a.v)); }
static inline Flop acos(Flop a) { flop::count[flop::TRIG]++; return Flop(::acos(a.v)); }
static inline Flop atan2(Flop a, Flop b) { flop::count[flop::TRIG]++; return Flop(::atan2(a.v, b.v)); }
static inline Flop fabs(Flop a) { flop::count[flop::ABSMINMAX]++; return Flop(::fabs(a.v)); }
static inline Flop fmin(Flop a, Flop b) { flop::count[flop::ABSMINMAX]++; return Flop(::fmin(a.v, b.v)); }
static inline Flop fmax(Flop a, Flop b) { flop::count[flop::ABSMINMAX]++; return Flop(::fmax(a.v, b.v)); }
static inline Flop floor(Flop a) { flop::count[flop::ABSMINMAX]++; return Flop(::floor(a.v)); }
static inline bool isnan(Flop a) { return ::isnan(a.v); }
static inline bool isfinite(Flop a) { return ::isfinite(a.v); }
static inline bool isnormal(Flop a) { return ::isnormal(a.v); }

#define double Flop
#include "spx_oracle.c"
#undef double

extern "C" {
// counters in the order ADD MUL DIV SQRT CMP NEG EXP LOG POW TANH TRIG ABSMINMAX
int orc_flops_read(unsigned long long* out, int n)
{
   for (int k = 0; k < n && k < flop::N_KIND; k++) out[k] = flop::count[k];
   return flop::N_KIND;
}
void orc_flops_reset(void) { memset(flop::count, 0, sizeof(flop::count)); }
}
