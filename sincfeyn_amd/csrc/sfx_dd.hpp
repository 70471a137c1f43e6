// sfx_dd.hpp -- double-double arithmetic for the high-precision fallback, usable on gfx950 and on the
// host (the host instantiation is what the CPU tests compare against the reference's dd).
//
// A value is hi + lo with |lo| <= ulp(hi)/2.  The integrand kernels are rational functions, so only
// + - * /, negation and comparisons are needed.  The operation sequence of every function is the one of
// the reference's driver/dd.h:18-75 (Dekker/Knuth two-sum, fma two-product, the "sloppy" add that the
// reference's fallback error model QEPS * sum|terms| assumes, dd/dd with three quotient digits, and the
// mixed dd-op-double forms that skip the promotion).  Keeping the sequence identical is what makes a
// device fallback value equal, bit for bit, to the reference's value computed without FMA contraction.
// Translation units that include this header are compiled with FP contraction off; the one fused
// operation is the explicit fma of two_prod.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#define SFX_HD __host__ __device__ __forceinline__

struct sfx_dd {
    double hi, lo;
    SFX_HD sfx_dd() : hi(0), lo(0) {}
    SFX_HD sfx_dd(double x) : hi(x), lo(0) {}
    SFX_HD sfx_dd(double h, double l) : hi(h), lo(l) {}
    SFX_HD explicit operator double() const { return hi + lo; }
};

namespace sfxdd {
SFX_HD sfx_dd two_sum(double a, double b){ double s = a + b; double bb = s - a; return sfx_dd(s, (a - (s - bb)) + (b - bb)); }
SFX_HD sfx_dd quick_two_sum(double a, double b){ double s = a + b; return sfx_dd(s, b - (s - a)); }
SFX_HD sfx_dd two_prod(double a, double b){ double p = a*b; return sfx_dd(p, fma(a, b, -p)); }
SFX_HD sfx_dd add(const sfx_dd &a, const sfx_dd &b){        // sloppy add
    sfx_dd s = two_sum(a.hi, b.hi); s.lo += a.lo + b.lo; return quick_two_sum(s.hi, s.lo);
}
SFX_HD sfx_dd add_d(const sfx_dd &a, double b){ sfx_dd s = two_sum(a.hi, b); s.lo += a.lo; return quick_two_sum(s.hi, s.lo); }
SFX_HD sfx_dd mul_d(const sfx_dd &a, double b){ sfx_dd p = two_prod(a.hi, b); p.lo += a.lo*b; return quick_two_sum(p.hi, p.lo); }
SFX_HD sfx_dd mul(const sfx_dd &a, const sfx_dd &b){
    sfx_dd p = two_prod(a.hi, b.hi); p.lo += a.hi*b.lo + a.lo*b.hi; return quick_two_sum(p.hi, p.lo);
}
SFX_HD sfx_dd neg(const sfx_dd &a){ return sfx_dd(-a.hi, -a.lo); }
SFX_HD sfx_dd div(const sfx_dd &a, const sfx_dd &b){
    double q1 = a.hi/b.hi; sfx_dd r = add(a, neg(mul(b, sfx_dd(q1))));
    double q2 = r.hi/b.hi; r = add(r, neg(mul(b, sfx_dd(q2))));
    double q3 = r.hi/b.hi;
    sfx_dd q = quick_two_sum(q1, q2); return add(q, sfx_dd(q3));
}
SFX_HD bool lt(const sfx_dd &a, const sfx_dd &b){ return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo); }
SFX_HD bool eq(const sfx_dd &a, const sfx_dd &b){ return a.hi == b.hi && a.lo == b.lo; }
}  // namespace sfxdd

SFX_HD sfx_dd operator+(const sfx_dd &a, const sfx_dd &b){ return sfxdd::add(a, b); }
SFX_HD sfx_dd operator-(const sfx_dd &a){ return sfxdd::neg(a); }
SFX_HD sfx_dd operator-(const sfx_dd &a, const sfx_dd &b){ return sfxdd::add(a, sfxdd::neg(b)); }
SFX_HD sfx_dd operator*(const sfx_dd &a, const sfx_dd &b){ return sfxdd::mul(a, b); }
SFX_HD sfx_dd operator/(const sfx_dd &a, const sfx_dd &b){ return sfxdd::div(a, b); }
SFX_HD sfx_dd &operator+=(sfx_dd &a, const sfx_dd &b){ a = a + b; return a; }
SFX_HD sfx_dd &operator-=(sfx_dd &a, const sfx_dd &b){ a = a - b; return a; }
SFX_HD sfx_dd &operator*=(sfx_dd &a, const sfx_dd &b){ a = a * b; return a; }
SFX_HD sfx_dd &operator/=(sfx_dd &a, const sfx_dd &b){ a = a / b; return a; }
SFX_HD bool operator<(const sfx_dd &a, const sfx_dd &b){ return sfxdd::lt(a, b); }
SFX_HD bool operator>(const sfx_dd &a, const sfx_dd &b){ return sfxdd::lt(b, a); }
SFX_HD bool operator<=(const sfx_dd &a, const sfx_dd &b){ return !sfxdd::lt(b, a); }
SFX_HD bool operator>=(const sfx_dd &a, const sfx_dd &b){ return !sfxdd::lt(a, b); }
SFX_HD bool operator==(const sfx_dd &a, const sfx_dd &b){ return sfxdd::eq(a, b); }
SFX_HD bool operator!=(const sfx_dd &a, const sfx_dd &b){ return !sfxdd::eq(a, b); }
// mixed forms: an int or double operand is not promoted for + - *, as in the reference (int / and dd / int
// promote the int, as there)
SFX_HD sfx_dd operator+(int a, const sfx_dd &b){ return sfxdd::add_d(b, (double)a); }
SFX_HD sfx_dd operator+(const sfx_dd &a, int b){ return sfxdd::add_d(a, (double)b); }
SFX_HD sfx_dd operator-(int a, const sfx_dd &b){ return sfxdd::add_d(-b, (double)a); }
SFX_HD sfx_dd operator-(const sfx_dd &a, int b){ return sfxdd::add_d(a, -(double)b); }
SFX_HD sfx_dd operator*(int a, const sfx_dd &b){ return sfxdd::mul_d(b, (double)a); }
SFX_HD sfx_dd operator*(const sfx_dd &a, int b){ return sfxdd::mul_d(a, (double)b); }
SFX_HD sfx_dd operator/(int a, const sfx_dd &b){ return sfx_dd((double)a) / b; }
SFX_HD sfx_dd operator/(const sfx_dd &a, int b){ return a / sfx_dd((double)b); }
SFX_HD sfx_dd operator+(double a, const sfx_dd &b){ return sfxdd::add_d(b, a); }
SFX_HD sfx_dd operator+(const sfx_dd &a, double b){ return sfxdd::add_d(a, b); }
SFX_HD sfx_dd operator-(double a, const sfx_dd &b){ return sfxdd::add_d(-b, a); }
SFX_HD sfx_dd operator-(const sfx_dd &a, double b){ return sfxdd::add_d(a, -b); }
SFX_HD sfx_dd operator*(double a, const sfx_dd &b){ return sfxdd::mul_d(b, a); }
SFX_HD sfx_dd operator*(const sfx_dd &a, double b){ return sfxdd::mul_d(a, b); }
SFX_HD bool operator<(const sfx_dd &a, int b){ return a < sfx_dd((double)b); }
SFX_HD bool operator>(const sfx_dd &a, int b){ return a > sfx_dd((double)b); }

// |x| as the kernels' RABS ((x) < 0 ? -(x) : (x)); for double the sign bit is cleared, which differs
// only for -0 (a zero either way in every sum it enters)
SFX_HD double sfx_abs(double x){ return fabs(x); }
SFX_HD sfx_dd sfx_abs(const sfx_dd &x){ return x < sfx_dd(0.0) ? -x : x; }
