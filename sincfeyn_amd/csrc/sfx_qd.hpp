// sfx_qd.hpp -- quad-double arithmetic (a value is the exact sum of four doubles, ~200 bits) for the
// all-QD error reference of the two-loop integrals, usable on gfx950 and on the host.
//
// Every operation forms the exact result as a floating-point expansion and rounds it back to four
// doubles:
//   * error-free transformations: two_sum (Knuth) and the fma two-product;
//   * the exact expansion of n terms by Shewchuk's Grow-Expansion without zero elimination (J. R.
//     Shewchuk, "Adaptive precision floating-point arithmetic and fast robust geometric predicates",
//     1997, Fig. 6): nonoverlapping components of increasing magnitude whose sum is exact; all indices
//     are compile-time constants, so the expansion lives in registers;
//   * rounding to four doubles by four cascaded error-free summations (the VecSum of Ogita, Rump and
//     Oishi, "Accurate sum and dot product", 2005), each leaving the next component as the rounded sum
//     of what the previous passes left; the rest is added into the fourth;
//   * a + b: the 8 components; a * b: the exact products a_i b_j with i + j <= 3 (value and fma error)
//     and the rounded ones with i + j = 4 (the rest are below 2^-260 of the product); a / b: five
//     quotient digits of long division, each remainder formed exactly.
// Measured against 420-bit arithmetic (the tests): every operation is correct to below 1e-64 of its
// operands' magnitude.  FP contraction must stay off in every translation unit that includes this
// header (the error-free transformations depend on it).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#define SFX_QHD __host__ __device__ __forceinline__
#define SFX_QOP static __host__ __device__ __noinline__   // the core operations: called, not inlined
#define SFX_QD_EPS 1e-62                                 // relative precision used in the fallback tests

struct sfx_qd {
    double x[4];
    SFX_QHD sfx_qd() : x{0, 0, 0, 0} {}
    SFX_QHD sfx_qd(double a) : x{a, 0, 0, 0} {}
    SFX_QHD sfx_qd(int a) : x{(double)a, 0, 0, 0} {}
    SFX_QHD sfx_qd(double a, double b, double c, double d) : x{a, b, c, d} {}
    SFX_QHD explicit operator double() const { return x[0] + (x[1] + (x[2] + x[3])); }
};

namespace sfxqd {
SFX_QHD void two_sum(double a, double b, double &s, double &e){ s = a + b; double bb = s - a; e = (a - (s - bb)) + (b - bb); }

// t[0..N-1] -> its exact expansion, in place (Grow-Expansion, one term at a time)
template<int N> SFX_QHD void expand(double *t){
#pragma unroll
    for(int k = 1; k < N; k++){
        double Q = t[k];
#pragma unroll
        for(int i = 0; i < k; i++){ double s, h; two_sum(Q, t[i], s, h); t[i] = h; Q = s; }
        t[k] = Q;
    }
}
// four cascaded error-free passes over e[0..N-1]; pass p leaves the rounded sum of e[0..N-1-p] in
// e[N-1-p] and the exact remainder below it
template<int N> SFX_QHD sfx_qd round4(double *e){
    sfx_qd r;
#pragma unroll
    for(int p = 0; p < 4; p++){
#pragma unroll
        for(int i = 0; i < N - 1 - p; i++){ double s, h; two_sum(e[i + 1], e[i], s, h); e[i + 1] = s; e[i] = h; }
        r.x[p] = e[N - 1 - p];
    }
    double rest = 0;
#pragma unroll
    for(int i = 0; i < N - 4; i++) rest += e[i];
    r.x[3] += rest;
    return r;
}
template<int N> SFX_QHD sfx_qd exact_sum(double *t){ expand<N>(t); return round4<N>(t); }

SFX_QOP sfx_qd add(sfx_qd a, sfx_qd b){
    double t[8] = {a.x[3], b.x[3], a.x[2], b.x[2], a.x[1], b.x[1], a.x[0], b.x[0]};
    return exact_sum<8>(t);
}
SFX_QOP sfx_qd mul(sfx_qd a, sfx_qd b){
    double t[23]; int n = 0;
#pragma unroll
    for(int s = 4; s >= 0; s--){
#pragma unroll
        for(int i = 0; i < 4; i++){
            int j = s - i;
            if(j < 0 || j > 3) continue;
            double p = a.x[i]*b.x[j];
            if(s == 4) t[n++] = p;
            else { t[n++] = fma(a.x[i], b.x[j], -p); t[n++] = p; }
        }
    }
    return exact_sum<23>(t);
}
// r - b q for a double q, exactly formed then rounded: the remainder step of the division
SFX_QHD sfx_qd sub_mul_d(sfx_qd r, sfx_qd b, double q){
    double t[12];
#pragma unroll
    for(int i = 0; i < 4; i++){ double p = -b.x[3 - i]*q; t[2*i] = fma(-b.x[3 - i], q, -p); t[2*i + 1] = p; }
#pragma unroll
    for(int i = 0; i < 4; i++) t[8 + i] = r.x[3 - i];
    return exact_sum<12>(t);
}
SFX_QHD sfx_qd renorm(const sfx_qd &a){ double t[4] = {a.x[3], a.x[2], a.x[1], a.x[0]}; return exact_sum<4>(t); }
SFX_QOP sfx_qd div(sfx_qd a, sfx_qd b){
    // the leading components must carry the values: both operands are rounded to four doubles first
    double q[5];
    sfx_qd r = renorm(a);
    b = renorm(b);
#pragma unroll
    for(int k = 0; k < 5; k++){
        q[k] = r.x[0]/b.x[0];
        if(k < 4) r = sub_mul_d(r, b, q[k]);
    }
    double t[5] = {q[4], q[3], q[2], q[1], q[0]};
    return exact_sum<5>(t);
}
SFX_QHD sfx_qd neg(const sfx_qd &a){ return sfx_qd(-a.x[0], -a.x[1], -a.x[2], -a.x[3]); }
// the sign of an exact sum: its rounded leading component
SFX_QHD int sign(const sfx_qd &a){ sfx_qd r = add(a, sfx_qd()); return (r.x[0] > 0) - (r.x[0] < 0); }
}  // namespace sfxqd

SFX_QHD sfx_qd operator+(const sfx_qd &a, const sfx_qd &b){ return sfxqd::add(a, b); }
SFX_QHD sfx_qd operator-(const sfx_qd &a){ return sfxqd::neg(a); }
SFX_QHD sfx_qd operator-(const sfx_qd &a, const sfx_qd &b){ return sfxqd::add(a, sfxqd::neg(b)); }
SFX_QHD sfx_qd operator*(const sfx_qd &a, const sfx_qd &b){ return sfxqd::mul(a, b); }
SFX_QHD sfx_qd operator/(const sfx_qd &a, const sfx_qd &b){ return sfxqd::div(a, b); }
SFX_QHD sfx_qd &operator+=(sfx_qd &a, const sfx_qd &b){ a = a + b; return a; }
SFX_QHD sfx_qd &operator-=(sfx_qd &a, const sfx_qd &b){ a = a - b; return a; }
SFX_QHD sfx_qd &operator*=(sfx_qd &a, const sfx_qd &b){ a = a * b; return a; }
SFX_QHD sfx_qd &operator/=(sfx_qd &a, const sfx_qd &b){ a = a / b; return a; }
SFX_QHD bool operator<(const sfx_qd &a, const sfx_qd &b){ return sfxqd::sign(a - b) < 0; }
SFX_QHD bool operator>(const sfx_qd &a, const sfx_qd &b){ return sfxqd::sign(a - b) > 0; }
SFX_QHD bool operator<=(const sfx_qd &a, const sfx_qd &b){ return sfxqd::sign(a - b) <= 0; }
SFX_QHD bool operator>=(const sfx_qd &a, const sfx_qd &b){ return sfxqd::sign(a - b) >= 0; }
// an int or double operand is exact as a quad-double: promoted
SFX_QHD sfx_qd operator+(int a, const sfx_qd &b){ return sfx_qd(a) + b; }
SFX_QHD sfx_qd operator+(const sfx_qd &a, int b){ return a + sfx_qd(b); }
SFX_QHD sfx_qd operator-(int a, const sfx_qd &b){ return sfx_qd(a) - b; }
SFX_QHD sfx_qd operator-(const sfx_qd &a, int b){ return a - sfx_qd(b); }
SFX_QHD sfx_qd operator*(int a, const sfx_qd &b){ return sfx_qd(a)*b; }
SFX_QHD sfx_qd operator*(const sfx_qd &a, int b){ return a*sfx_qd(b); }
SFX_QHD sfx_qd operator/(int a, const sfx_qd &b){ return sfx_qd(a)/b; }
SFX_QHD sfx_qd operator/(const sfx_qd &a, int b){ return a/sfx_qd(b); }
SFX_QHD sfx_qd operator+(double a, const sfx_qd &b){ return sfx_qd(a) + b; }
SFX_QHD sfx_qd operator+(const sfx_qd &a, double b){ return a + sfx_qd(b); }
SFX_QHD sfx_qd operator-(double a, const sfx_qd &b){ return sfx_qd(a) - b; }
SFX_QHD sfx_qd operator-(const sfx_qd &a, double b){ return a - sfx_qd(b); }
SFX_QHD sfx_qd operator*(double a, const sfx_qd &b){ return sfx_qd(a)*b; }
SFX_QHD sfx_qd operator*(const sfx_qd &a, double b){ return a*sfx_qd(b); }
SFX_QHD sfx_qd operator/(double a, const sfx_qd &b){ return sfx_qd(a)/b; }
SFX_QHD sfx_qd operator/(const sfx_qd &a, double b){ return a/sfx_qd(b); }
SFX_QHD bool operator<(const sfx_qd &a, int b){ return a < sfx_qd(b); }
SFX_QHD bool operator>(const sfx_qd &a, int b){ return a > sfx_qd(b); }
SFX_QHD sfx_qd sfx_abs(const sfx_qd &x){ return sfxqd::sign(x) < 0 ? -x : x; }
