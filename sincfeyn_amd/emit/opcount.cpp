// opcount.cpp -- per-point FP64 operation counts of one emitted integrand (host tool, built and run by
// sincfeyn_amd/build.py).  The emitted body k_<graph>.inc is instantiated with REAL = an op-counting
// double: every binary + - * / (with a double or int operand too, compound assignments included) counts
// one operation, division separately; negation, comparison and |x| count nothing.  The integrand is run
// at NP points s = e^z, z uniform in [-3, 3] (fixed seed), with the |.| sums on (SQ_AB = 1) and off
// (SQ_AB = 0); the counts can depend on the point where a kernel branches on values (v3_p4*, w3_*).
//   g++ -O1 -DSFX_INC='"k_g3_00.inc"' -DSFX_F=f_g3_00 -DSFX_D=7 opcount.cpp -> prints one JSON object
#include <cmath>
#include <cstdio>
#include <cstdint>
#define __constant__
#define SFX_FN inline
#define SFX_CHUNK inline
#define SFX_CHUNK_INLINE(R) true
// the straight-line point function (the source's operations); the node-split body of the FP64 kernels
// performs the same operations per node without the subexpressions shared between nodes
#define SFX_NODE_SPLIT(R) false
static inline void sfx_opaque(double &) {}
static inline void sfx_opaque_s(int &) {}
static long long N_ADD, N_MUL, N_DIV;
struct oc {
    double v;
    oc() : v(0) {}
    oc(double x) : v(x) {}
    explicit operator double() const { return v; }
};
static inline oc operator+(oc a, oc b){ N_ADD++; return oc(a.v + b.v); }
static inline oc operator-(oc a, oc b){ N_ADD++; return oc(a.v - b.v); }
static inline oc operator*(oc a, oc b){ N_MUL++; return oc(a.v*b.v); }
static inline oc operator/(oc a, oc b){ N_DIV++; return oc(a.v/b.v); }
static inline oc operator-(oc a){ return oc(-a.v); }
static inline oc operator+(int a, oc b){ return oc((double)a) + b; }
static inline oc operator+(oc a, int b){ return a + oc((double)b); }
static inline oc operator-(int a, oc b){ return oc((double)a) - b; }
static inline oc operator-(oc a, int b){ return a - oc((double)b); }
static inline oc operator*(int a, oc b){ return oc((double)a)*b; }
static inline oc operator*(oc a, int b){ return a*oc((double)b); }
static inline oc operator/(int a, oc b){ return oc((double)a)/b; }
static inline oc operator/(oc a, int b){ return a/oc((double)b); }
static inline oc &operator+=(oc &a, oc b){ a = a + b; return a; }
static inline oc &operator-=(oc &a, oc b){ a = a - b; return a; }
static inline oc &operator*=(oc &a, oc b){ a = a*b; return a; }
static inline oc &operator/=(oc &a, oc b){ a = a/b; return a; }
static inline bool operator<(oc a, oc b){ return a.v < b.v; }
static inline bool operator>(oc a, oc b){ return a.v > b.v; }
static inline bool operator<=(oc a, oc b){ return a.v <= b.v; }
static inline bool operator>=(oc a, oc b){ return a.v >= b.v; }
static inline bool operator==(oc a, oc b){ return a.v == b.v; }
static inline bool operator!=(oc a, oc b){ return a.v != b.v; }
static inline oc sfx_abs(oc x){ return oc(std::fabs(x.v)); }

namespace K {
#include SFX_INC
}

int main(){
    const int NP = 256;
    uint64_t y = 0x2545F4914F6CDD1DULL;
    double tot[2][3] = {{0, 0, 0}, {0, 0, 0}};
    for(int p = 0; p < NP; p++){
        oc sf[SFX_D], ab;
        for(int d = 0; d < SFX_D; d++){
            y ^= y << 13; y ^= y >> 7; y ^= y << 17;
            double z = -3.0 + 6.0*(double)(y >> 11)/9007199254740992.0;
            sf[d] = oc(std::exp(z));
        }
        for(int ab_on = 0; ab_on <= 1; ab_on++){
            N_ADD = N_MUL = N_DIV = 0;
            (void)K::SFX_F<oc>(sf, &ab, ab_on);
            tot[ab_on][0] += N_ADD; tot[ab_on][1] += N_MUL; tot[ab_on][2] += N_DIV;
        }
    }
    printf("{\"add1\": %.2f, \"mul1\": %.2f, \"div1\": %.2f, \"add0\": %.2f, \"mul0\": %.2f, \"div0\": %.2f, "
           "\"ops1\": %.2f, \"ops0\": %.2f}\n",
           tot[1][0]/NP, tot[1][1]/NP, tot[1][2]/NP, tot[0][0]/NP, tot[0][1]/NP, tot[0][2]/NP,
           (tot[1][0] + tot[1][1] + tot[1][2])/NP, (tot[0][0] + tot[0][1] + tot[0][2])/NP);
    return 0;
}
