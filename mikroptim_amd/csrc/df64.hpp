// df64: double-word ("double-float") numbers x = hi + lo of two floats with |lo| <= ulp(hi) / 2,
// about 48 significant bits (unit roundoff u^2 = 2^-48) at FP32 arithmetic. The number type of
// the reference's Precision::Df64 trace (include/mikroptim/math/df64.hpp:108).
//
// Algorithms (error bounds relative, u = 2^-24), the ones the reference names
// (include/mikroptim/math/df64.hpp:8-16), from the published double-word literature:
//   a + b, a - b  AccurateDWPlusDW     M. Joldes, J.-M. Muller, V. Popescu, "Tight and rigorous
//                                      error bounds for basic building blocks of double-word
//                                      arithmetic", ACM TOMS 44 (2017), Algorithm 6: 3u^2 + 13u^3
//   a * b         DWTimesDW3 (FMA)     ibid., Algorithm 12: 4u^2
//   a * f         DWTimesFP3 (FMA)     ibid., Algorithm 9: 2u^2
//   a / b         DWDivDW2             ibid., Algorithm 17: 15u^2 + 56u^3
//   sqrt(a)       SQRTDWtoDW           V. Lefevre, N. Louvet, J.-M. Muller, J. Picot, L. Rideau,
//                                      "Accurate calculation of Euclidean norms using
//                                      double-word arithmetic", ACM TOMS 49 (2023): 25/8 u^2
// Every error-free transformation is exact, so results are bitwise those of any other
// implementation of the same algorithms (checked against the oracle's host restatement).
// The error-free transformations (TwoSum, Fast2Sum, TwoProd) are exact only if every float
// operation is rounded on its own: every function below switches FMA contraction off, and
// products that must stay fused use fmaf explicitly.
//
// On MI355X the FP64 vector rate equals the non-packed FP32 rate, so df64 (10-20 FP32 operations
// per operation) is slower than native FP64 there; it is kept as the reference's second number
// type for the RESIDUALS / MERIT traces and measured against FP64.
#pragma once

#include <hip/hip_runtime.h>

#include <math.h>

namespace mkod {

struct df64 {
    float hi, lo;
};

namespace df64_detail {

__host__ __device__ inline df64 fast_two_sum(float a, float b) { // |a| >= |b| or a == 0
#pragma clang fp contract(off)
    const float s = a + b;
    const float z = s - a;
    return df64{s, b - z};
}

__host__ __device__ inline df64 two_sum(float a, float b) {
#pragma clang fp contract(off)
    const float s = a + b;
    const float ap = s - b;
    const float bp = s - ap;
    const float da = a - ap;
    const float db = b - bp;
    return df64{s, da + db};
}

__host__ __device__ inline df64 two_prod(float a, float b) {
#pragma clang fp contract(off)
    const float p = a * b;
    return df64{p, fmaf(a, b, -p)};
}

} // namespace df64_detail

// hi = float(x), lo = float(x - hi) (x - hi is exact in double); +-inf and NaN give lo = 0.
__host__ __device__ inline df64 df64_from(double x) {
#pragma clang fp contract(off)
    const float hi = float(x);
    const double rest = x - double(hi);
    return df64{hi, rest - rest == 0.0 ? float(rest) : 0.0f};
}
__host__ __device__ inline double df64_to(df64 x) { return double(x.hi) + double(x.lo); }

// AccurateDWPlusDW (Algorithm 6)
__host__ __device__ inline df64 operator+(df64 x, df64 y) {
#pragma clang fp contract(off)
    const df64 s = df64_detail::two_sum(x.hi, y.hi);
    const df64 t = df64_detail::two_sum(x.lo, y.lo);
    const float c = s.lo + t.hi;
    const df64 v = df64_detail::fast_two_sum(s.hi, c);
    const float w = t.lo + v.lo;
    return df64_detail::fast_two_sum(v.hi, w);
}
__host__ __device__ inline df64 operator-(df64 x) { return df64{-x.hi, -x.lo}; }
__host__ __device__ inline df64 operator-(df64 x, df64 y) { return x + (-y); }

// DWTimesDW3 (Algorithm 12)
__host__ __device__ inline df64 operator*(df64 x, df64 y) {
#pragma clang fp contract(off)
    const df64 c = df64_detail::two_prod(x.hi, y.hi);
    const float tl0 = x.lo * y.lo;
    const float tl1 = fmaf(x.hi, y.lo, tl0);
    const float cl2 = fmaf(x.lo, y.hi, tl1);
    const float cl3 = c.lo + cl2;
    return df64_detail::fast_two_sum(c.hi, cl3);
}

// DWTimesFP3 (Algorithm 9)
__host__ __device__ inline df64 mul_fp(df64 x, float y) {
#pragma clang fp contract(off)
    const df64 c = df64_detail::two_prod(x.hi, y);
    const float cl3 = fmaf(x.lo, y, c.lo);
    return df64_detail::fast_two_sum(c.hi, cl3);
}

// DWDivDW2 (Algorithm 17)
__host__ __device__ inline df64 operator/(df64 x, df64 y) {
#pragma clang fp contract(off)
    const float th = x.hi / y.hi;
    const df64 r = mul_fp(y, th);
    const float pih = x.hi - r.hi;
    const float dl = x.lo - r.lo;
    const float d = pih + dl;
    const float tl = d / y.hi;
    return df64_detail::fast_two_sum(th, tl);
}

// SQRTDWtoDW (Lefevre et al. 2023); sqrt(0) = 0, negative arguments give NaN, inf stays inf.
__host__ __device__ inline df64 sqrt(df64 x) {
#pragma clang fp contract(off)
    if (!(x.hi > 0.0f) || x.hi > 3.40282347e38f) return df64{sqrtf(x.hi), 0.0f};
    const float sh = sqrtf(x.hi);
    const float rh = fmaf(-sh, sh, x.hi); // x.hi - sh^2, exact
    const float rl = x.lo + rh;
    const float sl = rl / (sh + sh);
    return df64_detail::fast_two_sum(sh, sl);
}

__host__ __device__ inline df64& operator+=(df64& a, df64 b) { return a = a + b; }
__host__ __device__ inline df64& operator-=(df64& a, df64 b) { return a = a - b; }
__host__ __device__ inline df64& operator*=(df64& a, df64 b) { return a = a * b; }

// Comparisons of normalised values: by hi, then lo.
__host__ __device__ inline bool operator<(df64 a, df64 b) {
    return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo);
}
__host__ __device__ inline bool operator>(df64 a, df64 b) { return b < a; }
__host__ __device__ inline bool operator<=(df64 a, df64 b) {
    return a.hi < b.hi || (a.hi == b.hi && a.lo <= b.lo);
}
__host__ __device__ inline bool operator>=(df64 a, df64 b) { return b <= a; }

} // namespace mkod
