// dual_dev.hpp -- fixed-width forward-mode dual numbers for device code.
//
// DualV<W> carries a value and W directional derivatives in registers; the
// Jacobian kernels seed W unknowns of one block at a time (a "chunk") and
// read the W derivative columns of every residual row.  Branches compare
// values only, so each derivative is that of the branch the value takes.
#pragma once
#include <hip/hip_runtime.h>

namespace osr {

template <int W>
struct DualV {
    double v;
    double d[W];
    __device__ __forceinline__ DualV() : v(0) {
#pragma unroll
        for (int k = 0; k < W; k++) d[k] = 0;
    }
    __device__ __forceinline__ DualV(double x) : v(x) {
#pragma unroll
        for (int k = 0; k < W; k++) d[k] = 0;
    }
};

#define OSR_DUAL_BIN(op, expr_v, ca, cb)                                              \
    template <int W>                                                                  \
    __device__ __forceinline__ DualV<W> operator op(const DualV<W>& a, const DualV<W>& b) { \
        DualV<W> r;                                                                   \
        r.v = expr_v;                                                                 \
        _Pragma("unroll") for (int k = 0; k < W; k++) r.d[k] = ca * a.d[k] + cb * b.d[k]; \
        return r;                                                                     \
    }
OSR_DUAL_BIN(+, a.v + b.v, 1.0, 1.0)
OSR_DUAL_BIN(-, a.v - b.v, 1.0, -1.0)
OSR_DUAL_BIN(*, a.v * b.v, b.v, a.v)
#undef OSR_DUAL_BIN

template <int W>
__device__ __forceinline__ DualV<W> operator/(const DualV<W>& a, const DualV<W>& b) {
    DualV<W> r;
    const double ib = 1.0 / b.v;
    r.v = a.v * ib;
#pragma unroll
    for (int k = 0; k < W; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
template <int W>
__device__ __forceinline__ DualV<W> operator-(const DualV<W>& a) {
    DualV<W> r;
    r.v = -a.v;
#pragma unroll
    for (int k = 0; k < W; k++) r.d[k] = -a.d[k];
    return r;
}
// scalar-dual mixes avoid building a zero-derivative dual
template <int W> __device__ __forceinline__ DualV<W> operator+(const DualV<W>& a, double b) { DualV<W> r = a; r.v += b; return r; }
template <int W> __device__ __forceinline__ DualV<W> operator+(double b, const DualV<W>& a) { DualV<W> r = a; r.v += b; return r; }
template <int W> __device__ __forceinline__ DualV<W> operator-(const DualV<W>& a, double b) { DualV<W> r = a; r.v -= b; return r; }
template <int W> __device__ __forceinline__ DualV<W> operator-(double b, const DualV<W>& a) {
    DualV<W> r;
    r.v = b - a.v;
#pragma unroll
    for (int k = 0; k < W; k++) r.d[k] = -a.d[k];
    return r;
}
template <int W> __device__ __forceinline__ DualV<W> operator*(const DualV<W>& a, double b) {
    DualV<W> r;
    r.v = a.v * b;
#pragma unroll
    for (int k = 0; k < W; k++) r.d[k] = a.d[k] * b;
    return r;
}
template <int W> __device__ __forceinline__ DualV<W> operator*(double b, const DualV<W>& a) { return a * b; }
template <int W> __device__ __forceinline__ DualV<W> operator/(const DualV<W>& a, double b) { return a * (1.0 / b); }
template <int W> __device__ __forceinline__ DualV<W> operator/(double a, const DualV<W>& b) {
    DualV<W> r;
    const double ib = 1.0 / b.v;
    r.v = a * ib;
#pragma unroll
    for (int k = 0; k < W; k++) r.d[k] = -r.v * b.d[k] * ib;
    return r;
}
template <int W> __device__ __forceinline__ DualV<W>& operator+=(DualV<W>& a, const DualV<W>& b) { return a = a + b; }
template <int W> __device__ __forceinline__ DualV<W>& operator-=(DualV<W>& a, const DualV<W>& b) { return a = a - b; }
template <int W> __device__ __forceinline__ DualV<W>& operator+=(DualV<W>& a, double b) { a.v += b; return a; }

#define OSR_DUAL_CMP(op)                                                                                     \
    template <int W> __device__ __forceinline__ bool operator op(const DualV<W>& a, const DualV<W>& b) { return a.v op b.v; } \
    template <int W> __device__ __forceinline__ bool operator op(const DualV<W>& a, double b) { return a.v op b; }          \
    template <int W> __device__ __forceinline__ bool operator op(double a, const DualV<W>& b) { return a op b.v; }
OSR_DUAL_CMP(<)
OSR_DUAL_CMP(>)
OSR_DUAL_CMP(<=)
OSR_DUAL_CMP(>=)
OSR_DUAL_CMP(==)
OSR_DUAL_CMP(!=)
#undef OSR_DUAL_CMP

template <int W>
__device__ __forceinline__ DualV<W> dchain(const DualV<W>& a, double fv, double fp) {
    DualV<W> r;
    r.v = fv;
#pragma unroll
    for (int k = 0; k < W; k++) r.d[k] = fp * a.d[k];
    return r;
}

// Elementary functions, overloaded with the double versions under the same
// names so templated code can call fsqrt(x) etc. for either type.
__device__ __forceinline__ double fsqrt(double a) { return sqrt(a); }
__device__ __forceinline__ double flog(double a) { return log(a); }
__device__ __forceinline__ double fexp(double a) { return exp(a); }
__device__ __forceinline__ double fcos(double a) { return cos(a); }
__device__ __forceinline__ double facos(double a) { return acos(a); }
__device__ __forceinline__ double fpow(double a, double p) { return pow(a, p); }
__device__ __forceinline__ double fcbrt(double a) { return cbrt(a); }
__device__ __forceinline__ double dval(double a) { return a; }

template <int W> __device__ __forceinline__ DualV<W> fsqrt(const DualV<W>& a) { const double s = sqrt(a.v); return dchain(a, s, 0.5 / s); }
template <int W> __device__ __forceinline__ DualV<W> flog(const DualV<W>& a) { return dchain(a, log(a.v), 1.0 / a.v); }
template <int W> __device__ __forceinline__ DualV<W> fexp(const DualV<W>& a) { const double e = exp(a.v); return dchain(a, e, e); }
template <int W> __device__ __forceinline__ DualV<W> fcos(const DualV<W>& a) {
    double s, c;
    sincos(a.v, &s, &c);
    return dchain(a, c, -s);
}
template <int W> __device__ __forceinline__ DualV<W> facos(const DualV<W>& a) {
    return dchain(a, acos(a.v), -1.0 / sqrt(1 - a.v * a.v));
}
template <int W> __device__ __forceinline__ DualV<W> fpow(const DualV<W>& a, double p) {
    const double f = pow(a.v, p);
    return dchain(a, f, p * f / a.v);
}
template <int W> __device__ __forceinline__ DualV<W> fcbrt(const DualV<W>& a) {
    const double f = cbrt(a.v);
    return dchain(a, f, f / (3 * a.v));
}
template <int W> __device__ __forceinline__ double dval(const DualV<W>& a) { return a.v; }

}  // namespace osr
