// dual.hpp -- forward-mode dual numbers for the oracle's Jacobian.
//
// TEST INFRASTRUCTURE ONLY.  Dual<W> carries a value and up to W partial
// derivatives (runtime width n <= W); every elementary function applies its
// textbook derivative.  Branches compare values only, so a derivative is the
// one-sided derivative of the branch the value takes (frozen upstream
// weighting, clamps, table segments) -- the derivative a hand-written
// analytic Jacobian of the same residual produces.
#pragma once
#include <cmath>

namespace orad {

template <int W>
struct Dual {
    double v = 0;
    double d[W] = {};
    int n = 0;
    Dual() = default;
    Dual(double x) : v(x) {}
};

template <int W>
inline int wid(const Dual<W>& a, const Dual<W>& b) { return a.n > b.n ? a.n : b.n; }

#define ORAD_BIN(op, val, da, db)                                   \
    template <int W>                                                \
    inline Dual<W> operator op(const Dual<W>& a, const Dual<W>& b) { \
        Dual<W> r;                                                  \
        r.n = wid(a, b);                                            \
        r.v = val;                                                  \
        for (int k = 0; k < r.n; k++) r.d[k] = da * a.d[k] + db * b.d[k]; \
        return r;                                                   \
    }
ORAD_BIN(+, a.v + b.v, 1.0, 1.0)
ORAD_BIN(-, a.v - b.v, 1.0, -1.0)
ORAD_BIN(*, a.v * b.v, b.v, a.v)
#undef ORAD_BIN

template <int W>
inline Dual<W> operator/(const Dual<W>& a, const Dual<W>& b) {
    Dual<W> r;
    r.n = wid(a, b);
    r.v = a.v / b.v;
    const double ib = 1.0 / b.v;
    for (int k = 0; k < r.n; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
template <int W> inline Dual<W> operator-(const Dual<W>& a) {
    Dual<W> r = a;
    r.v = -a.v;
    for (int k = 0; k < r.n; k++) r.d[k] = -a.d[k];
    return r;
}
template <int W> inline Dual<W>& operator+=(Dual<W>& a, const Dual<W>& b) { return a = a + b; }
template <int W> inline Dual<W>& operator-=(Dual<W>& a, const Dual<W>& b) { return a = a - b; }
template <int W> inline Dual<W>& operator*=(Dual<W>& a, const Dual<W>& b) { return a = a * b; }
template <int W> inline Dual<W>& operator/=(Dual<W>& a, const Dual<W>& b) { return a = a / b; }

#define ORAD_MIX(op)                                                                          \
    template <int W> inline Dual<W> operator op(const Dual<W>& a, double b) { return a op Dual<W>(b); } \
    template <int W> inline Dual<W> operator op(double a, const Dual<W>& b) { return Dual<W>(a) op b; }
ORAD_MIX(+)
ORAD_MIX(-)
ORAD_MIX(*)
ORAD_MIX(/)
#undef ORAD_MIX

#define ORAD_CMP(op)                                                                  \
    template <int W> inline bool operator op(const Dual<W>& a, const Dual<W>& b) { return a.v op b.v; } \
    template <int W> inline bool operator op(const Dual<W>& a, double b) { return a.v op b; }        \
    template <int W> inline bool operator op(double a, const Dual<W>& b) { return a op b.v; }
ORAD_CMP(<)
ORAD_CMP(>)
ORAD_CMP(<=)
ORAD_CMP(>=)
ORAD_CMP(==)
ORAD_CMP(!=)
#undef ORAD_CMP

template <int W>
inline Dual<W> chain(const Dual<W>& a, double fv, double fp) {
    Dual<W> r;
    r.n = a.n;
    r.v = fv;
    for (int k = 0; k < r.n; k++) r.d[k] = fp * a.d[k];
    return r;
}
template <int W> inline Dual<W> sqrt(const Dual<W>& a) { double s = std::sqrt(a.v); return chain(a, s, 0.5 / s); }
template <int W> inline Dual<W> log(const Dual<W>& a) { return chain(a, std::log(a.v), 1.0 / a.v); }
template <int W> inline Dual<W> exp(const Dual<W>& a) { double e = std::exp(a.v); return chain(a, e, e); }
template <int W> inline Dual<W> cos(const Dual<W>& a) { return chain(a, std::cos(a.v), -std::sin(a.v)); }
template <int W> inline Dual<W> acos(const Dual<W>& a) {
    return chain(a, std::acos(a.v), -1.0 / std::sqrt(1 - a.v * a.v));
}
template <int W> inline Dual<W> pow(const Dual<W>& a, double p) {
    double f = std::pow(a.v, p);
    return chain(a, f, p * std::pow(a.v, p - 1));
}
template <int W> inline Dual<W> fabs(const Dual<W>& a) { return a.v < 0 ? -a : a; }

template <int W> inline double val(const Dual<W>& a) { return a.v; }
inline double val(double a) { return a; }

}  // namespace orad
