// Forward-mode dual numbers for the drift adjoint: a value and N tangent
// components, with the elementary functions the drift step uses.  The
// derivative of each function is its textbook one; branch choices (floor,
// max, min, comparisons) follow the value, so a dual evaluation takes the
// same path as the plain double evaluation and its value parts are
// bit-identical to it.
#pragma once

#include <hip/hip_runtime.h>

namespace rrea_dev {

// the double functions stay visible beside the dual overloads below
using ::atan2;
using ::cos;
using ::exp;
using ::hypot;
using ::log;
using ::sin;
using ::sqrt;

template <int N>
struct Dual {
    double v;
    double d[N];
    __device__ Dual() : v(0.0) { for (int i = 0; i < N; ++i) d[i] = 0.0; }
    __device__ Dual(double x) : v(x) { for (int i = 0; i < N; ++i) d[i] = 0.0; }  // NOLINT: implicit constant
    // the independent variable `slot`
    __device__ static Dual seed(double x, int slot)
    {
        Dual r(x);
        r.d[slot] = 1.0;
        return r;
    }
};

// value of a double or a dual
__device__ __forceinline__ double val(double x) { return x; }
template <int N>
__device__ __forceinline__ double val(Dual<N> const& x) { return x.v; }

// a*x + b*y on the tangents (the chain rule's common shape)
template <int N>
__device__ __forceinline__ Dual<N> combine(double v, double a, Dual<N> const& x, double b, Dual<N> const& y)
{
    Dual<N> r(v);
    for (int i = 0; i < N; ++i) r.d[i] = a * x.d[i] + b * y.d[i];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> scale(double v, double a, Dual<N> const& x)
{
    Dual<N> r(v);
    for (int i = 0; i < N; ++i) r.d[i] = a * x.d[i];
    return r;
}

template <int N>
__device__ __forceinline__ Dual<N> operator+(Dual<N> const& a, Dual<N> const& b) { return combine(a.v + b.v, 1.0, a, 1.0, b); }
template <int N>
__device__ __forceinline__ Dual<N> operator-(Dual<N> const& a, Dual<N> const& b) { return combine(a.v - b.v, 1.0, a, -1.0, b); }
template <int N>
__device__ __forceinline__ Dual<N> operator*(Dual<N> const& a, Dual<N> const& b) { return combine(a.v * b.v, b.v, a, a.v, b); }
template <int N>
__device__ __forceinline__ Dual<N> operator/(Dual<N> const& a, Dual<N> const& b)
{
    double const q = a.v / b.v;
    return combine(q, 1.0 / b.v, a, -q / b.v, b);
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(Dual<N> const& a) { return scale(-a.v, -1.0, a); }

template <int N>
__device__ __forceinline__ Dual<N> operator+(Dual<N> const& a, double b) { Dual<N> r = a; r.v = a.v + b; return r; }
template <int N>
__device__ __forceinline__ Dual<N> operator+(double a, Dual<N> const& b) { Dual<N> r = b; r.v = a + b.v; return r; }
template <int N>
__device__ __forceinline__ Dual<N> operator-(Dual<N> const& a, double b) { Dual<N> r = a; r.v = a.v - b; return r; }
template <int N>
__device__ __forceinline__ Dual<N> operator-(double a, Dual<N> const& b) { return scale(a - b.v, -1.0, b); }
template <int N>
__device__ __forceinline__ Dual<N> operator*(Dual<N> const& a, double b) { return scale(a.v * b, b, a); }
template <int N>
__device__ __forceinline__ Dual<N> operator*(double a, Dual<N> const& b) { return scale(a * b.v, a, b); }
template <int N>
__device__ __forceinline__ Dual<N> operator/(Dual<N> const& a, double b) { return scale(a.v / b, 1.0 / b, a); }
template <int N>
__device__ __forceinline__ Dual<N> operator/(double a, Dual<N> const& b)
{
    double const q = a / b.v;
    return scale(q, -q / b.v, b);
}

template <int N>
__device__ __forceinline__ Dual<N> sqrt(Dual<N> const& a)
{
    double const s = ::sqrt(a.v);
    return scale(s, s > 0.0 ? 0.5 / s : 0.0, a);
}
template <int N>
__device__ __forceinline__ Dual<N> hypot(Dual<N> const& a, Dual<N> const& b)
{
    double const h = ::hypot(a.v, b.v);
    if (!(h > 0.0)) return Dual<N>(h);
    return combine(h, a.v / h, a, b.v / h, b);
}
template <int N>
__device__ __forceinline__ Dual<N> atan2(Dual<N> const& y, Dual<N> const& x)
{
    double const t = ::atan2(y.v, x.v);
    double const r2 = x.v * x.v + y.v * y.v;
    if (!(r2 > 0.0)) return Dual<N>(t);
    return combine(t, -y.v / r2, x, x.v / r2, y);
}
template <int N>
__device__ __forceinline__ Dual<N> sin(Dual<N> const& a) { return scale(::sin(a.v), ::cos(a.v), a); }
template <int N>
__device__ __forceinline__ Dual<N> cos(Dual<N> const& a) { return scale(::cos(a.v), -::sin(a.v), a); }
template <int N>
__device__ __forceinline__ Dual<N> log(Dual<N> const& a) { return scale(::log(a.v), 1.0 / a.v, a); }
template <int N>
__device__ __forceinline__ Dual<N> exp(Dual<N> const& a)
{
    double const e = ::exp(a.v);
    return scale(e, e, a);
}

// max / min pick an operand by value, as dmax / dmin do (a tie keeps the first)
__device__ __forceinline__ double tmax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double tmin(double a, double b) { return (b < a) ? b : a; }
template <int N>
__device__ __forceinline__ Dual<N> tmax(Dual<N> const& a, Dual<N> const& b) { return (a.v < b.v) ? b : a; }
template <int N>
__device__ __forceinline__ Dual<N> tmax(double a, Dual<N> const& b) { return (a < b.v) ? b : Dual<N>(a); }
template <int N>
__device__ __forceinline__ Dual<N> tmax(Dual<N> const& a, double b) { return (a.v < b) ? Dual<N>(b) : a; }
template <int N>
__device__ __forceinline__ Dual<N> tmin(Dual<N> const& a, Dual<N> const& b) { return (b.v < a.v) ? b : a; }
template <int N>
__device__ __forceinline__ Dual<N> tmin(Dual<N> const& a, double b) { return (b < a.v) ? Dual<N>(b) : a; }
template <int N>
__device__ __forceinline__ Dual<N> tmin(double a, Dual<N> const& b) { return (b.v < a) ? b : Dual<N>(a); }

}  // namespace rrea_dev
