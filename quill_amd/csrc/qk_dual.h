// qk_dual.h -- forward-mode dual numbers for the adjoint kernels.
//
// Dual<N, T> carries a value and its partial derivatives with respect to N
// chosen inputs, all of type T.  The adjoint of one particle's push runs the
// forward code of the step (gather, pusher, displacement, deposit, all
// templated on the arithmetic type) on Dual numbers seeded at the inputs
// whose adjoints are wanted.  Branches (face crossings) follow the value, so
// the derivative is the one of the branch the forward step took.  Values are
// computed with the same operations as the forward code (the hardware
// reciprocal and square root in FP32), so the recomputed branches match.
#pragma once

#include "qk_device.h"

namespace qk {

// the library functions stay visible next to the Dual overloads below
using ::fabs;
using ::floor;
using ::sqrt;

template <int N, typename T>
struct Dual {
    T v;
    T d[N];

    __device__ Dual() : v(0)
    {
#pragma unroll
        for (int k = 0; k < N; ++k) d[k] = 0;
    }
    // constants (literals, integer casts, field values) carry no derivative
    __device__ Dual(double x) : v((T)x)
    {
#pragma unroll
        for (int k = 0; k < N; ++k) d[k] = 0;
    }
    __device__ Dual(float x) : Dual((double)x) {}
    __device__ Dual(int x) : Dual((double)x) {}
    // the k-th input variable
    __device__ static Dual var(T x, int k)
    {
        Dual r((double)x);
        r.v = x;
        r.d[k] = 1;
        return r;
    }

    __device__ Dual &operator+=(const Dual &b)
    {
        v += b.v;
#pragma unroll
        for (int k = 0; k < N; ++k) d[k] += b.d[k];
        return *this;
    }
    __device__ Dual &operator-=(const Dual &b)
    {
        v -= b.v;
#pragma unroll
        for (int k = 0; k < N; ++k) d[k] -= b.d[k];
        return *this;
    }
};

template <int N, typename T>
__device__ inline Dual<N, T> operator-(const Dual<N, T> &a)
{
    Dual<N, T> r;
    r.v = -a.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
    return r;
}

template <int N, typename T>
__device__ inline Dual<N, T> operator+(const Dual<N, T> &a, const Dual<N, T> &b)
{
    Dual<N, T> r = a;
    r += b;
    return r;
}

template <int N, typename T>
__device__ inline Dual<N, T> operator-(const Dual<N, T> &a, const Dual<N, T> &b)
{
    Dual<N, T> r = a;
    r -= b;
    return r;
}

template <int N, typename T>
__device__ inline Dual<N, T> operator*(const Dual<N, T> &a, const Dual<N, T> &b)
{
    Dual<N, T> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}

template <int N, typename T>
__device__ inline Dual<N, T> operator/(const Dual<N, T> &a, const Dual<N, T> &b)
{
    Dual<N, T> r;
    r.v = a.v / b.v;
    const T ib = T(1) / b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}

template <int N, typename T>
__device__ inline bool operator<(const Dual<N, T> &a, const Dual<N, T> &b) { return a.v < b.v; }
template <int N, typename T>
__device__ inline bool operator>(const Dual<N, T> &a, const Dual<N, T> &b) { return a.v > b.v; }
template <int N, typename T>
__device__ inline bool operator<=(const Dual<N, T> &a, const Dual<N, T> &b) { return a.v <= b.v; }
template <int N, typename T>
__device__ inline bool operator>=(const Dual<N, T> &a, const Dual<N, T> &b) { return a.v >= b.v; }
template <int N, typename T>
__device__ inline bool operator!=(const Dual<N, T> &a, const Dual<N, T> &b) { return a.v != b.v; }

// 1/a with the forward code's reciprocal for the value
template <int N, typename T>
__device__ inline Dual<N, T> qk_rcp(const Dual<N, T> &a)
{
    Dual<N, T> r;
    r.v = qk_rcp(a.v);
    const T m = -r.v * r.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * m;
    return r;
}

template <int N, typename T>
__device__ inline Dual<N, T> qk_sqrt(const Dual<N, T> &a)
{
    Dual<N, T> r;
    r.v = qk_sqrt(a.v);
    const T h = r.v > T(0) ? T(0.5) / r.v : T(0);
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * h;
    return r;
}

template <int N, typename T>
__device__ inline Dual<N, T> sqrt(const Dual<N, T> &a)
{
    Dual<N, T> r;
    r.v = ::sqrt(a.v);
    const T h = r.v > T(0) ? T(0.5) / r.v : T(0);
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * h;
    return r;
}

// piecewise constant: no derivative
template <int N, typename T>
__device__ inline T floor(const Dual<N, T> &a)
{
    return ::floor(a.v);
}

template <int N, typename T>
__device__ inline Dual<N, T> fabs(const Dual<N, T> &a)
{
    return a.v < T(0) ? -a : a;
}

}  // namespace qk
