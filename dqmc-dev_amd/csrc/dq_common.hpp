// dq_common.hpp -- scalar types and helpers shared by the DQMC kernels.
//
// Two scalar types: double (the reference's real build) and zd, an
// interleaved (re, im) pair of doubles (the reference's USE_CPLX build,
// src/util.h:36-40).  Every kernel is a template over the scalar type.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct zd {
    double x, y;
};

__host__ __device__ inline zd mkz(double a, double b) { zd r; r.x = a; r.y = b; return r; }
__host__ __device__ inline zd operator+(zd a, zd b) { return mkz(a.x + b.x, a.y + b.y); }
__host__ __device__ inline zd operator-(zd a, zd b) { return mkz(a.x - b.x, a.y - b.y); }
__host__ __device__ inline zd operator-(zd a) { return mkz(-a.x, -a.y); }
__host__ __device__ inline zd operator*(zd a, zd b)
{
    return mkz(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline zd operator*(double s, zd a) { return mkz(s * a.x, s * a.y); }
__host__ __device__ inline zd operator*(zd a, double s) { return mkz(s * a.x, s * a.y); }
__host__ __device__ inline zd operator+(double s, zd a) { return mkz(s + a.x, a.y); }
__host__ __device__ inline zd operator+(zd a, double s) { return mkz(a.x + s, a.y); }
__host__ __device__ inline zd operator-(double s, zd a) { return mkz(s - a.x, -a.y); }
__host__ __device__ inline zd operator-(zd a, double s) { return mkz(a.x - s, a.y); }
__host__ __device__ inline zd& operator+=(zd& a, zd b) { a.x += b.x; a.y += b.y; return a; }
__host__ __device__ inline zd& operator-=(zd& a, zd b) { a.x -= b.x; a.y -= b.y; return a; }
__host__ __device__ inline zd& operator*=(zd& a, zd b) { a = a * b; return a; }
__host__ __device__ inline zd& operator*=(zd& a, double s) { a.x *= s; a.y *= s; return a; }

// complex division a/b (Smith's algorithm: avoids overflow in |b|^2)
__host__ __device__ inline zd zdiv(zd a, zd b)
{
    if (fabs(b.x) >= fabs(b.y)) {
        const double r = b.y / b.x, den = b.x + b.y * r;
        return mkz((a.x + a.y * r) / den, (a.y - a.x * r) / den);
    }
    const double r = b.x / b.y, den = b.y + b.x * r;
    return mkz((a.x * r + a.y) / den, (a.y * r - a.x) / den);
}
__host__ __device__ inline zd operator/(zd a, zd b) { return zdiv(a, b); }

// generic helpers ----------------------------------------------------------
__host__ __device__ inline double re(double a) { return a; }
__host__ __device__ inline double re(zd a) { return a.x; }
__host__ __device__ inline double im(double) { return 0.0; }
__host__ __device__ inline double im(zd a) { return a.y; }
__host__ __device__ inline double cj(double a) { return a; }
__host__ __device__ inline zd cj(zd a) { return mkz(a.x, -a.y); }
__host__ __device__ inline double abs2(double a) { return a * a; }
__host__ __device__ inline double abs2(zd a) { return a.x * a.x + a.y * a.y; }
__host__ __device__ inline double absv(double a) { return fabs(a); }
__host__ __device__ inline double absv(zd a) { return hypot(a.x, a.y); }
__host__ __device__ inline double abs1(double a) { return fabs(a); }
__host__ __device__ inline double abs1(zd a) { return fabs(a.x) + fabs(a.y); }
__host__ __device__ inline double rdiv(double a, double b) { return a / b; }
__host__ __device__ inline zd rdiv(zd a, zd b) { return zdiv(a, b); }
template <typename T> __host__ __device__ inline T from_re(double a);
template <> __host__ __device__ inline double from_re<double>(double a) { return a; }
template <> __host__ __device__ inline zd from_re<zd>(double a) { return mkz(a, 0.0); }
__host__ __device__ inline bool is_zero(double a) { return a == 0.0; }
__host__ __device__ inline bool is_zero(zd a) { return a.x == 0.0 && a.y == 0.0; }

template <typename T> struct is_cplx { static const bool value = false; };
template <> struct is_cplx<zd> { static const bool value = true; };

// wave64 reductions ---------------------------------------------------------
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline zd wave_sum(zd v)
{
    v.x = wave_sum(v.x);
    v.y = wave_sum(v.y);
    return v;
}
__device__ inline double shfl(double v, int src) { return __shfl(v, src, 64); }
__device__ inline int shfl(int v, int src) { return __shfl(v, src, 64); }
__device__ inline zd shfl(zd v, int src) { return mkz(__shfl(v.x, src, 64), __shfl(v.y, src, 64)); }
