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
__device__ inline double readlane_d(double v, int l)
{
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)u, l);
    const int hi = __builtin_amdgcn_readlane((int)(u >> 32), l);
    return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}

// Elementwise kernels over a batch of m x n column-major blocks run on the grid
// (row blocks, columns, batch entries), so that a thread's row i, column j and
// entry b come from its indices with no integer division (64-bit divisions of
// a flat index made these kernels instruction-bound).
// Each thread takes row i of EW_COLS consecutive columns j (EW_FOR_J).
#define EW_COLS 4
#define EW_IB(m_)                                        \
    const int i = blockIdx.x * blockDim.x + threadIdx.x; \
    const int b = blockIdx.z;                            \
    if (i >= (m_)) return
#define EW_FOR_J(n_)                                                                  \
    for (int j = blockIdx.y * EW_COLS, j_end_ = (j + EW_COLS < (n_) ? j + EW_COLS : (n_)); \
         j < j_end_; j++)
static inline int ew_threads(int m) { return m > 128 ? 256 : (m > 64 ? 128 : 64); }
static inline dim3 ew_grid(int m, int n, int batch)
{
    const int t = ew_threads(m);
    return dim3((unsigned)((m + t - 1) / t), (unsigned)((n + EW_COLS - 1) / EW_COLS),
                (unsigned)batch);
}

// --- wave64 reductions in registers (no LDS round trips) -------------------
// gfx950 lane exchanges:  permlane32_swap(x, y) -> ([x lanes 0-31, y lanes
// 0-31], [x lanes 32-63, y lanes 32-63]);  permlane16_swap(x, y) ->
// ([x row 0, y row 0, x row 2, y row 2], [x row 1, y row 1, x row 3, y row 3]);
// DPP row_ror:n reads lane (i - n) mod 16 of the same 16-lane row.
__device__ inline void swap32_d(double& x, double& y)
{
    const long long ux = __double_as_longlong(x), uy = __double_as_longlong(y);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)ux, (unsigned)uy, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(ux >> 32), (unsigned)(uy >> 32),
                                                     false, false);
    x = __longlong_as_double(((long long)hi[0] << 32) | lo[0]);
    y = __longlong_as_double(((long long)hi[1] << 32) | lo[1]);
}
__device__ inline void swap16_d(double& x, double& y)
{
    const long long ux = __double_as_longlong(x), uy = __double_as_longlong(y);
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)ux, (unsigned)uy, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(ux >> 32), (unsigned)(uy >> 32),
                                                     false, false);
    x = __longlong_as_double(((long long)hi[0] << 32) | lo[0]);
    y = __longlong_as_double(((long long)hi[1] << 32) | lo[1]);
}
// The controls used here (row rotations and mirrors, quad permutations) read a
// valid lane for every lane, so the old-value operand is never taken: the
// plain DPP move needs no v_mov of a fill value before it.
template <int CTRL>
__device__ inline double dpp_d(double v)
{
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)u, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
#define DPP_ROR(n) (0x120 + (n))
#define DPP_MIRROR 0x140
#define DPP_HALF_MIRROR 0x141
#define DPP_QUAD_REV 0x1B   // quad_perm [3,2,1,0]
#define DPP_QUAD_XOR1 0xB1  // quad_perm [1,0,3,2]

// Sums K in {1, 2, 4, 8} per-lane values over the wave at once, transposed:
// each exchange halves the values a lane carries (lanes 32-63 take the upper
// half, then odd rows, then the upper 8 lanes of a row), the rest is summed
// within the lane groups.  Value i ends in lanes (64 / K) i ... ; on return
// every a[i] holds its full sum, uniform over the wave.
template <int K>
__device__ inline void wave_sum_many(double (&a)[K], int lane)
{
    static_assert(K == 1 || K == 2 || K == 4 || K == 8, "K in {1, 2, 4, 8}");
    constexpr int KA = K > 1 ? K / 2 : 1;
    constexpr int KB = KA > 1 ? KA / 2 : 1;
    double c1[KA], c2[KB];
#pragma unroll
    for (int i = 0; i < KA; i++) {
        double x = a[i], y = a[K > 1 ? i + KA : i];
        swap32_d(x, y);
        c1[i] = x + y;
    }
#pragma unroll
    for (int i = 0; i < KB; i++) {
        double x = c1[i], y = c1[KA > 1 ? i + KB : i];
        swap16_d(x, y);
        c2[i] = x + y;
    }
    double cur;
    if constexpr (K == 8) {
        const bool up = (lane & 8) != 0;
        const double keep = up ? c2[1] : c2[0], send = up ? c2[0] : c2[1];
        cur = keep + dpp_d<DPP_MIRROR>(send);
        cur += dpp_d<DPP_HALF_MIRROR>(cur);
        cur += dpp_d<DPP_QUAD_REV>(cur);
        cur += dpp_d<DPP_QUAD_XOR1>(cur);
    } else {
        cur = c2[0];
        cur += dpp_d<DPP_ROR(8)>(cur);
        cur += dpp_d<DPP_ROR(4)>(cur);
        cur += dpp_d<DPP_ROR(2)>(cur);
        cur += dpp_d<DPP_ROR(1)>(cur);
    }
#pragma unroll
    for (int i = 0; i < K; i++) a[i] = readlane_d(cur, (64 / K) * i);
}
__device__ inline double wave_sum_r(double v, int lane)
{
    double a[1] = {v};
    wave_sum_many<1>(a, lane);
    return a[0];
}
__device__ inline zd wave_sum_r(zd v, int lane)
{
    double a[2] = {v.x, v.y};
    wave_sum_many<2>(a, lane);
    return mkz(a[0], a[1]);
}

// component-wise select (a conditional over whole complex values would be
// lowered through memory)
__device__ inline double selv(bool c, double a, double b) { return c ? a : b; }
__device__ inline zd selv(bool c, zd a, zd b) { return mkz(c ? a.x : b.x, c ? a.y : b.y); }

// doubles per value, and moving values through double arrays
template <typename T> struct val_parts { static const int n = 1; };
template <> struct val_parts<zd> { static const int n = 2; };
__device__ inline void put_part(double* a, double v) { a[0] = v; }
__device__ inline void put_part(double* a, zd v) { a[0] = v.x; a[1] = v.y; }
__device__ inline void get_part(const double* a, double& v) { v = a[0]; }
__device__ inline void get_part(const double* a, zd& v) { v = mkz(a[0], a[1]); }

__device__ inline double shfl(double v, int src) { return __shfl(v, src, 64); }
// broadcast of lane l (wave-uniform) through the scalar unit: v_readlane, no
// LDS round trip (a shuffle from a uniform source costs a ds_bpermute)
__device__ inline double bcast(double v, int l) { return readlane_d(v, l); }
__device__ inline zd bcast(zd v, int l) { return mkz(readlane_d(v.x, l), readlane_d(v.y, l)); }
__device__ inline int shfl(int v, int src) { return __shfl(v, src, 64); }
__device__ inline zd shfl(zd v, int src) { return mkz(__shfl(v.x, src, 64), __shfl(v.y, src, 64)); }
