// dq_ue.hpp -- helper kernels for the unequal-time Green's functions
// (calc_o / bsofi / expand_g, src/greens.c:247-469): block copies with
// leading dimensions, triangle extraction, identity set-up, the triangular
// solve that inverts R (xtrtri), and row scalings that materialise B_l^-1.
#pragma once
#include "dq_common.hpp"

// dst[b] (m x n, ldd) = alpha * src[b] (m x n, lds)
template <typename T>
__global__ void k_copy_mn(int m, int n, const T* src, long ssrc, int lds, T alpha, T* dst,
                          long sdst, int ldd, int batch)
{
    EW_IB(m);
    EW_FOR_J(n) dst[b * sdst + i + (long)ldd * j] = alpha * src[b * ssrc + i + (long)lds * j];
}

// dst[b] (n x n, ldd) = upper triangle of src[b] (lds), zeros below
template <typename T>
__global__ void k_copy_upper(int n, const T* src, long ssrc, int lds, T* dst, long sdst, int ldd,
                             int batch)
{
    EW_IB(n);
    EW_FOR_J(n)
        dst[b * sdst + i + (long)ldd * j] = (i <= j) ? src[b * ssrc + i + (long)lds * j] : from_re<T>(0.0);
}

// upper triangle (incl. diagonal) of src written into dst; dst's strictly
// lower part is left untouched (xtrtri semantics)
template <typename T>
__global__ void k_store_upper(int n, const T* src, long ssrc, int lds, T* dst, long sdst, int ldd,
                              int batch)
{
    EW_IB(n);
    EW_FOR_J(n) if (i <= j) dst[b * sdst + i + (long)ldd * j] = src[b * ssrc + i + (long)lds * j];
}

// zero the strictly lower part of an m x n panel (ld)
template <typename T>
__global__ void k_zero_lower(int m, int n, T* A, long sA, int lda, int batch)
{
    EW_IB(m);
    EW_FOR_J(n) if (i > j) A[b * sA + i + (long)lda * j] = from_re<T>(0.0);
}

// A[b] (n x n, lda) = alpha * I + beta_keep * A  (beta_keep 0 or 1)
template <typename T>
__global__ void k_set_identity(int n, T* A, long sA, int lda, int keep, int batch)
{
    EW_IB(n);
    EW_FOR_J(n) {
        T* p = A + b * sA + i + (long)lda * j;
        const T d = from_re<T>(i == j ? 1.0 : 0.0);
        *p = keep ? (*p + d) : d;
    }
}

// X[b] = diag(s[b]) * A[b / adiv]  (B_l^-1 = diag(e'_l) inv_expK, src/dqmc.c:45-59)
template <typename T>
__global__ void k_rowscale(int N, const T* A, long sA, int adiv, const double* s, long ss, T* X,
                           long sX, int batch)
{
    EW_IB(N);
    EW_FOR_J(N) {
        const long r = i + (long)N * j;
        X[b * sX + r] = s[b * ss + i] * A[(b / adiv) * sA + r];
    }
}

// Solve U X = X in place for the n x n upper-triangular U[b] (ldu) and
// ncols right-hand sides X[b] (ldx): back substitution, TS_NC columns per
// wave held in registers (RPL rows per lane).  With X = I this is xtrtri.
#define TS_NC 4
template <typename T, int RPL>
__global__ __launch_bounds__(256) void k_trsm_upper(int n, int ncols, const T* U_, long sU, int ldu,
                                                    T* X_, long sX, int ldx)
{
    const int b = blockIdx.y;
    const T* U = U_ + b * sU;
    T* X = X_ + b * sX;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c0 = (blockIdx.x * (blockDim.x >> 6) + wave) * TS_NC;
    if (c0 >= ncols) return;
    T x[TS_NC][RPL];
#pragma unroll
    for (int q = 0; q < TS_NC; q++)
#pragma unroll
        for (int r = 0; r < RPL; r++) {
            const int i = lane + 64 * r;
            x[q][r] = (i < n && c0 + q < ncols) ? X[i + (long)ldx * (c0 + q)] : from_re<T>(0.0);
        }
    for (int kk = n - 1; kk >= 0; kk--) {
        const int owner = kk & 63, rk = kk >> 6;
        const T ukk = U[kk + (long)ldu * kk];
        T xk[TS_NC];
#pragma unroll
        for (int q = 0; q < TS_NC; q++) {
            T val = from_re<T>(0.0);
#pragma unroll
            for (int r = 0; r < RPL; r++)
                if (r == rk) val = x[q][r];
            val = rdiv(val, ukk);
            xk[q] = bcast(val, owner);
#pragma unroll
            for (int r = 0; r < RPL; r++)
                if (r == rk && lane == owner) x[q][r] = xk[q];
        }
#pragma unroll
        for (int r = 0; r < RPL; r++) {
            const int i = lane + 64 * r;
            if (i < kk) {
                const T u = U[i + (long)ldu * kk];
#pragma unroll
                for (int q = 0; q < TS_NC; q++) x[q][r] -= u * xk[q];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < TS_NC; q++)
#pragma unroll
        for (int r = 0; r < RPL; r++) {
            const int i = lane + 64 * r;
            if (i < n && c0 + q < ncols) X[i + (long)ldx * (c0 + q)] = x[q][r];
        }
}

// Y[b] = X[b]^T for the N x N matrices b = blockIdx.x / tn^2 (stride NN), one
// 32 x 32 tile per workgroup through LDS (tn = ceil(N / 32) tiles a side)
template <typename T>
__global__ __launch_bounds__(256) void k_transpose_batch(int N, int tn, const T* X, T* Y, long NN)
{
    __shared__ T s[32][33];
    const int tiles = tn * tn;
    const long b = blockIdx.x / tiles;
    const int t = blockIdx.x % tiles, ti = t % tn, tj = t / tn;
    const T* x = X + b * NN;
    T* y = Y + b * NN;
    const int lx = threadIdx.x & 31, ly = threadIdx.x >> 5;
    for (int k = ly; k < 32; k += 8) {
        const int r = 32 * ti + lx, c = 32 * tj + k;
        if (r < N && c < N) s[k][lx] = x[r + (long)N * c];
    }
    __syncthreads();
    for (int k = ly; k < 32; k += 8) {  // y(r, c) = x(c, r) = s[lx][k]
        const int r = 32 * tj + lx, c = 32 * ti + k;
        if (r < N && c < N) y[r + (long)N * c] = s[lx][k];
    }
}

// p[0:n) = 0 (a fill kernel rather than a memset node, so that the sweeps can
// be captured into HIP graphs)
__global__ void k_zero(long n, double* p)
{
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
        p[e] = 0.0;
}

// X[b][l] = Y[b][l] = I for the slices l >= Lc[b / 2] of every (chain, spin) b
template <typename T>
__global__ void k_ue_identity_slices(int N, int L, T* X, T* Y, const int* Lc, int B)
{
    const long NN = (long)N * N, LNN = (long)L * NN;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < LNN * B;
         e += (long)gridDim.x * blockDim.x) {
        const int b = (int)(e / LNN);
        const long r = e % LNN;
        const int l = (int)(r / NN);
        if (l < Lc[b >> 1]) continue;
        const long q = r % NN;
        const T v = from_re<T>((q % N) == (q / N) ? 1.0 : 0.0);
        X[b * LNN + r] = v;
        Y[b * LNN + r] = v;
    }
}
