// dq_gemm.hpp -- batched FP64 MFMA GEMM with fused diagonal scalings.
//
//   C[b] = alpha * diag(r[b]) * op(A[b]) * diag(u[b]) * op(B[b]) * diag(w[b]) + beta * C[b]
//
// This one kernel carries every dense product of the DQMC path:
//   B_l X   = expK * diag(e_l) * X                   (u = e_l)       src/dqmc.c:29-35, greens.c:8-41
//   X B_l^-1 = X * diag(e'_l) * invexpK              (u = e'_l)      src/dqmc.c:45-59, 294-295
//   X B_l   = X * expK * diag(e_l)                    (w = e_l)       src/greens.c:374-376
//   B_l^-1 X = diag(e'_l) * invexpK * X              (r = e'_l)      src/greens.c:414-416
//   rank-k flush G += A B^T, Q^H products, half-wraps, T updates.
// B_l is never materialised: the HS-field scaling is applied on the fly.
//
// Tiling for CDNA4: 256 threads = 4 wave64s, 64x64 C tile per workgroup, each
// wave owns 32x32 = 2x2 tiles of v_mfma_f64_16x16x4 (K step 16 staged through
// LDS).  Complex (zd) uses four real MFMAs per block on planar LDS copies.
// Operand lane layout of v_mfma_f64_16x16x4f64 (wave64):
//   A[16x4]: lane l holds A[l%16][l/16];  B[4x16]: lane l holds B[l/16][l%16]
//   D[16x16]: lane l, element r holds D[4*r + l/16][l%16]
// (input and output layouts verified on gfx950 with a single-element probe)
#pragma once
#include "dq_common.hpp"

typedef double dbl4 __attribute__((ext_vector_type(4)));

enum { OP_N = 0, OP_T = 1, OP_C = 2 };

template <typename T>
struct GemmArgs {
    int M, N, K;
    const T* A; int lda; long sA;
    const T* B; int ldb; long sB;
    T* C; int ldc; long sC;
    const double* r; long sr; // row scale of the result (or null)
    const double* u; long su; // inner scale (or null)
    const double* w; long sw; // column scale of the result (or null)
    T alpha, beta;
    int beta_zero;            // 1: C is write-only
    int adiv, bdiv;           // batch z uses A[z / adiv], B[z / bdiv] (shared operands)
};

#define GB_M 64
#define GB_N 64
#define GB_K 16

template <int OP, typename T>
__device__ inline T ld_op(const T* p, long idx)
{
    if (OP == OP_C) return cj(p[idx]);
    return p[idx];
}

__device__ inline dbl4 mfma(double a, double b, dbl4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// real ---------------------------------------------------------------------
template <int OPA, int OPB>
__global__ __launch_bounds__(256) void k_gemm_r(GemmArgs<double> g)
{
    __shared__ double As[GB_K][GB_M];
    __shared__ double Bs[GB_K][GB_N];
    const int b = blockIdx.z;
    const int m0 = blockIdx.x * GB_M, n0 = blockIdx.y * GB_N;
    const double* A = g.A + (b / g.adiv) * g.sA;
    const double* B = g.B + (b / g.bdiv) * g.sB;
    double* C = g.C + b * g.sC;
    const double* u = g.u ? g.u + b * g.su : nullptr;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    dbl4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = (dbl4){0.0, 0.0, 0.0, 0.0};

    for (int k0 = 0; k0 < g.K; k0 += GB_K) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int idx = tid + 256 * q;
            int i, k;
            if (OPA == OP_N) { i = idx & 63; k = idx >> 6; }
            else { k = idx & 15; i = idx >> 4; }
            const int gi = m0 + i;
            int gk = k0 + k;
            double v = 0.0;
            if (gi < g.M && gk < g.K)
                v = (OPA == OP_N) ? A[gi + (long)g.lda * gk] : A[gk + (long)g.lda * gi];
            As[k][i] = v;
            int j;
            if (OPB == OP_N) { k = idx & 15; j = idx >> 4; }
            else { j = idx & 63; k = idx >> 6; }
            const int gj = n0 + j;
            gk = k0 + k;
            double wv = 0.0;
            if (gj < g.N && gk < g.K) {
                wv = (OPB == OP_N) ? B[gk + (long)g.ldb * gj] : B[gj + (long)g.ldb * gk];
                if (u) wv *= u[gk];
            }
            Bs[k][j] = wv;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < GB_K; ks += 4) {
            const int kk = ks + (lane >> 4);
            double a0 = As[kk][wm + (lane & 15)];
            double a1 = As[kk][wm + 16 + (lane & 15)];
            double b0 = Bs[kk][wn + (lane & 15)];
            double b1 = Bs[kk][wn + 16 + (lane & 15)];
            acc[0][0] = mfma(a0, b0, acc[0][0]);
            acc[0][1] = mfma(a0, b1, acc[0][1]);
            acc[1][0] = mfma(a1, b0, acc[1][0]);
            acc[1][1] = mfma(a1, b1, acc[1][1]);
        }
        __syncthreads();
    }
    const double* rr = g.r ? g.r + b * g.sr : nullptr;
    const double* ww = g.w ? g.w + b * g.sw : nullptr;
#pragma unroll
    for (int ti = 0; ti < 2; ti++)
#pragma unroll
        for (int tj = 0; tj < 2; tj++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int row = m0 + wm + ti * 16 + 4 * e + (lane >> 4);
                const int col = n0 + wn + tj * 16 + (lane & 15);
                if (row < g.M && col < g.N) {
                    double v = g.alpha * acc[ti][tj][e];
                    if (rr) v *= rr[row];
                    if (ww) v *= ww[col];
                    const long o = row + (long)g.ldc * col;
                    if (!g.beta_zero) v += g.beta * C[o];
                    C[o] = v;
                }
            }
}

// complex ------------------------------------------------------------------
template <int OPA, int OPB>
__global__ __launch_bounds__(256) void k_gemm_z(GemmArgs<zd> g)
{
    __shared__ double Ar[GB_K][GB_M], Ai[GB_K][GB_M];
    __shared__ double Br[GB_K][GB_N], Bi[GB_K][GB_N];
    const int b = blockIdx.z;
    const int m0 = blockIdx.x * GB_M, n0 = blockIdx.y * GB_N;
    const zd* A = g.A + (b / g.adiv) * g.sA;
    const zd* B = g.B + (b / g.bdiv) * g.sB;
    zd* C = g.C + b * g.sC;
    const double* u = g.u ? g.u + b * g.su : nullptr;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    dbl4 cr[2][2], ci[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            cr[i][j] = (dbl4){0.0, 0.0, 0.0, 0.0};
            ci[i][j] = (dbl4){0.0, 0.0, 0.0, 0.0};
        }
    for (int k0 = 0; k0 < g.K; k0 += GB_K) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int idx = tid + 256 * q;
            int i, k;
            if (OPA == OP_N) { i = idx & 63; k = idx >> 6; }
            else { k = idx & 15; i = idx >> 4; }
            const int gi = m0 + i;
            int gk = k0 + k;
            zd v = mkz(0.0, 0.0);
            if (gi < g.M && gk < g.K)
                v = (OPA == OP_N) ? A[gi + (long)g.lda * gk]
                                  : ld_op<OPA>(A, gk + (long)g.lda * gi);
            Ar[k][i] = v.x;
            Ai[k][i] = v.y;
            int j;
            if (OPB == OP_N) { k = idx & 15; j = idx >> 4; }
            else { j = idx & 63; k = idx >> 6; }
            const int gj = n0 + j;
            gk = k0 + k;
            zd wv = mkz(0.0, 0.0);
            if (gj < g.N && gk < g.K) {
                wv = (OPB == OP_N) ? B[gk + (long)g.ldb * gj]
                                   : ld_op<OPB>(B, gj + (long)g.ldb * gk);
                if (u) wv = u[gk] * wv;
            }
            Br[k][j] = wv.x;
            Bi[k][j] = wv.y;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < GB_K; ks += 4) {
            const int kk = ks + (lane >> 4);
            const int l16 = lane & 15;
            double ar0 = Ar[kk][wm + l16], ai0 = Ai[kk][wm + l16];
            double ar1 = Ar[kk][wm + 16 + l16], ai1 = Ai[kk][wm + 16 + l16];
            double br0 = Br[kk][wn + l16], bi0 = Bi[kk][wn + l16];
            double br1 = Br[kk][wn + 16 + l16], bi1 = Bi[kk][wn + 16 + l16];
            double nai0 = -ai0, nai1 = -ai1;
            cr[0][0] = mfma(ar0, br0, cr[0][0]); cr[0][0] = mfma(nai0, bi0, cr[0][0]);
            ci[0][0] = mfma(ar0, bi0, ci[0][0]); ci[0][0] = mfma(ai0, br0, ci[0][0]);
            cr[0][1] = mfma(ar0, br1, cr[0][1]); cr[0][1] = mfma(nai0, bi1, cr[0][1]);
            ci[0][1] = mfma(ar0, bi1, ci[0][1]); ci[0][1] = mfma(ai0, br1, ci[0][1]);
            cr[1][0] = mfma(ar1, br0, cr[1][0]); cr[1][0] = mfma(nai1, bi0, cr[1][0]);
            ci[1][0] = mfma(ar1, bi0, ci[1][0]); ci[1][0] = mfma(ai1, br0, ci[1][0]);
            cr[1][1] = mfma(ar1, br1, cr[1][1]); cr[1][1] = mfma(nai1, bi1, cr[1][1]);
            ci[1][1] = mfma(ar1, bi1, ci[1][1]); ci[1][1] = mfma(ai1, br1, ci[1][1]);
        }
        __syncthreads();
    }
    const double* rr = g.r ? g.r + b * g.sr : nullptr;
    const double* ww = g.w ? g.w + b * g.sw : nullptr;
#pragma unroll
    for (int ti = 0; ti < 2; ti++)
#pragma unroll
        for (int tj = 0; tj < 2; tj++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int row = m0 + wm + ti * 16 + 4 * e + (lane >> 4);
                const int col = n0 + wn + tj * 16 + (lane & 15);
                if (row < g.M && col < g.N) {
                    zd v = g.alpha * mkz(cr[ti][tj][e], ci[ti][tj][e]);
                    if (rr) v = rr[row] * v;
                    if (ww) v = ww[col] * v;
                    const long o = row + (long)g.ldc * col;
                    if (!g.beta_zero) v += g.beta * C[o];
                    C[o] = v;
                }
            }
}

template <typename T> struct GemmLauncher;

template <> struct GemmLauncher<double> {
    static hipError_t launch(int opa, int opb, const GemmArgs<double>& g, int batch, hipStream_t s)
    {
        dim3 grid((g.M + GB_M - 1) / GB_M, (g.N + GB_N - 1) / GB_N, batch);
        if (opa == OP_N && opb == OP_N) k_gemm_r<OP_N, OP_N><<<grid, 256, 0, s>>>(g);
        else if (opa == OP_N) k_gemm_r<OP_N, OP_T><<<grid, 256, 0, s>>>(g);
        else if (opb == OP_N) k_gemm_r<OP_T, OP_N><<<grid, 256, 0, s>>>(g);
        else k_gemm_r<OP_T, OP_T><<<grid, 256, 0, s>>>(g);
        return hipGetLastError();
    }
};

template <> struct GemmLauncher<zd> {
    static hipError_t launch(int opa, int opb, const GemmArgs<zd>& g, int batch, hipStream_t s)
    {
        dim3 grid((g.M + GB_M - 1) / GB_M, (g.N + GB_N - 1) / GB_N, batch);
#define ZL(a, b) k_gemm_z<a, b><<<grid, 256, 0, s>>>(g)
        if (opa == OP_N) {
            if (opb == OP_N) ZL(OP_N, OP_N);
            else if (opb == OP_T) ZL(OP_N, OP_T);
            else ZL(OP_N, OP_C);
        } else if (opa == OP_T) {
            if (opb == OP_N) ZL(OP_T, OP_N);
            else if (opb == OP_T) ZL(OP_T, OP_T);
            else ZL(OP_T, OP_C);
        } else {
            if (opb == OP_N) ZL(OP_C, OP_N);
            else if (opb == OP_T) ZL(OP_C, OP_T);
            else ZL(OP_C, OP_C);
        }
#undef ZL
        return hipGetLastError();
    }
};
