// dq_update3.hpp -- Metropolis sweep over one time slice with delayed rank-KD
// updates (src/updates.c:7-84) for N <= 256, delay buffers split between
// registers and LDS so that the rank can be 32 (16 complex).
//
// Each flush G += A B^T reads and writes both spins' G (4 N^2 values per
// chain); with all chains flushing, that streams at HBM speed, and the number
// of flushes falls as 1/KD.  One 256-thread workgroup per chain, both spins,
// thread j owns row j:
//   * its row of A (A[j][0..KD)) lives in registers -- the flush needs exactly
//     that row, and a new column enters at a run-time slot by selects;
//   * B lives in LDS as B[m][j] (128 KiB for both spins): the flush and the
//     gemv read B[m][c] / B[m][i] as broadcasts and B[m][j] without bank
//     conflicts;
//   * on an acceptance at site i the owner of row i publishes A[i][0..k)
//     through LDS.
// Lane 0 scans the precomputed site order from LDS; only an acceptance
// touches G.  Slots past the filled ones hold zeros and are skipped.
#pragma once
#include "dq_common.hpp"
#include "dq_gemm.hpp"
#include "dq_update2.hpp"

#define UPD3_THREADS 256

// G[j][c] += sum_m A[j][m] B[m][c] for the thread's row j (j < N), all c;
// returns the new G[j][j].  FB columns of the row stay in flight at once.
template <typename T, int KD>
__device__ inline T upd3_flush(int N, int j, T* G, const T (&A)[KD], const T (*B)[UPD3_THREADS],
                               int kused)
{
    constexpr int FB = 8;
    T gjj = from_re<T>(0.0);
    for (int c0 = 0; c0 < N; c0 += FB) {
        T g[FB];
#pragma unroll
        for (int t = 0; t < FB; t++)
            if (c0 + t < N) g[t] = G[j + (long)N * (c0 + t)];
#pragma unroll
        for (int m = 0; m < KD; m++) {
            if (m < kused) {  // uniform
#pragma unroll
                for (int t = 0; t < FB; t++) g[t] += A[m] * B[m][(c0 + t) & (UPD3_THREADS - 1)];
            }
        }
#pragma unroll
        for (int t = 0; t < FB; t++) {
            if (c0 + t < N) {
                G[j + (long)N * (c0 + t)] = g[t];
                if (c0 + t == j) gjj = g[t];
            }
        }
    }
    return gjj;
}

// The same flush on the FP64 matrix cores (real builds): per wave the rows
// [64w, 64w + 64) and all columns in chunks of 32, as the transposed product
//   G^T[c][r] += sum_m B[m][c] A[r][m]
// in 16x16x4 blocks: the B fragment (lane l: B[m0 + l/16][c0 + l%16]) is one
// LDS read per block column, the A fragment (lane l: A[r0 + l%16][m0 + l/16])
// is gathered once per flush from the owning lanes' registers, and lane l of
// the result holds G[r0 + l%16][c0 + 4e + l/16], so 16 lanes move 128
// contiguous bytes of the column-major G.  The diagonal goes to d[].
// (Operand layouts as in dq_gemm.hpp.)  The VALU flush above needs one LDS
// broadcast per FMA and is bound by the LDS return path; this one needs one
// LDS read per 64 FMAs per lane.
template <int KD>
__device__ inline void upd3_flush_mfma(int N, int j, double* G, const double (&A)[KD],
                                       const double (*B)[UPD3_THREADS], int kused, double* d)
{
    constexpr int KB = KD / 4;
    const int lane = j & 63, w = j >> 6;
    double af[4][KB];
#pragma unroll
    for (int b = 0; b < 4; b++) {
        const int src = 16 * b + (lane & 15);
#pragma unroll
        for (int kb = 0; kb < KB; kb++) {
            double v = 0.0;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const double x = __shfl(A[4 * kb + q], src, 64);
                if ((lane >> 4) == q) v = x;
            }
            af[b][kb] = v;
        }
    }
    const int rl = 64 * w + (lane & 15);
    constexpr int MB = 2;  // 16-column blocks per chunk
    for (int c0 = 0; c0 < N; c0 += 16 * MB) {
        dbl4 acc[MB][4];
#pragma unroll
        for (int mb = 0; mb < MB; mb++)
#pragma unroll
            for (int b = 0; b < 4; b++)
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int r = rl + 16 * b, cc = c0 + 16 * mb + 4 * e + (lane >> 4);
                    acc[mb][b][e] = (r < N && cc < N) ? G[r + (long)N * cc] : 0.0;
                }
#pragma unroll
        for (int kb = 0; kb < KB; kb++) {
            if (4 * kb < kused) {  // uniform
#pragma unroll
                for (int mb = 0; mb < MB; mb++) {
                    const int cc = (c0 + 16 * mb + (lane & 15)) & (UPD3_THREADS - 1);
                    const double bf = B[4 * kb + (lane >> 4)][cc];
#pragma unroll
                    for (int b = 0; b < 4; b++) acc[mb][b] = mfma(bf, af[b][kb], acc[mb][b]);
                }
            }
        }
#pragma unroll
        for (int mb = 0; mb < MB; mb++)
#pragma unroll
            for (int b = 0; b < 4; b++)
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int r = rl + 16 * b, cc = c0 + 16 * mb + 4 * e + (lane >> 4);
                    if (r < N && cc < N) {
                        G[r + (long)N * cc] = acc[mb][b][e];
                        if (r == cc) d[r] = acc[mb][b][e];
                    }
                }
    }
}

template <typename T, int KD>
__global__ __launch_bounds__(UPD3_THREADS) void k_update_lds(Upd2Args<T> a)
{
    constexpr int NT = UPD3_THREADS;
    __shared__ T du[NT], dd[NT];
    __shared__ int hsl[NT], ordl[NT];
    __shared__ double ul[NT], dl0[NT], dl1[NT];
    __shared__ T Bu[KD][NT], Bd[KD][NT];
    __shared__ T rAu[KD], rAd[KD];
    __shared__ int s_site;
    __shared__ T s_fu, s_fd;
    const int c = blockIdx.x;
    if (!a.active[c]) return;
    if (a.Lc && a.l >= a.Lc[c]) return;  // identity slice of a padded chain
    const int N = a.N, j = threadIdx.x;
    const bool own = j < N;
    T* Gu = a.G + (2 * c) * a.sG;
    T* Gd = a.G + (2 * c + 1) * a.sG;
    int* hs = a.hs + (long)c * a.L * N + (long)N * a.l;
    const long ol = ((long)c * a.L + a.l) * N;
    const T zero = from_re<T>(0.0), one = from_re<T>(1.0);
    if (own) {
        du[j] = Gu[j + (long)N * j];
        dd[j] = Gd[j + (long)N * j];
        hsl[j] = hs[j];
        ordl[j] = a.order[ol + j];
        ul[j] = a.unif[ol + j];
        dl0[j] = a.del[2L * N * c + j];
        dl1[j] = a.del[2L * N * c + N + j];
    }
    T phase = zero;
    unsigned long long nacc = 0;
    if (j == 0) phase = a.phase[c];
    int ii = 0;  // lane 0's scan position
    T Au[KD], Ad[KD];
    for (;;) {
#pragma unroll
        for (int m = 0; m < KD; m++) {
            Au[m] = zero; Ad[m] = zero;
            Bu[m][j] = zero; Bd[m][j] = zero;
        }
        __syncthreads();
        int k = 0;
        for (; k < KD; k++) {
            if (j == 0) {
                int site = -1;
                for (; ii < N; ii++) {
                    const int i = ordl[ii];
                    const int sv = hsl[i];
                    const double delu = sv ? dl1[i] : dl0[i];
                    const double deld = sv ? dl0[i] : dl1[i];
                    if (delu == 0.0 && deld == 0.0) continue;
                    const T ru = one + (one - du[i]) * delu;
                    const T rd = one + (one - dd[i]) * deld;
                    const T prob = ru * rd;
                    const double absprob = absv(prob);
                    if (ul[ii] < absprob) {
                        site = i;
                        s_fu = rdiv(from_re<T>(delu), ru);
                        s_fd = rdiv(from_re<T>(deld), rd);
                        hsl[i] = 1 - sv;
                        phase = phase * (prob * (1.0 / absprob));
                        nacc++;
                        ii++;
                        break;
                    }
                }
                s_site = site;
            }
            __syncthreads();
            const int i = s_site;
            if (i < 0) break;
            // the owner of row i publishes its filled slots of A
            if (j == i) {
#pragma unroll
                for (int m = 0; m < KD; m++)
                    if (m < k) { rAu[m] = Au[m]; rAd[m] = Ad[m]; }
            }
            T au = zero, bu = zero, ad = zero, bd = zero;
            if (own) {
                au = Gu[j + (long)N * i];
                bu = Gu[i + (long)N * j];
                ad = Gd[j + (long)N * i];
                bd = Gd[i + (long)N * j];
            }
            __syncthreads();
            // a_k = G[:, i] + A B[i, :]^T,  b_k = G[i, :]^T + B A[i, :]^T
#pragma unroll
            for (int m = 0; m < KD; m++) {
                if (m < k) {  // uniform
                    au += Au[m] * Bu[m][i];
                    bu += Bu[m][j] * rAu[m];
                    ad += Ad[m] * Bd[m][i];
                    bd += Bd[m][j] * rAd[m];
                }
            }
            if (j == i) { au -= one; ad -= one; }
            au = au * s_fu;
            ad = ad * s_fd;
            if (own) {
                du[j] += au * bu;
                dd[j] += ad * bd;
            } else {
                au = zero; bu = zero; ad = zero; bd = zero;
            }
#pragma unroll
            for (int m = 0; m < KD; m++)
                if (m == k) { Au[m] = au; Ad[m] = ad; }
            Bu[k][j] = bu;
            Bd[k][j] = bd;
            __syncthreads();
        }
        if (k > 0) {
            if constexpr (is_cplx<T>::value) {
                if (own) {
                    du[j] = upd3_flush<T, KD>(N, j, Gu, Au, Bu, k);
                    dd[j] = upd3_flush<T, KD>(N, j, Gd, Ad, Bd, k);
                }
            } else {  // every lane takes part (the A fragments are gathered)
                upd3_flush_mfma<KD>(N, j, Gu, Au, Bu, k, du);
                upd3_flush_mfma<KD>(N, j, Gd, Ad, Bd, k, dd);
            }
            __syncthreads();
        }
        if (k < KD) break;
    }
    // final field of this slice: HS row and the scalings of B_l / B_l^-1
    const double* el = a.exp_lambda + (long)c * 2 * N;
    double* esu = a.escale + (2 * c) * (long)a.L * N + (long)N * a.l;
    double* esd = a.escale + (2 * c + 1) * (long)a.L * N + (long)N * a.l;
    double* isu = a.iscale + (2 * c) * (long)a.L * N + (long)N * a.l;
    double* isd = a.iscale + (2 * c + 1) * (long)a.L * N + (long)N * a.l;
    if (own) {
        const int sv = hsl[j];
        hs[j] = sv;
        esu[j] = el[j + N * sv];
        esd[j] = el[j + N * (1 - sv)];
        isu[j] = el[j + N * (1 - sv)];
        isd[j] = el[j + N * sv];
    }
    if (j == 0) {
        a.phase[c] = phase;
        a.n_acc[c] += nacc;
    }
}
