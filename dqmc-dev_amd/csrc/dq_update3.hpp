// dq_update3.hpp -- Metropolis sweep over one time slice with delayed rank-KD
// updates (src/updates.c:7-84) for N <= 256, delay buffers split between
// registers and LDS so that the rank can be 32 (16 complex).
//
// Each flush G += A B^T reads and writes both spins' G (4 N^2 values per
// chain); with all chains flushing, that streams at HBM speed, and the number
// of flushes falls as 1/KD.  One 256-thread workgroup per chain, both spins,
// thread t owns rows j = t + 256 r (r < R; R = 2 up to N = 512, rank 16):
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

// The rank-k flush G += A B^T on the FP64 matrix cores (real builds): per wave the rows
// row0 + [64w, 64w + 64) and all columns in chunks of 32, as the transposed
// product
//   G^T[c][r] += sum_m B[m][c] A[r][m]
// in 16x16x4 blocks: the B fragment (lane l: B[m0 + l/16][c0 + l%16]) is one
// LDS read per block column, the A fragment (lane l: A[r0 + l%16][m0 + l/16])
// is gathered once per flush from the owning lanes' registers, and lane l of
// the result holds G[r0 + l%16][c0 + 4e + l/16], so 16 lanes move 128
// contiguous bytes of the column-major G.  The diagonal goes to d[].
// (Operand layouts as in dq_gemm.hpp.)  A VALU flush (thread j updating its
// row) needs one LDS broadcast of B per FMA and is bound by the LDS return
// path; this one needs one LDS read per 64 FMAs per lane.
template <int KD, int W>
__device__ inline void upd3_flush_mfma(int N, int j, int row0, double* G, const double (&A)[KD],
                                       const double (*B)[W], int kused, double* d)
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
    const int rl = row0 + 64 * w + (lane & 15);
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
                    const int cc = (c0 + 16 * mb + (lane & 15)) & (W - 1);
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

// The complex flush on the matrix cores: as upd3_flush_mfma with the real and
// imaginary parts as separate real operands, four MFMAs per block,
//   re += Br Ar - Bi Ai,   im += Br Ai + Bi Ar,
// one 16-column block per chunk (the accumulators are twice the real ones).
template <int KD, int W>
__device__ inline void upd3_flush_mfma_z(int N, int j, int row0, zd* G, const zd (&A)[KD],
                                         const zd (*B)[W], int kused, zd* d)
{
    constexpr int KB = KD / 4;
    const int lane = j & 63, w = j >> 6;
    double ar[4][KB], ai[4][KB];
#pragma unroll
    for (int b = 0; b < 4; b++) {
        const int src = 16 * b + (lane & 15);
#pragma unroll
        for (int kb = 0; kb < KB; kb++) {
            double vr = 0.0, vi = 0.0;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const double xr = __shfl(A[4 * kb + q].x, src, 64);
                const double xi = __shfl(A[4 * kb + q].y, src, 64);
                if ((lane >> 4) == q) {
                    vr = xr;
                    vi = xi;
                }
            }
            ar[b][kb] = vr;
            ai[b][kb] = vi;
        }
    }
    const int rl = row0 + 64 * w + (lane & 15);
    for (int c0 = 0; c0 < N; c0 += 16) {
        dbl4 cr[4], ci[4];
#pragma unroll
        for (int b = 0; b < 4; b++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int r = rl + 16 * b, cc = c0 + 4 * e + (lane >> 4);
                const zd g = (r < N && cc < N) ? G[r + (long)N * cc] : mkz(0.0, 0.0);
                cr[b][e] = g.x;
                ci[b][e] = g.y;
            }
#pragma unroll
        for (int kb = 0; kb < KB; kb++) {
            if (4 * kb < kused) {  // uniform
                const zd bz = B[4 * kb + (lane >> 4)][(c0 + (lane & 15)) & (W - 1)];
                const double br = bz.x, bi = bz.y;
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    cr[b] = mfma(br, ar[b][kb], cr[b]);
                    cr[b] = mfma(-bi, ai[b][kb], cr[b]);
                    ci[b] = mfma(br, ai[b][kb], ci[b]);
                    ci[b] = mfma(bi, ar[b][kb], ci[b]);
                }
            }
        }
#pragma unroll
        for (int b = 0; b < 4; b++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int r = rl + 16 * b, cc = c0 + 4 * e + (lane >> 4);
                if (r < N && cc < N) {
                    const zd g = mkz(cr[b][e], ci[b][e]);
                    G[r + (long)N * cc] = g;
                    if (r == cc) d[r] = g;
                }
            }
    }
}

// NT threads, R rows per thread (thread t owns rows t + NT r), N <= NT R.
template <typename T, int R, int KD, int NT = UPD3_THREADS>
__global__ __launch_bounds__(NT) void k_update_lds(Upd2Args<T> a)
{
    constexpr int W = NT * R;
    __shared__ T du[W], dd[W];
    __shared__ int hsl[W], ordl[W];
    __shared__ double ul[W], dl0[W], dl1[W];
    __shared__ T Bu[KD][W], Bd[KD][W];
    __shared__ T rAu[KD], rAd[KD];
    __shared__ int s_site;
    __shared__ T s_fu, s_fd;
    const int c = blockIdx.x;
    if (!a.active[c]) return;
    if (a.Lc && a.l >= a.Lc[c]) return;  // identity slice of a padded chain
    const int N = a.N, tid = threadIdx.x;
    T* Gu = a.G + (2 * c) * a.sG;
    T* Gd = a.G + (2 * c + 1) * a.sG;
    int* hs = a.hs + (long)c * a.L * N + (long)N * a.l;
    const long ol = ((long)c * a.L + a.l) * N;
    const T zero = from_re<T>(0.0), one = from_re<T>(1.0);
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int j = tid + NT * r;
        if (j < N) {
            du[j] = Gu[j + (long)N * j];
            dd[j] = Gd[j + (long)N * j];
            hsl[j] = hs[j];
            ordl[j] = a.order[ol + j];
            ul[j] = a.unif[ol + j];
            dl0[j] = a.del[2L * N * c + j];
            dl1[j] = a.del[2L * N * c + N + j];
        }
    }
    T phase = zero;
    unsigned long long nacc = 0;
    if (tid == 0) phase = a.phase[c];
    int ii = 0;  // lane 0's scan position
    T Au[R][KD], Ad[R][KD];
    for (;;) {
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int m = 0; m < KD; m++) {
                Au[r][m] = zero; Ad[r][m] = zero;
                Bu[m][tid + NT * r] = zero; Bd[m][tid + NT * r] = zero;
            }
        __syncthreads();
        int k = 0;
        for (; k < KD; k++) {
            if (tid == 0) {
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
            if (tid == (i & (NT - 1))) {
#pragma unroll
                for (int r = 0; r < R; r++)
                    if (r == i / NT) {
#pragma unroll
                        for (int m = 0; m < KD; m++)
                            if (m < k) { rAu[m] = Au[r][m]; rAd[m] = Ad[r][m]; }
                    }
            }
            T au[R], bu[R], ad[R], bd[R];
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int j = tid + NT * r;
                au[r] = zero; bu[r] = zero; ad[r] = zero; bd[r] = zero;
                if (j < N) {
                    au[r] = Gu[j + (long)N * i];
                    bu[r] = Gu[i + (long)N * j];
                    ad[r] = Gd[j + (long)N * i];
                    bd[r] = Gd[i + (long)N * j];
                }
            }
            __syncthreads();
            // a_k = G[:, i] + A B[i, :]^T,  b_k = G[i, :]^T + B A[i, :]^T
#pragma unroll
            for (int m = 0; m < KD; m++) {
                if (m < k) {  // uniform
                    const T bui = Bu[m][i], bdi = Bd[m][i], rau = rAu[m], rad = rAd[m];
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        const int j = tid + NT * r;
                        au[r] += Au[r][m] * bui;
                        bu[r] += Bu[m][j] * rau;
                        ad[r] += Ad[r][m] * bdi;
                        bd[r] += Bd[m][j] * rad;
                    }
                }
            }
            const T fu = s_fu, fd = s_fd;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int j = tid + NT * r;
                if (j == i) { au[r] -= one; ad[r] -= one; }
                au[r] = au[r] * fu;
                ad[r] = ad[r] * fd;
                if (j < N) {
                    du[j] += au[r] * bu[r];
                    dd[j] += ad[r] * bd[r];
                } else {
                    au[r] = zero; bu[r] = zero; ad[r] = zero; bd[r] = zero;
                }
#pragma unroll
                for (int m = 0; m < KD; m++)
                    if (m == k) { Au[r][m] = au[r]; Ad[r][m] = ad[r]; }
                Bu[k][j] = bu[r];
                Bd[k][j] = bd[r];
            }
            __syncthreads();
        }
        if (k > 0) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                if constexpr (is_cplx<T>::value) {  // every lane takes part
                    upd3_flush_mfma_z<KD, W>(N, tid, NT * r, Gu, Au[r], Bu, k, du);
                    upd3_flush_mfma_z<KD, W>(N, tid, NT * r, Gd, Ad[r], Bd, k, dd);
                } else {  // every lane takes part (the A fragments are gathered)
                    upd3_flush_mfma<KD, W>(N, tid, NT * r, Gu, Au[r], Bu, k, du);
                    upd3_flush_mfma<KD, W>(N, tid, NT * r, Gd, Ad[r], Bd, k, dd);
                }
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
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int j = tid + NT * r;
        if (j < N) {
            const int sv = hsl[j];
            hs[j] = sv;
            esu[j] = el[j + N * sv];
            esd[j] = el[j + N * (1 - sv)];
            isu[j] = el[j + N * (1 - sv)];
            isd[j] = el[j + N * sv];
        }
    }
    if (tid == 0) {
        a.phase[c] = phase;
        a.n_acc[c] += nacc;
        if (a.n_acc_sl) a.n_acc_sl[(long)c * a.L + a.l] += nacc;
    }
}
