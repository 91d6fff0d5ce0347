// dq_update3.hpp -- Metropolis sweep over one time slice with delayed rank-KD
// updates (src/updates.c:7-84) for N <= 256, delay buffers split between
// registers and LDS so that the rank can be 32 (16 complex).
//
// Each flush G += A B^T reads and writes both spins' G (4 N^2 values per
// chain) and runs 4 N^2 KD flops on one CU's matrix cores; the number of
// flushes falls as 1/KD.  Per chain the time splits into the flushes (about
// 40%, bound by the CU's FP64 matrix rate and the latency of its G traffic)
// and the serial per-acceptance chain (site scan, column and row fetch of G,
// the gemv over the filled slots).  One 256-thread workgroup per chain, both spins,
// thread t owns rows j = t + 256 r (r < R; R = 2 up to N = 512, rank 16):
//   * its row of A (A[j][0..KD)) lives in registers -- the flush needs exactly
//     that row, and a new column enters at a run-time slot by selects;
//   * B lives in LDS as B[m][j] (128 KiB for both spins): the flush and the
//     gemv read B[m][c] / B[m][i] as broadcasts and B[m][j] without bank
//     conflicts;
//   * on an acceptance at site i the owner of row i publishes A[i][0..k)
//     through LDS.
// Wave 0 scans the precomputed site order from LDS, 64 positions at a time;
// only an acceptance touches G.  Slots past the filled ones hold zeros; the
// gemv walks the slots in unbranched chunks up to the one holding the last
// filled slot, from zero while G's column and row are in flight.
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
// The chunks are software-pipelined: the next chunk's G values are requested
// before the current chunk's MFMAs, so each wave keeps its matrix core busy
// while the loads are in flight (one wave per SIMD: nothing else hides the
// memory latency).  A rank-32 flush of both spins of a 256-site chain is
// 8.4 MFLOP, about 30 us of one CU's FP64 matrix rate.
// A chunk is in range when its 64 rows and 16 MB columns all lie inside the
// matrix (uniform per wave); then its loads and stores are straight-line code,
// which lets the next chunk's loads stay in flight across this chunk's MFMAs.
template <int MB>
__device__ inline void upd3_chunk_load(int N, int rl, int lane, int c0, bool full, const double* G,
                                       dbl4 (&acc)[MB][4])
{
    if (full) {
        const double* g = G + rl + N * (c0 + (lane >> 4));
#pragma unroll
        for (int mb = 0; mb < MB; mb++)
#pragma unroll
            for (int b = 0; b < 4; b++)
#pragma unroll
                for (int e = 0; e < 4; e++) acc[mb][b][e] = g[16 * b + N * (16 * mb + 4 * e)];
        return;
    }
#pragma unroll
    for (int mb = 0; mb < MB; mb++)
#pragma unroll
        for (int b = 0; b < 4; b++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int r = rl + 16 * b, cc = c0 + 16 * mb + 4 * e + (lane >> 4);
                acc[mb][b][e] = (r < N && cc < N) ? G[r + N * cc] : 0.0;
            }
}

template <int KD, int W, int MB>
__device__ inline void upd3_chunk_mfma(int lane, int c0, const double (&af)[4][KD / 4],
                                       const double (*B)[W], int kused, dbl4 (&acc)[MB][4])
{
#pragma unroll
    for (int kb = 0; kb < KD / 4; kb++) {
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
}

// Stores a chunk; when the chunk's columns meet the wave's rows (dg, uniform)
// the diagonal elements (at most one per row b of a lane) go to d[] after the
// stores, selected without branches.
template <int MB>
__device__ inline void upd3_chunk_store(int N, int rl, int lane, int c0, bool full, bool dg,
                                        double* G, double* d, const dbl4 (&acc)[MB][4])
{
    double dv[4] = {0.0, 0.0, 0.0, 0.0};
    bool df[4] = {false, false, false, false};
    if (full) {
        double* g = G + rl + N * (c0 + (lane >> 4));
#pragma unroll
        for (int mb = 0; mb < MB; mb++)
#pragma unroll
            for (int b = 0; b < 4; b++)
#pragma unroll
                for (int e = 0; e < 4; e++) g[16 * b + N * (16 * mb + 4 * e)] = acc[mb][b][e];
    } else {
#pragma unroll
        for (int mb = 0; mb < MB; mb++)
#pragma unroll
            for (int b = 0; b < 4; b++)
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int r = rl + 16 * b, cc = c0 + 16 * mb + 4 * e + (lane >> 4);
                    if (r < N && cc < N) G[r + N * cc] = acc[mb][b][e];
                }
    }
    if (!dg) return;
#pragma unroll
    for (int mb = 0; mb < MB; mb++)
#pragma unroll
        for (int b = 0; b < 4; b++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int r = rl + 16 * b, cc = c0 + 16 * mb + 4 * e + (lane >> 4);
                const bool on = (r == cc);
                dv[b] = on ? acc[mb][b][e] : dv[b];
                df[b] = df[b] || on;
            }
#pragma unroll
    for (int b = 0; b < 4; b++)
        if (df[b] && rl + 16 * b < N) d[rl + 16 * b] = dv[b];
}

template <int KD, int W, int MB, bool PIPE>
__device__ inline void upd3_flush_mfma(int N, int j, int row0, double* G, const double (&A)[KD],
                                       const double (*B)[W], int kused, double* d)
{
    constexpr int KB = KD / 4;
    const int lane = j & 63, w = __builtin_amdgcn_readfirstlane(j >> 6);  // wave-uniform
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
    constexpr int CW = 16 * MB;  // MB 16-column blocks per chunk
    const bool rows_in = row0 + 64 * w + 64 <= N;  // uniform per wave
    auto full = [&](int c0) { return rows_in && c0 + CW <= N; };
    const int wr0 = row0 + 64 * w;  // the wave's rows [wr0, wr0 + 64)
    auto dg = [&](int c0) { return c0 < wr0 + 64 && wr0 < c0 + CW; };
    dbl4 x0[MB][4], x1[MB][4];
    if constexpr (!PIPE) {  // one chunk in flight (two waves per SIMD: 256 registers)
        for (int c0 = 0; c0 < N; c0 += CW) {
            upd3_chunk_load<MB>(N, rl, lane, c0, full(c0), G, x0);
            upd3_chunk_mfma<KD, W, MB>(lane, c0, af, B, kused, x0);
            upd3_chunk_store<MB>(N, rl, lane, c0, full(c0), dg(c0), G, d, x0);
        }
        return;
    }
    upd3_chunk_load<MB>(N, rl, lane, 0, full(0), G, x0);
    // two chunks per trip, ping-ponging between the register sets
    for (int c0 = 0; c0 < N; c0 += 2 * CW) {
        if (c0 + CW < N) upd3_chunk_load<MB>(N, rl, lane, c0 + CW, full(c0 + CW), G, x1);
        upd3_chunk_mfma<KD, W, MB>(lane, c0, af, B, kused, x0);
        upd3_chunk_store<MB>(N, rl, lane, c0, full(c0), dg(c0), G, d, x0);
        if (c0 + CW >= N) break;
        if (c0 + 2 * CW < N) upd3_chunk_load<MB>(N, rl, lane, c0 + 2 * CW, full(c0 + 2 * CW), G, x0);
        upd3_chunk_mfma<KD, W, MB>(lane, c0 + CW, af, B, kused, x1);
        upd3_chunk_store<MB>(N, rl, lane, c0 + CW, full(c0 + CW), dg(c0 + CW), G, d, x1);
    }
}

// The complex flush on the matrix cores: as upd3_flush_mfma with the real and
// imaginary parts as separate real operands, four MFMAs per block,
//   re += Br Ar - Bi Ai,   im += Br Ai + Bi Ar,
// one 16-column block per chunk (the accumulators are twice the real ones),
// pipelined the same way.
__device__ inline void upd3z_chunk_load(int N, int rl, int lane, int c0, bool full, const zd* G,
                                        dbl4 (&cr)[4], dbl4 (&ci)[4])
{
    if (full) {
        const zd* g = G + rl + N * (c0 + (lane >> 4));
#pragma unroll
        for (int b = 0; b < 4; b++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const zd v = g[16 * b + N * 4 * e];
                cr[b][e] = v.x;
                ci[b][e] = v.y;
            }
        return;
    }
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int r = rl + 16 * b, cc = c0 + 4 * e + (lane >> 4);
            const zd v = (r < N && cc < N) ? G[r + N * cc] : mkz(0.0, 0.0);
            cr[b][e] = v.x;
            ci[b][e] = v.y;
        }
}

template <int KD, int W>
__device__ inline void upd3z_chunk_mfma(int lane, int c0, const double (&ar)[4][KD / 4],
                                        const double (&ai)[4][KD / 4], const zd (*B)[W],
                                        int kused, dbl4 (&cr)[4], dbl4 (&ci)[4])
{
#pragma unroll
    for (int kb = 0; kb < KD / 4; kb++) {
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
}

__device__ inline void upd3z_chunk_store(int N, int rl, int lane, int c0, bool full, bool dg, zd* G,
                                         zd* d, const dbl4 (&cr)[4], const dbl4 (&ci)[4])
{
    if (full) {
        zd* g = G + rl + N * (c0 + (lane >> 4));
#pragma unroll
        for (int b = 0; b < 4; b++)
#pragma unroll
            for (int e = 0; e < 4; e++) g[16 * b + N * 4 * e] = mkz(cr[b][e], ci[b][e]);
    } else {
#pragma unroll
        for (int b = 0; b < 4; b++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int r = rl + 16 * b, cc = c0 + 4 * e + (lane >> 4);
                if (r < N && cc < N) G[r + N * cc] = mkz(cr[b][e], ci[b][e]);
            }
    }
    if (!dg) return;
    double dr[4] = {0.0, 0.0, 0.0, 0.0}, di[4] = {0.0, 0.0, 0.0, 0.0};
    bool df[4] = {false, false, false, false};
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int r = rl + 16 * b, cc = c0 + 4 * e + (lane >> 4);
            const bool on = (r == cc);
            dr[b] = on ? cr[b][e] : dr[b];
            di[b] = on ? ci[b][e] : di[b];
            df[b] = df[b] || on;
        }
#pragma unroll
    for (int b = 0; b < 4; b++)
        if (df[b] && rl + 16 * b < N) d[rl + 16 * b] = mkz(dr[b], di[b]);
}

template <int KD, int W>
__device__ inline void upd3_flush_mfma_z(int N, int j, int row0, zd* G, const zd (&A)[KD],
                                         const zd (*B)[W], int kused, zd* d)
{
    constexpr int KB = KD / 4;
    const int lane = j & 63, w = __builtin_amdgcn_readfirstlane(j >> 6);  // wave-uniform
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
    constexpr int CW = 16;
    const bool rows_in = row0 + 64 * w + 64 <= N;  // uniform per wave
    auto full = [&](int c0) { return rows_in && c0 + CW <= N; };
    const int wr0 = row0 + 64 * w;
    auto dg = [&](int c0) { return c0 < wr0 + 64 && wr0 < c0 + CW; };
    dbl4 r0[4], i0[4], r1[4], i1[4];
    upd3z_chunk_load(N, rl, lane, 0, full(0), G, r0, i0);
    for (int c0 = 0; c0 < N; c0 += 2 * CW) {
        if (c0 + CW < N) upd3z_chunk_load(N, rl, lane, c0 + CW, full(c0 + CW), G, r1, i1);
        upd3z_chunk_mfma<KD, W>(lane, c0, ar, ai, B, kused, r0, i0);
        upd3z_chunk_store(N, rl, lane, c0, full(c0), dg(c0), G, d, r0, i0);
        if (c0 + CW >= N) break;
        if (c0 + 2 * CW < N) upd3z_chunk_load(N, rl, lane, c0 + 2 * CW, full(c0 + 2 * CW), G, r0, i0);
        upd3z_chunk_mfma<KD, W>(lane, c0 + CW, ar, ai, B, kused, r1, i1);
        upd3z_chunk_store(N, rl, lane, c0 + CW, full(c0 + CW), dg(c0 + CW), G, d, r1, i1);
    }
}

// NT threads, R rows per thread (thread t owns rows t + NT r), N <= NT R.
template <typename T, int R, int KD, int NT = UPD3_THREADS>
__global__ __launch_bounds__(NT) void k_update_lds(Upd2Args<T> a)
{
    constexpr int W = NT * R;
    constexpr int MC = KD >= 32 ? 8 : 4;  // slots per unbranched chunk
    static_assert(KD % MC == 0 && (MC & (MC - 1)) == 0, "delay rank");
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
    int ii = 0;  // wave 0's scan position
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
            // the scan: wave 0 tests 64 positions of the site order at once
            // with the current diagonal and takes the first accepting one
            // (the draws do not depend on acceptances); its lane computes the
            // factors, flips the field and hands the phase factor to lane 0
            if (tid < 64) {
                int site = -1;
                while (ii < N) {  // ii is uniform over wave 0
                    const int pos = ii + tid;
                    bool acc = false;
                    int i = 0, sv = 0;
                    double delu = 0.0, deld = 0.0, absprob = 1.0;
                    T ru = one, rd = one, prob = one;
                    if (pos < N) {
                        i = ordl[pos];
                        sv = hsl[i];
                        delu = sv ? dl1[i] : dl0[i];
                        deld = sv ? dl0[i] : dl1[i];
                        if (!(delu == 0.0 && deld == 0.0)) {
                            ru = one + (one - du[i]) * delu;
                            rd = one + (one - dd[i]) * deld;
                            prob = ru * rd;
                            absprob = absv(prob);
                            acc = ul[pos] < absprob;
                        }
                    }
                    const unsigned long long bal = __ballot(acc);
                    if (bal == 0ull) {
                        ii += 64;
                        continue;
                    }
                    const int first = __ffsll((long long)bal) - 1;  // uniform
                    T fac = zero;
                    if (tid == first) {
                        s_fu = rdiv(from_re<T>(delu), ru);
                        s_fd = rdiv(from_re<T>(deld), rd);
                        hsl[i] = 1 - sv;
                        fac = prob * (1.0 / absprob);
                    }
                    site = __builtin_amdgcn_readlane(i, first);
                    fac = bcast(fac, first);
                    if (tid == 0) {
                        phase = phase * fac;
                        nacc++;
                    }
                    ii += first + 1;
                    break;
                }
                if (tid == 0) s_site = site;
            }
            __syncthreads();
            const int i = s_site;
            if (i < 0) break;
            // the owner of row i publishes its slots of A in chunks of MC, the
            // filled ones and the zeros after them up to the chunk's end
            if (tid == (i & (NT - 1))) {
#pragma unroll
                for (int r = 0; r < R; r++)
                    if (r == i / NT) {
#pragma unroll
                        for (int mc = 0; mc < KD; mc += MC)
                            if (mc < k) {  // uniform
#pragma unroll
                                for (int m = mc; m < mc + MC; m++) {
                                    rAu[m] = Au[r][m];
                                    rAd[m] = Ad[r][m];
                                }
                            }
                    }
            }
            // G's column and row of site i; the gemv below runs from zero while
            // these loads are in flight, and adds them at the end
            T gau[R], gbu[R], gad[R], gbd[R];
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int j = tid + NT * r;
                gau[r] = zero; gbu[r] = zero; gad[r] = zero; gbd[r] = zero;
                if (j < N) {
                    gau[r] = Gu[j + (long)N * i];
                    gbu[r] = Gu[i + (long)N * j];
                    gad[r] = Gd[j + (long)N * i];
                    gbd[r] = Gd[i + (long)N * j];
                }
            }
            __syncthreads();
            // a_k = G[:, i] + A B[i, :]^T,  b_k = G[i, :]^T + B A[i, :]^T, over the
            // slots in chunks of MC with no branch inside a chunk, so that a
            // chunk's LDS reads are in flight together: the slots from k to the
            // chunk's end are zero in A, B and the published row, and add 0
            T au[R], bu[R], ad[R], bd[R];
#pragma unroll
            for (int r = 0; r < R; r++) { au[r] = zero; bu[r] = zero; ad[r] = zero; bd[r] = zero; }
#pragma unroll
            for (int mc = 0; mc < KD; mc += MC) {
                if (mc < k) {  // uniform
#pragma unroll
                    for (int m = mc; m < mc + MC; m++) {
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
            }
            const T fu = s_fu, fd = s_fd;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int j = tid + NT * r;
                au[r] += gau[r]; bu[r] += gbu[r]; ad[r] += gad[r]; bd[r] += gbd[r];
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
                for (int mc = 0; mc < KD; mc += MC)
                    if ((k & ~(MC - 1)) == mc) {  // uniform: only k's chunk selects
#pragma unroll
                        for (int m = mc; m < mc + MC; m++)
                            if (m == k) { Au[r][m] = au[r]; Ad[r][m] = ad[r]; }
                    }
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
                    // pipelined 32-column chunks with one wave per SIMD; two
                    // waves per SIMD (256 registers each) keep one chunk
                    upd3_flush_mfma<KD, W, 2, (NT <= 256)>(N, tid, NT * r, Gu, Au[r], Bu, k, du);
                    upd3_flush_mfma<KD, W, 2, (NT <= 256)>(N, tid, NT * r, Gd, Ad[r], Bd, k, dd);
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
