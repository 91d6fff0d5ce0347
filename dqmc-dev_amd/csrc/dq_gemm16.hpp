// dq_gemm16.hpp -- real products on v_mfma_f64_16x16x4f64 with 128-bit LDS
// traffic (tile GT_W of dq_gemm.hpp's launcher).
//
//   C[b] = alpha * diag(r) * op(A) * diag(u) * op(B) * diag(w) + beta * C,
//   op(A) = A or A^T, op(B) = B or B^T, M % BM == 0 (BM = 128 or 64),
//   N % 64 == 0, K % 16 == 0, 16-byte aligned columns.
//
// 256 threads per BM x 64 tile of C, four waves in a 2 x 2 arrangement, each
// wave (BM / 2) x 32 = MI x 2 blocks of 16 x 16.  K advances in steps of 16
// through two LDS buffers (the next step's global loads in flight during the
// current step's MFMAs, one barrier per step).
//
// K order inside a step: MFMA sub-step s (0..3) gives the lanes of quarter q
// (lane / 16) the k index 4q + s, for both operands, so a lane needs
// k = 4q .. 4q + 3 of its row of op(A) and of its column of B.  LDS holds the
// tiles in that fragment order: per 16-row (16-column) block, two 1 KiB
// regions, region j holding at slot l the pair k = 4(l/16) + 2j, +1 of row
// (column) l % 16.  Every LDS access of the kernel is then one 16-byte slot
// per lane at slot = lane: a wave moves 1 KiB of contiguous LDS, the form that
// reads at the full 256 bytes per clock (rows 144 bytes apart, conflict-free
// by banks, took twice as long per read on gfx950).
//
// Global -> LDS: B's columns, and op(A)'s rows for op(A) = A^T, are
// k-contiguous in memory, so a thread moves k = 4q .. 4q + 3 of one column
// (row) with two 16-byte loads into its slot of both regions (u is applied to
// B here).  For op(A) = A the rows are m-contiguous: a thread loads a 2 x 2
// block (rows 2p, 2p + 1 of columns k, k + 1: two 16-byte loads, the 8 lanes
// of a column pair reading 128 contiguous bytes), which is exactly the
// (k, k + 1) pair of two row slots; lanes are numbered so that the 16-byte
// slots of every 8 consecutive lanes of each of the two LDS writes fall in 8
// consecutive bank groups (no bank conflicts by the counters).
// op(B) = B^T (n-contiguous) moves the same way.
#pragma once
#include "dq_common.hpp"
// included by dq_gemm.hpp after GemmArgs

#define GW_BN 64
#define GW_BK 16

// MI: 16-row blocks per wave along M (BM = 32 MI); TA: op(A) = A^T;
// TBT: op(B) = B^T
// (two workgroups per CU for the 128-row beta products: their C tile
// prefetch needs 240 registers)
template <int MI, bool TA, bool BETA, bool TBT = false>
__global__ __launch_bounds__(256, (BETA && MI == 4) ? 2 : 3) void k_gemm_w(GemmArgs<double> g,
                                                                          int tiles_m, int tiles_n,
                                                                          int nwg)
{
    constexpr int BM = 32 * MI, TB = MI / 2;  // TB: A blocks loaded per wave
    __shared__ __align__(16) double As[2][BM * GW_BK];
    __shared__ __align__(16) double Bs[2][GW_BN * GW_BK];

    // Workgroup ids are dealt round-robin over the 8 XCDs; renumber so that
    // each XCD works through a contiguous range of tiles (one matrix's tiles
    // share an L2).
    const int id = blockIdx.x;
    const int xcd = id & 7, per = nwg >> 3, rem = nwg & 7;
    const int lid = xcd * per + (xcd < rem ? xcd : rem) + (id >> 3);
    const int tm = lid % tiles_m;
    const int rest = lid / tiles_m;
    const int tn = rest % tiles_n;
    const int b = rest / tiles_n;
    const int m0 = tm * BM, n0 = tn * GW_BN;
    const double* A = g.A + (long)(b / g.adiv) * g.sA;
    const double* B = g.B + (long)(b / g.bdiv) * g.sB;
    double* C = g.C + (long)b * g.sC;
    const double* u = g.u ? g.u + (long)b * g.su : nullptr;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (g.idLc && g.idl >= g.idLc[b >> 1]) {  // identity factor (uniform; op N only)
        const double* X = (g.idside == 1) ? B : A;
        const int ldx = (g.idside == 1) ? g.ldb : g.lda;
        for (int e = tid; e < BM * GW_BN; e += 256) {
            const int row = m0 + e % BM, col = n0 + e / BM;
            double v = g.alpha * X[row + (long)ldx * col];
            if constexpr (BETA) v += g.beta * C[row + (long)g.ldc * col];
            C[row + (long)g.ldc * col] = v;
        }
        return;
    }
    const int q = lane >> 4, l16 = lane & 15;

    // A loads: op(A) = A^T rows 16 (TB wave + t) + l16, k = 4q .. 4q + 3;
    // op(A) = A rows 16 (TB wave + t) + 2p, +1 of columns k = 2kp, +1
    const int p2 = (lane >> 1) & 7, kp = (lane & 1) | ((lane >> 4) << 1);
    const double* Ag = TA ? A + 4 * q + (long)g.lda * (m0 + 16 * TB * wave + l16)
                          : A + (m0 + 16 * TB * wave + 2 * p2) + (long)g.lda * (2 * kp);
    // B loads: op(B) = B column 16 wave + l16, k = 4q .. 4q + 3; op(B) = B^T
    // columns 16 wave + 2p, +1 of k = 2kp, +1
    const double* Bg = TBT ? B + (n0 + 16 * wave + 2 * p2) + (long)g.ldb * (2 * kp)
                           : B + 4 * q + (long)g.ldb * (n0 + 16 * wave + l16);

    double ra[TB][4];  // op(A) = A^T: k = 4q .. 4q + 3; op(A) = A: [col][row] of the 2 x 2 block
    double2 rb[2];
    double2 ub0 = make_double2(1.0, 1.0), ub1 = ub0;
    auto gload = [&](int k0) {
#pragma unroll
        for (int t = 0; t < TB; t++) {
            if constexpr (TA) {
                const double* aa = Ag + (long)g.lda * (16 * t) + k0;
                const double2 a0 = *(const double2*)(aa), a1 = *(const double2*)(aa + 2);
                ra[t][0] = a0.x; ra[t][1] = a0.y; ra[t][2] = a1.x; ra[t][3] = a1.y;
            } else {
                const double* aa = Ag + 16 * t + (long)g.lda * k0;
                const double2 a0 = *(const double2*)(aa), a1 = *(const double2*)(aa + g.lda);
                ra[t][0] = a0.x; ra[t][1] = a0.y; ra[t][2] = a1.x; ra[t][3] = a1.y;
            }
        }
        if constexpr (TBT) {  // rb[c]: k = 2kp + c of columns 2p, 2p + 1
            const double* bb = Bg + (long)g.ldb * k0;
            rb[0] = *(const double2*)(bb);
            rb[1] = *(const double2*)(bb + g.ldb);
            if (u) ub0 = *(const double2*)(u + k0 + 2 * kp);
        } else {
            const double* bb = Bg + k0;
            rb[0] = *(const double2*)(bb);
            rb[1] = *(const double2*)(bb + 2);
            if (u) {
                ub0 = *(const double2*)(u + k0 + 4 * q);
                ub1 = *(const double2*)(u + k0 + 4 * q + 2);
            }
        }
    };
    // slot `lane` of region j of block blk: double offset ((2 blk + j) 64 + lane) 2
    auto slot = [&](int blk, int j) { return ((2 * blk + j) * 64 + lane) * 2; };
    const bool odd = kp & 1;
    auto sstore = [&](int buf) {
#pragma unroll
        for (int t = 0; t < TB; t++) {
            if constexpr (TA) {
                *(double2*)(&As[buf][slot(TB * wave + t, 0)]) = make_double2(ra[t][0], ra[t][1]);
                *(double2*)(&As[buf][slot(TB * wave + t, 1)]) = make_double2(ra[t][2], ra[t][3]);
            } else {
                // row 2 p2 + e, k = 2 kp + c: region kp & 1, slot 2 p2 + e + 16 (kp >> 1)
                double* base = &As[buf][(((2 * (TB * wave + t) + (kp & 1)) * 64) + 16 * (kp >> 1) +
                                         2 * p2) * 2];
                // ra[t][2 c + e]: column k + c, row 2 p2 + e; first row 2 p2 + odd
                const double f0 = odd ? ra[t][1] : ra[t][0], f1 = odd ? ra[t][3] : ra[t][2];
                const double s0 = odd ? ra[t][0] : ra[t][1], s1 = odd ? ra[t][2] : ra[t][3];
                *(double2*)(base + 2 * odd) = make_double2(f0, f1);
                *(double2*)(base + 2 - 2 * odd) = make_double2(s0, s1);
            }
        }
        if constexpr (TBT) {  // as op(A) = A, the pair k = 2kp, +1 scaled by u
            double* base = &Bs[buf][(((2 * wave + (kp & 1)) * 64) + 16 * (kp >> 1) + 2 * p2) * 2];
            const double f0 = odd ? rb[0].y : rb[0].x, f1 = odd ? rb[1].y : rb[1].x;
            const double s0 = odd ? rb[0].x : rb[0].y, s1 = odd ? rb[1].x : rb[1].y;
            *(double2*)(base + 2 * odd) = make_double2(f0 * ub0.x, f1 * ub0.y);
            *(double2*)(base + 2 - 2 * odd) = make_double2(s0 * ub0.x, s1 * ub0.y);
        } else {
            *(double2*)(&Bs[buf][slot(wave, 0)]) = make_double2(rb[0].x * ub0.x, rb[0].y * ub0.y);
            *(double2*)(&Bs[buf][slot(wave, 1)]) = make_double2(rb[1].x * ub1.x, rb[1].y * ub1.y);
        }
    };

    const int wr = wave >> 1, wc = wave & 1;
    const int wm = 16 * MI * wr, wn = 32 * wc;
    dbl4 acc[MI][2];
#pragma unroll
    for (int i = 0; i < MI; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = (dbl4){0.0, 0.0, 0.0, 0.0};
    // beta != 0: the C tile is requested before the K loop, so that its
    // latency hides under the MFMAs (the beta products here are the QR
    // chain's short-K updates, whose epilogue read was exposed)
    dbl4 cpre[BETA ? MI : 1][2];
    if constexpr (BETA) {
#pragma unroll
        for (int i = 0; i < MI; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int e = 0; e < 4; e++)
                    cpre[i][j][e] = C[(m0 + wm + 16 * i + l16) +
                                      (long)g.ldc * (n0 + wn + 16 * j + 4 * e + q)];
    }
    double2 av[MI][2], bv[2][2];
    auto reads = [&](int buf, int jj) {  // region jj: sub-steps 2 jj, 2 jj + 1
#pragma unroll
        for (int i = 0; i < MI; i++) av[i][jj] = *(const double2*)(&As[buf][slot(MI * wr + i, jj)]);
#pragma unroll
        for (int j = 0; j < 2; j++) bv[j][jj] = *(const double2*)(&Bs[buf][slot(2 * wc + j, jj)]);
    };
    auto mf = [&](int s) {  // MFMA sub-step s: k = 4q + s
#pragma unroll
        for (int i = 0; i < MI; i++) {
            const double a = (s & 1) ? av[i][s >> 1].y : av[i][s >> 1].x;
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const double bq = (s & 1) ? bv[j][s >> 1].y : bv[j][s >> 1].x;
                acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bq, a, acc[i][j], 0, 0, 0);
            }
        }
    };

    const int nk = g.K / GW_BK;
    // an upper-triangular A contributes nothing below row m0's diagonal
    const int kt0 = (!TA && g.a_upper) ? (m0 < g.K ? m0 : g.K) / GW_BK : 0;
    gload(kt0 * GW_BK);
    sstore(kt0 & 1);
    __syncthreads();
    for (int kt = kt0; kt < nk; kt++) {
        const int cur = kt & 1;
        const bool more = kt + 1 < nk;
        if (more) gload((kt + 1) * GW_BK);
        // the second region's reads go out after the first sub-step's MFMAs
        // (twelve reads at once stall the wave's issue) and land during the
        // second sub-step's
        reads(cur, 0);
        mf(0);
        __builtin_amdgcn_sched_barrier(0);
        reads(cur, 1);
        __builtin_amdgcn_sched_barrier(0);
        mf(1);
        // the next step's LDS writes between the MFMA halves (2% fewer cycles
        // than after them: the loads have landed and the LDS is idle)
        if (more) sstore(cur ^ 1);
        mf(2);
        mf(3);
        __syncthreads();
    }

    const double* rr = g.r ? g.r + (long)b * g.sr : nullptr;
    const double* ww = g.w ? g.w + (long)b * g.sw : nullptr;
#pragma unroll
    for (int i = 0; i < MI; i++) {
        const int row = m0 + wm + 16 * i + l16;
        const double rs = rr ? rr[row] : 1.0;
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int col = n0 + wn + 16 * j + 4 * e + q;
                double v = g.alpha * acc[i][j][e];
                if (rr) v = v * rs;
                if (ww) v = v * ww[col];
                const long o = row + (long)g.ldc * col;
                if constexpr (BETA) v += g.beta * cpre[i][j][e];
                C[o] = v;
            }
    }
}

// The shapes and alignments k_gemm_w takes (op(A) = A^T also for OP_C: the
// data are real); an identity factor needs op N
static inline bool gemm_w_fits(int opa, int opb, const GemmArgs<double>& g)
{
    const auto al = [](const void* p) { return ((unsigned long)p & 15) == 0; };
    return (opa == 0 || opb == 0) && ((opa == 0 && opb == 0) || !g.idLc) && g.M % 64 == 0 &&
           g.N % GW_BN == 0 &&
           g.K % GW_BK == 0 && g.M > 0 && g.N > 0 && g.K > 0 && g.lda % 2 == 0 &&
           g.ldb % 2 == 0 && g.sA % 2 == 0 && g.sB % 2 == 0 && al(g.A) && al(g.B) &&
           (!g.u || (al(g.u) && g.su % 2 == 0));
}

// 128-row tiles where M allows, else 64
static inline hipError_t gemm_w_launch(int opa, int opb, const GemmArgs<double>& g, int batch,
                                       hipStream_t s)
{
    const int bm = g.M % 128 == 0 ? 128 : 64;
    const int tm = g.M / bm, tn = g.N / GW_BN;
    const long nwg = (long)tm * tn * batch;
    if (nwg <= 0) return hipSuccess;
    if (nwg > 0x7fffffffL) return hipErrorInvalidValue;
    const int n = (int)nwg;
#define GWL(MI, TA, TB)                                                            \
    (g.beta_zero ? k_gemm_w<MI, TA, false, TB><<<n, 256, 0, s>>>(g, tm, tn, n) \
                 : k_gemm_w<MI, TA, true, TB><<<n, 256, 0, s>>>(g, tm, tn, n))
    if (opb != 0) {  // op(A) = A only
        if (bm == 128) GWL(4, false, true);
        else GWL(2, false, true);
    } else if (opa == 0) {
        if (bm == 128) GWL(4, false, false);
        else GWL(2, false, false);
    } else {
        if (bm == 128) GWL(4, true, false);
        else GWL(2, true, false);
    }
#undef GWL
    return hipGetLastError();
}
