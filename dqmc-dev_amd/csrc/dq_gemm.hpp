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
// Tiling for CDNA4: 256 threads = 4 wave64s per C tile (64x64 or 128x64),
// each wave a 2-D array of FP64 MFMA blocks; K is staged
// through double-buffered LDS.  Complex (zd) uses four real MFMAs per block on
// planar (re/im) LDS copies.
// Operand lane layout of v_mfma_f64_16x16x4f64 (wave64):
//   A[16x4]: lane l holds A[l%16][l/16];  B[4x16]: lane l holds B[l/16][l%16]
//   D[16x16]: lane l, element r holds D[4*r + l/16][l%16]
// and of v_mfma_f64_4x4x4f64 (four independent 4x4x4 blocks b, one double
// per lane):
//   A_b[i][k] at lane i + 4b + 16k;  B_b[k][j] at lane j + 4b + 16k;
//   D_b[i][j] at lane j + 4b + 16i
// (layouts verified on gfx950 with probes of known operands).  On gfx950 the
// 4x4x4 form sustains ~74 TFLOP/s in a register-only loop against ~49 for the
// 16x16x4 form, so the main tiles use it: the four blocks form a 2 x 2 grid of
// 4 x 4 sub-tiles, one instruction an 8 x 8 x 4 product (each operand value
// read by two lanes).  Its CBSZ/ABID broadcast controls do not act as a
// block broadcast for this instruction and are not used.
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
    // identity factor: for batch z with idl >= idLc[z / 2] the operand on side
    // idside (1 = A, 2 = B) is the identity, C = alpha X + beta C with X the
    // other operand, scales ignored (an identity slice of a padded chain);
    // idLc null: none.  Square products with op N only.
    const int* idLc; int idl; int idside;
    // 1: op(A) = A is upper triangular (zeros below the diagonal), so that a
    // tile of rows m0.. needs only k >= m0 (the 16x16x4 kernel skips the rest;
    // the other kernels compute the zeros)
    int a_upper;
};

// Tile shapes (rows x columns of C per 256-thread workgroup).  The engine's
// autotuner times the shapes on its own product sizes and keeps the fastest.
// GT_W: real products op(A) B on the 16x16x4 form with 128-bit LDS traffic
// (dq_gemm16.hpp), for the shapes it takes; other products fall back to 64x64.
enum { GT_AUTO = 0, GT_64 = 1, GT_128 = 2, GT_128x64 = 3, GT_W = 4 };

__device__ inline dbl4 mfma(double a, double b, dbl4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ inline double mfma4(double a, double b, double c)
{
    return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
}

// LDS image of one operand tile, BO (rows of C or columns of C) x BK, in the
// orientation of its global storage so that the global->LDS copy is linear:
//   OFAST: the tile's o index is contiguous in global memory -> rows [k][o],
//          pitch BO + 16 (the 4 k-rows read by one MFMA step hit disjoint banks)
//   else:  k is contiguous -> rows [o][k], pitch BK + 1.
template <int BO, int BK, bool OFAST>
struct GTile {
    static constexpr int P = OFAST ? BO + 16 : BK + 1;
    static constexpr int SIZE = OFAST ? BK * P : BO * P;
    __device__ static inline int at(int o, int k) { return OFAST ? k * P + o : o * P + k; }
};

// C tile = 32TM x 32TN, 4 waves in a 2x2 arrangement, each 16TM x 16TN made
// of TM x TN MFMA blocks; K advanced in steps of BK with two LDS buffers (the
// next step's global loads are in flight while the current one computes).
// The MFMA computes the transposed block (op(B)^T op(A)^T), so that a lane's
// accumulator entries run down a column of C and 16 lanes store 128
// contiguous bytes of the column-major result.
template <typename T, int OPA, int OPB, int TM, int TN, int BK, bool BETA, int MF = 0>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs<T> g, int tiles_m, int tiles_n, int nwg)
{
    constexpr bool Z = is_cplx<T>::value;
    constexpr int NP = Z ? 2 : 1;
    constexpr int BM = 32 * TM, BN = 32 * TN;
    constexpr bool M4 = MF == 1;
    constexpr bool AF = (OPA == OP_N), BF = (OPB != OP_N);
    using LA = GTile<BM, BK, AF>;
    using LB = GTile<BN, BK, BF>;
    constexpr int NA = BM * BK / 256, NB = BN * BK / 256;
    static_assert(NA * 256 == BM * BK && NB * 256 == BN * BK, "tile/thread mismatch");
    static_assert(2 * NP * (LA::SIZE + LB::SIZE) * 8 <= 65536, "LDS budget");
    __shared__ double As[2][NP][LA::SIZE];
    __shared__ double Bs[2][NP][LB::SIZE];
    __shared__ double Us[2][BK];  // inner scale u of the K step

    // Workgroup ids are dealt round-robin over the 8 XCDs; renumber so that
    // each XCD works through a contiguous range of tiles, which keeps the
    // tiles of one matrix (and their shared operand panels) in one L2.
    const int id = blockIdx.x;
    const int xcd = id & 7, per = nwg >> 3, rem = nwg & 7;
    const int lid = xcd * per + (xcd < rem ? xcd : rem) + (id >> 3);
    const int tm = lid % tiles_m;
    const int rest = lid / tiles_m;
    const int tn = rest % tiles_n;
    const int b = rest / tiles_n;
    const int m0 = tm * BM, n0 = tn * BN;
    const T* A = g.A + (long)(b / g.adiv) * g.sA;
    const T* B = g.B + (long)(b / g.bdiv) * g.sB;
    T* C = g.C + (long)b * g.sC;
    const double* u = g.u ? g.u + (long)b * g.su : nullptr;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (g.idLc && g.idl >= g.idLc[b >> 1]) {  // uniform over the workgroup
        const T* X = (g.idside == 1) ? B : A;
        const int ldx = (g.idside == 1) ? g.ldb : g.lda;
        for (int e = tid; e < BM * BN; e += 256) {
            const int row = m0 + e % BM, col = n0 + e / BM;
            if (row < g.M && col < g.N) {
                T v = g.alpha * X[row + (long)ldx * col];
                if constexpr (BETA) v += g.beta * C[row + (long)g.ldc * col];
                C[row + (long)g.ldc * col] = v;
            }
        }
        return;
    }
    const int l16 = lane & 15, l4 = lane >> 4;
    const int wm = (wave >> 1) * 16 * TM, wn = (wave & 1) * 16 * TN;

    dbl4 acc[NP][TM][TN];
#pragma unroll
    for (int p = 0; p < NP; p++)
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
            for (int j = 0; j < TN; j++) acc[p][i][j] = (dbl4){0.0, 0.0, 0.0, 0.0};
    // 4x4x4 form: the wave tile is (8 * 2TM) x (8 * 2TN) sub-tiles of 8 x 8;
    // lane l works on row 4 * (b & 1) + (l & 3) of the C rows of a sub-tile and
    // column 4 * (b >> 1) + (l & 3) of its C columns for the operands, with
    // b = (l >> 2) & 3; it holds C[row 4 * (b & 1) + (l & 3)][col 4 * (b >> 1) + (l >> 4)].
    constexpr int QM = 2 * TM, QN = 2 * TN;
    double acc4[NP][QN][QM];
#pragma unroll
    for (int p = 0; p < NP; p++)
#pragma unroll
        for (int i = 0; i < QN; i++)
#pragma unroll
            for (int j = 0; j < QM; j++) acc4[p][i][j] = 0.0;
    const int b4 = (lane >> 2) & 3;
    const int r4 = 4 * (b4 & 1) + (lane & 3), c4 = 4 * (b4 >> 1) + (lane & 3);

    // Raw global values of the next K step.  Conjugation is applied when the
    // values go to LDS and the inner scale u when the B operand is read from
    // LDS, so that no arithmetic waits on loads still in flight.
    T ra[NA], rb[NB];
    double uk = 1.0;
    auto gload = [&](int k0) {
#pragma unroll
        for (int q = 0; q < NA; q++) {
            const int idx = tid + 256 * q;
            const int o = AF ? idx % BM : idx / BK;
            const int k = AF ? idx / BM : idx % BK;
            const int gm = m0 + o, gk = k0 + k;
            T v = from_re<T>(0.0);
            if (gm < g.M && gk < g.K) v = AF ? A[gm + (long)g.lda * gk] : A[gk + (long)g.lda * gm];
            ra[q] = v;
        }
#pragma unroll
        for (int q = 0; q < NB; q++) {
            const int idx = tid + 256 * q;
            const int o = BF ? idx % BN : idx / BK;
            const int k = BF ? idx / BN : idx % BK;
            const int gn = n0 + o, gk = k0 + k;
            T v = from_re<T>(0.0);
            if (gn < g.N && gk < g.K) v = BF ? B[gn + (long)g.ldb * gk] : B[gk + (long)g.ldb * gn];
            rb[q] = v;
        }
        if (u && tid < BK) uk = (k0 + tid < g.K) ? u[k0 + tid] : 0.0;
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int q = 0; q < NA; q++) {
            const int idx = tid + 256 * q;
            const int o = AF ? idx % BM : idx / BK;
            const int k = AF ? idx / BM : idx % BK;
            As[buf][0][LA::at(o, k)] = re(ra[q]);
            if constexpr (Z) As[buf][NP - 1][LA::at(o, k)] = (OPA == OP_C) ? -im(ra[q]) : im(ra[q]);
        }
#pragma unroll
        for (int q = 0; q < NB; q++) {
            const int idx = tid + 256 * q;
            const int o = BF ? idx % BN : idx / BK;
            const int k = BF ? idx / BN : idx % BK;
            Bs[buf][0][LB::at(o, k)] = re(rb[q]);
            if constexpr (Z) Bs[buf][NP - 1][LB::at(o, k)] = (OPB == OP_C) ? -im(rb[q]) : im(rb[q]);
        }
        if (u && tid < BK) Us[buf][tid] = uk;
    };
    auto compute4 = [&](int buf) {
#pragma unroll
        for (int ks = 0; ks < BK; ks += 4) {
            const int kk = ks + l4;
            double xr[QM], xi[QM], yr[QN], yi[QN];
#pragma unroll
            for (int q = 0; q < QM; q++) {
                xr[q] = As[buf][0][LA::at(wm + 8 * q + r4, kk)];
                if constexpr (Z) xi[q] = As[buf][NP - 1][LA::at(wm + 8 * q + r4, kk)];
            }
#pragma unroll
            for (int p = 0; p < QN; p++) {
                yr[p] = Bs[buf][0][LB::at(wn + 8 * p + c4, kk)];
                if constexpr (Z) yi[p] = Bs[buf][NP - 1][LB::at(wn + 8 * p + c4, kk)];
            }
            if (u) {
                const double us = Us[buf][kk];
#pragma unroll
                for (int p = 0; p < QN; p++) {
                    yr[p] *= us;
                    if constexpr (Z) yi[p] *= us;
                }
            }
            // D_b[i][j] = sum_k y[i][k] x[k][j]: i runs over C columns, j over C rows
#pragma unroll
            for (int p = 0; p < QN; p++)
#pragma unroll
                for (int q = 0; q < QM; q++) acc4[0][p][q] = mfma4(yr[p], xr[q], acc4[0][p][q]);
            if constexpr (Z) {
#pragma unroll
                for (int p = 0; p < QN; p++)
#pragma unroll
                    for (int q = 0; q < QM; q++) acc4[NP - 1][p][q] = mfma4(yi[p], xr[q], acc4[NP - 1][p][q]);
#pragma unroll
                for (int p = 0; p < QN; p++)
#pragma unroll
                    for (int q = 0; q < QM; q++) acc4[0][p][q] = mfma4(-yi[p], xi[q], acc4[0][p][q]);
#pragma unroll
                for (int p = 0; p < QN; p++)
#pragma unroll
                    for (int q = 0; q < QM; q++) acc4[NP - 1][p][q] = mfma4(yr[p], xi[q], acc4[NP - 1][p][q]);
            }
        }
    };
    auto compute = [&](int buf) {
        if constexpr (M4) {
            compute4(buf);
            return;
        }
#pragma unroll
        for (int ks = 0; ks < BK; ks += 4) {
            const int kk = ks + l4;
            double ar[TM], br[TN], ai[TM], bi[TN];
#pragma unroll
            for (int i = 0; i < TM; i++) {
                ar[i] = As[buf][0][LA::at(wm + 16 * i + l16, kk)];
                if constexpr (Z) ai[i] = As[buf][NP - 1][LA::at(wm + 16 * i + l16, kk)];
            }
#pragma unroll
            for (int j = 0; j < TN; j++) {
                br[j] = Bs[buf][0][LB::at(wn + 16 * j + l16, kk)];
                if constexpr (Z) bi[j] = Bs[buf][NP - 1][LB::at(wn + 16 * j + l16, kk)];
            }
            if (u) {
                const double us = Us[buf][kk];
#pragma unroll
                for (int j = 0; j < TN; j++) {
                    br[j] *= us;
                    if constexpr (Z) bi[j] *= us;
                }
            }
            // (AB)^T = B^T A^T; complex: re = Br Ar - Bi Ai, im = Bi Ar + Br Ai
#pragma unroll
            for (int i = 0; i < TM; i++)
#pragma unroll
                for (int j = 0; j < TN; j++) acc[0][i][j] = mfma(br[j], ar[i], acc[0][i][j]);
            if constexpr (Z) {
#pragma unroll
                for (int i = 0; i < TM; i++)
#pragma unroll
                    for (int j = 0; j < TN; j++) acc[NP - 1][i][j] = mfma(bi[j], ar[i], acc[NP - 1][i][j]);
#pragma unroll
                for (int i = 0; i < TM; i++)
#pragma unroll
                    for (int j = 0; j < TN; j++) acc[0][i][j] = mfma(-bi[j], ai[i], acc[0][i][j]);
#pragma unroll
                for (int i = 0; i < TM; i++)
#pragma unroll
                    for (int j = 0; j < TN; j++) acc[NP - 1][i][j] = mfma(br[j], ai[i], acc[NP - 1][i][j]);
            }
        }
    };

    const int nk = (g.K + BK - 1) / BK;
    const double* rr = g.r ? g.r + (long)b * g.sr : nullptr;
    const double* ww = g.w ? g.w + (long)b * g.sw : nullptr;
    gload(0);
    sstore(0);
    __syncthreads();
    for (int kt = 0; kt < nk; kt++) {
        const int cur = kt & 1;
        if (kt + 1 < nk) gload((kt + 1) * BK);
        compute(cur);
        if (kt + 1 < nk) sstore(cur ^ 1);
        __syncthreads();
    }

    if constexpr (M4) {
#pragma unroll
        for (int p = 0; p < QN; p++)
#pragma unroll
            for (int q = 0; q < QM; q++) {
                const int row = m0 + wm + 8 * q + r4;
                const int col = n0 + wn + 8 * p + 4 * (b4 >> 1) + l4;
                if (row < g.M && col < g.N) {
                    T v;
                    if constexpr (Z) v = mkz(acc4[0][p][q], acc4[NP - 1][p][q]);
                    else v = acc4[0][p][q];
                    v = g.alpha * v;
                    if (rr) v = v * rr[row];
                    if (ww) v = v * ww[col];
                    const long o = row + (long)g.ldc * col;
                    if constexpr (BETA) v += g.beta * C[o];
                    C[o] = v;
                }
            }
        return;
    }
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int row = m0 + wm + 16 * i + l16;
                const int col = n0 + wn + 16 * j + 4 * e + l4;
                if (row < g.M && col < g.N) {
                    T v;
                    if constexpr (Z) v = mkz(acc[0][i][j][e], acc[NP - 1][i][j][e]);
                    else v = acc[0][i][j][e];
                    v = g.alpha * v;
                    if (rr) v = v * rr[row];
                    if (ww) v = v * ww[col];
                    const long o = row + (long)g.ldc * col;
                    if constexpr (BETA) v += g.beta * C[o];
                    C[o] = v;
                }
            }
}

#include "dq_gemm16.hpp"

template <typename T, int TM, int TN, int BK, int MF>
static hipError_t gemm_launch_tile(int opa, int opb, const GemmArgs<T>& g, int batch, hipStream_t s)
{
    constexpr int BM = 32 * TM, BN = 32 * TN;
    const int tm = (g.M + BM - 1) / BM, tn = (g.N + BN - 1) / BN;
    const long nwg = (long)tm * tn * batch;
    if (nwg <= 0) return hipSuccess;
    if (nwg > 0x7fffffffL) return hipErrorInvalidValue;
    const int n = (int)nwg;
#define GL(a, b)                                                              \
    (g.beta_zero ? k_gemm<T, a, b, TM, TN, BK, false, MF><<<n, 256, 0, s>>>(g, tm, tn, n) \
                 : k_gemm<T, a, b, TM, TN, BK, true, MF><<<n, 256, 0, s>>>(g, tm, tn, n))
    if constexpr (!is_cplx<T>::value) {
        // conjugation is the identity on real data
        const bool ta = opa != OP_N, tb = opb != OP_N;
        if (!ta && !tb) GL(OP_N, OP_N);
        else if (!ta) GL(OP_N, OP_T);
        else if (!tb) GL(OP_T, OP_N);
        else GL(OP_T, OP_T);
    } else {
        if (opa == OP_N) {
            if (opb == OP_N) GL(OP_N, OP_N);
            else if (opb == OP_T) GL(OP_N, OP_T);
            else GL(OP_N, OP_C);
        } else if (opa == OP_T) {
            if (opb == OP_N) GL(OP_T, OP_N);
            else if (opb == OP_T) GL(OP_T, OP_T);
            else GL(OP_T, OP_C);
        } else {
            if (opb == OP_N) GL(OP_C, OP_N);
            else if (opb == OP_T) GL(OP_C, OP_T);
            else GL(OP_C, OP_C);
        }
    }
#undef GL
    return hipGetLastError();
}

// Launch with tile shape `tile` (GT_*).  GT_W takes the real op-N products
// whose shapes and alignments it fits (gemm_w_fits); GT_128 is the 128 x 64
// tile on the 4x4x4 form with K steps of 16 (GT_128x64 steps by 8); GT_AUTO
// and every product a shape does not take fall back to 64 x 64 on the 4x4x4
// form.  (A 128 x 128 tile on the 4x4x4 form needs 268 registers, one wave
// per SIMD, and was the slowest shape measured.)
template <typename T>
static hipError_t gemm_launch(int opa, int opb, const GemmArgs<T>& g, int batch, hipStream_t s,
                              int tile)
{
    constexpr bool Z = is_cplx<T>::value;
    if constexpr (!Z) {
        if (tile == GT_W && gemm_w_fits(opa, opb, g)) return gemm_w_launch(opa, opb, g, batch, s);
    }
    if (tile == GT_128 && g.M > 64)  // 128 x 64 with K steps of 16 (8 complex)
        return gemm_launch_tile<T, 4, 2, Z ? 8 : 16, 1>(opa, opb, g, batch, s);
    if (tile == GT_128x64 && g.M > 64)
        return gemm_launch_tile<T, 4, 2, 8, 1>(opa, opb, g, batch, s);
    return gemm_launch_tile<T, 2, 2, Z ? 8 : 16, 1>(opa, opb, g, batch, s);
}
